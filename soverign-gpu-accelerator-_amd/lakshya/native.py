"""ctypes binding of liblakshya.so (include/lakshya.h).  The product path: every call goes to the
HIP library; nothing here falls back to a CPU implementation -- a missing library raises."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "liblakshya.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p

ERRORS = {1: ValueError, 2: MemoryError, 3: RuntimeError, 5: RuntimeError}


class KKTStats(C.Structure):
    _fields_ = [("nsuper", C.c_int64), ("nnz_L", C.c_int64), ("nnz_L_exact", C.c_int64),
                ("npairs", C.c_int64), ("nsegments", C.c_int64), ("panel_elems", C.c_int64),
                ("levels", C.c_int32), ("launches_factor", C.c_int32),
                ("launches_solve", C.c_int32), ("pad", C.c_int32),
                ("flops_exact", C.c_double), ("flops_panel", C.c_double),
                ("symbolic_ms", C.c_double), ("factorizations", C.c_int64),
                ("solves", C.c_int64), ("refinement_steps", C.c_int64)]


class KKTSolveInfo(C.Structure):
    _fields_ = [("success", C.c_int32), ("retries", C.c_int32),
                ("refinement_steps", C.c_int32), ("pad", C.c_int32),
                ("delta", C.c_double), ("residual_inf", C.c_double)]


class IpmOptions(C.Structure):
    _fields_ = [("stop_rule", C.c_int32), ("max_iter", C.c_int32), ("tol", C.c_double),
                ("gap_tol", C.c_double), ("precision", C.c_int32),
                ("mixed_max_refine", C.c_int32), ("mixed_fp64_spread", C.c_double),
                ("ordering", C.c_int32), ("use_graphs", C.c_int32), ("verbose", C.c_int32),
                ("pad", C.c_int32), ("warm_floor", C.c_double)]


class IpmResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("objective", C.c_double),
                ("primal_residual", C.c_double), ("dual_residual", C.c_double),
                ("duality_gap", C.c_double), ("dual_objective", C.c_double),
                ("rel_gap", C.c_double), ("rel_primal_residual", C.c_double),
                ("rel_dual_residual", C.c_double), ("factorizations", C.c_int32),
                ("fp32_factorizations", C.c_int32), ("fp64_factorizations", C.c_int32),
                ("solves", C.c_int32), ("refinement_steps", C.c_int32), ("fail_code", C.c_int32),
                ("setup_ms", C.c_double), ("solve_ms", C.c_double), ("factor_ms", C.c_double),
                ("nnz_L", C.c_double), ("flops_per_factorization", C.c_double)]


class BatchOptions(C.Structure):
    _fields_ = [("stop_rule", C.c_int32), ("precision", C.c_int32), ("max_iter", C.c_int32),
                ("mixed_max_refine", C.c_int32), ("tol", C.c_double), ("gap_tol", C.c_double),
                ("mixed_fp64_spread", C.c_double), ("wg_per_cu", C.c_int32),
                ("iters_per_launch", C.c_int32), ("use_graph", C.c_int32), ("order", C.c_int32),
                ("warm_floor", C.c_double), ("warm_mu", C.c_double),
                ("warm_max_infeas", C.c_double)]


class BatchSummary(C.Structure):
    _fields_ = [("count", C.c_int32), ("optimal", C.c_int32), ("launches", C.c_int32),
                ("graph_replays", C.c_int32), ("ms", C.c_double), ("flops_chol", C.c_double),
                ("wg_per_cu", C.c_int32), ("iters_per_launch", C.c_int32), ("grid", C.c_int32),
                ("pad", C.c_int32)]


class BatchResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("fail_code", C.c_int32),
                ("factorizations", C.c_int32), ("fp32_factorizations", C.c_int32),
                ("solves", C.c_int32), ("pcg_steps", C.c_int32), ("pad", C.c_int32),
                ("objective", C.c_double), ("dual_objective", C.c_double),
                ("rel_gap", C.c_double), ("rel_primal", C.c_double), ("rel_dual", C.c_double),
                ("primal_residual", C.c_double), ("dual_residual", C.c_double), ("mu", C.c_double),
                ("ms_init", C.c_double), ("ms_other", C.c_double), ("ms_assemble", C.c_double),
                ("ms_factor", C.c_double), ("ms_solve", C.c_double),
                ("warm_infeas", C.c_double), ("warm", C.c_int32), ("pad2", C.c_int32)]


STATUS = {0: "Optimal", 1: "Infeasible", 2: "Unbounded", 3: "IterationLimit"}

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"liblakshya.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.lk_last_error.restype = C.c_char_p
    sig = {
        "lk_device_count": [C.POINTER(C.c_int)],
        "lk_arena_create": [C.c_size_t, C.POINTER(_vp)],
        "lk_arena_alloc": [_vp, C.c_size_t, C.POINTER(_vp)],
        "lk_arena_free": [_vp, _vp],
        "lk_arena_occupancy": [_vp, C.POINTER(C.c_size_t)],
        "lk_arena_destroy": [_vp],
        "lk_order_reference": [C.c_int32, C.c_int32, _i32p, _i32p, _i32p],
        "lk_order_amd": [C.c_int32, C.c_int32, _i32p, _i32p, _i32p],
        "lk_symbolic": [C.c_int32, C.c_int32, _i32p, _i32p, _i32p, _vp, _vp, _vp, _vp,
                        C.c_int64, C.POINTER(C.c_int64)],
        "lk_supernodal": [C.c_int32, C.c_int32, _i32p, _i32p, _vp, C.c_int32, C.c_double,
                          C.POINTER(KKTStats), _vp, _vp, C.POINTER(C.c_int32)],
        "lk_kkt_create": [_vp, C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _vp, C.c_int32,
                          C.c_int32, C.POINTER(_vp)],
        "lk_kkt_factor": [_vp, _vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32)],
        "lk_kkt_solve": [_vp, _vp, C.c_int32, C.c_double, C.c_int32,
                         C.POINTER(KKTSolveInfo)],
        "lk_kkt_get_L_csr": [_vp, _f64p, C.c_int64],
        "lk_kkt_permutation": [_vp, _i32p],
        "lk_kkt_get_stats": [_vp, C.POINTER(KKTStats)],
        "lk_kkt_stream": [_vp, C.POINTER(_vp)],
        "lk_kkt_destroy": [_vp],
        "lk_ipm_default_options": [C.POINTER(IpmOptions)],
        "lk_ruiz_scaling": [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                            _f64p],
        "lk_ipm_solve": [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _vp, _vp,
                         C.POINTER(IpmOptions), _vp, _vp, _vp, C.POINTER(IpmResult), _vp, _vp,
                         _vp],
        "lk_ipm_create": [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p,
                          C.POINTER(IpmOptions), C.POINTER(_vp)],
        "lk_ipm_run": [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(IpmResult), _vp, _vp, _vp],
        "lk_ipm_destroy": [_vp],
        "lk_batch_default_options": [C.POINTER(BatchOptions)],
        "lk_batch_create": [C.c_int32, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p,
                            C.POINTER(BatchOptions), C.POINTER(_vp)],
        "lk_batch_solve": [_vp, C.POINTER(BatchSummary)],
        "lk_batch_autotune": [_vp, C.c_int32, C.POINTER(BatchSummary)],
        "lk_batch_results": [_vp, _vp],
        "lk_batch_solution": [_vp, C.c_int32, _f64p, _f64p, _f64p],
        "lk_batch_crossover": [_vp, _i32p, _f64p, _i32p, _i32p],
        "lk_crossover": [C.c_int32, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                         _i32p, _f64p, _i32p, _i32p],
        "lk_simplex": [C.c_int32, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int32,
                       _i32p, _f64p, _f64p, _i32p, _i32p, _f64p],
        "lk_batch_destroy": [_vp],
        "lk_batch_set_warm_start": [_vp, _vp, _vp, _vp, _vp],
        "lk_batch_solve_noalloc": [C.c_int32, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p,
                                   C.POINTER(BatchOptions), _f64p, _f64p, _f64p, _vp,
                                   C.POINTER(BatchSummary)],
        "lk_mlp_create": [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                          C.POINTER(_vp)],
        "lk_mlp_param_count": [_vp, C.POINTER(C.c_int64)],
        "lk_mlp_set_params": [_vp, _f32p],
        "lk_mlp_get_params": [_vp, _f32p],
        "lk_mlp_forward": [_vp, _f32p, C.c_int32, _f32p],
        "lk_mlp_gradients": [_vp, _f32p, _f32p, C.c_int32, _f32p, C.POINTER(C.c_double)],
        "lk_mlp_set_dataset": [_vp, _f32p, _f32p, C.c_int32],
        "lk_mlp_train": [_vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _f64p],
        "lk_mlp_destroy": [_vp],
        "lk_mps_open": [C.c_char_p, C.POINTER(_vp), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                        C.POINTER(C.c_int64)],
        "lk_mps_get": [_vp, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _i32p],
        "lk_mps_close": [_vp],
        "lk_verify_optimal": [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _vp, _vp,
                              C.c_int32, _f64p, _f64p, C.c_double, C.POINTER(C.c_double),
                              C.POINTER(C.c_int32)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def _ck(rc: int):
    if rc != 0:
        msg = lib().lk_last_error().decode(errors="replace")
        raise ERRORS.get(rc, RuntimeError)(msg)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data


def device_count() -> int:
    c = C.c_int(0)
    lib().lk_device_count(C.byref(c))
    return c.value


# ------------------------------------------------------------------ arena
class Arena:
    def __init__(self, capacity: int):
        self.h = _vp()
        _ck(lib().lk_arena_create(capacity, C.byref(self.h)))

    def allocate(self, nbytes: int) -> int:
        p = _vp()
        _ck(lib().lk_arena_alloc(self.h, nbytes, C.byref(p)))
        return p.value or 0

    def free(self, ptr):
        _ck(lib().lk_arena_free(self.h, ptr))

    def occupancy_percentage(self) -> int:
        v = C.c_size_t(0)
        _ck(lib().lk_arena_occupancy(self.h, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            _ck(lib().lk_arena_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ symbolic
def order_reference(rp, ci, n):
    m = len(rp) - 1
    P = np.zeros(m, np.int32)
    _ck(lib().lk_order_reference(m, n, _i32(rp), _i32(ci), P))
    return P


def order_amd(rp, ci, n):
    m = len(rp) - 1
    P = np.zeros(m, np.int32)
    _ck(lib().lk_order_amd(m, n, _i32(rp), _i32(ci), P))
    return P


def symbolic(rp, ci, n, P):
    m = len(rp) - 1
    nnz = C.c_int64(0)
    parent = np.zeros(m, np.int32)
    pinv = np.zeros(m, np.int32)
    Lptr = np.zeros(m + 1, np.int32)
    _ck(lib().lk_symbolic(m, n, _i32(rp), _i32(ci), _i32(P), None, None, None, None, 0,
                          C.byref(nnz)))
    Lidx = np.zeros(max(nnz.value, 1), np.int32)
    _ck(lib().lk_symbolic(m, n, _i32(rp), _i32(ci), _i32(P), parent.ctypes.data,
                          pinv.ctypes.data, Lptr.ctypes.data, Lidx.ctypes.data, nnz.value,
                          C.byref(nnz)))
    return {"P_inv": pinv, "parent": parent, "Lptr": Lptr, "Lidx": Lidx[:nnz.value]}


def supernodal(rp, ci, n, P=None, relax_small=-1, relax_frac=-1.0):
    m = len(rp) - 1
    st = KKTStats()
    perm = np.zeros(m, np.int32)
    first = np.zeros(m + 1, np.int32)
    ns = C.c_int32(0)
    Pp = _i32(P) if P is not None else None
    _ck(lib().lk_supernodal(m, n, _i32(rp), _i32(ci), _ptr(Pp), relax_small, relax_frac,
                            C.byref(st), perm.ctypes.data, first.ctypes.data, C.byref(ns)))
    out = {f: getattr(st, f) for f, _ in KKTStats._fields_ if f != "pad"}
    out["perm"] = perm
    out["sn_first"] = first[:ns.value + 1]
    return out


# ------------------------------------------------------------------ KKT
class KKT:
    """gpu::GPUKKTCholeskySolver through the C ABI (A in CSR)."""

    def __init__(self, arena: Arena, rp, ci, val, n, P=None, reference_symbolic=True,
                 use_graphs=False):
        self.m = len(rp) - 1
        self.n = n
        self.arena = arena
        self.h = _vp()
        Pp = _i32(P) if P is not None else None
        _ck(lib().lk_kkt_create(arena.h, self.m, n, _i32(rp), _i32(ci), _f64(val), _ptr(Pp),
                                int(reference_symbolic), int(use_graphs), C.byref(self.h)))

    def factorize(self, theta, mode=0, device_ptr=None):
        """mode 0: reference factorize (M0 + 1e-9 I); 1: equilibrated FP64; 2: FP32.
        Returns True when positive definite."""
        ns = C.c_int32(0)
        if device_ptr is not None:
            _ck(lib().lk_kkt_factor(self.h, device_ptr, 1, mode, C.byref(ns)))
        else:
            t = _f64(theta)
            _ck(lib().lk_kkt_factor(self.h, t.ctypes.data, 0, mode, C.byref(ns)))
        return ns.value == 0

    def solve(self, rhs, tol=1e-10, max_refine=-1, device_ptr=None):
        info = KKTSolveInfo()
        if device_ptr is not None:
            _ck(lib().lk_kkt_solve(self.h, device_ptr, 1, tol, max_refine, C.byref(info)))
            return None, info
        r = _f64(rhs).copy()
        _ck(lib().lk_kkt_solve(self.h, r.ctypes.data, 0, tol, max_refine, C.byref(info)))
        return r, info

    def L_csr(self, nnz):
        v = np.zeros(nnz)
        _ck(lib().lk_kkt_get_L_csr(self.h, v, nnz))
        return v

    def permutation(self):
        p = np.zeros(self.m, np.int32)
        _ck(lib().lk_kkt_permutation(self.h, p))
        return p

    def stats(self):
        st = KKTStats()
        _ck(lib().lk_kkt_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in KKTStats._fields_ if f != "pad"}

    def stream(self):
        s = _vp()
        _ck(lib().lk_kkt_stream(self.h, C.byref(s)))
        return s.value

    def close(self):
        if self.h:
            _ck(lib().lk_kkt_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ IPM
def default_options(**kw) -> IpmOptions:
    o = IpmOptions()
    _ck(lib().lk_ipm_default_options(C.byref(o)))
    aliases = {"relative": ("stop_rule", 1)}
    for k, v in kw.items():
        if k in aliases:
            if v:
                setattr(o, aliases[k][0], aliases[k][1])
            continue
        setattr(o, k, v)
    return o


def _result(res: IpmResult, x, y, s):
    out = {f: getattr(res, f) for f, _ in IpmResult._fields_ if f != "pad"}
    out["status_name"] = STATUS.get(res.status, str(res.status))
    out.update(x=x, y=y, s=s)
    return out


def ipm_solve(lp, warm=None, **opts):
    """Solve one gen.LP (CSC) with MehrotraSolver; opts map onto IpmOptions fields."""
    o = default_options(**opts)
    res = IpmResult()
    x, y, s = np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n)
    w = [None, None, None] if warm is None else [_f64(a) for a in warm]
    _ck(lib().lk_ipm_solve(lp.m, lp.n, _i32(lp.indptr), _i32(lp.indices), _f64(lp.data),
                           _f64(lp.c), _f64(lp.b), None, None, C.byref(o), _ptr(w[0]),
                           _ptr(w[1]), _ptr(w[2]), C.byref(res), x.ctypes.data, y.ctypes.data,
                           s.ctypes.data))
    return _result(res, x, y, s)


class IPM:
    """Reusable MehrotraSolver for a fixed A (setup once, many (b, c))."""

    def __init__(self, lp, **opts):
        self.m, self.n = lp.m, lp.n
        o = default_options(**opts)
        self.h = _vp()
        _ck(lib().lk_ipm_create(lp.m, lp.n, _i32(lp.indptr), _i32(lp.indices), _f64(lp.data),
                                _f64(lp.c), _f64(lp.b), C.byref(o), C.byref(self.h)))

    def run(self, c=None, b=None, warm=None):
        res = IpmResult()
        x, y, s = np.zeros(self.n), np.zeros(self.m), np.zeros(self.n)
        cc = _f64(c) if c is not None else None
        bb = _f64(b) if b is not None else None
        w = [None, None, None] if warm is None else [_f64(a) for a in warm]
        _ck(lib().lk_ipm_run(self.h, _ptr(cc), _ptr(bb), _ptr(w[0]), _ptr(w[1]), _ptr(w[2]),
                             C.byref(res), x.ctypes.data, y.ctypes.data, s.ctypes.data))
        return _result(res, x, y, s)

    def profile_factor(self, fp32=False):
        L = lib()
        L.lk_ipm_profile_factor.argtypes = [_vp, C.c_int32, _f64p, _f64p, _f64p]
        ms, fl, la = np.zeros(6), np.zeros(6), np.zeros(6)
        _ck(L.lk_ipm_profile_factor(self.h, int(fp32), ms, fl, la))
        names = ["step", "diag", "update", "assembly", "factorization", "part"]
        return {k: {"ms": float(ms[i]), "flops": float(fl[i]), "launches": int(la[i])}
                for i, k in enumerate(names)}

    def set_launch_timing(self, on=True):
        """HIP-event timing of every factorization launch of the following runs (resets)."""
        L = lib()
        L.lk_ipm_launch_timing.argtypes = [_vp, C.c_int32]
        _ck(L.lk_ipm_launch_timing(self.h, int(on)))

    def launch_timing(self):
        L = lib()
        L.lk_ipm_read_launch_timing.argtypes = [_vp, _f64p, _f64p, _f64p]
        ms, fl, la = np.zeros(6), np.zeros(6), np.zeros(6)
        _ck(L.lk_ipm_read_launch_timing(self.h, ms, fl, la))
        names = ["step", "diag", "update", "assembly", "factorization", "part"]
        return {k: {"ms": float(ms[i]), "flops": float(fl[i]), "launches": int(la[i])}
                for i, k in enumerate(names)}

    def kkt_stats(self):
        L = lib()
        L.lk_ipm_kkt_stats.argtypes = [_vp, _f64p]
        v = np.zeros(12)
        _ck(L.lk_ipm_kkt_stats(self.h, v))
        keys = ["m", "n", "nnz_A", "nsuper", "nnz_L", "nnz_L_exact", "flops_exact",
                "flops_panel", "levels", "launches_factor", "launches_solve", "symbolic_ms"]
        return dict(zip(keys, v.tolist()))

    def close(self):
        if self.h:
            _ck(lib().lk_ipm_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ batched IPM
def batch_options(**kw) -> BatchOptions:
    o = BatchOptions()
    _ck(lib().lk_batch_default_options(C.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _summary(s: BatchSummary) -> dict:
    return {f: getattr(s, f) for f, _ in BatchSummary._fields_ if f != "pad"}


class Batch:
    """ipm::BatchedMehrotraSolver over a list of CSC LPs (lakshya.gen.LP)."""

    def __init__(self, lps, **opts):
        m = np.array([lp.m for lp in lps], np.int32)
        n = np.array([lp.n for lp in lps], np.int32)
        colptr = np.concatenate([np.asarray(lp.indptr, np.int32) for lp in lps])
        rowidx = np.concatenate([np.asarray(lp.indices, np.int32) for lp in lps])
        vals = np.concatenate([np.asarray(lp.data, np.float64) for lp in lps])
        b = np.concatenate([np.asarray(lp.b, np.float64) for lp in lps])
        c = np.concatenate([np.asarray(lp.c, np.float64) for lp in lps])
        self._create(m, n, colptr, rowidx, vals, b, c, opts)

    @classmethod
    def from_arrays(cls, m, n, colptr, rowidx, vals, b, c, **opts):
        """The batch from its concatenated arrays (the lk_batch_create layout)."""
        self = cls.__new__(cls)
        self._create(_i32(m), _i32(n), _i32(colptr), _i32(rowidx), _f64(vals), _f64(b), _f64(c), opts)
        return self

    def _create(self, m, n, colptr, rowidx, vals, b, c, opts):
        self.h = _vp()
        self.count = len(m)
        self.m, self.n = m, n
        o = batch_options(**opts)
        _ck(lib().lk_batch_create(self.count, self.m, self.n, colptr, rowidx, vals, b, c,
                                  C.byref(o), C.byref(self.h)))

    def solve(self) -> dict:
        s = BatchSummary()
        _ck(lib().lk_batch_solve(self.h, C.byref(s)))
        return _summary(s)

    def autotune(self, sample=512) -> dict:
        s = BatchSummary()
        _ck(lib().lk_batch_autotune(self.h, sample, C.byref(s)))
        return _summary(s)

    def results(self) -> list:
        arr = (BatchResult * self.count)()
        _ck(lib().lk_batch_results(self.h, C.cast(arr, _vp)))
        out = []
        for r in arr:
            d = {f: getattr(r, f) for f, _ in BatchResult._fields_ if f != "pad"}
            d["status_name"] = STATUS.get(r.status, str(r.status))
            out.append(d)
        return out

    def set_warm_start(self, x=None, y=None, s=None, flags=None):
        """Original-space start points (concatenated like the inputs); None clears them."""
        if x is None:
            _ck(lib().lk_batch_set_warm_start(self.h, None, None, None, None))
            return
        x, y, s = _f64(x), _f64(y), _f64(s)
        f = None if flags is None else _i32(flags)
        _ck(lib().lk_batch_set_warm_start(self.h, x.ctypes.data, y.ctypes.data, s.ctypes.data,
                                          None if f is None else f.ctypes.data))

    def solutions(self):
        """All solutions, concatenated (x, y, s)."""
        xs, ys, ss = zip(*[self.solution(i) for i in range(self.count)])
        return np.concatenate(xs), np.concatenate(ys), np.concatenate(ss)

    def solution(self, i):
        x, y, s = np.zeros(self.n[i]), np.zeros(self.m[i]), np.zeros(self.n[i])
        _ck(lib().lk_batch_solution(self.h, i, x, y, s))
        return x, y, s

    def crossover(self):
        """lk_batch_crossover: basic solutions from the last solve, one list entry per LP:
        {"basis": int32[m] (-1 artificial row), "x": float64[n], "status": int, "pivots": int}."""
        mo = np.concatenate([[0], np.cumsum(self.m)]).astype(np.int64)
        no = np.concatenate([[0], np.cumsum(self.n)]).astype(np.int64)
        basis = np.zeros(int(mo[-1]), np.int32)
        x = np.zeros(int(no[-1]))
        st = np.zeros(self.count, np.int32)
        pv = np.zeros(self.count, np.int32)
        _ck(lib().lk_batch_crossover(self.h, basis, x, st, pv))
        return [{"basis": basis[mo[i]:mo[i + 1]], "x": x[no[i]:no[i + 1]], "status": int(st[i]),
                 "pivots": int(pv[i])} for i in range(self.count)]

    def close(self):
        if self.h:
            _ck(lib().lk_batch_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SIMPLEX_STATUS = {0: "Optimal", 1: "Infeasible", 2: "Unbounded", 3: "IterationLimit"}


def simplex(lps, max_iter=10000):
    """lk_simplex: the device Phase I / Phase II simplex route on each LP, one dict per LP:
    {"status", "status_name", "objective", "x" (n), "y" (m), "basis" (m, -1 artificial),
    "iterations" (Phase I, Phase II)}."""
    m = np.array([lp.m for lp in lps], np.int32)
    n = np.array([lp.n for lp in lps], np.int32)
    cp = np.concatenate([np.asarray(lp.indptr, np.int32) for lp in lps])
    ri = np.concatenate([np.asarray(lp.indices, np.int32) for lp in lps])
    va = np.concatenate([np.asarray(lp.data, np.float64) for lp in lps])
    b = np.concatenate([np.asarray(lp.b, np.float64) for lp in lps])
    c = np.concatenate([np.asarray(lp.c, np.float64) for lp in lps])
    mo = np.concatenate([[0], np.cumsum(m)]).astype(np.int64)
    no = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    basis = np.zeros(int(mo[-1]), np.int32)
    x = np.zeros(int(no[-1]))
    y = np.zeros(int(mo[-1]))
    st = np.zeros(len(lps), np.int32)
    it = np.zeros(2 * len(lps), np.int32)
    obj = np.zeros(len(lps))
    _ck(lib().lk_simplex(len(lps), m, n, cp, ri, va, b, c, int(max_iter), basis, x, y, st, it, obj))
    return [{"status": int(st[i]), "status_name": SIMPLEX_STATUS[int(st[i])], "objective": float(obj[i]),
             "x": x[no[i]:no[i + 1]], "y": y[mo[i]:mo[i + 1]], "basis": basis[mo[i]:mo[i + 1]],
             "iterations": (int(it[2 * i]), int(it[2 * i + 1]))} for i in range(len(lps))]


def crossover_points(lps, xs):
    """lk_crossover: basic solutions of the LPs from the points xs (one per LP), in the form of
    Batch.crossover()."""
    m = np.array([lp.m for lp in lps], np.int32)
    n = np.array([lp.n for lp in lps], np.int32)
    colptr = np.concatenate([np.asarray(lp.indptr, np.int32) for lp in lps])
    rowidx = np.concatenate([np.asarray(lp.indices, np.int32) for lp in lps])
    vals = np.concatenate([np.asarray(lp.data, np.float64) for lp in lps])
    b = np.concatenate([np.asarray(lp.b, np.float64) for lp in lps])
    c = np.concatenate([np.asarray(lp.c, np.float64) for lp in lps])
    x = np.concatenate([np.asarray(v, np.float64) for v in xs])
    if x.size != int(n.sum()):
        raise ValueError("crossover_points: one point of n entries per LP")
    mo = np.concatenate([[0], np.cumsum(m)]).astype(np.int64)
    no = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    basis = np.zeros(int(mo[-1]), np.int32)
    xv = np.zeros(int(no[-1]))
    st = np.zeros(len(lps), np.int32)
    pv = np.zeros(len(lps), np.int32)
    _ck(lib().lk_crossover(len(lps), m, n, colptr, rowidx, vals, b, c, x, basis, xv, st, pv))
    return [{"basis": basis[mo[i]:mo[i + 1]], "x": xv[no[i]:no[i + 1]], "status": int(st[i]),
             "pivots": int(pv[i])} for i in range(len(lps))]


def batch_solve_noalloc(lps, **opts):
    """lk_batch_solve_noalloc: the batch solved in the library's fixed device scratch (no device
    allocation).  Returns (rc, x, y, s, results, summary); rc != 0 leaves the message in
    lk_last_error()."""
    m = np.array([lp.m for lp in lps], np.int32)
    n = np.array([lp.n for lp in lps], np.int32)
    colptr = np.concatenate([np.asarray(lp.indptr, np.int32) for lp in lps])
    rowidx = np.concatenate([np.asarray(lp.indices, np.int32) for lp in lps])
    vals = np.concatenate([np.asarray(lp.data, np.float64) for lp in lps])
    b = np.concatenate([np.asarray(lp.b, np.float64) for lp in lps])
    c = np.concatenate([np.asarray(lp.c, np.float64) for lp in lps])
    x, s, y = np.zeros(int(n.sum())), np.zeros(int(n.sum())), np.zeros(int(m.sum()))
    res = (BatchResult * len(lps))()
    summ = BatchSummary()
    o = batch_options(**opts)
    rc = lib().lk_batch_solve_noalloc(len(lps), m, n, colptr, rowidx, vals, b, c, C.byref(o),
                                      x, y, s, C.cast(res, _vp), C.byref(summ))
    out = []
    for r in res:
        d = {f: getattr(r, f) for f, _ in BatchResult._fields_ if f != "pad"}
        d["status_name"] = STATUS.get(r.status, str(r.status))
        out.append(d)
    return rc, x, y, s, out, _summary(summ)


def last_error() -> str:
    return lib().lk_last_error().decode(errors="replace")


# ------------------------------------------------------------------ MPS input
def read_mps(path):
    """parsers::read_mps: the file as an equality-form LP (lakshya.gen.LP, CSC); meta holds
    lb, ub and vtype (0 Continuous, 1 Integer, 2 Binary).  Parse errors raise RuntimeError."""
    from .gen import LP
    h = _vp()
    m, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
    _ck(lib().lk_mps_open(str(path).encode(), C.byref(h), C.byref(m), C.byref(n), C.byref(nnz)))
    try:
        m, n, nnz = m.value, n.value, nnz.value
        colptr, rowidx = np.zeros(n + 1, np.int32), np.zeros(max(nnz, 1), np.int32)
        vals, obj, rhs = np.zeros(max(nnz, 1)), np.zeros(max(n, 1)), np.zeros(max(m, 1))
        lb, ub, vt = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1), np.int32)
        _ck(lib().lk_mps_get(h, colptr, rowidx, vals, obj, rhs, lb, ub, vt))
    finally:
        lib().lk_mps_close(h)
    return LP(m, n, colptr, rowidx[:nnz], vals[:nnz], rhs[:m], obj[:n], str(path),
              {"lb": lb[:n], "ub": ub[:n], "vtype": vt[:n]})


# ------------------------------------------------------------------ warm-start network
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    """learn::WarmStartMLP: d_in -> h1 -> h2 -> d_out, ReLU, BF16 MFMA / FP32 accumulation."""

    def __init__(self, d_in, h1, h2, d_out, max_batch=256, seed=0):
        self.dims = (d_in, h1, h2, d_out)
        self.max_batch = max_batch
        self.h = _vp()
        _ck(lib().lk_mlp_create(d_in, h1, h2, d_out, max_batch, seed, C.byref(self.h)))
        n = C.c_int64()
        _ck(lib().lk_mlp_param_count(self.h, C.byref(n)))
        self.nparams = n.value

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        _ck(lib().lk_mlp_get_params(self.h, p))
        return p

    def set_params(self, p):
        _ck(lib().lk_mlp_set_params(self.h, _f32(p)))

    def forward(self, u):
        u = _f32(u)
        out = np.zeros((u.shape[0], self.dims[3]), np.float32)
        for b0 in range(0, u.shape[0], self.max_batch):
            ub = np.ascontiguousarray(u[b0:b0 + self.max_batch])
            yb = np.zeros((ub.shape[0], self.dims[3]), np.float32)
            _ck(lib().lk_mlp_forward(self.h, ub, ub.shape[0], yb))
            out[b0:b0 + ub.shape[0]] = yb
        return out

    def gradients(self, u, t):
        u, t = _f32(u), _f32(t)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_double()
        _ck(lib().lk_mlp_gradients(self.h, u, t, u.shape[0], g, C.byref(loss)))
        return g, loss.value

    def set_dataset(self, u, t):
        u, t = _f32(u), _f32(t)
        _ck(lib().lk_mlp_set_dataset(self.h, u, t, u.shape[0]))

    def train(self, epochs, batch, lr, seed=0):
        out = np.zeros(epochs)
        _ck(lib().lk_mlp_train(self.h, epochs, batch, lr, seed, out))
        return out

    def close(self):
        if self.h:
            _ck(lib().lk_mlp_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ruiz_scaling(rowptr, colidx, vals, b, c):
    """The solver's 10-pass Jacobi Ruiz equilibration (host code of the library).
    Returns (scaled values, scaled b, scaled c, R, C)."""
    rp, ci = _i32(rowptr), _i32(colidx)
    m, n = len(rp) - 1, len(c)
    v = np.array(vals, dtype=np.float64)
    bb, cc = np.array(b, dtype=np.float64), np.array(c, dtype=np.float64)
    R, Cs = np.empty(m), np.empty(n)
    _ck(lib().lk_ruiz_scaling(m, n, rp, ci, v, bb, cc, R, Cs))
    return v, bb, cc, R, Cs


def verify_optimal(lp, x, pi, tol=1e-6, lb=None, ub=None, maximize=False):
    res = C.c_double(0.0)
    ok = C.c_int32(0)
    lbp = _f64(lb) if lb is not None else None
    ubp = _f64(ub) if ub is not None else None
    _ck(lib().lk_verify_optimal(lp.m, lp.n, _i32(lp.indptr), _i32(lp.indices), _f64(lp.data),
                                _f64(lp.c), _f64(lp.b), _ptr(lbp), _ptr(ubp), int(maximize),
                                _f64(x), _f64(pi), tol, C.byref(res), C.byref(ok)))
    return bool(ok.value), res.value


# ------------------------------------------------------------------ plan inspection (host only)
TASK_DTYPE = np.dtype([("off", "<i8"), ("aux", "<i8"), ("aux2", "<i8"), ("ws", "<i8"),
                       ("ld", "<i4"), ("ncol", "<i4"), ("k0", "<i4"), ("kb", "<i4"),
                       ("r0", "<i4"), ("c0", "<i4"), ("kbeg", "<i4"), ("kend", "<i4"),
                       ("nsplit", "<i4"), ("sn", "<i4"), ("first", "<i4"), ("np", "<i4"),
                       ("px", "<i8"), ("prio", "<i4"), ("part", "<i4")])
PLAN_DTYPES = {"perm": np.int32, "sn_rowptr": np.int64, "sn_rows": np.int32,
               "sn_panel": np.int64, "col_to_sn": np.int32, "sn_first": np.int32,
               "flaunch": np.int64, "slaunch": np.int64, "diag_elems": np.int64,
               "colptr": np.int32, "prow": np.int32, "cval": np.float64, "prp": np.int32,
               "pci": np.int32, "pval": np.float64, "empty_row": np.uint8, "segptr": np.int64,
               "segdst": np.int64, "segi": np.int32, "segj": np.int32, "pp": np.float64,
               "pk": np.int32, "dseg": np.int64, "ftasks": TASK_DTYPE, "stasks": TASK_DTYPE,
               "csr_index": np.int64, "stage_elems": np.int64, "long_prows": np.int32,
               "long_cols": np.int32, "long_segs": np.int32, "long_diag": np.int32,
               "mid_segs": np.int32, "mid_diag": np.int32,
               "chunk_seg": np.int64, "fdeps": np.int32, "fsdep": np.int32,
               "fqueue": np.int32, "fqoff": np.int32, "n_xcd": np.int32,
               "squeue": np.int32, "sqoff": np.int32}


def plan_export(rp, ci, val, n, P=None, relax_small=-1, relax_frac=-1.0, step_dag=1):
    """Host-only device plan of the KKT solver (dict of numpy arrays).  step_dag: 0 one STEP
    launch per block column, 1 the solver's choice per level, 2 one STEPS launch per level."""
    L = lib()
    L.lk_plan_create.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _vp, C.c_int32,
                                 C.c_double, C.c_int32, C.POINTER(_vp)]
    L.lk_plan_array.argtypes = [_vp, C.c_char_p, _vp, C.c_size_t, C.POINTER(C.c_size_t)]
    L.lk_plan_destroy.argtypes = [_vp]
    m = len(rp) - 1
    h = _vp()
    Pp = _i32(P) if P is not None else None
    _ck(L.lk_plan_create(m, n, _i32(rp), _i32(ci), _f64(val), _ptr(Pp), relax_small,
                         relax_frac, int(step_dag), C.byref(h)))
    out = {}
    try:
        for name, dt in PLAN_DTYPES.items():
            nb = C.c_size_t(0)
            if L.lk_plan_array(h, name.encode(), None, 0, C.byref(nb)) != 0:
                continue
            arr = np.zeros(nb.value // np.dtype(dt).itemsize, dtype=dt)
            _ck(L.lk_plan_array(h, name.encode(), arr.ctypes.data, nb.value, C.byref(nb)))
            out[name] = arr
    finally:
        L.lk_plan_destroy(h)
    return out


# ------------------------------------------------------------------ raw device buffers (tests)
class DeviceBuffer:
    """Device memory from the library's own allocator (lkd_malloc), with host copies."""

    def __init__(self, arr_or_nbytes, dtype=None):
        L = lib()
        for name, args in (("lkd_malloc", [C.POINTER(_vp), C.c_size_t]), ("lkd_free", [_vp]),
                           ("lkd_memcpy_htod", [_vp, _vp, C.c_size_t, _vp]),
                           ("lkd_memcpy_dtoh", [_vp, _vp, C.c_size_t, _vp]),
                           ("lkd_stream_sync", [_vp])):
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_int
        if isinstance(arr_or_nbytes, np.ndarray):
            a = np.ascontiguousarray(arr_or_nbytes)
            self.nbytes, self.dtype, self.shape = a.nbytes, a.dtype, a.shape
        else:
            a = None
            self.nbytes, self.dtype, self.shape = int(arr_or_nbytes), np.dtype(dtype or np.uint8), None
        self.ptr = _vp()
        rc = L.lkd_malloc(C.byref(self.ptr), max(self.nbytes, 1))
        if rc:
            raise RuntimeError(f"lkd_malloc failed ({rc})")
        if a is not None:
            self.upload(a)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if lib().lkd_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes, None) or lib().lkd_stream_sync(None):
            raise RuntimeError("upload failed")

    def download(self, dtype=None, count=None):
        dt = np.dtype(dtype or self.dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.zeros(n, dt)
        if lib().lkd_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes, None):
            raise RuntimeError("download failed")
        return out

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            lib().lkd_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

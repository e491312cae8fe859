"""ctypes front end of the C oracle (oracle/oracle.c).  Test infrastructure only: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class IpmOptions(C.Structure):
    _fields_ = [("stop_rule", C.c_int), ("tol", C.c_double), ("gap_tol", C.c_double),
                ("max_iter", C.c_int)]


class IpmResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("objective", C.c_double),
                ("primal_residual", C.c_double), ("dual_residual", C.c_double),
                ("duality_gap", C.c_double), ("rel_gap", C.c_double),
                ("rel_primal_inf", C.c_double), ("rel_dual_inf", C.c_double),
                ("dual_objective", C.c_double), ("factorizations", C.c_int),
                ("solves", C.c_int), ("fail_code", C.c_int)]


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_ruiz.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.or_md_order.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p]
        L.or_symbolic.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p,
                                  C.c_int64]
        L.or_assemble_dense.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _i32p, _f64p]
        L.or_segment_sums.argtypes = [C.c_int64, _i64p, _f64p, _i32p, _f64p, _f64p]
        L.or_fma_vec.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p]
        L.or_kkt_factor_solve.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _i32p,
                                          C.c_void_p, C.c_void_p, C.c_double,
                                          C.POINTER(C.c_int)]
        L.or_ipm_solve.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                   C.c_void_p, C.POINTER(IpmOptions), C.POINTER(IpmResult),
                                   _f64p, _f64p, _f64p]
        L.or_verify_optimal.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                        _f64p, _f64p, C.c_int, _f64p, _f64p, C.c_double,
                                        C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def segment_sums(segptr, pp, pk, theta):
    """Pair-map segment sums in the device assembly's fixed order (exact fma)."""
    nseg = len(segptr) - 1
    out = np.zeros(nseg)
    lib().or_segment_sums(nseg, np.ascontiguousarray(segptr, np.int64), _f64(pp),
                          np.ascontiguousarray(pk, np.int32), _f64(theta), out)
    return out


def fma(a, b, c):
    """Elementwise a * b + c with one rounding (C fma)."""
    a, b, c = np.broadcast_arrays(np.asarray(a, np.float64), np.asarray(b, np.float64), np.asarray(c, np.float64))
    out = np.zeros(a.shape)
    lib().or_fma_vec(a.size, _f64(a.ravel()), _f64(b.ravel()), _f64(c.ravel()), out.reshape(-1))
    return out


def ruiz(rp, ci, val, b, c):
    m, n = len(rp) - 1, len(c)
    val, b, c = _f64(val).copy(), _f64(b).copy(), _f64(c).copy()
    R, Cs = np.zeros(m), np.zeros(n)
    lib().or_ruiz(m, n, _i32(rp), _i32(ci), val, b, c, R, Cs)
    return val, b, c, R, Cs


def md_order(rp, ci, n):
    m = len(rp) - 1
    P = np.zeros(m, np.int32)
    lib().or_md_order(m, n, _i32(rp), _i32(ci), P)
    return P


def symbolic(rp, ci, n, P):
    m = len(rp) - 1
    parent = np.zeros(m, np.int32)
    Lptr = np.zeros(m + 1, np.int32)
    cap = m * (m - 1) // 2 + 1
    Lidx = np.zeros(cap, np.int32)
    rc = lib().or_symbolic(m, n, _i32(rp), _i32(ci), _i32(P), parent, Lptr, Lidx, cap)
    assert rc == 0
    return parent, Lptr, Lidx[:Lptr[-1]].copy()


def assemble_dense(rp, ci, val, theta, P):
    m = len(rp) - 1
    M = np.zeros((m, m))
    lib().or_assemble_dense(m, _i32(rp), _i32(ci), _f64(val), _f64(theta), _i32(P), M)
    return M


def kkt_factor_solve(rp, ci, val, theta, P, rhs=None, tol=1e-10, want_L=False):
    """Returns (status, L_unequilibrated or None, dy or None, success)."""
    m = len(rp) - 1
    L = np.zeros((m, m)) if want_L else None
    r = _f64(rhs).copy() if rhs is not None else None
    ok = C.c_int(0)
    st = lib().or_kkt_factor_solve(m, _i32(rp), _i32(ci), _f64(val), _f64(theta), _i32(P),
                                   L.ctypes.data if want_L else None,
                                   r.ctypes.data if r is not None else None, tol, C.byref(ok))
    return st, (np.tril(L) if want_L else None), r, bool(ok.value)


def ipm_solve(lp_csr, b, c, n, P=None, stop_rule=0, tol=1e-6, gap_tol=1e-8, max_iter=200):
    rp, ci, val = (_i32(lp_csr[0]), _i32(lp_csr[1]), _f64(lp_csr[2]))
    m = len(rp) - 1
    opt = IpmOptions(stop_rule, tol, gap_tol, max_iter)
    res = IpmResult()
    x, y, s = np.zeros(n), np.zeros(m), np.zeros(n)
    Pp = _i32(P) if P is not None else None
    rc = lib().or_ipm_solve(m, n, rp, ci, val, _f64(b), _f64(c),
                            Pp.ctypes.data if Pp is not None else None,
                            C.byref(opt), C.byref(res), x, y, s)
    out = {f: getattr(res, f) for f, _ in IpmResult._fields_}
    out.update(rc=rc, x=x, y=y, s=s)
    return out


def verify_optimal(m, n, colptr, rowidx, val, obj, rhs, lb, ub, x, pi, tol=1e-6, maximize=False):
    res = C.c_double(0.0)
    ok = lib().or_verify_optimal(m, n, _i32(colptr), _i32(rowidx), _f64(val), _f64(obj),
                                 _f64(rhs), _f64(lb), _f64(ub), int(maximize), _f64(x), _f64(pi),
                                 tol, C.byref(res))
    return bool(ok), res.value

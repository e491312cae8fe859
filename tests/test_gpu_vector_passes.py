"""The IPM's fused sparse and vector passes in isolation, against closed forms in numpy (the
elementwise kernels, the reductions and the SpMVs of src/ipm/mehrotra_device.cu:93-176,191-259,
494-660, which the device fuses into row and column passes with block partials and a
fixed-order final reduction):
  lkd_ipm_residuals  rp = A x - b, rd = A^T y + s - c, theta = x / s and the slots
                     ||rp||^2, max |rp_i / R_i|, b^T y, ||rd||^2, x^T s, c^T x, max |rd_j / C_j|,
                     max |x_j (s_j - rd_j)|, max and min theta;
  lkd_ipm_direction  ds = -A^T dy - rd, dx = -theta ds + r_xs / s, qd = A^T dy + ds, qp = A dx
                     and the slots of verify_newton_direction (e1, e2, e3 sums of squares),
                     ||r_xs||^2, the ratio-test bounds min -x/dx, min -s/ds, ||dx||^2, ||ds||^2,
                     ||dy||^2.
The LP has rows and columns longer than LKD_LONG_ROW entries (summed in pieces by a pre-pass)
and several lane-group widths."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LONG_ROW, LONG_PIECE, MAX_SPLIT, MAX_NBLK = 512, 1024, 64, 4096
SLOTS = dict(RP2=0, RD2=1, XS=2, CX=3, BY=4, RPINF=5, RDINF=6, R1=7, R2=8, R3=9, RXS2=10, APMAX=11,
             ADMAX=12, DX2=13, DS2=14, MUAFF=15, SIGMAMU=16, AP=17, AD=18, DY2=19, KKTR=20,
             XSMAX=21, THMAX=22, THMIN=23)
LKS_TOTAL = 32 + 8 * 32


class Long(C.Structure):
    _fields_ = [("n", C.c_int32), ("idx", C.c_void_p), ("split", C.c_int32), ("part", C.c_void_p)]


class Ipm(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("arp", C.c_void_p), ("aci", C.c_void_p),
                ("aval", C.c_void_p), ("trp", C.c_void_p), ("tci", C.c_void_p), ("tval", C.c_void_p),
                ("b", C.c_void_p), ("c", C.c_void_p), ("R", C.c_void_p), ("C", C.c_void_p),
                ("grp_rows", C.c_int32), ("grp_cols", C.c_int32), ("long_rows", Long),
                ("long_cols", Long), ("part", C.c_void_p), ("nblk_rows", C.c_int32),
                ("nblk_cols", C.c_int32), ("S", C.c_void_p)]


def _group(avg):
    g = 1
    while g < 64 and g < avg:
        g <<= 1
    return g


class Passes:
    """lkd_ipm set up as the solver does (mehrotra.cpp): long-row lists with their piece
    scratch, lane groups from the mean short-row length / 4, blocks per pass, partials, slots."""

    def __init__(self, native, A, b, c, R, Cs):
        import scipy.sparse as sp
        self.native, self.keep = native, []
        Ar = sp.csr_matrix(A)
        Ar.sort_indices()
        At = sp.csr_matrix(Ar.T)
        At.sort_indices()
        m, n = Ar.shape
        P = Ipm()
        P.m, P.n = m, n
        P.arp, P.aci, P.aval = self.dev(Ar.indptr, np.int32), self.dev(Ar.indices, np.int32), self.dev(Ar.data, np.float64)
        P.trp, P.tci, P.tval = self.dev(At.indptr, np.int32), self.dev(At.indices, np.int32), self.dev(At.data, np.float64)
        P.b, P.c, P.R, P.C = self.dev(b, np.float64), self.dev(c, np.float64), self.dev(R, np.float64), self.dev(Cs, np.float64)
        mr = self.long_list(Ar.indptr, P, "long_rows")
        mc = self.long_list(At.indptr, P, "long_cols")
        P.grp_rows, P.grp_cols = _group(mr / 4.0), _group(mc / 4.0)
        P.nblk_rows = max(1, min(MAX_NBLK, (m * P.grp_rows + 255) // 256))
        P.nblk_cols = max(1, min(MAX_NBLK, (n * P.grp_cols + 255) // 256))
        P.part = self.buf(256 * MAX_NBLK)
        self.S = native.DeviceBuffer(8 * LKS_TOTAL, np.float64)
        P.S = self.S.ptr.value
        self.P, self.m, self.n = P, m, n

    def dev(self, a, dt):
        d = self.native.DeviceBuffer(np.ascontiguousarray(a, dt))
        self.keep.append(d)
        return d.ptr.value

    def buf(self, count):
        d = self.native.DeviceBuffer(8 * max(count, 1), np.float64)
        self.keep.append(d)
        return d.ptr.value

    def long_list(self, ptr, P, field):
        lens = np.diff(ptr)
        idx = np.flatnonzero(lens > LONG_ROW).astype(np.int32)
        lg = getattr(P, field)
        lg.n = len(idx)
        mean = lens[lens <= LONG_ROW].sum() / max(1, len(lens) - len(idx))
        lg.idx = self.dev(idx, np.int32) if len(idx) else None
        longest = int(lens[idx].max()) if len(idx) else 0
        lg.split = min(MAX_SPLIT, (longest + LONG_PIECE - 1) // LONG_PIECE)
        lg.part = self.buf(len(idx) * lg.split) if len(idx) and lg.split > 1 else None
        return mean

    def slots(self):
        return self.S.download()


def _lp(seed):
    """800 x 3,000 sparse A (4 entries per column) with 6 rows of about 1,500 entries and 4
    columns of about 700 (above LKD_LONG_ROW = 512: listed and summed in pieces)."""
    rng = np.random.default_rng(seed)
    m, n = 800, 3000
    A = np.zeros((m, n))
    for j in range(n):
        A[rng.choice(m, 4, replace=False), j] = rng.standard_normal(4)
    for i in rng.choice(m, 6, replace=False):
        A[i, rng.choice(n, 1500, replace=False)] = rng.uniform(-1, 1, 1500)
    for j in rng.choice(n, 4, replace=False):
        A[rng.choice(m, 700, replace=False), j] = rng.standard_normal(700)
    return A, rng


def _close(got, exp, rel):
    scale = max(1e-300, float(np.max(np.abs(exp))) if np.ndim(exp) else abs(exp))
    assert np.max(np.abs(np.asarray(got) - exp)) <= rel * scale, (got, exp)


def test_residual_pass_closed_form(gpu):
    A, rng = _lp(3)
    m, n = A.shape
    b, c = rng.standard_normal(m), rng.standard_normal(n)
    R, Cs = rng.uniform(0.5, 2, m), rng.uniform(0.5, 2, n)
    x, s, y = rng.uniform(0.1, 2, n), rng.uniform(0.1, 2, n), rng.standard_normal(m)
    ps = Passes(gpu, A, b, c, R, Cs)
    assert ps.P.long_rows.n == 6 and ps.P.long_cols.n == 4 and ps.P.long_rows.split == 2
    L = gpu.lib()
    vp = C.c_void_p
    L.lkd_ipm_residuals.argtypes = [C.POINTER(Ipm), vp, vp, vp, vp, vp, vp, vp]
    L.lkd_ipm_residuals.restype = C.c_int
    d_x, d_y, d_s = ps.dev(x, np.float64), ps.dev(y, np.float64), ps.dev(s, np.float64)
    rp, rd, th = ps.buf(m), ps.buf(n), ps.buf(n)
    assert L.lkd_ipm_residuals(C.byref(ps.P), d_x, d_y, d_s, rp, rd, th, None) == 0
    assert L.lkd_stream_sync(None) == 0
    rp_h = np.zeros(m); rd_h = np.zeros(n); th_h = np.zeros(n)
    for arr, p in ((rp_h, rp), (rd_h, rd), (th_h, th)):
        assert L.lkd_memcpy_dtoh(arr.ctypes.data, p, arr.nbytes, None) == 0
    S = ps.slots()
    e_rp, e_rd = A @ x - b, A.T @ y + s - c
    _close(rp_h, e_rp, 1e-13)
    _close(rd_h, e_rd, 1e-13)
    np.testing.assert_array_equal(th_h, x / s)
    _close(S[SLOTS["RP2"]], e_rp @ e_rp, 1e-12)
    _close(S[SLOTS["RD2"]], e_rd @ e_rd, 1e-12)
    _close(S[SLOTS["BY"]], b @ y, 1e-12)
    _close(S[SLOTS["XS"]], x @ s, 1e-13)
    _close(S[SLOTS["CX"]], c @ x, 1e-12)
    _close(S[SLOTS["RPINF"]], np.max(np.abs(e_rp / R)), 1e-13)
    _close(S[SLOTS["RDINF"]], np.max(np.abs(e_rd / Cs)), 1e-13)
    _close(S[SLOTS["XSMAX"]], np.max(np.abs(x * (s - e_rd))), 1e-13)
    assert S[SLOTS["THMAX"]] == np.max(x / s) and S[SLOTS["THMIN"]] == np.min(x / s)


def test_direction_pass_closed_form(gpu):
    A, rng = _lp(4)
    m, n = A.shape
    b, c = rng.standard_normal(m), rng.standard_normal(n)
    ones_m, ones_n = np.ones(m), np.ones(n)
    x, s = rng.uniform(0.1, 2, n), rng.uniform(0.1, 2, n)
    theta = x / s
    rdv, rpv = rng.standard_normal(n), rng.standard_normal(m)
    rxs, dy = rng.standard_normal(n), rng.standard_normal(m)
    ps = Passes(gpu, A, b, c, ones_m, ones_n)
    L = gpu.lib()
    vp = C.c_void_p
    L.lkd_ipm_direction.argtypes = [C.POINTER(Ipm)] + [vp] * 13
    L.lkd_ipm_direction.restype = C.c_int
    args = [ps.dev(v, np.float64) for v in (x, s, np.zeros(m), theta, rdv, rpv, rxs, dy)]
    dx, ds, qd, qp = ps.buf(n), ps.buf(n), ps.buf(n), ps.buf(m)
    assert L.lkd_ipm_direction(C.byref(ps.P), *args, dx, ds, qd, qp, None) == 0
    assert L.lkd_stream_sync(None) == 0
    out = {}
    for name, p, k in (("dx", dx, n), ("ds", ds, n), ("qd", qd, n), ("qp", qp, m)):
        out[name] = np.zeros(k)
        assert L.lkd_memcpy_dtoh(out[name].ctypes.data, p, 8 * k, None) == 0
    S = ps.slots()
    t = A.T @ dy
    e_ds = -t - rdv
    e_dx = -theta * e_ds + rxs / s
    _close(out["ds"], e_ds, 1e-13)
    _close(out["dx"], e_dx, 1e-12)
    _close(out["qd"], t + e_ds, 1e-12)
    _close(out["qp"], A @ out["dx"], 1e-13)
    e1 = A @ e_dx + rpv
    _close(S[SLOTS["R1"]], e1 @ e1, 1e-11)
    # e2 and e3 vanish in exact arithmetic: only rounding is left
    assert S[SLOTS["R2"]] <= n * (1e-14 * np.abs(t).max()) ** 2 * 16
    assert S[SLOTS["R3"]] <= n * (1e-13 * np.abs(rxs).max()) ** 2 * 16
    _close(S[SLOTS["RXS2"]], rxs @ rxs, 1e-13)
    neg = e_dx < 0
    _close(S[SLOTS["APMAX"]], np.min(-x[neg] / e_dx[neg]) if neg.any() else 1e30, 1e-12)
    neg = e_ds < 0
    _close(S[SLOTS["ADMAX"]], np.min(-s[neg] / e_ds[neg]) if neg.any() else 1e30, 1e-12)
    _close(S[SLOTS["DX2"]], e_dx @ e_dx, 1e-12)
    _close(S[SLOTS["DS2"]], e_ds @ e_ds, 1e-12)
    _close(S[SLOTS["DY2"]], dy @ dy, 1e-13)

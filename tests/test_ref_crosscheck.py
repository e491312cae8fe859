"""Cross-check against the reference's own CPU sources (oracle/_ref/libref.so, built from
src/ipm/symbolic.cpp, src/verifier/certificate.cpp and src/core by oracle/build_ref.sh where the
reference tree is present; skipped elsewhere): the exact minimum-degree ordering rule, the
elimination tree / L pattern, and verify_optimal, for both the C oracle and the library; and
the CPU simplex route timed as the benchmark's CPU baseline (solve_with_phase1 over the sparse
LU) against the oracle IPM's optimum."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference CPU sources not built here")


@pytest.fixture(scope="module")
def ref():
    L = C.CDLL(REF)
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.ref_order.argtypes = [C.c_int, C.c_int, ip, ip, ip]
    L.ref_symbolic.argtypes = [C.c_int, C.c_int, ip, ip, ip, ip, ip, ip, ip, C.c_longlong,
                               C.POINTER(C.c_longlong)]
    L.ref_verify.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, dp, dp, C.c_int, dp, dp,
                             C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.ref_phase1_solve.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), dp, C.POINTER(C.c_double)]
    return L


def random_csr(m, n, k, seed):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for j in range(n):
        r = rng.choice(m, size=min(k, m), replace=False)
        rows += r.tolist()
        cols += [j] * len(r)
    A = sp.csr_matrix((rng.standard_normal(len(rows)), (rows, cols)), shape=(m, n))
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A


CASES = [(12, 30, 2, 1), (40, 90, 3, 2), (80, 200, 3, 3), (120, 300, 4, 4), (150, 260, 2, 5)]


@pytest.mark.parametrize("m,n,k,seed", CASES)
def test_ordering_matches_reference(ref, native, oracle, m, n, k, seed):
    rp, ci, _ = random_csr(m, n, k, seed)
    P = np.zeros(m, np.int32)
    assert ref.ref_order(m, n, rp, ci, P) == 0
    np.testing.assert_array_equal(native.order_reference(rp, ci, n), P)
    np.testing.assert_array_equal(oracle.md_order(rp, ci, n), P)


@pytest.mark.parametrize("m,n,k,seed", CASES)
def test_symbolic_matches_reference(ref, native, oracle, m, n, k, seed):
    rp, ci, _ = random_csr(m, n, k, seed)
    P = np.zeros(m, np.int32)
    assert ref.ref_order(m, n, rp, ci, P) == 0
    parent, pinv, Lptr = (np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m + 1, np.int32))
    nnz = C.c_longlong(0)
    cap = m * m
    Lidx = np.zeros(cap, np.int32)
    assert ref.ref_symbolic(m, n, rp, ci, P, parent, pinv, Lptr, Lidx, cap, C.byref(nnz)) == 0
    Lidx = Lidx[:nnz.value]
    got = native.symbolic(rp, ci, n, P)
    np.testing.assert_array_equal(got["parent"], parent)
    np.testing.assert_array_equal(got["P_inv"], pinv)
    np.testing.assert_array_equal(got["Lptr"], Lptr)
    np.testing.assert_array_equal(got["Lidx"], Lidx)
    op, oL, oi = oracle.symbolic(rp, ci, n, P)
    np.testing.assert_array_equal(op, parent)
    np.testing.assert_array_equal(oL, Lptr)
    np.testing.assert_array_equal(oi, Lidx)


def test_verifier_matches_reference_on_random_certificates(ref, native, oracle):
    from lakshya.gen import LP
    rng = np.random.default_rng(9)
    agree = 0
    for trial in range(300):
        m, n = int(rng.integers(1, 5)), int(rng.integers(1, 7))
        A = sp.random(m, n, density=0.6, random_state=int(rng.integers(1 << 30)), format="csc")
        A.data = np.round(A.data * 4 - 2, 1)
        A.sort_indices()
        x = np.round(rng.uniform(-1, 2, n), 1)
        x[rng.random(n) < 0.3] = 0.0
        b = A @ x + (rng.random(m) < 0.2) * 1e-3
        pi = np.round(rng.uniform(-1, 1, m), 1)
        c = A.T @ pi + np.where(rng.random(n) < 0.5, 0.0, np.round(rng.uniform(-1, 1, n), 1))
        lb = np.where(rng.random(n) < 0.2, -np.inf, np.minimum(0.0, x))
        ub = np.where(rng.random(n) < 0.5, np.inf, np.maximum(1.0, x))
        at = rng.random(n) < 0.2                 # some variables exactly at a finite lower bound
        lb[at] = x[at]
        maximize = int(rng.random() < 0.2)
        cp, ri = A.indptr.astype(np.int32), A.indices.astype(np.int32)
        val = np.ascontiguousarray(A.data, np.float64)
        p, r = C.c_int(0), C.c_double(0.0)
        assert ref.ref_verify(m, n, cp, ri, val, np.ascontiguousarray(c), np.ascontiguousarray(b),
                              np.ascontiguousarray(lb), np.ascontiguousarray(ub), maximize,
                              np.ascontiguousarray(x), np.ascontiguousarray(pi), C.byref(p), C.byref(r)) == 0
        lp = LP(m, n, cp, ri, val, b, c, "rand")
        mine, _ = native.verify_optimal(lp, x, pi, 1e-6, lb=lb, ub=ub, maximize=bool(maximize))
        orc, _ = oracle.verify_optimal(m, n, cp, ri, val, c, b, lb, ub, x, pi, 1e-6, bool(maximize))
        assert mine == bool(p.value) == orc, trial
        agree += bool(p.value)
    assert 0 < agree < 300      # both outcomes exercised


@pytest.mark.parametrize("m,n,seed", [(40, 100, 3), (120, 300, 4)])
def test_reference_simplex_route_reaches_the_ipm_optimum(ref, oracle, m, n, seed):
    """The CPU baseline's route (Phase I + Phase II primal simplex, sparse LU) and the oracle
    IPM at a 1e-8 gap agree on the objective of a feasible, bounded random LP; the simplex
    point is feasible."""
    from lakshya import gen
    lp = gen.random_sparse(m, n, 6, seed=seed)
    st, obj, ms = C.c_int(), C.c_double(), C.c_double()
    x = np.zeros(n)
    assert ref.ref_phase1_solve(m, n, lp.indptr.astype(np.int32), lp.indices.astype(np.int32),
                                lp.data, lp.c, lp.b, C.byref(st), C.byref(obj), x, C.byref(ms)) == 0
    assert st.value == 0                                     # Optimal
    r = oracle.ipm_solve(lp.csr(), lp.b, lp.c, n, stop_rule=1)
    assert r["status"] == 0
    assert abs(obj.value - r["objective"]) <= 1e-8 * max(1.0, abs(r["objective"])) * 10
    A = lp.to_scipy()
    assert np.abs(A @ x - lp.b).max() <= 1e-7 * (1 + np.abs(lp.b).max())
    assert x.min() >= -1e-9

"""The simplex route's CPU oracle (oracle/pysimplex.py) pinned three ways before it judges the
device: the known answers of the reference's Phase-I unit tests (tests/unit/test_phase1.cpp:
123-430), enumeration of every basis of tiny LPs (the optimum is the smallest feasible basic
objective), and -- where oracle/_ref is built from the reference sources -- the reference's own
route, solve_with_phase1 over its sparse LU, on random feasible, infeasible and unbounded LPs
(same status, same objective)."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")


@pytest.fixture(scope="module")
def ps():
    import pysimplex
    return pysimplex


def lps_for_routes(seed, kind, m=30, n=75):
    """Random LP of the batch family: feasible ('opt'), infeasible ('inf': nonnegative A and a
    negative right-hand side) or unbounded ('unb': a column pair a, -a with negative cost sum)."""
    from lakshya import gen
    lp = gen.batch_lp(m, seed, density=0.15)
    A = lp.to_scipy().toarray()
    b, c = lp.b.copy(), lp.c.copy()
    if kind == "inf":
        A = np.abs(A)
        b = -np.abs(b) - 1.0
    elif kind == "unb":
        A[:, 1] = -A[:, 0]
        c[0], c[1] = -1.0, -1.0
    return A, b, c


def csc(A):
    import scipy.sparse as sp
    M = sp.csc_matrix(A)
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64)


@pytest.mark.parametrize("name", ["test1_slack_basis_123", "test2_equality_161", "test3_negative_rhs_198",
                                  "test4_phase1_zero_229", "test5_infeasible_264",
                                  "test6_redundant_row_294", "test8_original_basis_366",
                                  "test9_full_solve_410"])
def test_oracle_known_answers(ps, name):
    m, n, cp, ri, va, b, c, status, obj, x = ps.PHASE1_CASES[name]
    A, b, c = ps.case_dense(name)
    r = ps.solve(A, b, c)
    assert r["status"] == status, r["status_name"]
    if obj is not None:
        assert abs(r["objective"] - obj) <= 1e-9
        assert np.abs(A @ r["x"] - b).max() <= 1e-9 and r["x"].min() >= 0.0
    if x is not None:
        np.testing.assert_allclose(r["x"], x, atol=1e-9)
    if status == 0:
        assert (r["basis"] >= 0).all() and (r["basis"] < n).all()   # only structural columns


def test_oracle_mps_slack_row_is_unbounded(ps):
    """test_phase1.cpp:329-354: one row x1 + x2 + s = 4 with costs -1, -1, 0 and no upper
    bounds -- bounded here (the row caps x1 + x2): the reference accepts Optimal or Unbounded."""
    A = np.array([[1.0, 1.0, 1.0]])
    r = ps.solve(A, [4.0], [-1.0, -1.0, 0.0])
    assert r["status"] == 0 and abs(r["objective"] + 4.0) <= 1e-12


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_brute_force(ps, seed):
    """Tiny random LPs (b = A x* with x* >= 0 having zeros, so degenerate vertices occur; c >= 0
    keeps them bounded): the oracle's optimum is the best feasible basic solution."""
    rng = np.random.default_rng(seed)
    m, n = 3, 7
    A = rng.integers(-3, 4, size=(m, n)).astype(float)
    xs = rng.integers(0, 3, size=n).astype(float)
    b = A @ xs
    c = rng.integers(0, 5, size=n).astype(float) + rng.uniform(0, 0.5, n)
    r = ps.solve(A, b, c)
    best = ps.brute_force(A, b, c)
    if best is None:
        assert r["status"] in (1, 3)
        return
    if np.linalg.matrix_rank(A) < m:
        assert r["status"] in (0, 3)
        if r["status"] == 3:
            return
    assert r["status"] == 0
    assert abs(r["objective"] - best) <= 1e-9 * max(1.0, abs(best))
    assert np.abs(A @ r["x"] - b).max() <= 1e-9 * (1 + np.abs(b).max()) and r["x"].min() >= -1e-12


def test_brute_force_catches_a_wrong_vertex(ps):
    """The brute-force check is sharp: a suboptimal basic solution of a tiny LP differs from it."""
    A = np.array([[1.0, 1.0, 1.0, 0.0], [1.0, -1.0, 0.0, 1.0]])
    b = np.array([4.0, 1.0])
    c = np.array([-3.0, -2.0, 0.0, 0.0])
    best = ps.brute_force(A, b, c)
    assert abs(best - (-10.5)) <= 1e-12            # x = (2.5, 1.5)
    assert abs(ps.solve(A, b, c)["objective"] - best) <= 1e-12
    assert c @ np.array([1.0, 0.0, 3.0, 0.0]) > best + 1.0   # the vertex x1 = 1 (row 2 tight)


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("reference CPU sources not built here")
    L = C.CDLL(REF)
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.ref_phase1_solve.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), dp, C.POINTER(C.c_double)]
    return L


@pytest.mark.parametrize("kind,seed", [("opt", s) for s in range(6)] + [("inf", 1), ("inf", 2), ("unb", 3), ("unb", 4)])
def test_oracle_matches_reference_route(ps, ref, kind, seed):
    A, b, c = lps_for_routes(seed, kind)
    m, n = A.shape
    cp, ri, va = csc(A)
    st, obj, ms = C.c_int(), C.c_double(), C.c_double()
    x = np.zeros(n)
    assert ref.ref_phase1_solve(m, n, cp, ri, va, c, b, C.byref(st), C.byref(obj), x, C.byref(ms)) == 0
    r = ps.solve(A, b, c)
    assert r["status"] == st.value, (r["status_name"], st.value)
    assert r["status"] == {"opt": 0, "inf": 1, "unb": 2}[kind]
    if kind == "opt":
        assert abs(r["objective"] - obj.value) <= 1e-9 * max(1.0, abs(obj.value))

"""The CPU oracle pinned against the reference's own known-answer tests and HiGHS objectives."""
import json
import os

import numpy as np
import pytest

from conftest import lp_from_rows

HERE = os.path.dirname(os.path.abspath(__file__))


def test_symbolic_star_kat(oracle):
    """tests/unit/test_ipm_symbolic.cpp:10-78: star graph -> P = {1,2,3,0,4}, nnz(L) = 4."""
    rp = np.array([0, 4, 5, 6, 7, 8]); ci = np.array([0, 1, 2, 3, 0, 1, 2, 3])
    P = oracle.md_order(rp, ci, 5)
    assert list(P) == [1, 2, 3, 0, 4]
    parent, Lptr, Lidx = oracle.symbolic(rp, ci, 5, P)
    assert Lptr[5] == 4
    assert list(parent) == [3, 3, 3, 4, -1]


def test_symbolic_fill_kat(oracle):
    """tests/unit/test_ipm_symbolic.cpp:80-151: 4-cycle + isolated node, one fill edge."""
    rp = np.array([0, 2, 4, 6, 8, 9]); ci = np.array([0, 3, 0, 1, 1, 2, 2, 3, 4])
    P = oracle.md_order(rp, ci, 5)
    assert list(P) == [4, 0, 1, 2, 3]
    parent, Lptr, Lidx = oracle.symbolic(rp, ci, 5, P)
    assert list(Lptr) == [0, 0, 0, 1, 2, 5]
    assert list(Lidx) == [1, 2, 1, 2, 3]
    assert list(parent) == [-1, 2, 3, 4, -1]


def test_symbolic_matches_dense_fill(oracle):
    """L pattern = nonzero pattern of chol(P M P^T) for a random SPD M with the same graph."""
    rng = np.random.default_rng(5)
    for t in range(20):
        m, n = int(rng.integers(4, 25)), int(rng.integers(10, 60))
        A = (rng.random((m, n)) < 0.12) * rng.standard_normal((m, n))
        A[np.arange(m), rng.integers(0, n, m)] = 1.0
        import scipy.sparse as sp
        Ac = sp.csr_matrix(A); Ac.sort_indices()
        rp, ci = Ac.indptr, Ac.indices
        P = oracle.md_order(rp, ci, n)
        parent, Lptr, Lidx = oracle.symbolic(rp, ci, n, P)
        G = (np.abs(A) > 0).astype(float)
        M = G @ G.T + m * np.eye(m)                      # generic values on the A A^T pattern
        Mp = M[np.ix_(P, P)]
        L = np.linalg.cholesky(Mp)
        pat = set()
        for i in range(m):
            for j in range(i):
                if abs(L[i, j]) > 1e-12:
                    pat.add((i, j))
        got = {(i, int(Lidx[q])) for i in range(m) for q in range(Lptr[i], Lptr[i + 1])}
        assert pat <= got
        # structural fill can exceed numerical fill only through exact cancellation (none here)
        assert got == pat


def test_kkt_kat(oracle):
    """tests/unit/test_gpu_kkt.cu:14-123 on the oracle: P = {1,0,2}, L L^T = M_perm, solve."""
    rp = np.array([0, 3, 5, 7]); ci = np.array([0, 3, 4, 1, 3, 2, 4])
    va = np.array([2.0, 1, 1, 2, 1, 2, 1])
    P = oracle.md_order(rp, ci, 5)
    assert list(P) == [1, 0, 2]
    st, L, dy, ok = oracle.kkt_factor_solve(rp, ci, va, np.ones(5), P, rhs=[11.0, 11, 16], want_L=True)
    Mp = np.array([[5.0, 1, 0], [1, 6, 1], [0, 1, 5]])
    assert st == 0 and ok
    assert np.linalg.norm(L @ L.T - Mp) < 1e-6
    assert np.allclose(dy, [1, 2, 3], atol=1e-6)
    st, L, dy, ok = oracle.kkt_factor_solve(rp, ci, va, np.ones(5), P, rhs=np.array([11.0, 11, 16]) * 1e4)
    assert ok and np.allclose(dy, 1e4 * np.array([1, 2, 3]), rtol=1e-8)


def test_assembly_closed_form(oracle):
    """M0 = P A Theta A^T P^T (dense check of the merge restatement)."""
    rng = np.random.default_rng(1)
    from lakshya import gen
    lp = gen.random_sparse(30, 80, 4, seed=3)
    rp, ci, va = lp.csr()
    th = np.exp(rng.uniform(-2, 2, lp.n))
    P = rng.permutation(lp.m).astype(np.int32)
    M = oracle.assemble_dense(rp, ci, va, th, P)
    A = lp.to_scipy().toarray()
    ref = (A * th) @ A.T
    assert np.allclose(M, ref[np.ix_(P, P)], rtol=1e-13, atol=1e-13)


def test_ruiz_invariant(oracle):
    """Scaled matrix = R A C; row/column max-abs of the scaled matrix tend to 1."""
    from lakshya import gen
    lp = gen.random_sparse(50, 120, 4, seed=9)
    rp, ci, va = lp.csr()
    v2, b2, c2, R, Cs = oracle.ruiz(rp, ci, va, lp.b, lp.c)
    A = lp.to_scipy().toarray()
    import scipy.sparse as sp
    As = sp.csr_matrix((v2, ci, rp), shape=A.shape).toarray()
    assert np.allclose(As, R[:, None] * A * Cs[None, :], rtol=1e-12)
    assert np.allclose(b2, R * lp.b) and np.allclose(c2, Cs * lp.c)
    assert np.all(np.abs(np.abs(As).max(1) - 1) < 1e-2) and np.all(np.abs(np.abs(As).max(0) - 1) < 1e-2)


def _gauss_seidel_ruiz(A, passes=10):
    """A planted mistake for the KAT below: rows scaled first, then the columns of the
    row-scaled matrix (the ordering of src/presolve/scaling.cpp:39-115, not the IPM's)."""
    A = A.copy()
    R, Cs = np.ones(A.shape[0]), np.ones(A.shape[1])
    for _ in range(passes):
        fr = 1 / np.sqrt(np.abs(A).max(1))
        A *= fr[:, None]
        R *= fr
        fc = 1 / np.sqrt(np.abs(A).max(0))
        A *= fc[None, :]
        Cs *= fc
    return R, Cs


def test_ruiz_jacobi_kat(oracle, native):
    """Hand-computed Jacobi Ruiz (src/ipm/mehrotra_device.cu:293-323) on A = [[4,1],[1,1]]:
    pass 1 takes row max (4, 1) and column max (4, 1) from the same snapshot, so
    R = C = (1/2, 1) and the scaled A = [[1, 1/2], [1/2, 1]]; every later pass sees unit maxima
    and changes nothing.  A Gauss-Seidel Ruiz gives C_0 = 1/sqrt(2) after its first pass."""
    rp = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 0, 1], np.int32)
    va = np.array([4.0, 1.0, 1.0, 1.0])
    b, c = np.array([3.0, 5.0]), np.array([7.0, 11.0])
    for impl in (oracle.ruiz, native.ruiz_scaling):
        v2, b2, c2, R, Cs = impl(rp, ci, va.copy(), b.copy(), c.copy())
        assert R.tolist() == [0.5, 1.0] and Cs.tolist() == [0.5, 1.0]
        assert v2.tolist() == [1.0, 0.5, 0.5, 1.0]
        assert b2.tolist() == [1.5, 5.0] and c2.tolist() == [3.5, 11.0]
    Rg, Cg = _gauss_seidel_ruiz(np.array([[4.0, 1.0], [1.0, 1.0]]), passes=1)
    assert abs(Cg[0] - 2 ** -0.5) < 1e-15 and Cg[0] != 0.5


def test_ruiz_gauss_seidel_is_rejected(oracle):
    """The invariant check below also holds for Gauss-Seidel; this one separates the two on a
    random matrix: the library-order (Jacobi) R and C differ from Gauss-Seidel's."""
    from lakshya import gen
    lp = gen.random_sparse(50, 120, 4, seed=9)
    rp, ci, va = lp.csr()
    _, _, _, R, Cs = oracle.ruiz(rp, ci, va.copy(), lp.b.copy(), lp.c.copy())
    Rg, Cg = _gauss_seidel_ruiz(lp.to_scipy().toarray())
    assert np.max(np.abs(R / Rg - 1)) > 1e-3 or np.max(np.abs(Cs / Cg - 1)) > 1e-3


@pytest.mark.parametrize("seed", [9, 21])
def test_ruiz_library_matches_oracle_bitwise(oracle, native, seed):
    """The library's host equilibration (run in the MehrotraSolver constructor) gives the
    oracle's R, C, scaled A, b and c bit for bit."""
    from lakshya import gen
    lp = gen.random_sparse(300, 700, 5, seed=seed)
    rp, ci, va = lp.csr()
    va = va * np.exp(np.random.default_rng(seed).uniform(-6, 6, len(va)))   # badly scaled
    o = oracle.ruiz(rp, ci, va.copy(), lp.b.copy(), lp.c.copy())
    g = native.ruiz_scaling(rp, ci, va.copy(), lp.b.copy(), lp.c.copy())
    for a, b in zip(o, g):
        assert np.array_equal(a, b)


def test_textbook_lp(oracle):
    """tests/unit/test_mehrotra.cu:29-78: Optimal, obj -20 +- 1e-4, residuals and mu < 1e-6."""
    lp = lp_from_rows([[1, 1, 1, 0], [2, 3, 0, 1]], [4, 12], [-3, -5, 0, 0], 4)
    r = oracle.ipm_solve(lp.csr(), lp.b, lp.c, 4)
    assert r["status"] == 0 and r["iterations"] > 0
    assert r["primal_residual"] < 1e-6 and r["dual_residual"] < 1e-6 and r["duality_gap"] < 1e-6
    assert abs(r["objective"] + 20) <= 1e-4
    assert np.isfinite(r["x"]).all() and np.isfinite(r["y"]).all()


def test_ill_conditioned_fallback_model(oracle):
    """tests/unit/test_phase1.cpp:483-530: entries 1e-7; Optimal => certificate passes."""
    lp = lp_from_rows([[1e-7, 1.0], [1.0, 1e-7]], [1.0, 1.0], [1.0, 1.0], 2)
    r = oracle.ipm_solve(lp.csr(), lp.b, lp.c, 2)
    assert r["status"] in (0, 3)
    if r["status"] == 0:
        ok, _ = oracle.verify_optimal(2, 2, lp.indptr, lp.indices, lp.data, lp.c, lp.b,
                                      np.zeros(2), np.full(2, np.inf), r["x"], r["y"])
        assert ok


def _golden():
    with open(os.path.join(HERE, "golden", "objectives.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["rs_60_150_3_s1", "rs_200_500_6_s2", "ba_4x60x150_L5_s3",
                                  "tr_12x12_side30_s4", "dense_80_s5"])
def test_oracle_vs_highs_golden(oracle, name):
    from golden_cases import build_case
    g = _golden()[name]
    lp = build_case(name)
    assert lp.digest() == g["digest"]
    r = oracle.ipm_solve(lp.csr(), lp.b, lp.c, lp.n, stop_rule=1)
    assert r["status"] == 0
    assert abs(r["objective"] - g["objective"]) <= 1e-8 * max(1.0, abs(g["objective"]))
    ok, _ = oracle.verify_optimal(lp.m, lp.n, lp.indptr, lp.indices, lp.data, lp.c, lp.b,
                                  np.zeros(lp.n), np.full(lp.n, np.inf), r["x"], r["y"], tol=1e-6)
    assert ok

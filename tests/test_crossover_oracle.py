"""The crossover oracle (oracle/pycrossover.py), pinned without the device: every vertex it returns
against brute-force enumeration of all bases of small LPs, known answers for each status code, a
point in the relative interior of an optimal face (purification must reach one of its two
vertices) and an exact vertex (no purification pivot).  A wrong reduced-cost sign or ratio-test
direction ends at a non-optimal vertex or a wrong status and fails one of these."""
import itertools

import numpy as np
import pytest

import pycrossover as X


def _small_lp(rng, m, n):
    """Bounded, feasible LP: the last row bounds sum(x); b = A x0 for a positive x0."""
    A = rng.normal(size=(m, n))
    A[-1] = 1.0
    x0 = rng.uniform(0.2, 1.0, n)
    return A, A @ x0, rng.normal(size=n)


def _brute_force(A, b, c):
    """Best objective over the feasible bases (every m-column subset with a nonsingular B)."""
    m, n = A.shape
    best, verts = np.inf, []
    for cols in itertools.combinations(range(n), m):
        B = A[:, cols]
        if abs(np.linalg.det(B)) < 1e-10:
            continue
        xb = np.linalg.solve(B, b)
        if xb.min() < -1e-10:
            continue
        x = np.zeros(n)
        x[list(cols)] = xb
        verts.append(x)
        best = min(best, float(c @ x))
    return best, verts


def _interior_optimum(A, b, c):
    """An optimal point from the oracle IPM (a strictly complementary, non-vertex point)."""
    import pyoracle
    import scipy.sparse as sp
    Ac = sp.csr_matrix(A)
    r = pyoracle.ipm_solve((Ac.indptr, Ac.indices, Ac.data), b, c, A.shape[1], stop_rule=1)
    assert r["status"] == 0
    return r["x"]


@pytest.mark.parametrize("seed", range(12))
def test_vertices_match_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    m, n = 3 + seed % 2, 7 + seed % 3
    A, b, c = _small_lp(rng, m, n)
    best, verts = _brute_force(A, b, c)
    basis, xv, st, piv = X.crossover(A, b, c, _interior_optimum(A, b, c))
    assert st == 0
    assert (basis >= 0).all() and len(set(basis.tolist())) == m
    assert abs(np.linalg.det(A[:, basis])) > 1e-10
    assert np.abs(A @ xv - b).max() < 1e-9 and xv.min() >= -1e-12
    assert set(np.flatnonzero(xv).tolist()) <= set(basis.tolist())
    assert abs(c @ xv - best) <= 1e-8 * max(1.0, abs(best))
    assert min(np.abs(v - xv).max() for v in verts) < 1e-8      # a vertex of the polytope


def test_exact_vertex_needs_no_purification():
    rng = np.random.default_rng(7)
    A, b, c = _small_lp(rng, 4, 9)
    best, verts = _brute_force(A, b, c)
    v = min(verts, key=lambda x: c @ x)
    basis, xv, st, piv = X.crossover(A, b, c, v)
    assert st == 0 and piv == 4
    assert set(np.flatnonzero(v > 1e-12).tolist()) <= set(basis.tolist())
    np.testing.assert_allclose(xv, v, atol=1e-12)


def test_optimal_face_midpoint_purifies_to_a_face_vertex():
    # min -(x1 + x2)  s.t.  x1 + x2 + s1 = 2,  x1 + s2 = 1.5,  x2 + s3 = 1.5: the optimal face
    # is the segment (0.5, 1.5) -- (1.5, 0.5); its midpoint (1, 1) has support 4 > m = 3
    A = np.array([[1.0, 1, 1, 0, 0], [1, 0, 0, 1, 0], [0, 1, 0, 0, 1]])
    b = np.array([2.0, 1.5, 1.5])
    c = np.array([-1.0, -1, 0, 0, 0])
    x = np.array([1.0, 1.0, 0.0, 0.5, 0.5])
    basis, xv, st, piv = X.crossover(A, b, c, x)
    assert st == 0
    assert c @ xv == pytest.approx(-2.0, abs=1e-14)
    assert any(np.allclose(xv[:2], v, atol=1e-14) for v in ([0.5, 1.5], [1.5, 0.5]))
    assert np.count_nonzero(xv) <= 3 and xv.min() >= 0


def test_status_unbounded_direction():
    # min -x1  s.t.  x1 - x2 = 0: from (1, 1) the ray x1 = x2 -> inf lowers the cost
    basis, xv, st, piv = X.crossover(np.array([[1.0, -1.0]]), np.array([0.0]), np.array([-1.0, 0.0]),
                                     np.array([1.0, 1.0]))
    assert st == 3


def test_status_not_dual_feasible():
    # min -x1  s.t.  x1 + x2 = 1 from the vertex (0, 1): feasible basis {x2}, reduced cost of
    # x1 < 0; without the clean-up that is status 1, with it one primal pivot reaches (1, 0)
    A, b, c, x = np.array([[1.0, 1.0]]), np.array([1.0]), np.array([-1.0, 0.0]), np.array([0.0, 1.0])
    basis, xv, st, piv = X.crossover(A, b, c, x, cleanup=False)
    assert st == 1 and basis.tolist() == [1] and xv.tolist() == [0.0, 1.0]
    basis, xv, st, piv = X.crossover(A, b, c, x)
    assert st == 0 and basis.tolist() == [0] and xv.tolist() == [1.0, 0.0] and piv == 2


def test_status_basic_solution_infeasible():
    # x1 + x2 = 1, x1 - x2 = 3: the only basis gives x2 = -1
    basis, xv, st, piv = X.crossover(np.array([[1.0, 1.0], [1.0, -1.0]]), np.array([1.0, 3.0]),
                                     np.zeros(2), np.array([1.0, 1.0]))
    assert st == 4 and xv.tolist() == [2.0, -1.0]


def test_status_rank_deficient_rows():
    rng = np.random.default_rng(11)
    A, b, c = _small_lp(rng, 3, 8)
    x = _interior_optimum(A, b, c)
    A2, b2 = np.vstack([A, A[1]]), np.append(b, b[1])
    basis, xv, st, piv = X.crossover(A2, b2, c, x)
    assert st == 2 and (basis < 0).sum() == 1 and piv >= 3


def _degenerate_vertex_lp(seed, m=5, n=12, b_noise=0.0):
    """An optimal vertex v with m - 1 positive entries (primal degenerate); c makes v optimal
    with strictly positive reduced costs off its support.  b_noise perturbs b as an interior
    point's residual would, so the fillers' basic values are small and of either sign."""
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(m, n))
    S = rng.choice(n, m - 1, replace=False)
    v = np.zeros(n)
    v[S] = rng.uniform(0.5, 2.0, m - 1)
    b = A @ v + rng.normal(scale=b_noise, size=m) if b_noise else A @ v
    y = rng.normal(size=m)
    s = rng.uniform(0.5, 1.5, n)
    s[S] = 0.0
    x = v.copy()
    x[x == 0] = 1e-10
    return A, b, A.T @ y + s, x


@pytest.mark.parametrize("seed", [2, 4, 5, 8])
def test_clean_up_restores_dual_feasibility(seed):
    """At a degenerate vertex the basis completed by a filler column need not be dual feasible;
    the primal simplex clean-up (Bland) pivots to an optimal basis of the same vertex."""
    A, b, c, x = _degenerate_vertex_lp(seed)
    assert X.crossover(A, b, c, x, cleanup=False)[2] == 1
    basis, xv, st, piv = X.crossover(A, b, c, x)
    best, _ = _brute_force(A, b, c)
    assert st == 0 and abs(c @ xv - best) <= 1e-12 * max(1.0, abs(best))
    y = np.linalg.solve(A[:, basis].T, c[basis])
    assert (c - A.T @ y).min() >= -1e-9 * max(1.0, np.abs(c).max())


@pytest.mark.parametrize("seed", [2, 3, 5, 7])
def test_clean_up_restores_primal_feasibility(seed):
    """With b off the degenerate vertex by 1e-7 a filler's basic value can be negative; the dual
    simplex clean-up pivots it out and the result is the brute-force optimum."""
    A, b, c, x = _degenerate_vertex_lp(seed, b_noise=1e-7)
    assert X.crossover(A, b, c, x, cleanup=False)[2] == 4
    basis, xv, st, piv = X.crossover(A, b, c, x)
    best, _ = _brute_force(A, b, c)
    assert st == 0 and abs(c @ xv - best) <= 1e-8 * max(1.0, abs(best))
    assert xv.min() >= -1e-9 * max(1.0, np.abs(xv).max())

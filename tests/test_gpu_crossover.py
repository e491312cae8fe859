"""Batched crossover on the device (lk_batch_crossover) against the numpy oracle on the same
interior-point x: identical bases, statuses and pivot counts, basic values within 1e-9; and,
without the oracle, that every optimal-status output is a vertex (A x = b, x >= 0, support inside
a nonsingular basis) with the interior-point objective."""
import numpy as np
import pytest
import scipy.sparse as sp

import pycrossover as X

pytestmark = pytest.mark.gpu

# 16 LPs of the 4,096-LP batch (m from 50 to 800), plus two whose support is m - 1 and whose
# filler column comes out slightly negative before the dual simplex clean-up
MEMBERS = list(range(0, 4096, 273)) + [1043, 3294]


def _check_against_oracle(lp, x, out):
    A = lp.to_scipy().toarray()
    assert np.isfinite(x).all()
    basis, xv, st, piv = X.crossover(A, lp.b, lp.c, x)
    assert out["status"] == st, (lp.name, out["status"], st)
    assert out["pivots"] == piv, (lp.name, out["pivots"], piv)
    np.testing.assert_array_equal(out["basis"], basis)
    np.testing.assert_allclose(out["x"], xv, rtol=0, atol=1e-9 * max(1.0, np.abs(xv).max()))
    return A


def _check_vertex(A, lp, out, objective):
    xv, basis = out["x"], out["basis"]
    scale = max(1.0, np.abs(xv).max())
    assert (basis >= 0).all() and len(set(basis.tolist())) == lp.m
    assert np.linalg.matrix_rank(A[:, basis]) == lp.m
    assert np.abs(A @ xv - lp.b).max() <= 1e-8 * max(1.0, np.abs(lp.b).max())
    assert xv.min() >= -1e-9 * scale
    assert set(np.flatnonzero(xv).tolist()) <= set(basis.tolist())
    assert abs(lp.c @ xv - objective) <= 1e-6 * max(1.0, abs(objective))


@pytest.fixture(scope="module")
def solved(gpu):
    from lakshya import gen
    lps = gen.batch(4096, indices=MEMBERS)
    B = gpu.Batch(lps, stop_rule=1)
    B.solve()
    res = B.results()
    xs = [B.solution(i)[0] for i in range(len(lps))]
    outs = B.crossover()
    yield lps, res, xs, outs
    B.close()


def test_batch_members_match_oracle(solved):
    lps, res, xs, outs = solved
    for lp, x, out in zip(lps, xs, outs):
        _check_against_oracle(lp, x, out)


def test_given_points_match_the_batch_method(solved, gpu):
    """lk_crossover on the same x gives the bases and values lk_batch_crossover gave."""
    lps, res, xs, outs = solved
    again = gpu.crossover_points(lps, xs)
    for a, o in zip(again, outs):
        assert a["status"] == o["status"] and a["pivots"] == o["pivots"]
        np.testing.assert_array_equal(a["basis"], o["basis"])
        np.testing.assert_array_equal(a["x"], o["x"])


def test_status_known_answers(gpu):
    """Points whose outcome is fixed by hand (the oracle's known-answer cases): unbounded ray,
    a non-optimal vertex that the primal clean-up moves to (1, 0), an infeasible basis, a face
    midpoint."""
    from conftest import lp_from_rows
    cases = [
        (lp_from_rows([[1.0, -1.0]], [0.0], [-1.0, 0.0], 2), [1.0, 1.0], 3),
        (lp_from_rows([[1.0, 1.0]], [1.0], [-1.0, 0.0], 2), [0.0, 1.0], 0),
        (lp_from_rows([[1.0, 1.0], [1.0, -1.0]], [1.0, 3.0], [0.0, 0.0], 2), [1.0, 1.0], 4),
        (lp_from_rows([[1.0, 1, 1, 0, 0], [1, 0, 0, 1, 0], [0, 1, 0, 0, 1]], [2.0, 1.5, 1.5],
                      [-1.0, -1, 0, 0, 0], 5), [1.0, 1.0, 0.0, 0.5, 0.5], 0),
    ]
    outs = gpu.crossover_points([c[0] for c in cases], [c[1] for c in cases])
    for (lp, x, want), out in zip(cases, outs):
        assert out["status"] == want, (x, out)
        _check_against_oracle(lp, np.array(x), out)
    np.testing.assert_array_equal(outs[1]["x"], [1.0, 0.0])
    face = outs[3]["x"]
    assert face[0] + face[1] == pytest.approx(2.0, abs=1e-14)
    assert np.count_nonzero(face) <= 3


def test_clean_up_at_degenerate_vertices(gpu):
    """Degenerate optimal vertices (support m - 1) from the oracle's clean-up tests: exact b
    (the filler's basis is not dual feasible: primal clean-up) and b off by 1e-7 (the filler's
    value is negative: dual clean-up); device and oracle agree and end optimal."""
    from conftest import lp_from_rows
    from test_crossover_oracle import _degenerate_vertex_lp
    cases = [_degenerate_vertex_lp(s) for s in (2, 4, 5, 8)]
    cases += [_degenerate_vertex_lp(s, b_noise=1e-7) for s in (2, 3, 5, 7)]
    lps = [lp_from_rows(A, b, c, A.shape[1]) for A, b, c, x in cases]
    outs = gpu.crossover_points(lps, [x for A, b, c, x in cases])
    for lp, (A, b, c, x), out in zip(lps, cases, outs):
        assert X.crossover(A, b, c, x, cleanup=False)[2] in (1, 4)
        _check_against_oracle(lp, x, out)
        assert out["status"] == 0


def test_batch_members_reach_optimal_vertices(solved):
    lps, res, xs, outs = solved
    for lp, r, out in zip(lps, res, outs):
        assert r["status"] == 0 and out["status"] == 0, (lp.name, r["status"], out["status"])
        _check_vertex(lp.to_scipy().toarray(), lp, out, r["objective"])


def _degenerate_lp():
    """Dual degenerate and primal degenerate at once: b = A x* with x* supported on m + 25
    columns whose reduced costs c - A'y are zero, so the optimal face is not a vertex."""
    from lakshya.gen import LP
    rng = np.random.default_rng(3)
    m, n = 60, 150
    A = sp.random(m, n, density=0.1, random_state=4, data_rvs=rng.standard_normal).tocsc()
    A = (A + sp.csc_matrix((rng.standard_normal(n), (rng.integers(0, m, n), np.arange(n))),
                           shape=(m, n))).tocsc()
    A.sort_indices()
    xs = rng.uniform(0.5, 1.5, n)
    y = rng.standard_normal(m)
    s = rng.uniform(0.1, 1.0, n)
    zero = rng.choice(n, m + 25, replace=False)
    s[zero] = 0.0
    x0 = np.zeros(n)
    x0[zero] = xs[zero]
    return LP(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, A @ x0,
              A.T @ y + s, "degenerate_60x150")


def test_degenerate_face_is_purified(gpu):
    lp = _degenerate_lp()
    B = gpu.Batch([lp], stop_rule=1)
    try:
        B.solve()
        r = B.results()[0]
        x = B.solution(0)[0]
        out = B.crossover()[0]
    finally:
        B.close()
    assert r["status"] == 0
    assert int((x > 1e-6).sum()) > lp.m                # the interior point is not a vertex
    A = _check_against_oracle(lp, x, out)
    assert out["status"] == 0 and out["pivots"] > lp.m  # purification pivots beyond the basis
    _check_vertex(A, lp, out, r["objective"])


def test_duplicated_row_leaves_an_artificial_row(gpu):
    from lakshya import gen
    from lakshya.gen import LP
    lp = gen.batch(4096, indices=[300])[0]
    A = lp.to_scipy().tocsc()
    A2 = sp.vstack([A, A[3]]).tocsc()
    A2.sort_indices()
    dup = LP(lp.m + 1, lp.n, A2.indptr.astype(np.int32), A2.indices.astype(np.int32), A2.data,
             np.append(lp.b, lp.b[3]), lp.c, "duplicated_row")
    B = gpu.Batch([dup], stop_rule=1)
    try:
        B.solve()
        x = B.solution(0)[0]
        out = B.crossover()[0]
    finally:
        B.close()
    _check_against_oracle(dup, x, out)
    assert out["status"] == 2 and int((out["basis"] < 0).sum()) == 1


def test_crossover_before_solve_is_rejected(gpu):
    from lakshya import gen
    B = gpu.Batch(gen.batch(4096, indices=[1]), stop_rule=1)
    try:
        with pytest.raises(Exception, match="solve"):
            B.crossover()
    finally:
        B.close()

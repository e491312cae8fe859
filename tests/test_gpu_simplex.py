"""The device simplex route (lk_simplex: Phase I / Phase II primal simplex, one workgroup per LP)
against its CPU oracle (oracle/pysimplex.py, itself pinned to the reference's route and to
brute-force enumeration by tests/test_simplex_oracle.py): the known answers of the reference's
Phase-I tests, statuses and optima on random feasible / infeasible / unbounded LPs, and the
HiGHS optima of members of the 4,096-LP batch family."""
import json
import os

import numpy as np
import pytest

import pysimplex as PS
from test_simplex_oracle import lps_for_routes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class _LP:
    def __init__(self, A, b, c, name=""):
        import scipy.sparse as sp
        M = sp.csc_matrix(np.asarray(A, float))
        M.sort_indices()
        self.m, self.n = M.shape
        self.indptr, self.indices, self.data = M.indptr, M.indices, M.data
        self.b, self.c, self.name = np.asarray(b, float), np.asarray(c, float), name


def _certify(A, b, c, r):
    """An Optimal result: A x = b, x >= 0, the basis nonsingular, y = B^-T c_B dual feasible and
    c^T x = b^T y."""
    x, y, basis = r["x"], r["y"], r["basis"]
    assert (basis >= 0).all()
    assert np.abs(A @ x - b).max() <= 1e-8 * max(1.0, np.abs(b).max())
    assert x.min() >= -1e-9 * max(1.0, np.abs(x).max())
    assert np.linalg.matrix_rank(A[:, basis]) == A.shape[0]
    d = c - A.T @ y
    assert d.min() >= -1e-6 * max(1.0, np.abs(c).max())
    assert abs(c @ x - b @ y) <= 1e-8 * max(1.0, abs(c @ x))


def test_device_known_answers(gpu):
    names = list(PS.PHASE1_CASES)
    lps = [_LP(*PS.case_dense(nm), name=nm) for nm in names]
    out = gpu.simplex(lps)
    for nm, lp, r in zip(names, lps, out):
        m, n, cp, ri, va, b, c, status, obj, x = PS.PHASE1_CASES[nm]
        assert r["status"] == status, (nm, r["status_name"])
        if obj is not None:
            assert abs(r["objective"] - obj) <= 1e-9, nm
        if x is not None:
            np.testing.assert_allclose(r["x"], x, atol=1e-9)
        if status == 0:
            A, bb, cc = PS.case_dense(nm)
            _certify(A, bb, cc, r)


def test_device_matches_oracle(gpu):
    """One launch over 24 LPs of three kinds; statuses equal, optima within 1e-9, and the pivot
    counts of the two phases (the same rules on differently rounded inverses) equal on most."""
    cases = [(k, s) for k in ("opt",) for s in range(16)] + [("inf", s) for s in range(4)] + \
            [("unb", s) for s in range(4)]
    mats = [lps_for_routes(s, k, m=20 + 7 * s, n=int(2.5 * (20 + 7 * s))) for k, s in cases]
    lps = [_LP(A, b, c) for A, b, c in mats]
    out = gpu.simplex(lps)
    same_iters = 0
    for (kind, s), (A, b, c), r in zip(cases, mats, out):
        o = PS.solve(A, b, c)
        assert r["status"] == o["status"] == {"opt": 0, "inf": 1, "unb": 2}[kind], (kind, s, r["status_name"])
        if kind == "opt":
            assert abs(r["objective"] - o["objective"]) <= 1e-9 * max(1.0, abs(o["objective"])), (s,)
            _certify(A, b, c, r)
        same_iters += r["iterations"] == o["iterations"]
    assert same_iters >= 0.75 * len(cases), same_iters


def test_device_batch_members_reach_highs_optimum(gpu):
    """Members of the 4,096-LP family (m up to 800): Optimal, certified, objective within 1e-8 of
    the committed HiGHS value."""
    from lakshya import gen
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "batch_objectives.json")))
    sizes = gen.batch_sizes(4096)
    idx = sorted(gold, key=int)[:6]
    lps = [gen.batch_lp(int(sizes[int(i)]), int(i)) for i in idx]
    out = gpu.simplex(lps)
    for i, lp, r in zip(idx, lps, out):
        assert r["status"] == 0, (i, r["status_name"], r["iterations"])
        g = gold[i]["objective"]
        assert abs(r["objective"] - g) <= 1e-8 * max(1.0, abs(g)), (i, r["objective"], g)
        _certify(lp.to_scipy().toarray(), lp.b, lp.c, r)


def test_device_iteration_limit(gpu):
    """max_iter caps each phase: a limit below Phase I's pivot count ends in IterationLimit."""
    A, b, c = lps_for_routes(5, "opt", m=40, n=100)
    full = gpu.simplex([_LP(A, b, c)])[0]
    assert full["status"] == 0 and full["iterations"][0] > 3
    capped = gpu.simplex([_LP(A, b, c)], max_iter=2)[0]
    assert capped["status"] == 3 and capped["iterations"][0] == 2

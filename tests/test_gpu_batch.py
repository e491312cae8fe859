"""Batched interior-point solver on the device: parity with the oracle IPM per LP, certified
solutions, launch-shape independence (single launch, graph replays, eager replays), mixed
precision, the reference stop mode and HiGHS objectives of members of the 4,096-LP batch."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _small_batch(count=24, lo=50, hi=220, seed=7):
    from lakshya import gen
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=count)
    return [gen.batch_lp(int(m), seed * 1000 + i) for i, m in enumerate(sizes)]


def _certified(gpu, lp, B, i, r, tol=1e-8):
    assert r["status_name"] == "Optimal", (i, r["fail_code"], r["iterations"])
    assert r["rel_gap"] <= tol and r["rel_primal"] <= tol and r["rel_dual"] <= tol, (i, r)
    x, y, _ = B.solution(i)
    ok, _ = gpu.verify_optimal(lp, x, y, tol=tol)
    assert ok, i


def test_batch_matches_oracle(gpu, oracle):
    """Every LP of a mixed-size batch: Optimal, certified at 1e-8 in the original space, and the
    objective within 1e-8 (relative) of the oracle's dense IPM on the same LP."""
    lps = _small_batch()
    B = gpu.Batch(lps)
    s = B.solve()
    assert s["count"] == len(lps) and s["optimal"] == len(lps)
    res = B.results()
    for i, lp in enumerate(lps):
        o = oracle.ipm_solve(lp.csr(), lp.b, lp.c, lp.n, stop_rule=1)
        assert o["status"] == 0
        r = res[i]
        _certified(gpu, lp, B, i, r)
        assert abs(r["objective"] - o["objective"]) <= 1e-8 * max(1.0, abs(o["objective"])), i
        assert abs(r["iterations"] - o["iterations"]) <= 3, (i, r["iterations"], o["iterations"])
    B.close()


def test_batch_launch_shapes_agree_bitwise(gpu):
    """The LP state survives between launches: one persistent launch, graph replays of launches
    of 1 or 4 iterations, eager replays and either workgroup count give bit-identical results
    (every replay of the captured graph is checked against the eager result)."""
    lps = _small_batch(count=16, seed=11)
    ref = None
    shapes = [dict(), dict(iters_per_launch=1), dict(iters_per_launch=4),
              dict(iters_per_launch=1, use_graph=0), dict(wg_per_cu=1, iters_per_launch=3)]
    for kw in shapes:
        B = gpu.Batch(lps, **kw)
        for rep in range(2):              # two solves: the graph is replayed again
            s = B.solve()
            assert s["optimal"] == len(lps), kw
            if kw.get("iters_per_launch") and kw.get("use_graph", 1):
                assert s["graph_replays"] >= 1
            res = [(r["status"], r["iterations"], r["objective"], r["rel_gap"]) for r in B.results()]
            x0 = B.solution(0)[0].copy()
            if ref is None:
                ref, xref = res, x0
            assert res == ref, (kw, rep)
            assert np.array_equal(x0, xref), (kw, rep)
        B.close()


def test_batch_mixed_precision(gpu):
    """FP32 factors with FP64 PCG reach the same certified optimum as the FP64 batch."""
    lps = _small_batch(count=12, seed=13)
    B64 = gpu.Batch(lps)
    B64.solve()
    r64 = B64.results()
    B32 = gpu.Batch(lps, precision=1)
    s = B32.solve()
    assert s["optimal"] == len(lps)
    r32 = B32.results()
    assert sum(r["fp32_factorizations"] for r in r32) > 0
    for i, lp in enumerate(lps):
        _certified(gpu, lp, B32, i, r32[i])
        assert abs(r32[i]["objective"] - r64[i]["objective"]) <= 1e-8 * max(1.0, abs(r64[i]["objective"]))
    B64.close()
    B32.close()


def test_batch_reference_stop_mode(gpu, oracle):
    """Reference stop rule (scaled ||rp||, ||rd||, mu < 1e-6; delta 1e-9; refinement):
    Optimal like the oracle's reference mode, iteration counts within one, objectives within
    1e-6 relative."""
    lps = _small_batch(count=8, seed=17)
    B = gpu.Batch(lps, stop_rule=0)
    B.solve()
    for i, (lp, r) in enumerate(zip(lps, B.results())):
        o = oracle.ipm_solve(lp.csr(), lp.b, lp.c, lp.n, stop_rule=0)
        assert o["status"] == 0 and r["status_name"] == "Optimal", (i, r["fail_code"])
        assert r["primal_residual"] < 1e-6 and r["dual_residual"] < 1e-6 and r["mu"] < 1e-6
        assert abs(r["iterations"] - o["iterations"]) <= 1, (i, r["iterations"], o["iterations"])
        assert abs(r["objective"] - o["objective"]) <= 1e-6 * max(1.0, abs(o["objective"]))
    B.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_members_of_the_4096_family_match_highs(gpu, precision):
    """Members of the config-4 batch (indices spread over 0..4095 plus the largest and smallest
    m) solved together: certified at 1e-8 and within 1e-8 of the committed HiGHS objectives."""
    from lakshya import gen
    with open(os.path.join(HERE, "golden", "batch_objectives.json")) as f:
        gold = json.load(f)
    idx = [int(k) for k in gold]
    lps = gen.batch(indices=idx)
    for i, lp in zip(idx, lps):
        assert lp.digest() == gold[str(i)]["digest"], i
    B = gpu.Batch(lps, precision=precision)
    s = B.solve()
    assert s["optimal"] == len(lps)
    for k, (i, lp, r) in enumerate(zip(idx, lps, B.results())):
        _certified(gpu, lp, B, k, r)
        g = gold[str(i)]["objective"]
        assert abs(r["objective"] - g) <= 1e-8 * max(1.0, abs(g)), (i, r["objective"], g)
    B.close()

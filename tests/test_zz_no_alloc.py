"""The no-allocation batch entry point (lk_batch_solve_noalloc) against the allocating one
(lk_batch_create + lk_batch_solve): bit-identical results on two calls in a row, the second
with a smaller batch, and a clear rejection of a batch beyond the fixed scratch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(count, seed, lo=50, hi=200):
    from lakshya import gen
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=count)
    return [gen.batch_lp(int(m), seed * 1000 + i) for i, m in enumerate(sizes)]


def _allocating(native, lps, **opts):
    B = native.Batch(lps, **opts)
    B.solve()
    res = B.results()
    xs, ys, ss = zip(*[B.solution(i) for i in range(len(lps))])
    B.close()
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ss), res


def _same(native, lps, **opts):
    rc, x, y, s, res, summ = native.batch_solve_noalloc(lps, **opts)
    assert rc == 0, native.last_error()
    assert summ["count"] == len(lps) and summ["optimal"] == len(lps)
    xa, ya, sa, ra = _allocating(native, lps, **opts)
    assert np.array_equal(x, xa) and np.array_equal(y, ya) and np.array_equal(s, sa)
    for r, q in zip(res, ra):
        assert (r["status"], r["iterations"], r["objective"], r["rel_gap"]) == \
               (q["status"], q["iterations"], q["objective"], q["rel_gap"])


def test_noalloc_matches_allocating_entry_point(gpu):
    _same(gpu, _batch(12, 3))                # first call on the device
    _same(gpu, _batch(5, 4, hi=120))         # second call, smaller batch: stale scratch contents
    _same(gpu, _batch(6, 5), precision=1)    # mixed precision through the same scratch


def test_noalloc_rejects_a_batch_beyond_the_scratch(gpu):
    from lakshya import gen
    lps = [gen.batch_lp(1500, 7000 + i) for i in range(5)]   # dense factors alone ~ 90 MB
    rc, *_ = gpu.batch_solve_noalloc(lps)
    assert rc == 1
    assert "scratch" in gpu.last_error()

"""The device factorization and substitution schedules, executed on the CPU.

tests/plan_emulator.py runs the exported task lists with the semantics of the DIAG / STEP /
UPDATE / DIAGWB and substitution kernels; the result must reproduce a dense Cholesky of the same
equilibrated, regularized normal matrix.  This checks the host planner (tiles, staging slots,
levels, K splits of the ancestor updates, ancestor targets) without a GPU."""
import numpy as np
import pytest

from plan_emulator import Emulator


def _cases():
    from lakshya import gen
    return {
        "random_sparse": lambda: gen.random_sparse(40, 100, 3, seed=1),
        "block_angular": lambda: gen.block_angular(4, 60, 150, 5, 5, 0.2, seed=3),
        "transport": lambda: gen.transport(8, 8, 6, 3, seed=2),
        "dense": lambda: gen.dense_small(420, 900, 0.02, seed=5),
    }


@pytest.mark.parametrize("name", ["random_sparse", "block_angular", "transport", "dense"])
def test_schedule_reproduces_dense_cholesky(native, name):
    lp = _cases()[name]()
    rp, ci, va = lp.csr()
    m, n = lp.m, lp.n
    rng = np.random.default_rng(7)
    theta = np.exp(rng.uniform(-2, 2, n))
    P = native.order_amd(rp, ci, n)
    plan = native.plan_export(rp, ci, va, n, P=P)
    kinds = set(plan["flaunch"].reshape(-1, 3)[:, 0].tolist())
    assert kinds <= {1, 2, 3, 5, 6}
    if name == "dense":
        assert 6 in kinds                 # wide supernodes with few tiles split K
    em = Emulator(plan)
    D = em.assemble(theta, equil=True, delta=1e-9)
    assert em.factor(), f"emulated factorization failed (err {em.err})"
    L = em.dense_L()

    A = np.zeros((m, n))
    for i in range(m):
        A[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    perm = plan["perm"]
    M = (A * theta) @ A.T
    Mp = M[np.ix_(perm, perm)]
    Meq = D[:, None] * Mp * D[None, :] + 1e-9 * np.diag((plan["empty_row"] == 0).astype(float))
    scale = np.abs(Meq).max()
    assert np.abs(L @ L.T - Meq).max() <= 1e-12 * scale

    x = rng.standard_normal(m)
    got = em.solve(x.copy())
    exp = np.linalg.solve(L @ L.T, x)
    assert np.abs(got - exp).max() <= 1e-9 * max(1.0, np.abs(exp).max())


@pytest.mark.parametrize("name", ["block_angular", "dense"])
def test_pair_map_length_classes(native, name):
    """Every M0 segment is summed by exactly one class: short (<= 4 pairs, one thread), mid
    (<= 64, listed, 16 lanes) or long (listed, one block); the diagonal lists follow dseg."""
    lp = _cases()[name]()
    rp, ci, va = lp.csr()
    plan = native.plan_export(rp, ci, va, lp.n, P=native.order_amd(rp, ci, lp.n))
    L = np.diff(plan["segptr"])
    assert L.min() >= 1 and plan["pp"].shape == plan["pk"].shape == (plan["segptr"][-1],)
    np.testing.assert_array_equal(plan["mid_segs"], np.flatnonzero((L > 4) & (L <= 64)))
    np.testing.assert_array_equal(plan["long_segs"], np.flatnonzero(L > 64))
    dseg = plan["dseg"]
    Ld = np.where(dseg >= 0, L[np.maximum(dseg, 0)], 0)
    np.testing.assert_array_equal(plan["mid_diag"], np.flatnonzero((Ld > 4) & (Ld <= 64)))
    np.testing.assert_array_equal(plan["long_diag"], np.flatnonzero(Ld > 64))
    cs = plan["chunk_seg"]                # panel chunks of 4096 entries and their segments
    nchunk = -(-int(plan["sn_panel"][-1]) // 4096)
    np.testing.assert_array_equal(cs, np.searchsorted(plan["segdst"], np.arange(nchunk + 1) * 4096))
    assert cs[-1] == len(L) and np.all(np.diff(plan["segdst"]) > 0)
    if name == "block_angular":           # the linking rows' diagonals are long
        assert len(plan["mid_segs"]) and len(plan["long_segs"]) and len(plan["long_diag"])

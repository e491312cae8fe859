"""KKT solver on the device vs the reference's own KATs (tests/unit/test_gpu_kkt.cu) and the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# tests/unit/test_gpu_kkt.cu:21-46 -- A (3x5), Theta = 1, M = [[6,1,1],[1,5,0],[1,0,5]]
RP = np.array([0, 3, 5, 7], np.int32)
CI = np.array([0, 3, 4, 1, 3, 2, 4], np.int32)
VA = np.array([2.0, 1.0, 1.0, 2.0, 1.0, 2.0, 1.0])
M = np.array([[6.0, 1, 1], [1, 5, 0], [1, 0, 5]])


def test_kat_factor_and_solve(gpu):
    """test_gpu_kkt.cu:53-122: P = {1,0,2}, ||L L^T - M_perm||_F < 1e-6, solve -> [1,2,3]."""
    arena = gpu.Arena(1 << 20)
    P = gpu.order_reference(RP, CI, 5)
    assert list(P) == [1, 0, 2]
    sym = gpu.symbolic(RP, CI, 5, P)
    k = gpu.KKT(arena, RP, CI, VA, 5, P=P)
    assert k.factorize(np.ones(5), mode=0)
    nnz = len(sym["Lidx"]) + 3
    Lv = k.L_csr(nnz)
    Lrp = np.array([0, 1, 3, 5]); Lci = np.array([0, 0, 1, 1, 2])   # test_gpu_kkt.cu:66-67
    L = np.zeros((3, 3))
    for i in range(3):
        for p in range(Lrp[i], Lrp[i + 1]):
            L[i, Lci[p]] = Lv[p]
    Mp = M[np.ix_(P, P)]
    assert np.linalg.norm(L @ L.T - Mp) < 1e-6
    x, info = k.solve([11.0, 11.0, 16.0])
    assert np.allclose(x, [1, 2, 3], atol=1e-6)
    assert np.linalg.norm(M @ x - [11, 11, 16]) < 1e-6
    k.close(); arena.close()


def test_kat_relative_residual_contract(gpu):
    """test_gpu_kkt.cu:138-194: rhs x 1e4 solves to 1e-8 relative and the solve succeeds."""
    arena = gpu.Arena(1 << 20)
    P = gpu.order_reference(RP, CI, 5)
    k = gpu.KKT(arena, RP, CI, VA, 5, P=P)
    k.factorize(np.ones(5), mode=0)
    scale = 1e4
    x, info = k.solve(np.array([11.0, 11, 16]) * scale)
    assert info.success == 1
    assert np.allclose(x, scale * np.array([1, 2, 3]), rtol=1e-8)
    r = M @ x - scale * np.array([11, 11, 16])
    assert np.abs(r).max() / max(1.0, 16 * scale) < 1e-8
    k.close(); arena.close()


def _random_kkt(m, n, k, seed):
    from lakshya import gen
    lp = gen.random_sparse(m, n, k, seed=seed)
    rp, ci, va = lp.csr()
    rng = np.random.default_rng(seed)
    theta = np.exp(rng.uniform(-3, 3, n))
    return rp, ci, va, theta


@pytest.mark.parametrize("m,n,k,seed", [(40, 100, 3, 1), (200, 500, 6, 2), (700, 1800, 6, 3)])
def test_factor_matches_dense(gpu, oracle, m, n, k, seed):
    """Unequilibrated factor of M0 + 1e-9 I equals the dense Cholesky of the oracle's M0."""
    rp, ci, va, theta = _random_kkt(m, n, k, seed)
    arena = gpu.Arena(256 << 20)
    P = gpu.order_amd(rp, ci, n)
    kk = gpu.KKT(arena, rp, ci, va, n, P=P)
    assert kk.factorize(theta, mode=0)
    sym = gpu.symbolic(rp, ci, n, P)
    Lv = kk.L_csr(len(sym["Lidx"]) + m)
    L = np.zeros((m, m))
    p = 0
    for i in range(m):
        for q in range(sym["Lptr"][i], sym["Lptr"][i + 1]):
            L[i, sym["Lidx"][q]] = Lv[p]; p += 1
        L[i, i] = Lv[p]; p += 1
    M0 = oracle.assemble_dense(rp, ci, va, theta, P)
    Mr = M0 + 1e-9 * np.eye(m)
    Lref = np.linalg.cholesky(Mr)
    assert np.linalg.norm(L @ L.T - Mr) <= 1e-12 * np.linalg.norm(Mr)
    assert np.abs(L - Lref).max() <= 1e-9 * np.abs(Lref).max()
    kk.close(); arena.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_equilibrated_solve(gpu, oracle, mode):
    rp, ci, va, theta = _random_kkt(300, 800, 6, 7)
    m, n = 300, 800
    arena = gpu.Arena(256 << 20)
    kk = gpu.KKT(arena, rp, ci, va, n)
    assert kk.factorize(theta, mode=mode)
    M0 = oracle.assemble_dense(rp, ci, va, theta, np.arange(m, dtype=np.int32))
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal(m)
    x, info = kk.solve(rhs, tol=1e-9, max_refine=10 if mode == 1 else 60)
    assert info.success == 1, (info.residual_inf, info.refinement_steps)
    assert np.abs(M0 @ x - rhs).max() <= 1e-9
    kk.close(); arena.close()


def _structured_cases():
    from lakshya import gen
    return [gen.block_angular(4, 150, 400, 8, 10, 0.05, seed=0),
            gen.block_angular(2, 300, 800, 8, 20, 0.05, seed=1),
            gen.transport(12, 12, 30, 4, seed=4),
            gen.block_angular(3, 200, 500, 6, 60, 0.1, seed=2)]


@pytest.mark.parametrize("idx", range(4))
@pytest.mark.parametrize("mode", [1, 2])
def test_direct_solve_multi_supernode(gpu, oracle, idx, mode):
    """Supernodal triangular solves on multi-level structures (ancestor rows, several blocks)."""
    lp = _structured_cases()[idx]
    rp, ci, va = lp.csr()
    m, n = lp.m, lp.n
    theta = np.exp(np.random.default_rng(idx).uniform(-2, 2, n))
    arena = gpu.Arena(512 << 20)
    kk = gpu.KKT(arena, rp, ci, va, n, reference_symbolic=False)
    st = kk.stats()
    assert st["nsuper"] > 1
    assert kk.factorize(theta, mode=mode)
    M0 = oracle.assemble_dense(rp, ci, va, theta, np.arange(m, dtype=np.int32))
    rhs = np.random.default_rng(5).standard_normal(m)
    x, info = kk.solve(rhs, tol=1e-300, max_refine=0)
    rel = np.abs(M0 @ x - rhs).max() / np.abs(rhs).max()
    # the factor is of D M0 D + 1e-9 I: the direct solve is accurate to ~1e-9 * cond(D M0 D)
    assert rel <= (1e-7 if mode == 1 else 1e-4), rel
    kk.close(); arena.close()


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("blocks,mb,nb", [(4, 1500, 4000), (12, 1100, 2600)])
def test_direct_solve_wide_supernodes(gpu, mode, blocks, mb, nb):
    """Wide supernodes (17 to 26 row tiles each) and the linking rows below them.  Four of
    them: both level solves take the wide form (a workgroup per row tile / block column,
    ancestor rows added with atomics forward and read backward); twelve: the forward is the
    grouped level solve (four workgroups per supernode), the backward the wide form.  Residual
    of the direct solve with sparse products, M0 x = A (theta o (A^T x))."""
    import scipy.sparse as sp
    from lakshya import gen
    lp = gen.block_angular(blocks, mb, nb, 8, 50, 0.05, seed=0)
    rp, ci, va = lp.csr()
    kinds = set(gpu.plan_export(rp, ci, va, lp.n)["slaunch"].reshape(-1, 3)[:, 0].tolist())
    assert 23 in kinds and (22 in kinds) == (blocks <= 8)
    m, n = lp.m, lp.n
    A = sp.csr_matrix((va, ci, rp), shape=(m, n))
    theta = np.exp(np.random.default_rng(3).uniform(-2, 2, n))
    arena = gpu.Arena(2 << 30)
    kk = gpu.KKT(arena, rp, ci, va, n, reference_symbolic=False)
    assert kk.factorize(theta, mode=mode)
    rhs = np.random.default_rng(6).standard_normal(m)
    for _ in range(3):                    # repeated solves: flags and queues of every solve
        x, info = kk.solve(rhs, tol=1e-300, max_refine=0)
        rel = np.abs(A @ (theta * (A.T @ x)) - rhs).max() / np.abs(rhs).max()
        assert rel <= (1e-7 if mode == 1 else 1e-4), rel
    kk.close(); arena.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_direct_solve_dense_separator(gpu, oracle, mode):
    """One dense supernode of 1,200 columns (19 blocks, no rows below): the level solves take
    the wide form, a workgroup per row tile (forward) or block column (backward) on every XCD,
    solved blocks handed over through device-coherent flags.  Residual against the dense M0,
    repeated solves (new epochs)."""
    from lakshya import gen
    lp = gen.dense_small(1200, 2400, 0.02, seed=8)
    rp, ci, va = lp.csr()
    m, n = lp.m, lp.n
    plan = gpu.plan_export(rp, ci, va, n)
    assert {22, 23} <= set(plan["slaunch"].reshape(-1, 3)[:, 0].tolist())
    theta = np.exp(np.random.default_rng(4).uniform(-2, 2, n))
    arena = gpu.Arena(512 << 20)
    kk = gpu.KKT(arena, rp, ci, va, n, reference_symbolic=False)
    assert kk.factorize(theta, mode=mode)
    M0 = oracle.assemble_dense(rp, ci, va, theta, np.arange(m, dtype=np.int32))
    for seed in range(3):
        rhs = np.random.default_rng(10 + seed).standard_normal(m)
        x, info = kk.solve(rhs, tol=1e-300, max_refine=0)
        rel = np.abs(M0 @ x - rhs).max() / np.abs(rhs).max()
        assert rel <= (1e-7 if mode == 1 else 1e-4), (seed, rel)
    kk.close(); arena.close()


def test_graph_capture_refused(gpu):
    """Launch-graph capture is not offered: asking for it is an invalid argument."""
    arena = gpu.Arena(1 << 20)
    with pytest.raises(ValueError):
        gpu.KKT(arena, RP, CI, VA, 5, use_graphs=True)
    arena.close()

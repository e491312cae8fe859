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
        "block_angular_wide": lambda: gen.block_angular(3, 150, 400, 6, 8, 0.2, seed=3),
        "dense_wide": lambda: gen.dense_small(700, 1400, 0.02, seed=6),
        "dense_separator": lambda: gen.dense_small(1100, 2200, 0.02, seed=8),
        "block_angular_2x1100": lambda: gen.block_angular(2, 1100, 2600, 6, 8, 0.05, seed=4),
    }


@pytest.mark.parametrize("name,steps", [(c, 1) for c in ("random_sparse", "block_angular", "transport",
                                                         "dense", "block_angular_wide", "dense_wide")] +
                         [("dense_wide", 0), ("dense_wide", 2), ("block_angular_wide", 2),
                          ("dense_separator", 1), ("block_angular_2x1100", 1),
                          ("dense_separator", 2), ("block_angular_2x1100", 2)])
def test_schedule_reproduces_dense_cholesky(native, name, steps):
    lp = _cases()[name]()
    rp, ci, va = lp.csr()
    m, n = lp.m, lp.n
    rng = np.random.default_rng(7)
    theta = np.exp(rng.uniform(-2, 2, n))
    P = native.order_amd(rp, ci, n)
    plan = native.plan_export(rp, ci, va, n, P=P, step_dag=steps)
    kinds = set(plan["flaunch"].reshape(-1, 3)[:, 0].tolist())
    assert kinds <= {1, 2, 3, 5, 6, 7, 8}  # STEPS launches, or PART / PSUM / STEP launches per step
    em = Emulator(plan)
    D = em.assemble(theta, equil=True, delta=1e-9)
    assert em.factor(), f"emulated factorization failed (err {em.err})"
    L = em.dense_L()
    for seed in range(3):                 # any order the STEPS dependencies allow
        em2 = Emulator(plan)
        em2.assemble(theta, equil=True, delta=1e-9)
        assert em2.factor(np.random.default_rng(seed))
        assert np.abs(em2.dense_L() - L).max() <= 1e-12 * np.abs(L).max()

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


def _tile_phases(t):
    """Regions a STEPS tile touches, per phase: P (GEMM operands), S (solve, store, diagonal
    blocks of a tile that does not factor), B (a factoring tile: its diagonal block, factor,
    staging).  Region: (space, key, lo0, hi0, lo1, hi1, writes)."""
    s, k0, kb, r0, ra = (int(t[x]) for x in ("sn", "k0", "kb", "r0", "c0"))
    nc, first, nsplit = int(t["ncol"]), int(t["first"]), int(t["nsplit"])
    kbeg, kend, px = int(t["kbeg"]), int(t["kend"]), int(t["px"])
    P = [("panel", s, kbeg, kend, r0, r0 + ra, False),         # L(R, kbeg:kend)
         ("panel", s, kbeg, kend, k0, k0 + kb, False)]         # L(k, kbeg:kend)
    if int(t["part"]):                                         # PART: product into work
        return [P + [("work", 0, px, px + 128 * 64, 0, 1, True)]]
    if int(t["np"]):                                           # its PART's product
        P.append(("work", 0, px, px + 128 * 64, 0, 1, False))
    S = [("panel", s, k0, k0 + kb, r0, r0 + ra, True),         # C(R, k) -> L(R, k)
         ("stage", int(t["ws"]), 0, 1, 0, 1, False)]
    D = []
    for h in range(2):                                         # diagonal blocks in the tile
        d0 = r0 + 64 * h
        if 64 * h >= ra or d0 >= nc:
            break
        db = min(64, ra - 64 * h, nc - d0)
        D.append(("panel", s, d0, d0 + db, d0, d0 + db, True))
    if not nsplit:
        return [P, S + D]
    B = D + [("stage", int(t["aux2"]), 0, 1, 0, 1, True),
             ("diag", 0, int(t["aux"]), int(t["aux"]) + ra * ra, 0, 1, True),
             ("piv", 0, first + r0, first + r0 + ra, 0, 1, True)]
    return [P, S, B]


@pytest.mark.parametrize("name", ["random_sparse", "dense", "block_angular_wide", "dense_wide",
                                  "block_angular_2x1100"])
def test_steps_dependencies_order_every_conflict(native, name):
    """Inside a STEPS launch, two phases of tiles that touch the same data with at least one
    write must be ordered: by program order inside a tile, or through the hand-offs (a
    dependency's rows flag before the GEMM, the staging flag before the solve)."""
    lp = _cases()[name]()
    rp, ci, va = lp.csr()
    # every level in a STEPS launch (the solver's own choice keeps a dense separator's level
    # in launches per block column)
    plan = native.plan_export(rp, ci, va, lp.n, P=native.order_amd(rp, ci, lp.n), step_dag=2)
    T, deps, sdep = plan["ftasks"], plan["fdeps"].reshape(-1, 4), plan["fsdep"]
    nsteps = 0
    for kind, b, c in plan["flaunch"].reshape(-1, 3):
        if kind != 7:
            continue
        nsteps += c
        ph = [_tile_phases(T[b + i]) for i in range(c)]
        node = {}                                   # (task, phase) -> node id
        for i in range(c):
            for p_ in range(len(ph[i])):
                node[(i, p_)] = len(node)
        anc = [0] * len(node)                       # nodes in task order are topological
        for i in range(c):
            for p_ in range(len(ph[i])):
                n_ = node[(i, p_)]
                a = 0
                if p_ > 0:                          # program order
                    a |= anc[node[(i, p_ - 1)]] | (1 << node[(i, p_ - 1)])
                if p_ == 0:                         # flags[j]: after j's S phase (a PART: its only one)
                    for j in deps[b + i]:
                        if j >= 0:
                            assert j < i
                            m = node[(int(j), min(1, len(ph[int(j)]) - 1))]
                            a |= anc[m] | (1 << m)
                if p_ == 1 and sdep[b + i] >= 0:    # flags2[j]: after j's B phase
                    j = int(sdep[b + i])
                    assert j < i and len(ph[j]) == 3
                    m = node[(j, 2)]
                    a |= anc[m] | (1 << m)
                anc[n_] = a
        items = [(node[(i, p_)], i, reg) for i in range(c) for p_ in range(len(ph[i])) for reg in ph[i][p_]]
        by_key = {}
        for it in items:
            by_key.setdefault((it[2][0], it[2][1]), []).append(it)
        for lst in by_key.values():
            for u in range(len(lst)):
                for v in range(u + 1, len(lst)):
                    (nx, i, x), (ny, j, y) = lst[u], lst[v]
                    if i == j or not (x[6] or y[6]):
                        continue
                    if not (x[2] < y[3] and y[2] < x[3] and x[4] < y[5] and y[4] < x[5]):
                        continue
                    assert (anc[ny] >> nx) & 1 or (anc[nx] >> ny) & 1, (name, b, i, j, x, y)
    assert nsteps > 0


def _level_kinds(plan):
    """Launch kinds of the factor plan between consecutive DIAG launches (one group per level)."""
    groups, cur = [], []
    for kind, b, c in plan["flaunch"].reshape(-1, 3):
        if kind == 1 and cur:
            groups.append(cur)
            cur = []
        cur.append((int(kind), int(c)))
    groups.append(cur)
    return groups


def test_steps_launch_only_where_the_xcds_share_the_work(native):
    """A STEPS launch keeps each supernode on one XCD: 16 equal diagonal blocks (two per XCD)
    are factored in one STEPS launch; one dense supernode, which would occupy a single XCD,
    gets a PART / STEP launch pair per block column (its tiles spread over every XCD)."""
    from lakshya import gen
    lp = gen.block_angular(16, 150, 400, 6, 8, 0.2, seed=3)
    rp, ci, va = lp.csr()
    plan = native.plan_export(rp, ci, va, lp.n, P=native.order_amd(rp, ci, lp.n))
    sf, rpt = plan["sn_first"], plan["sn_rowptr"]
    assert (np.diff(sf) >= 100).sum() == 16
    lv0 = _level_kinds(plan)[0]
    assert [k for k, _ in lv0 if k in (5, 6, 7)] == [7]
    lp = gen.dense_small(700, 1400, 0.02, seed=6)
    rp, ci, va = lp.csr()
    plan = native.plan_export(rp, ci, va, lp.n, P=native.order_amd(rp, ci, lp.n))
    ncol = np.diff(plan["sn_first"])
    assert ncol.max() >= 0.9 * lp.m                      # one dense supernode
    kinds = [k for g in _level_kinds(plan) for k, _ in g]
    assert 7 not in kinds and kinds.count(5) >= ncol.max() // 64 - 1


@pytest.mark.parametrize("name", ["dense_separator", "block_angular_2x1100"])
def test_wide_level_solves(native, name):
    """Levels of at most 8 supernodes, one of at least 16 blocks, take the wide solve form
    (kinds 22 / 23): a task per row tile / block column in an order where every task's
    producers come first (the plan checks it), flags of 2 nb per supernode."""
    lp = _cases()[name]()
    rp, ci, va = lp.csr()
    plan = native.plan_export(rp, ci, va, lp.n, P=native.order_amd(rp, ci, lp.n))
    sl = plan["slaunch"].reshape(-1, 3)
    kinds = [int(k) for k in sl[:, 0]]
    assert 22 in kinds and 23 in kinds
    T = plan["stasks"]
    for kind, b, c in sl:
        if kind < 22:
            continue
        tl = T[b:b + c]
        pos = {(int(t["sn"]), int(t["r0"])): i for i, t in enumerate(tl)}
        for i, t in enumerate(tl):
            nb = (int(t["ncol"]) + 63) // 64
            deps = range(min(int(t["r0"]), nb)) if kind == 22 else range(int(t["r0"]) + 1, nb)
            assert all(pos[(int(t["sn"]), j)] < i for j in deps)

"""Regression tests of fixed host-side defects and the reference (absolute 1e-6) stop mode."""
import json
import os
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _csr_case(m, n, k, seed):
    from lakshya import gen
    lp = gen.random_sparse(m, n, k, seed=seed)
    rp, ci, va = lp.csr()
    theta = np.exp(np.random.default_rng(seed).uniform(-2, 2, n))
    return lp, rp, ci, va, theta


def test_arena_zero_fill_ordered_with_solver_streams(gpu, oracle):
    """The arena zero-fills on allocation.  A fill queued on the null stream is not ordered with
    the solvers' non-blocking streams, so it could land after a solver's first upload (a host
    Theta or rhs) and wipe it.  Here a second solver is created, factored from a host Theta and
    solved right away while another solver keeps the device busy; every solve must match the
    dense solve of M0."""
    from lakshya import gen
    busy = gen.block_angular(4, 1500, 4000, 8, 50, 0.05, seed=0)
    brp, bci, bva = busy.csr()
    btheta = np.exp(np.random.default_rng(1).uniform(-2, 2, busy.n))
    barena = gpu.Arena(2 << 30)
    bk = gpu.KKT(barena, brp, bci, bva, busy.n, reference_symbolic=False)
    stop = threading.Event()
    errors = []

    def keep_busy():
        try:
            rhs = np.ones(busy.m)
            while not stop.is_set():
                bk.factorize(btheta, mode=1)
                bk.solve(rhs, tol=1e-300, max_refine=0)
        except Exception as e:   # reported by the main thread
            errors.append(e)

    t = threading.Thread(target=keep_busy)
    t.start()
    try:
        for seed in (3, 4, 5):
            lp, rp, ci, va, theta = _csr_case(300, 800, 6, seed)
            arena = gpu.Arena(256 << 20)
            kk = gpu.KKT(arena, rp, ci, va, lp.n, reference_symbolic=False)
            assert kk.factorize(theta, mode=1)
            M0 = oracle.assemble_dense(rp, ci, va, theta, np.arange(lp.m, dtype=np.int32))
            rhs = np.random.default_rng(seed).standard_normal(lp.m)
            x, info = kk.solve(rhs, tol=1e-9, max_refine=10)
            assert info.success == 1, (seed, info.residual_inf)
            assert np.abs(M0 @ x - rhs).max() <= 1e-9, seed
            kk.close()
            arena.close()
    finally:
        stop.set()
        t.join()
        bk.close()
        barena.close()
    assert not errors, errors


def test_arena_handle_destroyed_before_kkt(gpu):
    """Destroying the arena handle first (an order a garbage collector may pick) keeps the
    KKT solver usable: it shares ownership of the slab it allocated from."""
    # test_gpu_kkt.cu:21-46: A (3x5), Theta = 1, M = [[6,1,1],[1,5,0],[1,0,5]]
    rp = np.array([0, 3, 5, 7], np.int32)
    ci = np.array([0, 3, 4, 1, 3, 2, 4], np.int32)
    va = np.array([2.0, 1.0, 1.0, 2.0, 1.0, 2.0, 1.0])
    arena = gpu.Arena(1 << 20)
    kk = gpu.KKT(arena, rp, ci, va, 5)
    arena.close()
    assert kk.factorize(np.ones(5), mode=0)
    x, info = kk.solve([11.0, 11.0, 16.0])
    assert info.success == 1
    assert np.allclose(x, [1, 2, 3], atol=1e-6)
    kk.close()


def _golden(name):
    with open(os.path.join(HERE, "golden", "objectives.json")) as f:
        return json.load(f)[name]["objective"]


@pytest.mark.parametrize("name", ["rs_200_500_6_s2", "ba_4x60x150_L5_s3", "tr_12x12_side30_s4"])
def test_reference_stop_mode_matches_oracle(gpu, oracle, name):
    """Default (drop-in) mode: Ruiz scaling, delta = 1e-9 factors with refinement against M0,
    and the absolute 1e-6 stop on the scaled ||rp||_2, ||rd||_2 and mu
    (src/ipm/mehrotra_device.cu:286-323,746-748).  Device and oracle run the same algorithm:
    same status, iteration counts within one, both stop inside the rule, objectives agree to
    1e-6 relative and sit within n * 1e-6 of the HiGHS optimum."""
    from golden_cases import build_case
    lp = build_case(name)
    g = gpu.ipm_solve(lp)
    o = oracle.ipm_solve(lp.csr(), lp.b, lp.c, lp.n, stop_rule=0)
    assert o["status"] == 0 and g["status_name"] == "Optimal", g["fail_code"]
    assert abs(g["iterations"] - o["iterations"]) <= 1, (g["iterations"], o["iterations"])
    assert g["primal_residual"] < 1e-6 and g["dual_residual"] < 1e-6 and g["duality_gap"] < 1e-6
    scale = max(1.0, abs(o["objective"]))
    assert abs(g["objective"] - o["objective"]) <= 1e-6 * scale, (g["objective"], o["objective"])
    gold = _golden(name)
    assert abs(g["objective"] - gold) <= lp.n * 1e-6 * max(1.0, abs(gold))

"""MehrotraSolver on the device: reference KATs, oracle parity, HiGHS golden objectives."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_textbook_lp(gpu):
    """tests/unit/test_mehrotra.cu:29-78: min -3x1-5x2, x1+x2+x3=4, 2x1+3x2+x4=12 -> -20."""
    from conftest import lp_from_rows
    lp = lp_from_rows([[1, 1, 1, 0], [2, 3, 0, 1]], [4, 12], [-3, -5, 0, 0], 4)
    r = gpu.ipm_solve(lp)
    assert r["status_name"] == "Optimal"
    assert r["iterations"] > 0
    assert r["primal_residual"] < 1e-6 and r["dual_residual"] < 1e-6 and r["duality_gap"] < 1e-6
    assert abs(r["objective"] - (-20.0)) <= 1e-4
    assert r["x"].shape == (4,) and r["y"].shape == (2,)
    assert np.isfinite(r["x"]).all() and np.isfinite(r["y"]).all()
    r = gpu.ipm_solve(lp, relative=True)
    assert abs(r["objective"] + 20.0) <= 1e-8 * 20


@pytest.mark.parametrize("m,n,k,seed", [(60, 150, 3, 11), (200, 500, 6, 12), (400, 1000, 6, 13)])
def test_matches_oracle(gpu, oracle, m, n, k, seed):
    from lakshya import gen
    lp = gen.random_sparse(m, n, k, seed=seed)
    g = gpu.ipm_solve(lp, relative=True)
    o = oracle.ipm_solve(lp.csr(), lp.b, lp.c, lp.n, stop_rule=1)
    assert g["status_name"] == "Optimal" and o["status"] == 0
    assert abs(g["objective"] - o["objective"]) <= 1e-8 * max(1.0, abs(o["objective"]))
    assert g["rel_gap"] <= 1e-8 and g["rel_primal_residual"] <= 1e-8 and g["rel_dual_residual"] <= 1e-8
    ok, _ = gpu.verify_optimal(lp, g["x"], g["y"], tol=1e-6)
    assert ok


@pytest.mark.parametrize("case", ["linking_rows", "dense_columns", "linking_rows_split",
                                  "dense_columns_split"])
def test_long_rows_and_columns_match_oracle(gpu, oracle, case):
    """Rows / columns above the whole-block threshold (512 entries) and M0 entries summed from more
    than 64 column pairs go through the block-per-row / block-per-segment paths; above 1,024
    entries the row passes sum them in pieces (pre-pass)."""
    from lakshya import gen
    if case == "linking_rows":
        lp = gen.block_angular(4, 60, 200, 4, 5, 0.9, seed=7)      # 5 linking rows of ~720 entries
        assert np.diff(lp.csr()[0]).max() > 512
    elif case == "dense_columns":
        lp = gen.dense_columns(600, 1500, 4, 3, seed=8)            # 3 columns of 600 entries
        assert np.diff(lp.indptr).max() > 512
    elif case == "linking_rows_split":
        lp = gen.block_angular(4, 60, 800, 4, 5, 0.9, seed=9)      # 5 linking rows of ~2,900 entries
        assert np.diff(lp.csr()[0]).max() > 2048
    else:
        lp = gen.dense_columns(1100, 2200, 4, 3, seed=10)          # 3 columns of 1,100 entries
        assert np.diff(lp.indptr).max() > 1024
    g = gpu.ipm_solve(lp, relative=True)
    o = oracle.ipm_solve(lp.csr(), lp.b, lp.c, lp.n, stop_rule=1)
    assert g["status_name"] == "Optimal" and o["status"] == 0
    assert abs(g["objective"] - o["objective"]) <= 1e-8 * max(1.0, abs(o["objective"]))
    ok, _ = gpu.verify_optimal(lp, g["x"], g["y"], tol=1e-8)
    assert ok


GOLDEN_GPU = ["rs_60_150_3_s1", "rs_200_500_6_s2", "ba_4x60x150_L5_s3", "tr_12x12_side30_s4",
              "dense_80_s5", "tr_100x100_side2000_s0", "rs_2000_5000_6_s0",
              "ba_8x1500x4000_L50_s0", "ba_64x1500x4000_L50_s0", "tr_700x700_side20000_s0"]


def _golden(name):
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "objectives.json")) as f:
        return json.load(f)[name]["objective"]


def _check_certified(gpu, lp, r, gold):
    assert r["status_name"] == "Optimal", r.get("fail_code")
    assert abs(r["objective"] - gold) <= 1e-8 * max(1.0, abs(gold)), (r["objective"], gold)
    assert r["rel_gap"] <= 1e-8 and r["rel_primal_residual"] <= 1e-8 and r["rel_dual_residual"] <= 1e-8
    ok, _ = gpu.verify_optimal(lp, r["x"], r["y"], tol=1e-8)
    assert ok


@pytest.mark.parametrize("name", GOLDEN_GPU)
def test_golden_objective_fp64(gpu, name):
    """The plain FP64 device IPM on the seeded instances whose HiGHS objectives are committed
    (rs_2000_5000: the 2,000 x 5,000 random LP; ba_8 / ba_64: the block-angular LP with 8 / 64
    blocks of 1,500 x 4,000; tr_700x700: the 700 x 700 transport LP with 20,000 side rows)."""
    from golden_cases import build_case
    lp = build_case(name)
    _check_certified(gpu, lp, gpu.ipm_solve(lp, relative=True), _golden(name))


@pytest.mark.parametrize("name", ["rs_200_500_6_s2", "tr_100x100_side2000_s0", "rs_2000_5000_6_s0",
                                  "ba_8x1500x4000_L50_s0", "tr_700x700_side20000_s0"])
def test_golden_objective_mixed(gpu, name):
    """FP32 factorization with FP64 PCG / refinement reaches the same certified optimum."""
    from golden_cases import build_case
    lp = build_case(name)
    r = gpu.ipm_solve(lp, relative=True, precision=1)
    _check_certified(gpu, lp, r, _golden(name))
    assert r["fp32_factorizations"] > 0

"""Named seeded instances shared by the golden-objective fixtures and the parity tests."""
from lakshya import gen

CASES = {
    "rs_60_150_3_s1": lambda: gen.random_sparse(60, 150, 3, seed=1),
    "rs_200_500_6_s2": lambda: gen.random_sparse(200, 500, 6, seed=2),
    "ba_4x60x150_L5_s3": lambda: gen.block_angular(4, 60, 150, 4, 5, 0.2, seed=3),
    "tr_12x12_side30_s4": lambda: gen.transport(12, 12, 30, 4, seed=4),
    "dense_80_s5": lambda: gen.dense_small(80, seed=5),
    "rs_2000_5000_6_s0": lambda: gen.random_sparse(2000, 5000, 6, seed=0),
    "ba_64x1500x4000_L50_s0": lambda: gen.block_angular(seed=0),
    "ba_8x1500x4000_L50_s0": lambda: gen.block_angular(8, seed=0),
    "tr_700x700_side20000_s0": lambda: gen.transport(seed=0),
    "tr_100x100_side2000_s0": lambda: gen.transport(100, 100, 2000, seed=0),
    "rs_20000_50000_8_s0": lambda: gen.parametric_base(seed=0),
}


def build_case(name):
    return CASES[name]()

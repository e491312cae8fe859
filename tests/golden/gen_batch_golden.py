"""Writes tests/golden/batch_objectives.json: HiGHS (scipy.optimize.linprog, method="highs")
optimal objectives of a sample of the 4,096-LP batch family (lakshya.gen.batch, seed 0: LP i
has m_i ~ U{50..800}, n_i = round(2.5 m_i), Bernoulli(0.05) N(0,1) entries), with digests.

Usage: python tests/golden/gen_batch_golden.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "soverign-gpu-accelerator-_amd"))

from lakshya import gen  # noqa: E402
from scipy.optimize import linprog  # noqa: E402

OUT = os.path.join(HERE, "batch_objectives.json")
# members of the full batch: evenly spread indices plus the largest and smallest m
SAMPLE = list(range(0, 4096, 256)) + [4095]


def main():
    sizes = gen.batch_sizes()
    idx = SAMPLE + [int(sizes.argmax()), int(sizes.argmin())]
    out = {}
    for i, lp in zip(idx, gen.batch(indices=idx)):
        t = time.time()
        h = linprog(lp.c, A_eq=lp.to_scipy(), b_eq=lp.b, bounds=(0, None), method="highs",
                    options=dict(primal_feasibility_tolerance=1e-10,
                                 dual_feasibility_tolerance=1e-10, presolve=True))
        out[str(i)] = {"digest": lp.digest(), "m": lp.m, "n": lp.n, "nnz": lp.nnz,
                       "objective": float(h.fun) if h.status == 0 else None,
                       "highs_status": int(h.status), "highs_seconds": round(time.time() - t, 2)}
        print(i, out[str(i)], flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

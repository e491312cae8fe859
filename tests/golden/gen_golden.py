"""Writes tests/golden/objectives.json: HiGHS (scipy.optimize.linprog, method="highs") optimal
objectives of the named seeded instances in tests/golden_cases.py, with instance digests.

Usage: python tests/golden/gen_golden.py [--method=highs|highs-ipm] [case ...]
       (default: dual simplex; every case not yet recorded)
The LP optimal value is unique, so these pin the interior-point objective to 1e-8 relative even
where the optimal point is not unique.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "soverign-gpu-accelerator-_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from golden_cases import CASES, build_case  # noqa: E402
from scipy.optimize import linprog  # noqa: E402

OUT = os.path.join(HERE, "objectives.json")


def solve(name, method):
    lp = build_case(name)
    t = time.time()
    opts = dict(primal_feasibility_tolerance=1e-10, dual_feasibility_tolerance=1e-10, presolve=True)
    if method == "highs-ipm":
        opts["ipm_optimality_tolerance"] = 1e-12   # HiGHS IPM followed by its crossover
    h = linprog(lp.c, A_eq=lp.to_scipy(), b_eq=lp.b, bounds=(0, None), method=method, options=opts)
    return {"digest": lp.digest(), "m": lp.m, "n": lp.n, "nnz": lp.nnz,
            "objective": float(h.fun) if h.status == 0 else None, "highs_status": int(h.status),
            "highs_message": h.message, "highs_method": method,
            "highs_seconds": round(time.time() - t, 2)}


def main(argv):
    method = "highs"
    if argv and argv[0].startswith("--method="):
        method = argv[0].split("=", 1)[1]
        argv = argv[1:]
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    names = argv or [n for n in CASES if n not in data]
    for n in names:
        print("solving", n, flush=True)
        data[n] = solve(n, method)
        print(" ", data[n], flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])

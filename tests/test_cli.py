"""The command-line contract of the reference's sankhya_benchmark_cli
(tests/benchmark/sankhya_benchmark_cli.cpp:15-100): usage, the parse-error object, and on the
GPU a solve whose last {...} on stdout carries every key app.py / harness.py read."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "soverign-gpu-accelerator-_amd", "bin", "lakshya_benchmark_cli")
KEYS = {"solver", "status", "execution_path", "fallback_triggered", "solve_time_ms", "iterations",
        "objective", "primal_residual", "variables", "constraints", "nnz", "certificate_pass"}

# min -x1 - 2 x2  s.t.  x1 + x2 + s1 = 4,  x1 + 3 x2 + s2 = 6  (optimum -5 at x = (3, 1)), the
# LP of tests/unit/test_mehrotra.cu:29-78 written with L rows
TEXTBOOK = """NAME TEXTBOOK
ROWS
 N  COST
 L  C1
 L  C2
COLUMNS
    X1  COST  -1  C1  1
    X1  C2  1
    X2  COST  -2  C1  1
    X2  C2  3
RHS
    RHS  C1  4  C2  6
ENDATA
"""


def _run(*args):
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout, p.stderr


def _last_object(out):
    return json.loads(out[out.rfind("{"):out.rfind("}") + 1])


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_usage_and_parse_error_object(tmp_path):
    rc, _, err = _run()
    assert rc == 1 and "Usage" in err
    bad = tmp_path / "bad.mps"
    bad.write_text("NAME X\nROWS\n N OBJ\n E R\nRANGES\nENDATA\n")
    rc, out, _ = _run(str(bad))
    assert rc == 0 and "unsupported section" in _last_object(out)["error"]


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [[], ["--relative"]])
def test_solve_reports_the_contract(tmp_path, flags):
    f = tmp_path / "textbook.mps"
    f.write_text(TEXTBOOK)
    rc, out, err = _run(*flags, str(f))
    assert rc == 0, err
    r = _last_object(out)
    assert set(r) == KEYS
    assert r["solver"] == "SANKHYA" and r["status"] == "Optimal"
    assert r["execution_path"] == "GPU_ONLY" and r["fallback_triggered"] == "NO"
    # the reference stop rule ends at mu < 1e-6, where a nonbasic x_j of about 1e-6 can still
    # fail verify_optimal's 1e-6 bound check (the reference CLI reports the same); the 1e-8
    # relative-gap rule certifies
    if flags:
        assert r["certificate_pass"] is True
    assert (r["variables"], r["constraints"], r["nnz"]) == (4, 2, 6)
    assert abs(r["objective"] + 5.0) <= (1e-5 if not flags else 1e-7)
    assert r["iterations"] > 0 and r["solve_time_ms"] > 0


@pytest.mark.gpu
def test_generated_lp_through_mps(tmp_path, native):
    """A seeded 200 x 500 LP written as MPS (E rows): the CLI's relative-gap objective equals the
    library's on the same arrays."""
    from lakshya import gen
    lp = gen.random_sparse(200, 500, 6, seed=2)
    A = lp.to_scipy().tocsc()
    lines = ["NAME GEN", "ROWS", " N  OBJ"] + [f" E  R{i}" for i in range(lp.m)] + ["COLUMNS"]
    for j in range(lp.n):
        lines.append(f"    X{j}  OBJ  {float(lp.c[j])!r}")
        lines += [f"    X{j}  R{A.indices[p]}  {float(A.data[p])!r}" for p in range(A.indptr[j], A.indptr[j + 1])]
    lines += ["RHS"] + [f"    RHS  R{i}  {float(lp.b[i])!r}" for i in range(lp.m)] + ["ENDATA"]
    f = tmp_path / "gen.mps"
    f.write_text("\n".join(lines) + "\n")
    rc, out, err = _run("--relative", str(f))
    assert rc == 0, err
    r = _last_object(out)
    assert r["status"] == "Optimal" and r["certificate_pass"] is True
    ref = native.ipm_solve(lp, relative=True)
    # two certified solves of one LP (the factor's atomics may change the last bits of a path)
    assert abs(r["objective"] - ref["objective"]) <= 2e-8 * max(1.0, abs(ref["objective"]))


# x1 + x2 = 1 and x1 + x2 = 2 (tests/unit/test_phase1.cpp:264-283): no feasible point
INFEASIBLE = """NAME INFEAS
ROWS
 N  COST
 E  R1
 E  R2
COLUMNS
    X1  R1  1  R2  1
    X2  R1  1  R2  1
RHS
    RHS  R1  1  R2  2
ENDATA
"""


@pytest.mark.gpu
def test_iteration_limit_takes_the_simplex_route(tmp_path):
    """The interior-point solve of an infeasible LP ends in IterationLimit (the reference never
    reports Infeasible from the IPM); the fallback then runs the Phase-I / Phase-II simplex on the
    device, whose Phase I finds no feasible point: FailedRecovery, as the reference CLI reports
    a non-Optimal fallback (sankhya_benchmark_cli.cpp:48-75)."""
    f = tmp_path / "infeasible.mps"
    f.write_text(INFEASIBLE)
    rc, out, err = _run(str(f))
    assert rc == 0, err
    r = _last_object(out)
    assert set(r) == KEYS
    assert r["fallback_triggered"] == "YES" and r["execution_path"] == "GPU+GPU_SIMPLEX"
    assert r["status"] == "FailedRecovery" and r["certificate_pass"] is False

"""The reference-side C++ caller (tests/cpp/ref_callers.cpp), compiled against include/ and
linked to liblakshya.so: the class layout on the CPU, the KKT and IPM unit-test flows on the GPU."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_bin", "ref_callers")


def _run(what):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_bin/ref_callers is not built (__graft_entry__.build())")
    p = subprocess.run([BIN, what], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.strip().endswith("OK"), p.stdout
    return p.stdout


def test_model_layout_compiles_and_runs():
    """Model carries bilinear_terms (src/core/problem.hpp:26-49): the reference's MILP and
    McCormick code compiles against this header."""
    _run("model")


@pytest.mark.gpu
def test_kkt_four_argument_constructor_flow(gpu):
    """tests/unit/test_gpu_kkt.cu:21-121 through GPUKKTCholeskySolver(arena, A, sym, M_pattern),
    gpu_cholesky_factorize, get_device_L_values (reference CSR layout) and gpu_cholesky_solve."""
    _run("kkt")


@pytest.mark.gpu
def test_mehrotra_model_solve_flow(gpu):
    """tests/unit/test_mehrotra.cu:29-78 through MehrotraSolver(model).solve(), then the
    reference-signature verify_optimal(model, x, pi, &res)."""
    _run("ipm")


@pytest.mark.gpu
def test_crossover_device_model_flow(gpu):
    """tests/unit/test_crossover.cu:15-133 through upload_to_device and
    generate_crossover_basis(DeviceModel, x): a feasible nonsingular basis from (1, 1, 1, 1), and
    the rank-deficiency message for dependent rows."""
    _run("crossover")


@pytest.mark.gpu
def test_simplex_route_flow(gpu):
    """simplex::solve_with_phase1(model, basis, x) on the models of tests/unit/test_phase1.cpp:
    123-430 (statuses and objectives), a Maximize model, a shifted lower bound and the finite
    upper bound the reference reports as IterationLimit; each optimum certified by the
    reference-signature verify_optimal with pi = B^-T c_B."""
    _run("simplex")

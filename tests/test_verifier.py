"""verify_optimal through the C ABI (host code, no GPU): the reference's certificate cases
(tests/unit/test_certificate.cpp:30-219) at the reference tolerance 1e-6, plus the tolerance
argument this build adds (the IPM acceptance runs it at 1e-8)."""
import numpy as np
import pytest

from conftest import lp_from_rows

INF = np.inf


def single(c, lb, ub, rhs):
    """min c x  s.t.  x = rhs,  lb <= x <= ub  (test_certificate.cpp:14-24)."""
    lp = lp_from_rows([[1.0]], [rhs], [c], 1)
    return lp, np.array([lb], float), np.array([ub], float)


@pytest.mark.parametrize("c,lb,ub,rhs,x,pi,expect", [
    (2.0, 0.0, INF, 0.0, 0.0, 1.0, True),              # case 1: x = lb, r > 0
    (0.0, 0.0, INF, 5.3e-6, 5.3e-6, -0.0416, True),    # case 2: r (x - lb) = 2.2e-7 < eps
    (0.0, 0.0, INF, 2e-5, 2e-5, -1.0, False),          # case 3: r (x - lb) = 2e-5 > eps
    (-2.0, 0.0, 1.0, 1.0, 1.0, -1.0, True),            # case 4: x = ub, r < 0
    (0.0, 0.0, 1.0, 1.0 - 5e-7, 1.0 - 5e-7, 0.04, True),   # case 5: (-r)(ub - x) = 2e-8
    (0.0, 0.0, INF, 0.5, 0.5, 0.1, False),             # case 6: ub = +inf, r < 0
    (0.0, -INF, INF, 0.0, 0.0, -0.1, False),           # case 7: lb = -inf, r > 0
    (0.0, -INF, INF, 0.0, 0.0, -0.5, False),           # case 8: free variable, r != 0
])
def test_single_variable_cases(native, c, lb, ub, rhs, x, pi, expect):
    lp, lbv, ubv = single(c, lb, ub, rhs)
    ok, _ = native.verify_optimal(lp, np.array([x]), np.array([pi]), 1e-6, lb=lbv, ub=ubv)
    assert ok is expect


def two_slack_model():
    """min -x1 - x2  s.t.  x1 + s1 = 1,  x2 + s2 = 1,  x, s >= 0 (test_certificate.cpp:146-160)."""
    return lp_from_rows([[1, 0, 1, 0], [0, 1, 0, 1]], [1.0, 1.0], [-1.0, -1.0, 0.0, 0.0], 4)


@pytest.mark.parametrize("x,pi,expect", [
    ([1.0, 1.0, 0.0, 0.0], [-1.0, -1.0], True),        # known optimum
    ([1.0, 0.5, 0.0, 0.0], [-1.0, -1.0], False),       # Ax != b
    ([1.0, 1.0, 0.0, 0.0], [1.0, 1.0], False),         # r < 0 with ub = +inf
    ([0.5, 0.5, 0.5, 0.5], [-1.0, -1.0], False),       # complementarity product 0.5
])
def test_two_constraint_cases(native, x, pi, expect):
    ok, res = native.verify_optimal(two_slack_model(), np.array(x), np.array(pi), 1e-6)
    assert ok is expect
    if expect:
        assert res == 0.0


def test_tolerance_argument(native):
    # case 2 passes at 1e-6 (product 2.2e-7) but not at 1e-8
    lp, lbv, ubv = single(0.0, 0.0, INF, 5.3e-6)
    args = (lp, np.array([5.3e-6]), np.array([-0.0416]))
    assert native.verify_optimal(*args, 1e-6, lb=lbv, ub=ubv)[0]
    assert not native.verify_optimal(*args, 1e-8, lb=lbv, ub=ubv)[0]


def test_maximize_flips_reduced_cost_sign(native):
    # max x  s.t. x = 1, 0 <= x <= 1: r = -c - pi; pi = -1 gives r = 0 -> pass; pi = 0 gives
    # r = -1 < 0 at x = ub -> pass; as a minimization the same pi = 0 gives r = 1 > 0 at x = ub
    # with x - lb = 1 -> fail
    lp = lp_from_rows([[1.0]], [1.0], [1.0], 1)
    lb, ub = np.array([0.0]), np.array([1.0])
    x = np.array([1.0])
    assert native.verify_optimal(lp, x, np.array([-1.0]), 1e-6, lb=lb, ub=ub, maximize=True)[0]
    assert native.verify_optimal(lp, x, np.array([0.0]), 1e-6, lb=lb, ub=ub, maximize=True)[0]
    assert not native.verify_optimal(lp, x, np.array([0.0]), 1e-6, lb=lb, ub=ub, maximize=False)[0]


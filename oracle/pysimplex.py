"""CPU oracle of the device simplex route (test infrastructure only: the library never loads it).

A numpy restatement of the reference's CPU route for  min c^T x, A x = b, x >= 0:
solve_with_phase1 (src/simplex/phase1.cpp:124-449) around primal_simplex_phase2
(src/simplex/primal.cpp:27-423):
  * Phase I: rows with b_i < 0 negated, one artificial column per row (cost 1, the structural
    columns cost 0), the artificial basis with x_B = |b|  (phase1.cpp:188-284);
  * each phase: Devex pricing -- the largest r_j^2 / gamma_j over nonbasic r_j < -1e-6, gamma = 1
    at the start and every ncols iterations, the leaving column's gamma = max(1, |B^-1 a_q|^2)
    (primal.cpp:56-70,140-176,372-392) -- or, after more than 50 consecutive steps with
    theta < 1e-6, Bland's rule for the rest of the phase (primal.cpp:46-53,108-135,320-332);
    Harris two-pass ratio test over u_i > 1e-7 (primal.cpp:250-301) or, under Bland, the smallest
    ratio with ties to the lowest row (primal.cpp:228-249); at most 10,000 iterations per phase;
  * between the phases: IterationLimit ends the route, an artificial sum above 1e-6 is
    Infeasible, every row still holding an artificial takes the lowest nonbasic structural column
    with |(B^-1 a_j)_i| > 1e-7 (none: IterationLimit) (phase1.cpp:294-415); Phase II starts with a
    basic value below -1e-6 as Infeasible (primal.cpp:78-87).
Basis solves use a fresh dense LU per iteration (numpy.linalg.solve): the reference's sparse LU
with eta updates and the device's explicit Gauss-Jordan inverse are two other roundings of the
same B^-1, so the three agree on statuses and optima, not bit for bit.
"""
import numpy as np

FEAS = 1e-6     # math::kDefaultFeasibilityTol (include/sankhya/types.hpp)
PIVOT = 1e-7    # math::kDefaultPivotTol
BLAND = 50      # kBlandActivationThreshold (primal.cpp:52)
MAX_ITER = 10000
STATUS = {0: "Optimal", 1: "Infeasible", 2: "Unbounded", 3: "IterationLimit"}


def _primal(A, cost, basic, xb, max_iter):
    """Primal simplex on the dense A (m x ncols) from the feasible basis `basic` (row -> column)
    with basic values xb; returns (status, iterations); basic and xb are updated in place."""
    m, ncols = A.shape
    gamma = np.ones(ncols)
    deg, bland, it = 0, False, 0
    for iteration in range(max_iter):
        Bm = A[:, basic]
        y = np.linalg.solve(Bm.T, cost[basic])
        isb = np.zeros(ncols, bool)
        isb[basic] = True
        rc = cost - A.T @ y
        elig = (~isb) & (rc < -FEAS)
        if not elig.any():
            return 0, it
        if bland:
            q = int(np.flatnonzero(elig)[0])
        else:
            score = np.where(elig, rc * rc / gamma, -np.inf)
            q = int(np.argmax(score))
        u = np.linalg.solve(Bm, A[:, q])
        pos = u > PIVOT
        if not pos.any():
            return 2, it
        if bland:
            ratio = np.where(pos, xb / np.where(pos, u, 1.0), np.inf)
            p = int(np.argmin(ratio))
            theta = ratio[p]
        else:
            delta = np.min(np.where(pos, (xb + FEAS) / np.where(pos, u, 1.0), np.inf))
            cand = pos & (xb / np.where(pos, u, 1.0) <= delta)
            p = int(np.argmax(np.where(cand, np.abs(u), -1.0)))
            theta = max(xb[p] / u[p], 0.0)
        it += 1
        if theta < FEAS:
            deg += 1
            if deg > BLAND:
                bland = True
        else:
            deg = 0
        leave = basic[p]
        xb -= theta * u
        xb[p] = theta
        gamma[leave] = max(1.0, float(u @ u))
        basic[p] = q
        if (iteration + 1) % ncols == 0:
            gamma[:] = 1.0
    return 3, it


def solve(Acsc, b, c, max_iter=MAX_ITER):
    """The route on one LP (Acsc: scipy sparse or dense m x n).  Returns a dict with status,
    status_name, objective, x (n), y (m: B^-T c_B of the final basis, artificial rows cost 0),
    basis (m; -1 for a row left artificial) and iterations (Phase I, Phase II)."""
    A = np.asarray(Acsc.toarray() if hasattr(Acsc, "toarray") else Acsc, dtype=np.float64)
    b = np.asarray(b, np.float64)
    c = np.asarray(c, np.float64)
    m, n = A.shape
    sg = np.where(b < 0.0, -1.0, 1.0)
    A1 = np.hstack([sg[:, None] * A, np.eye(m)])
    cost1 = np.concatenate([np.zeros(n), np.ones(m)])
    basic = list(range(n, n + m))
    xb = np.abs(b).astype(np.float64)
    st, it1 = _primal(A1, cost1, basic, xb, max_iter)
    it2 = 0

    def result(status):
        x = np.zeros(n)
        bas = np.array([j if j < n else -1 for j in basic], np.int32)
        for i, j in enumerate(basic):
            if j < n:
                x[j] = xb[i]
        cb = np.array([c[j] if j < n else 0.0 for j in basic])
        Bm = np.hstack([A, np.eye(m)])[:, basic]
        try:
            y = np.linalg.solve(Bm.T, cb)
        except np.linalg.LinAlgError:
            y = np.full(m, np.nan)
        return {"status": status, "status_name": STATUS[status], "objective": float(c @ x),
                "x": x, "y": y, "basis": bas, "iterations": (it1, it2)}

    if st == 3:
        return result(3)
    art = sum(xb[i] for i in range(m) if basic[i] >= n and xb[i] > 0.0)
    if art > FEAS:
        return result(1)
    for r in range(m):                                  # artificials out (phase1.cpp:330-415)
        if basic[r] < n:
            continue
        Bm = A1[:, basic]
        q = -1
        for j in range(n):
            if j in basic:
                continue
            if abs(np.linalg.solve(Bm, A1[:, j])[r]) > PIVOT:
                q = j
                break
        if q < 0:
            return result(3)
        basic[r] = q
        xb[r] = 0.0
    if np.min(xb) < -FEAS:
        return result(1)
    st, it2 = _primal(A, c.copy(), basic, xb, max_iter)
    return result(st)


# The LPs of the reference's tests/unit/test_phase1.cpp (TEST_CASE line, CSC data, expected):
# each (m, n, colptr, rowidx, vals, b, c, status, objective or None, x or None)
PHASE1_CASES = {
    "test1_slack_basis_123": (2, 4, [0, 2, 3, 4, 5], [0, 1, 0, 0, 1], [1.0] * 5, [4.0, 2.0],
                              [-1.0, -1.0, 0.0, 0.0], 0, -4.0, None),
    "test2_equality_161": (1, 2, [0, 1, 2], [0, 0], [1.0, 1.0], [3.0], [1.0, 1.0], 0, 3.0, None),
    "test3_negative_rhs_198": (1, 1, [0, 1], [0], [-1.0], [-5.0], [1.0], 0, 5.0, [5.0]),
    "test4_phase1_zero_229": (2, 2, [0, 2, 4], [0, 1, 0, 1], [1.0, 1.0, 1.0, -1.0], [1.0, 0.0],
                              [1.0, 2.0], 0, 1.5, [0.5, 0.5]),
    "test5_infeasible_264": (2, 2, [0, 2, 4], [0, 1, 0, 1], [1.0, 1.0, 1.0, 1.0], [1.0, 2.0],
                             [0.0, 0.0], 1, None, None),
    "test6_redundant_row_294": (2, 2, [0, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 1.0, 2.0], [3.0, 6.0],
                                [1.0, 1.0], 3, None, None),
    "test8_original_basis_366": (1, 2, [0, 1, 2], [0, 0], [1.0, 1.0], [5.0], [1.0, 0.0], 0, 0.0, None),
    "test9_full_solve_410": (2, 4, [0, 2, 3, 4, 5], [0, 1, 0, 0, 1], [1.0] * 5, [4.0, 2.0],
                             [-3.0, -2.0, 0.0, 0.0], 0, -10.0, None),
}


def case_dense(name):
    m, n, cp, ri, va, b, c, *_ = PHASE1_CASES[name]
    A = np.zeros((m, n))
    for j in range(n):
        for k in range(cp[j], cp[j + 1]):
            A[ri[k], j] = va[k]
    return A, np.array(b, float), np.array(c, float)


def brute_force(A, b, c):
    """Optimum of min c^T x, A x = b, x >= 0 by enumeration of every basis (tiny LPs): the
    smallest objective over feasible basic solutions, or None when there is none (an unbounded
    LP is not detected here: use bounded instances)."""
    from itertools import combinations
    A = np.asarray(A, float)
    m, n = A.shape
    best = None
    for cols in combinations(range(n), m):
        Bm = A[:, cols]
        if abs(np.linalg.det(Bm)) < 1e-10:
            continue
        xb = np.linalg.solve(Bm, b)
        if np.min(xb) < -1e-9:
            continue
        val = float(np.asarray(c)[list(cols)] @ xb)
        best = val if best is None else min(best, val)
    return best

"""Crossover restated in numpy (TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the
device crossover, never as part of the product path).

From an interior-point solution x of  min c^T x, A x = b, x >= 0  to a basic solution: the same
steps, thresholds and orders as the device kernel (k_crossover), in float64 with a dense
explicit basis inverse.  The reference's crossover (src/ipm/crossover.cu:322-460) purifies
column by column from an artificial basis; this one builds the basis from the solution's
support first and purifies what remains:
  1. xb = max(x, 0); tau = 1e-7 max(1, max xb); the support S = {j: xb_j > tau}, ordered by xb
     descending (ties: lower index first), then the other columns in index order;
  2. basis: B^-1 = I with every row artificial; each candidate q in that order (while an
     artificial row remains) gets u = B^-1 a_q and the artificial row r of largest |u_r|
     (lowest r on ties); if |u_r| > 1e-9 max(1, max |u|) column q replaces the artificial:
     Gauss-Jordan update of B^-1 on pivot (r, u);
  3. purification: while a nonbasic column q with xb_q > tau exists (lowest q first):
     u = B^-1 a_q, reduced cost d_q = c_q - sum_i c_B(i) u_i (artificial rows cost 0); the step
     moves x_q by theta and the basic values by -theta u: theta > 0 when
     d_q <= 1e-11 (1 + |c_q|) (ratio test over rows with u_i > 1e-12, basic values floored at
     0; none: unbounded), else theta = -min(xb_q, ratios over rows with u_i < -1e-12); a
     blocking basic row r (the first minimum) pivots q in (its variable leaves at 0); x_q
     blocking sets xb_q = 0;
  4. x_B = B^-1 b, x_N = 0; y = B^-T c_B; reduced costs; status 0 optimal vertex, 1 feasible
     but not dual feasible, 2 rank deficient (an artificial row left), 3 unbounded direction,
     4 basic solution infeasible (min x_B < -1e-9 max(1, max|x_B|)).
"""
import numpy as np


def crossover(A, b, c, x):
    """A dense (m x n).  Returns (basis[m] with -1 for artificial rows, xv[n], status, pivots)."""
    A = np.asarray(A, np.float64)
    m, n = A.shape
    xb = np.maximum(np.asarray(x, np.float64), 0.0)
    tau = 1e-7 * max(1.0, float(xb.max()) if n else 1.0)
    sup = [j for j in np.argsort(-xb, kind="stable") if xb[j] > tau]
    rest = [j for j in range(n) if not xb[j] > tau]
    Binv = np.eye(m)
    basis = -np.ones(m, dtype=np.int64)
    isb = np.zeros(n, dtype=bool)
    pivots = 0

    def pivot(r, u):
        row = Binv[r, :] / u[r]
        Binv[:, :] -= np.outer(u, row)
        Binv[r, :] = row

    for q in sup + rest:
        if not (basis < 0).any():
            break
        u = Binv @ A[:, q]
        art = np.where(basis < 0)[0]
        r = art[np.argmax(np.abs(u[art]))]
        if abs(u[r]) > 1e-9 * max(1.0, float(np.abs(u).max())):
            pivot(r, u)
            basis[r] = q
            isb[q] = True
            pivots += 1
    status = 0
    # current basic values (the support's values for basic columns; artificial rows 0)
    xv = xb.copy()
    while True:
        cand = [q for q in range(n) if not isb[q] and xv[q] > tau]
        if not cand:
            break
        q = cand[0]
        u = Binv @ A[:, q]
        cb = np.array([c[j] if j >= 0 else 0.0 for j in basis])
        d = c[q] - cb @ u
        vals = np.array([max(xv[j], 0.0) if j >= 0 else 0.0 for j in basis])
        if d <= 1e-11 * (1.0 + abs(c[q])):
            rows = np.where(u > 1e-12)[0]
            if rows.size == 0:
                status = 3
                break
            ratios = vals[rows] / u[rows]
            k = int(np.argmin(ratios))
            theta, r = float(ratios[k]), int(rows[k])
        else:
            rows = np.where(u < -1e-12)[0]
            theta, r = float(xv[q]), -1
            if rows.size:
                ratios = vals[rows] / (-u[rows])
                k = int(np.argmin(ratios))
                if ratios[k] < theta:
                    theta, r = float(ratios[k]), int(rows[k])
            theta = -theta
        for i in range(m):
            if basis[i] >= 0:
                xv[basis[i]] -= theta * u[i]
        xv[q] += theta
        if r >= 0:
            leave = basis[r]
            pivot(r, u)
            if leave >= 0:
                isb[leave] = False
                xv[leave] = 0.0
            basis[r] = q
            isb[q] = True
            pivots += 1
        else:
            xv[q] = 0.0
    cb = np.array([c[j] if j >= 0 else 0.0 for j in basis])
    xB = Binv @ b
    out = np.zeros(n)
    for i in range(m):
        if basis[i] >= 0:
            out[basis[i]] = xB[i]
    if status == 0:
        if (basis < 0).any():
            status = 2
        elif xB.min() < -1e-9 * max(1.0, float(np.abs(xB).max())):
            status = 4
        else:
            y = Binv.T @ cb
            red = c - A.T @ y
            if red.min() < -1e-9 * max(1.0, float(np.abs(c).max())):
                status = 1
    return basis, out, status, pivots

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
  4. clean-up when no artificial row is left, with x_B = B^-1 b, y = B^-T c_B, d = c - A'y
     (0 on basic columns), ptol = 1e-9 max(1, max|x_B|), dtol = 1e-9 max(1, max|c|):
     a. dual simplex, at most m pivots: while the most negative x_B(r) (lowest r on ties) is
        below -ptol, the row alpha = (B^-1)_r A over nonbasic columns with
        alpha_j < -1e-9 max(1, max|alpha|) chooses the entering q = argmin d'_j / -alpha_j,
        d'_j = d_j if d_j > dtol else 0 (lowest j on ties); no such column ends this phase;
     b. primal simplex with Bland's rule, at most m pivots: the lowest nonbasic q with
        d_q < -dtol enters; u = B^-1 a_q; over rows with u_i > 1e-9 max(1, max|u|) the ratio
        x'_i / u_i, x'_i = x_B(i) if x_B(i) > ptol else 0, is minimised, ties to the lowest basic
        column; no such row: unbounded (status 3);
     q replaces the basic column of the chosen row (Gauss-Jordan update as in 2);
  5. x_B = B^-1 b, x_N = 0; y = B^-T c_B; reduced costs; status 0 optimal vertex, 1 feasible
     but not dual feasible, 2 rank deficient (an artificial row left), 3 unbounded direction,
     4 basic solution infeasible (min x_B < -1e-9 max(1, max|x_B|)).
Step 4 acts at primal-degenerate optimal vertices (support below m): the columns that fill the
last rows take B^-1 b values of the size of the point's residual, of either sign (4a), and the
reduced costs of that basis need not all be >= 0 (4b).
"""
import numpy as np


def crossover(A, b, c, x, cleanup=True):
    """A dense (m x n).  Returns (basis[m] with -1 for artificial rows, xv[n], status, pivots).
    cleanup=False skips step 4 (tests use it to show the step is needed)."""
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
    if cleanup and status == 0 and not (basis < 0).any():
        dtol = 1e-9 * max(1.0, float(np.abs(c).max()))

        def state():
            xB = Binv @ b
            y = Binv.T @ np.array([c[j] for j in basis])
            d = c - A.T @ y
            d[isb] = 0.0
            return xB, d, 1e-9 * max(1.0, float(np.abs(xB).max()))

        def swap(r, q):
            u = Binv @ A[:, q]
            pivot(r, u)
            isb[basis[r]] = False
            basis[r] = q
            isb[q] = True

        for _ in range(m):                       # dual simplex: primal infeasibilities
            xB, d, ptol = state()
            r = int(np.argmin(xB))
            if not xB[r] < -ptol:
                break
            alpha = Binv[r, :] @ A
            alpha[isb] = 0.0
            cand = np.where(alpha < -1e-9 * max(1.0, float(np.abs(alpha).max())))[0]
            if cand.size == 0:
                break
            ratios = np.where(d[cand] > dtol, d[cand], 0.0) / (-alpha[cand])
            swap(r, int(cand[int(np.argmin(ratios))]))
            pivots += 1
        for _ in range(m):                       # primal simplex (Bland): dual infeasibilities
            xB, d, ptol = state()
            neg = np.where(d < -dtol)[0]
            if neg.size == 0:
                break
            q = int(neg[0])
            u = Binv @ A[:, q]
            rows = np.where(u > 1e-9 * max(1.0, float(np.abs(u).max())))[0]
            if rows.size == 0:
                status = 3
                break
            ratios = np.where(xB[rows] > ptol, xB[rows], 0.0) / u[rows]
            best = ratios.min()
            tied = rows[ratios == best]
            r = int(tied[np.argmin(basis[tied])])
            swap(r, q)
            pivots += 1
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

/*
 * CPU oracle for the Mehrotra interior-point path.  TEST INFRASTRUCTURE ONLY: it is linked and
 * called by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker (or the timed CPU baseline), never as part of the product path.
 *
 * Plain C99 restatement of the reference algorithm.  Each function cites what it follows:
 *   or_ruiz              src/ipm/mehrotra_device.cu:286-323   (10 Jacobi Ruiz passes)
 *   or_md_order          src/ipm/symbolic.cpp:11-98           (explicit-graph minimum degree)
 *   or_symbolic          src/ipm/symbolic.cpp:100-218         (etree + L pattern, set unions)
 *   or_assemble_dense    src/cuda/kkt.cu:177-215              (M0 = P A Theta A^T P^T by row merge)
 *   or_kkt_factor        src/cuda/kkt.cu:556-603              (M0 + 1e-9 I, throws if not SPD)
 *   or_kkt_solve         src/cuda/kkt.cu:620-870              (equilibration, delta retries,
 *                                                              refinement against M0, rollback)
 *   or_ipm_solve         src/ipm/mehrotra_device.cu:695-1047  (predictor-corrector, merit
 *                                                              backtracking, centred fallback)
 *   or_verify_optimal    src/verifier/certificate.cpp:8-155
 * Dense Cholesky (O(m^3)) -- sized for m up to a few hundred.
 * Deliberate differences from the reference: the factorization of D*M0*D + delta*I is cached per
 * IPM iteration (the reference recomputes the identical factor in every solve call), and nothing
 * is printed.  Relative-gap mode (stop_rule 1, an extension with no reference counterpart) uses
 * PCG instead of plain refinement, regularizes from delta = 1e-14 (not 1e-9), accepts a solve at
 * the backward-error level only when PCG stagnates, skips the e1 re-check of the KKT residual,
 * stops on absolute residual and per-column complementarity targets as well (at REL_ABS_MARGIN
 * of gap_tol), rejects trials that take a residual already inside its absolute target back
 * out of it, tries dual-only steps of the Mehrotra direction before the centred one (the
 * primal residual cannot grow; near the optimum the normal equations limit the primal
 * accuracy of a direction), and takes the longest finite, positive centred trial when all
 * merit searches fail (non-monotone fallback, never leaving a met residual target), and, once
 * both residuals are inside their targets, lets the dual-only trials compete with a corrector
 * step shorter than SHORT_STEP (better merit wins): the reference's absolute 1e-9 acceptance and
 * monotone merit search stall before a 1e-8 gap (its own loop stops at 1e-6).  The device path
 * implements the same rules.  Sums use the reference's term order; multiply-adds use fma() so the assembled
 * M0 is bit-identical to an fma-based device kernel with the same order.
 *
 * Parity pins (tests/test_oracle.py): the symbolic KATs of tests/unit/test_ipm_symbolic.cpp,
 * the KKT KATs of tests/unit/test_gpu_kkt.cu, the textbook LP of tests/unit/test_mehrotra.cu,
 * the 12 certificate KATs of tests/unit/test_certificate.cpp, HiGHS golden objectives, and the
 * reference's own symbolic/verifier sources compiled into oracle/_ref where available.
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* relative-gap mode: the absolute primal, dual and complementarity targets are this fraction of
 * gap_tol (the FP64 certificate checks them absolutely at gap_tol) */
#define REL_ABS_MARGIN 0.9
#define SHORT_STEP 1e-2

#define OR_OK 0
#define OR_NOT_SPD 1
#define OR_BAD_ARG 2
#define OR_NOMEM 3

typedef struct {
    int stop_rule;      /* 0: reference (scaled 2-norms and mu < tol), 1: relative 1e-8 gap */
    double tol;         /* reference-mode threshold (1e-6) */
    double gap_tol;     /* relative-mode threshold (1e-8) */
    int max_iter;       /* 200 */
} or_ipm_options;

typedef struct {
    int status;         /* 0 Optimal, 3 IterationLimit, -1 exception (not SPD in factorize) */
    int iterations;
    double objective;   /* c_s^T x_s = c^T x */
    double primal_residual, dual_residual, duality_gap;  /* scaled ||rp||2, ||rd||2, mu */
    double rel_gap, rel_primal_inf, rel_dual_inf, dual_objective;
    int factorizations, solves;
    int fail_code;      /* 0 none, 1 factorization, 2 predictor solve, 3 predictor check,
                           4 corrector solve, 5 no step accepted */
} or_ipm_result;

/* ------------------------------------------------------------------------------------------ */
void or_ruiz(int m, int n, const int* rp, const int* ci, double* val, double* b, double* c,
             double* R, double* C) {
    double *rmax = malloc(sizeof(double) * (m ? m : 1)), *cmax = malloc(sizeof(double) * (n ? n : 1));
    for (int i = 0; i < m; ++i) R[i] = 1.0;
    for (int j = 0; j < n; ++j) C[j] = 1.0;
    for (int it = 0; it < 10; ++it) {
        for (int i = 0; i < m; ++i) rmax[i] = 0.0;
        for (int j = 0; j < n; ++j) cmax[j] = 0.0;
        for (int i = 0; i < m; ++i)
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                double a = fabs(val[p]);
                if (a > rmax[i]) rmax[i] = a;
                if (a > cmax[ci[p]]) cmax[ci[p]] = a;
            }
        for (int i = 0; i < m; ++i) {
            double f = rmax[i] > 1e-8 ? 1.0 / sqrt(rmax[i]) : 1.0;
            R[i] *= f;
            b[i] *= f;
        }
        for (int j = 0; j < n; ++j) {
            double f = cmax[j] > 1e-8 ? 1.0 / sqrt(cmax[j]) : 1.0;
            C[j] *= f;
            c[j] *= f;
        }
        for (int i = 0; i < m; ++i)
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                double fr = rmax[i] > 1e-8 ? 1.0 / sqrt(rmax[i]) : 1.0;
                double fc = cmax[ci[p]] > 1e-8 ? 1.0 / sqrt(cmax[ci[p]]) : 1.0;
                val[p] *= fr * fc;
            }
    }
    free(rmax);
    free(cmax);
}

/* ------------------------------------------------------------------------------------------ */
/* adjacency of A*A^T as a dense boolean matrix (small m only) */
static unsigned char* aat_graph(int m, int n, const int* rp, const int* ci) {
    unsigned char* g = calloc((size_t)m * m + 1, 1);
    int* cnt = calloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < m; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) cnt[ci[p] + 1]++;
    for (int k = 0; k < n; ++k) cnt[k + 1] += cnt[k];
    int* rows = malloc(sizeof(int) * (cnt[n] + 1));
    int* nx = malloc(sizeof(int) * (n + 1));
    memcpy(nx, cnt, sizeof(int) * n);
    for (int i = 0; i < m; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) rows[nx[ci[p]]++] = i;
    for (int k = 0; k < n; ++k)
        for (int a = cnt[k]; a < cnt[k + 1]; ++a)
            for (int b = cnt[k]; b < cnt[k + 1]; ++b)
                if (rows[a] != rows[b]) g[(size_t)rows[a] * m + rows[b]] = 1;
    free(cnt); free(rows); free(nx);
    return g;
}

int or_md_order(int m, int n, const int* rp, const int* ci, int* P) {
    unsigned char* g = aat_graph(m, n, rp, ci);
    unsigned char* done = calloc((size_t)m + 1, 1);
    int* nb = malloc(sizeof(int) * (m + 1));
    for (int step = 0; step < m; ++step) {
        int best = -1, bestd = m + 1;
        for (int i = 0; i < m; ++i) {
            if (done[i]) continue;
            int d = 0;
            for (int j = 0; j < m; ++j) d += g[(size_t)i * m + j];
            if (d < bestd) { bestd = d; best = i; }
        }
        done[best] = 1;
        P[step] = best;
        int k = 0;
        for (int j = 0; j < m; ++j)
            if (g[(size_t)best * m + j]) nb[k++] = j;
        for (int a = 0; a < k; ++a) {
            g[(size_t)nb[a] * m + best] = 0;
            for (int b = 0; b < k; ++b)
                if (a != b) g[(size_t)nb[a] * m + nb[b]] = 1;
        }
        for (int j = 0; j < m; ++j) g[(size_t)best * m + j] = 0;
    }
    free(g); free(done); free(nb);
    return OR_OK;
}

/* L pattern by the reference's column-union construction: struct(j) = lower M(:,j) U
 * struct(children with parent j) restricted to rows > j; parent(j) = min struct(j).
 * Outputs parent[m], Lptr[m+1] and Lidx (capacity cap) in CSR (row lists ascending). */
int or_symbolic(int m, int n, const int* rp, const int* ci, const int* P, int* parent,
                int* Lptr, int* Lidx, int64_t cap) {
    unsigned char* g = aat_graph(m, n, rp, ci);
    int* pinv = malloc(sizeof(int) * (m + 1));
    for (int i = 0; i < m; ++i) pinv[P[i]] = i;
    unsigned char* L = calloc((size_t)m * m + 1, 1);   /* L[i*m+j], i > j */
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
            if (g[(size_t)a * m + b]) {
                int i = pinv[a], j = pinv[b];
                if (i > j) L[(size_t)i * m + j] = 1;
            }
    for (int j = 0; j < m; ++j) {
        for (int c = 0; c < j; ++c)
            if (parent[c] == j)
                for (int i = j + 1; i < m; ++i)
                    if (L[(size_t)i * m + c]) L[(size_t)i * m + j] = 1;
        parent[j] = -1;
        for (int i = j + 1; i < m; ++i)
            if (L[(size_t)i * m + j]) { parent[j] = i; break; }
    }
    int64_t nz = 0;
    Lptr[0] = 0;
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < i; ++j)
            if (L[(size_t)i * m + j]) {
                if (nz < cap) Lidx[nz] = j;
                nz++;
            }
        Lptr[i + 1] = (int)nz;
    }
    free(g); free(pinv); free(L);
    return nz <= cap ? OR_OK : OR_BAD_ARG;
}

/* ------------------------------------------------------------------------------------------ */
/* M0(i,j) = sum_k (A[P_i,k] * Theta_k) * A[P_j,k], k ascending (merge of two sorted rows). */
void or_assemble_dense(int m, const int* rp, const int* ci, const double* val,
                       const double* theta, const int* P, double* M) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            int oi = P[i], oj = P[j];
            int pi = rp[oi], ei = rp[oi + 1], pj = rp[oj], ej = rp[oj + 1];
            double s = 0.0;
            while (pi < ei && pj < ej) {
                if (ci[pi] == ci[pj]) { s = fma(val[pi] * theta[ci[pi]], val[pj], s); ++pi; ++pj; }
                else if (ci[pi] < ci[pj]) ++pi;
                else ++pj;
            }
            M[(size_t)i * m + j] = s;
        }
}

/* in-place dense Cholesky of the lower triangle (row-major); returns 0 or 1 when a pivot <= 0 */
static int dense_chol(int m, double* a) {
    for (int j = 0; j < m; ++j) {
        double d = a[(size_t)j * m + j];
        for (int k = 0; k < j; ++k) d -= a[(size_t)j * m + k] * a[(size_t)j * m + k];
        if (!(d > 0.0)) return 1;
        d = sqrt(d);
        a[(size_t)j * m + j] = d;
        for (int i = j + 1; i < m; ++i) {
            double s = a[(size_t)i * m + j];
            for (int k = 0; k < j; ++k) s -= a[(size_t)i * m + k] * a[(size_t)j * m + k];
            a[(size_t)i * m + j] = s / d;
        }
    }
    return 0;
}

static void chol_solve(int m, const double* L, double* x) {
    for (int i = 0; i < m; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * m + k] * x[k];
        x[i] = s / L[(size_t)i * m + i];
    }
    for (int i = m - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < m; ++k) s -= L[(size_t)k * m + i] * x[k];
        x[i] = s / L[(size_t)i * m + i];
    }
}

static double inf_norm(int n, const double* v) {
    double r = 0.0;
    for (int i = 0; i < n; ++i) if (fabs(v[i]) > r) r = fabs(v[i]);
    return r;
}

/* KKT state over the permuted system */
typedef struct {
    int m;
    const int* P;
    const unsigned char* has_diag;   /* reference M pattern holds a diagonal only for non-empty rows */
    double* M0;      /* assembled, permuted, unregularised */
    double* Leq;     /* factor of D M0 D + delta I (cached for the current M0) */
    double* D;
    int have_eq;     /* Leq valid */
    double delta_eq;
} or_kkt;

/* src/cuda/kkt.cu:556-603: factor M0 + 1e-9 I unequilibrated; not SPD -> error */
static int kkt_factor_check(or_kkt* k) {
    int m = k->m;
    double* a = malloc(sizeof(double) * ((size_t)m * m + 1));
    memcpy(a, k->M0, sizeof(double) * (size_t)m * m);
    for (int i = 0; i < m; ++i)
        if (k->has_diag[i]) a[(size_t)i * m + i] += 1e-9;
    int bad = dense_chol(m, a);
    free(a);
    k->have_eq = 0;
    return bad ? OR_NOT_SPD : OR_OK;
}

static int kkt_factor_eq(or_kkt* k, double delta) {
    int m = k->m;
    for (int i = 0; i < m; ++i) {
        double d = fabs(k->M0[(size_t)i * m + i]);
        if (!k->has_diag[i]) d = 0.0;
        if (d < 1e-12) d = 1e-12;
        k->D[i] = 1.0 / sqrt(d);
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double v = k->M0[(size_t)i * m + j] * k->D[i] * k->D[j];
            if (i == j && k->has_diag[i]) v += delta;
            k->Leq[(size_t)i * m + j] = v;
        }
    return dense_chol(m, k->Leq);
}

/* residual r = rhs - M0 dy  (permuted) */
static void m0_residual(const or_kkt* k, const double* dy, const double* rhs, double* r) {
    int m = k->m;
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) s += k->M0[(size_t)i * m + j] * dy[j];
        r[i] = rhs[i] - s;
    }
}

/* src/cuda/kkt.cu:620-870.  rhs (original order) is overwritten by dy even on failure. */
/* Relative-gap mode only: a solve is also accepted when its residual is at the level a backward
 * stable solve can reach, ||r||_inf <= 1e-12 (||rhs||_inf + max_i M0_ii ||dy||_inf). */
static int backward_ok(const or_kkt* k, const double* rhs, const double* dy, double rinf) {
    int m = k->m;
    double mx = 0.0;
    for (int i = 0; i < m; ++i) if (k->M0[(size_t)i * m + i] > mx) mx = k->M0[(size_t)i * m + i];
    return rinf <= 1e-12 * (inf_norm(m, rhs) + mx * inf_norm(m, dy));
}

/* Relative-gap mode: preconditioned conjugate gradients on M0 dy = rhs (permuted), started from
 * the direct solve, preconditioner D (L L^T)^{-1} D with the cached equilibrated factor.  The
 * true residual is recomputed whenever the recurrence claims convergence and at the end; the
 * best iterate (smallest true ||r||_inf) is returned.  Returns 1 on convergence. */
static int kkt_pcg(or_kkt* k, const double* rhs, double* dy, double tol, int max_it) {
    int m = k->m;
    double *r = malloc(sizeof(double) * (m + 1)), *z = malloc(sizeof(double) * (m + 1)),
           *p = malloc(sizeof(double) * (m + 1)), *q = malloc(sizeof(double) * (m + 1)),
           *best = malloc(sizeof(double) * (m + 1));
    m0_residual(k, dy, rhs, r);
    double rbest = inf_norm(m, r);
    memcpy(best, dy, sizeof(double) * m);
    int ok = rbest <= tol;
    int stall = 0;      /* true-residual checks without a 2x improvement of the best */
    double rz = 0.0;
    if (!ok) {
        for (int i = 0; i < m; ++i) z[i] = r[i] * k->D[i];
        chol_solve(m, k->Leq, z);
        for (int i = 0; i < m; ++i) { z[i] *= k->D[i]; p[i] = z[i]; rz += r[i] * z[i]; }
    }
    for (int it = 0; it < max_it && !ok; ++it) {
        for (int i = 0; i < m; ++i) {
            double t = 0.0;
            for (int j = 0; j < m; ++j) t += k->M0[(size_t)i * m + j] * p[j];
            q[i] = t;
        }
        double pq = 0.0;
        for (int i = 0; i < m; ++i) pq += p[i] * q[i];
        if (!(pq > 0.0)) break;
        double alpha = rz / pq;
        for (int i = 0; i < m; ++i) { dy[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        double rr = inf_norm(m, r);
        if (rr <= tol || it == max_it - 1 || (it + 1) % 5 == 0) {
            m0_residual(k, dy, rhs, r);          /* true residual */
            rr = inf_norm(m, r);
            if (rr < 0.5 * rbest) stall = 0; else stall++;
            if (rr < rbest) { rbest = rr; memcpy(best, dy, sizeof(double) * m); }
            if (rr <= tol) { ok = 1; break; }
            if (stall >= 2) break;              /* stagnated: accept at the backward level */
        }
        for (int i = 0; i < m; ++i) z[i] = r[i] * k->D[i];
        chol_solve(m, k->Leq, z);
        double rz2 = 0.0;
        for (int i = 0; i < m; ++i) { z[i] *= k->D[i]; rz2 += r[i] * z[i]; }
        double beta = rz2 / rz;
        rz = rz2;
        for (int i = 0; i < m; ++i) p[i] = z[i] + beta * p[i];
    }
    if (!ok) {
        memcpy(dy, best, sizeof(double) * m);
        ok = backward_ok(k, rhs, dy, rbest);
    }
    free(r); free(z); free(p); free(q); free(best);
    return ok;
}

static int kkt_solve(or_kkt* k, double* rhs_orig, double tol, int relmode) {
    int m = k->m;
    double* rhs = malloc(sizeof(double) * (m + 1));
    double* dy = calloc((size_t)m + 1, sizeof(double));
    double* bak = malloc(sizeof(double) * (m + 1));
    double* r = malloc(sizeof(double) * (m + 1));
    double* corr = malloc(sizeof(double) * (m + 1));
    for (int i = 0; i < m; ++i) rhs[i] = rhs_orig[k->P[i]];
    int success = 0;
    /* reference: 1e-9 x 10^r, r < 5; relative-gap mode starts at 1e-14 (PCG then needs a step or
     * two instead of tens) and keeps the same ceiling 1e-5 */
    double delta = relmode ? 1e-14 : 1e-9;
    const int nretry = relmode ? 10 : 5;
    for (int retry = 0; retry < nretry; ++retry) {
        if (retry > 0) delta *= 10.0;
        if (!(k->have_eq && k->delta_eq == delta)) {
            if (kkt_factor_eq(k, delta)) { k->have_eq = 0; continue; }
            k->have_eq = 1;
            k->delta_eq = delta;
        }
        for (int i = 0; i < m; ++i) dy[i] = rhs[i] * k->D[i];
        chol_solve(m, k->Leq, dy);
        for (int i = 0; i < m; ++i) dy[i] *= k->D[i];
        m0_residual(k, dy, rhs, r);
        double before = inf_norm(m, r);
        if (before <= tol) { success = 1; break; }
        if (relmode) { success = kkt_pcg(k, rhs, dy, tol, 60); break; }
        int converged = 0;
        double after = before;
        for (int it = 0; it < 10; ++it) {
            memcpy(bak, dy, sizeof(double) * m);
            for (int i = 0; i < m; ++i) corr[i] = r[i] * k->D[i];
            chol_solve(m, k->Leq, corr);
            for (int i = 0; i < m; ++i) dy[i] += corr[i] * k->D[i];
            m0_residual(k, dy, rhs, r);
            after = inf_norm(m, r);
            if (after <= tol) { converged = 1; break; }
            if (after >= before) {
                memcpy(dy, bak, sizeof(double) * m);
                m0_residual(k, dy, rhs, r);
                break;
            }
            before = after;
        }
        if (converged) success = 1;
        break;
    }
    for (int i = 0; i < m; ++i) rhs_orig[k->P[i]] = dy[i];
    free(rhs); free(dy); free(bak); free(r); free(corr);
    return success;
}

/* ---- exported KKT helpers (dense, small m): factor + solve on an explicit A, Theta, P ---- */
/* Sums of the pair-map segments in the device's documented fixed order (the summation contract
 * of the assembly kernels, include/lakshya_device.h, for the normal-equations entries of
 * src/cuda/kkt.cu:177-215 M0_ij = sum_k A_ik theta_k A_jk): segment t holds pairs
 * q = segptr[t] .. segptr[t+1]-1 with products pp[q] and columns pk[q];
 *   L <= 4:   acc = fma(pp[q], theta[pk[q]], acc) over q in order from 0;
 *   L <= 64:  16 lanes, lane l: fma over q = b+l, b+l+16, b+l+32, b+l+48 (past the end the
 *             product is 0 * theta[pk[e-1]]), then a xor butterfly v_l += v_{l^o}, o = 8,4,2,1;
 *   L > 64:   256 threads, thread t: fma over q = b+t, b+t+256, ..; a xor butterfly over each
 *             group of 64 (o = 32 .. 1), then (w0 + w1) + (w2 + w3).
 * Test infrastructure: the checker of the device assembly, bit for bit. */
static double seg_fixed(int64_t b, int64_t e, const double* pp, const int* pk, const double* theta) {
    const int64_t L = e - b;
    if (L <= 4) {
        double acc = 0.0;
        for (int64_t q = b; q < e; ++q) acc = fma(pp[q], theta[pk[q]], acc);
        return acc;
    }
    if (L <= 64) {
        double v[16];
        for (int l = 0; l < 16; ++l) {
            double acc = 0.0;
            for (int u = 0; u < 4; ++u) {
                const int64_t q0 = b + l + 16 * u;
                const int64_t q = q0 < e ? q0 : e - 1;
                acc = fma(q0 < e ? pp[q] : 0.0, theta[pk[q]], acc);
            }
            v[l] = acc;
        }
        for (int o = 8; o > 0; o >>= 1) {
            double w[16];
            for (int l = 0; l < 16; ++l) w[l] = v[l] + v[l ^ o];
            for (int l = 0; l < 16; ++l) v[l] = w[l];
        }
        return v[0];
    }
    double v[256];
    for (int t = 0; t < 256; ++t) {
        double acc = 0.0;
        for (int64_t q = b + t; q < e; q += 256) acc = fma(pp[q], theta[pk[q]], acc);
        v[t] = acc;
    }
    double sw[4];
    for (int wv = 0; wv < 4; ++wv) {
        double* x = v + 64 * wv;
        for (int o = 32; o > 0; o >>= 1) {
            double w[64];
            for (int l = 0; l < 64; ++l) w[l] = x[l] + x[l ^ o];
            for (int l = 0; l < 64; ++l) x[l] = w[l];
        }
        sw[wv] = x[0];
    }
    return (sw[0] + sw[1]) + (sw[2] + sw[3]);
}

void or_fma_vec(int64_t n, const double* a, const double* b, const double* c, double* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = fma(a[i], b[i], c[i]);
}

void or_segment_sums(int64_t nseg, const int64_t* segptr, const double* pp, const int* pk,
                     const double* theta, double* out) {
    for (int64_t t = 0; t < nseg; ++t) out[t] = seg_fixed(segptr[t], segptr[t + 1], pp, pk, theta);
}

int or_kkt_factor_solve(int m, const int* rp, const int* ci, const double* val,
                        const double* theta, const int* P, double* Lout_eq_free,
                        double* rhs, double tol, int* success) {
    or_kkt k;
    k.m = m;
    k.P = P;
    unsigned char* hd = malloc((size_t)m + 1);
    for (int i = 0; i < m; ++i) hd[i] = rp[P[i] + 1] > rp[P[i]];
    k.has_diag = hd;
    k.M0 = malloc(sizeof(double) * ((size_t)m * m + 1));
    k.Leq = malloc(sizeof(double) * ((size_t)m * m + 1));
    k.D = malloc(sizeof(double) * (m + 1));
    k.have_eq = 0;
    or_assemble_dense(m, rp, ci, val, theta, P, k.M0);
    int st = kkt_factor_check(&k);
    if (st == OR_OK && Lout_eq_free) {
        /* unequilibrated factor of M0 + 1e-9 I, as the reference exposes after factorize() */
        memcpy(Lout_eq_free, k.M0, sizeof(double) * (size_t)m * m);
        for (int i = 0; i < m; ++i) if (hd[i]) Lout_eq_free[(size_t)i * m + i] += 1e-9;
        dense_chol(m, Lout_eq_free);
    }
    if (st == OR_OK && rhs) *success = kkt_solve(&k, rhs, tol, 0);
    free(hd); free(k.M0); free(k.Leq); free(k.D);
    return st;
}

/* ------------------------------------------------------------------------------------------ */
static void spmv(int m, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * x[ci[p]];
        y[i] += s;
    }
}
static void spmv_t(int m, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    for (int i = 0; i < m; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) y[ci[p]] += v[p] * x[i];
}
static double nrm2(int n, const double* v) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += v[i] * v[i];
    return sqrt(s);
}
static double dot(int n, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double step_len(int n, const double* v, const double* dv) {
    double r = 1e30;
    for (int i = 0; i < n; ++i) {
        double t = dv[i] >= 0.0 ? 1e30 : -v[i] / dv[i];
        if (t < r) r = t;
    }
    return r;
}
static double vmin(int n, const double* v) {
    double r = 1e30;
    for (int i = 0; i < n; ++i) if (v[i] < r) r = v[i];
    return r;
}

/* The IPM.  A in CSR (original rows), b, c.  x0/y0/s0 optional warm start (scaled space if
 * given; NULL -> x = s = 1, y = 0).  Outputs x (n), y (m), s (n) in the original space. */
int or_ipm_solve(int m, int n, const int* rp_in, const int* ci, const double* val_in,
                 const double* b_in, const double* c_in, const int* P_in,
                 const or_ipm_options* opt, or_ipm_result* res, double* x_out, double* y_out,
                 double* s_out) {
    int nnz = rp_in[m];
    int* rp = malloc(sizeof(int) * (m + 1));
    memcpy(rp, rp_in, sizeof(int) * (m + 1));
    double* A = malloc(sizeof(double) * (nnz + 1));
    memcpy(A, val_in, sizeof(double) * nnz);
    double *b = malloc(sizeof(double) * (m + 1)), *c = malloc(sizeof(double) * (n + 1));
    memcpy(b, b_in, sizeof(double) * m);
    memcpy(c, c_in, sizeof(double) * n);
    double *R = malloc(sizeof(double) * (m + 1)), *C = malloc(sizeof(double) * (n + 1));
    or_ruiz(m, n, rp, ci, A, b, c, R, C);

    int* P = malloc(sizeof(int) * (m + 1));
    if (P_in) memcpy(P, P_in, sizeof(int) * m);
    else or_md_order(m, n, rp, ci, P);

    or_kkt k;
    k.m = m;
    k.P = P;
    unsigned char* hd = malloc((size_t)m + 1);
    for (int i = 0; i < m; ++i) hd[i] = rp[P[i] + 1] > rp[P[i]];
    k.has_diag = hd;
    k.M0 = malloc(sizeof(double) * ((size_t)m * m + 1));
    k.Leq = malloc(sizeof(double) * ((size_t)m * m + 1));
    k.D = malloc(sizeof(double) * (m + 1));
    k.have_eq = 0;

#define NV(name, len) double* name = calloc((size_t)(len) + 1, sizeof(double))
    NV(x, n); NV(y, m); NV(s, n); NV(rpv, m); NV(rd, n); NV(th, n); NV(rxs, n); NV(v, n);
    NV(dy, m); NV(dxa, n); NV(dsa, n); NV(dx, n); NV(ds, n); NV(tx, n); NV(ts, n);
    NV(bx, n); NV(by, m); NV(bs, n); NV(brp, m); NV(brd, n); NV(r1, m); NV(r2, n); NV(r3, n);
#undef NV
    for (int j = 0; j < n; ++j) { x[j] = 1.0; s[j] = 1.0; }
    res->factorizations = 0;
    res->solves = 0;
    res->fail_code = 0;
    res->status = 0;
    res->iterations = 0;
    double rp0 = 1, rd0 = 1, mu0 = 1;
    int rc = OR_OK;
    int max_iter = opt ? opt->max_iter : 200;
    int relmode = opt && opt->stop_rule == 1;
    int done = 0;
    /* OR_TRACE set: one line per iteration on stderr (route: 0 corrector, 1 centred, 3 dual-only,
     * 4 non-monotone centred fallback) */
    const int trace = getenv("OR_TRACE") != NULL;

#define RESID(X, Y, S, RP, RD)                                                   \
    do {                                                                          \
        for (int i_ = 0; i_ < m; ++i_) RP[i_] = -b[i_];                           \
        spmv(m, rp, ci, A, X, RP);                                                \
        for (int j_ = 0; j_ < n; ++j_) RD[j_] = S[j_] - c[j_];                    \
        spmv_t(m, rp, ci, A, Y, RD);                                              \
    } while (0)

    /* verify_newton_direction (mehrotra_device.cu:494-587) */
#define VERIFY(DX, DY, DS, OK)                                                    \
    do {                                                                          \
        memcpy(r1, rpv, sizeof(double) * m);                                      \
        spmv(m, rp, ci, A, DX, r1);                                               \
        for (int j_ = 0; j_ < n; ++j_) r2[j_] = rd[j_] + DS[j_];                  \
        spmv_t(m, rp, ci, A, DY, r2);                                             \
        for (int j_ = 0; j_ < n; ++j_) r3[j_] = s[j_] * DX[j_] + x[j_] * DS[j_] - rxs[j_]; \
        double e1 = nrm2(m, r1) / fmax(1.0, nrp);                                 \
        double e2 = nrm2(n, r2) / fmax(1.0, nrd);                                 \
        double e3 = nrm2(n, r3) / fmax(1.0, nrm2(n, rxs));                        \
        if (relmode) e1 = 0.0;                                                    \
        OK = !(e1 > 1e-8 || e2 > 1e-8 || e3 > 1e-8);                              \
    } while (0)

    /* direction: rkkt = -rp + A v, solve, ds = -rd - A^T dy, dx = -Theta ds + r_xs / s */
#define DIRECTION(DXO, DSO, OK)                                                   \
    do {                                                                          \
        for (int j_ = 0; j_ < n; ++j_) v[j_] = -th[j_] * rd[j_] - rxs[j_] / s[j_]; \
        for (int i_ = 0; i_ < m; ++i_) dy[i_] = -rpv[i_];                         \
        spmv(m, rp, ci, A, v, dy);                                                \
        res->solves++;                                                            \
        OK = kkt_solve(&k, dy, tol_kkt, relmode);                                          \
        for (int j_ = 0; j_ < n; ++j_) DSO[j_] = -rd[j_];                         \
        for (int i_ = 0; i_ < m; ++i_)                                            \
            for (int p_ = rp[i_]; p_ < rp[i_ + 1]; ++p_) DSO[ci[p_]] += -1.0 * A[p_] * dy[i_]; \
        for (int j_ = 0; j_ < n; ++j_) DXO[j_] = -th[j_] * DSO[j_] + rxs[j_] / s[j_]; \
    } while (0)

    for (int iter = 0; iter < max_iter && !done; ++iter) {
        RESID(x, y, s, rpv, rd);
        double nrp = nrm2(m, rpv), nrd = nrm2(n, rd), mu = dot(n, x, s) / n;
        if (iter == 0) { rp0 = fmax(1.0, nrp); rd0 = fmax(1.0, nrd); mu0 = fmax(1.0, mu); }
        int stop;
        if (!opt || opt->stop_rule == 0) {
            double t = opt ? opt->tol : 1e-6;
            stop = nrp < t && nrd < t && mu < t;
        } else {
            double pobj = dot(n, c, x), dobj = dot(m, b, y);
            double gap = fabs(pobj - dobj) / fmax(1.0, fmax(fabs(pobj), fabs(dobj)));
            double rpi = 0, rdi = 0, bi = 0, cinf = 0;
            for (int i = 0; i < m; ++i) { rpi = fmax(rpi, fabs(rpv[i] / R[i])); bi = fmax(bi, fabs(b_in[i])); }
            double xsmax = 0;
            for (int j = 0; j < n; ++j) {
                rdi = fmax(rdi, fabs(rd[j] / C[j])); cinf = fmax(cinf, fabs(c_in[j]));
                xsmax = fmax(xsmax, fabs(x[j] * (s[j] - rd[j])));   /* x_j (c - A^T y)_j */
            }
            /* absolute targets as well (the FP64 certificate checks |Ax-b| and the reduced costs
             * absolutely), floored where FP64 cannot resolve them */
            stop = gap <= opt->gap_tol && rpi / (1 + bi) <= opt->gap_tol && rdi / (1 + cinf) <= opt->gap_tol &&
                   rpi <= fmax(REL_ABS_MARGIN * opt->gap_tol, 1e-12 * (1 + bi)) &&
                   rdi <= fmax(REL_ABS_MARGIN * opt->gap_tol, 1e-12 * (1 + cinf)) &&
                   xsmax <= fmax(REL_ABS_MARGIN * opt->gap_tol, 1e-14 * fmax(1.0, fmax(fabs(pobj), fabs(dobj))));
        }
        if (trace && relmode) {
            double pobj = dot(n, c, x), dobj = dot(m, b, y), xsm = 0;
            for (int j = 0; j < n; ++j) xsm = fmax(xsm, fabs(x[j] * (s[j] - rd[j])));
            fprintf(stderr, "or_ipm %3d gap %.3e xsmax %.3e\n", iter,
                    fabs(pobj - dobj) / fmax(1.0, fmax(fabs(pobj), fabs(dobj))), xsm);
        }
        if (stop) { res->iterations = iter; res->status = 0; done = 1; break; }

        for (int j = 0; j < n; ++j) th[j] = x[j] / s[j];
        or_assemble_dense(m, rp, ci, A, th, P, k.M0);
        res->factorizations++;
        /* the reference's unequilibrated factorization of M0 + 1e-9 I only decides whether its
         * solve throws; the relative-gap mode (an extension) does without it, like the device */
        if (relmode) k.have_eq = 0;
        else if (kkt_factor_check(&k) != OR_OK) { rc = OR_NOT_SPD; res->status = -1; res->iterations = iter; res->fail_code = 1; break; }
        double tol_kkt = 1e-9 * fmax(1.0, nrp);

        /* predictor */
        for (int j = 0; j < n; ++j) rxs[j] = -x[j] * s[j] + 0.0;
        int ok;
        DIRECTION(dxa, dsa, ok);
        if (!ok) { res->status = 3; res->iterations = iter; res->fail_code = 2; break; }
        VERIFY(dxa, dy, dsa, ok);
        if (!ok) { res->status = 3; res->iterations = iter; res->fail_code = 3; break; }
        double ap_aff = fmin(1.0, step_len(n, x, dxa)), ad_aff = fmin(1.0, step_len(n, s, dsa));
        for (int j = 0; j < n; ++j) { tx[j] = x[j] + ap_aff * dxa[j]; ts[j] = s[j] + ad_aff * dsa[j]; }
        double mu_aff = dot(n, tx, ts) / n;
        double sigma = fmin(1.0, pow(fmax(0.0, mu_aff) / fmax(1e-16, mu), 3.0));

        /* corrector */
        for (int j = 0; j < n; ++j) rxs[j] = -x[j] * s[j] + sigma * mu - dxa[j] * dsa[j];
        DIRECTION(dx, ds, ok);
        if (!ok) { res->status = 3; res->iterations = iter; res->fail_code = 4; break; }
        int corr_valid;
        VERIFY(dx, dy, ds, corr_valid);
        double ap = fmin(1.0, 0.995 * step_len(n, x, dx)), ad = fmin(1.0, 0.995 * step_len(n, s, ds));
        double merit0 = nrp / rp0 + nrd / rd0 + mu / mu0;
        /* relative-gap mode: a residual already inside its absolute target must stay inside */
        double rpi0 = 0, rdi0 = 0, ptar = 0, dtar = 0;
        if (relmode) {
            double bi = 0, ci = 0;
            for (int i = 0; i < m; ++i) { rpi0 = fmax(rpi0, fabs(rpv[i] / R[i])); bi = fmax(bi, fabs(b_in[i])); }
            for (int j = 0; j < n; ++j) { rdi0 = fmax(rdi0, fabs(rd[j] / C[j])); ci = fmax(ci, fabs(c_in[j])); }
            ptar = fmax(REL_ABS_MARGIN * opt->gap_tol, 1e-12 * (1 + bi));
            dtar = fmax(REL_ABS_MARGIN * opt->gap_tol, 1e-12 * (1 + ci));
        }
#define KEEPS_TARGETS(OKV)                                                        \
        do {                                                                      \
            OKV = 1;                                                              \
            if (relmode) {                                                        \
                double rpt_ = 0, rdt_ = 0;                                        \
                for (int i_ = 0; i_ < m; ++i_) rpt_ = fmax(rpt_, fabs(brp[i_] / R[i_])); \
                for (int j_ = 0; j_ < n; ++j_) rdt_ = fmax(rdt_, fabs(brd[j_] / C[j_])); \
                if ((rpi0 <= ptar && rpt_ > ptar) || (rdi0 <= dtar && rdt_ > dtar)) OKV = 0; \
            }                                                                     \
        } while (0)
        int accepted = 0, route = -1;
        double tr_ap = 0, tr_ad = 0, acc_merit = 0;
        for (int dir = corr_valid ? 0 : 1; dir < 2 && !accepted; ++dir) {
            if (dir == 1) {
                for (int j = 0; j < n; ++j) rxs[j] = -x[j] * s[j] + 0.5 * mu;
                DIRECTION(dx, ds, ok);
                if (!ok) break;
                int vok;
                VERIFY(dx, dy, ds, vok);
                if (!vok) break;
                double bnd = fmin(step_len(n, x, dx), step_len(n, s, ds));
                ap = ad = fmin(1.0, 0.995 * bnd);
            }
            int maxbt = dir == 0 ? 15 : 30;
            int fallback_bt = -1;   /* relative mode: first finite, positive centred trial that
                                       keeps the met residual targets */
            for (int bt = 0; bt < maxbt; ++bt) {
                for (int j = 0; j < n; ++j) { bx[j] = x[j] + ap * dx[j]; bs[j] = s[j] + ad * ds[j]; }
                for (int i = 0; i < m; ++i) by[i] = y[i] + ad * dy[i];
                RESID(bx, by, bs, brp, brd);
                double nrp2 = nrm2(m, brp), nrd2 = nrm2(n, brd), mu2 = dot(n, bx, bs) / n;
                double mx = vmin(n, bx), ms = vmin(n, bs);
                double merit = nrp2 / rp0 + nrd2 / rd0 + mu2 / mu0;
                double st = fmin(ap, ad);
                int finite = isfinite(nrp2) && isfinite(nrd2) && isfinite(mu2) && isfinite(merit) &&
                             isfinite(merit0) && isfinite(mx) && isfinite(ms) && isfinite(ap) &&
                             isfinite(ad) && isfinite(nrm2(m, by));
                int pos = mx > 0.0 && ms > 0.0 && mu2 > 0.0 && st > 0.0;
                int prog = merit < merit0 && merit <= (1.0 - 1e-4 * st) * merit0;
                int keeps;
                KEEPS_TARGETS(keeps);
                prog = prog && keeps;
                if (finite && pos && prog) { accepted = 1; route = dir; tr_ap = ap; tr_ad = ad; acc_merit = merit; break; }
                if (relmode && dir == 1 && finite && pos && keeps && fallback_bt < 0) fallback_bt = bt;
                ap *= 0.5;
                ad *= 0.5;
            }
            /* relative-gap mode: a corrector step shorter than SHORT_STEP taken while both
             * residuals are inside their targets competes with the dual-only trials (the better
             * merit wins) */
            const int short_step = relmode && dir == 0 && accepted && fmin(tr_ap, tr_ad) < SHORT_STEP &&
                                   rpi0 <= ptar && rdi0 <= dtar;
            if (short_step) {        /* keep the corrector's point in case the dual-only loses */
                memcpy(tx, bx, sizeof(double) * n);
                memcpy(ts, bs, sizeof(double) * n);
                memcpy(r1, by, sizeof(double) * m);
            }
            if ((!accepted || short_step) && relmode && dir == 0) {
                /* relative-gap mode: dual-only trials of the same direction (the primal point,
                 * and with it the primal residual, stays), ad halving from 0.995 of its bound */
                double ad2 = fmin(1.0, 0.995 * step_len(n, s, ds));
                int took = 0;
                for (int bt = 0; bt < 15 && !took; ++bt) {
                    for (int j = 0; j < n; ++j) { bx[j] = x[j]; bs[j] = s[j] + ad2 * ds[j]; }
                    for (int i = 0; i < m; ++i) by[i] = y[i] + ad2 * dy[i];
                    RESID(bx, by, bs, brp, brd);
                    double nrp2 = nrm2(m, brp), nrd2 = nrm2(n, brd), mu2 = dot(n, bx, bs) / n;
                    double ms = vmin(n, bs);
                    double merit = nrp2 / rp0 + nrd2 / rd0 + mu2 / mu0;
                    int finite = isfinite(nrp2) && isfinite(nrd2) && isfinite(mu2) && isfinite(merit) &&
                                 isfinite(merit0) && isfinite(ms) && isfinite(ad2) && isfinite(nrm2(m, by));
                    int pos = ms > 0.0 && mu2 > 0.0 && ad2 > 0.0;
                    int prog = merit < merit0 && merit <= (1.0 - 1e-4 * ad2) * merit0;
                    int keeps;
                    KEEPS_TARGETS(keeps);
                    if (finite && pos && prog && keeps && (!short_step || merit < acc_merit)) {
                        took = accepted = 1; route = 3; tr_ap = 0; tr_ad = ad2;
                    } else ad2 *= 0.5;
                }
                if (short_step && !took) {
                    memcpy(bx, tx, sizeof(double) * n);
                    memcpy(bs, ts, sizeof(double) * n);
                    memcpy(by, r1, sizeof(double) * m);
                }
            }
            if (!accepted && relmode && dir == 1 && fallback_bt >= 0) {
                /* non-monotone fallback: take the longest admissible centred step */
                double a2 = fmin(1.0, 0.995 * fmin(step_len(n, x, dx), step_len(n, s, ds)));
                for (int t = 0; t < fallback_bt; ++t) a2 *= 0.5;
                for (int j = 0; j < n; ++j) { bx[j] = x[j] + a2 * dx[j]; bs[j] = s[j] + a2 * ds[j]; }
                for (int i = 0; i < m; ++i) by[i] = y[i] + a2 * dy[i];
                accepted = 1;
                route = 4;
                tr_ap = tr_ad = a2;
            }
        }
        if (trace)
            fprintf(stderr, "or_ipm %3d |rp| %.3e |rd| %.3e mu %.3e rp_inf %.3e sigma %.3e route %d ap %.3e ad %.3e\n",
                    iter, nrp, nrd, mu, rpi0, sigma, route, tr_ap, tr_ad);
        if (!accepted) { res->status = 3; res->iterations = iter; res->fail_code = 5; break; }
        memcpy(x, bx, sizeof(double) * n);
        memcpy(y, by, sizeof(double) * m);
        memcpy(s, bs, sizeof(double) * n);
        if (iter == max_iter - 1) { res->status = 3; res->iterations = max_iter; }
    }
    if (!done && res->status == 0) { res->status = 3; res->iterations = max_iter; }

    RESID(x, y, s, rpv, rd);
    res->primal_residual = nrm2(m, rpv);
    res->dual_residual = nrm2(n, rd);
    res->duality_gap = dot(n, x, s) / n;
    res->objective = dot(n, c, x);
    res->dual_objective = dot(m, b, y);
    {
        double pobj = res->objective, dobj = res->dual_objective;
        res->rel_gap = fabs(pobj - dobj) / fmax(1.0, fmax(fabs(pobj), fabs(dobj)));
        double rpi = 0, rdi = 0, bi = 0, cinf = 0;
        for (int i = 0; i < m; ++i) { rpi = fmax(rpi, fabs(rpv[i] / R[i])); bi = fmax(bi, fabs(b_in[i])); }
        for (int j = 0; j < n; ++j) { rdi = fmax(rdi, fabs(rd[j] / C[j])); cinf = fmax(cinf, fabs(c_in[j])); }
        res->rel_primal_inf = rpi / (1 + bi);
        res->rel_dual_inf = rdi / (1 + cinf);
    }
    for (int j = 0; j < n; ++j) {
        if (x_out) x_out[j] = x[j] * C[j];
        if (s_out) s_out[j] = s[j] / C[j];
    }
    for (int i = 0; i < m; ++i)
        if (y_out) y_out[i] = y[i] * R[i];

    free(x); free(y); free(s); free(rpv); free(rd); free(th); free(rxs); free(v); free(dy);
    free(dxa); free(dsa); free(dx); free(ds); free(tx); free(ts); free(bx); free(by); free(bs);
    free(brp); free(brd); free(r1); free(r2); free(r3);
    free(k.M0); free(k.Leq); free(k.D); free(hd); free(P); free(rp); free(A); free(b); free(c);
    free(R); free(C);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
int or_verify_optimal(int m, int n, const int* colptr, const int* rowidx, const double* val,
                      const double* obj, const double* rhs, const double* lb, const double* ub,
                      int maximize, const double* x, const double* pi, double tol,
                      double* out_primal_residual) {
    double sign = maximize ? -1.0 : 1.0;
    for (int j = 0; j < n; ++j) {
        if (lb[j] > -INFINITY && x[j] <= lb[j] - tol) return 0;
        if (ub[j] < INFINITY && x[j] >= ub[j] + tol) return 0;
    }
    double* ax = calloc((size_t)m + 1, sizeof(double));
    for (int j = 0; j < n; ++j)
        for (int k = colptr[j]; k < colptr[j + 1]; ++k) {
            if (rowidx[k] < 0 || rowidx[k] >= m) { free(ax); return 0; }
            ax[rowidx[k]] += val[k] * x[j];
        }
    double worst = 0.0;
    for (int i = 0; i < m; ++i) {
        double r = fabs(ax[i] - rhs[i]);
        if (r > worst) worst = r;
        if (r >= tol) { free(ax); return 0; }
    }
    free(ax);
    if (out_primal_residual) *out_primal_residual = worst;
    for (int j = 0; j < n; ++j) {
        double atpi = 0.0;
        for (int k = colptr[j]; k < colptr[j + 1]; ++k) atpi += val[k] * pi[rowidx[k]];
        double r = sign * obj[j] - atpi;
        if (r >= tol) {
            if (lb[j] <= -INFINITY) return 0;
            if (r * (x[j] - lb[j]) >= tol) return 0;
        } else if (r <= -tol) {
            if (ub[j] >= INFINITY) return 0;
            if ((-r) * (ub[j] - x[j]) >= tol) return 0;
        }
    }
    return 1;
}

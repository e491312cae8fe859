// Batched dense interior-point kernels for gfx950: thousands of small LPs in one persistent
// launch, one 256-thread workgroup per LP at a time.
//
// Per LP the workgroup runs the Mehrotra predictor-corrector loop of the single-LP solver (the
// reference loop src/ipm/mehrotra_device.cu:695-1047 with the same constants, in the
// relative-gap or the reference stop mode), entirely on the device:
//   * 10 Jacobi Ruiz passes (mehrotra_device.cu:286-323) over both the CSR and the CSC copy,
//     with the same per-entry products, so the two copies stay bitwise equal;
//   * fused sparse passes (G lanes per row of A or of A^T) with workgroup reductions in a
//     fixed order: residuals, right-hand sides, directions, Newton-direction checks;
//   * M = D (A Theta A^T) D + delta I assembled densely, one column per wavefront: the column
//     is accumulated in LDS from the CSC segments at or below the diagonal of every column of
//     A that row i touches (flattened over the lanes, so loads stay coalesced);
//   * a left-looking blocked Cholesky on MFMA (64-column block columns, 128-row tiles, the
//     GEMM / potrf64 primitives of dense_tile.hpp), FP64 or FP32;
//   * blocked forward / backward substitution through LDS;
//   * PCG on M0 dy = rhs (preconditioned by the factor) or the reference's refinement.
// The launch is persistent: a workgroup takes tickets from ctl[0] until all LPs are visited;
// every visit runs up to iters_per_visit iterations and leaves the LP's state in lkb_state, so
// the same launch can be repeated (for instance from a captured graph) until all LPs finish.
#include <mutex>
#include "dense_tile.hpp"

// The no-allocation entry point's device memory: a fixed array (each GPU holds its own copy).
// Host code never uses its name as an address: lkd_batch_scratch resolves the device address
// once per device (hipGetSymbolAddress with that device current) and every host use -- kernel
// arguments, copies -- goes through that pointer.
static __device__ __align__(256) unsigned char scratch[LKB_SCRATCH_BYTES];

namespace {

constexpr int BT = 256;
constexpr int GR = 16;                 // lanes per row of A in the row passes
constexpr int GC = 8;                  // lanes per column of A in the column passes
constexpr int LDS_D = 6656;            // doubles in the workgroup's shared union
static_assert(sizeof(GemmLds<double>) == LDS_D * sizeof(double), "GEMM staging fits the union");
static_assert(sizeof(GemmLds<float>) <= LDS_D * sizeof(double), "GEMM staging fits the union");
static_assert(64 * DP + 1024 <= LDS_D, "diagonal block and inverses fit the union");
static_assert(4 * LKB_MAX_M <= LDS_D, "one assembly column per wavefront");
static_assert(LKB_MAX_M + 64 * DP + 64 + 4 * 64 <= LDS_D, "substitution buffers");
static_assert(LKB_STAGE == 64 * DP + 1024, "staged diagonal block");

// The workgroup's shared memory, at file scope so that the large phases can be separate
// (non-inlined) functions with their own register allocation and still address it as LDS:
// one union (GEMM staging / diagonal block / assembly columns / substitution buffers), the
// reduction scratch and a flag word.
__shared__ double g_lds[LDS_D];
__shared__ double g_red[4 * 10];
__shared__ int g_flag;

// n-vectors and m-vectors of one LP in its work area
enum { VX = 0, VS, VC, VCS, VRD, VTH, VRXS, VDXA, VDSA, VDX, VDS, VQD, VTN, VV };
enum { WY = 0, WB, WR, WRP, WDY, WRHS, WD, WQP, WPR, WPZ, WPP, WPQ, WBEST, WTM };
static_assert(VV + 1 == LKB_NVN && WTM + 1 == LKB_NVM, "work vector counts");

enum { RSUM = 0, RMIN = 1, RMAX = 2 };
__device__ __forceinline__ double rop(int op, double a, double b) {
    return op == RSUM ? a + b : (op == RMIN ? fmin(a, b) : fmax(a, b));
}

struct Lp {
    int m, n, nnz, ld;
    const int32_t *rp, *ci, *pos, *cp, *ri;
    const double* v0;
    double *av, *cv;
    const double *b0, *c0;
    double *vn, *vm, *stage;
    double* fac;
    __device__ __forceinline__ double* N(int k) const { return vn + (int64_t)k * n; }
    __device__ __forceinline__ double* M(int k) const { return vm + (int64_t)k * m; }
};

__device__ __forceinline__ Lp make_lp(const lkb_args& A, const lkb_lp& d) {
    Lp L;
    L.m = d.m;
    L.n = d.n;
    L.nnz = d.nnz;
    L.ld = d.ld;
    const int32_t* ip = A.idx + d.a;
    L.rp = ip;
    L.ci = ip + d.m + 1;
    L.pos = L.ci + d.nnz;
    L.cp = L.pos + d.nnz;
    L.ri = L.cp + d.n + 1;
    L.v0 = A.val0 + d.v;
    L.av = A.valw + 2 * d.v;
    L.cv = L.av + d.nnz;
    L.b0 = A.rhs + d.bc;
    L.c0 = L.b0 + d.m;
    L.vn = A.work + d.w;
    L.vm = L.vn + (int64_t)LKB_NVN * d.n;
    L.stage = L.vm + (int64_t)LKB_NVM * d.m;
    L.fac = static_cast<double*>(A.fac) + d.f;
    return L;
}

// K values reduced over the workgroup in a fixed order; every thread receives the results.
template <int K>
__device__ __forceinline__ void wg_reduce(double (&v)[K], const int (&op)[K], double* red) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[q] = rop(op[q], v[q], __shfl_xor(v[q], o, 64));
    if (l == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) red[w * K + q] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q)
        v[q] = rop(op[q], rop(op[q], red[q], red[K + q]), rop(op[q], red[2 * K + q], red[3 * K + q]));
    __syncthreads();
}

template <int G>
__device__ __forceinline__ double group_op(int op, double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = rop(op, v, __shfl_xor(v, o, 64));
    return v;
}

// Segments [ptr[r], ptr[r+1]) of a compressed matrix, G lanes per segment: the lanes combine
// term(p) with `op` and the first lane of the group calls f(r, value).  Every thread runs the
// same number of rounds (the group reductions are wave-wide shuffles).
template <int G, typename Term, typename F>
__device__ __forceinline__ void seg_pass(int nseg, const int32_t* __restrict__ ptr, int op,
                                         Term&& term, F&& f) {
    const int g = threadIdx.x & (G - 1);
    const double ident = op == RSUM ? 0.0 : (op == RMIN ? 1e300 : 0.0);
    for (int r0 = 0; r0 < nseg; r0 += BT / G) {
        const int r = r0 + (int)threadIdx.x / G;
        double s = ident;
        if (r < nseg) {
            // four independent partial results per lane keep four gathers in flight
            const int e = ptr[r + 1];
            int p = ptr[r] + g;
            double s1 = ident, s2 = ident, s3 = ident;
            for (; p + 3 * G < e; p += 4 * G) {
                const double t0 = term(p), t1 = term(p + G), t2 = term(p + 2 * G), t3 = term(p + 3 * G);
                s = rop(op, s, t0);
                s1 = rop(op, s1, t1);
                s2 = rop(op, s2, t2);
                s3 = rop(op, s3, t3);
            }
            for (; p < e; p += G) s = rop(op, s, term(p));
            s = rop(op, rop(op, s, s1), rop(op, s2, s3));
        }
        s = group_op<G>(op, s);
        if (r < nseg && g == 0) f(r, s);
    }
}

// f(r, p) for every entry p of every segment r, G lanes per segment
template <int G, typename F>
__device__ __forceinline__ void seg_for(int nseg, const int32_t* __restrict__ ptr, F&& f) {
    const int g = threadIdx.x & (G - 1);
    for (int r = (int)threadIdx.x / G; r < nseg; r += BT / G) {
        const int e = ptr[r + 1];
        for (int p = ptr[r] + g; p < e; p += G) f(r, p);
    }
}

__device__ __forceinline__ double step_len(double v, double dv) {
    return dv >= 0.0 ? 1e30 : -v / dv;
}

// Wave-uniform copies of values that arrive in vector registers at a call boundary (function
// arguments are passed in VGPRs): in scalar registers the address arithmetic derived from them
// stays scalar instead of occupying (and spilling) vector registers.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename P>
__device__ __forceinline__ P* uni(P* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<P*>(((uint64_t)hi << 32) | lo);
}

// potrf64 on the diagonal block at the start of the shared union (Ds, then Lv), as a call of
// its own: it takes the whole register file, which the tile code around it then keeps
template <typename T>
__device__ __attribute__((noinline)) bool potrf_lds() {
    T* Ds = reinterpret_cast<T*>(g_lds);
    return potrf64<T>(Ds, Ds + 64 * DP, &g_flag);
}

// ------------------------------------------------------------------ dense factor
// Left-looking blocked Cholesky of the m x m lower matrix at F (column-major, ld), in place,
// by 64-column block columns k = columns k0 .. k0+kb:
//   diagonal block   C(k, k) - L(k, 0:k0) L(k, 0:k0)^T on MFMA (waves 0 and 1 hold its rows),
//                    factored in LDS (potrf64), stored, and staged in the global slot `stage`
//                    (the Ds + inverses layout of the shared union)
//   128-row tiles R  below it: X = C(R, k) - L(R, 0:k0) L(k, 0:k0)^T (gemm_tn: accumulators in
//                    the solve layout), then X <- X Lkk^{-T} on MFMA, 16 columns at a time
// Returns false (uniformly) when a pivot is not positive.
template <typename T>
__device__ __attribute__((noinline)) bool dense_chol(T* __restrict__ F_, int m_, int ld_,
                                                     T* __restrict__ stage_) {
    using R = Mfma<T>;
    T* __restrict__ F = uni(F_);
    T* __restrict__ stage = uni(stage_);
    const int m = uni(m_), ld = uni(ld_);
    GemmLds<T>& sm = *reinterpret_cast<GemmLds<T>*>(g_lds);
    T* Ds = reinterpret_cast<T*>(g_lds);
    T* Lv = Ds + 64 * DP;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15;
    const int rl0 = R::row(l, 0);
    for (int k0 = 0; k0 < m; k0 += TS) {
        const int kb = min(TS, m - k0), kbm = kb - 1;
        {   // ---- the diagonal block
            Acc<T> acc[4][2];
            zero_acc<T, 4, 2>(acc);
            gemm_tn<T, true>(F + k0, ld, kb, F + k0, ld, kb, 0, k0, sm, acc);
            __syncthreads();
            if (w < 2) {
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int rl = 32 * w + 16 * b + li;
                    const int rc = min(rl, kbm);
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int c = 16 * p + R::row(l, e);
                            const T cv = F[(int64_t)(k0 + min(c, kbm)) * ld + k0 + rc] - acc[p][b][e];
                            T v = rl == c ? T(1) : T(0);
                            if (rl < kb && c < kb) v = rl >= c ? cv : T(0);
                            Ds[c * DP + rl] = v;
                        }
                }
            }
            __syncthreads();
            if (!potrf_lds<T>()) return false;
            for (int idx = t; idx < LKB_STAGE; idx += BT) stage[idx] = Ds[idx];
            for (int idx = t; idx < TS * TS; idx += BT) {
                const int r = idx & 63, c = idx >> 6;
                if (r < kb && c <= r) F[(int64_t)(k0 + c) * ld + k0 + r] = Ds[c * DP + r];
            }
            __syncthreads();
        }
        for (int r0 = k0 + kb; r0 < m; r0 += MT) {   // ---- the tiles below it
            const int ra = min(MT, m - r0);
            Acc<T> acc[4][2];
            zero_acc<T, 4, 2>(acc);
            gemm_tn<T, true>(F + r0, ld, ra, F + k0, ld, kb, 0, k0, sm, acc);
            Acc<T> y[2][4];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int rl = 32 * w + 16 * b + li;
                const int64_t rc = rl < ra ? (int64_t)r0 + rl : (int64_t)r0;
                if (kb == TS) {
                    // full block column: one base per lane, wave-uniform element offsets
                    const T* pb = F + (int64_t)(k0 + rl0) * ld + rc;
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            y[b][p][e] = pb[(int64_t)(16 * p + R::row(l, e) - rl0) * ld] - acc[p][b][e];
                } else {
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int c = min(16 * p + R::row(l, e), kbm);   // padding columns
                            y[b][p][e] = F[(int64_t)(k0 + c) * ld + rc] - acc[p][b][e];
                        }
                }
            }
            // the staged factor of block k into LDS (the GEMM used the space)
            for (int idx = t; idx < LKB_STAGE; idx += BT) Ds[idx] = stage[idx];
            __syncthreads();
            // Y_p = Linv_pp (X_p^T - sum_{q<p} L_pq Y_q)
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                Acc<T> tp[2] = {y[0][p], y[1][p]};
#pragma unroll
                for (int q = 0; q < p; ++q)
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const T a = -Ds[(16 * q + R::row(l, s)) * DP + 16 * p + li];
#pragma unroll
                        for (int b = 0; b < 2; ++b) tp[b] = R::mma(a, y[b][q][s], tp[b]);
                    }
#pragma unroll
                for (int b = 0; b < 2; ++b) y[b][p] = Acc<T>{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const T a = Lv[256 * p + R::row(l, s) * 16 + li];
#pragma unroll
                    for (int b = 0; b < 2; ++b) y[b][p] = R::mma(a, tp[b][s], y[b][p]);
                }
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int rl = 32 * w + 16 * b + li;
                if (rl >= ra) continue;
                T* out = F + (int64_t)(k0 + rl0) * ld + r0 + rl;
#pragma unroll
                for (int p = 0; p < 4; ++p)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int c = 16 * p + R::row(l, e);
                        if (c < kb) out[(int64_t)(c - rl0) * ld] = y[b][p][e];
                    }
            }
            __syncthreads();   // LDS reads done before the next tile's GEMM; stores issued
        }
    }
    return true;
}

// x <- (L L^T)^{-1} x for the factor at F (lower, column-major, ld), x an m-vector in global
// memory; blocked by 64 through LDS (z, one diagonal block, reciprocal pivots, partial sums).
template <typename T>
__device__ __attribute__((noinline)) void dense_solve(const T* __restrict__ F_, int m_, int ld_,
                                                      double* __restrict__ x_) {
    const T* __restrict__ F = uni(F_);
    double* __restrict__ x = uni(x_);
    const int m = uni(m_), ld = uni(ld_);
    double* z = g_lds;
    double* Db = g_lds + LKB_MAX_M;
    double* rinv = Db + 64 * DP;
    double* part = rinv + 64;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    for (int i = t; i < m; i += BT) z[i] = x[i];
    __syncthreads();
    auto stage_block = [&](int jb, int kb) {
        for (int idx = t; idx < 64 * 64; idx += BT) {
            const int r = idx & 63, c = idx >> 6;
            if (r < kb && c <= r) Db[c * DP + r] = (double)F[(int64_t)(jb + c) * ld + jb + r];
        }
        __syncthreads();
        if (t < kb) rinv[t] = 1.0 / Db[t * DP + t];
    };
    // forward: L z = x
    for (int jb = 0; jb < m; jb += 64) {
        const int kb = min(64, m - jb);
        {   // part[g][r] = sum over columns c = g, g + 4, .. < jb of L(jb + r, c) z[c]
            // (four columns per round in flight, rows past kb read row kb - 1 and are unused)
            const int r = min(l, kb - 1), g = w;
            const T* col = F + jb + r;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int c = g;
            for (; c + 12 < jb; c += 16) {
                const double a0 = (double)col[(int64_t)c * ld], a1 = (double)col[(int64_t)(c + 4) * ld];
                const double a2 = (double)col[(int64_t)(c + 8) * ld], a3 = (double)col[(int64_t)(c + 12) * ld];
                s0 += a0 * z[c];
                s1 += a1 * z[c + 4];
                s2 += a2 * z[c + 8];
                s3 += a3 * z[c + 12];
            }
            for (; c < jb; c += 4) s0 += (double)col[(int64_t)c * ld] * z[c];
            part[g * 64 + l] = (s0 + s1) + (s2 + s3);
        }
        stage_block(jb, kb);
        __syncthreads();
        if (w == 0) {
            double v = l < kb ? z[jb + l] - ((part[l] + part[64 + l]) + (part[128 + l] + part[192 + l])) : 0.0;
            for (int j = 0; j < kb; ++j) {
                if (l == j) v *= rinv[j];
                const double zj = __shfl(v, j, 64);
                if (l > j && l < kb) v -= Db[j * DP + l] * zj;
            }
            if (l < kb) z[jb + l] = v;
        }
        __syncthreads();
    }
    // backward: L^T x = z
    const int nblk = (m + 63) / 64;
    for (int bi = nblk - 1; bi >= 0; --bi) {
        const int jb = 64 * bi, kb = min(64, m - jb), r1 = jb + kb;
        // part[c] = sum over rows r >= r1 of L(r, jb + c) x[r], column c on wave c % 4
        for (int c = w; c < kb; c += 4) {
            const T* col = F + (int64_t)(jb + c) * ld;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int r = r1 + l;
            for (; r + 192 < m; r += 256) {
                const double a0 = (double)col[r], a1 = (double)col[r + 64];
                const double a2 = (double)col[r + 128], a3 = (double)col[r + 192];
                s0 += a0 * z[r];
                s1 += a1 * z[r + 64];
                s2 += a2 * z[r + 128];
                s3 += a3 * z[r + 192];
            }
            for (; r < m; r += 64) s0 += (double)col[r] * z[r];
            const double s = wave_sum((s0 + s1) + (s2 + s3));
            if (l == 0) part[c] = s;
        }
        stage_block(jb, kb);
        __syncthreads();
        if (w == 0) {
            double v = l < kb ? z[jb + l] - part[l] : 0.0;
            for (int j = kb - 1; j >= 0; --j) {
                if (l == j) v *= rinv[j];
                const double xj = __shfl(v, j, 64);
                if (l < j) v -= Db[l * DP + j] * xj;
            }
            if (l < kb) z[jb + l] = v;
        }
        __syncthreads();
    }
    for (int i = t; i < m; i += BT) x[i] = z[i];
    __syncthreads();
}

// ------------------------------------------------------------------ the per-LP solver
struct Ctx {
    const lkb_args& A;
    Lp L;
    lkb_state st;
    double* red;
    bool rel;
    bool fp32;        // the current factor is FP32
    double m0max;     // max_i M0_ii of the current Theta
    long long mark;   // wall clock at the end of the last timed phase
    // the time since the last mark goes to phase slot `slot` of the LP's state
    __device__ __forceinline__ void tick(double& slot) {
        const long long now = wall_clock64();
        slot += (double)(now - mark);
        mark = now;
    }
};

// D_i = 1 / sqrt(max(|M0_ii|, 1e-12)), M0_ii = sum_k a_ik^2 theta_k (src/cuda/kkt.cu:75 floor)
__device__ void diag_scaling(Ctx& c) {
    const Lp& L = c.L;
    const double* th = L.N(VTH);
    double* D = L.M(WD);
    double mx = 0.0;
    seg_pass<GR>(L.m, L.rp, RSUM,
                 [&](int p) { const double a = L.av[p]; return a * a * th[L.ci[p]]; },
                 [&](int i, double d) {
                     mx = fmax(mx, d);
                     D[i] = 1.0 / sqrt(fmax(fabs(d), 1e-12));
                 });
    double v[1] = {mx};
    const int op[1] = {RMAX};
    wg_reduce<1>(v, op, c.red);
    c.m0max = v[0];
}

// The lower part of D M0 D + delta I (delta on rows that are not empty) into the factor
// storage, one column per wavefront accumulated in LDS.
template <typename T>
__device__ __attribute__((noinline)) void assemble(Ctx& c, double delta) {
    const Lp& L = c.L;
    const double* th = L.N(VTH);
    const double* D = L.M(WD);
    T* F = reinterpret_cast<T*>(L.fac);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double* buf = g_lds + w * LKB_MAX_M;
    for (int i = w; i < L.m; i += 4) {
        for (int r = i + l; r < L.m; r += 64) buf[r] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int e0 = L.rp[i], e1 = L.rp[i + 1];
        for (int eb = e0; eb < e1; eb += 64) {
            const int e = eb + l;
            int p0 = 0, len = 0;
            double coef = 0.0;
            if (e < e1) {
                const int k = L.ci[e];
                p0 = L.pos[e];
                len = L.cp[k + 1] - p0;
                coef = L.av[e] * th[k];
            }
            int incl = len;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(incl, o, 64);
                if (l >= o) incl += u;
            }
            const int excl = incl - len;
            const int total = __shfl(incl, 63, 64);
            // four rounds of 64 flattened entries at a time: all their loads are issued before
            // the first accumulation (the LDS adds of one wavefront stay in program order)
            for (int f0 = 0; f0 < total; f0 += 256) {
                int q[4];
                double cj[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int fq = min(f0 + 64 * u + l, total - 1);
                    int j = 0;
#pragma unroll
                    for (int b = 32; b > 0; b >>= 1)
                        if (__shfl(incl, j + b - 1, 64) <= fq) j += b;
                    q[u] = __shfl(p0, j, 64) + (fq - __shfl(excl, j, 64));
                    cj[u] = __shfl(coef, j, 64);
                }
                int rr[4];
                double vv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    rr[u] = L.ri[q[u]];
                    vv[u] = L.cv[q[u]];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (f0 + 64 * u + l < total) atomicAdd(&buf[rr[u]], vv[u] * cj[u]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const bool has_diag = e1 > e0;
        const double di = D[i];
        for (int r = i + l; r < L.m; r += 64) {
            double v = buf[r] * D[r] * di;
            if (r == i && has_diag) v += delta;
            F[(int64_t)i * L.ld + r] = (T)v;
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
}

// Factor D M0 D + delta I for the current Theta in precision T: delta = delta0 * 10^r for the
// first r that succeeds (src/cuda/kkt.cu:652-690).  false when every delta fails.
template <typename T>
__device__ bool factor(Ctx& c, double delta0, int retries) {
    double delta = delta0;
    for (int r = 0; r < retries; ++r) {
        if (r > 0) delta *= 10.0;
        c.tick(c.st.t_other);
        assemble<T>(c, delta);
        c.tick(c.st.t_assemble);
        const bool ok = dense_chol<T>(reinterpret_cast<T*>(c.L.fac), c.L.m, c.L.ld,
                                      reinterpret_cast<T*>(c.L.stage));
        c.tick(c.st.t_factor);
        if (ok) return true;
        __syncthreads();
    }
    return false;
}

__device__ bool factor_any(Ctx& c, bool fp32) {
    c.st.factorizations++;
    if (fp32) {
        c.st.fp32_factorizations++;
        c.fp32 = true;
        if (factor<float>(c, c.rel ? 1e-6 : 1e-9, c.rel ? 10 : 5)) return true;
    }
    c.fp32 = false;
    return factor<double>(c, c.rel ? 1e-14 : 1e-9, c.rel ? 10 : 5);
}

// x <- D (L L^T)^{-1} D x  (the preconditioner / direct solve, permutation-free)
__device__ void precond(Ctx& c, double* x) {
    const Lp& L = c.L;
    const double* D = L.M(WD);
    for (int i = threadIdx.x; i < L.m; i += BT) x[i] *= D[i];
    __syncthreads();
    c.tick(c.st.t_other);
    if (c.fp32) dense_solve<float>(reinterpret_cast<const float*>(L.fac), L.m, L.ld, x);
    else dense_solve<double>(L.fac, L.m, L.ld, x);
    c.tick(c.st.t_solve);
    for (int i = threadIdx.x; i < L.m; i += BT) x[i] *= D[i];
    __syncthreads();
}

// r = rhs - M0 v with M0 v = A (Theta (A^T v)); returns ||r||_inf
__device__ double m0_residual(Ctx& c, const double* v, const double* rhs, double* r) {
    const Lp& L = c.L;
    const double* th = L.N(VTH);
    double* tn = L.N(VTN);
    seg_pass<GC>(L.n, L.cp, RSUM, [&](int q) { return L.cv[q] * v[L.ri[q]]; },
                 [&](int j, double s) { tn[j] = th[j] * s; });
    __syncthreads();
    double mx = 0.0;
    seg_pass<GR>(L.m, L.rp, RSUM, [&](int p) { return L.av[p] * tn[L.ci[p]]; },
                 [&](int i, double s) {
                     const double ri = rhs[i] - s;
                     r[i] = ri;
                     mx = fmax(mx, fabs(ri));
                 });
    double vv[1] = {mx};
    const int op[1] = {RMAX};
    wg_reduce<1>(vv, op, c.red);
    return vv[0];
}

// q = M0 p; returns p^T q
__device__ double m0_apply(Ctx& c, const double* p, double* q) {
    const Lp& L = c.L;
    const double* th = L.N(VTH);
    double* tn = L.N(VTN);
    seg_pass<GC>(L.n, L.cp, RSUM, [&](int k) { return L.cv[k] * p[L.ri[k]]; },
                 [&](int j, double s) { tn[j] = th[j] * s; });
    __syncthreads();
    double pq = 0.0;
    seg_pass<GR>(L.m, L.rp, RSUM, [&](int k) { return L.av[k] * tn[L.ci[k]]; },
                 [&](int i, double s) {
                     q[i] = s;
                     pq += p[i] * s;
                 });
    double vv[1] = {pq};
    const int op[1] = {RSUM};
    wg_reduce<1>(vv, op, c.red);
    return vv[0];
}

__device__ void copy_m(const Ctx& c, double* dst, const double* src) {
    for (int i = threadIdx.x; i < c.L.m; i += BT) dst[i] = src[i];
    __syncthreads();
}

// KKT solve M0 dy = rhs (src/cuda/kkt.cu:620-870 contract: success iff ||rhs - M0 dy||_inf <=
// tol; relative-gap mode and FP32 factors: PCG from the direct solve with backward-error
// acceptance in relative-gap mode; reference mode: refinement, <= max_refine steps with
// rollback on stagnation).  dy overwritten even on failure.
__device__ bool kkt_solve(Ctx& c, double tol, int max_it) {
    const Lp& L = c.L;
    const double* rhs = L.M(WRHS);
    double* dy = L.M(WDY);
    double* r = L.M(WPR);
    double* z = L.M(WPZ);
    double* p = L.M(WPP);
    double* q = L.M(WPQ);
    double* best = L.M(WBEST);
    copy_m(c, dy, rhs);
    precond(c, dy);
    double before = m0_residual(c, dy, rhs, r);
    c.st.solves++;
    if (before <= tol) return true;
    if (!(c.rel || c.fp32)) {
        for (int it = 0; it < max_it; ++it) {
            copy_m(c, best, dy);
            copy_m(c, z, r);
            precond(c, z);
            for (int i = threadIdx.x; i < L.m; i += BT) dy[i] += z[i];
            __syncthreads();
            c.st.pcg_steps++;
            const double after = m0_residual(c, dy, rhs, r);
            if (after <= tol) return true;
            if (after >= before) {
                copy_m(c, dy, best);
                m0_residual(c, dy, rhs, r);
                return false;
            }
            before = after;
        }
        return false;
    }
    // PCG, the best iterate (smallest true ||r||_inf) kept
    double rbest = before;
    copy_m(c, best, dy);
    bool ok = false;
    int stall = 0;
    copy_m(c, z, r);
    precond(c, z);
    double rz;
    {
        double s = 0.0;
        for (int i = threadIdx.x; i < L.m; i += BT) {
            p[i] = z[i];
            s += r[i] * z[i];
        }
        double v[1] = {s};
        const int op[1] = {RSUM};
        wg_reduce<1>(v, op, c.red);
        rz = v[0];
    }
    for (int it = 0; it < max_it && !ok; ++it) {
        const double pq = m0_apply(c, p, q);
        if (!(pq > 0.0)) break;
        const double alpha = rz / pq;
        double mx = 0.0;
        for (int i = threadIdx.x; i < L.m; i += BT) {
            dy[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            mx = fmax(mx, fabs(r[i]));
        }
        double v1[1] = {mx};
        const int opm[1] = {RMAX};
        wg_reduce<1>(v1, opm, c.red);
        double rr = v1[0];
        c.st.pcg_steps++;
        if (rr <= tol || it == max_it - 1 || (it + 1) % 5 == 0) {
            rr = m0_residual(c, dy, rhs, r);
            stall = rr < 0.5 * rbest ? 0 : stall + 1;
            if (rr < rbest) {
                rbest = rr;
                copy_m(c, best, dy);
            }
            if (rr <= tol) { ok = true; break; }
            if (stall >= 2) break;
        }
        copy_m(c, z, r);
        precond(c, z);
        double s = 0.0;
        for (int i = threadIdx.x; i < L.m; i += BT) s += r[i] * z[i];
        double v2[1] = {s};
        const int ops[1] = {RSUM};
        wg_reduce<1>(v2, ops, c.red);
        const double beta = v2[0] / rz;
        rz = v2[0];
        for (int i = threadIdx.x; i < L.m; i += BT) p[i] = z[i] + beta * p[i];
        __syncthreads();
    }
    if (!ok) {
        copy_m(c, dy, best);
        if (c.rel) {
            // accepted at the level a backward-stable solve reaches:
            // ||r||_inf <= 1e-12 (||rhs||_inf + max_i M0_ii ||dy||_inf)
            double a = 0.0, b = 0.0;
            for (int i = threadIdx.x; i < L.m; i += BT) {
                a = fmax(a, fabs(rhs[i]));
                b = fmax(b, fabs(dy[i]));
            }
            double v[2] = {a, b};
            const int op[2] = {RMAX, RMAX};
            wg_reduce<2>(v, op, c.red);
            ok = rbest <= 1e-12 * (v[0] + c.m0max * v[1]);
        }
    }
    return ok;
}

struct Res {
    double nrp, nrd, mu, cx, by, rpinf, rdinf, xsmax, thmax, thmin;
};

// rp = A x - b, rd = A^T y + s - c, Theta = x / s and their norms (compute_residuals,
// mehrotra_device.cu:589-624)
__device__ Res residuals(Ctx& c) {
    const Lp& L = c.L;
    const double *x = L.N(VX), *s = L.N(VS), *cc = L.N(VC), *Cs = L.N(VCS);
    const double *y = L.M(WY), *b = L.M(WB), *R = L.M(WR);
    double *rp = L.M(WRP), *rd = L.N(VRD), *th = L.N(VTH);
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, -1e300, 1e300};
    seg_pass<GR>(L.m, L.rp, RSUM, [&](int p) { return L.av[p] * x[L.ci[p]]; },
                 [&](int i, double ax) {
                     const double r = ax + (-b[i]);
                     rp[i] = r;
                     v[0] += r * r;
                     v[1] = fmax(v[1], fabs(r / R[i]));
                     v[2] += b[i] * y[i];
                 });
    seg_pass<GC>(L.n, L.cp, RSUM, [&](int q) { return L.cv[q] * y[L.ri[q]]; },
                 [&](int j, double aty) {
                     const double r = aty + (s[j] - cc[j]);
                     const double tj = x[j] / s[j];
                     rd[j] = r;
                     th[j] = tj;
                     v[3] += r * r;
                     v[4] += x[j] * s[j];
                     v[5] += cc[j] * x[j];
                     v[6] = fmax(v[6], fabs(r / Cs[j]));
                     v[7] = fmax(v[7], fabs(x[j] * (s[j] - r)));
                     v[8] = fmax(v[8], tj);
                     v[9] = fmin(v[9], tj);
                 });
    const int op[10] = {RSUM, RMAX, RSUM, RSUM, RSUM, RSUM, RMAX, RMAX, RMAX, RMIN};
    wg_reduce<10>(v, op, c.red);
    Res r;
    r.nrp = sqrt(v[0]);
    r.rpinf = v[1];
    r.by = v[2];
    r.nrd = sqrt(v[3]);
    r.mu = v[4] / L.n;
    r.cx = v[5];
    r.rdinf = v[6];
    r.xsmax = v[7];
    r.thmax = v[8];
    r.thmin = v[9];
    return r;
}

struct Dir { double e1, e2, e3, apmax, admax; };

// v = -Theta rd - rxs / s ; rhs = -rp + A v ; KKT solve ; ds = -rd - A^T dy ;
// dx = -Theta ds + rxs / s (mehrotra_device.cu:111-176,627-660).  false when the solve failed.
__device__ bool direction(Ctx& c, double* dx, double* ds, double tol) {
    const Lp& L = c.L;
    const double *x = L.N(VX), *s = L.N(VS), *th = L.N(VTH), *rd = L.N(VRD), *rxs = L.N(VRXS);
    const double* rp = L.M(WRP);
    double *v = L.N(VV), *rhs = L.M(WRHS), *dy = L.M(WDY);
    (void)x;
    for (int j = threadIdx.x; j < L.n; j += BT) v[j] = -th[j] * rd[j] - rxs[j] / s[j];
    __syncthreads();
    seg_pass<GR>(L.m, L.rp, RSUM, [&](int p) { return L.av[p] * v[L.ci[p]]; },
                 [&](int i, double av) { rhs[i] = -rp[i] + av; });
    __syncthreads();
    const int budget = c.fp32 ? c.A.mixed_max_refine : (c.rel ? 60 : 10);
    bool ok = kkt_solve(c, tol, budget);
    if (!ok && c.fp32) {
        // the FP32 factor could not deliver: FP64 factor of the same matrix, same right-hand side
        if (!factor_any(c, false)) return false;
        ok = kkt_solve(c, tol, c.rel ? 60 : 10);
    }
    if (!ok) return false;
    seg_pass<GC>(L.n, L.cp, RSUM, [&](int q) { return L.cv[q] * dy[L.ri[q]]; },
                 [&](int j, double aty) {
                     const double d = -rd[j] - aty;
                     ds[j] = d;
                     dx[j] = -th[j] * d + rxs[j] / s[j];
                 });
    __syncthreads();
    return true;
}

// verify_newton_direction (mehrotra_device.cu:494-587) with the line-search products:
// qp = A dx, qd = A^T dy + ds; e1 = ||rp + qp|| / max(1, ||rp||) (0 in relative-gap mode),
// e2 = ||rd + qd|| / max(1, ||rd||), e3 = ||s dx + x ds - rxs|| / max(1, ||rxs||); step bounds
__device__ Dir verify(Ctx& c, const double* dx, const double* ds, const Res& nr) {
    const Lp& L = c.L;
    const double *x = L.N(VX), *s = L.N(VS), *rd = L.N(VRD), *rxs = L.N(VRXS);
    const double *rp = L.M(WRP), *dy = L.M(WDY);
    double *qp = L.M(WQP), *qd = L.N(VQD);
    double v[6] = {0, 0, 0, 0, 1e30, 1e30};
    seg_pass<GR>(L.m, L.rp, RSUM, [&](int p) { return L.av[p] * dx[L.ci[p]]; },
                 [&](int i, double a) {
                     qp[i] = a;
                     const double r1 = rp[i] + a;
                     v[0] += r1 * r1;
                 });
    seg_pass<GC>(L.n, L.cp, RSUM, [&](int q) { return L.cv[q] * dy[L.ri[q]]; },
                 [&](int j, double a) {
                     const double qj = a + ds[j];
                     qd[j] = qj;
                     const double r2 = rd[j] + qj;
                     const double r3 = s[j] * dx[j] + x[j] * ds[j] - rxs[j];
                     v[1] += r2 * r2;
                     v[2] += r3 * r3;
                     v[3] += rxs[j] * rxs[j];
                     v[4] = fmin(v[4], step_len(x[j], dx[j]));
                     v[5] = fmin(v[5], step_len(s[j], ds[j]));
                 });
    const int op[6] = {RSUM, RSUM, RSUM, RSUM, RMIN, RMIN};
    wg_reduce<6>(v, op, c.red);
    Dir d;
    d.e1 = c.rel ? 0.0 : sqrt(v[0]) / fmax(1.0, nr.nrp);
    d.e2 = sqrt(v[1]) / fmax(1.0, nr.nrd);
    d.e3 = sqrt(v[2]) / fmax(1.0, sqrt(v[3]));
    d.apmax = v[4];
    d.admax = v[5];
    return d;
}

__device__ __forceinline__ bool dir_ok(const Dir& d) {
    return !(d.e1 > 1e-8 || d.e2 > 1e-8 || d.e3 > 1e-8);
}


// Original-space relative infeasibility of the warm point of LP d:
// max(||A x0 - b||_inf / (1 + ||b||_inf), ||A^T y0 + s0 - c||_inf / (1 + ||c||_inf)), with the
// working values still equal to the original ones.
__device__ double warm_infeasibility(Ctx& c, const lkb_lp& d) {
    const Lp& L = c.L;
    const double *x0 = c.A.x0 + d.on, *s0 = c.A.s0 + d.on, *y0 = c.A.y0 + d.om;
    double v[2] = {0.0, 0.0};
    seg_pass<GR>(L.m, L.rp, RSUM, [&](int p) { return L.av[p] * x0[L.ci[p]]; },
                 [&](int i, double ax) { v[0] = fmax(v[0], fabs(ax - L.b0[i])); });
    seg_pass<GC>(L.n, L.cp, RSUM, [&](int q) { return L.cv[q] * y0[L.ri[q]]; },
                 [&](int j, double aty) { v[1] = fmax(v[1], fabs(aty + s0[j] - L.c0[j])); });
    const int op[2] = {RMAX, RMAX};
    wg_reduce<2>(v, op, c.red);
    return fmax(v[0] / (1.0 + c.st.binf), v[1] / (1.0 + c.st.cinf));
}

// Scaling and start point (mehrotra_device.cu:93-102,286-323): the working CSR and CSC values
// from the original ones, 10 Jacobi Ruiz passes, x = s = 1, y = 0 -- or, for an LP with an
// accepted warm start, the given original-space point scaled (x / C, y / R, s C), x and s
// floored, then each pair lifted to x_j s_j >= warm_mu by raising its smaller member (the
// central-path point at mu of a strictly complementary pair is about (x + mu / s, s + mu / x)).
__device__ void init_lp(Ctx& c, const lkb_lp& d, bool warm) {
    const Lp& L = c.L;
    double *b = L.M(WB), *R = L.M(WR), *fr = L.M(WTM), *y = L.M(WY);
    double *cc = L.N(VC), *Cs = L.N(VCS), *fc = L.N(VTN), *x = L.N(VX), *s = L.N(VS);
    double bi = 0.0, ci = 0.0;
    for (int p = threadIdx.x; p < L.nnz; p += BT) {
        L.cv[p] = L.v0[p];
        L.av[p] = L.v0[L.pos[p]];
    }
    for (int i = threadIdx.x; i < L.m; i += BT) {
        b[i] = L.b0[i];
        R[i] = 1.0;
        y[i] = 0.0;
        bi = fmax(bi, fabs(L.b0[i]));
    }
    for (int j = threadIdx.x; j < L.n; j += BT) {
        cc[j] = L.c0[j];
        Cs[j] = 1.0;
        x[j] = 1.0;
        s[j] = 1.0;
        ci = fmax(ci, fabs(L.c0[j]));
    }
    {
        double v[2] = {bi, ci};
        const int op[2] = {RMAX, RMAX};
        wg_reduce<2>(v, op, c.red);   // its barriers also order the copies above
        c.st.binf = v[0];
        c.st.cinf = v[1];
    }
    c.st.warm = 0;
    c.st.warm_infeas = 0.0;
    if (warm) {
        c.st.warm_infeas = warm_infeasibility(c, d);
        warm = c.st.warm_infeas <= c.A.warm_max_infeas;
        c.st.warm = warm ? 1 : 2;
    }
    for (int pass = 0; pass < 10; ++pass) {
        seg_pass<GR>(L.m, L.rp, RMAX, [&](int p) { return fabs(L.av[p]); },
                     [&](int i, double mx) { fr[i] = mx > 1e-8 ? 1.0 / sqrt(mx) : 1.0; });
        seg_pass<GC>(L.n, L.cp, RMAX, [&](int q) { return fabs(L.cv[q]); },
                     [&](int j, double mx) { fc[j] = mx > 1e-8 ? 1.0 / sqrt(mx) : 1.0; });
        __syncthreads();
        for (int i = threadIdx.x; i < L.m; i += BT) {
            R[i] *= fr[i];
            b[i] *= fr[i];
        }
        for (int j = threadIdx.x; j < L.n; j += BT) {
            Cs[j] *= fc[j];
            cc[j] *= fc[j];
        }
        // the same product fr_i fc_j scales the CSR and the CSC copy of every entry
        seg_for<GR>(L.m, L.rp, [&](int i, int p) { L.av[p] *= fr[i] * fc[L.ci[p]]; });
        seg_for<GC>(L.n, L.cp, [&](int j, int q) { L.cv[q] *= fr[L.ri[q]] * fc[j]; });
        __syncthreads();
    }
    if (warm) {
        const double fl = c.A.warm_floor, mu = c.A.warm_mu;
        for (int j = threadIdx.x; j < L.n; j += BT) {
            double xj = fmax(c.A.x0[d.on + j] / Cs[j], fl);
            double sj = fmax(c.A.s0[d.on + j] * Cs[j], fl);
            if (xj * sj < mu) {
                if (xj >= sj) sj = mu / xj;
                else xj = mu / sj;
            }
            x[j] = xj;
            s[j] = sj;
        }
        for (int i = threadIdx.x; i < L.m; i += BT) y[i] = c.A.y0[d.om + i] / R[i];
        __syncthreads();
    }
}

// Final figures of the LP: objective c_s^T x_s (= c^T x), b_s^T y_s, the relative gap and the
// original-space relative residuals (the solver's result fields), then the solution unscaled:
// x = C x_s, y = R y_s, s = s_s / C.
__device__ void finish(Ctx& c, const lkb_lp& d, const Res& nr, int status) {
    const Lp& L = c.L;
    if (threadIdx.x == 0) atomicAdd(c.A.ctl + 1, 1u);   // LPs finished in this solve
    c.st.phase = 2;
    c.st.status = status;
    c.st.objective = nr.cx;
    c.st.dual_objective = nr.by;
    c.st.rel_gap = fabs(nr.cx - nr.by) / fmax(1.0, fmax(fabs(nr.cx), fabs(nr.by)));
    c.st.rel_primal = nr.rpinf / (1.0 + c.st.binf);
    c.st.rel_dual = nr.rdinf / (1.0 + c.st.cinf);
    c.st.primal_residual = nr.nrp;
    c.st.dual_residual = nr.nrd;
    c.st.mu = nr.mu;
    const double *x = L.N(VX), *s = L.N(VS), *Cs = L.N(VCS), *y = L.M(WY), *R = L.M(WR);
    for (int j = threadIdx.x; j < L.n; j += BT) {
        c.A.x_out[d.on + j] = x[j] * Cs[j];
        c.A.s_out[d.on + j] = s[j] / Cs[j];
    }
    for (int i = threadIdx.x; i < L.m; i += BT) c.A.y_out[d.om + i] = y[i] * R[i];
}

// relative-gap mode: absolute targets of ||Ax - b||_inf, ||A^T y + s - c||_inf (original space)
// and max_j x_j (c - A^T y)_j, a fraction REL_ABS_MARGIN of gap_tol (the FP64 certificate
// checks them absolutely)
constexpr double REL_ABS_MARGIN = 0.9;
constexpr double SHORT_STEP = 1e-2;
__device__ __forceinline__ double primal_target(const Ctx& c) {
    return fmax(REL_ABS_MARGIN * c.A.gap_tol, 1e-12 * (1.0 + c.st.binf));
}
__device__ __forceinline__ double dual_target(const Ctx& c) {
    return fmax(REL_ABS_MARGIN * c.A.gap_tol, 1e-12 * (1.0 + c.st.cinf));
}

__device__ bool stop_rule(const Ctx& c, const Res& nr) {
    if (!c.rel) return nr.nrp < c.A.tol && nr.nrd < c.A.tol && nr.mu < c.A.tol;
    const double g = c.A.gap_tol;
    const double big = fmax(1.0, fmax(fabs(nr.cx), fabs(nr.by)));
    const double gap = fabs(nr.cx - nr.by) / big;
    return gap <= g && nr.rpinf / (1.0 + c.st.binf) <= g && nr.rdinf / (1.0 + c.st.cinf) <= g &&
           nr.rpinf <= primal_target(c) && nr.rdinf <= dual_target(c) &&
           nr.xsmax <= fmax(REL_ABS_MARGIN * g, 1e-14 * big);
}

// One backtracking trial x + a dx, s + b ds, y + b dy evaluated through the residual updates
// rp + a (A dx), rd + b (A^T dy + ds): 2-norms, mu, minima, ||y||, and the original-space
// infinity norms of both residuals.
struct Trial { double nrp, nrd, mu, mx, ms, ny, rpinf, rdinf; };
__device__ Trial trial(Ctx& c, double a, double b) {
    const Lp& L = c.L;
    const double *x = L.N(VX), *s = L.N(VS), *dx = L.N(VDX), *ds = L.N(VDS), *qd = L.N(VQD);
    const double *rd = L.N(VRD), *Cs = L.N(VCS);
    const double *y = L.M(WY), *dy = L.M(WDY), *qp = L.M(WQP), *rp = L.M(WRP), *R = L.M(WR);
    double v[8] = {0, 0, 0, 1e300, 1e300, 0, 0, 0};
    for (int j = threadIdx.x; j < L.n; j += BT) {
        const double xn = x[j] + a * dx[j], sn = s[j] + b * ds[j], r2 = rd[j] + b * qd[j];
        v[1] += r2 * r2;
        v[2] += xn * sn;
        v[3] = fmin(v[3], xn);
        v[4] = fmin(v[4], sn);
        v[7] = fmax(v[7], fabs(r2 / Cs[j]));
    }
    for (int i = threadIdx.x; i < L.m; i += BT) {
        const double r1 = rp[i] + a * qp[i], yn = y[i] + b * dy[i];
        v[0] += r1 * r1;
        v[5] += yn * yn;
        v[6] = fmax(v[6], fabs(r1 / R[i]));
    }
    const int op[8] = {RSUM, RSUM, RSUM, RMIN, RMIN, RSUM, RMAX, RMAX};
    wg_reduce<8>(v, op, c.red);
    return Trial{sqrt(v[0]), sqrt(v[1]), v[2] / L.n, v[3], v[4], sqrt(v[5]), v[6], v[7]};
}

// One Mehrotra iteration (mehrotra_device.cu:695-1047).  Returns true when the LP finished.
__device__ bool iterate(Ctx& c, const lkb_lp& d) {
    const Lp& L = c.L;
    double *x = L.N(VX), *s = L.N(VS), *rxs = L.N(VRXS), *dxa = L.N(VDXA), *dsa = L.N(VDSA);
    double *dx = L.N(VDX), *ds = L.N(VDS), *qd = L.N(VQD), *rd = L.N(VRD);
    double *y = L.M(WY), *dy = L.M(WDY), *qp = L.M(WQP), *rp = L.M(WRP);
    const int iter = c.st.iterations;
    const Res nr = residuals(c);
    if (iter == 0) {
        c.st.rp0 = fmax(1.0, nr.nrp);
        c.st.rd0 = fmax(1.0, nr.nrd);
        c.st.mu0 = fmax(1.0, nr.mu);
    }
    if (iter >= c.A.max_iter) { finish(c, d, nr, 3); return true; }
    if (stop_rule(c, nr)) { finish(c, d, nr, 0); return true; }
    auto fail = [&](int code) { c.st.fail_code = code; finish(c, d, nr, 3); return true; };

    // ---- factorization of this iteration (FP32 while the Theta spread allows it)
    const bool want32 = c.A.precision == 1 && nr.thmin > 0.0 && nr.thmax / nr.thmin <= c.A.mixed_spread;
    diag_scaling(c);
    if (!factor_any(c, want32)) return fail(1);
    const double tol = 1e-9 * fmax(1.0, nr.nrp);

    // ---- predictor
    for (int j = threadIdx.x; j < L.n; j += BT) rxs[j] = -x[j] * s[j] + 0.0;
    __syncthreads();
    if (!direction(c, dxa, dsa, tol)) return fail(2);
    const Dir dp = verify(c, dxa, dsa, nr);
    if (!dir_ok(dp)) return fail(3);
    double mu_aff;
    {
        const double ap = fmin(1.0, dp.apmax), ad = fmin(1.0, dp.admax);
        double v[1] = {0.0};
        for (int j = threadIdx.x; j < L.n; j += BT) v[0] += (x[j] + ap * dxa[j]) * (s[j] + ad * dsa[j]);
        const int op[1] = {RSUM};
        wg_reduce<1>(v, op, c.red);
        mu_aff = v[0] / L.n;
    }
    const double sigma = fmin(1.0, pow(fmax(0.0, mu_aff) / fmax(1e-16, nr.mu), 3.0));

    // ---- corrector
    for (int j = threadIdx.x; j < L.n; j += BT) rxs[j] = -x[j] * s[j] + sigma * nr.mu - dxa[j] * dsa[j];
    __syncthreads();
    if (!direction(c, dx, ds, tol)) return fail(4);
    const Dir dc = verify(c, dx, ds, nr);
    double ap = fmin(1.0, 0.995 * dc.apmax), ad = fmin(1.0, 0.995 * dc.admax);
    const double merit0 = nr.nrp / c.st.rp0 + nr.nrd / c.st.rd0 + nr.mu / c.st.mu0;
    // relative-gap mode: a residual already inside its absolute target must stay inside
    const bool keep_p = c.rel && nr.rpinf <= primal_target(c);
    const bool keep_d = c.rel && nr.rdinf <= dual_target(c);
    // `usable`: finite, positive and keeping the met residual targets (merit aside); `merit`
    // receives the trial's merit
    auto admissible = [&](const Trial& T, double a, double b, bool& usable, double& merit) {
        merit = T.nrp / c.st.rp0 + T.nrd / c.st.rd0 + T.mu / c.st.mu0;
        const double stp = fmin(a, b);
        const bool finite = isfinite(T.nrp) && isfinite(T.nrd) && isfinite(T.mu) &&
                            isfinite(merit) && isfinite(merit0) && isfinite(T.mx) &&
                            isfinite(T.ms) && isfinite(a) && isfinite(b) && isfinite(T.ny);
        const bool positive = T.mx > 0.0 && T.ms > 0.0 && T.mu > 0.0 && stp > 0.0;
        const bool keeps = !(keep_p && T.rpinf > primal_target(c)) &&
                           !(keep_d && T.rdinf > dual_target(c));
        usable = finite && positive && keeps;
        return usable && merit < merit0 && merit <= (1.0 - 1e-4 * stp) * merit0;
    };
    bool accepted = false;
    for (int dir = dir_ok(dc) ? 0 : 1; dir < 2 && !accepted; ++dir) {
        if (dir == 1) {
            // centred fallback: sigma = 0.5, common step (mehrotra_device.cu:905-995)
            for (int j = threadIdx.x; j < L.n; j += BT) rxs[j] = -x[j] * s[j] + 0.5 * nr.mu;
            __syncthreads();
            if (!direction(c, dx, ds, tol)) break;
            const Dir dz = verify(c, dx, ds, nr);
            if (!dir_ok(dz)) break;
            ap = ad = fmin(1.0, 0.995 * fmin(dz.apmax, dz.admax));
        }
        const int ntr = dir == 0 ? 15 : 30;
        double a = ap, b = ad, fa = -1.0, fb = -1.0, acc_merit = 0.0;
        for (int t = 0; t < ntr; ++t) {
            bool usable;
            double mt;
            if (admissible(trial(c, a, b), a, b, usable, mt)) { accepted = true; acc_merit = mt; break; }
            if (c.rel && dir == 1 && usable && fa < 0) { fa = a; fb = b; }
            a *= 0.5;
            b *= 0.5;
        }
        // relative-gap mode: a corrector step shorter than SHORT_STEP taken with both residuals
        // inside their targets competes with the dual-only trials (the better merit wins)
        const bool short_step = accepted && c.rel && dir == 0 && fmin(a, b) < SHORT_STEP && keep_p && keep_d;
        if ((!accepted || short_step) && c.rel && dir == 0) {
            // relative-gap mode: dual-only trials of the same direction -- the primal point (and
            // its residual) stays; near the optimum the normal equations limit the primal
            // accuracy of a direction, not its dual accuracy
            double b2 = fmin(1.0, 0.995 * dc.admax);
            for (int t = 0; t < 15; ++t) {
                bool usable;
                double mt;
                if (admissible(trial(c, 0.0, b2), b2, b2, usable, mt) && (!short_step || mt < acc_merit)) {
                    accepted = true;
                    a = 0.0;
                    b = b2;
                    break;
                }
                b2 *= 0.5;
            }
        }
        if (!accepted && c.rel && dir == 1 && fa > 0) {
            // relative-gap mode: the longest usable centred step (non-monotone; it may not break
            // a met residual target)
            a = fa;
            b = fb;
            accepted = true;
        }
        if (accepted) {
            for (int j = threadIdx.x; j < L.n; j += BT) {
                x[j] += a * dx[j];
                s[j] += b * ds[j];
            }
            for (int i = threadIdx.x; i < L.m; i += BT) y[i] += b * dy[i];
            __syncthreads();
        }
    }
    if (!accepted) return fail(5);
    c.st.iterations = iter + 1;
    return false;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_batch(lkb_args A) {
    __shared__ int s_ticket;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = (int)atomicAdd(A.ctl, 1u);
        __syncthreads();
        const int ticket = s_ticket;
        __syncthreads();
        if (ticket >= A.count) break;
        const int li = A.order[ticket];
        const lkb_lp d = A.lps[li];
        Ctx c{A, make_lp(A, d), A.state[li], g_red, A.stop_rule == 1, false, 0.0, wall_clock64()};
        if (c.st.phase == 2) continue;
        if (c.st.phase == 0) {
            c.st.phase = 1;
            c.st.iterations = 0;
            c.st.fail_code = 0;
            c.st.factorizations = c.st.fp32_factorizations = c.st.solves = c.st.pcg_steps = 0;
            c.st.t_init = c.st.t_other = c.st.t_assemble = c.st.t_factor = c.st.t_solve = 0.0;
            init_lp(c, d, A.warm_flag && A.warm_flag[li] && A.x0 && A.y0 && A.s0);
            c.tick(c.st.t_init);
        }
        for (int k = 0; k < A.iters_per_visit; ++k)
            if (iterate(c, d)) break;
        c.tick(c.st.t_other);
        __syncthreads();
        if (threadIdx.x == 0) A.state[li] = c.st;
        __syncthreads();
    }
}

__global__ void k_batch_reset(uint32_t* ctl, lkb_state* state, const int32_t* order, int count,
                              int restart) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctl[0] = 0u;
        if (restart) ctl[1] = 0u;
    }
    if (restart && i < count) state[order[i]].phase = 0;
}

// ----------------------------------------------------------------------------- crossover
// One workgroup per LP (src/ipm/crossover.cu:322-460 is the reference's column-by-column
// purification from an artificial basis; here the basis comes from the support first):
//  1. xv = max(x, 0), tau = 1e-7 max(1, max xv); candidates: the support {xv > tau} by xv
//     descending (ties: lower index), then the other columns by index;
//  2. B^-1 = I with every row artificial (the LP's factor slot, column-major, leading dimension
//     ld); a candidate q takes the artificial row r of largest |u_r| (lowest r on ties),
//     u = B^-1 a_q, when |u_r| > 1e-9 max(1, max |u|): Gauss-Jordan update on pivot (r, u);
//  3. purification: the lowest nonbasic q with xv_q > tau moves by theta and the basic values by
//     -theta u, theta > 0 when its reduced cost d_q <= 1e-11 (1 + |c_q|) (ratio test over
//     u_i > 1e-12 with the basic values floored at 0; none: unbounded, status 3), else
//     theta = -min(xv_q, ratios over u_i < -1e-12); a blocking basic row (first minimum) pivots
//     q in, x_q blocking ends at 0;
//  4. clean-up when no row is artificial (x_B = B^-1 b, y = B^-T c_B, d = c - A'y,
//     ptol = 1e-9 max(1, max|x_B|), dtol = 1e-9 max(1, max|c|)): dual simplex while the lowest
//     x_B(r) < -ptol (entering: alpha = (B^-1)_r A below -1e-9 max(1, max|alpha|), least
//     d'_j / -alpha_j with d' = d above dtol else 0, lowest j on ties), then primal simplex
//     with Bland's rule while some d_q < -dtol (lowest q; ratio x'_i / u_i over u_i above
//     1e-9 max(1, max|u|), x' = x_B above ptol else 0, ties to the lowest basic column; no row:
//     status 3); at most m pivots each;
//  5. x_B = B^-1 b, x_N = 0, y = B^-T c_B, reduced costs; the status.
// Work vectors: N(0) values, N(1) candidate order (int), N(2) basic flags, M(0) u, M(1) c_B,
// M(2) x_B, M(3) y.
constexpr int XO_T = 256;

// thread-local (value, index) candidates reduced over the workgroup: the smallest key wins,
// ties by the lower index; every thread receives the result
struct KeyIdx { double key; int idx; };
__device__ KeyIdx wg_argmin(KeyIdx v, KeyIdx* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = XO_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const KeyIdx a = sh[threadIdx.x], b = sh[threadIdx.x + w];
            if (b.idx >= 0 && (a.idx < 0 || b.key < a.key || (b.key == a.key && b.idx < a.idx)))
                sh[threadIdx.x] = b;
        }
        __syncthreads();
    }
    const KeyIdx r = sh[0];
    __syncthreads();
    return r;
}
__device__ double wg_sum_max(double v, bool is_max, double* shd) {
    shd[threadIdx.x] = v;
    __syncthreads();
    for (int w = XO_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            shd[threadIdx.x] = is_max ? fmax(shd[threadIdx.x], shd[threadIdx.x + w])
                                      : shd[threadIdx.x] + shd[threadIdx.x + w];
        __syncthreads();
    }
    const double r = shd[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_crossover(lkb_args A, lkx_args X) {
    __shared__ int brow[LKB_MAX_M];          // basic column of each row, -1 artificial
    __shared__ double prow[LKB_MAX_M];       // pivot row of B^-1 (scaled)
    __shared__ KeyIdx ski[XO_T];
    __shared__ double sd[XO_T];
    const int li = blockIdx.x;
    if (li >= A.count) return;
    const lkb_lp d = A.lps[li];
    const Lp L = make_lp(A, d);
    const int m = L.m, n = L.n, t = threadIdx.x;
    const int64_t ld = L.ld;
    double* B = L.fac;
    double* xv = L.N(0);
    int* ord = reinterpret_cast<int*>(L.N(1));
    double* isb = L.N(2);
    double* u = L.M(0);
    double* cb = L.M(1);
    const double* x = X.x + d.on;

    // 1. values, threshold, candidate order
    double mx = 0.0;
    for (int j = t; j < n; j += XO_T) {
        const double v = fmax(x[j], 0.0);
        xv[j] = v;
        isb[j] = 0.0;
        mx = fmax(mx, v);
    }
    const double tau = 1e-7 * fmax(1.0, wg_sum_max(mx, true, sd));
    for (int j = t; j < n; j += XO_T) {      // rank of a support column: larger values first
        const double v = xv[j];
        if (!(v > tau)) continue;
        int rank = 0;
        for (int k = 0; k < n; ++k) {
            const double w = xv[k];
            rank += (w > tau && (w > v || (w == v && k < j))) ? 1 : 0;
        }
        ord[rank] = j;
    }
    __syncthreads();
    if (t == 0) {
        int p = 0;
        for (int j = 0; j < n; ++j) p += xv[j] > tau ? 1 : 0;
        for (int j = 0; j < n; ++j)
            if (!(xv[j] > tau)) ord[p++] = j;
    }
    for (int i = t; i < m; i += XO_T) brow[i] = -1;
    for (int64_t e = t; e < ld * m; e += XO_T) {
        const int64_t j = e / ld, i = e - j * ld;
        B[e] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();

    // u = B^-1 a_q
    auto ftran = [&](int q) {
        const int p0 = L.cp[q], p1 = L.cp[q + 1];
        for (int i = t; i < m; i += XO_T) {
            double s = 0.0;
            for (int p = p0; p < p1; ++p) s += B[i + ld * L.ri[p]] * L.v0[p];
            u[i] = s;
        }
        __syncthreads();
    };
    // Gauss-Jordan update of B^-1 on pivot (r, u)
    auto pivot = [&](int r) {
        const double ur = u[r];
        for (int j = t; j < m; j += XO_T) prow[j] = B[r + ld * j] / ur;
        __syncthreads();
        for (int i = t; i < m; i += XO_T) {
            const double ui = u[i];
            if (i == r) {
                for (int j = 0; j < m; ++j) B[i + ld * j] = prow[j];
            } else if (ui != 0.0) {
                for (int j = 0; j < m; ++j) B[i + ld * j] -= ui * prow[j];
            }
        }
        __syncthreads();
    };

    // 2. basis from the candidates
    int pivots = 0, nart = m;
    for (int c = 0; c < n && nart > 0; ++c) {
        const int q = ord[c];
        ftran(q);
        double um = 0.0;
        KeyIdx best{0.0, -1};
        for (int i = t; i < m; i += XO_T) {
            const double a = fabs(u[i]);
            um = fmax(um, a);
            if (brow[i] < 0 && (best.idx < 0 || -a < best.key)) best = KeyIdx{-a, i};
        }
        um = wg_sum_max(um, true, sd);
        best = wg_argmin(best, ski);
        if (best.idx >= 0 && -best.key > 1e-9 * fmax(1.0, um)) {
            pivot(best.idx);
            if (t == 0) {
                brow[best.idx] = q;
                isb[q] = 1.0;
            }
            __syncthreads();
            ++pivots;
            --nart;
        }
    }

    // 3. purification
    int status = 0;
    for (int guard = 0; guard < n + m; ++guard) {
        KeyIdx qk{0.0, -1};
        for (int j = t; j < n; j += XO_T)
            if (isb[j] == 0.0 && xv[j] > tau && (qk.idx < 0 || j < qk.idx)) qk = KeyIdx{(double)j, j};
        qk = wg_argmin(qk, ski);
        if (qk.idx < 0) break;
        const int q = qk.idx;
        ftran(q);
        double dsum = 0.0;
        for (int i = t; i < m; i += XO_T) {
            const double ci = brow[i] >= 0 ? L.c0[brow[i]] : 0.0;
            cb[i] = ci;
            dsum += ci * u[i];
        }
        const double dq = L.c0[q] - wg_sum_max(dsum, false, sd);
        const bool up = dq <= 1e-11 * (1.0 + fabs(L.c0[q]));   // no cost increase moving up
        KeyIdx rt{0.0, -1};
        for (int i = t; i < m; i += XO_T) {
            const double ui = u[i];
            const double vi = brow[i] >= 0 ? fmax(xv[brow[i]], 0.0) : 0.0;
            double ratio = -1.0;
            if (up && ui > 1e-12) ratio = vi / ui;
            if (!up && ui < -1e-12) ratio = vi / (-ui);
            if (ratio >= 0.0 && (rt.idx < 0 || ratio < rt.key)) rt = KeyIdx{ratio, i};
        }
        rt = wg_argmin(rt, ski);
        double theta;
        int r = rt.idx;
        if (up) {
            if (r < 0) { status = 3; break; }
            theta = rt.key;
        } else {
            theta = xv[q];
            if (r >= 0 && rt.key < theta) theta = rt.key;
            else r = -1;
            theta = -theta;
        }
        for (int i = t; i < m; i += XO_T)
            if (brow[i] >= 0) xv[brow[i]] -= theta * u[i];
        __syncthreads();
        if (t == 0) xv[q] += theta;
        __syncthreads();
        if (r >= 0) {
            const int leave = brow[r];
            pivot(r);
            if (t == 0) {
                if (leave >= 0) { isb[leave] = 0.0; xv[leave] = 0.0; }
                brow[r] = q;
                isb[q] = 1.0;
            }
            ++pivots;
        } else if (t == 0) {
            xv[q] = 0.0;
        }
        __syncthreads();
    }

    // 4. clean-up at degenerate vertices: dual simplex on primal infeasibilities, then primal
    //    simplex (Bland's rule) on dual infeasibilities; both only with no artificial row left
    int nart_left = 0;
    for (int i = t; i < m; i += XO_T) nart_left += brow[i] < 0 ? 1 : 0;
    nart_left = (int)wg_sum_max((double)nart_left, false, sd);
    if (status == 0 && nart_left == 0) {
        __shared__ int s_row;
        double* xB = L.M(2);
        double* y = L.M(3);
        double cmax = 0.0;
        for (int j = t; j < n; j += XO_T) cmax = fmax(cmax, fabs(L.c0[j]));
        const double dtol = 1e-9 * fmax(1.0, wg_sum_max(cmax, true, sd));
        // x_B = B^-1 b and y = B^-T c_B; returns the primal tolerance 1e-9 max(1, max|x_B|)
        auto basic_state = [&]() -> double {
            double xa = 0.0;
            for (int i = t; i < m; i += XO_T) {
                double s = 0.0;
                for (int j = 0; j < m; ++j) s += B[i + ld * j] * L.b0[j];
                xB[i] = s;
                cb[i] = L.c0[brow[i]];
                xa = fmax(xa, fabs(s));
            }
            __syncthreads();
            for (int j = t; j < m; j += XO_T) {
                double s = 0.0;
                for (int i = 0; i < m; ++i) s += B[i + ld * j] * cb[i];
                y[j] = s;
            }
            return 1e-9 * fmax(1.0, wg_sum_max(xa, true, sd));
        };
        auto reduced_cost = [&](int j) {
            double s = 0.0;
            for (int p = L.cp[j]; p < L.cp[j + 1]; ++p) s += L.v0[p] * y[L.ri[p]];
            return L.c0[j] - s;
        };
        auto row_alpha = [&](int j) {        // (B^-1)_r a_j with row r staged in prow
            double a = 0.0;
            for (int p = L.cp[j]; p < L.cp[j + 1]; ++p) a += prow[L.ri[p]] * L.v0[p];
            return a;
        };
        auto swap_in = [&](int r, int q) {
            ftran(q);
            const int leave = brow[r];
            pivot(r);
            if (t == 0) {
                isb[leave] = 0.0;
                brow[r] = q;
                isb[q] = 1.0;
            }
            __syncthreads();
            ++pivots;
        };
        for (int it = 0; it < m; ++it) {            // dual simplex
            const double ptol = basic_state();
            KeyIdx lo{0.0, -1};
            for (int i = t; i < m; i += XO_T)
                if (lo.idx < 0 || xB[i] < lo.key) lo = KeyIdx{xB[i], i};
            lo = wg_argmin(lo, ski);
            if (!(lo.key < -ptol)) break;
            const int r = lo.idx;
            for (int j = t; j < m; j += XO_T) prow[j] = B[r + ld * j];
            __syncthreads();
            double amax = 0.0;
            for (int j = t; j < n; j += XO_T)
                if (isb[j] == 0.0) amax = fmax(amax, fabs(row_alpha(j)));
            const double atol = 1e-9 * fmax(1.0, wg_sum_max(amax, true, sd));
            KeyIdx best{0.0, -1};
            for (int j = t; j < n; j += XO_T) {
                if (isb[j] != 0.0) continue;
                const double a = row_alpha(j);
                if (!(a < -atol)) continue;
                const double d = reduced_cost(j);
                const double ratio = (d > dtol ? d : 0.0) / (-a);
                if (best.idx < 0 || ratio < best.key) best = KeyIdx{ratio, j};
            }
            best = wg_argmin(best, ski);
            if (best.idx < 0) break;
            swap_in(r, best.idx);
        }
        for (int it = 0; it < m; ++it) {            // primal simplex, Bland's rule
            const double ptol = basic_state();
            __syncthreads();
            KeyIdx qk{0.0, -1};
            for (int j = t; j < n; j += XO_T)
                if (isb[j] == 0.0 && reduced_cost(j) < -dtol) { qk = KeyIdx{(double)j, j}; break; }
            qk = wg_argmin(qk, ski);
            if (qk.idx < 0) break;
            const int q = qk.idx;
            ftran(q);
            double um = 0.0;
            for (int i = t; i < m; i += XO_T) um = fmax(um, fabs(u[i]));
            const double utol = 1e-9 * fmax(1.0, wg_sum_max(um, true, sd));
            KeyIdx rt{0.0, -1};                       // (ratio, basic column): ties to the lowest column
            for (int i = t; i < m; i += XO_T) {
                if (!(u[i] > utol)) continue;
                const double ratio = (xB[i] > ptol ? xB[i] : 0.0) / u[i];
                if (rt.idx < 0 || ratio < rt.key || (ratio == rt.key && brow[i] < rt.idx))
                    rt = KeyIdx{ratio, brow[i]};
            }
            rt = wg_argmin(rt, ski);
            if (rt.idx < 0) { status = 3; break; }
            for (int i = t; i < m; i += XO_T)
                if (brow[i] == rt.idx) s_row = i;
            __syncthreads();
            const int r = s_row;
            __syncthreads();
            swap_in(r, q);
        }
    }

    // 5. basic solution, dual check, outputs
    double* xout = X.xv + d.on;
    for (int j = t; j < n; j += XO_T) xout[j] = 0.0;
    __syncthreads();
    double xbmin = 1e300, xbabs = 0.0;
    for (int i = t; i < m; i += XO_T) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) s += B[i + ld * j] * L.b0[j];
        u[i] = s;
        cb[i] = brow[i] >= 0 ? L.c0[brow[i]] : 0.0;
        if (brow[i] >= 0) {
            xout[brow[i]] = s;
            xbmin = fmin(xbmin, s);
        }
        xbabs = fmax(xbabs, fabs(s));
    }
    xbmin = -wg_sum_max(-xbmin, true, sd);
    xbabs = wg_sum_max(xbabs, true, sd);
    int nleft = 0;
    for (int i = t; i < m; i += XO_T) nleft += brow[i] < 0 ? 1 : 0;
    nleft = (int)wg_sum_max((double)nleft, false, sd);
    if (status == 0) {
        if (nleft > 0) status = 2;
        else if (xbmin < -1e-9 * fmax(1.0, xbabs)) status = 4;
    }
    if (status == 0) {
        double* y = prow;                    // y = B^-T c_B (m <= LKB_MAX_M)
        for (int j = t; j < m; j += XO_T) {
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += B[i + ld * j] * cb[i];
            y[j] = s;
        }
        __syncthreads();
        double rmin = 1e300, cabs = 0.0;
        for (int j = t; j < n; j += XO_T) {
            double s = 0.0;
            for (int p = L.cp[j]; p < L.cp[j + 1]; ++p) s += L.v0[p] * y[L.ri[p]];
            rmin = fmin(rmin, L.c0[j] - s);
            cabs = fmax(cabs, fabs(L.c0[j]));
        }
        rmin = -wg_sum_max(-rmin, true, sd);
        cabs = wg_sum_max(cabs, true, sd);
        if (rmin < -1e-9 * fmax(1.0, cabs)) status = 1;
    }
    for (int i = t; i < m; i += XO_T) X.basis[d.om + i] = brow[i];
    if (t == 0) {
        X.status[li] = status;
        X.pivots[li] = pivots;
    }
}

// ----------------------------------------------------------------------------- simplex
// The reference's CPU route, solve_with_phase1 + primal_simplex_phase2 (src/simplex/phase1.cpp:
// 124-449, src/simplex/primal.cpp:27-423), for every LP of a batch (min c^T x, A x = b, x >= 0),
// one workgroup per LP, with a dense explicit basis inverse in the LP's factor slot (Gauss-Jordan
// updates) in place of the sparse LU with eta updates:
//  Phase I: rows with b_i < 0 negated (sign s_i), one artificial column per row (cost 1, the
//    structural columns cost 0), the artificial basis: B^-1 = I, x_B = |b|.
//  Each phase is the primal simplex, at most LKS_MAX_ITER iterations:
//    duals y = B^-T c_B (updated by y += (r_q / u_p) rho_p after every pivot, rho_p the pivot
//      row of the old inverse, and recomputed every 50 iterations); every 50 iterations and at
//      the end x_B and y are also refined against the basis matrix itself (one step of
//      x_B += B^-1 (b - B x_B), y += B^-T (c_B - B^T y)), since the Gauss-Jordan inverse drifts
//      over thousands of pivots where the reference refactorizes its LU;
//    entering column: Devex, the largest r_j^2 / gamma_j over nonbasic r_j < -1e-6 (lowest j
//      on ties; gamma = 1 at the start and every ncols iterations, the leaving column's gamma =
//      max(1, |u|^2)), or after more than 50 consecutive degenerate steps (theta < 1e-6) for the
//      rest of the phase, Bland's rule: the lowest j with r_j < -1e-6; none: Optimal;
//    u = B^-1 a_q; ratio test over u_i > 1e-7: Harris, Delta = min (x_i + 1e-6) / u_i, then the
//      largest |u_i| among x_i / u_i <= Delta (lowest i on ties), theta = max(0, x_p / u_p); under
//      Bland the smallest x_i / u_i (lowest i on ties); no eligible row: Unbounded;
//    x_B -= theta u, x_p = theta, pivot.
//  Between the phases: IterationLimit ends the route; artificial sum > 1e-6: Infeasible; every
//  row still holding an artificial takes the lowest nonbasic structural column j with
//  |(B^-1 a_j)_i| > 1e-7 (theta = 0; none: IterationLimit, the reference's report for a
//  rank-deficient model); B^-1 of the unnegated rows = B^-1 diag(s); a basic value below -1e-6
//  then ends as Infeasible.  Phase II runs on the structural columns with the costs c.
// Outputs: status (0 Optimal, 1 Infeasible, 2 Unbounded, 3 IterationLimit), the basis (y
// layout), x (x layout; nonbasic 0), c^T x, the duals B^-T c_B of the final basis (artificial
// rows cost 0), and the iterations of each phase.
// Work vectors: N(0) basic flags of the structural columns, N(1) their Devex weights, N(2) x of
// the basic columns; M(0) u, M(1) x_B, M(2) y, M(3) row signs, M(4) basic flags of the
// artificials, M(5) their weights, M(6) residuals.
constexpr double kSxFeas = 1e-6;     // math::kDefaultFeasibilityTol
constexpr double kSxPivot = 1e-7;    // math::kDefaultPivotTol
constexpr int kSxBland = 50;         // consecutive degenerate steps before Bland's rule
constexpr int kSxRefresh = 50;       // iterations between recomputations of y

__global__ __launch_bounds__(256) void k_simplex(lkb_args A, lks_args S) {
    __shared__ int brow[LKB_MAX_M];          // basic column of each row (n + i: artificial i)
    __shared__ double rho[LKB_MAX_M];        // pivot row of B^-1
    __shared__ KeyIdx ski[XO_T];
    __shared__ double sd[XO_T];
    __shared__ double s_rq;
    const int li = blockIdx.x;
    if (li >= A.count) return;
    const lkb_lp d = A.lps[li];
    const Lp L = make_lp(A, d);
    const int m = L.m, n = L.n, t = threadIdx.x;
    const int64_t ld = L.ld;
    double* B = L.fac;
    double* isb = L.N(0);
    double* gam = L.N(1);
    double* u = L.M(0);
    double* xb = L.M(1);
    double* y = L.M(2);
    double* sg = L.M(3);
    double* isa = L.M(4);
    double* gma = L.M(5);
    const int max_iter = S.max_iter;

    // Phase I start
    for (int j = t; j < n; j += XO_T) { isb[j] = 0.0; gam[j] = 1.0; }
    for (int i = t; i < m; i += XO_T) {
        const double bi = L.b0[i];
        sg[i] = bi < 0.0 ? -1.0 : 1.0;
        xb[i] = fabs(bi);
        isa[i] = 1.0;
        gma[i] = 1.0;
        brow[i] = n + i;
    }
    for (int64_t e = t; e < ld * m; e += XO_T) {
        const int64_t j = e / ld, i = e - j * ld;
        B[e] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();

    auto cost = [&](int j, bool p1) { return j >= n ? (p1 ? 1.0 : 0.0) : (p1 ? 0.0 : L.c0[j]); };
    auto dot_col = [&](int j, const double* v) {       // a_j^T v (current row signs)
        if (j >= n) return v[j - n];
        double s = 0.0;
        for (int k = L.cp[j]; k < L.cp[j + 1]; ++k) s += L.v0[k] * sg[L.ri[k]] * v[L.ri[k]];
        return s;
    };
    auto ftran = [&](int q) {                          // u = B^-1 a_q
        for (int i = t; i < m; i += XO_T) {
            double s = 0.0;
            if (q >= n) {
                s = B[i + ld * (q - n)];
            } else {
                for (int k = L.cp[q]; k < L.cp[q + 1]; ++k)
                    s += B[i + ld * L.ri[k]] * (L.v0[k] * sg[L.ri[k]]);
            }
            u[i] = s;
        }
        __syncthreads();
    };
    auto load_rho = [&](int r) {
        for (int j = t; j < m; j += XO_T) rho[j] = B[r + ld * j];
        __syncthreads();
    };
    auto pivot = [&](int r) {                          // Gauss-Jordan on (r, u), rho = old row r
        const double ur = u[r];
        for (int i = t; i < m; i += XO_T) {
            const double ui = u[i];
            if (i == r) {
                for (int j = 0; j < m; ++j) B[i + ld * j] = rho[j] / ur;
            } else if (ui != 0.0) {
                const double f = ui / ur;
                for (int j = 0; j < m; ++j) B[i + ld * j] -= f * rho[j];
            }
        }
        __syncthreads();
    };
    auto duals = [&](bool p1) {                        // y = B^-T c_B
        for (int j = t; j < m; j += XO_T) {
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += B[i + ld * j] * cost(brow[i], p1);
            y[j] = s;
        }
        __syncthreads();
    };
    // x_B and y refined against the basis itself (the Gauss-Jordan inverse drifts over thousands
    // of pivots): r = b' - B x_B, x_B += B^-1 r; r = c_B - B^T y, y += B^-T r
    double* xcol = L.N(2);
    double* rr = L.M(6);
    auto refine = [&](bool p1) {
        for (int j = t; j < n; j += XO_T) xcol[j] = 0.0;
        __syncthreads();
        for (int i = t; i < m; i += XO_T)
            if (brow[i] < n) xcol[brow[i]] = xb[i];
        __syncthreads();
        for (int k = t; k < m; k += XO_T) {
            double s = 0.0;
            for (int e = L.rp[k]; e < L.rp[k + 1]; ++e) s += L.v0[L.pos[e]] * xcol[L.ci[e]];
            rr[k] = sg[k] * (L.b0[k] - s);
        }
        __syncthreads();
        for (int i = t; i < m; i += XO_T)
            if (brow[i] >= n) rr[brow[i] - n] -= xb[i];   // artificial columns are unit vectors
        __syncthreads();
        for (int i = t; i < m; i += XO_T) {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += B[i + ld * j] * rr[j];
            xb[i] += s;
        }
        __syncthreads();
        for (int i = t; i < m; i += XO_T) {
            const int j = brow[i];
            double s = j >= n ? y[j - n] : 0.0;
            if (j < n)
                for (int k = L.cp[j]; k < L.cp[j + 1]; ++k) s += L.v0[k] * sg[L.ri[k]] * y[L.ri[k]];
            rr[i] = cost(j, p1) - s;
        }
        __syncthreads();
        for (int j = t; j < m; j += XO_T) {
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += B[i + ld * j] * rr[i];
            y[j] += s;
        }
        __syncthreads();
    };
    auto is_basic = [&](int j) { return j >= n ? isa[j - n] != 0.0 : isb[j] != 0.0; };
    auto gamma_of = [&](int j) -> double& { return j >= n ? gma[j - n] : gam[j]; };

    // the primal simplex on columns 0 .. ncols; returns its status, adds its iterations to *it
    auto primal = [&](bool p1, int* it) -> int {
        const int ncols = p1 ? n + m : n;
        int deg = 0;
        bool bland = false;
        for (int iter = 0; iter < max_iter; ++iter) {
            if (iter % kSxRefresh == 0) {
                duals(p1);
                refine(p1);
            }
            KeyIdx best{0.0, -1};
            for (int j = t; j < ncols; j += XO_T) {
                if (is_basic(j)) continue;
                const double rc = cost(j, p1) - dot_col(j, y);
                if (!(rc < -kSxFeas)) continue;
                if (bland) {
                    if (best.idx < 0) best = KeyIdx{(double)j, j};
                } else {
                    double g = gamma_of(j);
                    if (g <= 0.0) { g = 1.0; gamma_of(j) = 1.0; }
                    const double key = -(rc * rc) / g;
                    if (best.idx < 0 || key < best.key) best = KeyIdx{key, j};
                }
            }
            best = wg_argmin(best, ski);
            if (best.idx < 0) {                                    // Optimal
                refine(p1);
                return 0;
            }
            const int q = best.idx;
            if (t == 0) s_rq = cost(q, p1) - dot_col(q, y);
            ftran(q);
            int pr = -1;
            double theta = 0.0;
            if (bland) {
                KeyIdx rt{0.0, -1};
                for (int i = t; i < m; i += XO_T)
                    if (u[i] > kSxPivot) {
                        const double ratio = xb[i] / u[i];
                        if (rt.idx < 0 || ratio < rt.key) rt = KeyIdx{ratio, i};
                    }
                rt = wg_argmin(rt, ski);
                pr = rt.idx;
                if (pr >= 0) theta = rt.key;
            } else {
                KeyIdx dl{0.0, -1};
                for (int i = t; i < m; i += XO_T)
                    if (u[i] > kSxPivot) {
                        const double rr = (xb[i] + kSxFeas) / u[i];
                        if (dl.idx < 0 || rr < dl.key) dl = KeyIdx{rr, i};
                    }
                dl = wg_argmin(dl, ski);
                if (dl.idx >= 0) {
                    const double Delta = dl.key;
                    KeyIdx bp{0.0, -1};
                    for (int i = t; i < m; i += XO_T)
                        if (u[i] > kSxPivot && xb[i] / u[i] <= Delta) {
                            const double key = -fabs(u[i]);
                            if (bp.idx < 0 || key < bp.key) bp = KeyIdx{key, i};
                        }
                    bp = wg_argmin(bp, ski);
                    pr = bp.idx;
                    if (pr >= 0) theta = fmax(xb[pr] / u[pr], 0.0);
                }
            }
            if (pr < 0) return 2;                                  // Unbounded
            ++*it;
            if (theta < kSxFeas) {
                if (++deg > kSxBland) bland = true;
            } else {
                deg = 0;
            }
            const int leave = brow[pr];
            double uu = 0.0;
            for (int i = t; i < m; i += XO_T) uu += u[i] * u[i];
            uu = wg_sum_max(uu, false, sd);
            load_rho(pr);
            const double f = s_rq / u[pr];
            for (int j = t; j < m; j += XO_T) y[j] += f * rho[j];
            for (int i = t; i < m; i += XO_T) xb[i] = i == pr ? theta : xb[i] - theta * u[i];
            pivot(pr);
            if (t == 0) {
                gamma_of(leave) = uu > 1.0 ? uu : 1.0;
                if (leave >= n) isa[leave - n] = 0.0; else isb[leave] = 0.0;
                if (q >= n) isa[q - n] = 1.0; else isb[q] = 1.0;
                brow[pr] = q;
            }
            __syncthreads();
            if ((iter + 1) % ncols == 0) {
                for (int j = t; j < n; j += XO_T) gam[j] = 1.0;
                for (int i = t; i < m; i += XO_T) gma[i] = 1.0;
                __syncthreads();
            }
        }
        return 3;                                                  // IterationLimit
    };

    int it1 = 0, it2 = 0;
    int status = primal(true, &it1);
    if (status == 3) goto done;
    {
        double art = 0.0;
        for (int i = t; i < m; i += XO_T)
            if (brow[i] >= n && xb[i] > 0.0) art += xb[i];
        art = wg_sum_max(art, false, sd);
        if (art > kSxFeas) { status = 1; goto done; }
    }
    for (int r = 0; r < m; ++r) {                                  // artificials out
        if (brow[r] < n) continue;
        load_rho(r);
        KeyIdx pc{0.0, -1};
        for (int j = t; j < n; j += XO_T) {
            if (isb[j] != 0.0 || pc.idx >= 0) continue;
            if (fabs(dot_col(j, rho)) > kSxPivot) pc = KeyIdx{(double)j, j};
        }
        pc = wg_argmin(pc, ski);
        if (pc.idx < 0) { status = 3; goto done; }
        const int q = pc.idx;
        ftran(q);
        pivot(r);
        if (t == 0) {
            isa[brow[r] - n] = 0.0;
            isb[q] = 1.0;
            brow[r] = q;
            xb[r] = 0.0;
        }
        __syncthreads();
    }
    // Phase II on the unnegated rows: B^-1 diag(s), then s = 1
    for (int64_t e = t; e < ld * m; e += XO_T) {
        const int64_t j = e / ld, i = e - j * ld;
        if (i < m) B[e] *= sg[j];
    }
    __syncthreads();
    {
        double lo = 0.0;
        for (int i = t; i < m; i += XO_T) { sg[i] = 1.0; lo = fmin(lo, xb[i]); }
        for (int j = t; j < n; j += XO_T) gam[j] = 1.0;
        lo = -wg_sum_max(-lo, true, sd);
        if (lo < -kSxFeas) { status = 1; goto done; }
    }
    status = primal(false, &it2);

done:
    __syncthreads();
    duals(false);
    for (int pass = 0; pass < 2; ++pass) refine(false);
    double* xo = S.x + d.on;
    for (int j = t; j < n; j += XO_T) xo[j] = 0.0;
    __syncthreads();
    double obj = 0.0;
    for (int i = t; i < m; i += XO_T) {
        const int j = brow[i];
        S.basis[d.om + i] = j < n ? j : -1;
        if (j < n) {
            xo[j] = xb[i];
            obj += L.c0[j] * xb[i];
        }
    }
    obj = wg_sum_max(obj, false, sd);
    for (int i = t; i < m; i += XO_T) S.y[d.om + i] = y[i];   // B^-T c_B of the final basis
    if (t == 0) {
        S.status[li] = status;
        S.iterations[2 * li] = it1;
        S.iterations[2 * li + 1] = it2;
        S.objective[li] = obj;
    }
}

}  // namespace

extern "C" {

int lkd_batch_scratch(void** ptr, size_t* bytes) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    *bytes = (size_t)LKB_SCRATCH_BYTES;
    return 0;
}

int lkd_batch_reset(uint32_t* ctl, lkb_state* state, const int32_t* order, int32_t count,
                    int32_t restart, void* stream) {
    if (!ctl || count < 0 || (restart && (!state || !order))) return (int)hipErrorInvalidValue;
    const int grid = count > 0 ? (count + 255) / 256 : 1;
    LK_LAUNCH(k_batch_reset, grid, 256, 0, stream, ctl, state, order, count, restart);
    return lk_last_error();
}

int lkd_batch_crossover(const lkb_args* args, const lkx_args* xo, void* stream) {
    if (!args || !xo || !xo->x || !xo->basis || !xo->xv || !xo->status || !xo->pivots || args->count < 0)
        return (int)hipErrorInvalidValue;
    if (args->count == 0) return 0;
    LK_LAUNCH(k_crossover, args->count, XO_T, 0, stream, *args, *xo);
    return lk_last_error();
}

int lkd_batch_simplex(const lkb_args* args, const lks_args* sx, void* stream) {
    if (!args || !sx || !sx->x || !sx->y || !sx->basis || !sx->status || !sx->iterations || !sx->objective ||
        args->count < 0 || sx->max_iter < 0)
        return (int)hipErrorInvalidValue;
    if (args->count == 0) return 0;
    LK_LAUNCH(k_simplex, args->count, XO_T, 0, stream, *args, *sx);
    return lk_last_error();
}

int lkd_batch_solve(const lkb_args* args, int32_t grid, void* stream) {
    if (!args || grid < 1 || args->count < 0 || args->iters_per_visit < 1) return (int)hipErrorInvalidValue;
    if (args->count == 0) return 0;
    LK_LAUNCH(k_batch, grid, 256, 0, stream, *args);
    return lk_last_error();
}

}  // extern "C"

// Normal-equations assembly, KKT-solve vector kernels and the fused interior-point vector kernels.
//
// Reductions are deterministic: every fused kernel writes one partial per block and quantity
// (part[q * nblk + block]), and k_finalize reduces each quantity with a fixed tree.  The IPM
// scalars stay on the device (slots in S) and the host reads them once per decision point,
// replacing the ~20-60 Thrust reductions per iteration of src/ipm/mehrotra_device.cu:191-259.
#include "common.hpp"

namespace {

constexpr int NT = 256;
enum { OP_SUM = 0, OP_MIN = 1, OP_MAX = 2 };

__device__ __forceinline__ double op_apply(int op, double a, double b) {
    return op == OP_SUM ? a + b : (op == OP_MIN ? fmin(a, b) : fmax(a, b));
}
__device__ __forceinline__ double op_ident(int op) {
    return op == OP_SUM ? 0.0 : (op == OP_MIN ? 1e300 : -1e300);
}

// Block-wide reduction of v (all 256 threads participate); result valid in thread 0.
__device__ __forceinline__ double block_reduce(double v, int op, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v = op_apply(op, v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = op_apply(op, op_apply(op, sh[0], sh[1]), op_apply(op, sh[2], sh[3]));
    return r;
}

struct FinMeta {
    int32_t q;            // number of quantities
    int32_t nblk;
    int8_t op[200];
    int16_t slot[200];
};

__global__ __launch_bounds__(NT) void k_finalize(const double* __restrict__ part, FinMeta meta,
                                                 double* __restrict__ S) {
    __shared__ double sh[4];
    const int q = blockIdx.x;
    const int op = meta.op[q];
    double v = op_ident(op);
    for (int b = threadIdx.x; b < meta.nblk; b += NT) v = op_apply(op, v, part[q * meta.nblk + b]);
    v = block_reduce(v, op, sh);
    if (threadIdx.x == 0) S[meta.slot[q]] = v;
}

int finalize(const double* part, int nblk, int q, const int* ops, const int* slots, double* S,
             void* stream) {
    FinMeta meta;
    meta.q = q;
    meta.nblk = nblk;
    for (int i = 0; i < q; ++i) { meta.op[i] = (int8_t)ops[i]; meta.slot[i] = (int16_t)slots[i]; }
    LK_LAUNCH(k_finalize, q, NT, 0, stream, part, meta, S);
    return lk_last_error();
}

constexpr int LONG_U = 8;

// sum of val[p] x[ci[p]] over p in [b, e) by one block (valid in thread 0): LONG_U entries per
// thread and round, their loads all in flight together, summed in entry order, then the tree
__device__ __forceinline__ double long_sum(int32_t b, int32_t e, const int32_t* __restrict__ ci,
                                           const double* __restrict__ val,
                                           const double* __restrict__ x, double* sh) {
    double s = 0.0;
    for (int32_t p = b + threadIdx.x; p < e; p += NT * LONG_U) {
        double v[LONG_U];
#pragma unroll
        for (int u = 0; u < LONG_U; ++u) {
            const int32_t q = p + u * NT < e ? p + u * NT : e - 1;
            v[u] = p + u * NT < e ? val[q] * x[ci[q]] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < LONG_U; ++u) s += v[u];
    }
    return block_reduce(s, OP_SUM, sh);
}

// Pieces of the listed long rows of a row pass: block k * split + p sums the entries
// [b + len p / split, b + len (p + 1) / split) of listed row k.
__global__ __launch_bounds__(NT) void k_long_part(const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ x, lkd_long lg) {
    __shared__ double sh[4];
    const int32_t k = blockIdx.x / lg.split, p = blockIdx.x % lg.split;
    const int32_t i = lg.idx[k];
    const int64_t b0 = rp[i], len = rp[i + 1] - b0;
    const double s = long_sum((int32_t)(b0 + len * p / lg.split), (int32_t)(b0 + len * (p + 1) / lg.split),
                              ci, val, x, sh);
    if (threadIdx.x == 0) lg.part[blockIdx.x] = s;
}

// the pre-pass of a row pass over (rp, ci, val) with vector x, when its long rows are split
void long_pre(const lkd_long& lg, const int32_t* rp, const int32_t* ci, const double* val,
              const double* x, void* st) {
    if (lg.n > 0 && lg.split > 1) LK_LAUNCH(k_long_part, lg.n * lg.split, NT, 0, st, rp, ci, val, x, lg);
}

// Sparse row pass y_i = sum_p val[p] x[ci[p]] with the load balanced: rows longer than
// LKD_LONG_ROW entries (listed in `lg`, e.g. the linking rows of a block-angular matrix) take
// a whole block each, in a fixed thread-stride + tree order, or, split, the pieces of the
// pre-pass (long_pre, launched before the pass) added in order; the others take G lanes.
// f(i, y_i) runs once per row, in thread 0 of the block (long rows) or lane 0 of the group.
template <int G, typename F>
__device__ __forceinline__ void for_rows(int64_t m, const int32_t* __restrict__ rp,
                                         const int32_t* __restrict__ ci,
                                         const double* __restrict__ val,
                                         const double* __restrict__ x, const lkd_long& lg,
                                         double* sh, F&& f) {
    for (int32_t k = blockIdx.x; k < lg.n; k += gridDim.x) {
        const int32_t i = lg.idx[k];
        double s = 0.0;
        if (lg.split > 1) {
            if (threadIdx.x == 0)
                for (int p = 0; p < lg.split; ++p) s += lg.part[(int64_t)k * lg.split + p];
        } else {
            s = long_sum(rp[i], rp[i + 1], ci, val, x, sh);
        }
        if (threadIdx.x == 0) f((int64_t)i, s);
    }
    const int lane = threadIdx.x & (G - 1);
    const int64_t g0 = (blockIdx.x * (int64_t)NT + threadIdx.x) / G;
    const int64_t ng = (int64_t)gridDim.x * NT / G;
    for (int64_t i = g0; i < m; i += ng) {
        const int32_t b = rp[i], e = rp[i + 1];
        if (e - b > LKD_LONG_ROW) continue;
        double s = 0.0;
        for (int32_t p = b + lane; p < e; p += G) s += val[p] * x[ci[p]];
        s = group_sum<G>(s);
        if (lane == 0) f(i, s);
    }
}

// ------------------------------------------------------------------------------------------
// Assembly: M0_ij = sum over the pairs q of entry (i,j) of pp[q] * theta[pk[q]], each segment
// summed in the fixed order of its length class (short: one thread; mid: 16 lanes; long: one
// block).  The diagonal pass and the panel pass run the same code, so they agree bitwise.
constexpr int SEG_LANES = 16;

__device__ __forceinline__ double seg_short(int64_t b, int64_t e, const double* __restrict__ pp,
                                            const int32_t* __restrict__ pk,
                                            const double* __restrict__ theta) {
    // lanes past their segment's end issue no loads (most segments hold a single pair)
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < LKD_SHORT_SEG; ++u)
        if (b + u < e) acc = fma(pp[b + u], theta[pk[b + u]], acc);
    return acc;
}

// Valid in every lane of the 16-lane group; all 16 lanes must call it.
__device__ __forceinline__ double seg_mid(int64_t b, int64_t e, int l,
                                          const double* __restrict__ pp,
                                          const int32_t* __restrict__ pk,
                                          const double* __restrict__ theta) {
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < LKD_LONG_SEG / SEG_LANES; ++u) {
        const int64_t q0 = b + l + u * SEG_LANES;
        const int64_t q = q0 < e ? q0 : e - 1;
        const double v = q0 < e ? pp[q] : 0.0;
        acc = fma(v, theta[pk[q]], acc);
    }
    return group_sum<SEG_LANES>(acc);
}

// Result valid in thread 0; all threads of the block must call it.
__device__ __forceinline__ double seg_block(int64_t b, int64_t e, const double* __restrict__ pp,
                                            const int32_t* __restrict__ pk,
                                            const double* __restrict__ theta, double* sh) {
    double acc = 0.0;
    for (int64_t q = b + threadIdx.x; q < e; q += NT) acc = fma(pp[q], theta[pk[q]], acc);
    return block_reduce(acc, OP_SUM, sh);
}

struct SegOut {              // where a segment's sum goes
    // diagonal pass (D != nullptr and panel == nullptr)
    int equil;
    double* D;
    double* m0d;
    // entry pass
    const int32_t* segi;
    const int32_t* segj;
    const double* Dp;
    const uint8_t* empty;
    double delta;
    double* mval;
};

template <bool DIAG>
__device__ __forceinline__ void seg_store(const SegOut& o, int64_t r, double acc) {
    if constexpr (DIAG) {                   // r: row
        if (o.m0d) o.m0d[r] = acc;
        double a = fabs(acc);
        if (a < 1e-12) a = 1e-12;
        o.D[r] = o.equil ? 1.0 / sqrt(a) : 1.0;
    } else {                                // r: segment
        const int32_t i = o.segi[r], j = o.segj[r];
        double v = acc * o.Dp[i] * o.Dp[j];
        if (i == j && !o.empty[i]) v += o.delta;
        o.mval[r] = v;
    }
}

// Blocks [0, nbl) take the long list, [nbl, nbl + nbm) the mid list (16 groups of 16 lanes per
// block), the rest every item of the short class.  Items are rows (DIAG, segment dseg[r]) or
// segments.
template <bool DIAG>
__global__ __launch_bounds__(NT) void k_segs(int64_t count, const int64_t* __restrict__ dseg,
                                             const int64_t* __restrict__ segptr,
                                             const double* __restrict__ pp,
                                             const int32_t* __restrict__ pk,
                                             const double* __restrict__ theta, lkd_long mid,
                                             lkd_long lng, int nbl, int nbm, SegOut o) {
    __shared__ double sh[4];
    auto seg_of = [&](int64_t r) -> int64_t { return DIAG ? dseg[r] : r; };
    const int bx = blockIdx.x;
    if (bx < nbl) {
        for (int32_t k = bx; k < lng.n; k += nbl) {
            const int64_t r = lng.idx[k], t = seg_of(r);
            const double acc = seg_block(segptr[t], segptr[t + 1], pp, pk, theta, sh);
            if (threadIdx.x == 0) seg_store<DIAG>(o, r, acc);
        }
        return;
    }
    if (bx < nbl + nbm) {
        const int l = threadIdx.x % SEG_LANES;
        for (int32_t k = (bx - nbl) * (NT / SEG_LANES) + threadIdx.x / SEG_LANES; k < mid.n;
             k += nbm * (NT / SEG_LANES)) {
            const int64_t r = mid.idx[k], t = seg_of(r);
            const double acc = seg_mid(segptr[t], segptr[t + 1], l, pp, pk, theta);
            if (l == 0) seg_store<DIAG>(o, r, acc);
        }
        return;
    }
    // two items per thread and round, written out so that both items' loads are in flight
    const int64_t stride = (int64_t)(gridDim.x - nbl - nbm) * NT;
    for (int64_t r0 = (bx - nbl - nbm) * (int64_t)NT + threadIdx.x; r0 < count; r0 += 2 * stride) {
        const int64_t r1 = r0 + stride;
        const bool has1 = r1 < count;
        const int64_t t0 = seg_of(r0), t1 = has1 ? seg_of(r1) : -1;
        const int64_t b0 = t0 >= 0 ? segptr[t0] : 0, e0 = t0 >= 0 ? segptr[t0 + 1] : 0;
        const int64_t b1 = t1 >= 0 ? segptr[t1] : 0, e1 = t1 >= 0 ? segptr[t1 + 1] : 0;
        const bool s0 = t0 >= 0 && e0 - b0 <= LKD_SHORT_SEG, s1 = t1 >= 0 && e1 - b1 <= LKD_SHORT_SEG;
        const double v0 = s0 ? seg_short(b0, e0, pp, pk, theta) : 0.0;
        const double v1 = s1 ? seg_short(b1, e1, pp, pk, theta) : 0.0;
        if (s0 || t0 < 0) seg_store<DIAG>(o, r0, v0);      // t < 0: empty row (DIAG only)
        if (has1 && (s1 || t1 < 0)) seg_store<DIAG>(o, r1, v1);
    }
}

template <bool DIAG>
void launch_segs(int64_t count, const int64_t* dseg, const int64_t* segptr, const double* pp,
                const int32_t* pk, const double* theta, lkd_long mid, lkd_long lng,
                const SegOut& o, void* stream) {
    const int nbl = lng.n < 1024 ? lng.n : 1024;
    const int64_t gm = (mid.n + NT / SEG_LANES - 1) / (NT / SEG_LANES);
    const int nbm = (int)(gm < 4096 ? gm : 4096);
    int64_t gs = (count + NT - 1) / NT;
    gs = gs < 1 ? 1 : (gs > 8192 ? 8192 : gs);
    LK_LAUNCH(k_segs<DIAG>, (int)(nbl + nbm + gs), NT, 0, stream, count, dseg, segptr, pp, pk,
              theta, mid, lng, nbl, nbm, o);
}

// One block per panel chunk: zeros and the chunk's entries in LDS, then whole-line stores, so
// every HBM line of the panel is written once and in full.
template <typename T>
__global__ __launch_bounds__(NT) void k_panel_chunks(int64_t panel_elems,
                                                     const int64_t* __restrict__ chunk_seg,
                                                     const int64_t* __restrict__ segdst,
                                                     const double* __restrict__ mval,
                                                     T* __restrict__ panel) {
    constexpr int CH = LKD_PANEL_CHUNK;
    constexpr int V = 16 / sizeof(T);                 // elements per 16-byte store
    typedef T vec __attribute__((ext_vector_type(V)));
    __shared__ __align__(16) T buf[CH];
    const int64_t c = blockIdx.x, base = c * CH;
    vec* bv = reinterpret_cast<vec*>(buf);
    for (int e = threadIdx.x; e < CH / V; e += NT) bv[e] = vec{};
    __syncthreads();
    const int64_t t1 = chunk_seg[c + 1];
    for (int64_t t = chunk_seg[c] + threadIdx.x; t < t1; t += NT)
        buf[segdst[t] - base] = (T)mval[t];
    __syncthreads();
    const int64_t rem = panel_elems - base;
    if (rem >= CH) {
        vec* pv = reinterpret_cast<vec*>(panel + base);
        for (int e = threadIdx.x; e < CH / V; e += NT) pv[e] = bv[e];
    } else {
        for (int e = threadIdx.x; e < rem; e += NT) panel[base + e] = buf[e];
    }
}

int grid_for(int64_t n, int64_t cap = 8192) {
    int64_t b = (n + NT - 1) / NT;
    if (b < 1) b = 1;
    return (int)(b < cap ? b : cap);
}

// ------------------------------------------------------------------------------------------
// KKT solve vectors
__global__ __launch_bounds__(NT) void k_gather(int32_t m, const int32_t* __restrict__ P,
                                               const double* __restrict__ rhs,
                                               const double* __restrict__ D,
                                               double* __restrict__ r, double* __restrict__ req) {
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < m; i += (int64_t)gridDim.x * NT) {
        const double v = rhs[P[i]];
        r[i] = v;
        req[i] = v * D[i];
    }
}

__global__ __launch_bounds__(NT) void k_kupdate(int32_t m, const double* __restrict__ D,
                                                const double* __restrict__ z,
                                                double* __restrict__ dy,
                                                double* __restrict__ backup, int acc) {
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < m; i += (int64_t)gridDim.x * NT) {
        const double corr = z[i] * D[i];
        if (acc) {
            const double old = dy[i];
            if (backup) backup[i] = old;
            dy[i] = old + corr;
        } else {
            dy[i] = corr;
        }
    }
}

// u_j = theta_j * sum_i A_p[i][j] dy_i   (column pass over the CSC with permuted row indices)
template <int G>
__global__ __launch_bounds__(NT) void k_res_cols(int32_t n, const int32_t* __restrict__ colptr,
                                                 const int32_t* __restrict__ prow,
                                                 const double* __restrict__ cval,
                                                 const double* __restrict__ theta,
                                                 const double* __restrict__ dy,
                                                 double* __restrict__ u, lkd_long lc) {
    __shared__ double sh[4];
    for_rows<G>(n, colptr, prow, cval, dy, lc, sh,
                [&](int64_t j, double s) { u[j] = theta[j] * s; });
}

// r_i = rhs_i - sum_j A_p[i][j] u_j ; req_i = D_i r_i ; partial max |r|, |rhs|, |dy|, M0_ii
template <int G>
__global__ __launch_bounds__(NT) void k_res_rows(int32_t m, const int32_t* __restrict__ prp,
                                                 const int32_t* __restrict__ pci,
                                                 const double* __restrict__ pval,
                                                 const double* __restrict__ u,
                                                 const double* __restrict__ rhs,
                                                 const double* __restrict__ dy,
                                                 const double* __restrict__ m0d,
                                                 double* __restrict__ r,
                                                 const double* __restrict__ D,
                                                 double* __restrict__ req,
                                                 double* __restrict__ part, lkd_long lr) {
    __shared__ double sh[4];
    double mr = 0.0, mb = 0.0, md = 0.0, mm = 0.0;
    for_rows<G>(m, prp, pci, pval, u, lr, sh, [&](int64_t i, double s) {
        const double v = rhs[i] - s;
        r[i] = v;
        req[i] = v * D[i];
        mr = fmax(mr, fabs(v));
        mb = fmax(mb, fabs(rhs[i]));
        md = fmax(md, fabs(dy[i]));
        mm = fmax(mm, m0d[i]);
    });
    const int nb = gridDim.x;
    double t;
    t = block_reduce(mr, OP_MAX, sh); if (threadIdx.x == 0) part[0 * nb + blockIdx.x] = t;
    t = block_reduce(mb, OP_MAX, sh); if (threadIdx.x == 0) part[1 * nb + blockIdx.x] = t;
    t = block_reduce(md, OP_MAX, sh); if (threadIdx.x == 0) part[2 * nb + blockIdx.x] = t;
    t = block_reduce(mm, OP_MAX, sh); if (threadIdx.x == 0) part[3 * nb + blockIdx.x] = t;
}

// q_i = sum_j A_p[i][j] u_j ; partial p^T q
template <int G>
__global__ __launch_bounds__(NT) void k_pcg_rows(int32_t m, const int32_t* __restrict__ prp,
                                                 const int32_t* __restrict__ pci,
                                                 const double* __restrict__ pval,
                                                 const double* __restrict__ u,
                                                 const double* __restrict__ p,
                                                 double* __restrict__ q,
                                                 double* __restrict__ part, lkd_long lr) {
    __shared__ double sh[4];
    double acc = 0.0;
    for_rows<G>(m, prp, pci, pval, u, lr, sh, [&](int64_t i, double s) {
        q[i] = s;
        acc += p[i] * s;
    });
    acc = block_reduce(acc, OP_SUM, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(NT) void k_pcg_update(int32_t m, const double* __restrict__ kS,
                                                   int32_t srz, int32_t spq,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ q,
                                                   double* __restrict__ dy,
                                                   double* __restrict__ r,
                                                   const double* __restrict__ D,
                                                   double* __restrict__ req,
                                                   double* __restrict__ part) {
    __shared__ double sh[4];
    const double a = kS[srz] / kS[spq];
    double mr = 0.0, md = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < m; i += (int64_t)gridDim.x * NT) {
        const double d = dy[i] + a * p[i];
        const double v = r[i] - a * q[i];
        dy[i] = d;
        r[i] = v;
        req[i] = v * D[i];
        mr = fmax(mr, fabs(v));
        md = fmax(md, fabs(d));
    }
    const int nb = gridDim.x;
    double t;
    t = block_reduce(mr, OP_MAX, sh); if (threadIdx.x == 0) part[0 * nb + blockIdx.x] = t;
    t = block_reduce(md, OP_MAX, sh); if (threadIdx.x == 0) part[1 * nb + blockIdx.x] = t;
}

__global__ __launch_bounds__(NT) void k_pcg_precdot(int32_t m, const double* __restrict__ D,
                                                    double* __restrict__ z,
                                                    const double* __restrict__ r,
                                                    double* __restrict__ part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < m; i += (int64_t)gridDim.x * NT) {
        const double v = z[i] * D[i];
        z[i] = v;
        acc += r[i] * v;
    }
    acc = block_reduce(acc, OP_SUM, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(NT) void k_pcg_dir(int32_t m, const double* __restrict__ kS,
                                                int32_t snew, int32_t sold,
                                                const double* __restrict__ z,
                                                double* __restrict__ p) {
    const double beta = kS[snew] / kS[sold];
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < m; i += (int64_t)gridDim.x * NT)
        p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(NT) void k_scatter(int32_t m, const int32_t* __restrict__ P,
                                                const double* __restrict__ dy,
                                                double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < m; i += (int64_t)gridDim.x * NT)
        out[P[i]] = dy[i];
}

template <typename T>
__global__ __launch_bounds__(NT) void k_panel_csr(int64_t nnz, const int64_t* __restrict__ idx,
                                                  const T* __restrict__ panel,
                                                  double* __restrict__ vals) {
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * NT)
        vals[i] = (double)panel[idx[i]];
}

// ------------------------------------------------------------------------------------------
// Interior-point vector kernels.  Row passes use G lanes per row of A, column passes G lanes per
// row of A^T.  Blocks write partials part[q * nblk + block].
template <int G>
__global__ __launch_bounds__(NT) void k_resid_rows(lkd_ipm P, const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   double* __restrict__ rp) {
    __shared__ double sh[4];
    double s2 = 0.0, inf = 0.0, by = 0.0;
    for_rows<G>(P.m, P.arp, P.aci, P.aval, x, P.long_rows, sh, [&](int64_t i, double ax) {
        const double v = ax + (-P.b[i]);
        rp[i] = v;
        s2 += v * v;
        inf = fmax(inf, fabs(v / P.R[i]));
        by += P.b[i] * y[i];
    });
    const int nb = gridDim.x;
    double r;
    r = block_reduce(s2, OP_SUM, sh); if (threadIdx.x == 0) P.part[0 * nb + blockIdx.x] = r;
    r = block_reduce(inf, OP_MAX, sh); if (threadIdx.x == 0) P.part[1 * nb + blockIdx.x] = r;
    r = block_reduce(by, OP_SUM, sh); if (threadIdx.x == 0) P.part[2 * nb + blockIdx.x] = r;
}

template <int G>
__global__ __launch_bounds__(NT) void k_resid_cols(lkd_ipm P, const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ s,
                                                   double* __restrict__ rd,
                                                   double* __restrict__ theta, int qoff) {
    __shared__ double sh[4];
    double s2 = 0.0, xs = 0.0, cx = 0.0, inf = 0.0, xsmax = 0.0, thmax = -1e300, thmin = 1e300;
    for_rows<G>(P.n, P.trp, P.tci, P.tval, y, P.long_cols, sh, [&](int64_t j, double aty) {
        const double v = aty + (s[j] - P.c[j]);
        const double t = x[j] / s[j];
        rd[j] = v;
        if (theta) theta[j] = t;
        thmax = fmax(thmax, t);
        thmin = fmin(thmin, t);
        s2 += v * v;
        xs += x[j] * s[j];
        cx += P.c[j] * x[j];
        inf = fmax(inf, fabs(v / P.C[j]));
        xsmax = fmax(xsmax, fabs(x[j] * (s[j] - v)));   // x_j (c - A^T y)_j, scale-free
    });
    const int nb = gridDim.x;
    double* part = P.part + (int64_t)qoff * nb;
    double r;
    r = block_reduce(s2, OP_SUM, sh); if (threadIdx.x == 0) part[0 * nb + blockIdx.x] = r;
    r = block_reduce(xs, OP_SUM, sh); if (threadIdx.x == 0) part[1 * nb + blockIdx.x] = r;
    r = block_reduce(cx, OP_SUM, sh); if (threadIdx.x == 0) part[2 * nb + blockIdx.x] = r;
    r = block_reduce(inf, OP_MAX, sh); if (threadIdx.x == 0) part[3 * nb + blockIdx.x] = r;
    r = block_reduce(xsmax, OP_MAX, sh); if (threadIdx.x == 0) part[4 * nb + blockIdx.x] = r;
    r = block_reduce(thmax, OP_MAX, sh); if (threadIdx.x == 0) part[5 * nb + blockIdx.x] = r;
    r = block_reduce(thmin, OP_MIN, sh); if (threadIdx.x == 0) part[6 * nb + blockIdx.x] = r;
}

__global__ __launch_bounds__(NT) void k_rhs_cols(lkd_ipm P, const double* __restrict__ x,
                                                 const double* __restrict__ s,
                                                 const double* __restrict__ theta,
                                                 const double* __restrict__ rd,
                                                 const double* __restrict__ dxa,
                                                 const double* __restrict__ dsa, double smu,
                                                 double* __restrict__ rxs,
                                                 double* __restrict__ v) {
    for (int64_t j = blockIdx.x * (int64_t)NT + threadIdx.x; j < P.n; j += (int64_t)gridDim.x * NT) {
        const double aff = dxa ? dxa[j] * dsa[j] : 0.0;
        const double r = (-x[j] * s[j] + smu) - aff;
        rxs[j] = r;
        v[j] = -theta[j] * rd[j] - r / s[j];
    }
}

template <int G>
__global__ __launch_bounds__(NT) void k_rkkt_rows(lkd_ipm P, const double* __restrict__ v,
                                                  const double* __restrict__ rp,
                                                  double* __restrict__ rkkt) {
    __shared__ double sh[4];
    for_rows<G>(P.m, P.arp, P.aci, P.aval, v, P.long_rows, sh,
                [&](int64_t i, double av) { rkkt[i] = av + (-rp[i]); });
}

template <int G>
__global__ __launch_bounds__(NT) void k_dir_cols(lkd_ipm P, const double* __restrict__ x,
                                                 const double* __restrict__ s,
                                                 const double* __restrict__ theta,
                                                 const double* __restrict__ rd,
                                                 const double* __restrict__ rxs,
                                                 const double* __restrict__ dy,
                                                 double* __restrict__ dx, double* __restrict__ ds,
                                                 double* __restrict__ qd) {
    __shared__ double sh[4];
    double r2 = 0, r3 = 0, rx = 0, apm = 1e30, adm = 1e30, dx2 = 0, ds2 = 0;
    for_rows<G>(P.n, P.trp, P.tci, P.tval, dy, P.long_cols, sh, [&](int64_t j, double t) {
        const double dsj = (-t) + (-rd[j]);
        const double dxj = -theta[j] * dsj + rxs[j] / s[j];
        ds[j] = dsj;
        dx[j] = dxj;
        qd[j] = t + dsj;
        const double e2 = t + (dsj + rd[j]);
        const double e3 = s[j] * dxj + x[j] * dsj - rxs[j];
        r2 += e2 * e2;
        r3 += e3 * e3;
        rx += rxs[j] * rxs[j];
        if (dxj < 0.0) apm = fmin(apm, -x[j] / dxj);
        if (dsj < 0.0) adm = fmin(adm, -s[j] / dsj);
        dx2 += dxj * dxj;
        ds2 += dsj * dsj;
    });
    const int nb = gridDim.x;
    double r;
    r = block_reduce(r2, OP_SUM, sh); if (threadIdx.x == 0) P.part[0 * nb + blockIdx.x] = r;
    r = block_reduce(r3, OP_SUM, sh); if (threadIdx.x == 0) P.part[1 * nb + blockIdx.x] = r;
    r = block_reduce(rx, OP_SUM, sh); if (threadIdx.x == 0) P.part[2 * nb + blockIdx.x] = r;
    r = block_reduce(apm, OP_MIN, sh); if (threadIdx.x == 0) P.part[3 * nb + blockIdx.x] = r;
    r = block_reduce(adm, OP_MIN, sh); if (threadIdx.x == 0) P.part[4 * nb + blockIdx.x] = r;
    r = block_reduce(dx2, OP_SUM, sh); if (threadIdx.x == 0) P.part[5 * nb + blockIdx.x] = r;
    r = block_reduce(ds2, OP_SUM, sh); if (threadIdx.x == 0) P.part[6 * nb + blockIdx.x] = r;
}

template <int G>
__global__ __launch_bounds__(NT) void k_dir_rows(lkd_ipm P, const double* __restrict__ dx,
                                                 const double* __restrict__ rp,
                                                 const double* __restrict__ dy,
                                                 double* __restrict__ qp, int qoff) {
    __shared__ double sh[4];
    double r1 = 0, dy2 = 0;
    for_rows<G>(P.m, P.arp, P.aci, P.aval, dx, P.long_rows, sh, [&](int64_t i, double adx) {
        qp[i] = adx;
        const double e1 = adx + rp[i];
        r1 += e1 * e1;
        dy2 += dy[i] * dy[i];
    });
    const int nb = gridDim.x;
    double* part = P.part + (int64_t)qoff * nb;
    double r;
    r = block_reduce(r1, OP_SUM, sh); if (threadIdx.x == 0) part[0 * nb + blockIdx.x] = r;
    r = block_reduce(dy2, OP_SUM, sh); if (threadIdx.x == 0) part[1 * nb + blockIdx.x] = r;
}

__global__ __launch_bounds__(NT) void k_muaff(lkd_ipm P, const double* __restrict__ x,
                                              const double* __restrict__ s,
                                              const double* __restrict__ dxa,
                                              const double* __restrict__ dsa, double ap,
                                              double ad, int from_slots) {
    __shared__ double sh[4];
    if (from_slots) {   // the affine step lengths min(1, APMAX), min(1, ADMAX) from the slots
        const double a = P.S[LKS_APMAX], b = P.S[LKS_ADMAX];
        ap = a < 1.0 ? a : 1.0;
        ad = b < 1.0 ? b : 1.0;
    }
    double acc = 0.0;
    for (int64_t j = blockIdx.x * (int64_t)NT + threadIdx.x; j < P.n; j += (int64_t)gridDim.x * NT) {
        const double tx = x[j] + ap * dxa[j];
        const double ts = s[j] + ad * dsa[j];
        acc += tx * ts;
    }
    acc = block_reduce(acc, OP_SUM, sh);
    if (threadIdx.x == 0) P.part[blockIdx.x] = acc;
}

// All backtracking trials in one pass (rp_t = rp + ap_t A dx, rd_t = rd + ad_t (A^T dy + ds)).
// Per trial, every wavefront reduces its lanes' values at once (fixed butterfly) and lane 0
// folds them into the wavefront's LDS row; one barrier at the end combines the four rows in a
// fixed order.  No per-thread arrays indexed by the trial.
__device__ __forceinline__ void trial_fold(double* row, int q, double v, int op, bool first) {
    v = op == OP_SUM ? wave_sum(v) : (op == OP_MIN ? wave_min(v) : wave_max(v));
    if ((threadIdx.x & 63) == 0) row[q] = first ? v : op_apply(op, row[q], v);
}

__global__ __launch_bounds__(NT) void k_trials_cols(lkd_ipm P, const double* __restrict__ x,
                                                    const double* __restrict__ s,
                                                    const double* __restrict__ dx,
                                                    const double* __restrict__ ds,
                                                    const double* __restrict__ rd,
                                                    const double* __restrict__ qd, double ap,
                                                    double ad, int nt) {
    __shared__ double acc[4][5 * LKD_MAX_TRIALS];
    double* row = acc[threadIdx.x >> 6];
    bool first = true;
    const int64_t stride = (int64_t)gridDim.x * NT;
    for (int64_t j0 = blockIdx.x * (int64_t)NT; j0 < P.n; j0 += stride) {   // uniform per block
        const int64_t j = j0 + threadIdx.x;
        const bool ok = j < P.n;
        const int64_t jc = ok ? j : P.n - 1;
        const double xj = x[jc], sj = s[jc], dxj = dx[jc], dsj = ds[jc], rdj = rd[jc], qdj = qd[jc];
        const double cj = P.C[jc];
        double a = ap, b = ad;
        for (int t = 0; t < nt; ++t) {
            const double tx = xj + a * dxj, ts = sj + b * dsj, tr = rdj + b * qdj;
            trial_fold(row, 5 * t + 0, ok ? tr * tr : 0.0, OP_SUM, first);
            trial_fold(row, 5 * t + 1, ok ? tx * ts : 0.0, OP_SUM, first);
            trial_fold(row, 5 * t + 2, ok ? tx : 1e30, OP_MIN, first);
            trial_fold(row, 5 * t + 3, ok ? ts : 1e30, OP_MIN, first);
            trial_fold(row, 5 * t + 4, ok ? fabs(tr / cj) : 0.0, OP_MAX, first);   // original space
            a *= 0.5;
            b *= 0.5;
        }
        first = false;
    }
    __syncthreads();
    const int nb = gridDim.x;
    for (int q = threadIdx.x; q < 5 * nt; q += NT) {
        const int k = q % 5;
        const int op = k < 2 ? OP_SUM : (k < 4 ? OP_MIN : OP_MAX);
        const double r = first ? op_ident(op)
                               : op_apply(op, op_apply(op, acc[0][q], acc[1][q]), op_apply(op, acc[2][q], acc[3][q]));
        P.part[(int64_t)q * nb + blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(NT) void k_trials_rows(lkd_ipm P, const double* __restrict__ rp,
                                                    const double* __restrict__ qp,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ dy, double ap,
                                                    double ad, int nt, int qoff) {
    __shared__ double acc[4][3 * LKD_MAX_TRIALS];
    double* row = acc[threadIdx.x >> 6];
    bool first = true;
    const int64_t stride = (int64_t)gridDim.x * NT;
    for (int64_t i0 = blockIdx.x * (int64_t)NT; i0 < P.m; i0 += stride) {   // uniform per block
        const int64_t i = i0 + threadIdx.x;
        const bool ok = i < P.m;
        const int64_t ic = ok ? i : P.m - 1;
        const double rpi = rp[ic], qpi = qp[ic], yi = y[ic], dyi = dy[ic], ri = P.R[ic];
        double a = ap, b = ad;
        for (int t = 0; t < nt; ++t) {
            const double tr = rpi + a * qpi, ty = yi + b * dyi;
            trial_fold(row, 3 * t + 0, ok ? tr * tr : 0.0, OP_SUM, first);
            trial_fold(row, 3 * t + 1, ok ? ty * ty : 0.0, OP_SUM, first);
            trial_fold(row, 3 * t + 2, ok ? fabs(tr / ri) : 0.0, OP_MAX, first);   // original space
            a *= 0.5;
            b *= 0.5;
        }
        first = false;
    }
    __syncthreads();
    const int nb = gridDim.x;
    double* part = P.part + (int64_t)qoff * nb;
    for (int q = threadIdx.x; q < 3 * nt; q += NT) {
        const int op = q % 3 < 2 ? OP_SUM : OP_MAX;
        const double r = first ? op_ident(op)
                               : op_apply(op, op_apply(op, acc[0][q], acc[1][q]), op_apply(op, acc[2][q], acc[3][q]));
        part[(int64_t)q * nb + blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(NT) void k_commit(int32_t m, int32_t n, double* __restrict__ x,
                                               double* __restrict__ s, double* __restrict__ y,
                                               const double* __restrict__ dx,
                                               const double* __restrict__ ds,
                                               const double* __restrict__ dy, double ap,
                                               double ad) {
    const int64_t len = m > n ? m : n;
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < len; i += (int64_t)gridDim.x * NT) {
        if (i < n) { x[i] += ap * dx[i]; s[i] += ad * ds[i]; }
        if (i < m) y[i] += ad * dy[i];
    }
}

__global__ __launch_bounds__(NT) void k_init(int32_t m, int32_t n, double* __restrict__ x,
                                             double* __restrict__ y, double* __restrict__ s) {
    const int64_t len = m > n ? m : n;
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < len; i += (int64_t)gridDim.x * NT) {
        if (i < n) { x[i] = 1.0; s[i] = 1.0; }
        if (i < m) y[i] = 0.0;
    }
}

#define LK_GSWITCH(G, KERNEL, ...)                                                    \
    switch (G) {                                                                      \
        case 1: KERNEL<1>(__VA_ARGS__); break;                                        \
        case 2: KERNEL<2>(__VA_ARGS__); break;                                        \
        case 4: KERNEL<4>(__VA_ARGS__); break;                                        \
        case 8: KERNEL<8>(__VA_ARGS__); break;                                        \
        case 16: KERNEL<16>(__VA_ARGS__); break;                                      \
        case 32: KERNEL<32>(__VA_ARGS__); break;                                      \
        default: KERNEL<64>(__VA_ARGS__); break;                                      \
    }

template <int G> void l_resid_rows(const lkd_ipm* P, const double* x, const double* y, double* rp, void* st) {
    long_pre(P->long_rows, P->arp, P->aci, P->aval, x, st);
    LK_LAUNCH(k_resid_rows<G>, P->nblk_rows, NT, 0, st, *P, x, y, rp);
}
template <int G> void l_resid_cols(const lkd_ipm* P, const double* x, const double* y, const double* s,
                                   double* rd, double* th, int qoff, void* st) {
    long_pre(P->long_cols, P->trp, P->tci, P->tval, y, st);
    LK_LAUNCH(k_resid_cols<G>, P->nblk_cols, NT, 0, st, *P, x, y, s, rd, th, qoff);
}
template <int G> void l_rkkt(const lkd_ipm* P, const double* v, const double* rp, double* rk, void* st) {
    long_pre(P->long_rows, P->arp, P->aci, P->aval, v, st);
    LK_LAUNCH(k_rkkt_rows<G>, P->nblk_rows, NT, 0, st, *P, v, rp, rk);
}
template <int G> void l_dir_cols(const lkd_ipm* P, const double* x, const double* s, const double* th,
                                 const double* rd, const double* rxs, const double* dy, double* dx,
                                 double* ds, double* qd, void* st) {
    long_pre(P->long_cols, P->trp, P->tci, P->tval, dy, st);
    LK_LAUNCH(k_dir_cols<G>, P->nblk_cols, NT, 0, st, *P, x, s, th, rd, rxs, dy, dx, ds, qd);
}
template <int G> void l_dir_rows(const lkd_ipm* P, const double* dx, const double* rp, const double* dy,
                                 double* qp, int qoff, void* st) {
    long_pre(P->long_rows, P->arp, P->aci, P->aval, dx, st);
    LK_LAUNCH(k_dir_rows<G>, P->nblk_rows, NT, 0, st, *P, dx, rp, dy, qp, qoff);
}
template <int G> void l_res_rows(int32_t m, const int32_t* prp, const int32_t* pci, const double* pval,
                                 const double* u, const double* rhs, const double* dy, const double* m0d,
                                 double* r, const double* D, double* req, double* part, int nblk,
                                 lkd_long lr, void* st) {
    long_pre(lr, prp, pci, pval, u, st);
    LK_LAUNCH(k_res_rows<G>, nblk, NT, 0, st, m, prp, pci, pval, u, rhs, dy, m0d, r, D, req, part, lr);
}
template <int G> void l_pcg_rows(int32_t m, const int32_t* prp, const int32_t* pci, const double* pval,
                                 const double* u, const double* p, double* q, double* part, int nblk,
                                 lkd_long lr, void* st) {
    long_pre(lr, prp, pci, pval, u, st);
    LK_LAUNCH(k_pcg_rows<G>, nblk, NT, 0, st, m, prp, pci, pval, u, p, q, part, lr);
}

}  // namespace

extern "C" {

int lkd_assemble_diag(int32_t m, const int64_t* dseg, const int64_t* segptr, const double* pp,
                      const int32_t* pk, const double* theta, int equilibrate, double* D,
                      double* m0diag, lkd_long mid_diag, lkd_long long_diag, void* stream) {
    if (m <= 0) return 0;
    SegOut o{};
    o.equil = equilibrate;
    o.D = D;
    o.m0d = m0diag;
    launch_segs<true>(m, dseg, segptr, pp, pk, theta, mid_diag, long_diag, o, stream);
    return lk_last_error();
}

int lkd_assemble_pairs(int prec, int64_t nseg, const int64_t* segptr, const int64_t* segdst,
                       const int32_t* segi, const int32_t* segj, const double* pp,
                       const int32_t* pk, const double* theta, const double* D,
                       const uint8_t* empty_row, double delta, lkd_long mid_segs,
                       lkd_long long_segs, double* mval, int64_t panel_elems,
                       const int64_t* chunk_seg, void* panel, void* stream) {
    if (nseg > 0) {
        SegOut o{};
        o.segi = segi;
        o.segj = segj;
        o.Dp = D;
        o.empty = empty_row;
        o.delta = delta;
        o.mval = mval;
        launch_segs<false>(nseg, nullptr, segptr, pp, pk, theta, mid_segs, long_segs, o, stream);
    }
    const int64_t nchunk = (panel_elems + LKD_PANEL_CHUNK - 1) / LKD_PANEL_CHUNK;
    if (nchunk >= (int64_t(1) << 31)) return (int)hipErrorInvalidValue;
    if (prec == LKD_F32)
        LK_LAUNCH(k_panel_chunks<float>, (int)nchunk, NT, 0, stream, panel_elems, chunk_seg,
                  segdst, mval, (float*)panel);
    else
        LK_LAUNCH(k_panel_chunks<double>, (int)nchunk, NT, 0, stream, panel_elems, chunk_seg,
                  segdst, mval, (double*)panel);
    return lk_last_error();
}

int lkd_kkt_gather(int32_t m, const int32_t* P, const double* rhs, const double* D, double* r,
                   double* req, void* stream) {
    LK_LAUNCH(k_gather, grid_for(m), NT, 0, stream, m, P, rhs, D, r, req);
    return lk_last_error();
}

int lkd_kkt_update(int32_t m, const double* D, const double* z, double* dy, double* backup,
                   int accumulate, void* stream) {
    LK_LAUNCH(k_kupdate, grid_for(m), NT, 0, stream, m, D, z, dy, backup, accumulate);
    return lk_last_error();
}

int lkd_kkt_residual(int32_t m, int32_t n, const int32_t* colptr, const int32_t* prow,
                     const double* cval, const int32_t* prp, const int32_t* pci,
                     const double* pval, const double* theta, const double* dy,
                     const double* rhs_perm, const double* m0d, double* u, double* r,
                     const double* D, double* req, double* kS, lkd_long long_rows,
                     lkd_long long_cols, void* stream) {
    long_pre(long_cols, colptr, prow, cval, dy, stream);
    LK_LAUNCH(k_res_cols<4>, grid_for(4 * (int64_t)n), NT, 0, stream, n, colptr, prow, cval, theta, dy, u, long_cols);
    const int nblk = grid_for(8 * (int64_t)m, LKD_MAX_NBLK);   // 8 lanes per row
    double* part = kS + LKD_KS_PART;
    l_res_rows<8>(m, prp, pci, pval, u, rhs_perm, dy, m0d, r, D, req, part, nblk, long_rows, stream);
    int ops[4] = {OP_MAX, OP_MAX, OP_MAX, OP_MAX}, slots[4] = {0, 1, 2, 3};
    return finalize(part, nblk, 4, ops, slots, kS, stream);
}

int lkd_pcg_matvec(int32_t m, int32_t n, const int32_t* colptr, const int32_t* prow,
                   const double* cval, const int32_t* prp, const int32_t* pci,
                   const double* pval, const double* theta, const double* p, double* u,
                   double* q, double* kS, int32_t slot_pq, lkd_long long_rows, lkd_long long_cols,
                   void* stream) {
    long_pre(long_cols, colptr, prow, cval, p, stream);
    LK_LAUNCH(k_res_cols<4>, grid_for(4 * (int64_t)n), NT, 0, stream, n, colptr, prow, cval, theta, p, u, long_cols);
    const int nblk = grid_for(8 * (int64_t)m, LKD_MAX_NBLK);   // 8 lanes per row
    double* part = kS + LKD_KS_PART;
    l_pcg_rows<8>(m, prp, pci, pval, u, p, q, part, nblk, long_rows, stream);
    int op = OP_SUM;
    return finalize(part, nblk, 1, &op, &slot_pq, kS, stream);
}

int lkd_pcg_update(int32_t m, double* kS, int32_t slot_rz, int32_t slot_pq, const double* p,
                   const double* q, double* dy, double* r, const double* D, double* req,
                   void* stream) {
    const int nblk = grid_for(m, LKD_MAX_NBLK);
    double* part = kS + LKD_KS_PART;
    LK_LAUNCH(k_pcg_update, nblk, NT, 0, stream, m, kS, slot_rz, slot_pq, p, q, dy, r, D, req, part);
    int ops[2] = {OP_MAX, OP_MAX}, slots[2] = {4, 2};
    return finalize(part, nblk, 2, ops, slots, kS, stream);
}

int lkd_pcg_precdot(int32_t m, const double* D, double* z, const double* r, double* kS,
                    int32_t slot_rz, void* stream) {
    const int nblk = grid_for(m, LKD_MAX_NBLK);
    double* part = kS + LKD_KS_PART;
    LK_LAUNCH(k_pcg_precdot, nblk, NT, 0, stream, m, D, z, r, part);
    int op = OP_SUM;
    return finalize(part, nblk, 1, &op, &slot_rz, kS, stream);
}

int lkd_pcg_dir(int32_t m, const double* kS, int32_t slot_new, int32_t slot_old,
                const double* z, double* p, void* stream) {
    LK_LAUNCH(k_pcg_dir, grid_for(m), NT, 0, stream, m, kS, slot_new, slot_old, z, p);
    return lk_last_error();
}

int lkd_kkt_scatter(int32_t m, const int32_t* P, const double* dy, double* out, void* stream) {
    LK_LAUNCH(k_scatter, grid_for(m), NT, 0, stream, m, P, dy, out);
    return lk_last_error();
}

int lkd_panel_to_csr(int prec, int64_t nnz, const int64_t* src_index, const void* panel,
                     double* vals, void* stream) {
    if (nnz <= 0) return 0;
    if (prec == LKD_F32)
        LK_LAUNCH(k_panel_csr<float>, grid_for(nnz), NT, 0, stream, nnz, src_index, (const float*)panel, vals);
    else
        LK_LAUNCH(k_panel_csr<double>, grid_for(nnz), NT, 0, stream, nnz, src_index, (const double*)panel, vals);
    return lk_last_error();
}

int lkd_ipm_init(int32_t m, int32_t n, double* x, double* y, double* s, void* stream) {
    LK_LAUNCH(k_init, grid_for(m > n ? m : n), NT, 0, stream, m, n, x, y, s);
    return lk_last_error();
}

int lkd_ipm_residuals(const lkd_ipm* P, const double* x, const double* y, const double* s,
                      double* rp, double* rd, double* theta, void* stream) {
    LK_GSWITCH(P->grp_rows, l_resid_rows, P, x, y, rp, stream);
    int ops_r[3] = {OP_SUM, OP_MAX, OP_SUM}, slots_r[3] = {LKS_RP2, LKS_RPINF, LKS_BY};
    int e = finalize(P->part, P->nblk_rows, 3, ops_r, slots_r, P->S, stream);
    if (e) return e;
    // column pass writes its partials after the three row quantities
    const int qoff = (3 * P->nblk_rows + P->nblk_cols - 1) / P->nblk_cols;
    LK_GSWITCH(P->grp_cols, l_resid_cols, P, x, y, s, rd, theta, qoff, stream);
    int ops_c[7] = {OP_SUM, OP_SUM, OP_SUM, OP_MAX, OP_MAX, OP_MAX, OP_MIN};
    int slots_c[7] = {LKS_RD2, LKS_XS, LKS_CX, LKS_RDINF, LKS_XSMAX, LKS_THMAX, LKS_THMIN};
    return finalize(P->part + (int64_t)qoff * P->nblk_cols, P->nblk_cols, 7, ops_c, slots_c, P->S, stream);
}

int lkd_ipm_rhs(const lkd_ipm* P, const double* x, const double* s, const double* theta,
                const double* rd, const double* rp, const double* dxa, const double* dsa,
                double sigma_mu, double* rxs, double* v, double* rkkt, void* stream) {
    LK_LAUNCH(k_rhs_cols, P->nblk_cols, NT, 0, stream, *P, x, s, theta, rd, dxa, dsa, sigma_mu, rxs, v);
    LK_GSWITCH(P->grp_rows, l_rkkt, P, v, rp, rkkt, stream);
    return lk_last_error();
}

int lkd_ipm_direction(const lkd_ipm* P, const double* x, const double* s, const double* y,
                      const double* theta, const double* rd, const double* rp,
                      const double* rxs, const double* dy, double* dx, double* ds, double* qd,
                      double* qp, void* stream) {
    (void)y;
    LK_GSWITCH(P->grp_cols, l_dir_cols, P, x, s, theta, rd, rxs, dy, dx, ds, qd, stream);
    int ops_c[7] = {OP_SUM, OP_SUM, OP_SUM, OP_MIN, OP_MIN, OP_SUM, OP_SUM};
    int slots_c[7] = {LKS_R2, LKS_R3, LKS_RXS2, LKS_APMAX, LKS_ADMAX, LKS_DX2, LKS_DS2};
    int e = finalize(P->part, P->nblk_cols, 7, ops_c, slots_c, P->S, stream);
    if (e) return e;
    const int qoff = (7 * P->nblk_cols + P->nblk_rows - 1) / P->nblk_rows;
    LK_GSWITCH(P->grp_rows, l_dir_rows, P, dx, rp, dy, qp, qoff, stream);
    int ops_r[2] = {OP_SUM, OP_SUM}, slots_r[2] = {LKS_R1, LKS_DY2};
    return finalize(P->part + (int64_t)qoff * P->nblk_rows, P->nblk_rows, 2, ops_r, slots_r, P->S, stream);
}

int lkd_ipm_mu_aff(const lkd_ipm* P, const double* x, const double* s, const double* dxa,
                   const double* dsa, double ap, double ad, void* stream) {
    LK_LAUNCH(k_muaff, P->nblk_cols, NT, 0, stream, *P, x, s, dxa, dsa, ap, ad, 0);
    int op = OP_SUM, slot = LKS_MUAFF;
    return finalize(P->part, P->nblk_cols, 1, &op, &slot, P->S, stream);
}

int lkd_ipm_mu_aff_step(const lkd_ipm* P, const double* x, const double* s, const double* dxa,
                        const double* dsa, void* stream) {
    LK_LAUNCH(k_muaff, P->nblk_cols, NT, 0, stream, *P, x, s, dxa, dsa, 0.0, 0.0, 1);
    int op = OP_SUM, slot = LKS_MUAFF;
    return finalize(P->part, P->nblk_cols, 1, &op, &slot, P->S, stream);
}

int lkd_ipm_trials(const lkd_ipm* P, const double* x, const double* s, const double* y,
                   const double* dx, const double* ds, const double* dy, const double* rp,
                   const double* rd, const double* qp, const double* qd, double ap, double ad,
                   int32_t ntrial, void* stream) {
    if (ntrial > LKD_MAX_TRIALS) ntrial = LKD_MAX_TRIALS;
    LK_LAUNCH(k_trials_cols, P->nblk_cols, NT, 0, stream, *P, x, s, dx, ds, rd, qd, ap, ad, ntrial);
    int ops[5 * LKD_MAX_TRIALS], slots[5 * LKD_MAX_TRIALS];
    for (int t = 0; t < ntrial; ++t) {
        const int base = LKS_TRIAL_BASE + t * LKS_TRIAL_STRIDE;
        ops[5 * t + 0] = OP_SUM; slots[5 * t + 0] = base + 1;   // ||rd_t||^2
        ops[5 * t + 1] = OP_SUM; slots[5 * t + 1] = base + 2;   // x_t^T s_t
        ops[5 * t + 2] = OP_MIN; slots[5 * t + 2] = base + 3;   // min x_t
        ops[5 * t + 3] = OP_MIN; slots[5 * t + 3] = base + 4;   // min s_t
        ops[5 * t + 4] = OP_MAX; slots[5 * t + 4] = base + 6;   // ||rd_t / C||_inf
    }
    int e = finalize(P->part, P->nblk_cols, 5 * ntrial, ops, slots, P->S, stream);
    if (e) return e;
    const int qoff = (5 * ntrial * P->nblk_cols + P->nblk_rows - 1) / P->nblk_rows;
    LK_LAUNCH(k_trials_rows, P->nblk_rows, NT, 0, stream, *P, rp, qp, y, dy, ap, ad, ntrial, qoff);
    for (int t = 0; t < ntrial; ++t) {
        const int base = LKS_TRIAL_BASE + t * LKS_TRIAL_STRIDE;
        ops[3 * t + 0] = OP_SUM; slots[3 * t + 0] = base + 0;   // ||rp_t||^2
        ops[3 * t + 1] = OP_SUM; slots[3 * t + 1] = base + 5;   // ||y_t||^2
        ops[3 * t + 2] = OP_MAX; slots[3 * t + 2] = base + 7;   // ||rp_t / R||_inf
    }
    return finalize(P->part + (int64_t)qoff * P->nblk_rows, P->nblk_rows, 3 * ntrial, ops, slots, P->S, stream);
}

int lkd_ipm_commit(const lkd_ipm* P, double* x, double* s, double* y, const double* dx,
                   const double* ds, const double* dy, double ap, double ad, void* stream) {
    LK_LAUNCH(k_commit, grid_for(P->m > P->n ? P->m : P->n), NT, 0, stream, P->m, P->n, x, s, y,
              dx, ds, dy, ap, ad);
    return lk_last_error();
}

}  // extern "C"

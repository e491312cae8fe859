// Dense 64-column tile primitives shared by the supernodal factorization (chol.hip) and the
// batched dense interior-point kernels (batch.hip): the double-buffered MFMA GEMM loop over
// LDS-staged K chunks, its NT / TN products, and the in-LDS Cholesky of a 64x64 diagonal block.
// Every function assumes a 256-thread workgroup (4 wavefronts of 64).
#pragma once
#include "common.hpp"

namespace {

constexpr int TS = 64;        // block-column width, diagonal block edge
constexpr int DP = 65;        // LDS pitch of a 64x64 block
constexpr int MT = 128;       // GEMM tile rows
constexpr int NC = 64;        // GEMM tile columns
constexpr int PA = MT + 8;    // LDS pitch of the A chunk (k-major)
constexpr int PB = NC + 8;    // LDS pitch of the B chunk


template <typename T> struct KChunk;
template <> struct KChunk<double> { static constexpr int v = 16; };
template <> struct KChunk<float> { static constexpr int v = 32; };

template <typename T> struct GemmLds {
    T a[2][KChunk<T>::v * PA];
    T b[2][KChunk<T>::v * PB];
};

template <typename T> using Acc = typename Mfma<T>::acc;

template <typename T, int X, int Y>
__device__ __forceinline__ void zero_acc(Acc<T> (&acc)[X][Y]) {
#pragma unroll
    for (int x = 0; x < X; ++x)
#pragma unroll
        for (int y = 0; y < Y; ++y) acc[x][y] = Acc<T>{0, 0, 0, 0};
}

// Double-buffered K loop shared by the GEMM variants: chunks of KC columns of
// A (rows 0..ra of pa, column-major, lda) and B (rows 0..rb of pb, ldb) are staged k-major in
// LDS (columns past kend read as zero, rows past ra / rb as below); compute(As, Bs, k) consumes the
// staged chunk for k = 0, 4, .., KC-4.  Global loads run two chunks ahead in two register
// sets (chunk j in set j % 2): while chunk c is computed, chunks c + 1 and c + 2 are in
// flight, and chunk c + 1 goes to LDS right after.  One barrier per chunk; the LDS is free
// again when it returns.
// Clamp = false (the supernodal kernels): rows past ra / rb are loaded from row ra - 1 / rb - 1
// (always valid addresses) and not zeroed: a row of A only feeds the same row of the product
// and a row of B the same column, so they only reach product rows / columns the callers
// discard.  Whole chunks (the only kind when K ranges are multiples of KC, as in the supernodal
// plans) load without any per-element condition, so the loads carry no exec-mask branches and
// the waits before the LDS stores count only the chunk being stored; a partial last chunk
// zeroes its columns past kend.  Clamp = true (the batched kernel, at its register limit):
// every element goes through a clamped column and a select that zeroes it outside the range.
template <typename T, bool Clamp, typename F>
__device__ __forceinline__ void gemm_loop(const T* __restrict__ pa, int64_t lda, int ra,
                                          const T* __restrict__ pb, int64_t ldb, int rb,
                                          int kbeg, int kend, GemmLds<T>& sm, F&& compute) {
    constexpr int KC = KChunk<T>::v;
    constexpr int NA = KC * MT / 256, NB = KC * NC / 256;
    const int t = threadIdx.x;
    const int ar = t & (MT - 1), ak = t >> 7;
    const int br = t & (NC - 1), bk = t >> 6;
    const bool arow = ar < ra, brow = br < rb;
    T va[2][NA], vb[2][NB];
    const T* qa = pa + (Clamp ? (arow ? ar : 0) : min(ar, max(ra, 1) - 1));
    const T* qb = pb + (Clamp ? (brow ? br : 0) : min(br, max(rb, 1) - 1));
    auto load = [&](int set, int kc) {
        if (!Clamp && kc + KC <= kend) {
#pragma unroll
            for (int q = 0; q < NA; ++q) va[set][q] = qa[(int64_t)(kc + ak + 2 * q) * lda];
#pragma unroll
            for (int q = 0; q < NB; ++q) vb[set][q] = qb[(int64_t)(kc + bk + 4 * q) * ldb];
            return;
        }
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int k = kc + ak + 2 * q;
            const T v = qa[(int64_t)min(k, kend - 1) * lda];
            va[set][q] = ((Clamp ? arow : true) && k < kend) ? v : T(0);
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int k = kc + bk + 4 * q;
            const T v = qb[(int64_t)min(k, kend - 1) * ldb];
            vb[set][q] = ((Clamp ? brow : true) && k < kend) ? v : T(0);
        }
    };
    auto store = [&](int set, int buf) {
#pragma unroll
        for (int q = 0; q < NA; ++q) sm.a[buf][(ak + 2 * q) * PA + ar] = va[set][q];
#pragma unroll
        for (int q = 0; q < NB; ++q) sm.b[buf][(bk + 4 * q) * PB + br] = vb[set][q];
    };
    const int nch = (kend - kbeg + KC - 1) / KC;
    if (nch <= 0) return;
    load(0, kbeg);
    store(0, 0);
    if (nch > 1) load(1, kbeg + KC);
    if (nch > 2) load(0, kbeg + 2 * KC);
    __syncthreads();
    // unrolled by two so that register sets and LDS buffers are compile-time indices
    for (int c = 0; c < nch; c += 2) {
#pragma unroll
        for (int k = 0; k < KC; k += 4) compute(sm.a[0], sm.b[0], k);
        if (c + 1 < nch) store(1, 1);
        __syncthreads();
        if (c + 3 < nch) load(1, kbeg + (c + 3) * KC);
        if (c + 1 >= nch) break;
#pragma unroll
        for (int k = 0; k < KC; k += 4) compute(sm.a[1], sm.b[1], k);
        if (c + 2 < nch) store(0, 0);
        __syncthreads();
        if (c + 4 < nch) load(0, kbeg + (c + 4) * KC);
    }
}

// NT: acc[x][y] (wave w) = tile rows 32w + 16x .., columns 16y ..:
//   acc += sum_k A(r, k) B(c, k);  element e of lane l: row 32w + 16x + row(l,e), col 16y + l%16
template <typename T, bool Clamp = false>
__device__ __forceinline__ void gemm_nt(const T* __restrict__ pa, int64_t lda, int ra,
                                        const T* __restrict__ pb, int64_t ldb, int rb, int kbeg,
                                        int kend, GemmLds<T>& sm, Acc<T> (&acc)[2][4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    gemm_loop<T, Clamp>(pa, lda, ra, pb, ldb, rb, kbeg, kend, sm, [&](const T* As, const T* Bs, int k) {
        const T a0 = As[(k + lk) * PA + 32 * w + li];
        const T a1 = As[(k + lk) * PA + 32 * w + 16 + li];
        T b[4];
#pragma unroll
        for (int y = 0; y < 4; ++y) b[y] = Bs[(k + lk) * PB + 16 * y + li];
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            acc[0][y] = Mfma<T>::mma(a0, b[y], acc[0][y]);
            acc[1][y] = Mfma<T>::mma(a1, b[y], acc[1][y]);
        }
    });
}

// TN (the transposed product): acc[p][b] (wave w) holds D = B A^T for B columns 16p .. and tile
// rows 32w + 16b ..:  element e of lane l = sum_k A(32w + 16b + l%16, k) B(16p + row(l,e), k),
// i.e. tile row from the lane, column from the element -- the B-operand layout of the solve.
// Clamp = false (supernodal kernels): a wavefront skips the products of its 16-row blocks that
// lie past row ra (product rows the callers discard; their accumulators keep their initial
// values); the test is wave-uniform, so whole MFMA sequences are skipped, not masked.
template <typename T, bool Clamp = false>
__device__ __forceinline__ void gemm_tn(const T* __restrict__ pa, int64_t lda, int ra,
                                        const T* __restrict__ pb, int64_t ldb, int rb, int kbeg,
                                        int kend, GemmLds<T>& sm, Acc<T> (&acc)[4][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const bool v0 = Clamp || 32 * wu < ra, v1 = Clamp || 32 * wu + 16 < ra;
    gemm_loop<T, Clamp>(pa, lda, ra, pb, ldb, rb, kbeg, kend, sm, [&](const T* As, const T* Bs, int k) {
        if (!v0) return;
        const T a0 = As[(k + lk) * PA + 32 * w + li];
        T b[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) b[p] = Bs[(k + lk) * PB + 16 * p + li];
        if (v1) {
            const T a1 = As[(k + lk) * PA + 32 * w + 16 + li];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                acc[p][0] = Mfma<T>::mma(b[p], a0, acc[p][0]);
                acc[p][1] = Mfma<T>::mma(b[p], a1, acc[p][1]);
            }
        } else {
#pragma unroll
            for (int p = 0; p < 4; ++p) acc[p][0] = Mfma<T>::mma(b[p], a0, acc[p][0]);
        }
    });
}

// FP64 TN product with direct-to-LDS staging, for the supernodal STEP and PART tiles: each
// wavefront moves its share of every K chunk with global_load_lds_dwordx4 (16 bytes per lane
// straight into LDS, no staging registers, no LDS store instructions), three LDS stages deep,
// so chunks c + 1 and c + 2 are in flight while chunk c is computed.  A chunk is 16 columns of
// A (128 rows each: one instruction per column, lane l moves rows 2l and 2l + 1) and of B (64
// rows: one instruction per two columns).  Each wavefront waits for its own loads of chunk c
// (vmcnt: the six loads of chunk c + 1 may still run), then a barrier makes chunk c complete
// for everyone and chunk c - 1's stage free for chunk c + 2.  The loads are issued in inline
// assembly (m0 = the LDS address, set in the same statement): the instruction has no output
// register, and the waits above are the only ones its LDS writes need; m0 is saved and restored
// around it.
// Rows: a lane whose pair starts at or past ra (rb) loads rows 0, 1 instead (product rows the
// callers discard, as in gemm_loop); a pair starting at ra - 1 also reads row ra, which lies in
// the panel: every column of A and B here is left of the block column k0 <= ncol - 1, so the
// element after its last row is row 0 of the next column.  K ranges must be whole chunks
// (multiples of 16; the supernodal plans' are).  Results are bit for bit those of gemm_tn.
struct GemmLdsDma {
    double a[3][16 * PA];
    double b[3][16 * NC];
};

// m0 is saved and restored around the load, so the statement leaves it as it found it.
__device__ __forceinline__ void dma16(const double* g, double* l) {
    typedef __attribute__((address_space(3))) void lds_void;
    const uint32_t lds = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(lds_void*)l);
    uint32_t saved;
    asm volatile("s_mov_b32 %0, m0\n\ts_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off\n\t"
                 "s_mov_b32 m0, %0"
                 : "=&s"(saved) : "v"(g), "s"(lds) : "memory");
}

__device__ __forceinline__ void gemm_tn_dma(const double* __restrict__ pa, int64_t lda, int ra,
                                            const double* __restrict__ pb, int64_t ldb, int rb,
                                            int kbeg, int kend, GemmLdsDma& sm, Acc<double> (&acc)[4][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const bool v0 = 32 * wu < ra, v1 = 32 * wu + 16 < ra;
    const int ar = 2 * lane < ra ? 2 * lane : 0;
    const int hb = lane & 31, bkk = lane >> 5;
    const int br = 2 * hb < rb ? 2 * hb : 0;
    const double* qa = pa + ar + (int64_t)(4 * w) * lda;
    const double* qb = pb + br + (int64_t)(4 * w + bkk) * ldb;
    auto issue = [&](int st, int kc) {
#pragma unroll
        for (int i = 0; i < 4; ++i) dma16(qa + (int64_t)(kc + i) * lda, sm.a[st] + (4 * w + i) * PA);
#pragma unroll
        for (int i = 0; i < 2; ++i) dma16(qb + (int64_t)(kc + 2 * i) * ldb, sm.b[st] + (4 * w + 2 * i) * NC);
    };
    const int nch = (kend - kbeg) / 16;
    if (nch <= 0) return;
    issue(0, kbeg);
    if (nch > 1) issue(1, kbeg + 16);
    int st = 0;
#pragma unroll 1
    for (int c = 0; c < nch; ++c) {
        if (c + 1 < nch) asm volatile("s_waitcnt vmcnt(6)" ::: "memory");
        else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        asm volatile("s_barrier" ::: "memory");
        if (c + 2 < nch) issue(st == 0 ? 2 : st - 1, kbeg + (c + 2) * 16);
        const double* As = sm.a[st];
        const double* Bs = sm.b[st];
        if (v0) {
#pragma unroll
            for (int k = 0; k < 16; k += 4) {
                const double a0 = As[(k + lk) * PA + 32 * w + li];
                double b[4];
#pragma unroll
                for (int p = 0; p < 4; ++p) b[p] = Bs[(k + lk) * NC + 16 * p + li];
                if (v1) {
                    const double a1 = As[(k + lk) * PA + 32 * w + 16 + li];
#pragma unroll
                    for (int p = 0; p < 4; ++p) {
                        acc[p][0] = Mfma<double>::mma(b[p], a0, acc[p][0]);
                        acc[p][1] = Mfma<double>::mma(b[p], a1, acc[p][1]);
                    }
                } else {
#pragma unroll
                    for (int p = 0; p < 4; ++p) acc[p][0] = Mfma<double>::mma(b[p], a0, acc[p][0]);
                }
            }
        }
        st = st == 2 ? 0 : st + 1;
    }
    __syncthreads();   // every wavefront done with the stages: the caller may reuse the LDS
}

// gemm_tn_dma for two block columns at once (128 B rows: a 128 x 128 product, accumulators
// acc[p][b] for B columns 16 p .., p < 8): the A tile is read once for both.  Two LDS stages of
// 16 columns of A and B (pitch PA both; one instruction per column and wavefront share); chunk
// c + 1 loads while chunk c is computed.  Same row and K conventions as gemm_tn_dma, except
// that every wavefront computes all its rows (skipping the 16-row blocks past ra, a rare partial
// tile here, makes the compiler keep two MFMA sequences and spill).
struct GemmLdsDma2 {
    double a[2][16 * PA];
    double b[2][16 * PA];
};

__device__ __forceinline__ void gemm_tn_dma2(const double* __restrict__ pa, int64_t lda, int ra,
                                             const double* __restrict__ pb, int64_t ldb, int rb,
                                             int kbeg, int kend, GemmLdsDma2& sm, Acc<double> (&acc)[8][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
    const int ar = 2 * lane < ra ? 2 * lane : 0;
    const int br = 2 * lane < rb ? 2 * lane : 0;
    const double* qa = pa + ar + (int64_t)(4 * w) * lda;
    const double* qb = pb + br + (int64_t)(4 * w) * ldb;
    auto issue = [&](int st, int kc) {
#pragma unroll
        for (int i = 0; i < 4; ++i) dma16(qa + (int64_t)(kc + i) * lda, sm.a[st] + (4 * w + i) * PA);
#pragma unroll
        for (int i = 0; i < 4; ++i) dma16(qb + (int64_t)(kc + i) * ldb, sm.b[st] + (4 * w + i) * PA);
    };
    const int nch = (kend - kbeg) / 16;
    if (nch <= 0) return;
    issue(0, kbeg);
#pragma unroll 1
    for (int c = 0; c < nch; ++c) {
        const int st = c & 1;
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        asm volatile("s_barrier" ::: "memory");
        if (c + 1 < nch) issue(st ^ 1, kbeg + (c + 1) * 16);
        const double* As = sm.a[st];
        const double* Bs = sm.b[st];
#pragma unroll
        for (int k = 0; k < 16; k += 4) {
            const double a0 = As[(k + lk) * PA + 32 * w + li];
            const double a1 = As[(k + lk) * PA + 32 * w + 16 + li];
#pragma unroll
            for (int p = 0; p < 8; ++p) {
                const double bp = Bs[(k + lk) * PA + 16 * p + li];
                acc[p][0] = Mfma<double>::mma(bp, a0, acc[p][0]);
                acc[p][1] = Mfma<double>::mma(bp, a1, acc[p][1]);
            }
        }
    }
    __syncthreads();
}

// LDS of the supernodal STEP / PART kernels and their TN product: direct-to-LDS staging in
// FP64 (whole 16-column chunks), gemm_tn otherwise.
template <typename T> struct StepLdsT { using type = GemmLds<T>; };
template <> struct StepLdsT<double> { using type = GemmLdsDma; };
template <typename T> using StepLds = typename StepLdsT<T>::type;
static_assert(sizeof(GemmLdsDma) >= sizeof(GemmLds<double>), "fallback layout");
static_assert(sizeof(GemmLdsDma) >= sizeof(GemmLdsDma2), "two-column product layout");

template <typename T>
__device__ __forceinline__ void gemm_tn_step(const T* __restrict__ pa, int64_t lda, int ra,
                                             const T* __restrict__ pb, int64_t ldb, int rb, int kbeg,
                                             int kend, StepLds<T>& sm, Acc<T> (&acc)[4][2]) {
    if constexpr (sizeof(T) == 8) {
        if ((kend - kbeg) % 16 == 0) {
            gemm_tn_dma(pa, lda, ra, pb, ldb, rb, kbeg, kend, sm, acc);
            return;
        }
        gemm_tn<T>(pa, lda, ra, pb, ldb, rb, kbeg, kend, reinterpret_cast<GemmLds<T>&>(sm), acc);
    } else {
        gemm_tn<T>(pa, lda, ra, pb, ldb, rb, kbeg, kend, sm, acc);
    }
}

// f(row, col, value) for every NT accumulator element of this thread (tile coordinates).
template <typename T, typename F>
__device__ __forceinline__ void for_acc(Acc<T> (&acc)[2][4], F&& f) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                f(32 * w + 16 * x + Mfma<T>::row(lane, e), 16 * y + (lane & 15), acc[x][y][e]);
}

// Wave-uniform broadcast of lane `src`'s value (src is a compile-time constant at every use).
__device__ __forceinline__ double bcast_lane(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float bcast_lane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// 1/sqrt(d) for d > 0: hardware estimate refined by two Newton steps (each squares the relative
// error), avoiding the long correctly-rounded sqrt + divide sequences on the critical path.
__device__ __forceinline__ double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    y = y * fma(-0.5 * d * y, y, 1.5);
    y = y * fma(-0.5 * d * y, y, 1.5);
    return y;
}
__device__ __forceinline__ float rsqrt_nr(float d) {
    float y = __builtin_amdgcn_rsqf(d);
    y = y * fmaf(-0.5f * d * y, y, 1.5f);
    return y;
}

// One wavefront factors the 16x16 block at Ds(o, o) and inverts its factor.  Lane r (mod 16)
// holds row r in registers.  Columns are factored four at a time: every lane forms the
// Cholesky factor of the 4x4 diagonal block from broadcast (readlane) values and solves its
// own row's four entries against it; the four finished columns go to LDS once (cb, column-
// major, pitch 16) and are read back with wave-uniform (broadcast) addresses for the update of
// the remaining columns -- four LDS round trips instead of sixteen.  Lane i then forms row i
// of L^{-1} by substitution from the end, y L = e_i, reading L's columns from cb.  Every LDS
// store puts consecutive lanes on consecutive addresses (no bank conflicts).  Writes L (lower)
// back to Ds and L^{-1} to Lv (column-major, pitch 16, upper part zero); cb needs 272 values.
// Returns false (wave-uniform) when a pivot is not positive.
template <typename T>
__device__ __forceinline__ bool potrf16_wave(T* Ds, int o, T* Lv, T* cb) {
    const int l = threadIdx.x & 63, r = l & 15;
    T* isd = cb + 256;                            // 1 / L[j][j]
    T a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = Ds[(o + k) * DP + o + r];
    bool ok = true;
#pragma unroll
    for (int jb = 0; jb < 16; jb += 4) {
        T m[4][4], L4[4][4], is4[4];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q <= p; ++q) m[p][q] = bcast_lane(a[jb + q], jb + p);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T d = m[q][q];
#pragma unroll
            for (int s = 0; s < q; ++s) d = fma(-L4[q][s], L4[q][s], d);
            ok = ok && (d > T(0));
            is4[q] = rsqrt_nr(d);
#pragma unroll
            for (int p = q + 1; p < 4; ++p) {
                T v = m[p][q];
#pragma unroll
                for (int s = 0; s < q; ++s) v = fma(-L4[p][s], L4[q][s], v);
                L4[p][q] = v * is4[q];
            }
        }
        // row r in columns jb .. jb+3 (rows of the block reproduce L4, the diagonal included;
        // entries right of a lane's diagonal are never used)
        T lq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T v = a[jb + q];
#pragma unroll
            for (int s = 0; s < q; ++s) v = fma(-lq[s], L4[q][s], v);
            lq[q] = v * is4[q];
        }
        if (l < 16) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                cb[16 * (jb + q) + r] = lq[q];
                if (jb + q <= r) Ds[(o + jb + q) * DP + o + r] = lq[q];
            }
        }
        if (l < 4) isd[jb + l] = l == 0 ? is4[0] : l == 1 ? is4[1] : l == 2 ? is4[2] : is4[3];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = jb + 4; k < 16; ++k) {
            T v = a[k];
#pragma unroll
            for (int q = 0; q < 4; ++q) v = fma(-lq[q], cb[16 * (jb + q) + k], v);
            a[k] = v;
        }
    }
    __builtin_amdgcn_wave_barrier();
    // row r of L^{-1}: y[k] = (delta_rk - sum_{j>k} y[j] L[j][k]) / L[k][k], k = 15 .. 0
    // (zero for k > r)
    T y[16];
#pragma unroll
    for (int k = 15; k >= 0; --k) {
        T s0 = k == r ? T(1) : T(0), s1 = T(0);
#pragma unroll
        for (int j = k + 1; j < 16; ++j) {
            const T ljk = cb[16 * k + j];
            if (j & 1) s1 = fma(-y[j], ljk, s1);
            else s0 = fma(-y[j], ljk, s0);
        }
        y[k] = (s0 + s1) * isd[k];
    }
    if (l < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) Lv[c * 16 + r] = c <= r ? y[c] : T(0);
    }
    __builtin_amdgcn_wave_barrier();              // cb reads done before the caller reuses it
    return ok;
}

// In-place Cholesky of the 64x64 block Ds[c*DP + r] (lower part; rows/columns past the valid
// block must hold the identity), blocked by 16: wave 0 factors and inverts the diagonal block,
// the 16-row panel below is solved on MFMA (L_ip^T = Linv_pp A_ip^T) and the trailing blocks
// are updated on MFMA (A_ij -= L_ip L_jp^T).  Lv receives the four diagonal inverses
// (Lv + 256 p, see potrf16_wave).  Three barriers per 16 columns.  Returns false, uniformly
// across the workgroup, when a pivot is not positive.
template <typename T>
__device__ __forceinline__ bool potrf64(T* Ds, T* Lv, int* s_ok) {
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15;
    using R = Mfma<T>;
    __shared__ T cb[256 + 16];
    // not unrolled: one copy of the 16-column code keeps the kernels that inline this small
    // enough for the instruction cache
#pragma unroll 1
    for (int p = 0; p < 4; ++p) {
        const int o = 16 * p;
        if (w == 0) {
            const bool ok = potrf16_wave<T>(Ds, o, Lv + 256 * p, cb);
            if (l == 0) *s_ok = ok ? 1 : 0;
        }
        __syncthreads();
        if (!*s_ok) return false;
        if (w < 3 - p) {                          // panel block i: L_ip^T = Linv_pp * A_ip^T
            const int i = p + 1 + w;
            Acc<T> d = Acc<T>{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = R::row(l, s);
                d = R::mma(Lv[256 * p + k * 16 + li], Ds[(o + k) * DP + 16 * i + li], d);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) Ds[(o + R::row(l, e)) * DP + 16 * i + li] = d[e];
        }
        __syncthreads();
        // trailing blocks (i, jj), p < jj <= i <= 3, round-robin over the waves
        int b = 0;
#pragma unroll 1
        for (int i = p + 1; i < 4; ++i)
#pragma unroll 1
            for (int jj = p + 1; jj <= i; ++jj, ++b) {
                if ((b & 3) != w) continue;
                Acc<T> d = Acc<T>{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int k = R::row(l, s);
                    d = R::mma(Ds[(o + k) * DP + 16 * i + li], Ds[(o + k) * DP + 16 * jj + li], d);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) Ds[(16 * jj + li) * DP + 16 * i + R::row(l, e)] -= d[e];
            }
        __syncthreads();
    }
    return true;
}


}  // namespace

// Supernodal Cholesky kernels for gfx950 (replace cholesky_level_kernel and its per-level
// launch loop, src/cuda/kkt.cu:217-285,582-596, and the repeated refactorization in
// gpu_cholesky_solve_device, src/cuda/kkt.cu:671-689).
//
// Storage: every supernode owns a column-major panel (ld = rows of its structure, ncol columns)
// at panel + off.  The host plan drives a left-looking blocked factorization per etree level
// with 64-column block columns k (columns k0 .. k0+kb), one launch per step:
//   DIAG     (block 0 only) one workgroup per supernode factors its first 64x64 diagonal block
//            and stages it (Lkk, the inverses of its 16x16 diagonal blocks), writes the row-major
//            copy with reciprocal diagonal (for the substitution kernels) and the pivots.
//   STEP k   every 128-row tile R below the diagonal block:
//              X = C(R, k) - L(R, 0:k0) L(k, 0:k0)^T   on MFMA (16x16x4, FP64 / FP32), computed
//                  transposed so that the accumulators are already in the layout of the solve
//              X <- X Lkk^{-T}                         on MFMA (transposed blocked solve)
//            and the tile holding the next diagonal block also forms, factors and stages it.
//            Nothing waits for another workgroup: the launch boundary is the only barrier.
//   PART     (steps with few tiles) the K range of each tile is split: a launch before the STEP
//            computes the partial products, the STEP adds them in a fixed order.
//   UPDATE   below-diagonal rows:  U = L21 * L21^T (split K) scattered (atomic add) into the
//            ancestor panels.
//   DIAGWB   after the last step, Lkk of every block column is copied into the panel (the
//            steps never write the panel's diagonal blocks: they live in the staging slots).
// All workgroups are 256 threads (4 wavefronts of 64).  Task i of a launch runs on workgroup
// (i % per) * 8 + i / per (per = grid / 8): consecutive tasks -- the tiles of one supernode --
// share an XCD and its L2.
#include "dense_tile.hpp"

namespace {
// Staging slot of a factored diagonal block: L (64x64 column-major, pitch 64, identity past kb)
// followed by the inverses of its four 16x16 diagonal blocks (column-major, pitch 16).
constexpr int STAGE = TS * TS + 4 * 256;
static_assert(STAGE == LKD_STAGE, "staging slot size");
static_assert(LKD_PART == 32 * 256, "partial product: 32 accumulator elements per thread");

// the GEMM buffers are reused afterwards: 64 x DP staging tile (UPDATE), staged factor (STEP),
// diagonal block + inverses (STEP, next diagonal)
static_assert(sizeof(GemmLds<double>) >= STAGE * sizeof(double), "staged factor");
static_assert(sizeof(GemmLds<double>) >= (64 * DP + 1024) * sizeof(double), "diagonal block");
static_assert(sizeof(GemmLds<float>) >= (64 * DP + 1024) * sizeof(float), "diagonal block");

__device__ __forceinline__ int task_index() {
    const int per = gridDim.x >> 3;   // grid is a multiple of 8
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// Error flag read once per workgroup: every thread takes the same early exit.
__device__ __forceinline__ bool wg_error(const int32_t* err, int* s) {
    if (threadIdx.x == 0) *s = *(volatile const int32_t*)err;
    __syncthreads();
    return *s != 0;
}

// Row-major copy of a factored diagonal block for the substitution kernels:
// out[j*kb + k] = L[j][k] (k < j), out[j*kb + j] = 1 / L[j][j].
template <typename T>
__device__ __forceinline__ void write_diag_copy(const T* Ds, int kb, T* out) {
    for (int idx = threadIdx.x; idx < kb * kb; idx += 256) {
        const int j = idx / kb, k = idx - j * kb;
        T v = T(0);
        if (k < j) v = Ds[k * DP + j];
        else if (k == j) v = T(1) / Ds[j * DP + j];
        out[idx] = v;
    }
}

// Publish a factored diagonal block (Ds, Lv as left by potrf64): staging slot S, pivots of the
// permuted columns first .. first+kb, row-major copy with reciprocal diagonal at dcopy.
template <typename T>
__device__ __forceinline__ void publish_diag(const T* Ds, const T* Lv, int kb, T* __restrict__ S,
                                             T* __restrict__ piv, T* __restrict__ dcopy) {
    const int t = threadIdx.x;
    for (int idx = t; idx < TS * TS; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        S[idx] = r >= c ? Ds[c * DP + r] : T(0);
    }
    for (int idx = t; idx < 4 * 256; idx += 256) S[TS * TS + idx] = Lv[idx];
    if (t < kb) piv[t] = Ds[t * DP + t];
    write_diag_copy<T>(Ds, kb, dcopy);
}

// Factor the diagonal block of columns k0 .. k0+kb of the task's supernode as it stands in the
// panel (block 0: nothing to subtract; a later block: every earlier block column's update was
// applied by the STEP tiles); staging slot at stage + ws, row-major copy at diag + aux.
template <typename T>
__device__ __forceinline__ void diag_task(const lkd_task& tk, const T* __restrict__ L,
                                          T* __restrict__ diag, T* __restrict__ piv,
                                          T* __restrict__ stage, int32_t* err, T* Ds, T* Lv,
                                          int* s_ok) {
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int k0 = tk.k0, kb = tk.kb;
    for (int idx = threadIdx.x; idx < TS * TS; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        T v = r == c ? T(1) : T(0);
        if (r < kb && c < kb) v = r >= c ? P[(int64_t)(k0 + c) * ld + k0 + r] : T(0);
        Ds[c * DP + r] = v;
    }
    __syncthreads();
    if (!potrf64<T>(Ds, Lv, s_ok)) {
        if (threadIdx.x == 0) atomicOr(err, 1);
        return;
    }
    publish_diag<T>(Ds, Lv, kb, stage + tk.ws, piv + tk.first + k0, diag + tk.aux);
}

// DIAG: the first diagonal block of each task's supernode (diag_task).
template <typename T>
__global__ __launch_bounds__(256) void k_diag(const lkd_task* __restrict__ tasks, int ntask,
                                              const T* __restrict__ L, T* __restrict__ diag,
                                              T* __restrict__ piv, T* __restrict__ stage,
                                              int32_t* err) {
    const int ti = task_index();
    if (ti >= ntask) return;
    __shared__ int s_err;
    if (wg_error(err, &s_err)) return;
    __shared__ T Ds[TS * DP];
    __shared__ T Lv[4 * 256];
    __shared__ int s_ok;
    diag_task<T>(tasks[ti], L, diag, piv, stage, err, Ds, Lv, &s_ok);
}

// STEP k (one task = tile R = rows r0 .. r0+c0-1, all below the diagonal block of block column k):
//   X = C(R, k) - L(R, 0:k0) L(k, 0:k0)^T          (gemm_tn: accumulators in solve layout)
//   X <- X Lkk^{-T}: Y = X^T solves Lkk Y = X^T, by 16-row blocks of Y
//       Y_p = Linv_pp (X_p^T - sum_{q<p} L_pq Y_q)  (A operands from the staged factor in LDS;
//       an accumulator is exactly the B operand of the next product, so Y stays in registers)
//   every diagonal block j whose rows lie in R: C(j, j) -= L(j, k) L(j, k)^T (the diagonal
//       blocks are kept up to date step by step, so none needs a long update when its turn
//       comes)
// Task fields: ws = staging slot of block k (written by the previous launch).  nsplit = 1 marks
// the tile that is exactly the next diagonal block k1 = k0 + kb (c0 = kbeg rows): after its own
// rank-kb update it factors C(k1, k1) and publishes it (slot aux2, copy aux).
// PART: one K range [kbeg, kend) of a STEP tile's GEMM, L(R, kbeg:kend) L(k, kbeg:kend)^T,
// stored as this thread's accumulator elements: work[px + i * 256 + thread], i < 32.  The STEP
// launch that follows reads them back in the same layout (coalesced) and adds them in order.
template <typename T>
__device__ __forceinline__ void part_tile(const lkd_task& tk, const T* __restrict__ L,
                                          T* __restrict__ work, StepLds<T>& sm) {
    const T* P = L + tk.off;
    Acc<T> acc[4][2];
    zero_acc<T, 4, 2>(acc);
    gemm_tn_step<T>(P + tk.r0, tk.ld, tk.c0, P + tk.k0, tk.ld, tk.kb, tk.kbeg, tk.kend, sm, acc);
    T* out = work + tk.px + threadIdx.x;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) out[((p * 2 + b) * 4 + e) * 256] = acc[p][b][e];
}

// PART launch of the one-launch-per-step form: PART tiles (part = 2: two block columns, see
// part_tile2), and (nsplit = 2, listed first) the factorization of the step's diagonal block,
// which runs beside the partial products (the STEP launch after this one is the first to read
// it).
// PART of two consecutive block columns (task part = 2: the one-launch-per-step form pairs an
// even step k with step k + 1 on the same 128-row tile): L(R, kbeg:kend) [L(k, kbeg:kend);
// L(k + 1, kbeg:kend)]^T, block column k's half to work + px and block column k + 1's to
// work + ws, each in part_tile's layout: one 128 x 128 product, the A tile read once (FP64;
// k_part runs FP32 paired tasks as two PART tiles).
template <typename T>
__device__ __forceinline__ void part_tile2(const lkd_task& tk, const T* __restrict__ L,
                                           T* __restrict__ work, StepLds<T>& sm) {
    const int rb = min(2 * TS, tk.ncol - tk.k0);
    if constexpr (sizeof(T) == 8) {
        const T* P = L + tk.off;
        T* out1 = work + tk.px + threadIdx.x;
        T* out2 = work + tk.ws + threadIdx.x;
        Acc<T> acc[8][2];
        zero_acc<T, 8, 2>(acc);
        gemm_tn_dma2(P + tk.r0, tk.ld, tk.c0, P + tk.k0, tk.ld, rb, tk.kbeg, tk.kend,
                     reinterpret_cast<GemmLdsDma2&>(sm), acc);
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    out1[((p * 2 + b) * 4 + e) * 256] = acc[p][b][e];
                    out2[((p * 2 + b) * 4 + e) * 256] = acc[p + 4][b][e];
                }
    }
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_part(const lkd_task* __restrict__ tasks, int ntask, const T* __restrict__ L,
            T* __restrict__ diag, T* __restrict__ piv, T* __restrict__ stage,
            T* __restrict__ work, int32_t* err, int halves) {
    // FP32 launches run every task as two halves (workgroup 2 i + h): a paired task's block
    // columns one per workgroup, the second half of any other task has nothing to do
    const int vi = task_index();
    if (vi >= ntask * halves) return;
    const int ti = halves == 2 ? vi >> 1 : vi, h = halves == 2 ? (vi & 1) : 0;
    __shared__ int s_err;
    if (wg_error(err, &s_err)) return;
    __shared__ StepLds<T> sm;
    __shared__ int s_ok;
    const lkd_task tk = tasks[ti];
    if (tk.nsplit == 2) {
        if (h) return;
        T* Ds = reinterpret_cast<T*>(&sm);
        diag_task<T>(tk, L, diag, piv, stage, err, Ds, Ds + TS * DP, &s_ok);
        return;
    }
    if constexpr (sizeof(T) == 8) {
        if (tk.part == 2) part_tile2<T>(tk, L, work, sm);
        else part_tile<T>(tk, L, work, sm);
    } else {
        if (tk.part != 2 && h) return;
        lkd_task u = tk;
        if (tk.part == 2) {
            u.k0 = tk.k0 + TS * h;
            u.kb = min(TS, tk.ncol - u.k0);
            u.px = h ? tk.ws : tk.px;
        }
        part_tile<T>(u, L, work, sm);
    }
}

// PSUM: the np partial products of a STEP tile (work + px, LKD_PART elements each) summed in
// order into the first one; task kbeg = slice of 512 elements (two per thread), so a tile with
// many K ranges is summed by LKD_PART / 512 workgroups, each with many loads in flight.
template <typename T>
__global__ __launch_bounds__(256) void k_psum(const lkd_task* __restrict__ tasks, int ntask,
                                              T* __restrict__ work) {
    const int ti = task_index();
    if (ti >= ntask) return;
    const lkd_task tk = tasks[ti];
    typedef T T2 __attribute__((ext_vector_type(2)));
    T2* p = reinterpret_cast<T2*>(work + tk.px + (int64_t)tk.kbeg * 512) + threadIdx.x;
    constexpr int64_t stride = LKD_PART / 2;
    T2 acc = p[0];
#pragma unroll 16
    for (int k = 1; k < tk.np; ++k) acc += p[(int64_t)k * stride];
    p[0] = acc;
}

// Wait (one thread) until flag == ep; false on an error flag or after kSpinLimit polls (then
// bit 0x20 of err is set).
constexpr long long kSpinLimit = 1ll << 22;
__device__ __forceinline__ bool flag_wait(const uint32_t* flag, uint32_t ep, int32_t* err) {
    long long n = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != ep) {
        if ((n & 255) == 255 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        if (++n >= kSpinLimit) {
            atomicOr(err, 0x20);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

// Hand-offs of a tile run inside a STEPS launch (absent for the one-launch-per-step form):
// the tile waits for flags2[sdep] (its diagonal block staged) only after its GEMM, and a tile
// that factors the next diagonal block sets flags[ti] as soon as its own rows are stored (the
// next step's GEMMs read them) and flags2[ti] once the factor is staged.
struct DagSync {
    uint32_t* flags;
    uint32_t* flags2;
    int sdep, ti;
    uint32_t ep;
};

template <typename T>
__device__ __forceinline__ void step_tile(const lkd_task& tk, T* __restrict__ L,
                                          T* __restrict__ diag, T* __restrict__ piv,
                                          T* __restrict__ stage, const T* __restrict__ work,
                                          int32_t* err, StepLds<T>& sm, int* s_ok,
                                          const DagSync* ds = nullptr) {
    using R = Mfma<T>;
    T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int k0 = tk.k0, kb = tk.kb, kbm = kb - 1;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15;
    const int ra = tk.c0;
    // the tile that factors the next diagonal block is the launch's critical path: its waves
    // get issue priority over the other workgroup sharing the compute unit
    if (tk.nsplit || tk.prio) __builtin_amdgcn_s_setprio(3);
    // The accumulators start at (sum of the PART products) - C(R, k): the PART tasks are
    // dependencies (their products are final), and the tile's block of the assembled panel is
    // modified by no task of the factorization, so both loads are issued ahead of the GEMM and
    // their latency overlaps its first chunks.  After the GEMM, X = C - L L^T - parts = -acc.
    Acc<T> acc[4][2];
    int64_t row[2];
    bool valid[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int rl = 32 * w + 16 * b + li;
        row[b] = (int64_t)tk.r0 + rl;
        valid[b] = rl < ra;
        const int64_t rc = valid[b] ? row[b] : tk.r0;     // a row inside the panel
        if (kb == TS) {
            // full block: one per-lane base, the element offsets are wave-uniform multiples of ld
            const T* pb = P + (int64_t)(k0 + R::row(l, 0)) * ld + rc;
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    acc[p][b][e] = -pb[(int64_t)(16 * p + R::row(l, e) - R::row(l, 0)) * ld];
        } else {
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int c = min(16 * p + R::row(l, e), kbm);   // columns past kb: padding
                    acc[p][b][e] = -P[(int64_t)(k0 + c) * ld + rc];
                }
        }
    }
    for (int s = 0; s < tk.np; ++s) {                 // K ranges computed by PART tasks
        const T* in = work + tk.px + (int64_t)s * LKD_PART + t;
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[p][b][e] += in[((p * 2 + b) * 4 + e) * 256];
    }
    gemm_tn_step<T>(P + tk.r0, ld, ra, P + k0, ld, kb, tk.kbeg, tk.kend, sm, acc);
    Acc<T> y[2][4];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int p = 0; p < 4; ++p) y[b][p] = -acc[p][b];
    // the staged factor of block k (in a STEPS launch: once its tile has published it)
    if (ds && ds->sdep >= 0) {
        if (t == 0) *s_ok = flag_wait(ds->flags2 + ds->sdep, ds->ep, err) ? 1 : 0;
        __syncthreads();
        if (!*s_ok) return;
        asm volatile("buffer_inv sc0" ::: "memory");
    }
    T* S = reinterpret_cast<T*>(&sm);
    {
        const T* Sg = stage + tk.ws;
#pragma unroll 4
        for (int idx = t; idx < STAGE; idx += 256) S[idx] = Sg[idx];
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        Acc<T> tp[2] = {y[0][p], y[1][p]};
#pragma unroll
        for (int q = 0; q < p; ++q)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const T a = -S[(16 * q + R::row(l, s)) * TS + 16 * p + li];   // -L_pq[li][k]
#pragma unroll
                for (int b = 0; b < 2; ++b) tp[b] = R::mma(a, y[b][q][s], tp[b]);
            }
#pragma unroll
        for (int b = 0; b < 2; ++b) y[b][p] = Acc<T>{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const T a = S[TS * TS + 256 * p + R::row(l, s) * 16 + li];       // Linv_pp[li][k]
#pragma unroll
            for (int b = 0; b < 2; ++b) y[b][p] = R::mma(a, tp[b][s], y[b][p]);
        }
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        if (!valid[b]) continue;
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = 16 * p + R::row(l, e);
                if (c < kb) P[(int64_t)(k0 + c) * ld + row[b]] = y[b][p][e];
            }
    }
    if (ds && tk.nsplit) {                            // rows stored: the next step may read them
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (t == 0) __hip_atomic_store(ds->flags + ds->ti, ds->ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // ---- rank-kb updates of the diagonal blocks inside the tile (rows below ncol only): waves
    //      2h, 2h+1 hold the rows of block h in y; they go to LDS (Ys[c * 65 + r] = Y[r][c],
    //      zero past kb columns and db rows), and C(d, d) -= Y Y^T is formed on MFMA, wave w
    //      computing the 16-row stripe w, lower blocks only
    for (int h = 0; h < 2; ++h) {
        const int d0 = tk.r0 + 64 * h;                    // first row = first column of block
        if (64 * h >= ra || d0 >= tk.ncol) break;         // uniform
        const int db = min(64, min(ra - 64 * h, tk.ncol - d0));
        T* Ys = S;
        // this wave's stripe of C(d, d) (clamped addresses), in flight during the product
        T cold[4][4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = min(16 * w + R::row(l, e), db - 1), c = min(16 * q + li, db - 1);
                cold[q][e] = q <= w ? P[(int64_t)(d0 + c) * ld + d0 + r] : T(0);
            }
        __syncthreads();                                  // the staged factor / last stripe is read
        if ((w >> 1) == h) {
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int p = 0; p < 4; ++p)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int rl = 32 * (w & 1) + 16 * b + li, c = 16 * p + R::row(l, e);
                        Ys[c * DP + rl] = (c < kb && rl < db) ? y[b][p][e] : T(0);
                    }
        }
        __syncthreads();
        Acc<T> dq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) dq[q] = Acc<T>{0, 0, 0, 0};
#pragma unroll 1
        for (int kk = 0; kk < TS; kk += 16)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = kk + R::row(l, s);
                const T a = Ys[k * DP + 16 * w + li];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q <= w) dq[q] = R::mma(a, Ys[k * DP + 16 * q + li], dq[q]);
            }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (q > w) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = 16 * w + R::row(l, e), c = 16 * q + li;
                if (r < db && c <= r) P[(int64_t)(d0 + c) * ld + d0 + r] = cold[q][e] - dq[q][e];
            }
        }
    }
    if (!tk.nsplit) return;
    // ---- the next diagonal block (this tile) is complete: factor and publish it
    const int k1 = tk.r0, kb1 = ra;
    __syncthreads();
    T* Ds = S;
    T* Lv = S + TS * DP;
    for (int idx = t; idx < TS * TS; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        T v = r == c ? T(1) : T(0);
        if (r < kb1 && c < kb1) v = r >= c ? P[(int64_t)(k1 + c) * ld + k1 + r] : T(0);
        Ds[c * DP + r] = v;
    }
    __syncthreads();
    if (!potrf64<T>(Ds, Lv, s_ok)) {
        if (t == 0) atomicOr(err, 1);
        return;
    }
    publish_diag<T>(Ds, Lv, kb1, stage + tk.aux2, piv + tk.first + k1, diag + tk.aux);
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_step(const lkd_task* __restrict__ tasks, int ntask, T* __restrict__ L, T* __restrict__ diag,
            T* __restrict__ piv, T* __restrict__ stage, const T* __restrict__ work, int32_t* err) {
    const int ti = task_index();
    if (ti >= ntask) return;
    __shared__ int s_err;
    if (wg_error(err, &s_err)) return;
    __shared__ StepLds<T> sm;
    __shared__ int s_ok;
    step_tile<T>(tasks[ti], L, diag, piv, stage, work, err, sm, &s_ok);
}

// All STEP tiles of one etree level in one launch.  Every supernode's tiles are assigned to one
// XCD (its L2): queue x lists them in an order where every task's dependencies come first
// (queue[qoff[x] .. qoff[x+1]), task indices).  A workgroup reads the XCD it runs on
// (XCC_ID mod nx), takes the next ticket of that queue from counter cnt[x] (none left: it
// exits), waits until the tasks deps[4 i .. 4 i + 3] (-1 = none) carry this factorization's
// epoch ep_ctl[0] in flags, runs task i as a STEP tile and sets flags[i].  Tickets go to
// running workgroups in order and a task waits only for tasks earlier in its queue, so the
// earliest unfinished task of a queue can always run.  Workgroups are dealt to the XCDs in
// turn, so a grid of nx times the longest queue gives every queue a workgroup per ticket; if
// one were left without, its waiters end at the poll limit with error bit 0x20.  Producer
// and consumer share an L2, so a release is the wait for the wavefront's stores (they reach
// the L2) and an acquire is an L1 invalidate.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_step_dag(const lkd_task* __restrict__ tasks, const int32_t* __restrict__ deps,
                const int32_t* __restrict__ sdeps, const int32_t* __restrict__ queue,
                const int32_t* __restrict__ qoff, int nx, uint32_t* __restrict__ flags,
                uint32_t* __restrict__ flags2, const uint32_t* __restrict__ ep_ctl,
                uint32_t* __restrict__ cnt, T* __restrict__ L, T* __restrict__ diag,
                T* __restrict__ piv, T* __restrict__ stage, T* __restrict__ work, int32_t* err) {
    __shared__ StepLds<T> sm;
    __shared__ int s_ok, s_task;
    const uint32_t ep = ep_ctl[0];
    if (threadIdx.x == 0) {
        unsigned xcc;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
        const int x = (int)(xcc % (unsigned)nx);
        const int qb = qoff[x], qn = qoff[x + 1] - qb;
        const int t = (int)atomicAdd(cnt + x, 1u);
        int ti = t < qn ? queue[qb + t] : -1;
        if (ti >= 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ti = -1;
        for (int d = 0; d < 4 && ti >= 0; ++d) {
            const int j = deps[4 * ti + d];
            if (j >= 0 && !flag_wait(flags + j, ep, err)) ti = -1;
        }
        s_task = ti;
    }
    __syncthreads();
    const int ti = s_task;
    if (ti < 0) return;
    asm volatile("buffer_inv sc0" ::: "memory");        // this CU's L1
    const lkd_task tk = tasks[ti];
    if (tk.part) {
        part_tile<T>(tk, L, work, sm);
    } else {
        const DagSync ds{flags, flags2, sdeps[ti], ti, ep};
        step_tile<T>(tk, L, diag, piv, stage, work, err, sm, &s_ok, &ds);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");    // this wavefront's stores are in the L2
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store((tk.nsplit ? flags2 : flags) + ti, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ctl[0] += 1 (the new epoch), ctl[1 .. 1 + nq) = 0 (queue counters)
__global__ void k_dag_begin(uint32_t* ctl, int nq) {
    for (int i = threadIdx.x; i <= nq; i += blockDim.x) ctl[i] = i == 0 ? ctl[0] + 1 : 0u;
}

// Lkk of each task's block column into the panel: strictly lower part from the row-major copy,
// the diagonal from the pivots (exact, not the stored reciprocal).
template <typename T>
__global__ __launch_bounds__(256) void k_diagwb(const lkd_task* __restrict__ tasks, int ntask,
                                                T* __restrict__ L, const T* __restrict__ diag,
                                                const T* __restrict__ piv) {
    const int ti = task_index();
    if (ti >= ntask) return;
    const lkd_task tk = tasks[ti];
    T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int k0 = tk.k0, kb = tk.kb;
    const T* Lr = diag + tk.aux;
    for (int idx = threadIdx.x; idx < kb * kb; idx += 256) {
        const int c = idx / kb, r = idx - c * kb;
        if (r > c) P[(int64_t)(k0 + c) * ld + k0 + r] = Lr[r * kb + c];
        else if (r == c) P[(int64_t)(k0 + c) * ld + k0 + r] = piv[tk.first + k0 + r];
    }
}

__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* a, int64_t lo, int64_t hi,
                                                   int32_t key) {
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Contribution of the below-diagonal rows of a supernode to its ancestors, K range
// [kbeg, kend): U = L21[r0 rows] * L21[c0 rows]^T subtracted (atomic add) from the ancestor
// panels.  The tile is staged through LDS 64 rows at a time for the scatter.  A target
// supernode's row list starts with its own columns in order, so a row inside them sits at
// position row - first; only rows below it are searched in the rest of the list.
template <typename T>
__global__ __launch_bounds__(256) void k_update(const lkd_task* __restrict__ tasks, int ntask,
                                                T* __restrict__ L, lkd_snplan plan,
                                                int32_t* err) {
    const int ti = task_index();
    if (ti >= ntask) return;
    __shared__ int s_err;
    if (wg_error(err, &s_err)) return;
    const lkd_task tk = tasks[ti];
    __shared__ GemmLds<T> sm;
    __shared__ int64_t tbase[NC];
    __shared__ int64_t trow0[NC];
    __shared__ int32_t tlo[NC];
    __shared__ int32_t tld[NC];
    __shared__ int32_t tfirst[NC];
    __shared__ int32_t tnc[NC];
    T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int ra = min(MT, tk.ld - tk.r0), rb = min(NC, tk.ld - tk.c0);
    Acc<T> acc[2][4];
    zero_acc<T, 2, 4>(acc);
    gemm_nt<T>(P + tk.r0, ld, ra, P + tk.c0, ld, rb, tk.kbeg, tk.kend, sm, acc);
    const int t = threadIdx.x, w = t >> 6;
    const int32_t* rows = plan.sn_rows + tk.aux2;
    if (t < rb) {
        const int32_t gj = rows[tk.c0 + t];
        const int32_t s = plan.col_to_sn[gj];
        const int32_t f = plan.sn_first[s];
        const int64_t rp = plan.sn_rowptr[s];
        const int32_t ldt = (int32_t)(plan.sn_rowptr[s + 1] - rp);
        tbase[t] = plan.sn_panel[s] + (int64_t)(gj - f) * ldt;
        trow0[t] = rp;
        tlo[t] = gj - f;
        tld[t] = ldt;
        tfirst[t] = f;
        tnc[t] = plan.sn_first[s + 1] - f;
    }
    T* Cs = reinterpret_cast<T*>(&sm);   // 64 x DP staging tile (the GEMM buffers are free now)
    for (int half = 0; half < 2; ++half) {
        __syncthreads();
        if ((w >> 1) == half)
            for_acc<T>(acc, [&](int r, int c, T v) { Cs[c * DP + (r - 64 * half)] = v; });
        __syncthreads();
        const int rl = t & 63, cg = t >> 6, rr = 64 * half + rl;
        if (rr >= ra) continue;
        const int32_t gi = rows[tk.r0 + rr];
        for (int q = 0; q < 16; ++q) {
            const int c = cg + 4 * q;
            if (c >= rb || tk.r0 + rr < tk.c0 + c) continue;
            const int32_t d = gi - tfirst[c];
            int64_t pos = d;
            if (d < tlo[c] || d >= tnc[c]) {              // below the target's columns: search
                const int32_t* trows = plan.sn_rows + trow0[c];
                pos = lower_bound_i32(trows, max(tlo[c], tnc[c]), tld[c], gi);
                if (pos >= tld[c] || trows[pos] != gi) {   // structure violated: report, never write
                    atomicOr(err, 2);
                    continue;
                }
            }
            atomic_add(L + tbase[c] + pos, -Cs[c * DP + rl]);
        }
    }
}

int grid8(int32_t ntask) { return (ntask + 7) & ~7; }

struct CholArgs {
    void *panel, *diag, *piv, *stage;
    const lkd_snplan* plan;
    int32_t* err;
    void* work;
};

template <typename T>
int launch_chol(int kind, int32_t ntask, const lkd_task* tasks, const CholArgs& a, void* stream) {
    T* L = static_cast<T*>(a.panel);
    switch (kind) {
        case 1: LK_LAUNCH(k_diag<T>, grid8(ntask), 256, 0, stream, tasks, ntask, L, (T*)a.diag, (T*)a.piv, (T*)a.stage, a.err); break;
        case 2: LK_LAUNCH(k_update<T>, grid8(ntask), 256, 0, stream, tasks, ntask, L, *a.plan, a.err); break;
        case 3: LK_LAUNCH(k_diagwb<T>, grid8(ntask), 256, 0, stream, tasks, ntask, L, (const T*)a.diag, (const T*)a.piv); break;
        case 8: LK_LAUNCH(k_psum<T>, grid8(ntask), 256, 0, stream, tasks, ntask, (T*)a.work); break;
        case 6: {
            const int halves = sizeof(T) == 8 ? 1 : 2;
            LK_LAUNCH(k_part<T>, grid8(ntask * halves), 256, 0, stream, tasks, ntask, L, (T*)a.diag, (T*)a.piv,
                      (T*)a.stage, (T*)a.work, a.err, halves);
            break;
        }
        default: LK_LAUNCH(k_step<T>, grid8(ntask), 256, 0, stream, tasks, ntask, L, (T*)a.diag, (T*)a.piv, (T*)a.stage, (const T*)a.work, a.err); break;
    }
    return lk_last_error();
}

int dispatch(int prec, int kind, int32_t ntask, const lkd_task* tasks, const CholArgs& a,
             void* stream) {
    if (ntask <= 0) return 0;
    return prec == LKD_F32 ? launch_chol<float>(kind, ntask, tasks, a, stream)
                           : launch_chol<double>(kind, ntask, tasks, a, stream);
}

template <typename T>
__global__ __launch_bounds__(256) void k_probe_gemm(const T* pa, int64_t lda, int ra, const T* pb,
                                                    int64_t ldb, int rb, int K, T* out) {
    __shared__ GemmLds<T> sm;
    Acc<T> acc[2][4];
    zero_acc<T, 2, 4>(acc);
    gemm_nt<T>(pa, lda, ra, pb, ldb, rb, 0, K, sm, acc);
    // rows >= ra and columns >= rb of the product are unspecified (the callers discard them):
    // written as zero here
    for_acc<T>(acc, [&](int r, int c, T v) { out[c * MT + r] = r < ra && c < rb ? v : T(0); });
    // the transposed variant must produce the same product
    Acc<T> tr[4][2];
    zero_acc<T, 4, 2>(tr);
    gemm_tn<T>(pa, lda, ra, pb, ldb, rb, 0, K, sm, tr);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                const int c = 16 * p + Mfma<T>::row(lane, e), r = 32 * w + 16 * b + (lane & 15);
                out[MT * NC + c * MT + r] = r < ra && c < rb ? tr[p][b][e] : T(0);
            }
}

// The FP64 direct-to-LDS products against the register-staged TN core: out[0 ..) gemm_tn,
// out[8192 ..) gemm_tn_dma, out[16384 ..) and out[24576 ..) the two 64-column halves of
// gemm_tn_dma2 (B rows 0 .. 63 and 64 .. 127); each 128 x 64 column-major, zero past ra / the
// half's B rows.
__global__ __launch_bounds__(256) void k_probe_dma(const double* pa, int64_t lda, int ra, const double* pb,
                                                   int64_t ldb, int rb, int K, double* out) {
    __shared__ GemmLdsDma sm;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    auto put = [&](double* o, Acc<double> (&tr)[4][2], int rows_b) {
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int c = 16 * p + Mfma<double>::row(lane, e), r = 32 * w + 16 * b + (lane & 15);
                    o[c * MT + r] = r < ra && c < rows_b ? tr[p][b][e] : 0.0;
                }
    };
    {
        Acc<double> tr[4][2];
        zero_acc<double, 4, 2>(tr);
        gemm_tn<double>(pa, lda, ra, pb, ldb, min(rb, NC), 0, K, reinterpret_cast<GemmLds<double>&>(sm), tr);
        put(out, tr, min(rb, NC));
        __syncthreads();
    }
    {
        Acc<double> tr[4][2];
        zero_acc<double, 4, 2>(tr);
        gemm_tn_dma(pa, lda, ra, pb, ldb, min(rb, NC), 0, K, sm, tr);
        put(out + MT * NC, tr, min(rb, NC));
    }
    {
        Acc<double> tr2[8][2];
        zero_acc<double, 8, 2>(tr2);
        gemm_tn_dma2(pa, lda, ra, pb, ldb, rb, 0, K, reinterpret_cast<GemmLdsDma2&>(sm), tr2);
        Acc<double> h0[4][2], h1[4][2];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                h0[p][b] = tr2[p][b];
                h1[p][b] = tr2[p + 4][b];
            }
        put(out + 2 * MT * NC, h0, min(rb, NC));
        put(out + 3 * MT * NC, h1, max(rb - NC, 0));
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_probe_potrf(T* a, int32_t* ok) {
    __shared__ T Ds[TS * DP];
    __shared__ T Lv[4 * 256];
    __shared__ int s_ok;
    for (int idx = threadIdx.x; idx < TS * TS; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        Ds[c * DP + r] = a[c * TS + r];
    }
    __syncthreads();
    const bool good = potrf64<T>(Ds, Lv, &s_ok);
    if (threadIdx.x == 0) *ok = good ? 1 : 0;
    if (!good) return;
    for (int idx = threadIdx.x; idx < TS * TS; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        a[c * TS + r] = r >= c ? Ds[c * DP + r] : T(0);
    }
}

}  // namespace

extern "C" {

// out[0 .. 128*64) (column-major 128 x 64) = A(ra x K) * B(rb x K)^T with the NT GEMM core;
// out[128*64 .. 2*128*64) = the same product from the transposed (TN) core
int lkd_probe_tile_gemm(int prec, const void* a, int64_t lda, int ra, const void* b, int64_t ldb,
                        int rb, int K, void* out, void* stream) {
    if (ra < 0 || ra > MT || rb < 0 || rb > NC || K < 0) return (int)hipErrorInvalidValue;
    if (prec == LKD_F32)
        LK_LAUNCH(k_probe_gemm<float>, 1, 256, 0, stream, (const float*)a, lda, ra, (const float*)b, ldb, rb, K, (float*)out);
    else
        LK_LAUNCH(k_probe_gemm<double>, 1, 256, 0, stream, (const double*)a, lda, ra, (const double*)b, ldb, rb, K, (double*)out);
    return lk_last_error();
}

// FP64 direct-to-LDS TN products against the register-staged one (k_probe_dma); K a multiple
// of 16, ra <= 128, rb <= 128, lda > ra and ldb > rb (a row pair may read one row past ra / rb)
int lkd_probe_tile_gemm_dma(const double* a, int64_t lda, int ra, const double* b, int64_t ldb, int rb,
                            int K, double* out, void* stream) {
    if (ra < 1 || ra > MT || rb < 1 || rb > 2 * NC || K < 0 || K % 16 || lda <= ra || ldb <= rb)
        return (int)hipErrorInvalidValue;
    LK_LAUNCH(k_probe_dma, 1, 256, 0, stream, a, lda, ra, b, ldb, rb, K, out);
    return lk_last_error();
}

// In-place Cholesky of one 64 x 64 column-major block with the diagonal factorization
int lkd_probe_potrf(int prec, void* a, int32_t* ok, void* stream) {
    if (prec == LKD_F32) LK_LAUNCH(k_probe_potrf<float>, 1, 256, 0, stream, (float*)a, ok);
    else LK_LAUNCH(k_probe_potrf<double>, 1, 256, 0, stream, (double*)a, ok);
    return lk_last_error();
}

int lkd_chol_diag(int prec, int32_t ntask, const lkd_task* tasks, const void* panel, void* diag,
                  void* piv, void* stage, int32_t* err, void* stream) {
    return dispatch(prec, 1, ntask, tasks, {const_cast<void*>(panel), diag, piv, stage, nullptr, err, nullptr}, stream);
}
int lkd_chol_part(int prec, int32_t ntask, const lkd_task* tasks, const void* panel, void* diag,
                  void* piv, void* stage, void* work, int32_t* err, void* stream) {
    return dispatch(prec, 6, ntask, tasks,
                    {const_cast<void*>(panel), diag, piv, stage, nullptr, err, work}, stream);
}
int lkd_chol_psum(int prec, int32_t ntask, const lkd_task* tasks, void* work, void* stream) {
    return dispatch(prec, 8, ntask, tasks, {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, work}, stream);
}
int lkd_chol_step(int prec, int32_t ntask, const lkd_task* tasks, void* panel, void* diag,
                  void* piv, void* stage, const void* work, int32_t* err, void* stream) {
    return dispatch(prec, 5, ntask, tasks, {panel, diag, piv, stage, nullptr, err, const_cast<void*>(work)}, stream);
}
int lkd_chol_step_dag(int prec, int32_t ntask, const lkd_task* tasks, const int32_t* deps,
                      const int32_t* sdeps, const int32_t* queue, const int32_t* qoff,
                      int32_t nx, int32_t grid, uint32_t* flags, uint32_t* flags2, uint32_t* ctl,
                      int32_t q, void* panel, void* diag, void* piv, void* stage, void* work,
                      int32_t* err, void* stream) {
    if (ntask <= 0) return 0;
    if (q < 0 || nx < 1 || grid < 1) return (int)hipErrorInvalidValue;
    uint32_t* cnt = ctl + 1 + (int64_t)q * nx;
    if (prec == LKD_F32)
        LK_LAUNCH(k_step_dag<float>, grid, 256, 0, stream, tasks, deps, sdeps, queue, qoff, nx,
                  flags, flags2, ctl, cnt, (float*)panel, (float*)diag, (float*)piv,
                  (float*)stage, (float*)work, err);
    else
        LK_LAUNCH(k_step_dag<double>, grid, 256, 0, stream, tasks, deps, sdeps, queue, qoff, nx,
                  flags, flags2, ctl, cnt, (double*)panel, (double*)diag, (double*)piv,
                  (double*)stage, (double*)work, err);
    return lk_last_error();
}
int lkd_chol_dag_begin(uint32_t* ctl, int32_t nq, void* stream) {
    if (nq < 0) return (int)hipErrorInvalidValue;
    LK_LAUNCH(k_dag_begin, 1, 256, 0, stream, ctl, nq);
    return lk_last_error();
}
int lkd_chol_update(int prec, int32_t ntask, const lkd_task* tasks, void* panel,
                    lkd_snplan plan, int32_t* err, void* stream) {
    return dispatch(prec, 2, ntask, tasks, {panel, nullptr, nullptr, nullptr, &plan, err, nullptr}, stream);
}
int lkd_chol_diagwb(int prec, int32_t ntask, const lkd_task* tasks, void* panel, const void* diag,
                    const void* piv, void* stream) {
    return dispatch(prec, 3, ntask, tasks,
                    {panel, const_cast<void*>(diag), const_cast<void*>(piv), nullptr, nullptr, nullptr, nullptr}, stream);
}

}  // extern "C"

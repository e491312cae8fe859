// Supernodal Cholesky tile kernels for gfx950 (replace cholesky_level_kernel and its per-level
// launch loop, src/cuda/kkt.cu:217-285,582-596, and the repeated refactorization in
// gpu_cholesky_solve_device, src/cuda/kkt.cu:671-689).
//
// Storage: every supernode owns a column-major panel (ld = rows of its structure, ncol columns)
// at panel + off.  The host plan drives a right-looking blocked factorization per etree level:
//   POTRF  first 64-column diagonal block of each supernode (one workgroup, LDS)
//   TRSM   64..256-row tiles of the current block column:  X <- X * Lkk^{-T}
//   SYRK   64x64 trailing tiles  C -= B_r * B_c^T  on MFMA (16x16x4 FP64 / FP32); the workgroup
//          owning the next diagonal tile also factors it (look-ahead, no extra launch)
//   UPDATE below-diagonal rows:  U = L21 * L21^T  scattered (atomic add) into ancestor panels
// All workgroups are 256 threads (4 wavefronts of 64).
#include "common.hpp"

namespace {

constexpr int TS = 64;   // tile edge
constexpr int DP = 65;   // LDS pitch of a 64x64 tile (one element of padding)

template <typename T> struct KChunk { static constexpr int v = 32; };
template <> struct KChunk<float> { static constexpr int v = 64; };

template <typename T> constexpr int gemm_lds_elems() {
    return (2 * KChunk<T>::v * TS > TS * DP) ? 2 * KChunk<T>::v * TS : TS * DP;
}

// acc (this wave's 2x2 blocks of 16x16 in the 64x64 tile) += sum_k A(r,k) * B(c,k)
// A(r,k) = pa[k*lda + r] (r < ra), B(c,k) = pb[k*ldb + c] (c < rb); other rows read as zero.
template <typename T>
__device__ __forceinline__ void tile_gemm_nt(const T* __restrict__ pa, int64_t lda, int ra,
                                             const T* __restrict__ pb, int64_t ldb, int rb,
                                             int K, T* lds, typename Mfma<T>::acc (&acc)[2][2]) {
    constexpr int KC = KChunk<T>::v;
    T* As = lds;
    T* Bs = lds + KC * TS;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const int lr = t & 63, lq = t >> 6;
    for (int kc = 0; kc < K; kc += KC) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KC / 4; ++q) {
            const int k = lq + 4 * q;
            const int kg = kc + k;
            As[k * TS + lr] = (lr < ra && kg < K) ? pa[(int64_t)kg * lda + lr] : T(0);
            Bs[k * TS + lr] = (lr < rb && kg < K) ? pb[(int64_t)kg * ldb + lr] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KC; k += 4) {
            const T a0 = As[(k + lk) * TS + wr + li], a1 = As[(k + lk) * TS + wr + 16 + li];
            const T b0 = Bs[(k + lk) * TS + wc + li], b1 = Bs[(k + lk) * TS + wc + 16 + li];
            acc[0][0] = Mfma<T>::mma(a0, b0, acc[0][0]);
            acc[0][1] = Mfma<T>::mma(a0, b1, acc[0][1]);
            acc[1][0] = Mfma<T>::mma(a1, b0, acc[1][0]);
            acc[1][1] = Mfma<T>::mma(a1, b1, acc[1][1]);
        }
    }
    __syncthreads();
}

// Accumulators -> LDS tile Cs[c * DP + r].
template <typename T>
__device__ __forceinline__ void stage_acc(typename Mfma<T>::acc (&acc)[2][2], T* Cs) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = wr + 16 * x + Mfma<T>::row(lane, e);
                const int c = wc + 16 * y + (lane & 15);
                Cs[c * DP + r] = acc[x][y][e];
            }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void zero_acc(typename Mfma<T>::acc (&acc)[2][2]) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = typename Mfma<T>::acc{0, 0, 0, 0};
}

// In-place right-looking Cholesky of the kb x kb lower block Ds[c*DP + r] (kb <= 64).
// Thread (r = t % 64, cg = t / 64) owns rows r of columns cg, cg+4, ...  One barrier per column:
// column j is scaled during iteration j+1, when nobody reads it any more.
// Returns false (uniformly across the workgroup) when a pivot is not positive.
template <typename T>
__device__ bool lds_potrf(T* Ds, int kb) {
    const int t = threadIdx.x;
    const int r = t & 63, cg = t >> 6;
    T sd_prev = T(1);
    for (int j = 0; j < kb; ++j) {
        __syncthreads();
        const T d = Ds[j * DP + j];
        if (!(d > T(0))) return false;
        if (j > 0 && ((j - 1) & 3) == cg && r >= j - 1 && r < kb) {
            Ds[(j - 1) * DP + r] = (r == j - 1) ? sd_prev : Ds[(j - 1) * DP + r] / sd_prev;
        }
        const T invd = T(1) / d;
        if (r < kb && r > j) {
            const T arj = Ds[j * DP + r] * invd;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int k = cg + 4 * q;
                if (k > j && k <= r) Ds[k * DP + r] -= arj * Ds[j * DP + k];
            }
        }
        sd_prev = sqrt(d);
    }
    __syncthreads();
    if (kb > 0 && ((kb - 1) & 3) == cg && r == kb - 1) Ds[(kb - 1) * DP + r] = sd_prev;
    __syncthreads();
    return true;
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

template <typename T>
__global__ __launch_bounds__(256) void k_potrf(const lkd_task* __restrict__ tasks,
                                               T* __restrict__ L, T* __restrict__ diag,
                                               int32_t* err) {
    if (*(volatile int32_t*)err) return;
    const lkd_task tk = tasks[blockIdx.x];
    __shared__ T Ds[TS * DP];
    T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int k0 = tk.k0, kb = tk.kb;
    for (int idx = threadIdx.x; idx < TS * TS; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        if (r < kb && c < kb && r >= c) Ds[c * DP + r] = P[(int64_t)(k0 + c) * ld + k0 + r];
    }
    if (!lds_potrf<T>(Ds, kb)) {
        if (threadIdx.x == 0) atomicOr(err, 1);
        return;
    }
    for (int idx = threadIdx.x; idx < TS * TS; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        if (r < kb && c < kb && r >= c) P[(int64_t)(k0 + c) * ld + k0 + r] = Ds[c * DP + r];
    }
    write_diag_copy<T>(Ds, kb, diag + tk.aux);
}

// X (256 rows starting at r0, columns k0..k0+kb) <- X * Lkk^{-T}: one thread per row, forward
// substitution with the unknowns in registers.  Lkk comes from the row-major diagonal copy
// (uniform addresses -> scalar loads, so the coefficients never occupy vector registers).
template <typename T, int KB>
__device__ __forceinline__ void trsm_rows(const T* __restrict__ Lr, T* __restrict__ P, int64_t ld,
                                          int k0, int64_t row, bool valid) {
    T x[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) x[k] = valid ? P[(int64_t)(k0 + k) * ld + row] : T(0);
#pragma unroll
    for (int j = 0; j < KB; ++j) {
        T s0 = x[j], s1 = T(0);
#pragma unroll
        for (int k = 0; k < j; k += 2) {
            s0 -= x[k] * Lr[j * KB + k];
            if (k + 1 < j) s1 -= x[k + 1] * Lr[j * KB + k + 1];
        }
        x[j] = (s0 + s1) * Lr[j * KB + j];
    }
    if (valid) {
#pragma unroll
        for (int k = 0; k < KB; ++k) P[(int64_t)(k0 + k) * ld + row] = x[k];
    }
}

template <typename T>
__device__ void trsm_rows_dyn(const T* __restrict__ Lr, int kb, T* __restrict__ P, int64_t ld,
                              int k0, int64_t row, bool valid) {
    // generic width (kb < 64, last block column): unknowns kept in the panel itself
    if (!valid) return;
    for (int j = 0; j < kb; ++j) {
        T s = P[(int64_t)(k0 + j) * ld + row];
        for (int k = 0; k < j; ++k) s -= P[(int64_t)(k0 + k) * ld + row] * Lr[j * kb + k];
        P[(int64_t)(k0 + j) * ld + row] = s * Lr[j * kb + j];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_trsm(const lkd_task* __restrict__ tasks,
                                              T* __restrict__ L, const T* __restrict__ diag,
                                              const int32_t* err) {
    if (*(volatile const int32_t*)err) return;
    const lkd_task tk = tasks[blockIdx.x];
    T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int64_t row = (int64_t)tk.r0 + threadIdx.x;
    const bool valid = row < ld;
    const T* Lr = diag + tk.aux;
    if (tk.kb == TS) trsm_rows<T, TS>(Lr, P, ld, tk.k0, row, valid);
    else trsm_rows_dyn<T>(Lr, tk.kb, P, ld, tk.k0, row, valid);
}

// Trailing update of one 64x64 tile (rows r0.., columns c0..) with block column k0..k0+kb.
// The tile at (k0+kb, k0+kb) is the next diagonal block: its workgroup factors it afterwards.
template <typename T>
__global__ __launch_bounds__(256) void k_syrk(const lkd_task* __restrict__ tasks,
                                              T* __restrict__ L, T* __restrict__ diag,
                                              int32_t* err) {
    if (*(volatile int32_t*)err) return;
    const lkd_task tk = tasks[blockIdx.x];
    __shared__ T lds[gemm_lds_elems<T>()];
    T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int ra = min(TS, tk.ld - tk.r0), rb = min(TS, tk.ncol - tk.c0);
    typename Mfma<T>::acc acc[2][2];
    zero_acc<T>(acc);
    tile_gemm_nt<T>(P + (int64_t)tk.k0 * ld + tk.r0, ld, ra, P + (int64_t)tk.k0 * ld + tk.c0, ld,
                    rb, tk.kb, lds, acc);
    stage_acc<T>(acc, lds);
    T* C = P + (int64_t)tk.c0 * ld + tk.r0;
    const int r = threadIdx.x & 63, cg = threadIdx.x >> 6;
    const bool next_diag = (tk.r0 == tk.c0) && (tk.c0 == tk.k0 + tk.kb);
    if (!next_diag) {
        if (r < ra) {
#pragma unroll 4
            for (int q = 0; q < 16; ++q) {
                const int c = cg + 4 * q;
                if (c < rb) C[(int64_t)c * ld + r] -= lds[c * DP + r];
            }
        }
        return;
    }
    const int kb2 = rb;
    if (r < ra) {
        for (int q = 0; q < 16; ++q) {
            const int c = cg + 4 * q;
            if (c < rb) {
                const T v = C[(int64_t)c * ld + r] - lds[c * DP + r];
                if (r < kb2) lds[c * DP + r] = v;
                else C[(int64_t)c * ld + r] = v;
            }
        }
    }
    if (!lds_potrf<T>(lds, kb2)) {
        if (threadIdx.x == 0) atomicOr(err, 1);
        return;
    }
    if (r < kb2) {
        for (int q = 0; q < 16; ++q) {
            const int c = cg + 4 * q;
            if (c <= r) C[(int64_t)c * ld + r] = lds[c * DP + r];
        }
    }
    write_diag_copy<T>(lds, kb2, diag + tk.aux);
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

// Contribution of the below-diagonal rows of supernode sn to its ancestors:
// U = L21[r0 rows] * L21[c0 rows]^T (K = ncol) subtracted from the ancestor panels.
template <typename T>
__global__ __launch_bounds__(256) void k_update(const lkd_task* __restrict__ tasks,
                                                T* __restrict__ L, lkd_snplan plan,
                                                int32_t* err) {
    if (*(volatile int32_t*)err) return;
    const lkd_task tk = tasks[blockIdx.x];
    __shared__ T lds[gemm_lds_elems<T>()];
    __shared__ int64_t tbase[TS];
    __shared__ int64_t trow0[TS];
    __shared__ int64_t tlo[TS];
    __shared__ int32_t tld[TS];
    T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int ra = min(TS, tk.ld - tk.r0), rb = min(TS, tk.ld - tk.c0);
    typename Mfma<T>::acc acc[2][2];
    zero_acc<T>(acc);
    tile_gemm_nt<T>(P + tk.r0, ld, ra, P + tk.c0, ld, rb, tk.ncol, lds, acc);
    stage_acc<T>(acc, lds);
    const int32_t* rows = plan.sn_rows + tk.aux2;
    if (threadIdx.x < rb) {
        const int32_t gj = rows[tk.c0 + threadIdx.x];
        const int32_t t = plan.col_to_sn[gj];
        const int32_t f = plan.sn_first[t];
        const int64_t rp = plan.sn_rowptr[t];
        const int32_t ldt = (int32_t)(plan.sn_rowptr[t + 1] - rp);
        tbase[threadIdx.x] = plan.sn_panel[t] + (int64_t)(gj - f) * ldt;
        trow0[threadIdx.x] = rp;
        tlo[threadIdx.x] = gj - f;
        tld[threadIdx.x] = ldt;
    }
    __syncthreads();
    const int r = threadIdx.x & 63, cg = threadIdx.x >> 6;
    if (r >= ra) return;
    const int32_t gi = rows[tk.r0 + r];
    for (int q = 0; q < 16; ++q) {
        const int c = cg + 4 * q;
        if (c >= rb || tk.r0 + r < tk.c0 + c) continue;
        const int32_t* trows = plan.sn_rows + trow0[c];
        const int64_t pos = lower_bound_i32(trows, tlo[c], tld[c], gi);
        if (pos >= tld[c] || trows[pos] != gi) {   // structure violated: report, never write
            atomicOr(err, 2);
            continue;
        }
        atomic_add(L + tbase[c] + pos, -lds[c * DP + r]);
    }
}

template <typename T>
int launch_chol(int kind, int32_t ntask, const lkd_task* tasks, void* panel, void* diagbuf,
                const lkd_snplan* plan, int32_t* err, void* stream) {
    T* L = static_cast<T*>(panel);
    T* D = static_cast<T*>(diagbuf);
    switch (kind) {
        case 0: LK_LAUNCH(k_potrf<T>, ntask, 256, 0, stream, tasks, L, D, err); break;
        case 1: LK_LAUNCH(k_trsm<T>, ntask, 256, 0, stream, tasks, L, D, err); break;
        case 2: LK_LAUNCH(k_syrk<T>, ntask, 256, 0, stream, tasks, L, D, err); break;
        default: LK_LAUNCH(k_update<T>, ntask, 256, 0, stream, tasks, L, *plan, err); break;
    }
    return lk_last_error();
}

int dispatch(int prec, int kind, int32_t ntask, const lkd_task* tasks, void* panel, void* diag,
             const lkd_snplan* plan, int32_t* err, void* stream) {
    if (ntask <= 0) return 0;
    return prec == LKD_F32
               ? launch_chol<float>(kind, ntask, tasks, panel, diag, plan, err, stream)
               : launch_chol<double>(kind, ntask, tasks, panel, diag, plan, err, stream);
}

template <typename T>
__global__ __launch_bounds__(256) void k_probe_gemm(const T* pa, int64_t lda, int ra, const T* pb,
                                                    int64_t ldb, int rb, int K, T* out) {
    __shared__ T lds[gemm_lds_elems<T>()];
    typename Mfma<T>::acc acc[2][2];
    zero_acc<T>(acc);
    tile_gemm_nt<T>(pa, lda, ra, pb, ldb, rb, K, lds, acc);
    stage_acc<T>(acc, lds);
    for (int idx = threadIdx.x; idx < TS * TS; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        out[c * TS + r] = lds[c * DP + r];
    }
}

}  // namespace

extern "C" {

// out (64x64 column-major) = A(64 x K) * B(64 x K)^T with the tile GEMM of the SYRK/UPDATE kernels
int lkd_probe_tile_gemm(int prec, const void* a, int64_t lda, int ra, const void* b, int64_t ldb,
                        int rb, int K, void* out, void* stream) {
    if (prec == LKD_F32)
        LK_LAUNCH(k_probe_gemm<float>, 1, 256, 0, stream, (const float*)a, lda, ra, (const float*)b, ldb, rb, K, (float*)out);
    else
        LK_LAUNCH(k_probe_gemm<double>, 1, 256, 0, stream, (const double*)a, lda, ra, (const double*)b, ldb, rb, K, (double*)out);
    return lk_last_error();
}


int lkd_chol_potrf(int prec, int32_t ntask, const lkd_task* tasks, void* panel, void* diag,
                   int32_t* err, void* stream) {
    return dispatch(prec, 0, ntask, tasks, panel, diag, nullptr, err, stream);
}
int lkd_chol_trsm(int prec, int32_t ntask, const lkd_task* tasks, void* panel, void* diag,
                  int32_t* err, void* stream) {
    return dispatch(prec, 1, ntask, tasks, panel, diag, nullptr, err, stream);
}
int lkd_chol_syrk(int prec, int32_t ntask, const lkd_task* tasks, void* panel, void* diag,
                  int32_t* err, void* stream) {
    return dispatch(prec, 2, ntask, tasks, panel, diag, nullptr, err, stream);
}
int lkd_chol_update(int prec, int32_t ntask, const lkd_task* tasks, void* panel,
                    lkd_snplan plan, int32_t* err, void* stream) {
    return dispatch(prec, 3, ntask, tasks, panel, nullptr, &plan, err, stream);
}

}  // extern "C"

// Supernodal triangular solves L z = r and L^T x = z on the factor panels (replace the
// cusparseSpSV analysis + solve pairs of src/cuda/kkt.cu:731-734,788-789, which re-ran the
// analysis on every call).  Levels of the supernodal etree are processed leaves-first (forward)
// and root-first (backward); inside a supernode, 64-column blocks are processed in order.
//  * diagonal blocks: one wavefront, lane i keeps row (forward) or column (backward) i of the
//    row-major diagonal copy in registers; broadcasts by lane shuffles, no barriers;
//  * off-diagonal parts: coalesced GEMV (forward, one thread per row) or LDS-staged GEMV^T
//    (backward); rows owned by ancestors are updated with atomic adds (forward).
// Vectors are FP64 for both FP64 and FP32 panels.
#include "common.hpp"

namespace {

constexpr int TS = 64;
constexpr int DP = 65;

// Forward substitution on a diagonal block by wave 0.  b: this lane's right-hand side entry
// (lane < kb), Lr: row-major copy (stride kb, reciprocal diagonal).  Returns z for this lane.
template <typename T>
__device__ __forceinline__ double wave_fsub(double b, const T* __restrict__ Lr, int kb) {
    const int lane = threadIdx.x & 63;
    T row[TS];
#pragma unroll
    for (int k = 0; k < TS; ++k) row[k] = (k < lane && lane < kb) ? Lr[lane * kb + k] : T(0);
#pragma unroll
    for (int j = 0; j < TS; ++j) {
        if (j < kb) {
            const double zj = lk_shfl(b, j) * (double)Lr[j * kb + j];
            if (lane == j) b = zj;
            else if (lane > j) b -= (double)row[j] * zj;
        }
    }
    return b;
}

// Backward substitution L^T x = r on a diagonal block by wave 0 (lane j keeps column j).
template <typename T>
__device__ __forceinline__ double wave_bsub(double r, const T* __restrict__ Lr, int kb) {
    const int lane = threadIdx.x & 63;
    T col[TS];
#pragma unroll
    for (int i = 0; i < TS; ++i) col[i] = (i > lane && i < kb) ? Lr[i * kb + lane] : T(0);
#pragma unroll
    for (int i = TS - 1; i >= 0; --i) {
        if (i < kb) {
            const double xi = lk_shfl(r, i) * (double)Lr[i * kb + i];
            if (lane == i) r = xi;
            else if (lane < i) r -= (double)col[i] * xi;
        }
    }
    return r;
}

template <typename T>
__global__ __launch_bounds__(64) void k_fwd_first(const lkd_task* __restrict__ tasks,
                                                  const T* __restrict__ diag,
                                                  double* __restrict__ x) {
    const lkd_task tk = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    double* xs = x + tk.first + tk.k0;
    const double b = lane < tk.kb ? xs[lane] : 0.0;
    const double z = wave_fsub<T>(b, diag + tk.aux, tk.kb);
    if (lane < tk.kb) xs[lane] = z;
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_step(const lkd_task* __restrict__ tasks,
                                                  const T* __restrict__ L,
                                                  const T* __restrict__ diag,
                                                  const int32_t* __restrict__ sn_rows,
                                                  double* __restrict__ x) {
    const lkd_task tk = tasks[blockIdx.x];
    __shared__ double zs[TS];
    __shared__ double nb[TS];
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    if (threadIdx.x < tk.kb) zs[threadIdx.x] = x[tk.first + tk.k0 + threadIdx.x];
    __syncthreads();
    const int64_t row = (int64_t)tk.r0 + threadIdx.x;
    const bool next_diag = tk.r0 == tk.k0 + tk.kb && tk.r0 < tk.ncol;
    const int kb2 = next_diag ? min(TS, tk.ncol - tk.r0) : 0;
    if (row < ld) {
        double acc0 = 0.0, acc1 = 0.0;
        const T* col = P + (int64_t)tk.k0 * ld + row;
        int j = 0;
        for (; j + 1 < tk.kb; j += 2) {
            acc0 += (double)col[(int64_t)j * ld] * zs[j];
            acc1 += (double)col[(int64_t)(j + 1) * ld] * zs[j + 1];
        }
        if (j < tk.kb) acc0 += (double)col[(int64_t)j * ld] * zs[j];
        const double acc = acc0 + acc1;
        if (row < tk.ncol) {
            const double v = x[tk.first + row] - acc;
            if (threadIdx.x < kb2) nb[threadIdx.x] = v;
            else x[tk.first + row] = v;
        } else {
            atomic_add(x + sn_rows[tk.aux2 + row], -acc);
        }
    }
    if (!next_diag) return;
    __syncthreads();
    if (threadIdx.x < TS) {
        const int lane = threadIdx.x;
        const double b = lane < kb2 ? nb[lane] : 0.0;
        const double z = wave_fsub<T>(b, diag + tk.aux, kb2);
        if (lane < kb2) x[tk.first + tk.r0 + lane] = z;
    }
}

// Backward: columns c0..c0+kbc of the supernode subtract L(rows, cols)^T x(rows), where rows are
// the below rows (below = 1, rows ncol..ld, x gathered through the row list) or the rows of the
// already-solved block k0..k0+kb (below = 0).  The block that is next in line then solves.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd(const lkd_task* __restrict__ tasks,
                                             const T* __restrict__ L,
                                             const T* __restrict__ diag,
                                             const int32_t* __restrict__ sn_rows,
                                             double* __restrict__ x, int below) {
    const lkd_task tk = tasks[blockIdx.x];
    __shared__ T Ts[TS * DP];
    __shared__ double xr[TS];
    __shared__ double part[4][TS];
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const int c0 = tk.c0;
    const int kbc = min(TS, tk.ncol - c0);
    const int t = threadIdx.x, r = t & 63, g = t >> 6;
    int64_t rbeg, rend;
    if (below) { rbeg = tk.ncol; rend = ld; }
    else { rbeg = tk.k0; rend = tk.k0 + tk.kb; }
    double acc = 0.0;
    for (int64_t rb = rbeg; rb < rend; rb += TS) {
        const int nr = (int)min((int64_t)TS, rend - rb);
        __syncthreads();
        if (t < nr) xr[t] = below ? x[sn_rows[tk.aux2 + rb + t]] : x[tk.first + rb + t];
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const int c = g + 4 * q;
            Ts[c * DP + r] = (r < nr && c < kbc) ? P[(int64_t)(c0 + c) * ld + rb + r] : T(0);
        }
        __syncthreads();
        // thread (column r, group g) sums rows g, g+4, ...
        if (r < kbc) {
            for (int i = g; i < nr; i += 4) acc += (double)Ts[r * DP + i] * xr[i];
        }
    }
    part[g][r] = acc;
    __syncthreads();
    const bool last = below ? (c0 + kbc == tk.ncol) : (c0 + kbc == tk.k0);
    if (t < TS) {
        double v = 0.0;
        if (t < kbc) v = x[tk.first + c0 + t] - ((part[0][t] + part[1][t]) + (part[2][t] + part[3][t]));
        if (!last) {
            if (t < kbc) x[tk.first + c0 + t] = v;
        } else {
            const double sol = wave_bsub<T>(v, diag + tk.aux, kbc);
            if (t < kbc) x[tk.first + c0 + t] = sol;
        }
    }
}

template <typename T>
int fwd(int32_t ntask, const lkd_task* tasks, const void* panel, const void* diag,
        const int32_t* rows, double* x, int first, void* stream) {
    if (first)
        LK_LAUNCH(k_fwd_first<T>, ntask, 64, 0, stream, tasks, (const T*)diag, x);
    else
        LK_LAUNCH(k_fwd_step<T>, ntask, 256, 0, stream, tasks, (const T*)panel, (const T*)diag,
                  rows, x);
    return lk_last_error();
}

template <typename T>
int bwd(int32_t ntask, const lkd_task* tasks, const void* panel, const void* diag,
        const int32_t* rows, double* x, int below, void* stream) {
    LK_LAUNCH(k_bwd<T>, ntask, 256, 0, stream, tasks, (const T*)panel, (const T*)diag, rows, x,
              below);
    return lk_last_error();
}

}  // namespace

extern "C" {

int lkd_trsv_fwd(int prec, int32_t ntask, const lkd_task* tasks, const void* panel,
                 const void* diag, const int32_t* sn_rows, double* x, int first_step,
                 void* stream) {
    if (ntask <= 0) return 0;
    return prec == LKD_F32 ? fwd<float>(ntask, tasks, panel, diag, sn_rows, x, first_step, stream)
                           : fwd<double>(ntask, tasks, panel, diag, sn_rows, x, first_step, stream);
}

int lkd_trsv_bwd(int prec, int32_t ntask, const lkd_task* tasks, const void* panel,
                 const void* diag, const int32_t* sn_rows, double* x, int below_step,
                 void* stream) {
    if (ntask <= 0) return 0;
    return prec == LKD_F32 ? bwd<float>(ntask, tasks, panel, diag, sn_rows, x, below_step, stream)
                           : bwd<double>(ntask, tasks, panel, diag, sn_rows, x, below_step, stream);
}

}  // extern "C"

// Training and inference kernels of the warm-start network (a multilayer perceptron mapping the
// compressed (b, c) perturbation of an LP to its optimal point) for gfx950:
//   * one strided GEMM  C = alpha A B^T + beta C (+ bias, ReLU of A on load, mask of C) on
//     MFMA v_mfma_f32_16x16x16_bf16: operands are rounded to BF16 as they are staged in LDS,
//     products accumulate in FP32;
//   * column sums (bias gradients), the squared-error loss and its gradient, row gathers of a
//     device-resident data set, and the Adam update on FP32 master weights.
// All reductions run in a fixed order, so a training step is deterministic.
#include <algorithm>

#include "common.hpp"

namespace {

constexpr int BM = 64, BN = 64, BK = 64, KP = BK + 8;   // tile, K chunk, LDS pitch (bf16)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short to_bf16(float f) {
    const unsigned u = __float_as_uint(f);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct GemmArgs {
    int M, N, K;
    const float* A;            // A(m, k) = A[m * sam + k * sak]
    int64_t sam, sak;
    const float* B;            // B(n, k) = B[n * sbn + k * sbk]
    int64_t sbn, sbk;
    float* C;                  // C(m, n) = C[m * ldc + n]
    int64_t ldc;
    float alpha, beta;
    const float* bias;         // bias[n] added (nullable)
    const float* mask;         // C(m, n) zeroed where mask[m * ldc + n] <= 0 (nullable)
    int relu_a, relu_b;        // operand read as max(., 0)
    float* part;               // split K (gridDim.z > 1): raw sums of split z at part[z M N + m N + n]
    int kchunk;                // K columns per split (a multiple of BK)
};

// One 64 x 64 tile of C per workgroup (split z of K: columns z kchunk ..); wave w computes rows
// 32 (w / 2) .., columns 32 (w % 2) .. as 2 x 2 blocks of 16 x 16.  Operand element e of a
// 64 x 64 chunk goes to thread e % 256: along k when k is the unit stride, along the row
// otherwise, so global loads coalesce; all 16 loads of an operand are issued before the LDS
// stores.  Without a split the epilogue (alpha, beta, bias, mask) is applied here.
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    __shared__ unsigned short As[BM * KP];
    __shared__ unsigned short Bs[BN * KP];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = 32 * (w >> 1), wn = 32 * (w & 1);
    const int kb0 = blockIdx.z * g.kchunk;
    const int kb1 = min(g.K, kb0 + g.kchunk);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
    const bool a_k = g.sak == 1, b_k = g.sbk == 1;
    for (int k0 = kb0; k0 < kb1; k0 += BK) {
        float va[(BM * BK) / 256], vb[(BN * BK) / 256];
#pragma unroll
        for (int j = 0; j < (BM * BK) / 256; ++j) {
            const int e = t + 256 * j;
            const int r = a_k ? e / BK : e % BM, k = a_k ? e % BK : e / BM;
            const int gm = m0 + r, gk = k0 + k;
            va[j] = (gm < g.M && gk < kb1) ? g.A[gm * g.sam + gk * g.sak] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < (BN * BK) / 256; ++j) {
            const int e = t + 256 * j;
            const int r = b_k ? e / BK : e % BN, k = b_k ? e % BK : e / BN;
            const int gn = n0 + r, gk = k0 + k;
            vb[j] = (gn < g.N && gk < kb1) ? g.B[gn * g.sbn + gk * g.sbk] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < (BM * BK) / 256; ++j) {
            const int e = t + 256 * j;
            const int r = a_k ? e / BK : e % BM, k = a_k ? e % BK : e / BM;
            As[r * KP + k] = to_bf16(g.relu_a ? fmaxf(va[j], 0.0f) : va[j]);
        }
#pragma unroll
        for (int j = 0; j < (BN * BK) / 256; ++j) {
            const int e = t + 256 * j;
            const int r = b_k ? e / BK : e % BN, k = b_k ? e % BK : e / BN;
            Bs[r * KP + k] = to_bf16(g.relu_b ? fmaxf(vb[j], 0.0f) : vb[j]);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int ko = kk + 4 * (l >> 4);
            bf16x4 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const unsigned short* pa = As + (wm + 16 * i + (l & 15)) * KP + ko;
                const unsigned short* pb = Bs + (wn + 16 * i + (l & 15)) * KP + ko;
                a[i] = bf16x4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
                b[i] = bf16x4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // element e of lane l: row 4 (l / 16) + e, column l % 16 of its 16 x 16 block
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wm + 16 * i + 4 * (l >> 4) + e;
                const int n = n0 + wn + 16 * j + (l & 15);
                if (m >= g.M || n >= g.N) continue;
                if (gridDim.z > 1) {
                    g.part[((int64_t)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][e];
                    continue;
                }
                float* cp = g.C + m * g.ldc + n;
                float v = g.alpha * acc[i][j][e];
                if (g.beta != 0.0f) v += g.beta * *cp;
                if (g.bias) v += g.bias[n];
                if (g.mask && !(g.mask[m * g.ldc + n] > 0.0f)) v = 0.0f;
                *cp = v;
            }
}

// split-K epilogue: C(m, n) = alpha sum_z part[z][m][n] (z in order) + beta C + bias, mask
__global__ __launch_bounds__(256) void k_gemm_splitk_epi(GemmArgs g, int nsplit) {
    const int64_t total = (int64_t)g.M * g.N;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
        const int m = (int)(i / g.N), n = (int)(i - (int64_t)m * g.N);
        float s = g.part[i];
        for (int z = 1; z < nsplit; ++z) s += g.part[(int64_t)z * total + i];
        float* cp = g.C + m * g.ldc + n;
        float v = g.alpha * s;
        if (g.beta != 0.0f) v += g.beta * *cp;
        if (g.bias) v += g.bias[n];
        if (g.mask && !(g.mask[m * g.ldc + n] > 0.0f)) v = 0.0f;
        *cp = v;
    }
}

// out[n] = beta out[n] + sum_m G[m * ld + n]: a workgroup takes 32 columns, its 8 row groups
// sum rows g, g + 8, .. in order and the 8 partials are added in order (fixed summation order)
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ G, int M, int N, int64_t ld,
                                                float beta, float* __restrict__ out) {
    __shared__ float part[8][32];
    const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int n = blockIdx.x * 32 + c;
    float s = 0.0f;
    if (n < N)
        for (int m = g; m < M; m += 8) s += G[m * ld + n];
    part[g][c] = s;
    __syncthreads();
    if (g == 0 && n < N) {
        float t = part[0][c];
        for (int k = 1; k < 8; ++k) t += part[k][c];
        out[n] = beta * out[n] + t;
    }
}

// g = scale (Y - T); per-block partial sums of (Y - T)^2 (block order fixed)
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             int64_t count, float scale, float* __restrict__ g,
                                             float* __restrict__ part) {
    __shared__ float sh[4];
    float s = 0.0f;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) {
        const float d = Y[i] - T[i];
        g[i] = scale * d;
        s += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// out[0] = sum of part[0 .. n): 256 threads sum strided slices in FP64, then a fixed tree
__global__ __launch_bounds__(256) void k_sum_parts(const float* __restrict__ part, int n,
                                                   float* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)sh[0];
}

// dst[r, :] = src[idx[r], :]
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                                int rows, int cols, float* __restrict__ dst) {
    const int64_t total = (int64_t)rows * cols;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
        const int64_t r = i / cols, c = i - r * cols;
        dst[i] = src[(int64_t)idx[r] * cols + c];
    }
}

// Adam on FP32 master weights (bias-corrected moments; step t >= 1)
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              int64_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// one MFMA on 64 lanes: lane l supplies 4 bf16 of A and of B, out[4 l + e] = D[e]
__global__ void k_probe_bf16(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out) {
    const int l = threadIdx.x;
    bf16x4 av, bv;
    for (int i = 0; i < 4; ++i) {
        av[i] = (short)to_bf16(a[4 * l + i]);
        bv[i] = (short)to_bf16(b[4 * l + i]);
    }
    const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, f32x4{0, 0, 0, 0}, 0, 0, 0);
    for (int e = 0; e < 4; ++e) out[4 * l + e] = d[e];
}

int blocks_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace

extern "C" {

int lkd_mlp_gemm(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
                 int64_t sbn, int64_t sbk, float* C, int64_t ldc, float alpha, float beta,
                 const float* bias, const float* mask, int relu_a, int relu_b, float* work,
                 int64_t work_elems, void* stream) {
    if (M <= 0 || N <= 0) return 0;
    if (K < 0 || !A || !B || !C) return (int)hipErrorInvalidValue;
    const int tiles = ((N + BN - 1) / BN) * ((M + BM - 1) / BM);
    // few tiles and a long K: split K so that about 256 workgroups run, at least 2 chunks each,
    // within the workspace; the partial sums are added in split order by the epilogue kernel
    int ns = 1;
    if (work && tiles < 192 && K >= 4 * BK) {
        ns = std::min({(256 + tiles - 1) / tiles, K / (2 * BK), 16});
        while (ns > 1 && (int64_t)ns * M * N > work_elems) --ns;
    }
    const int kchunk = ns > 1 ? ((K + ns - 1) / ns + BK - 1) / BK * BK : (K > 0 ? K : 1);
    ns = ns > 1 ? (K + kchunk - 1) / kchunk : 1;
    const GemmArgs g{M, N, K, A, sam, sak, B, sbn, sbk, C, ldc, alpha, beta, bias, mask, relu_a, relu_b,
                     work, kchunk};
    hipLaunchKernelGGL(k_gemm_bf16, dim3((N + BN - 1) / BN, (M + BM - 1) / BM, ns), dim3(256), 0,
                       (hipStream_t)stream, g);
    if (ns > 1) {
        const int64_t total = (int64_t)M * N;
        const int blocks = (int)std::min<int64_t>((total + 255) / 256, 2048);
        hipLaunchKernelGGL(k_gemm_splitk_epi, dim3(blocks), dim3(256), 0, (hipStream_t)stream, g, ns);
    }
    return lk_last_error();
}

int lkd_mlp_colsum(const float* G, int M, int N, int64_t ld, float beta, float* out, void* stream) {
    if (N <= 0) return 0;
    LK_LAUNCH(k_colsum, (N + 31) / 32, 256, 0, stream, G, M, N, ld, beta, out);
    return lk_last_error();
}

int lkd_mlp_mse(const float* Y, const float* T, int64_t count, float scale, float* g, float* part,
                int32_t nparts, float* loss, void* stream) {
    if (count <= 0 || nparts < 1) return (int)hipErrorInvalidValue;
    const int nb = blocks_for(count) < nparts ? blocks_for(count) : nparts;
    LK_LAUNCH(k_mse, nb, 256, 0, stream, Y, T, count, scale, g, part);
    LK_LAUNCH(k_sum_parts, 1, 256, 0, stream, part, nb, loss);
    return lk_last_error();
}

int lkd_mlp_gather(const float* src, const int32_t* idx, int rows, int cols, float* dst, void* stream) {
    if (rows <= 0 || cols <= 0) return 0;
    LK_LAUNCH(k_gather, blocks_for((int64_t)rows * cols), 256, 0, stream, src, idx, rows, cols, dst);
    return lk_last_error();
}

int lkd_mlp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                 float b2, float eps, int64_t step, void* stream) {
    if (n <= 0 || step < 1) return (int)hipErrorInvalidValue;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    LK_LAUNCH(k_adam, blocks_for(n), 256, 0, stream, w, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return lk_last_error();
}

int lkd_probe_mfma_bf16(const float* a, const float* b, float* out, void* stream) {
    LK_LAUNCH(k_probe_bf16, 1, 64, 0, stream, a, b, out);
    return lk_last_error();
}

}  // extern "C"

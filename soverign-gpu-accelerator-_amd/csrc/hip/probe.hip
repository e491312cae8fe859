// Small in-bounds probes of compiler builtins and tile routines, used by the GPU unit tests to
// check MFMA operand/result lane layouts and the 64x64 tile GEMM against numpy.
#include "common.hpp"

namespace {

// out[lane * 4 + e] = D element e of `lane` for D = A (16x4) * B (4x16),
// lane l supplies a = A_in[l], b = B_in[l] (the layout under test decides their meaning).
__global__ void k_mfma64(const double* A_in, const double* B_in, double* out) {
    const int l = threadIdx.x;
    v4d acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A_in[l], B_in[l], acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) out[l * 4 + e] = acc[e];
}
__global__ void k_mfma32(const float* A_in, const float* B_in, float* out) {
    const int l = threadIdx.x;
    v4f acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A_in[l], B_in[l], acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) out[l * 4 + e] = acc[e];
}

}  // namespace

extern "C" {

int lkd_probe_mfma(int prec, const void* a, const void* b, void* out, void* stream) {
    if (prec == LKD_F32)
        LK_LAUNCH(k_mfma32, 1, 64, 0, stream, (const float*)a, (const float*)b, (float*)out);
    else
        LK_LAUNCH(k_mfma64, 1, 64, 0, stream, (const double*)a, (const double*)b, (double*)out);
    return lk_last_error();
}

}  // extern "C"

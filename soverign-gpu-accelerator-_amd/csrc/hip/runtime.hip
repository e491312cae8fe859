// HIP runtime plumbing behind the device C ABI (allocation, copies, streams, events, graphs).
#include "common.hpp"

namespace {
__global__ void k_fill(double* __restrict__ p, double v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// Plain streaming copy (16-byte vectors): the reference kernel for HBM traffic counters.
__global__ void k_copy(double2* __restrict__ d, const double2* __restrict__ s, int64_t n2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}
// One thread per (queue, XCD) counter: a queue whose counter stayed below its length had
// tickets no workgroup took, so those tasks never ran.
__global__ void k_queue_check(const uint32_t* __restrict__ cnt, const int32_t* __restrict__ qoff,
                              int nq, int nx, int bit, int32_t* err) {
    for (int i = threadIdx.x; i < nq * nx; i += blockDim.x) {
        const int q = i / nx, x = i % nx;
        const int64_t len = qoff[(int64_t)q * (nx + 1) + x + 1] - qoff[(int64_t)q * (nx + 1) + x];
        if ((int64_t)cnt[i] < len) atomicOr(err, bit);
    }
}
}  // namespace

extern "C" {

const char* lkd_error_string(int code) { return hipGetErrorString((hipError_t)code); }
int lkd_device_count(int* count) { return (int)hipGetDeviceCount(count); }
int lkd_device_cus(int* cus) {
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    return (int)hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
}
int lkd_set_device(int dev) { return (int)hipSetDevice(dev); }
int lkd_wallclock_khz(int* khz) {
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    return (int)hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, dev);
}
int lkd_malloc(void** ptr, size_t bytes) { return (int)hipMalloc(ptr, bytes ? bytes : 1); }
int lkd_free(void* ptr) { return ptr ? (int)hipFree(ptr) : 0; }
int lkd_host_alloc(void** ptr, size_t bytes) {
    return (int)hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
}
int lkd_host_free(void* ptr) { return ptr ? (int)hipHostFree(ptr) : 0; }
int lkd_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
    if (!bytes) return 0;
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
}
int lkd_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
    if (!bytes) return 0;
    int e = (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e) return e;
    return (int)hipStreamSynchronize((hipStream_t)stream);
}
int lkd_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    if (!bytes) return 0;
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
}
int lkd_memset(void* dst, int value, size_t bytes, void* stream) {
    if (!bytes) return 0;
    return (int)hipMemsetAsync(dst, value, bytes, (hipStream_t)stream);
}
int lkd_stream_create(void** stream) {
    hipStream_t s;
    int e = (int)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    *stream = s;
    return e;
}
int lkd_stream_destroy(void* stream) { return stream ? (int)hipStreamDestroy((hipStream_t)stream) : 0; }
int lkd_stream_sync(void* stream) { return (int)hipStreamSynchronize((hipStream_t)stream); }
int lkd_event_create(void** ev) {
    hipEvent_t e;
    int r = (int)hipEventCreate(&e);
    *ev = e;
    return r;
}
int lkd_event_destroy(void* ev) { return ev ? (int)hipEventDestroy((hipEvent_t)ev) : 0; }
int lkd_event_record(void* ev, void* stream) {
    return (int)hipEventRecord((hipEvent_t)ev, (hipStream_t)stream);
}
int lkd_event_elapsed_ms(float* ms, void* start, void* stop) {
    int e = (int)hipEventSynchronize((hipEvent_t)stop);
    if (e) return e;
    return (int)hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
}
int lkd_fill_f64(double* dst, double value, int64_t n, void* stream) {
    if (n <= 0) return 0;
    const int64_t blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    LK_LAUNCH(k_fill, (int)blocks, 256, 0, stream, dst, value, n);
    return lk_last_error();
}
int lkd_copy_f64(double* dst, const double* src, int64_t n, void* stream) {
    if (n <= 0) return 0;
    if ((n & 1) || (((uintptr_t)dst | (uintptr_t)src) & 15)) return (int)hipErrorInvalidValue;
    const int64_t n2 = n / 2;
    const int64_t blocks = (n2 + 255) / 256 < 8192 ? (n2 + 255) / 256 : 8192;
    LK_LAUNCH(k_copy, (int)blocks, 256, 0, stream, (double2*)dst, (const double2*)src, n2);
    return lk_last_error();
}

int lkd_graph_begin(void* stream) {
    return (int)hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
}
int lkd_graph_end(void* stream, void** exec) {
    hipGraph_t g = nullptr;
    int e = (int)hipStreamEndCapture((hipStream_t)stream, &g);
    if (e) return e;
    hipGraphExec_t x = nullptr;
    e = (int)hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    const int e2 = (int)hipGraphDestroy(g);
    if (e) return e;
    *exec = x;
    return e2;
}
int lkd_graph_launch(void* exec, void* stream) {
    return (int)hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream);
}
int lkd_graph_destroy(void* exec) { return exec ? (int)hipGraphExecDestroy((hipGraphExec_t)exec) : 0; }

int lkd_queue_check(const uint32_t* cnt, const int32_t* qoff, int32_t nq, int32_t nx,
                    int32_t bit, int32_t* err, void* stream) {
    if (nq <= 0) return 0;
    if (nx < 1 || !cnt || !qoff || !err) return (int)hipErrorInvalidValue;
    LK_LAUNCH(k_queue_check, 1, 256, 0, stream, cnt, qoff, nq, nx, bit, err);
    return lk_last_error();
}

}  // extern "C"

// Host driver of the warm-start network (include/learn/mlp.hpp): parameters, moments and
// activations in one arena; every product is one lkd_mlp_gemm launch (BF16 MFMA, FP32
// accumulation), everything else small elementwise / reduction kernels.
//
// Shapes (batch B, row-major):
//   forward   z1 = u W1^T + b1 (B x h1); z2 = relu(z1) W2^T + b2 (B x h2); y = relu(z2) W3^T + b3
//   backward  g = 2 (y - t) / (B d_out);  dW3 = g^T relu(z2), db3 = sum_rows g
//             g2 = (g W3) * [z2 > 0];     dW2 = g2^T relu(z1), db2 = sum_rows g2
//             g1 = (g2 W2) * [z1 > 0];    dW1 = g1^T u,        db1 = sum_rows g1
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include "gpu/vram_arena.hpp"
#include "learn/mlp.hpp"
#include "lakshya_device.h"

namespace sankhya {
namespace learn {

namespace {
void check(int e, const char* what) {
    if (e != 0)
        throw std::runtime_error(std::string("HIP call failed (") + what + "): " +
                                 lkd_error_string(e));
}
constexpr int kParts = 1024;   // loss partials
}  // namespace

struct WarmStartMLP::Impl {
    int d_in, h1, h2, d_out, max_batch;
    int64_t np = 0;
    int64_t oW1, ob1, oW2, ob2, oW3, ob3;
    std::unique_ptr<gpu::VRAMArena> arena;
    void* stream = nullptr;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    float *U = nullptr, *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *T = nullptr, *GY = nullptr;
    float *G2 = nullptr, *G1 = nullptr, *part = nullptr, *loss = nullptr;
    float* W = nullptr;                      // split-K partial sums of the products
    int64_t work_elems = 0;
    int32_t* idx = nullptr;
    float *DU = nullptr, *DT = nullptr;      // device-resident training set
    int ndata = 0;
    int64_t step = 0;

    Impl(int di, int a, int b, int dout, int mb, uint64_t seed)
        : d_in(di), h1(a), h2(b), d_out(dout), max_batch(mb) {
        if (di < 1 || a < 1 || b < 1 || dout < 1 || mb < 1)
            throw std::invalid_argument("WarmStartMLP: dimensions and batch must be positive");
        oW1 = 0;
        ob1 = oW1 + int64_t(h1) * d_in;
        oW2 = ob1 + h1;
        ob2 = oW2 + int64_t(h2) * h1;
        oW3 = ob2 + h2;
        ob3 = oW3 + int64_t(d_out) * h2;
        np = ob3 + d_out;
        const size_t act = size_t(mb) * (size_t(d_in) + 2 * h1 + 2 * h2 + 3 * size_t(d_out));
        // split-K workspace: up to 16 splits of the largest product output
        const int64_t big = std::max({int64_t(mb) * h1, int64_t(mb) * h2, int64_t(mb) * dout,
                                      int64_t(dout) * h2, int64_t(h2) * h1, int64_t(h1) * di});
        work_elems = std::min<int64_t>(16 * big, int64_t(1) << 26);
        const size_t bytes = sizeof(float) * (4 * size_t(np) + act + kParts + 64 + size_t(work_elems)) +
                             sizeof(int32_t) * mb + (16u << 10);
        arena.reset(new gpu::VRAMArena(bytes + (1u << 20)));
        check(lkd_stream_create(&stream), "stream");
        auto f = [&](int64_t n) { return static_cast<float*>(arena->allocate(sizeof(float) * std::max<int64_t>(n, 1))); };
        P = f(np); G = f(np); Mo = f(np); Vo = f(np);
        U = f(int64_t(mb) * d_in); Z1 = f(int64_t(mb) * h1); Z2 = f(int64_t(mb) * h2);
        Y = f(int64_t(mb) * d_out); T = f(int64_t(mb) * d_out); GY = f(int64_t(mb) * d_out);
        G2 = f(int64_t(mb) * h2); G1 = f(int64_t(mb) * h1); part = f(kParts); loss = f(16);
        idx = static_cast<int32_t*>(arena->allocate(sizeof(int32_t) * mb));
        W = static_cast<float*>(arena->allocate_uninitialized(sizeof(float) * size_t(work_elems)));
        // He initialization (normal, std sqrt(2 / fan_in)), biases zero
        std::vector<float> p(static_cast<size_t>(np), 0.0f);
        std::mt19937_64 rng(seed);
        auto fill = [&](int64_t off, int64_t rows, int64_t fan_in) {
            std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / double(fan_in)));
            for (int64_t i = 0; i < rows * fan_in; ++i) p[static_cast<size_t>(off + i)] = static_cast<float>(nd(rng));
        };
        fill(oW1, h1, d_in);
        fill(oW2, h2, h1);
        fill(oW3, d_out, h2);
        set_params(p.data());
    }
    ~Impl() {
        if (stream) lkd_stream_sync(stream);
        arena.reset();
        if (stream) lkd_stream_destroy(stream);
    }

    void set_params(const float* p) {
        check(lkd_memcpy_htod(P, p, sizeof(float) * np, stream), "params");
        // fresh optimizer state
        std::vector<float> z(static_cast<size_t>(np), 0.0f);
        check(lkd_memcpy_htod(Mo, z.data(), sizeof(float) * np, stream), "moments");
        check(lkd_memcpy_htod(Vo, z.data(), sizeof(float) * np, stream), "moments");
        check(lkd_stream_sync(stream), "params");
        step = 0;
    }

    void gemm(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
              int64_t sbn, int64_t sbk, float* C, int64_t ldc, const float* bias, const float* mask,
              int relu_a, int relu_b) {
        check(lkd_mlp_gemm(M, N, K, A, sam, sak, B, sbn, sbk, C, ldc, 1.0f, 0.0f, bias, mask, relu_a,
                           relu_b, W, work_elems, stream), "gemm");
    }

    // forward on the batch in U (rows B)
    void forward_dev(int B) {
        gemm(B, h1, d_in, U, d_in, 1, P + oW1, d_in, 1, Z1, h1, P + ob1, nullptr, 0, 0);
        gemm(B, h2, h1, Z1, h1, 1, P + oW2, h1, 1, Z2, h2, P + ob2, nullptr, 1, 0);
        gemm(B, d_out, h2, Z2, h2, 1, P + oW3, h2, 1, Y, d_out, P + ob3, nullptr, 1, 0);
    }

    // loss and gradients for the batch in U, T (rows B) into G; returns the loss
    void backward_dev(int B) {
        const int64_t cnt = int64_t(B) * d_out;
        check(lkd_mlp_mse(Y, T, cnt, 2.0f / float(cnt), GY, part, kParts, loss, stream), "loss");
        // layer 3
        gemm(d_out, h2, B, GY, 1, d_out, Z2, 1, h2, G + oW3, h2, nullptr, nullptr, 0, 1);
        check(lkd_mlp_colsum(GY, B, d_out, d_out, 0.0f, G + ob3, stream), "colsum");
        gemm(B, h2, d_out, GY, d_out, 1, P + oW3, 1, h2, G2, h2, nullptr, Z2, 0, 0);
        // layer 2
        gemm(h2, h1, B, G2, 1, h2, Z1, 1, h1, G + oW2, h1, nullptr, nullptr, 0, 1);
        check(lkd_mlp_colsum(G2, B, h2, h2, 0.0f, G + ob2, stream), "colsum");
        gemm(B, h1, h2, G2, h2, 1, P + oW2, 1, h1, G1, h1, nullptr, Z1, 0, 0);
        // layer 1
        gemm(h1, d_in, B, G1, 1, h1, U, 1, d_in, G + oW1, d_in, nullptr, nullptr, 0, 0);
        check(lkd_mlp_colsum(G1, B, h1, h1, 0.0f, G + ob1, stream), "colsum");
    }

    void check_batch(int B) const {
        if (B < 1 || B > max_batch)
            throw std::invalid_argument("WarmStartMLP: batch " + std::to_string(B) + " outside [1, " +
                                        std::to_string(max_batch) + "]");
    }

    void forward(const float* u, int B, float* y) {
        check_batch(B);
        check(lkd_memcpy_htod(U, u, sizeof(float) * int64_t(B) * d_in, stream), "input");
        forward_dev(B);
        check(lkd_memcpy_dtoh(y, Y, sizeof(float) * int64_t(B) * d_out, stream), "output");
    }

    double gradients(const float* u, const float* t, int B, float* grad) {
        check_batch(B);
        check(lkd_memcpy_htod(U, u, sizeof(float) * int64_t(B) * d_in, stream), "input");
        check(lkd_memcpy_htod(T, t, sizeof(float) * int64_t(B) * d_out, stream), "target");
        forward_dev(B);
        backward_dev(B);
        check(lkd_memcpy_dtoh(grad, G, sizeof(float) * np, stream), "gradient");
        float l = 0;
        check(lkd_memcpy_dtoh(&l, loss, sizeof(float), stream), "loss");
        return double(l) / (double(B) * d_out);
    }

    void set_dataset(const float* u, const float* t, int n) {
        if (n < 1) throw std::invalid_argument("WarmStartMLP: empty data set");
        DU = DT = nullptr;
        // the data set lives in a slab of its own, replaced by the next set
        data_arena.reset(new gpu::VRAMArena(sizeof(float) * (size_t(n) * (d_in + d_out)) + (2u << 20)));
        DU = static_cast<float*>(data_arena->allocate_uninitialized(sizeof(float) * size_t(n) * d_in));
        DT = static_cast<float*>(data_arena->allocate_uninitialized(sizeof(float) * size_t(n) * d_out));
        check(lkd_memcpy_htod(DU, u, sizeof(float) * size_t(n) * d_in, stream), "data");
        check(lkd_memcpy_htod(DT, t, sizeof(float) * size_t(n) * d_out, stream), "data");
        check(lkd_stream_sync(stream), "data");
        ndata = n;
    }
    std::unique_ptr<gpu::VRAMArena> data_arena;

    std::vector<double> train(int epochs, int B, float lr, uint64_t seed) {
        if (!DU || !DT) throw std::logic_error("WarmStartMLP::train: no data set");
        check_batch(B);
        std::mt19937_64 rng(seed);
        std::vector<int32_t> perm(static_cast<size_t>(ndata));
        std::iota(perm.begin(), perm.end(), 0);
        std::vector<double> out;
        for (int e = 0; e < epochs; ++e) {
            std::shuffle(perm.begin(), perm.end(), rng);
            double sum = 0.0;
            int batches = 0;
            for (int b0 = 0; b0 + B <= ndata; b0 += B) {   // full batches only
                check(lkd_memcpy_htod(idx, perm.data() + b0, sizeof(int32_t) * B, stream), "batch");
                check(lkd_mlp_gather(DU, idx, B, d_in, U, stream), "gather");
                check(lkd_mlp_gather(DT, idx, B, d_out, T, stream), "gather");
                forward_dev(B);
                backward_dev(B);
                ++step;
                check(lkd_mlp_adam(P, G, Mo, Vo, np, lr, 0.9f, 0.999f, 1e-8f, step, stream), "adam");
                float l = 0;
                check(lkd_memcpy_dtoh(&l, loss, sizeof(float), stream), "loss");
                sum += double(l) / (double(B) * d_out);
                ++batches;
            }
            out.push_back(batches ? sum / batches : 0.0);
        }
        return out;
    }
};

WarmStartMLP::WarmStartMLP(int d_in, int h1, int h2, int d_out, int max_batch, uint64_t seed)
    : impl_(new Impl(d_in, h1, h2, d_out, max_batch, seed)) {}
WarmStartMLP::~WarmStartMLP() = default;
int64_t WarmStartMLP::param_count() const { return impl_->np; }
void WarmStartMLP::set_params(const float* p) { impl_->set_params(p); }
void WarmStartMLP::get_params(float* p) const {
    check(lkd_memcpy_dtoh(p, impl_->P, sizeof(float) * impl_->np, impl_->stream), "params");
}
void WarmStartMLP::forward(const float* u, int batch, float* y) { impl_->forward(u, batch, y); }
double WarmStartMLP::gradients(const float* u, const float* t, int batch, float* grad) {
    return impl_->gradients(u, t, batch, grad);
}
void WarmStartMLP::set_dataset(const float* u, const float* t, int n) { impl_->set_dataset(u, t, n); }
std::vector<double> WarmStartMLP::train(int epochs, int batch, float lr, uint64_t seed) {
    return impl_->train(epochs, batch, lr, seed);
}

}  // namespace learn
}  // namespace sankhya

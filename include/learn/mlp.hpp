// Warm-start network for the interior-point method: a three-layer perceptron
//   z1 = W1 u + b1, z2 = W2 relu(z1) + b2, y = W3 relu(z2) + b3
// mapping the compressed (b, c) perturbation u of an LP of a parametric family (fixed A) to a
// normalized prediction y of its optimal point (log x, y, log s).  Forward, backward and the
// Adam update run on the device: BF16 MFMA products with FP32 accumulation and FP32 master
// weights.  Extension with no reference counterpart: the start it replaces is the standard
// x = s = 1, y = 0 (src/ipm/mehrotra_device.cu:93-102).
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

namespace sankhya {
namespace learn {

class WarmStartMLP {
public:
    // Weights He-initialized from `seed` (normal, std sqrt(2 / fan_in)), biases zero.
    // max_batch bounds the rows of every forward / training call.
    WarmStartMLP(int d_in, int h1, int h2, int d_out, int max_batch, uint64_t seed);
    ~WarmStartMLP();
    WarmStartMLP(const WarmStartMLP&) = delete;
    WarmStartMLP& operator=(const WarmStartMLP&) = delete;

    // parameters, concatenated: W1 (h1 x d_in, row-major), b1, W2 (h2 x h1), b2, W3 (d_out x h2), b3
    int64_t param_count() const;
    void set_params(const float* p);
    void get_params(float* p) const;

    // y (batch x d_out) for inputs u (batch x d_in), host arrays, row-major
    void forward(const float* u, int batch, float* y);
    // Gradient of L = (1 / (batch d_out)) sum (y - t)^2 with respect to the parameters (same
    // layout as the parameters); returns L.  No update.
    double gradients(const float* u, const float* t, int batch, float* grad);

    // Training set resident on the device (n rows of inputs and targets).
    void set_dataset(const float* u, const float* t, int n);
    // Adam (lr, beta1 0.9, beta2 0.999, eps 1e-8) over `epochs` passes of shuffled mini-batches
    // (shuffle seeded by `seed`); returns the mean training loss of every epoch.
    std::vector<double> train(int epochs, int batch, float lr, uint64_t seed);

    struct Impl;

private:
    std::unique_ptr<Impl> impl_;
};

}  // namespace learn
}  // namespace sankhya

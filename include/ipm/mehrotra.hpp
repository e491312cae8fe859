// Mehrotra predictor-corrector interior-point method for  min c^T x, A x = b, x >= 0  on one
// MI355X.
//
// Reference interface (kept): MehrotraResult, MehrotraSolver(const core::Model&), solve()
// -- src/ipm/mehrotra.hpp:10-31.  The solver keeps a reference to the model, which must outlive
// it.  Default options reproduce the reference algorithm and stop rule
// (src/ipm/mehrotra_device.cu:695-1047): Jacobi Ruiz x10, start x = s = 1, y = 0, stop when the
// scaled ||rp||_2, ||rd||_2 and mu are all < 1e-6, sigma = min(1, (mu_aff/mu)^3), eta = 0.995,
// merit backtracking (Armijo 1e-4, halving, 15 trials) with a centred fallback (sigma_c = 0.5,
// 30 trials), 200 iterations; numerical failures end in SimplexStatus::IterationLimit.
//
// Extensions (IpmOptions): a 1e-8 relative-gap stop rule evaluated in the original space,
// FP32-factor mixed precision with FP64 refinement and an FP64 fallback, a warm start and the
// ordering choice.
#pragma once
#include <vector>
#include "core/problem.hpp"
#include "sankhya/types.hpp"
#include "simplex/primal.hpp"

namespace sankhya {
namespace ipm {

struct MehrotraResult {
    simplex::SimplexStatus status;
    Float objective_value;      // c^T x (= scaled c_s^T x_s)
    Index iterations;
    Float primal_residual;      // scaled ||A x - b||_2 at the last evaluated iterate
    Float dual_residual;        // scaled ||A^T y + s - c||_2
    Float duality_gap;          // mu = x^T s / n (reference meaning)
    std::vector<Float> x;       // original space
    std::vector<Float> pi;      // original space

    // ---- extensions ----
    std::vector<Float> s;                 // original-space reduced costs
    Float dual_objective = 0;             // b^T y
    Float rel_gap = 0;                    // |c^T x - b^T y| / max(1, |c^T x|, |b^T y|)
    Float rel_primal_residual = 0;        // ||A x - b||_inf / (1 + ||b||_inf), original space
    Float rel_dual_residual = 0;          // ||A^T y + s - c||_inf / (1 + ||c||_inf)
    int factorizations = 0, fp32_factorizations = 0, fp64_factorizations = 0;
    int solves = 0, refinement_steps = 0;
    int fail_code = 0;                    // 0; 1 factorization, 2 predictor solve, 3 predictor
                                          // check, 4 corrector solve, 5 no step accepted
    double setup_ms = 0, solve_ms = 0, factor_ms = 0;
    double nnz_L = 0, flops_per_factorization = 0;
};

struct IpmOptions {
    enum class StopRule { Reference = 0, RelativeGap = 1 };
    enum class Precision { FP64 = 0, Mixed = 1 };
    enum class Ordering { Amd = 0, ReferenceMinimumDegree = 1 };
    StopRule stop_rule = StopRule::Reference;
    Float tol = 1e-6;                 // reference rule threshold
    Float gap_tol = 1e-8;             // relative rule threshold (gap and both residuals)
    Index max_iter = 200;
    Precision precision = Precision::FP64;
    Float mixed_fp64_spread = 1e12;   // max(Theta)/min(Theta) beyond which FP64 factors are used
    int mixed_max_refine = 60;        // PCG budget with an FP32 factor before the FP64 retry
    Ordering ordering = Ordering::Amd;
    bool use_graphs = false;          // launch-graph capture: not supported (true is rejected)
    bool verbose = false;
    const Float* warm_x = nullptr;    // optional original-space warm start (all three or none)
    const Float* warm_y = nullptr;
    const Float* warm_s = nullptr;
    Float warm_floor = 1e-3;          // lower clamp of warm-start x and s (scaled space)
};

// The solver's equilibration, run once in its constructor: 10 Jacobi Ruiz passes
// (src/ipm/mehrotra_device.cu:286-323).  Each pass takes every row and column max-abs from the
// same snapshot of A, then scales row i by 1/sqrt(rmax_i) and column j by 1/sqrt(cmax_j) (a
// factor 1 where the max is <= 1e-8); b and c follow their rows / columns.  On return A holds
// R A C, b holds R b, c holds C c, and R (m) and C (n) the accumulated factors.
void ruiz_equilibrate(core::CSRMatrix& A, std::vector<Float>& b, std::vector<Float>& c,
                      std::vector<Float>& R, std::vector<Float>& C);

class MehrotraSolver {
public:
    MehrotraSolver(const core::Model& model);
    MehrotraSolver(const core::Model& model, const IpmOptions& options);
    ~MehrotraSolver();

    MehrotraResult solve();

    // Extension: replace c and b (same A, same dimensions) for the next solve(); the Ruiz
    // scaling, ordering and symbolic analysis of A are kept.
    void update_objective_and_rhs(const std::vector<Float>& c, const std::vector<Float>& b);
    // Extension: set / clear an original-space warm start for the next solve().
    void set_warm_start(const std::vector<Float>& x, const std::vector<Float>& y,
                        const std::vector<Float>& s);
    void clear_warm_start();
    // Extension: factorization profile at the last iterate (see GPUKKTCholeskySolver) and the
    // symbolic statistics of the normal equations.
    void profile_factor(bool fp32, double ms[6], double flops[6], double launches[6]);
    void set_launch_timing(bool on);
    void launch_timing(double ms[6], double flops[6], double launches[6]);
    void kkt_statistics(double out[12]);

    class Impl;

private:
    Impl* impl_;
};

}  // namespace ipm
}  // namespace sankhya

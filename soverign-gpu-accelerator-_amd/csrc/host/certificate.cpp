// FP64 optimality certificate (src/verifier/certificate.cpp:8-155, reference): dimension
// checks, bound feasibility, ||Ax-b||_inf < tol, reduced-cost sign against the finite bounds and
// complementarity < tol.  The reference hard-codes tol = 1e-6; here it is the trailing argument.
#include <cmath>
#include <cstdio>
#include "verifier/certificate.hpp"

namespace sankhya {
namespace verifier {

bool verify_optimal(const core::Model& model, const std::vector<Float>& x,
                    const std::vector<Float>& pi, Float* out_primal_residual, Float tol) {
    const auto& A = model.A;
    if (x.size() != static_cast<size_t>(A.cols) || pi.size() != static_cast<size_t>(A.rows))
        return false;
    if (model.obj.size() != x.size() || model.lb.size() != x.size() ||
        model.ub.size() != x.size() || A.col_ptrs.size() != x.size() + 1 ||
        A.col_ptrs.back() < 0 ||
        A.row_indices.size() < static_cast<size_t>(A.col_ptrs.back()) ||
        A.values.size() < static_cast<size_t>(A.col_ptrs.back()) ||
        model.rhs.size() != pi.size())
        return false;

    const double sign = model.sense == OptimizationSense::Maximize ? -1.0 : 1.0;
    const double inf = math::kInfinity;
    const size_t n = x.size(), m = pi.size();

    for (size_t j = 0; j < n; ++j) {
        if (model.lb[j] > -inf && x[j] <= model.lb[j] - tol) return false;
        if (model.ub[j] < inf && x[j] >= model.ub[j] + tol) return false;
    }
    std::vector<double> ax(m, 0.0);
    for (size_t j = 0; j < n; ++j)
        for (Index k = A.col_ptrs[j]; k < A.col_ptrs[j + 1]; ++k) {
            const Index i = A.row_indices[static_cast<size_t>(k)];
            if (i < 0 || static_cast<size_t>(i) >= m) return false;
            ax[static_cast<size_t>(i)] += A.values[static_cast<size_t>(k)] * x[j];
        }
    double worst = 0.0;
    for (size_t i = 0; i < m; ++i) {
        const double r = std::abs(ax[i] - model.rhs[i]);
        if (r > worst) worst = r;
        if (r >= tol) return false;
    }
    if (out_primal_residual) *out_primal_residual = worst;

    for (size_t j = 0; j < n; ++j) {
        double atpi = 0.0;
        for (Index k = A.col_ptrs[j]; k < A.col_ptrs[j + 1]; ++k) {
            const Index i = A.row_indices[static_cast<size_t>(k)];
            if (i < 0 || static_cast<size_t>(i) >= m) return false;
            atpi += A.values[static_cast<size_t>(k)] * pi[static_cast<size_t>(i)];
        }
        const double r = sign * model.obj[j] - atpi;
        if (r >= tol) {                       // must sit at a finite lower bound
            if (model.lb[j] <= -inf) return false;
            if (r * (x[j] - model.lb[j]) >= tol) return false;
        } else if (r <= -tol) {               // must sit at a finite upper bound
            if (model.ub[j] >= inf) return false;
            if ((-r) * (model.ub[j] - x[j]) >= tol) return false;
        }
    }
    return true;
}

bool verify_optimal(const core::Model& model, const std::vector<Float>& x,
                    const std::vector<Float>& pi, Float* out_primal_residual) {
    return verify_optimal(model, x, pi, out_primal_residual, 1e-6);
}

}  // namespace verifier
}  // namespace sankhya

// solve_with_phase1 on a core::Model through the device simplex (k_simplex via
// simplex::solve_batch): the model handling of src/simplex/phase1.cpp:124-170 (bounds policy,
// lower-bound shift) and primal.cpp:39-42 (objective sense), then one LP on the device.
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "lakshya_device.h"
#include "simplex/primal.hpp"

namespace sankhya {
namespace simplex {

static_assert(kDeviceSimplexMaxRows == LKB_MAX_M, "device simplex row limit");

SimplexStatus solve_with_phase1(const core::Model& model, Basis& basis, std::vector<Float>& x,
                                std::vector<Float>* pi) {
    const core::CSCMatrix& A = model.A;
    const Index m = A.rows, n = A.cols;
    if (static_cast<Index>(model.rhs.size()) != m || static_cast<Index>(model.obj.size()) != n)
        throw std::invalid_argument("solve_with_phase1: rhs / obj sizes do not match A");
    if (m < 1 || m > LKB_MAX_M)
        throw std::invalid_argument("solve_with_phase1: " + std::to_string(m) +
                                    " rows; the device simplex takes 1 to " + std::to_string(LKB_MAX_M));
    const bool has_lb = static_cast<Index>(model.lb.size()) == n;
    const bool has_ub = static_cast<Index>(model.ub.size()) == n;
    std::vector<Float> lb(n, 0.0);
    for (Index j = 0; j < n; ++j) {
        if (has_lb && std::isfinite(model.lb[j])) lb[j] = model.lb[j];
        const Float ub = has_ub ? model.ub[j] : math::kInfinity;
        const Float shifted_ub = ub - lb[j];
        if (shifted_ub < math::kInfinity * 0.5 && shifted_ub > math::kDefaultFeasibilityTol)
            return SimplexStatus::IterationLimit;
    }
    std::vector<double> b(model.rhs.begin(), model.rhs.end());
    for (Index j = 0; j < n; ++j)
        if (lb[j] != 0.0)
            for (Index k = A.col_ptrs[j]; k < A.col_ptrs[j + 1]; ++k) b[A.row_indices[k]] -= A.values[k] * lb[j];
    const double sign = model.sense == OptimizationSense::Minimize ? 1.0 : -1.0;
    std::vector<double> c(n);
    for (Index j = 0; j < n; ++j) c[j] = sign * model.obj[j];
    std::vector<int32_t> bas(m);
    std::vector<double> xs(n), y(m);
    int32_t status = 0, iters[2] = {0, 0};
    double obj = 0.0;
    const int32_t mm = m, nn = n;
    solve_batch(1, &mm, &nn, A.col_ptrs.data(), A.row_indices.data(), A.values.data(), b.data(),
                c.data(), LKS_MAX_ITER, bas.data(), xs.data(), y.data(), &status, iters, &obj);
    basis.basic_indices.assign(m, 0);
    basis.col_status.assign(n, BasisStatus::AtLower);
    for (Index i = 0; i < m; ++i) {
        basis.basic_indices[i] = bas[i] >= 0 ? bas[i] : n + i;
        if (bas[i] >= 0) basis.col_status[bas[i]] = BasisStatus::Basic;
    }
    x.assign(n, 0.0);
    for (Index j = 0; j < n; ++j) x[j] = xs[j] + lb[j];
    if (pi) pi->assign(y.begin(), y.end());
    switch (status) {
        case 0: return SimplexStatus::Optimal;
        case 1: return SimplexStatus::Infeasible;
        case 2: return SimplexStatus::Unbounded;
        default: return SimplexStatus::IterationLimit;
    }
}

}  // namespace simplex
}  // namespace sankhya

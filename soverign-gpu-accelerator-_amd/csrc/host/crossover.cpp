// generate_crossover_basis on a core::Model or a gpu::DeviceModel: the device crossover of one LP
// (k_crossover through crossover_points) and the reference's Basis result and failure messages
// (src/ipm/crossover.cu:416-459).
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "ipm/crossover.hpp"
#include "lakshya_device.h"

namespace sankhya {
namespace ipm {

namespace {

simplex::Basis basis_of(Index m, Index n, const Index* colptr, const Index* rowidx, const Float* vals,
                        const Float* rhs, const Float* obj, const Float* ub, const Float* x) {
    if (m < 1 || m > LKB_MAX_M)
        throw std::invalid_argument("generate_crossover_basis: " + std::to_string(m) +
                                    " rows; the device crossover takes 1 to " +
                                    std::to_string(LKB_MAX_M));
    std::vector<int32_t> basis(m);
    std::vector<double> xv(n);
    int32_t status = 0, pivots = 0;
    const int32_t mm = m, nn = n;
    crossover_points(1, &mm, &nn, colptr, rowidx, vals, rhs, obj, x, basis.data(), xv.data(),
                     &status, &pivots);
    if (status == kCrossoverUnbounded)
        throw std::runtime_error("Crossover failed: Invalid pivot encountered");
    simplex::Basis out;
    out.basic_indices.resize(m);
    out.col_status.assign(n, simplex::BasisStatus::AtLower);
    bool rank_deficient = false;
    for (Index i = 0; i < m; ++i) {
        if (basis[i] >= 0) {
            out.basic_indices[i] = basis[i];
            out.col_status[basis[i]] = simplex::BasisStatus::Basic;
        } else {
            out.basic_indices[i] = n + i;
            rank_deficient = true;
        }
    }
    for (Index j = 0; j < n; ++j)
        if (out.col_status[j] != simplex::BasisStatus::Basic && ub && std::isfinite(ub[j]) &&
            std::abs(xv[j] - ub[j]) < 1e-6)
            out.col_status[j] = simplex::BasisStatus::AtUpper;
    if (rank_deficient) throw std::runtime_error("Crossover failed: Matrix is rank deficient");
    return out;
}

template <typename T>
std::vector<T> download(const T* src, std::size_t count) {
    std::vector<T> h(count);
    if (count == 0) return h;
    if (!src) throw std::invalid_argument("generate_crossover_basis: the device model lacks an array");
    const int e = lkd_memcpy_dtoh(h.data(), src, sizeof(T) * count, nullptr);
    if (e) throw std::runtime_error(std::string("generate_crossover_basis: ") + lkd_error_string(e));
    return h;
}

void check_x(std::size_t size, Index n) {
    if (static_cast<Index>(size) != n)
        throw std::invalid_argument("generate_crossover_basis: x has " + std::to_string(size) +
                                    " entries for " + std::to_string(n) + " columns");
}

}  // namespace

simplex::Basis generate_crossover_basis(const core::Model& model, const std::vector<Float>& x_ipm) {
    const core::CSCMatrix& A = model.A;
    const Index m = A.rows, n = A.cols;
    check_x(x_ipm.size(), n);
    if (static_cast<Index>(model.rhs.size()) != m || static_cast<Index>(model.obj.size()) != n)
        throw std::invalid_argument("generate_crossover_basis: rhs / obj sizes do not match A");
    const Float* ub = static_cast<Index>(model.ub.size()) == n ? model.ub.data() : nullptr;
    return basis_of(m, n, A.col_ptrs.data(), A.row_indices.data(), A.values.data(), model.rhs.data(),
                    model.obj.data(), ub, x_ipm.data());
}

simplex::Basis generate_crossover_basis(const gpu::DeviceModel& model, const std::vector<Float>& x_ipm) {
    const Index m = model.rows, n = model.cols;
    check_x(x_ipm.size(), n);
    if (m < 1 || m > LKB_MAX_M) return basis_of(m, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    const auto cp = download(model.col_ptrs, static_cast<std::size_t>(n) + 1);
    const auto ri = download(model.row_indices, static_cast<std::size_t>(model.nnz));
    const auto v = download(model.values, static_cast<std::size_t>(model.nnz));
    const auto rhs = download(model.rhs, static_cast<std::size_t>(m));
    const auto obj = download(model.obj, static_cast<std::size_t>(n));
    const auto ub = download(model.ub, static_cast<std::size_t>(n));
    return basis_of(m, n, cp.data(), ri.data(), v.data(), rhs.data(), obj.data(), ub.data(), x_ipm.data());
}

}  // namespace ipm
}  // namespace sankhya

// LP model and sparse-matrix containers.
// Behaviour follows src/core/problem.cpp:84-132 and src/core/sparse_matrix.cpp:49-85 (reference):
// finalize() buckets rows into columns in row order; to_csr() is a stable counting sort.
#include <algorithm>
#include <stdexcept>
#include <string>
#include <utility>
#include "core/problem.hpp"
#include "core/sparse_matrix.hpp"

namespace sankhya {
namespace core {

CSCMatrix::CSCMatrix(Index r, Index c, std::vector<Float> vals, std::vector<Index> rows_idx,
                     std::vector<Index> ptrs)
    : rows(r), cols(c), col_ptrs(std::move(ptrs)), row_indices(std::move(rows_idx)),
      values(std::move(vals)) {}

Index CSCMatrix::nnz() const { return static_cast<Index>(values.size()); }

Float CSCMatrix::get(Index r, Index c) const {
    if (r < 0 || c < 0 || r >= rows || c >= cols) return 0.0;
    auto first = row_indices.begin() + col_ptrs[c];
    auto last = row_indices.begin() + col_ptrs[c + 1];
    auto it = std::lower_bound(first, last, r);
    return (it != last && *it == r) ? values[static_cast<size_t>(it - row_indices.begin())] : 0.0;
}

CSRMatrix to_csr(const CSCMatrix& csc) {
    CSRMatrix out;
    out.rows = csc.rows;
    out.cols = csc.cols;
    const size_t nz = static_cast<size_t>(csc.nnz());
    out.row_ptrs.assign(static_cast<size_t>(csc.rows) + 1, 0);
    for (size_t k = 0; k < nz; ++k) out.row_ptrs[static_cast<size_t>(csc.row_indices[k]) + 1]++;
    for (Index i = 0; i < csc.rows; ++i) out.row_ptrs[i + 1] += out.row_ptrs[i];
    out.col_indices.resize(nz);
    out.values.resize(nz);
    std::vector<Index> next(out.row_ptrs.begin(), out.row_ptrs.end() - 1);
    for (Index c = 0; c < csc.cols; ++c) {
        for (Index k = csc.col_ptrs[c]; k < csc.col_ptrs[c + 1]; ++k) {
            const Index dst = next[static_cast<size_t>(csc.row_indices[k])]++;
            out.col_indices[static_cast<size_t>(dst)] = c;
            out.values[static_cast<size_t>(dst)] = csc.values[static_cast<size_t>(k)];
        }
    }
    return out;
}

Model::Model() : sense(OptimizationSense::Minimize), A(0, 0, {}, {}, {0}) {}

void Model::add_variable(Float obj_coeff, Float lower, Float upper, VariableType vt) {
    obj.push_back(obj_coeff);
    lb.push_back(lower);
    ub.push_back(upper);
    vtype.push_back(vt);
}

static void check_row(const std::vector<Index>& cols, const std::vector<Float>& vals,
                      Index nvars, const char* who) {
    if (cols.size() != vals.size())
        throw std::invalid_argument(std::string(who) + ": " + std::to_string(cols.size()) + " column indices but " + std::to_string(vals.size()) + " values");
    for (Index c : cols)
        if (c < 0 || c >= nvars)
            throw std::invalid_argument(std::string(who) + ": column " + std::to_string(c) + " is not a variable");
}

void Model::add_constraint(const std::vector<Index>& cols, const std::vector<Float>& vals,
                           Float b) {
    check_row(cols, vals, static_cast<Index>(vtype.size()), "add_constraint");
    constraint_cols_.push_back(cols);
    constraint_vals_.push_back(vals);
    rhs.push_back(b);
}

void Model::update_constraint(Index row, const std::vector<Index>& cols,
                              const std::vector<Float>& vals, Float b) {
    if (row < 0 || static_cast<size_t>(row) >= rhs.size())
        throw std::out_of_range("Model::update_constraint: no row " + std::to_string(row));
    check_row(cols, vals, static_cast<Index>(vtype.size()), "update_constraint");
    constraint_cols_[static_cast<size_t>(row)] = cols;
    constraint_vals_[static_cast<size_t>(row)] = vals;
    rhs[static_cast<size_t>(row)] = b;
}

const std::vector<Index>& Model::get_constraint_cols(Index row) const {
    if (row < 0 || static_cast<size_t>(row) >= constraint_cols_.size())
        throw std::out_of_range("Model::get_constraint_cols: no row " + std::to_string(row));
    return constraint_cols_[static_cast<size_t>(row)];
}

const std::vector<Float>& Model::get_constraint_vals(Index row) const {
    if (row < 0 || static_cast<size_t>(row) >= constraint_vals_.size())
        throw std::out_of_range("Model::get_constraint_vals: no row " + std::to_string(row));
    return constraint_vals_[static_cast<size_t>(row)];
}

void Model::finalize() {
    const Index m = static_cast<Index>(rhs.size());
    const Index n = static_cast<Index>(vtype.size());
    if (obj.size() != vtype.size() || lb.size() != vtype.size() || ub.size() != vtype.size())
        throw std::invalid_argument("finalize: objective/bounds/type vectors differ in size");
    std::vector<Index> ptr(static_cast<size_t>(n) + 1, 0);
    for (Index r = 0; r < m; ++r)
        for (Index c : constraint_cols_[static_cast<size_t>(r)]) ptr[static_cast<size_t>(c) + 1]++;
    for (Index c = 0; c < n; ++c) ptr[c + 1] += ptr[c];
    std::vector<Index> ri(static_cast<size_t>(ptr[n]));
    std::vector<Float> va(static_cast<size_t>(ptr[n]));
    std::vector<Index> next(ptr.begin(), ptr.end() - 1);
    for (Index r = 0; r < m; ++r) {
        const auto& cr = constraint_cols_[static_cast<size_t>(r)];
        const auto& vr = constraint_vals_[static_cast<size_t>(r)];
        for (size_t e = 0; e < cr.size(); ++e) {
            const Index dst = next[static_cast<size_t>(cr[e])]++;
            ri[static_cast<size_t>(dst)] = r;
            va[static_cast<size_t>(dst)] = vr[e];
        }
    }
    A = CSCMatrix(m, n, std::move(va), std::move(ri), std::move(ptr));
}

void Model::set_from_csc(Index m, Index n, const Index* col_ptrs, const Index* row_idx,
                         const Float* vals, const Float* c, const Float* b, const Float* lower,
                         const Float* upper) {
    if (m < 0 || n < 0) throw std::invalid_argument("set_from_csc: negative dimension");
    const Index nz = col_ptrs[n];
    std::vector<Index> ptr(col_ptrs, col_ptrs + n + 1);
    std::vector<Index> ri(row_idx, row_idx + nz);
    std::vector<Float> va(vals, vals + nz);
    for (Index k = 0; k < nz; ++k)
        if (ri[k] < 0 || ri[k] >= m) throw std::invalid_argument("set_from_csc: row out of range");
    sense = OptimizationSense::Minimize;
    obj.assign(c, c + n);
    rhs.assign(b, b + m);
    lb.assign(static_cast<size_t>(n), 0.0);
    ub.assign(static_cast<size_t>(n), math::kInfinity);
    if (lower) lb.assign(lower, lower + n);
    if (upper) ub.assign(upper, upper + n);
    vtype.assign(static_cast<size_t>(n), VariableType::Continuous);
    constraint_cols_.assign(static_cast<size_t>(m), {});
    constraint_vals_.assign(static_cast<size_t>(m), {});
    for (Index j = 0; j < n; ++j)
        for (Index k = ptr[j]; k < ptr[j + 1]; ++k) {
            constraint_cols_[static_cast<size_t>(ri[k])].push_back(j);
            constraint_vals_[static_cast<size_t>(ri[k])].push_back(va[k]);
        }
    A = CSCMatrix(m, n, std::move(va), std::move(ri), std::move(ptr));
}

}  // namespace core
}  // namespace sankhya

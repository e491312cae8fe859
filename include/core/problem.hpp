// LP model in equality form: min/max c^T x  s.t.  A x = b,  lb <= x <= ub.
// Interface and member layout of src/core/problem.hpp:26-119 (reference), so reference code
// built against this header (the MILP / McCormick units included) sees the same class.  The
// bilinear-term list is carried but nothing on the interior-point path reads it.
#pragma once
#include <array>
#include <stdexcept>
#include <vector>
#include "sankhya/types.hpp"
#include "core/sparse_matrix.hpp"

namespace sankhya {
namespace core {

// w = x * y with its four McCormick envelope rows (src/core/problem.hpp:26-31)
struct BilinearTerm {
    Index x_col;
    Index y_col;
    Index w_col;
    std::array<Index, 4> mccormick_rows;
};

class Model {
public:
    OptimizationSense sense;
    std::vector<Float> obj;
    std::vector<Float> lb;
    std::vector<Float> ub;
    std::vector<VariableType> vtype;
    std::vector<Float> rhs;
    CSCMatrix A;
    std::vector<BilinearTerm> bilinear_terms;
    std::vector<Float> row_scale;   // optional presolve scaling (empty = unscaled)
    std::vector<Float> col_scale;

    Model();

    void add_variable(Float obj_coeff, Float lower = 0.0, Float upper = math::kInfinity,
                      VariableType vt = VariableType::Continuous);
    // Appends row  sum_e vals[e] * x[cols[e]] = b.  Throws std::invalid_argument on bad input.
    void add_constraint(const std::vector<Index>& cols, const std::vector<Float>& vals, Float b);
    void update_constraint(Index row, const std::vector<Index>& cols,
                           const std::vector<Float>& vals, Float b);
    const std::vector<Index>& get_constraint_cols(Index row) const;
    const std::vector<Float>& get_constraint_vals(Index row) const;
    // Builds the CSC matrix A from the accumulated rows (entries keep row order per column).
    void finalize();

    // Direct construction from CSC arrays (used by the C ABI; equivalent to add_* + finalize()).
    void set_from_csc(Index m, Index n, const Index* col_ptrs, const Index* row_idx,
                      const Float* vals, const Float* c, const Float* b,
                      const Float* lower, const Float* upper);

private:
    std::vector<std::vector<Index>> constraint_cols_;
    std::vector<std::vector<Float>> constraint_vals_;
};

}  // namespace core
}  // namespace sankhya

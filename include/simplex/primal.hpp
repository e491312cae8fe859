// The reference's CPU simplex route (src/simplex/primal.hpp:26-31, src/simplex/phase1.cpp:124-449,
// src/simplex/primal.cpp:27-423) as a device route: Phase I from the artificial basis, then the
// primal simplex with Devex pricing, Harris two-pass ratio test and Bland's rule after 50
// consecutive degenerate steps, at most LKS_MAX_ITER (10,000) iterations per phase, on a dense
// explicit basis inverse (one workgroup per LP, k_simplex).  Limit: m <= 1,600 rows per LP.
#pragma once

#include <cstdint>
#include <vector>

#include "core/problem.hpp"
#include "sankhya/types.hpp"
#include "simplex/basis.hpp"

namespace sankhya {
namespace simplex {

// Same enumerators as src/simplex/primal.hpp:26-31 (reference); also the IPM's status type.
enum class SimplexStatus { Optimal, Infeasible, Unbounded, IterationLimit };

// Rows per LP the device route takes (its basis inverse and row arrays; LKB_MAX_M).
constexpr Index kDeviceSimplexMaxRows = 1600;

// count LPs  min c^T x, A x = b, x >= 0  in the concatenated CSC layout of
// ipm::BatchedMehrotraSolver.  Outputs (any may be null): basis (sum m; the basic column of each
// row, -1 for a row left artificial), x (sum n; nonbasic 0), y (sum m; B^-T c_B of the final
// basis), status (count; SimplexStatus values), iterations (2 count; Phase I, Phase II pivots),
// objective (count; c^T x).  max_iter: iterations per phase (the reference's 10,000 for the
// route below).  Device memory for the call only.  Extension: the reference solves one LP.
void solve_batch(int count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                 const int32_t* rowidx, const double* vals, const double* b, const double* c,
                 int32_t max_iter, int32_t* basis, double* x, double* y, int32_t* status,
                 int32_t* iterations, double* objective);

// solve_with_phase1 (src/simplex/phase1.cpp:124) on the device, without the reference's
// BasisFactorization argument (the basis inverse lives on the device).  Same model handling:
// a column with a finite upper bound more than 1e-6 above its lower bound returns
// IterationLimit (the reference's unsupported case); other columns are shifted by their lower
// bounds (x = x' + lb); a Maximize model minimizes -c.  basis: basic_indices per row and
// col_status Basic / AtLower; x: the basic solution.  When pi is not null it receives the duals
// of the minimization the route solves, B^-T (s c)_B with s = -1 for a Maximize model -- the
// convention verify_optimal checks (r = s c - A^T pi); for a Minimize model this is the
// reference CLI's pi = B^-T c_B (sankhya_benchmark_cli.cpp:62-69).  std::invalid_argument for m
// above 1,600.
SimplexStatus solve_with_phase1(const core::Model& model, Basis& basis, std::vector<Float>& x,
                                std::vector<Float>* pi = nullptr);

}  // namespace simplex
}  // namespace sankhya

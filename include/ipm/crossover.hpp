// Crossover: from an interior-point solution of  min c^T x, A x = b, x >= 0  to a basic
// solution.  One workgroup per LP on the device (k_crossover), dense explicit basis inverse in
// HBM, so m <= LKB_MAX_M (1,600) per LP.  The basis is built from the solution's support (its
// largest values first), then every remaining positive nonbasic column is purified to zero or
// pivoted in; the final basic values are B^-1 b.
#pragma once

#include <cstdint>
#include <vector>

#include "core/problem.hpp"
#include "gpu/device_model.hpp"
#include "sankhya/types.hpp"
#include "simplex/basis.hpp"

namespace sankhya {
namespace ipm {

// Per-LP outcome codes of the device crossover.
enum CrossoverStatus : int32_t {
    kCrossoverOptimal = 0,          // primal and dual feasible basis: an optimal vertex
    kCrossoverNotDualFeasible = 1,  // feasible vertex whose reduced costs are not all >= 0
    kCrossoverRankDeficient = 2,    // a row left artificial (A does not have full row rank)
    kCrossoverUnbounded = 3,        // a purification direction with no blocking row
    kCrossoverInfeasible = 4,       // B^-1 b has a negative entry
};

// count LPs in the concatenated CSC layout of BatchedMehrotraSolver, with one point x per LP
// (sum n).  Outputs: basis (sum m; basic column of each row, -1 artificial), xv (sum n, x_N =
// 0), status and pivots (count); any output may be null.  Device memory for the call only.
// Extension: the reference has no batched crossover (its per-LP one is the next function).
void crossover_points(int count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                      const int32_t* rowidx, const double* vals, const double* b, const double* c,
                      const double* x, int32_t* basis, double* xv, int32_t* status,
                      int32_t* pivots);

// The reference's entry point (src/ipm/crossover.cuh:26-29, crossover.cu:322-460) on a host
// core::Model: basic_indices per row (cols + i for a row left artificial), col_status Basic,
// AtUpper when the purified value is within 1e-6 of a finite ub, else AtLower.  Throws
// std::runtime_error("Crossover failed: Matrix is rank deficient") when a row stays artificial
// and "Crossover failed: Invalid pivot encountered" when purification finds no blocking row, as
// the reference does for its two failure cases; std::invalid_argument for x of the wrong size
// or m above the device limit.
simplex::Basis generate_crossover_basis(const core::Model& model, const std::vector<Float>& x_ipm);
// The reference's exact signature: the model's arrays are read back from the device (m <=
// 1,600, so at most a few MB), then as above.
simplex::Basis generate_crossover_basis(const gpu::DeviceModel& model, const std::vector<Float>& x_ipm);

}  // namespace ipm
}  // namespace sankhya

// C entry points over the reference's own CPU sources (compiled from where they lie by
// oracle/build_ref.sh into oracle/_ref/libref.so, never copied): the symbolic phase
// (src/ipm/symbolic.cpp:11-218), the optimality certificate (src/verifier/certificate.cpp:
// 8-155) and the CPU simplex route (Phase I + Phase II over the sparse LU basis factorization,
// src/simplex/phase1.cpp:124-449, src/simplex/primal.cpp:27-423, src/numerics/sparse_lu.cpp).
// Test and measurement infrastructure only: the CPU tests use it to cross-check this
// repository's restatements, and the benchmark times the simplex route as the CPU baseline.
#include <chrono>
#include <cstring>
#include <exception>
#include <vector>

#include "core/problem.hpp"
#include "ipm/symbolic.hpp"
#include "numerics/sparse_lu.hpp"
#include "simplex/phase1.hpp"
#include "verifier/certificate.hpp"

using namespace sankhya;

namespace {
core::CSRMatrix csr_of(int m, int n, const int* rp, const int* ci) {
    core::CSRMatrix A;
    A.rows = m;
    A.cols = n;
    A.row_ptrs.assign(rp, rp + m + 1);
    A.col_indices.assign(ci, ci + rp[m]);
    A.values.assign(static_cast<size_t>(rp[m]), 1.0);
    return A;
}
}  // namespace

extern "C" {

// P[m] = compute_amd_ordering(A).  0 on success, 1 on an exception.
int ref_order(int m, int n, const int* rp, const int* ci, int* P) {
    try {
        const auto p = ipm::compute_amd_ordering(csr_of(m, n, rp, ci));
        std::memcpy(P, p.data(), p.size() * sizeof(int));
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// compute_symbolic_factorization(A, P): parent[m], P_inv[m], Lptr[m+1] and, when cap allows,
// Lidx[nnz]; *nnz = nnz of the strictly lower pattern.
int ref_symbolic(int m, int n, const int* rp, const int* ci, const int* P, int* parent,
                 int* P_inv, int* Lptr, int* Lidx, long long cap, long long* nnz) {
    try {
        ipm::SymbolicFactorization s;
        ipm::compute_symbolic_factorization(csr_of(m, n, rp, ci), std::vector<Index>(P, P + m), s);
        std::memcpy(parent, s.parent.data(), s.parent.size() * sizeof(int));
        std::memcpy(P_inv, s.P_inv.data(), s.P_inv.size() * sizeof(int));
        std::memcpy(Lptr, s.L_pattern.row_ptrs.data(), s.L_pattern.row_ptrs.size() * sizeof(int));
        *nnz = static_cast<long long>(s.L_pattern.col_indices.size());
        if (cap >= *nnz)
            std::memcpy(Lidx, s.L_pattern.col_indices.data(), s.L_pattern.col_indices.size() * sizeof(int));
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// verify_optimal on a model given in CSC; lb / ub may be NULL (0, +inf).
int ref_verify(int m, int n, const int* colptr, const int* rowidx, const double* vals,
               const double* c, const double* b, const double* lb, const double* ub,
               int maximize, const double* x, const double* pi, int* pass, double* residual) {
    try {
        core::Model model;
        model.sense = maximize ? OptimizationSense::Maximize : OptimizationSense::Minimize;
        model.A = core::CSCMatrix(m, n, std::vector<Float>(vals, vals + colptr[n]),
                                  std::vector<Index>(rowidx, rowidx + colptr[n]),
                                  std::vector<Index>(colptr, colptr + n + 1));
        model.obj.assign(c, c + n);
        model.lb = lb ? std::vector<Float>(lb, lb + n) : std::vector<Float>(static_cast<size_t>(n), 0.0);
        model.ub = ub ? std::vector<Float>(ub, ub + n)
                      : std::vector<Float>(static_cast<size_t>(n), math::kInfinity);
        model.vtype.assign(static_cast<size_t>(n), VariableType::Continuous);
        model.rhs.assign(b, b + m);
        Float r = 0.0;
        *pass = verifier::verify_optimal(model, std::vector<Float>(x, x + n),
                                         std::vector<Float>(pi, pi + m), &r) ? 1 : 0;
        *residual = r;
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// The reference CLI's recovery route (tests/benchmark/sankhya_benchmark_cli.cpp:48-75) on
// min c^T x, A x = b, x >= 0 (CSC): solve_with_phase1 with SparseLUFactorization, then the
// objective.  *status: 0 Optimal, 1 Infeasible, 2 Unbounded, 3 IterationLimit (SimplexStatus);
// x (n) receives the final point; *ms the wall time of the call.  Returns 1 on an exception.
int ref_phase1_solve(int m, int n, const int* colptr, const int* rowidx, const double* vals,
                     const double* c, const double* b, int* status, double* objective,
                     double* x, double* ms) {
    try {
        core::Model model;
        model.sense = OptimizationSense::Minimize;
        model.A = core::CSCMatrix(m, n, std::vector<Float>(vals, vals + colptr[n]),
                                  std::vector<Index>(rowidx, rowidx + colptr[n]),
                                  std::vector<Index>(colptr, colptr + n + 1));
        model.obj.assign(c, c + n);
        model.lb.assign(static_cast<size_t>(n), 0.0);
        model.ub.assign(static_cast<size_t>(n), math::kInfinity);
        model.vtype.assign(static_cast<size_t>(n), VariableType::Continuous);
        model.rhs.assign(b, b + m);
        simplex::Basis basis;
        std::vector<Float> xs;
        numerics::SparseLUFactorization factorizer;
        const auto t0 = std::chrono::steady_clock::now();
        const simplex::SimplexStatus st = simplex::solve_with_phase1(model, basis, xs, factorizer);
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        *status = static_cast<int>(st);
        double obj = 0.0;
        for (size_t j = 0; j < xs.size() && j < static_cast<size_t>(n); ++j) {
            obj += xs[j] * c[j];
            x[j] = xs[j];
        }
        *objective = obj;
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

}  // extern "C"

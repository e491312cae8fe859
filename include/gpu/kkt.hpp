// GPU normal-equations solver for M = A * Theta * A^T (interior-point KKT systems).
//
// Reference interface (kept, same names and semantics): src/cuda/kkt.cuh:38-71
//   GPUKKTCholeskySolver(arena, A, sym, M_pattern)
//   gpu_cholesky_factorize(A, Theta)          host Theta; throws std::runtime_error if not SPD
//   gpu_cholesky_solve(rhs, tol = 1e-10)      host rhs, overwritten with dy; false on failure
//   gpu_cholesky_factorize_device(d_Theta)    M = P A Theta A^T P^T + 1e-9 I factored in place
//   gpu_cholesky_solve_device(d_rhs, tol=-1)  equilibrated solve with delta retries (1e-9 x10,
//                                             5 tries) and refinement against M0 (<= 10 steps,
//                                             rollback on stagnation); success iff
//                                             ||rhs - M0 dy||_inf <= tol; d_rhs <- dy always
//   get_device_L_values()                     L in the reference CSR layout (lower, diagonal
//                                             last in each row, pattern of sym.L_pattern)
//
// MI355X design: a supernodal factorization planned once from the symbolic structure
// (ordering sym.P, postordered), numeric-only assembly over a precomputed pair map, MFMA tile
// kernels, and one factorization of D*M0*D + delta*I per Theta reused by every solve (the
// reference refactorizes inside every solve call; the factor it recomputes is identical).
// M_pattern is accepted for interface compatibility; the structure is derived from A and P.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>
#include "core/sparse_matrix.hpp"
#include "gpu/vram_arena.hpp"
#include "ipm/symbolic.hpp"
#include "sankhya/types.hpp"

namespace sankhya {
namespace gpu {

enum class FactorPrecision { FP64 = 0, FP32 = 1 };

struct KKTOptions {
    ipm::SupernodeOptions supernodes;
    bool use_graphs = false;        // launch-graph capture: not supported (true is rejected)
    // STEP tiles: 0 one launch per block column; 1 all tiles of a level in one dependency-driven
    // launch unless one supernode dominates the level (then one launch per block column);
    // 2 one dependency-driven launch for every level
    int step_dag = 1;
    void* stream = nullptr;         // hipStream_t to run on (nullptr: the solver creates one)
    // factor_equilibrated regularization: delta_r = delta0 * 10^r, r < delta_retries
    // (src/cuda/kkt.cu:652-690 uses 1e-9 and 5 for both precisions)
    double delta0_fp64 = 1e-9;
    double delta0_fp32 = 1e-9;
    int delta_retries = 5;
};

struct KKTSolveInfo {
    int retries = 0;               // delta increases used (0 = delta0)
    int refinement_steps = 0;
    double delta = 0.0;
    double residual_inf = 0.0;     // final ||rhs - M0 dy||_inf
    int pcg_iterations = 0;
};

struct KKTStats {
    int64_t m = 0, n = 0, nnz_A = 0;
    int64_t nsuper = 0, nnz_L = 0, nnz_L_exact = 0, npairs = 0, nsegments = 0;
    int32_t levels = 0, launches_factor = 0, launches_solve = 0;
    double flops_exact = 0, flops_panel = 0;
    int64_t panel_elems = 0;
    int64_t factorizations = 0, solves = 0, refinement_steps = 0;
    double symbolic_ms = 0;
};

class GPUKKTCholeskySolver {
public:
    GPUKKTCholeskySolver(VRAMArena& arena, const core::CSRMatrix& A,
                         const ipm::SymbolicFactorization& sym, const core::CSRMatrix& M_pattern);
    GPUKKTCholeskySolver(VRAMArena& arena, const core::CSRMatrix& A,
                         const ipm::SymbolicFactorization& sym, const KKTOptions& opt);
    // From a precomputed supernodal structure (its ordering S.perm is used as is).
    GPUKKTCholeskySolver(VRAMArena& arena, const core::CSRMatrix& A,
                         ipm::SupernodalStructure&& structure, const KKTOptions& opt);
    // Device bytes the solver will take from its arena for this structure (upper bound).
    static std::size_t estimate_bytes(const core::CSRMatrix& A,
                                      const ipm::SupernodalStructure& S);
    // Host-only construction of the complete device plan (structure, pair map, task lists) for
    // inspection and CPU emulation; keys name the arrays the device would receive.
    static std::map<std::string, std::vector<char>> export_plan(
        const core::CSRMatrix& A, const std::vector<Index>& P, const ipm::SupernodeOptions& so,
        int step_dag = 1);
    ~GPUKKTCholeskySolver();
    GPUKKTCholeskySolver(const GPUKKTCholeskySolver&) = delete;
    GPUKKTCholeskySolver& operator=(const GPUKKTCholeskySolver&) = delete;

    void gpu_cholesky_factorize(const core::CSRMatrix& A, const std::vector<Float>& Theta);
    bool gpu_cholesky_solve(std::vector<Float>& rhs, Float required_abs_tol = 1e-10);
    void gpu_cholesky_factorize_device(const Float* d_Theta);
    bool gpu_cholesky_solve_device(Float* d_rhs, Float required_abs_tol = -1.0);
    Float* get_device_L_values() const;

    // ---- extensions used by the interior-point loop ----
    // Assemble D*M0*D + delta*I for d_Theta and factor it (delta = 1e-9 * 10^r, r < 5).
    // Returns the retry index used, or -1 when every delta failed.
    // Regularization of the equilibrated factor (defaults: the reference's 1e-9 x10, 5 tries).
    void set_regularization(double delta0_fp64, double delta0_fp32, int retries);
    int factor_equilibrated(const Float* d_Theta, FactorPrecision prec = FactorPrecision::FP64);
    // Solve with the cached factor (refining against M0 in FP64).  d_rhs is in original row
    // order and is overwritten with dy.  Same contract as gpu_cholesky_solve_device.
    bool solve_cached(Float* d_rhs, Float required_abs_tol, KKTSolveInfo* info = nullptr,
                      int max_refine = 10);
    // Direct solve + preconditioned CG on M0 dy = rhs with the cached (FP64 or FP32) factor as
    // preconditioner.  Success when ||rhs - M0 dy||_inf <= tol or, if backward_error_acceptance,
    // when ||rhs - M0 dy||_inf <= 1e-12 (||rhs||_inf + max_i M0_ii ||dy||_inf).
    bool solve_pcg(Float* d_rhs, Float required_abs_tol, int max_iter, KKTSolveInfo* info,
                   bool backward_error_acceptance);
    // Raw triangular solves with the cached factor on a permuted device vector (tests).
    void triangular_solve_permuted(Float* d_x);

    const KKTStats& stats() const;
    // One factorization of the last Theta with events around every launch: per kind (0 STEP,
    // 1 DIAG + DIAGWB, 2 UPDATE, 3 assembly, 5 PART) device ms, stored-panel flops and launch
    // count; index 4 = the whole factorization including assembly (flops = sum_j c_j^2 of the
    // exact factor).
    void profile_factor(FactorPrecision prec, double ms[6], double flops[6], double launches[6]);
    // Extension: live timing of the factorizations this object runs (HIP events around every
    // launch on its stream, slots as profile_factor; slot 4 counts factorizations and their
    // exact flops).  set_launch_timing resets the totals; launch_timing drains the stream.
    void set_launch_timing(bool on);
    void launch_timing(double ms[6], double flops[6], double launches[6]);
    const std::vector<Index>& permutation() const;     // ordering used (postordered)
    void* stream() const;
    Index rows() const;
    Index cols() const;

    struct Impl;

private:
    std::unique_ptr<Impl> impl_;
};

}  // namespace gpu
}  // namespace sankhya

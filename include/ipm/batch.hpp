// Batched interior-point solver: thousands of small, independent LPs
//   min c_i^T x, A_i x = b_i, x >= 0
// solved together on one MI355X in a persistent launch, one 256-thread workgroup per LP at a
// time (dense normal matrix per LP, MFMA Cholesky, PCG / refinement), each LP running the
// single-LP MehrotraSolver algorithm (src/ipm/mehrotra_device.cu:695-1047) in the relative-gap
// or the reference stop mode.
//
// Extension: the reference has no batched interior-point path; its one-workgroup-per-problem
// pattern for batches of small problems is src/gpu/batched_dual_simplex.cu:7,55-58.
//
// Launch shape: `wg_per_cu` persistent workgroups per compute unit take LPs from a ticket
// counter in `order` (largest m first by default); each visit runs `iters_per_launch` IPM
// iterations; the sequence of launches can be captured once in a hipGraph and replayed until
// every LP has finished.  autotune() times the candidate shapes on a sample of the batch and
// keeps the fastest.
#pragma once
#include <cstdint>
#include <vector>
#include "sankhya/types.hpp"

namespace sankhya {
namespace ipm {

struct BatchOptions {
    int stop_rule = 1;              // 0 reference (scaled norms < tol), 1 relative gap (gap_tol)
    int precision = 0;              // 0 FP64 factor, 1 FP32 factor + FP64 PCG (FP64 fallback)
    int max_iter = 200;
    double tol = 1e-6, gap_tol = 1e-8;
    double mixed_fp64_spread = 1e12;
    int mixed_max_refine = 60;
    int wg_per_cu = 0;              // persistent workgroups per CU (0: autotuned or 2)
    int iters_per_launch = 0;       // IPM iterations per LP visit (0: autotuned or all)
    int use_graph = 1;              // replay the launch sequence from a captured hipGraph
    int order = 0;                  // 0 largest m first, 1 input order
    double warm_floor = 1e-3;       // lower clamp of warm-start x and s (scaled space)
    double warm_mu = 0.0;           // lift warm pairs to x_j s_j >= warm_mu (scaled space; 0: none)
    // a warm point whose relative infeasibility max(|Ax-b|_inf / (1+|b|_inf),
    // |A^T y+s-c|_inf / (1+|c|_inf)) exceeds this is rejected for the standard start
    double warm_max_infeas = 1e-2;
};

struct BatchLPResult {
    int status = 3;                 // 0 Optimal, 3 IterationLimit
    int iterations = 0, fail_code = 0;
    int factorizations = 0, fp32_factorizations = 0, solves = 0, pcg_steps = 0;
    double objective = 0, dual_objective = 0, rel_gap = 0, rel_primal = 0, rel_dual = 0;
    double primal_residual = 0, dual_residual = 0, mu = 0;
    // device time of the LP's workgroup (ms): scaling, sparse passes and vector work, assembly,
    // dense factorization, dense triangular solves
    double ms_init = 0, ms_other = 0, ms_assemble = 0, ms_factor = 0, ms_solve = 0;
    int warm = 0;                   // 0 standard start, 1 warm point used, 2 warm point rejected
    double warm_infeas = 0;         // relative infeasibility of the warm point offered
};

struct BatchSummary {
    int count = 0, optimal = 0, launches = 0, graph_replays = 0;
    double ms = 0;                  // device time of the solve (events around the launches)
    double flops_chol = 0;          // sum over LPs and factorizations of m^3 / 3
    int wg_per_cu = 0, iters_per_launch = 0, grid = 0;
};

class BatchedMehrotraSolver {
public:
    // count LPs in CSC; LP i has m[i] rows and n[i] columns, its column pointers are the n[i]+1
    // entries of colptr starting at sum_{k<i} (n[k] + 1) (each starting at 0), its rows and
    // values start at sum_{k<i} nnz_k in rowidx / vals, b and c at sum m[k] / sum n[k].
    BatchedMehrotraSolver(int count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                          const int32_t* rowidx, const double* vals, const double* b,
                          const double* c, const BatchOptions& opt);
    ~BatchedMehrotraSolver();
    BatchedMehrotraSolver(const BatchedMehrotraSolver&) = delete;
    BatchedMehrotraSolver& operator=(const BatchedMehrotraSolver&) = delete;

    // Solve every LP from the standard start point (x = s = 1, y = 0); inputs stay resident.
    BatchSummary solve();
    // Time the candidate launch shapes on `sample` LPs spread over the size range and keep the
    // fastest; returns the chosen wg_per_cu and iters_per_launch in the summary.
    BatchSummary autotune(int sample);
    const std::vector<BatchLPResult>& results() const;
    // Original-space start points for the next solves (x, s: sum n; y: sum m, the layout of
    // solution()); LP i uses its point when flags[i] != 0 (flags null: all).  x0 null: every LP
    // back to the standard start.
    void set_warm_start(const double* x0, const double* y0, const double* s0, const int32_t* flags);
    // original-space solution of LP i (after solve())
    void solution(int i, double* x, double* y, double* s) const;
    // Basic solutions from the last solve's x, one workgroup per LP (the reference's counterpart
    // for one LP is generate_crossover_basis, src/ipm/crossover.cu:322).  Host outputs in the
    // layouts of the inputs: basis (sum m; the basic column of each row, -1 for a row left
    // artificial), x (sum n; x_N = 0), and per LP status (0 optimal vertex, 1 primal feasible
    // basis that is not dual feasible, 2 rank deficient, 3 unbounded purification direction,
    // 4 basic solution infeasible) and pivots; any output may be null.  The solver's factor and
    // work buffers serve as the crossover's workspace: the next solve() rebuilds them.
    void crossover(int32_t* basis, double* x, int32_t* status, int32_t* pivots);
    int count() const;

    struct Impl;

private:
    Impl* impl_;
};

// For callers that cannot allocate device memory (code under graph capture, serving loops,
// harnesses that forbid hipMalloc): the same solve with every device buffer carved from a fixed
// device scratch array of LKB_SCRATCH_BYTES (64 MiB) instead of an arena; no hipMalloc / hipFree.
// Host inputs and outputs as the class (x, s: sum n; y: sum m; results: count).  A batch whose
// device footprint exceeds the scratch is rejected (std::invalid_argument).  Calls on one GPU
// share its scratch: they are serialized within a process, and the scratch keeps the last call's
// contents (every buffer is rewritten before it is read).  Extension with no reference
// counterpart.
BatchSummary solve_batch_noalloc(int count, const int32_t* m, const int32_t* n,
                                 const int32_t* colptr, const int32_t* rowidx, const double* vals,
                                 const double* b, const double* c, const BatchOptions& opt,
                                 double* x, double* y, double* s, BatchLPResult* results);

}  // namespace ipm
}  // namespace sankhya

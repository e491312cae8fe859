/*
 * C ABI of the MI355X interior-point path (liblakshya.so).  Plain pointers and sizes only.
 * All functions return LK_OK (0) or an error code; lk_last_error() gives the message of the
 * last failure on the calling thread.  Host arrays unless a parameter says "device".
 *
 * Each entry point names the reference interface it replaces.  The reference has no C binding
 * for this path (its C API, include/sankhya/sankhya.h:30-63, only reaches the CPU simplex), so
 * these are the functions a ctypes / cgo / JNI binding of the C++ interface would need.
 */
#ifndef LAKSHYA_H
#define LAKSHYA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LK_OK 0
#define LK_ERR_INVALID 1      /* std::invalid_argument (dimension mismatch, bad permutation, ...) */
#define LK_ERR_NOMEM 2        /* std::bad_alloc (arena exhausted or fragmented) */
#define LK_ERR_RUNTIME 3      /* std::runtime_error (HIP failure, not positive definite) */
#define LK_ERR_INTERNAL 5

typedef struct lk_arena lk_arena;
typedef struct lk_kkt lk_kkt;
typedef struct lk_ipm lk_ipm;

const char* lk_last_error(void);
int lk_version(void);
int lk_device_count(int* count);

/* ---- VRAM arena: gpu::VRAMArena (src/gpu/vram_arena.cuh:21-75) ---- */
int lk_arena_create(size_t capacity_bytes, lk_arena** out);
int lk_arena_alloc(lk_arena* a, size_t bytes, void** device_ptr);   /* 0 bytes -> NULL */
int lk_arena_free(lk_arena* a, void* device_ptr);
int lk_arena_occupancy(lk_arena* a, size_t* percent);
int lk_arena_destroy(lk_arena* a);

/* ---- host symbolic phase: ipm::compute_amd_ordering / compute_symbolic_factorization
 *      (src/ipm/symbolic.hpp:50-63) and the supernodal structure ---- */
int lk_order_reference(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                       int32_t* P);
int lk_order_amd(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                 int32_t* P);
/* parent[m], Lptr[m+1]; Lidx capacity `cap` (query the size with cap = 0: *nnz is set) */
int lk_symbolic(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                const int32_t* P, int32_t* parent, int32_t* P_inv, int32_t* Lptr, int32_t* Lidx,
                int64_t cap, int64_t* nnz);

typedef struct {
    int64_t nsuper, nnz_L, nnz_L_exact, npairs, nsegments, panel_elems;
    int32_t levels, launches_factor, launches_solve, pad;
    double flops_exact, flops_panel, symbolic_ms;
    int64_t factorizations, solves, refinement_steps;
} lk_kkt_stats;

/* Supernodal structure statistics for an ordering (NULL P -> AMD); perm_out[m] optional. */
int lk_supernodal(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                  const int32_t* P, int32_t relax_small, double relax_frac, lk_kkt_stats* out,
                  int32_t* perm_out, int32_t* sn_first_out, int32_t* nsuper_out);

/* ---- plan inspection (host only): the device plan of the KKT solver for an ordering P
 * (NULL -> AMD).  step_dag: 0 one STEP launch per block column, 1 the solver's choice per level,
 * 2 one dependency-driven launch per level.  lk_plan_array copies the named array (call with
 * dst NULL for its size). */
typedef struct lk_plan lk_plan;
int lk_plan_create(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                   const double* vals, const int32_t* P, int32_t relax_small, double relax_frac,
                   int32_t step_dag, lk_plan** out);
int lk_plan_array(lk_plan* p, const char* name, void* dst, size_t capacity, size_t* bytes);
int lk_plan_destroy(lk_plan* p);

/* ---- KKT solver: gpu::GPUKKTCholeskySolver (src/cuda/kkt.cuh:38-71) ----
 * A in CSR.  P: ordering (NULL -> AMD); when `reference_symbolic` != 0 the reference ordering
 * rule is used and L is exposed in the reference CSR layout by lk_kkt_get_L_csr. */
int lk_kkt_create(lk_arena* arena, int32_t m, int32_t n, const int32_t* rowptr,
                  const int32_t* colidx, const double* vals, const int32_t* P,
                  int32_t reference_symbolic, int32_t use_graphs, lk_kkt** out);   /* use_graphs must be 0 */
/* mode 0: gpu_cholesky_factorize (M0 + 1e-9 I, unequilibrated; *not_spd set instead of throwing)
 * mode 1: equilibrated FP64 factor (delta retries), mode 2: equilibrated FP32 factor.
 * theta: n values, device pointer when on_device != 0. */
int lk_kkt_factor(lk_kkt* k, const double* theta, int32_t on_device, int32_t mode,
                  int32_t* not_spd);
typedef struct {
    int32_t success, retries, refinement_steps, pad;
    double delta, residual_inf;
} lk_kkt_solve_info;
/* gpu_cholesky_solve[_device]: rhs (m) overwritten with dy; info->success as the reference's
 * bool return. max_refine < 0 -> reference budget (10). */
int lk_kkt_solve(lk_kkt* k, double* rhs, int32_t on_device, double tol, int32_t max_refine,
                 lk_kkt_solve_info* info);
/* L values in the reference layout (rows of sym.L_pattern, diagonal last), after mode 0. */
int lk_kkt_get_L_csr(lk_kkt* k, double* vals, int64_t nnz);
int lk_kkt_permutation(lk_kkt* k, int32_t* perm);
int lk_kkt_get_stats(lk_kkt* k, lk_kkt_stats* out);
int lk_kkt_stream(lk_kkt* k, void** stream);
int lk_kkt_destroy(lk_kkt* k);

/* ---- interior-point solver: ipm::MehrotraSolver (src/ipm/mehrotra.hpp:10-31) ---- */
typedef struct {
    int32_t stop_rule;          /* 0 reference (scaled norms < tol), 1 relative gap (gap_tol) */
    int32_t max_iter;
    double tol, gap_tol;
    int32_t precision;          /* 0 FP64, 1 mixed (FP32 factor + FP64 refinement) */
    int32_t mixed_max_refine;
    double mixed_fp64_spread;
    int32_t ordering;           /* 0 AMD, 1 reference minimum degree */
    int32_t use_graphs;         /* must be 0: launch-graph capture is not supported */
    int32_t verbose;
    int32_t pad;
    double warm_floor;
} lk_ipm_options;

typedef struct {
    int32_t status;             /* 0 Optimal, 1 Infeasible, 2 Unbounded, 3 IterationLimit */
    int32_t iterations;
    double objective, primal_residual, dual_residual, duality_gap;
    double dual_objective, rel_gap, rel_primal_residual, rel_dual_residual;
    int32_t factorizations, fp32_factorizations, fp64_factorizations, solves;
    int32_t refinement_steps;
    int32_t fail_code;          /* 0 none; 1 factorization, 2 predictor solve, 3 predictor check,
                                   4 corrector solve, 5 no step accepted */
    double setup_ms, solve_ms, factor_ms, nnz_L, flops_per_factorization;
} lk_ipm_result;

int lk_ipm_default_options(lk_ipm_options* opt);
/* The solver's equilibration (host only): 10 Jacobi Ruiz passes of the MehrotraSolver
 * constructor (src/ipm/mehrotra_device.cu:286-323).  A in CSR; vals (nnz), b (m), c (n) are
 * scaled in place; R (m) and C (n) receive the factors (scaled A = R A C). */
int lk_ruiz_scaling(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                    double* vals, double* b, double* c, double* R, double* C);
/* One LP in CSC (rows ascending per column): construct + solve.  lb/ub may be NULL (0, +inf);
 * warm_* (original space) optional; x (n), y (m), s (n) outputs optional. */
int lk_ipm_solve(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                 const double* vals, const double* c, const double* b, const double* lb,
                 const double* ub, const lk_ipm_options* opt, const double* warm_x,
                 const double* warm_y, const double* warm_s, lk_ipm_result* res, double* x,
                 double* y, double* s);
/* Reusable solver for a fixed A: setup once, solve for many (b, c). */
int lk_ipm_create(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                  const double* vals, const double* c, const double* b,
                  const lk_ipm_options* opt, lk_ipm** out);
int lk_ipm_run(lk_ipm* h, const double* c, const double* b, const double* warm_x,
               const double* warm_y, const double* warm_s, lk_ipm_result* res, double* x,
               double* y, double* s);
int lk_ipm_destroy(lk_ipm* h);
/* Factorization profile at the last iterate, one extra factorization with events around every
 * launch (slot 0 STEP, 1 DIAG + DIAGWB, 2 UPDATE, 3 assembly, 4 whole factorization, 5 PART):
 * device ms, flops (stored panel work; slot 4: sum_j c_j^2 of the exact factor), launches
 * (6 each).
 * Extension: the reference's profiler (src/ipm/mehrotra_device.cu:45-91) times host phases. */
int lk_ipm_profile_factor(lk_ipm* h, int32_t fp32, double* ms, double* flops, double* launches);
/* Live timing of the factorizations run by lk_ipm_run: on != 0 starts (and resets) HIP-event
 * timing of every factorization launch; read returns the totals so far in the slots above
 * (slot 3 unused; slot 4: factorizations timed end to end).  Extension, no reference
 * counterpart. */
int lk_ipm_launch_timing(lk_ipm* h, int32_t on);
int lk_ipm_read_launch_timing(lk_ipm* h, double* ms, double* flops, double* launches);
/* m, n, nnz(A), supernodes, nnz(L) stored, nnz(L) exact, flops exact, flops stored, levels,
 * factor launches, solve launches, symbolic ms */
int lk_ipm_kkt_stats(lk_ipm* h, double* out12);

/* ---- batched interior-point solver: ipm::BatchedMehrotraSolver (include/ipm/batch.hpp) ----
 * Extension with no reference counterpart (the reference solves one LP per MehrotraSolver; its
 * batch pattern, one block per problem, is src/gpu/batched_dual_simplex.cu:7,55-58).  count LPs
 * in CSC, concatenated: LP i's column pointers are n[i] + 1 entries of colptr (starting at 0)
 * following those of LP i-1; its rows / values follow in rowidx / vals, b and c likewise. */
typedef struct lk_batch lk_batch;
typedef struct {
    int32_t stop_rule;          /* 0 reference (scaled norms < tol), 1 relative gap (default) */
    int32_t precision;          /* 0 FP64 factor, 1 FP32 factor + FP64 PCG */
    int32_t max_iter, mixed_max_refine;
    double tol, gap_tol, mixed_fp64_spread;
    int32_t wg_per_cu;          /* persistent workgroups per CU (0: 2, or the autotuned value) */
    int32_t iters_per_launch;   /* IPM iterations per LP visit (0: all, or the autotuned value) */
    int32_t use_graph;          /* replay the launch sequence from a captured hipGraph */
    int32_t order;              /* 0 largest m first, 1 input order */
    double warm_floor;          /* lower clamp of warm-start x and s (scaled space) */
    double warm_mu;             /* lift of warm-start pairs to x_j s_j >= warm_mu (0: none) */
    double warm_max_infeas;     /* warm point rejected (standard start) above this relative
                                   infeasibility, max(|Ax-b|_inf/(1+|b|_inf), |A'y+s-c|_inf/(1+|c|_inf)) */
} lk_batch_options;
typedef struct {
    int32_t count, optimal, launches, graph_replays;
    double ms, flops_chol;      /* device ms of the solve; sum over LPs of factorizations m^3/3 */
    int32_t wg_per_cu, iters_per_launch, grid, pad;
} lk_batch_summary;
typedef struct {
    int32_t status, iterations, fail_code, factorizations;
    int32_t fp32_factorizations, solves, pcg_steps, pad;
    double objective, dual_objective, rel_gap, rel_primal, rel_dual;
    double primal_residual, dual_residual, mu;
    double ms_init, ms_other, ms_assemble, ms_factor, ms_solve;   /* LP's workgroup time */
    double warm_infeas;         /* relative infeasibility of the warm point offered */
    int32_t warm, pad2;         /* 0 standard start, 1 warm point used, 2 warm point rejected */
} lk_batch_result;
int lk_batch_default_options(lk_batch_options* opt);
int lk_batch_create(int32_t count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                    const int32_t* rowidx, const double* vals, const double* b, const double* c,
                    const lk_batch_options* opt, lk_batch** out);
int lk_batch_solve(lk_batch* h, lk_batch_summary* sum);
/* time the launch shapes on `sample` LPs and keep the fastest for later solves */
int lk_batch_autotune(lk_batch* h, int32_t sample, lk_batch_summary* chosen);
int lk_batch_results(lk_batch* h, lk_batch_result* out);          /* count entries */
int lk_batch_solution(lk_batch* h, int32_t i, double* x, double* y, double* s);
/* Basic solutions from the last lk_batch_solve's x, one workgroup per LP (the reference's
 * counterpart for one LP: generate_crossover_basis, src/ipm/crossover.cu:322).  Host outputs:
 * basis (sum m; basic column of each row, -1 for a row left artificial), x (sum n, x_N = 0),
 * status and pivots (count; status 0 optimal vertex, 1 primal feasible basis that is not dual
 * feasible, 2 rank deficient, 3 unbounded purification direction, 4 basic solution
 * infeasible); any output may be NULL.  LK_ERR_INVALID before the first solve. */
int lk_batch_crossover(lk_batch* h, int32_t* basis, double* x, int32_t* status, int32_t* pivots);
/* The same crossover from caller-given points: count LPs in the lk_batch_create layout and one
 * point per LP in x (sum n), outputs as lk_batch_crossover (ipm::crossover_points,
 * include/ipm/crossover.hpp).  Device memory is allocated for the call only. */
int lk_crossover(int32_t count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                 const int32_t* rowidx, const double* vals, const double* b, const double* c,
                 const double* x, int32_t* basis, double* xv, int32_t* status, int32_t* pivots);

/* ---- device simplex route: the reference's CPU route solve_with_phase1 + primal_simplex_phase2
 * (src/simplex/phase1.cpp:124-449, src/simplex/primal.cpp:27-423) on the device, one workgroup
 * per LP (simplex::solve_batch, include/simplex/primal.hpp): count LPs min c^T x, A x = b,
 * x >= 0 in the lk_batch_create layout (m <= 1,600 each); max_iter iterations per phase (the
 * reference: 10,000).  Outputs (any may be NULL): basis (sum m; -1 for a row left artificial),
 * x (sum n), y (sum m; B^-T c_B), status (count; 0 Optimal, 1 Infeasible, 2 Unbounded,
 * 3 IterationLimit), iterations (2 count: Phase I, Phase II), objective (count).  Device memory
 * is allocated for the call only. */
int lk_simplex(int32_t count, const int32_t* m, const int32_t* n, const int32_t* colptr,
               const int32_t* rowidx, const double* vals, const double* b, const double* c,
               int32_t max_iter, int32_t* basis, double* x, double* y, int32_t* status,
               int32_t* iterations, double* objective);
/* Warm start for the next solves: original-space x, s (sum n) and y (sum m) in the layout of
 * the inputs; LP i uses it when flags[i] != 0 (flags NULL: every LP); x NULL clears it.
 * Extension (the single-LP counterpart is lk_ipm_run's warm_* arguments). */
int lk_batch_set_warm_start(lk_batch* h, const double* x, const double* y, const double* s,
                            const int32_t* flags);
int lk_batch_destroy(lk_batch* h);
/* No reference counterpart.  For callers that cannot allocate device memory (code under graph
 * capture, serving loops, harnesses that forbid hipMalloc): the lk_batch_create + lk_batch_solve
 * computation in one call that allocates no device memory -- every device buffer lies in a fixed
 * device scratch array of LKB_SCRATCH_BYTES (64 MiB, lakshya_device.h) held per GPU until the
 * process ends.  Host inputs as lk_batch_create; host outputs x, s (sum n), y (sum m), results
 * (count), sum optional.  A batch needing more scratch is rejected with LK_ERR_INVALID.  Calls on
 * one GPU share its scratch: they are serialized within the process. */
int lk_batch_solve_noalloc(int32_t count, const int32_t* m, const int32_t* n,
                           const int32_t* colptr, const int32_t* rowidx, const double* vals,
                           const double* b, const double* c, const lk_batch_options* opt,
                           double* x, double* y, double* s, lk_batch_result* results,
                           lk_batch_summary* sum);

/* ---- warm-start network: learn::WarmStartMLP (include/learn/mlp.hpp) ----
 * Extension with no reference counterpart (the start point it replaces is x = s = 1, y = 0,
 * src/ipm/mehrotra_device.cu:93-102).  Three-layer perceptron d_in -> h1 -> h2 -> d_out with
 * ReLU, BF16 MFMA products with FP32 accumulation, FP32 master weights and Adam.  Host arrays,
 * row-major; parameters concatenated W1 (h1 x d_in), b1, W2 (h2 x h1), b2, W3 (d_out x h2), b3. */
typedef struct lk_mlp lk_mlp;
int lk_mlp_create(int32_t d_in, int32_t h1, int32_t h2, int32_t d_out, int32_t max_batch,
                  uint64_t seed, lk_mlp** out);
int lk_mlp_param_count(lk_mlp* h, int64_t* count);
int lk_mlp_set_params(lk_mlp* h, const float* params);
int lk_mlp_get_params(lk_mlp* h, float* params);
int lk_mlp_forward(lk_mlp* h, const float* u, int32_t batch, float* y);
/* gradient of the mean squared error (1 / (batch d_out)) sum (y - t)^2; *loss = that error */
int lk_mlp_gradients(lk_mlp* h, const float* u, const float* t, int32_t batch, float* grad,
                     double* loss);
int lk_mlp_set_dataset(lk_mlp* h, const float* u, const float* t, int32_t n);
/* Adam over `epochs` shuffled passes of full mini-batches; epoch_loss[epochs] optional */
int lk_mlp_train(lk_mlp* h, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                 double* epoch_loss);
int lk_mlp_destroy(lk_mlp* h);

/* ---- MPS input: parsers::read_mps (src/parsers/mps_reader.hpp:8-20, mps_reader.cpp:65-355) ----
 * lk_mps_open parses the file (L / G rows get slack / surplus columns after the structural
 * ones; the model is in equality form, Minimize) and returns its sizes; lk_mps_get copies it
 * out -- A in CSC (colptr n + 1, rowidx and vals nnz), obj, lb, ub and vtype (0 Continuous,
 * 1 Integer, 2 Binary) n, rhs m; lk_mps_close frees it.  Parse errors: LK_ERR_RUNTIME with
 * the message in lk_last_error(). */
typedef struct lk_mps lk_mps;
int lk_mps_open(const char* path, lk_mps** out, int32_t* m, int32_t* n, int64_t* nnz);
int lk_mps_get(lk_mps* h, int32_t* colptr, int32_t* rowidx, double* vals, double* obj,
               double* rhs, double* lb, double* ub, int32_t* vtype);
int lk_mps_close(lk_mps* h);

/* ---- verifier::verify_optimal (src/verifier/certificate.hpp:26-31) with a tolerance ---- */
int lk_verify_optimal(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                      const double* vals, const double* c, const double* b, const double* lb,
                      const double* ub, int32_t maximize, const double* x, const double* pi,
                      double tol, double* primal_residual, int32_t* pass);

#ifdef __cplusplus
}
#endif
#endif /* LAKSHYA_H */

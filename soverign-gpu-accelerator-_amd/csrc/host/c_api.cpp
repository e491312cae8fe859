// C ABI over the C++ solver classes (include/lakshya.h).  Exceptions never cross the boundary:
// they become error codes with the message kept per thread for lk_last_error().
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>
#include "core/problem.hpp"
#include "gpu/kkt.hpp"
#include "gpu/vram_arena.hpp"
#include "ipm/batch.hpp"
#include "ipm/crossover.hpp"
#include "ipm/mehrotra.hpp"
#include "ipm/symbolic.hpp"
#include "lakshya.h"
#include "learn/mlp.hpp"
#include "parsers/mps_reader.hpp"
#include "simplex/primal.hpp"
#include "lakshya_device.h"
#include "verifier/certificate.hpp"

using namespace sankhya;

namespace {

thread_local std::string g_err;

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return LK_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return LK_ERR_INVALID;
    } catch (const std::bad_alloc& e) {
        g_err = e.what();
        return LK_ERR_NOMEM;
    } catch (const std::runtime_error& e) {
        g_err = e.what();
        return LK_ERR_RUNTIME;
    } catch (const std::exception& e) {
        g_err = e.what();
        return LK_ERR_INTERNAL;
    } catch (...) {
        g_err = "unknown exception";
        return LK_ERR_INTERNAL;
    }
}

core::CSRMatrix make_csr(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                         const double* vals) {
    if (m < 0 || n < 0 || !rowptr) throw std::invalid_argument("bad CSR arguments");
    core::CSRMatrix A;
    A.rows = m;
    A.cols = n;
    A.row_ptrs.assign(rowptr, rowptr + m + 1);
    const int32_t nz = rowptr[m];
    A.col_indices.assign(colidx, colidx + nz);
    if (vals) A.values.assign(vals, vals + nz);
    else A.values.assign(static_cast<size_t>(nz), 1.0);
    for (int32_t i = 0; i < m; ++i) {
        if (rowptr[i + 1] < rowptr[i]) throw std::invalid_argument("row pointers not monotone");
        for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            if (colidx[p] < 0 || colidx[p] >= n) throw std::invalid_argument("column index out of range");
            if (p > rowptr[i] && colidx[p] <= colidx[p - 1])
                throw std::invalid_argument("column indices must increase within a row");
        }
    }
    return A;
}

void make_model(core::Model& model, int32_t m, int32_t n, const int32_t* colptr,
                const int32_t* rowidx, const double* vals, const double* c, const double* b,
                const double* lb, const double* ub) {
    for (int32_t j = 0; j < n; ++j)
        for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p)
            if (p > colptr[j] && rowidx[p] <= rowidx[p - 1])
                throw std::invalid_argument("row indices must increase within a column");
    model.set_from_csc(m, n, colptr, rowidx, vals, c, b, lb, ub);
}

ipm::IpmOptions to_options(const lk_ipm_options* o) {
    ipm::IpmOptions r;
    if (!o) return r;
    r.stop_rule = o->stop_rule == 1 ? ipm::IpmOptions::StopRule::RelativeGap
                                    : ipm::IpmOptions::StopRule::Reference;
    r.max_iter = o->max_iter;
    r.tol = o->tol;
    r.gap_tol = o->gap_tol;
    r.precision = o->precision == 1 ? ipm::IpmOptions::Precision::Mixed
                                    : ipm::IpmOptions::Precision::FP64;
    r.mixed_max_refine = o->mixed_max_refine;
    r.mixed_fp64_spread = o->mixed_fp64_spread;
    r.ordering = o->ordering == 1 ? ipm::IpmOptions::Ordering::ReferenceMinimumDegree
                                  : ipm::IpmOptions::Ordering::Amd;
    r.use_graphs = o->use_graphs != 0;
    r.verbose = o->verbose != 0;
    r.warm_floor = o->warm_floor;
    return r;
}

void fill_result(const ipm::MehrotraResult& r, lk_ipm_result* out, double* x, double* y,
                 double* s) {
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->status = static_cast<int32_t>(r.status);
        out->iterations = r.iterations;
        out->objective = r.objective_value;
        out->primal_residual = r.primal_residual;
        out->dual_residual = r.dual_residual;
        out->duality_gap = r.duality_gap;
        out->dual_objective = r.dual_objective;
        out->rel_gap = r.rel_gap;
        out->rel_primal_residual = r.rel_primal_residual;
        out->rel_dual_residual = r.rel_dual_residual;
        out->factorizations = r.factorizations;
        out->fp32_factorizations = r.fp32_factorizations;
        out->fp64_factorizations = r.fp64_factorizations;
        out->solves = r.solves;
        out->refinement_steps = r.refinement_steps;
        out->fail_code = r.fail_code;
        out->setup_ms = r.setup_ms;
        out->solve_ms = r.solve_ms;
        out->factor_ms = r.factor_ms;
        out->nnz_L = r.nnz_L;
        out->flops_per_factorization = r.flops_per_factorization;
    }
    if (x) std::copy(r.x.begin(), r.x.end(), x);
    if (y) std::copy(r.pi.begin(), r.pi.end(), y);
    if (s) std::copy(r.s.begin(), r.s.end(), s);
}

void copy_stats(const gpu::KKTStats& s, lk_kkt_stats* o) {
    std::memset(o, 0, sizeof(*o));
    o->nsuper = s.nsuper;
    o->nnz_L = s.nnz_L;
    o->nnz_L_exact = s.nnz_L_exact;
    o->npairs = s.npairs;
    o->nsegments = s.nsegments;
    o->panel_elems = s.panel_elems;
    o->levels = s.levels;
    o->launches_factor = s.launches_factor;
    o->launches_solve = s.launches_solve;
    o->flops_exact = s.flops_exact;
    o->flops_panel = s.flops_panel;
    o->symbolic_ms = s.symbolic_ms;
    o->factorizations = s.factorizations;
    o->solves = s.solves;
    o->refinement_steps = s.refinement_steps;
}

}  // namespace

// The arena is shared by the handles created from it, so destroying the lk_arena handle before
// a KKT solver that allocated from it (any order a garbage collector may pick) stays safe.
struct lk_arena {
    std::shared_ptr<gpu::VRAMArena> ptr;
    gpu::VRAMArena& arena;
    explicit lk_arena(size_t b) : ptr(std::make_shared<gpu::VRAMArena>(b)), arena(*ptr) {}
};
struct lk_kkt {
    std::unique_ptr<gpu::GPUKKTCholeskySolver> solver;
    core::CSRMatrix A;
    ipm::SymbolicFactorization sym;
    bool reference_layout = false;
    double* d_rhs = nullptr;
    std::shared_ptr<gpu::VRAMArena> arena;
};
struct lk_plan {
    std::map<std::string, std::vector<char>> arrays;
};
struct lk_ipm {
    core::Model model;
    std::unique_ptr<ipm::MehrotraSolver> solver;
};
struct lk_batch {
    std::unique_ptr<ipm::BatchedMehrotraSolver> solver;
};
struct lk_mlp {
    std::unique_ptr<learn::WarmStartMLP> net;
};

namespace {
ipm::BatchOptions to_batch_options(const lk_batch_options* o) {
    ipm::BatchOptions r;
    if (!o) return r;
    r.stop_rule = o->stop_rule;
    r.precision = o->precision;
    r.max_iter = o->max_iter;
    r.mixed_max_refine = o->mixed_max_refine;
    r.tol = o->tol;
    r.gap_tol = o->gap_tol;
    r.mixed_fp64_spread = o->mixed_fp64_spread;
    r.wg_per_cu = o->wg_per_cu;
    r.iters_per_launch = o->iters_per_launch;
    r.use_graph = o->use_graph;
    r.order = o->order;
    r.warm_floor = o->warm_floor;
    r.warm_mu = o->warm_mu;
    r.warm_max_infeas = o->warm_max_infeas;
    return r;
}
void copy_summary(const ipm::BatchSummary& s, lk_batch_summary* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->count = s.count;
    o->optimal = s.optimal;
    o->launches = s.launches;
    o->graph_replays = s.graph_replays;
    o->ms = s.ms;
    o->flops_chol = s.flops_chol;
    o->wg_per_cu = s.wg_per_cu;
    o->iters_per_launch = s.iters_per_launch;
    o->grid = s.grid;
}
}  // namespace

extern "C" {

const char* lk_last_error(void) { return g_err.c_str(); }
int lk_version(void) { return 100; }
int lk_device_count(int* count) {
    const int e = lkd_device_count(count);
    if (e) { g_err = lkd_error_string(e); *count = 0; return LK_ERR_RUNTIME; }
    return LK_OK;
}

int lk_arena_create(size_t capacity_bytes, lk_arena** out) {
    return guarded([&] { *out = new lk_arena(capacity_bytes); });
}
int lk_arena_alloc(lk_arena* a, size_t bytes, void** ptr) {
    return guarded([&] { *ptr = a->arena.allocate(bytes); });
}
int lk_arena_free(lk_arena* a, void* ptr) { return guarded([&] { a->arena.free(ptr); }); }
int lk_arena_occupancy(lk_arena* a, size_t* pct) {
    return guarded([&] { *pct = a->arena.occupancy_percentage(); });
}
int lk_arena_destroy(lk_arena* a) { return guarded([&] { delete a; }); }

int lk_order_reference(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                       int32_t* P) {
    return guarded([&] {
        auto A = make_csr(m, n, rowptr, colidx, nullptr);
        auto p = ipm::compute_amd_ordering(A);
        std::copy(p.begin(), p.end(), P);
    });
}
int lk_order_amd(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                 int32_t* P) {
    return guarded([&] {
        auto A = make_csr(m, n, rowptr, colidx, nullptr);
        auto p = ipm::compute_fill_reducing_ordering(A);
        std::copy(p.begin(), p.end(), P);
    });
}
int lk_symbolic(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                const int32_t* P, int32_t* parent, int32_t* P_inv, int32_t* Lptr, int32_t* Lidx,
                int64_t cap, int64_t* nnz) {
    return guarded([&] {
        auto A = make_csr(m, n, rowptr, colidx, nullptr);
        ipm::SymbolicFactorization sym;
        ipm::compute_symbolic_factorization(A, std::vector<Index>(P, P + m), sym);
        *nnz = static_cast<int64_t>(sym.L_pattern.col_indices.size());
        if (parent) std::copy(sym.parent.begin(), sym.parent.end(), parent);
        if (P_inv) std::copy(sym.P_inv.begin(), sym.P_inv.end(), P_inv);
        if (Lptr) std::copy(sym.L_pattern.row_ptrs.begin(), sym.L_pattern.row_ptrs.end(), Lptr);
        if (Lidx && cap >= *nnz)
            std::copy(sym.L_pattern.col_indices.begin(), sym.L_pattern.col_indices.end(), Lidx);
    });
}
int lk_supernodal(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                  const int32_t* P, int32_t relax_small, double relax_frac, lk_kkt_stats* out,
                  int32_t* perm_out, int32_t* sn_first_out, int32_t* nsuper_out) {
    return guarded([&] {
        auto A = make_csr(m, n, rowptr, colidx, nullptr);
        std::vector<Index> p = P ? std::vector<Index>(P, P + m) : ipm::compute_fill_reducing_ordering(A);
        ipm::SupernodeOptions so;
        if (relax_small >= 0) so.relax_small = relax_small;
        if (relax_frac >= 0) so.relax_frac = relax_frac;
        ipm::SupernodalStructure S;
        ipm::compute_supernodal_structure(A, p, so, S);
        if (out) {
            std::memset(out, 0, sizeof(*out));
            out->nsuper = S.nsuper;
            out->nnz_L = S.nnz_L;
            out->nnz_L_exact = S.nnz_L_exact;
            out->panel_elems = S.sn_panel[S.nsuper];
            out->levels = S.max_level + 1;
            out->flops_exact = S.flops_exact;
            out->flops_panel = S.flops_panel;
        }
        if (perm_out) std::copy(S.perm.begin(), S.perm.end(), perm_out);
        if (sn_first_out) std::copy(S.sn_first.begin(), S.sn_first.end(), sn_first_out);
        if (nsuper_out) *nsuper_out = S.nsuper;
    });
}

int lk_plan_create(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                   const double* vals, const int32_t* P, int32_t relax_small, double relax_frac,
                   int32_t step_dag, lk_plan** out) {
    return guarded([&] {
        auto A = make_csr(m, n, rowptr, colidx, vals);
        std::vector<Index> p = P ? std::vector<Index>(P, P + m) : ipm::compute_fill_reducing_ordering(A);
        ipm::SupernodeOptions so;
        if (relax_small >= 0) so.relax_small = relax_small;
        if (relax_frac >= 0) so.relax_frac = relax_frac;
        std::unique_ptr<lk_plan> h(new lk_plan);
        if (step_dag < 0 || step_dag > 2) throw std::invalid_argument("lk_plan_create: step_dag must be 0, 1 or 2");
        h->arrays = gpu::GPUKKTCholeskySolver::export_plan(A, p, so, step_dag);
        *out = h.release();
    });
}
int lk_plan_array(lk_plan* p, const char* name, void* dst, size_t capacity, size_t* bytes) {
    return guarded([&] {
        auto it = p->arrays.find(name);
        if (it == p->arrays.end()) throw std::invalid_argument(std::string("no plan array ") + name);
        *bytes = it->second.size();
        if (dst) {
            if (capacity < it->second.size()) throw std::invalid_argument("plan array: buffer too small");
            std::memcpy(dst, it->second.data(), it->second.size());
        }
    });
}
int lk_plan_destroy(lk_plan* p) { return guarded([&] { delete p; }); }

int lk_kkt_create(lk_arena* arena, int32_t m, int32_t n, const int32_t* rowptr,
                  const int32_t* colidx, const double* vals, const int32_t* P,
                  int32_t reference_symbolic, int32_t use_graphs, lk_kkt** out) {
    return guarded([&] {
        std::unique_ptr<lk_kkt> k(new lk_kkt);
        k->A = make_csr(m, n, rowptr, colidx, vals);
        std::vector<Index> p;
        if (P) p.assign(P, P + m);
        else if (reference_symbolic) p = ipm::compute_amd_ordering(k->A);
        else p = ipm::compute_fill_reducing_ordering(k->A);
        ipm::compute_symbolic_factorization(k->A, p, k->sym);
        k->reference_layout = true;
        gpu::KKTOptions ko;
        ko.use_graphs = use_graphs != 0;
        k->arena = arena->ptr;
        k->solver.reset(new gpu::GPUKKTCholeskySolver(*k->arena, k->A, k->sym, ko));
        k->d_rhs = static_cast<double*>(k->arena->allocate(sizeof(double) * (m > 0 ? m : 1)));
        *out = k.release();
    });
}

int lk_kkt_factor(lk_kkt* k, const double* theta, int32_t on_device, int32_t mode,
                  int32_t* not_spd) {
    return guarded([&] {
        const Index n = k->solver->cols();
        *not_spd = 0;
        if (mode == 0) {
            try {
                if (on_device) k->solver->gpu_cholesky_factorize_device(theta);
                else k->solver->gpu_cholesky_factorize(k->A, std::vector<Float>(theta, theta + n));
            } catch (const std::runtime_error& e) {
                if (std::string(e.what()).find("not positive definite") == std::string::npos) throw;
                *not_spd = 1;
            }
            return;
        }
        const double* d = theta;
        std::vector<double> tmp;
        double* buf = nullptr;
        if (!on_device) {
            buf = static_cast<double*>(k->arena->allocate(sizeof(double) * (n > 0 ? n : 1)));
            if (lkd_memcpy_htod(buf, theta, sizeof(double) * n, k->solver->stream()))
                throw std::runtime_error("theta upload failed");
            d = buf;
        }
        const int r = k->solver->factor_equilibrated(
            d, mode == 2 ? gpu::FactorPrecision::FP32 : gpu::FactorPrecision::FP64);
        lkd_stream_sync(k->solver->stream());
        if (buf) k->arena->free(buf);
        *not_spd = r < 0 ? 1 : 0;
    });
}

int lk_kkt_solve(lk_kkt* k, double* rhs, int32_t on_device, double tol, int32_t max_refine,
                 lk_kkt_solve_info* info) {
    return guarded([&] {
        const Index m = k->solver->rows();
        double* d = rhs;
        if (!on_device) {
            d = k->d_rhs;
            if (lkd_memcpy_htod(d, rhs, sizeof(double) * m, k->solver->stream()))
                throw std::runtime_error("rhs upload failed");
        }
        bool ok;
        gpu::KKTSolveInfo si;
        if (max_refine < 0) {
            ok = k->solver->gpu_cholesky_solve_device(d, tol);
        } else {
            ok = k->solver->solve_cached(d, tol, &si, max_refine);
        }
        if (!on_device) {
            if (lkd_memcpy_dtoh(rhs, d, sizeof(double) * m, k->solver->stream()))
                throw std::runtime_error("rhs download failed");
        } else {
            lkd_stream_sync(k->solver->stream());
        }
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->success = ok ? 1 : 0;
            info->retries = si.retries;
            info->refinement_steps = si.refinement_steps;
            info->delta = si.delta;
            info->residual_inf = si.residual_inf;
        }
    });
}

int lk_kkt_get_L_csr(lk_kkt* k, double* vals, int64_t nnz) {
    return guarded([&] {
        const int64_t expect = static_cast<int64_t>(k->sym.L_pattern.col_indices.size()) + k->solver->rows();
        if (nnz != expect) throw std::invalid_argument("lk_kkt_get_L_csr: nnz mismatch");
        const double* d = k->solver->get_device_L_values();
        if (!d) throw std::runtime_error("lk_kkt_get_L_csr: no factor available");
        if (lkd_memcpy_dtoh(vals, d, sizeof(double) * nnz, k->solver->stream()))
            throw std::runtime_error("L download failed");
    });
}
int lk_kkt_permutation(lk_kkt* k, int32_t* perm) {
    return guarded([&] {
        const auto& p = k->solver->permutation();
        std::copy(p.begin(), p.end(), perm);
    });
}
int lk_kkt_get_stats(lk_kkt* k, lk_kkt_stats* out) {
    return guarded([&] { copy_stats(k->solver->stats(), out); });
}
int lk_kkt_stream(lk_kkt* k, void** stream) { return guarded([&] { *stream = k->solver->stream(); }); }
int lk_kkt_destroy(lk_kkt* k) {
    return guarded([&] {
        if (!k) return;
        std::shared_ptr<gpu::VRAMArena> a = k->arena;
        double* d = k->d_rhs;
        k->solver.reset();
        if (a && d) a->free(d);
        delete k;
    });
}

int lk_ipm_default_options(lk_ipm_options* o) {
    return guarded([&] {
        ipm::IpmOptions d;
        std::memset(o, 0, sizeof(*o));
        o->stop_rule = 0;
        o->max_iter = d.max_iter;
        o->tol = d.tol;
        o->gap_tol = d.gap_tol;
        o->precision = 0;
        o->mixed_max_refine = d.mixed_max_refine;
        o->mixed_fp64_spread = d.mixed_fp64_spread;
        o->ordering = 0;
        o->use_graphs = 0;
        o->verbose = 0;
        o->warm_floor = d.warm_floor;
    });
}

int lk_ruiz_scaling(int32_t m, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                    double* vals, double* b, double* c, double* R, double* C) {
    return guarded([&] {
        if (!vals || !b || !c || !R || !C) throw std::invalid_argument("lk_ruiz_scaling: null argument");
        auto A = make_csr(m, n, rowptr, colidx, vals);
        std::vector<Float> bv(b, b + m), cv(c, c + n), Rv, Cv;
        ipm::ruiz_equilibrate(A, bv, cv, Rv, Cv);
        std::copy(A.values.begin(), A.values.end(), vals);
        std::copy(bv.begin(), bv.end(), b);
        std::copy(cv.begin(), cv.end(), c);
        std::copy(Rv.begin(), Rv.end(), R);
        std::copy(Cv.begin(), Cv.end(), C);
    });
}

int lk_ipm_solve(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                 const double* vals, const double* c, const double* b, const double* lb,
                 const double* ub, const lk_ipm_options* opt, const double* warm_x,
                 const double* warm_y, const double* warm_s, lk_ipm_result* res, double* x,
                 double* y, double* s) {
    return guarded([&] {
        core::Model model;
        make_model(model, m, n, colptr, rowidx, vals, c, b, lb, ub);
        ipm::IpmOptions o = to_options(opt);
        o.warm_x = warm_x;
        o.warm_y = warm_y;
        o.warm_s = warm_s;
        ipm::MehrotraSolver solver(model, o);
        fill_result(solver.solve(), res, x, y, s);
    });
}

int lk_ipm_create(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                  const double* vals, const double* c, const double* b,
                  const lk_ipm_options* opt, lk_ipm** out) {
    return guarded([&] {
        std::unique_ptr<lk_ipm> h(new lk_ipm);
        make_model(h->model, m, n, colptr, rowidx, vals, c, b, nullptr, nullptr);
        h->solver.reset(new ipm::MehrotraSolver(h->model, to_options(opt)));
        *out = h.release();
    });
}

int lk_ipm_run(lk_ipm* h, const double* c, const double* b, const double* warm_x,
               const double* warm_y, const double* warm_s, lk_ipm_result* res, double* x,
               double* y, double* s) {
    return guarded([&] {
        const Index m = h->model.A.rows, n = h->model.A.cols;
        if (c && b)
            h->solver->update_objective_and_rhs(std::vector<Float>(c, c + n),
                                                std::vector<Float>(b, b + m));
        if (warm_x && warm_y && warm_s)
            h->solver->set_warm_start(std::vector<Float>(warm_x, warm_x + n),
                                      std::vector<Float>(warm_y, warm_y + m),
                                      std::vector<Float>(warm_s, warm_s + n));
        else
            h->solver->clear_warm_start();
        fill_result(h->solver->solve(), res, x, y, s);
    });
}

int lk_ipm_destroy(lk_ipm* h) { return guarded([&] { delete h; }); }
int lk_ipm_profile_factor(lk_ipm* h, int32_t fp32, double* ms, double* flops, double* launches) {
    return guarded([&] { h->solver->profile_factor(fp32 != 0, ms, flops, launches); });
}
int lk_ipm_launch_timing(lk_ipm* h, int32_t on) {
    return guarded([&] { h->solver->set_launch_timing(on != 0); });
}
int lk_ipm_read_launch_timing(lk_ipm* h, double* ms, double* flops, double* launches) {
    return guarded([&] { h->solver->launch_timing(ms, flops, launches); });
}
int lk_ipm_kkt_stats(lk_ipm* h, double* out12) {
    return guarded([&] { h->solver->kkt_statistics(out12); });
}

int lk_batch_default_options(lk_batch_options* o) {
    return guarded([&] {
        ipm::BatchOptions d;
        std::memset(o, 0, sizeof(*o));
        o->stop_rule = d.stop_rule;
        o->precision = d.precision;
        o->max_iter = d.max_iter;
        o->mixed_max_refine = d.mixed_max_refine;
        o->tol = d.tol;
        o->gap_tol = d.gap_tol;
        o->mixed_fp64_spread = d.mixed_fp64_spread;
        o->wg_per_cu = d.wg_per_cu;
        o->iters_per_launch = d.iters_per_launch;
        o->use_graph = d.use_graph;
        o->order = d.order;
        o->warm_floor = d.warm_floor;
        o->warm_mu = d.warm_mu;
        o->warm_max_infeas = d.warm_max_infeas;
    });
}
int lk_batch_create(int32_t count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                    const int32_t* rowidx, const double* vals, const double* b, const double* c,
                    const lk_batch_options* opt, lk_batch** out) {
    return guarded([&] {
        std::unique_ptr<lk_batch> h(new lk_batch);
        h->solver.reset(new ipm::BatchedMehrotraSolver(count, m, n, colptr, rowidx, vals, b, c,
                                                       to_batch_options(opt)));
        *out = h.release();
    });
}
int lk_batch_solve(lk_batch* h, lk_batch_summary* sum) {
    return guarded([&] { copy_summary(h->solver->solve(), sum); });
}
int lk_batch_autotune(lk_batch* h, int32_t sample, lk_batch_summary* chosen) {
    return guarded([&] { copy_summary(h->solver->autotune(sample), chosen); });
}
namespace {
void copy_batch_result(const ipm::BatchLPResult& r, lk_batch_result& o) {
    std::memset(&o, 0, sizeof(o));
    o.status = r.status;
    o.iterations = r.iterations;
    o.fail_code = r.fail_code;
    o.factorizations = r.factorizations;
    o.fp32_factorizations = r.fp32_factorizations;
    o.solves = r.solves;
    o.pcg_steps = r.pcg_steps;
    o.objective = r.objective;
    o.dual_objective = r.dual_objective;
    o.rel_gap = r.rel_gap;
    o.rel_primal = r.rel_primal;
    o.rel_dual = r.rel_dual;
    o.primal_residual = r.primal_residual;
    o.dual_residual = r.dual_residual;
    o.mu = r.mu;
    o.ms_init = r.ms_init;
    o.ms_other = r.ms_other;
    o.ms_assemble = r.ms_assemble;
    o.ms_factor = r.ms_factor;
    o.ms_solve = r.ms_solve;
    o.warm = r.warm;
    o.warm_infeas = r.warm_infeas;
}
}  // namespace

int lk_batch_solve_noalloc(int32_t count, const int32_t* m, const int32_t* n,
                           const int32_t* colptr, const int32_t* rowidx, const double* vals,
                           const double* b, const double* c, const lk_batch_options* opt,
                           double* x, double* y, double* s, lk_batch_result* results,
                           lk_batch_summary* sum) {
    return guarded([&] {
        if (count < 1 || !results) throw std::invalid_argument("lk_batch_solve_noalloc: bad arguments");
        std::vector<ipm::BatchLPResult> r(static_cast<size_t>(count));
        const ipm::BatchSummary su = ipm::solve_batch_noalloc(count, m, n, colptr, rowidx, vals, b, c,
                                                              to_batch_options(opt), x, y, s, r.data());
        for (int32_t i = 0; i < count; ++i) copy_batch_result(r[i], results[i]);
        copy_summary(su, sum);
    });
}

int lk_batch_results(lk_batch* h, lk_batch_result* out) {
    return guarded([&] {
        const auto& r = h->solver->results();
        for (size_t i = 0; i < r.size(); ++i) copy_batch_result(r[i], out[i]);
    });
}
int lk_batch_solution(lk_batch* h, int32_t i, double* x, double* y, double* s) {
    return guarded([&] { h->solver->solution(i, x, y, s); });
}
int lk_batch_crossover(lk_batch* h, int32_t* basis, double* x, int32_t* status, int32_t* pivots) {
    return guarded([&] { h->solver->crossover(basis, x, status, pivots); });
}
int lk_crossover(int32_t count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                 const int32_t* rowidx, const double* vals, const double* b, const double* c,
                 const double* x, int32_t* basis, double* xv, int32_t* status, int32_t* pivots) {
    return guarded([&] {
        ipm::crossover_points(count, m, n, colptr, rowidx, vals, b, c, x, basis, xv, status, pivots);
    });
}
int lk_simplex(int32_t count, const int32_t* m, const int32_t* n, const int32_t* colptr,
               const int32_t* rowidx, const double* vals, const double* b, const double* c,
               int32_t max_iter, int32_t* basis, double* x, double* y, int32_t* status,
               int32_t* iterations, double* objective) {
    return guarded([&] {
        simplex::solve_batch(count, m, n, colptr, rowidx, vals, b, c, max_iter, basis, x, y, status,
                             iterations, objective);
    });
}
int lk_batch_destroy(lk_batch* h) { return guarded([&] { delete h; }); }
int lk_batch_set_warm_start(lk_batch* h, const double* x, const double* y, const double* s,
                            const int32_t* flags) {
    return guarded([&] { h->solver->set_warm_start(x, y, s, flags); });
}

int lk_mlp_create(int32_t d_in, int32_t h1, int32_t h2, int32_t d_out, int32_t max_batch,
                  uint64_t seed, lk_mlp** out) {
    return guarded([&] {
        std::unique_ptr<lk_mlp> h(new lk_mlp);
        h->net.reset(new learn::WarmStartMLP(d_in, h1, h2, d_out, max_batch, seed));
        *out = h.release();
    });
}
int lk_mlp_param_count(lk_mlp* h, int64_t* count) {
    return guarded([&] { *count = h->net->param_count(); });
}
int lk_mlp_set_params(lk_mlp* h, const float* params) {
    return guarded([&] { h->net->set_params(params); });
}
int lk_mlp_get_params(lk_mlp* h, float* params) {
    return guarded([&] { h->net->get_params(params); });
}
int lk_mlp_forward(lk_mlp* h, const float* u, int32_t batch, float* y) {
    return guarded([&] { h->net->forward(u, batch, y); });
}
int lk_mlp_gradients(lk_mlp* h, const float* u, const float* t, int32_t batch, float* grad,
                     double* loss) {
    return guarded([&] {
        const double l = h->net->gradients(u, t, batch, grad);
        if (loss) *loss = l;
    });
}
int lk_mlp_set_dataset(lk_mlp* h, const float* u, const float* t, int32_t n) {
    return guarded([&] { h->net->set_dataset(u, t, n); });
}
int lk_mlp_train(lk_mlp* h, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                 double* epoch_loss) {
    return guarded([&] {
        const auto l = h->net->train(epochs, batch, lr, seed);
        if (epoch_loss) std::copy(l.begin(), l.end(), epoch_loss);
    });
}
int lk_mlp_destroy(lk_mlp* h) { return guarded([&] { delete h; }); }

int lk_verify_optimal(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                      const double* vals, const double* c, const double* b, const double* lb,
                      const double* ub, int32_t maximize, const double* x, const double* pi,
                      double tol, double* primal_residual, int32_t* pass) {
    return guarded([&] {
        core::Model model;
        model.set_from_csc(m, n, colptr, rowidx, vals, c, b, lb, ub);
        model.sense = maximize ? OptimizationSense::Maximize : OptimizationSense::Minimize;
        double res = 0;
        *pass = verifier::verify_optimal(model, std::vector<Float>(x, x + n),
                                         std::vector<Float>(pi, pi + m), &res, tol)
                    ? 1 : 0;
        if (primal_residual) *primal_residual = res;
    });
}

struct lk_mps {
    core::Model model;
};

int lk_mps_open(const char* path, lk_mps** out, int32_t* m, int32_t* n, int64_t* nnz) {
    return guarded([&] {
        if (!path || !out) throw std::invalid_argument("lk_mps_open: null argument");
        std::unique_ptr<lk_mps> h(new lk_mps{parsers::read_mps(path)});
        if (m) *m = h->model.A.rows;
        if (n) *n = h->model.A.cols;
        if (nnz) *nnz = static_cast<int64_t>(h->model.A.values.size());
        *out = h.release();
    });
}

int lk_mps_get(lk_mps* h, int32_t* colptr, int32_t* rowidx, double* vals, double* obj,
               double* rhs, double* lb, double* ub, int32_t* vtype) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("lk_mps_get: null handle");
        const core::Model& md = h->model;
        const auto& A = md.A;
        if (colptr) std::copy(A.col_ptrs.begin(), A.col_ptrs.end(), colptr);
        if (rowidx) std::copy(A.row_indices.begin(), A.row_indices.end(), rowidx);
        if (vals) std::copy(A.values.begin(), A.values.end(), vals);
        if (obj) std::copy(md.obj.begin(), md.obj.end(), obj);
        if (rhs) std::copy(md.rhs.begin(), md.rhs.end(), rhs);
        if (lb) std::copy(md.lb.begin(), md.lb.end(), lb);
        if (ub) std::copy(md.ub.begin(), md.ub.end(), ub);
        if (vtype)
            for (size_t j = 0; j < md.vtype.size(); ++j) vtype[j] = static_cast<int32_t>(md.vtype[j]);
    });
}

int lk_mps_close(lk_mps* h) {
    return guarded([&] { delete h; });
}

}  // extern "C"

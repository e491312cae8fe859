// Host side of the batched interior-point solver (include/ipm/batch.hpp): packs the LPs into
// the device layout of lkb_lp (CSR copy and CSR -> CSC position map built here, in parallel),
// uploads everything once into one plan-sized arena, and drives the persistent launch -- one
// launch, or a replayed hipGraph of launches that each advance every unfinished LP by a few
// iterations -- reading back one counter per replay.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include "gpu/vram_arena.hpp"
#include "ipm/batch.hpp"
#include "lakshya_device.h"

namespace sankhya {
namespace ipm {

namespace {
void check(int e, const char* what) {
    if (e != 0)
        throw std::runtime_error(std::string("HIP call failed (") + what + "): " +
                                 lkd_error_string(e));
}
template <typename F>
void parallel_for(int n, F&& f) {
    const int nt = std::max(1, std::min<int>(16, static_cast<int>(std::thread::hardware_concurrency())));
    if (n < 64 || nt == 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) f(i);
        });
    for (auto& x : th) x.join();
}
}  // namespace

struct BatchedMehrotraSolver::Impl {
    BatchOptions opt;
    int count = 0;
    std::vector<lkb_lp> lps;
    std::vector<int32_t> order;       // largest m first (or input order)
    int64_t n_idx = 0, n_val = 0, n_rhs = 0, n_work = 0, n_fac = 0, n_xn = 0, n_ym = 0;
    std::unique_ptr<gpu::VRAMArena> arena;
    void* stream = nullptr;
    lkb_lp* d_lps = nullptr;
    int32_t *d_order = nullptr, *d_idx = nullptr;
    double *d_val0 = nullptr, *d_valw = nullptr, *d_rhs = nullptr, *d_work = nullptr;
    double *d_fac = nullptr, *d_x = nullptr, *d_y = nullptr, *d_s = nullptr;
    lkb_state* d_state = nullptr;
    uint32_t* d_ctl = nullptr;
    int cus = 256;
    double clock_khz = 100000.0;     // device wall clock ticks per ms
    std::vector<BatchLPResult> res;
    std::vector<double> hx, hy, hs;
    bool have_solution = false;
    // cached graph of `graph_len` launches for one (grid, iterations, count) shape
    void* graph = nullptr;
    int graph_grid = 0, graph_iters = 0, graph_len = 0, graph_count = -1;
    const int32_t* graph_order = nullptr;

    Impl(int cnt, const int32_t* m, const int32_t* n, const int32_t* colptr, const int32_t* rowidx,
         const double* vals, const double* b, const double* c, const BatchOptions& o)
        : opt(o), count(cnt) {
        if (cnt < 1) throw std::invalid_argument("BatchedMehrotraSolver: empty batch");
        if (!m || !n || !colptr || !rowidx || !vals || !b || !c)
            throw std::invalid_argument("BatchedMehrotraSolver: null argument");
        if (opt.max_iter < 1 || opt.stop_rule < 0 || opt.stop_rule > 1 || opt.precision < 0 ||
            opt.precision > 1)
            throw std::invalid_argument("BatchedMehrotraSolver: bad options");
        lps.resize(cnt);
        std::vector<int64_t> cp_off(cnt), nz_off(cnt);
        int64_t cpo = 0, nzo = 0, mo = 0, no = 0;
        for (int i = 0; i < cnt; ++i) {
            if (m[i] < 1 || m[i] > LKB_MAX_M || n[i] < 1)
                throw std::invalid_argument("BatchedMehrotraSolver: LP " + std::to_string(i) +
                                            " has m outside [1, " + std::to_string(LKB_MAX_M) +
                                            "] or no columns");
            const int32_t* cp = colptr + cpo;
            if (cp[0] != 0) throw std::invalid_argument("BatchedMehrotraSolver: column pointers must start at 0");
            for (int j = 0; j < n[i]; ++j)
                if (cp[j + 1] < cp[j]) throw std::invalid_argument("BatchedMehrotraSolver: column pointers not monotone");
            const int64_t nz = cp[n[i]];
            lkb_lp& d = lps[i];
            d.m = m[i];
            d.n = n[i];
            d.nnz = static_cast<int32_t>(nz);
            d.ld = (m[i] + 3) & ~3;
            d.a = n_idx;
            n_idx += 3 * nz + m[i] + n[i] + 2;
            d.v = n_val;
            n_val += nz;
            d.bc = n_rhs;
            n_rhs += m[i] + n[i];
            d.w = n_work;
            n_work += static_cast<int64_t>(LKB_NVN) * n[i] + static_cast<int64_t>(LKB_NVM) * m[i] + LKB_STAGE;
            n_work = (n_work + 31) & ~int64_t(31);
            d.f = n_fac;
            n_fac += static_cast<int64_t>(d.ld) * m[i];
            n_fac = (n_fac + 31) & ~int64_t(31);
            d.on = no;
            d.om = mo;
            cp_off[i] = cpo;
            nz_off[i] = nzo;
            cpo += n[i] + 1;
            nzo += nz;
            no += n[i];
            mo += m[i];
        }
        n_xn = no;
        n_ym = mo;

        // host images: index block per LP, original values, b / c
        std::vector<int32_t> idx(static_cast<size_t>(n_idx));
        std::vector<double> rhs(static_cast<size_t>(n_rhs));
        std::vector<int> bad(cnt, 0);
        parallel_for(cnt, [&](int i) {
            const lkb_lp& d = lps[i];
            const int32_t* cp = colptr + cp_off[i];
            const int32_t* ri = rowidx + nz_off[i];
            const int mm = d.m, nn = d.n, nz = d.nnz;
            int32_t* rp = idx.data() + d.a;
            int32_t* ci = rp + mm + 1;
            int32_t* pos = ci + nz;
            int32_t* ccp = pos + nz;
            int32_t* cri = ccp + nn + 1;
            for (int r = 0; r <= mm; ++r) rp[r] = 0;
            for (int j = 0; j < nn; ++j)
                for (int q = cp[j]; q < cp[j + 1]; ++q) {
                    const int r = ri[q];
                    if (r < 0 || r >= mm || (q > cp[j] && r <= ri[q - 1])) { bad[i] = 1; return; }
                    rp[r + 1]++;
                }
            for (int r = 0; r < mm; ++r) rp[r + 1] += rp[r];
            std::vector<int32_t> next(rp, rp + mm);
            // columns ascending, so every CSR row comes out with ascending columns
            for (int j = 0; j < nn; ++j)
                for (int q = cp[j]; q < cp[j + 1]; ++q) {
                    const int e = next[ri[q]]++;
                    ci[e] = j;
                    pos[e] = q;
                }
            std::copy(cp, cp + nn + 1, ccp);
            std::copy(ri, ri + nz, cri);
            std::copy(b + d.om, b + d.om + mm, rhs.data() + d.bc);
            std::copy(c + d.on, c + d.on + nn, rhs.data() + d.bc + mm);
        });
        for (int i = 0; i < cnt; ++i)
            if (bad[i])
                throw std::invalid_argument("BatchedMehrotraSolver: LP " + std::to_string(i) +
                                            ": row indices must increase within each column");
        order.resize(cnt);
        std::iota(order.begin(), order.end(), 0);
        if (opt.order == 0)
            std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return lps[a].m > lps[b2].m; });

        check(lkd_device_cus(&cus), "device attributes");
        if (cus < 1) cus = 256;
        int khz = 0;
        check(lkd_wallclock_khz(&khz), "device attributes");
        clock_khz = khz > 0 ? khz : 100000.0;
        const size_t al = gpu::VRAMArena::kAlignment;
        auto rnd = [&](size_t bytes) { return (bytes + al - 1) / al * al; };
        const size_t bytes = rnd(sizeof(lkb_lp) * cnt) + rnd(sizeof(int32_t) * cnt) +
                             rnd(sizeof(int32_t) * n_idx) + rnd(sizeof(double) * n_val) +
                             rnd(sizeof(double) * 2 * n_val) + rnd(sizeof(double) * n_rhs) +
                             rnd(sizeof(double) * n_work) + rnd(sizeof(double) * n_fac) +
                             rnd(sizeof(lkb_state) * cnt) + 2 * rnd(sizeof(double) * n_xn) +
                             rnd(sizeof(double) * n_ym) + rnd(64) + (1u << 20);
        arena.reset(new gpu::VRAMArena(bytes));
        check(lkd_stream_create(&stream), "stream");
        auto up = [&](const void* src, size_t nb) {
            void* p = arena->allocate_uninitialized(nb);
            check(lkd_memcpy_htod(p, src, nb, stream), "upload");
            return p;
        };
        d_lps = static_cast<lkb_lp*>(up(lps.data(), sizeof(lkb_lp) * cnt));
        d_order = static_cast<int32_t*>(up(order.data(), sizeof(int32_t) * cnt));
        d_idx = static_cast<int32_t*>(up(idx.data(), sizeof(int32_t) * n_idx));
        d_val0 = static_cast<double*>(up(vals, sizeof(double) * n_val));
        d_valw = static_cast<double*>(arena->allocate_uninitialized(sizeof(double) * 2 * n_val));
        d_rhs = static_cast<double*>(up(rhs.data(), sizeof(double) * n_rhs));
        d_work = static_cast<double*>(arena->allocate_uninitialized(sizeof(double) * n_work));
        d_fac = static_cast<double*>(arena->allocate_uninitialized(sizeof(double) * n_fac));
        d_state = static_cast<lkb_state*>(arena->allocate(sizeof(lkb_state) * cnt));
        d_x = static_cast<double*>(arena->allocate(sizeof(double) * n_xn));
        d_s = static_cast<double*>(arena->allocate(sizeof(double) * n_xn));
        d_y = static_cast<double*>(arena->allocate(sizeof(double) * n_ym));
        d_ctl = static_cast<uint32_t*>(arena->allocate(64));
        check(lkd_stream_sync(stream), "upload");
        res.assign(cnt, BatchLPResult{});
    }

    ~Impl() {
        if (stream) lkd_stream_sync(stream);
        if (graph) lkd_graph_destroy(graph);
        arena.reset();
        if (stream) lkd_stream_destroy(stream);
    }

    lkb_args args(const int32_t* ord, int cnt, int iters) const {
        lkb_args a;
        std::memset(&a, 0, sizeof(a));
        a.count = cnt;
        a.stop_rule = opt.stop_rule;
        a.precision = opt.precision;
        a.max_iter = opt.max_iter;
        a.iters_per_visit = iters;
        a.mixed_max_refine = opt.mixed_max_refine;
        a.tol = opt.tol;
        a.gap_tol = opt.gap_tol;
        a.mixed_spread = opt.mixed_fp64_spread;
        a.lps = d_lps;
        a.order = ord;
        a.idx = d_idx;
        a.val0 = d_val0;
        a.valw = d_valw;
        a.rhs = d_rhs;
        a.work = d_work;
        a.fac = d_fac;
        a.state = d_state;
        a.x_out = d_x;
        a.y_out = d_y;
        a.s_out = d_s;
        a.ctl = d_ctl;
        return a;
    }

    // Solve LPs ord[0 .. cnt) with `wg` workgroups per CU and `iters` iterations per visit.
    BatchSummary run(const int32_t* ord, int cnt, int wg, int iters, bool use_graph) {
        BatchSummary sum;
        sum.count = cnt;
        sum.wg_per_cu = wg;
        sum.iters_per_launch = iters;
        sum.grid = wg * cus;
        const lkb_args a = args(ord, cnt, iters);
        const int max_launches = (opt.max_iter + 1 + iters - 1) / iters + 1;
        void *e0 = nullptr, *e1 = nullptr;
        check(lkd_event_create(&e0), "event");
        check(lkd_event_create(&e1), "event");
        check(lkd_event_record(e0, stream), "event");
        check(lkd_batch_reset(d_ctl, d_state, ord, cnt, 1, stream), "reset");
        uint32_t done = 0;
        if (iters > opt.max_iter) {
            // every LP finishes in its one visit: a single persistent launch
            check(lkd_batch_solve(&a, sum.grid, stream), "batch launch");
            sum.launches = 1;
        } else {
            // launches of `iters` iterations each, replayed (from a graph) until all finished
            const int len = std::min(4, max_launches);
            if (use_graph && !(graph && graph_grid == sum.grid && graph_iters == iters &&
                               graph_len == len && graph_count == cnt && graph_order == ord)) {
                if (graph) check(lkd_graph_destroy(graph), "graph");
                graph = nullptr;
                check(lkd_graph_begin(stream), "capture");
                for (int k = 0; k < len; ++k) {
                    check(lkd_batch_reset(d_ctl, d_state, ord, cnt, 0, stream), "reset");
                    check(lkd_batch_solve(&a, sum.grid, stream), "batch launch");
                }
                check(lkd_graph_end(stream, &graph), "capture");
                graph_grid = sum.grid;
                graph_iters = iters;
                graph_len = len;
                graph_count = cnt;
                graph_order = ord;
            }
            while (sum.launches < max_launches) {
                if (use_graph) {
                    check(lkd_graph_launch(graph, stream), "graph launch");
                    sum.graph_replays++;
                    sum.launches += len;
                } else {
                    check(lkd_batch_reset(d_ctl, d_state, ord, cnt, 0, stream), "reset");
                    check(lkd_batch_solve(&a, sum.grid, stream), "batch launch");
                    sum.launches++;
                }
                check(lkd_memcpy_dtoh(&done, d_ctl + 1, sizeof(uint32_t), stream), "progress");
                if (static_cast<int>(done) >= cnt) break;
            }
        }
        check(lkd_event_record(e1, stream), "event");
        float ms = 0;
        check(lkd_event_elapsed_ms(&ms, e0, e1), "elapsed");
        sum.ms = ms;
        lkd_event_destroy(e0);
        lkd_event_destroy(e1);
        check(lkd_memcpy_dtoh(&done, d_ctl + 1, sizeof(uint32_t), stream), "progress");
        if (static_cast<int>(done) < cnt)
            throw std::runtime_error("BatchedMehrotraSolver: " + std::to_string(cnt - static_cast<int>(done)) +
                                     " LPs did not finish within the launch budget");
        std::vector<lkb_state> st(count);
        check(lkd_memcpy_dtoh(st.data(), d_state, sizeof(lkb_state) * count, stream), "states");
        std::vector<int32_t> ho(cnt);
        if (ord == d_order) ho = order;
        else check(lkd_memcpy_dtoh(ho.data(), ord, sizeof(int32_t) * cnt, stream), "order");
        for (int t = 0; t < cnt; ++t) {
            const int i = ho[t];
            const lkb_state& s = st[i];
            BatchLPResult& r = res[i];
            r.status = s.status;
            r.iterations = s.iterations;
            r.fail_code = s.fail_code;
            r.factorizations = s.factorizations;
            r.fp32_factorizations = s.fp32_factorizations;
            r.solves = s.solves;
            r.pcg_steps = s.pcg_steps;
            r.objective = s.objective;
            r.dual_objective = s.dual_objective;
            r.rel_gap = s.rel_gap;
            r.rel_primal = s.rel_primal;
            r.rel_dual = s.rel_dual;
            r.primal_residual = s.primal_residual;
            r.dual_residual = s.dual_residual;
            r.mu = s.mu;
            r.ms_init = s.t_init / clock_khz;
            r.ms_other = s.t_other / clock_khz;
            r.ms_assemble = s.t_assemble / clock_khz;
            r.ms_factor = s.t_factor / clock_khz;
            r.ms_solve = s.t_solve / clock_khz;
            if (s.status == 0) sum.optimal++;
            const double mm = lps[i].m;
            sum.flops_chol += s.factorizations * mm * mm * mm / 3.0;
        }
        have_solution = false;
        return sum;
    }

    int default_wg() const { return opt.wg_per_cu > 0 ? opt.wg_per_cu : 2; }
    int default_iters() const { return opt.iters_per_launch > 0 ? opt.iters_per_launch : opt.max_iter + 1; }

    BatchSummary solve() {
        return run(d_order, count, default_wg(), default_iters(), opt.use_graph != 0);
    }

    BatchSummary autotune(int sample) {
        sample = std::max(1, std::min(sample, count));
        // every (count / sample)-th LP of the size order: the sample spans the size range
        std::vector<int32_t> so(sample);
        for (int t = 0; t < sample; ++t)
            so[t] = order[static_cast<size_t>(static_cast<int64_t>(t) * count / sample)];
        int32_t* d_so = static_cast<int32_t*>(arena->allocate_uninitialized(sizeof(int32_t) * sample));
        check(lkd_memcpy_htod(d_so, so.data(), sizeof(int32_t) * sample, stream), "sample");
        const int wgs[2] = {1, 2};
        const int its[3] = {1, 4, opt.max_iter + 1};
        BatchSummary best;
        best.ms = 1e300;
        try {
            for (int wg : wgs)
                for (int it : its) {
                    run(d_so, sample, wg, it, opt.use_graph != 0);    // warm (graph capture)
                    BatchSummary s = run(d_so, sample, wg, it, opt.use_graph != 0);
                    if (s.ms < best.ms) best = s;
                }
        } catch (...) {
            if (graph) { lkd_graph_destroy(graph); graph = nullptr; }
            arena->free(d_so);
            throw;
        }
        if (graph) { check(lkd_graph_destroy(graph), "graph"); graph = nullptr; }
        arena->free(d_so);
        opt.wg_per_cu = best.wg_per_cu;
        opt.iters_per_launch = best.iters_per_launch;
        return best;
    }

    void fetch_solution() {
        if (have_solution) return;
        hx.resize(static_cast<size_t>(n_xn));
        hs.resize(static_cast<size_t>(n_xn));
        hy.resize(static_cast<size_t>(n_ym));
        check(lkd_memcpy_dtoh(hx.data(), d_x, sizeof(double) * n_xn, stream), "x");
        check(lkd_memcpy_dtoh(hs.data(), d_s, sizeof(double) * n_xn, stream), "s");
        check(lkd_memcpy_dtoh(hy.data(), d_y, sizeof(double) * n_ym, stream), "y");
        have_solution = true;
    }
};

BatchedMehrotraSolver::BatchedMehrotraSolver(int count, const int32_t* m, const int32_t* n,
                                             const int32_t* colptr, const int32_t* rowidx,
                                             const double* vals, const double* b, const double* c,
                                             const BatchOptions& opt)
    : impl_(new Impl(count, m, n, colptr, rowidx, vals, b, c, opt)) {}
BatchedMehrotraSolver::~BatchedMehrotraSolver() { delete impl_; }
BatchSummary BatchedMehrotraSolver::solve() { return impl_->solve(); }
BatchSummary BatchedMehrotraSolver::autotune(int sample) { return impl_->autotune(sample); }
const std::vector<BatchLPResult>& BatchedMehrotraSolver::results() const { return impl_->res; }
int BatchedMehrotraSolver::count() const { return impl_->count; }
void BatchedMehrotraSolver::solution(int i, double* x, double* y, double* s) const {
    if (i < 0 || i >= impl_->count) throw std::invalid_argument("solution: LP index out of range");
    impl_->fetch_solution();
    const lkb_lp& d = impl_->lps[i];
    if (x) std::copy(impl_->hx.begin() + d.on, impl_->hx.begin() + d.on + d.n, x);
    if (s) std::copy(impl_->hs.begin() + d.on, impl_->hs.begin() + d.on + d.n, s);
    if (y) std::copy(impl_->hy.begin() + d.om, impl_->hy.begin() + d.om + d.m, y);
}

}  // namespace ipm
}  // namespace sankhya

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
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include "gpu/vram_arena.hpp"
#include "ipm/batch.hpp"
#include "ipm/crossover.hpp"
#include "simplex/primal.hpp"
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
// Host images of a batch in the device layout (lkb_lp): descriptors, the index block of every
// LP (CSR built here from the CSC input, with the CSC position of each CSR entry), b / c, and
// the visit order.  Validates the input.
struct BatchImage {
    int count = 0;
    std::vector<lkb_lp> lps;
    std::vector<int32_t> order;
    std::vector<int32_t> idx;
    std::vector<double> rhs;
    const double* vals = nullptr;     // the caller's CSC values (uploaded as they are)
    int64_t n_idx = 0, n_val = 0, n_rhs = 0, n_work = 0, n_fac = 0, n_xn = 0, n_ym = 0;

    static size_t rnd(size_t bytes) {
        const size_t al = gpu::VRAMArena::kAlignment;
        return (bytes + al - 1) / al * al;
    }
    // device bytes of all buffers, each rounded to the 256-byte granule
    size_t device_bytes() const {
        return rnd(sizeof(lkb_lp) * count) + rnd(sizeof(int32_t) * count) +
               rnd(sizeof(int32_t) * n_idx) + rnd(sizeof(double) * n_val) +
               rnd(sizeof(double) * 2 * n_val) + rnd(sizeof(double) * n_rhs) +
               rnd(sizeof(double) * n_work) + rnd(sizeof(double) * n_fac) +
               rnd(sizeof(lkb_state) * count) + 2 * rnd(sizeof(double) * n_xn) +
               rnd(sizeof(double) * n_ym) + rnd(64);
    }
    // the optional warm-start buffers (x0, s0, y0, flags)
    size_t warm_bytes() const {
        return 2 * rnd(sizeof(double) * n_xn) + rnd(sizeof(double) * n_ym) + rnd(sizeof(int32_t) * count);
    }
    // the optional crossover outputs (basis, values, status + pivots)
    size_t crossover_bytes() const {
        return rnd(sizeof(int32_t) * n_ym) + rnd(sizeof(double) * n_xn) + rnd(sizeof(int32_t) * 2 * count);
    }
};

BatchImage pack_batch(int cnt, const int32_t* m, const int32_t* n, const int32_t* colptr,
                      const int32_t* rowidx, const double* vals, const double* b, const double* c,
                      int order_rule) {
    if (cnt < 1) throw std::invalid_argument("BatchedMehrotraSolver: empty batch");
    if (!m || !n || !colptr || !rowidx || !vals || !b || !c)
        throw std::invalid_argument("BatchedMehrotraSolver: null argument");
    BatchImage img;
    img.count = cnt;
    img.vals = vals;
    img.lps.resize(cnt);
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
        lkb_lp& d = img.lps[i];
        d.m = m[i];
        d.n = n[i];
        d.nnz = static_cast<int32_t>(nz);
        d.ld = (m[i] + 3) & ~3;
        d.a = img.n_idx;
        img.n_idx += 3 * nz + m[i] + n[i] + 2;
        d.v = img.n_val;
        img.n_val += nz;
        d.bc = img.n_rhs;
        img.n_rhs += m[i] + n[i];
        d.w = img.n_work;
        img.n_work += static_cast<int64_t>(LKB_NVN) * n[i] + static_cast<int64_t>(LKB_NVM) * m[i] + LKB_STAGE;
        img.n_work = (img.n_work + 31) & ~int64_t(31);
        d.f = img.n_fac;
        img.n_fac += static_cast<int64_t>(d.ld) * m[i];
        img.n_fac = (img.n_fac + 31) & ~int64_t(31);
        d.on = no;
        d.om = mo;
        cp_off[i] = cpo;
        nz_off[i] = nzo;
        cpo += n[i] + 1;
        nzo += nz;
        no += n[i];
        mo += m[i];
    }
    img.n_xn = no;
    img.n_ym = mo;
    img.idx.resize(static_cast<size_t>(img.n_idx));
    img.rhs.resize(static_cast<size_t>(img.n_rhs));
    std::vector<int> bad(cnt, 0);
    parallel_for(cnt, [&](int i) {
        const lkb_lp& d = img.lps[i];
        const int32_t* cp = colptr + cp_off[i];
        const int32_t* ri = rowidx + nz_off[i];
        const int mm = d.m, nn = d.n, nz = d.nnz;
        int32_t* rp = img.idx.data() + d.a;
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
        std::copy(b + d.om, b + d.om + mm, img.rhs.data() + d.bc);
        std::copy(c + d.on, c + d.on + nn, img.rhs.data() + d.bc + mm);
    });
    for (int i = 0; i < cnt; ++i)
        if (bad[i])
            throw std::invalid_argument("BatchedMehrotraSolver: LP " + std::to_string(i) +
                                        ": row indices must increase within each column");
    img.order.resize(cnt);
    std::iota(img.order.begin(), img.order.end(), 0);
    if (order_rule == 0)
        std::stable_sort(img.order.begin(), img.order.end(),
                         [&](int a, int b2) { return img.lps[a].m > img.lps[b2].m; });
    return img;
}

// Device buffers of a batch, carved by `take(bytes)` (arena or fixed scratch), inputs uploaded
// on `stream`.
struct BatchDevice {
    lkb_lp* lps = nullptr;
    int32_t *order = nullptr, *idx = nullptr;
    double *val0 = nullptr, *valw = nullptr, *rhs = nullptr, *work = nullptr;
    double *fac = nullptr, *x = nullptr, *y = nullptr, *s = nullptr;
    lkb_state* state = nullptr;
    uint32_t* ctl = nullptr;
};

template <typename Take>
BatchDevice place_batch(const BatchImage& img, Take&& take, void* stream) {
    BatchDevice d;
    auto up = [&](const void* src, size_t nb) {
        void* p = take(nb);
        check(lkd_memcpy_htod(p, src, nb, stream), "upload");
        return p;
    };
    const int cnt = img.count;
    d.lps = static_cast<lkb_lp*>(up(img.lps.data(), sizeof(lkb_lp) * cnt));
    d.order = static_cast<int32_t*>(up(img.order.data(), sizeof(int32_t) * cnt));
    d.idx = static_cast<int32_t*>(up(img.idx.data(), sizeof(int32_t) * img.n_idx));
    d.val0 = static_cast<double*>(up(img.vals, sizeof(double) * img.n_val));
    d.valw = static_cast<double*>(take(sizeof(double) * 2 * img.n_val));
    d.rhs = static_cast<double*>(up(img.rhs.data(), sizeof(double) * img.n_rhs));
    d.work = static_cast<double*>(take(sizeof(double) * img.n_work));
    d.fac = static_cast<double*>(take(sizeof(double) * img.n_fac));
    d.state = static_cast<lkb_state*>(take(sizeof(lkb_state) * cnt));
    d.x = static_cast<double*>(take(sizeof(double) * img.n_xn));
    d.s = static_cast<double*>(take(sizeof(double) * img.n_xn));
    d.y = static_cast<double*>(take(sizeof(double) * img.n_ym));
    d.ctl = static_cast<uint32_t*>(take(64));
    return d;
}

lkb_args make_args(const BatchOptions& opt, const BatchDevice& d, const int32_t* ord, int cnt,
                   int iters) {
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
    a.lps = d.lps;
    a.order = ord;
    a.idx = d.idx;
    a.val0 = d.val0;
    a.valw = d.valw;
    a.rhs = d.rhs;
    a.work = d.work;
    a.fac = d.fac;
    a.state = d.state;
    a.x_out = d.x;
    a.y_out = d.y;
    a.s_out = d.s;
    a.ctl = d.ctl;
    return a;
}

void check_options(const BatchOptions& opt) {
    if (opt.max_iter < 1 || opt.stop_rule < 0 || opt.stop_rule > 1 || opt.precision < 0 ||
        opt.precision > 1)
        throw std::invalid_argument("BatchedMehrotraSolver: bad options");
}

BatchLPResult to_result(const lkb_state& s, double clock_khz) {
    BatchLPResult r;
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
    r.warm = s.warm;
    r.warm_infeas = s.warm_infeas;
    return r;
}

double device_clock_khz() {
    int khz = 0;
    check(lkd_wallclock_khz(&khz), "device attributes");
    return khz > 0 ? khz : 100000.0;
}

int device_cus() {
    int cus = 0;
    check(lkd_device_cus(&cus), "device attributes");
    return cus > 0 ? cus : 256;
}

}  // namespace

struct BatchedMehrotraSolver::Impl {
    BatchOptions opt;
    int count = 0;
    BatchImage img;
    std::unique_ptr<gpu::VRAMArena> arena;
    void* stream = nullptr;
    BatchDevice dev;
    int cus = 256;
    double clock_khz = 100000.0;     // device wall clock ticks per ms
    std::vector<BatchLPResult> res;
    std::vector<double> hx, hy, hs;
    bool have_solution = false;
    // warm start (original space) and per-LP flags, allocated on first use
    double *d_x0 = nullptr, *d_y0 = nullptr, *d_s0 = nullptr;
    int32_t* d_wflag = nullptr;
    // crossover outputs, allocated on first use
    int32_t *d_basis = nullptr, *d_xstat = nullptr;
    double* d_xv = nullptr;
    bool solved = false;
    // cached graph of `graph_len` launches for one (grid, iterations, count) shape
    void* graph = nullptr;
    int graph_grid = 0, graph_iters = 0, graph_len = 0, graph_count = -1;
    const int32_t* graph_order = nullptr;

    Impl(int cnt, const int32_t* m, const int32_t* n, const int32_t* colptr, const int32_t* rowidx,
         const double* vals, const double* b, const double* c, const BatchOptions& o)
        : opt(o), count(cnt) {
        check_options(opt);
        img = pack_batch(cnt, m, n, colptr, rowidx, vals, b, c, opt.order);
        cus = device_cus();
        clock_khz = device_clock_khz();
        arena.reset(new gpu::VRAMArena(img.device_bytes() + img.warm_bytes() + img.crossover_bytes() +
                                       (1u << 20)));
        check(lkd_stream_create(&stream), "stream");
        dev = place_batch(img, [&](size_t nb) { return arena->allocate_uninitialized(nb); }, stream);
        check(lkd_stream_sync(stream), "upload");
        res.assign(cnt, BatchLPResult{});
    }

    ~Impl() {
        if (stream) lkd_stream_sync(stream);
        if (graph) lkd_graph_destroy(graph);
        arena.reset();
        if (stream) lkd_stream_destroy(stream);
    }

    // Solve LPs ord[0 .. cnt) with `wg` workgroups per CU and `iters` iterations per visit.
    BatchSummary run(const int32_t* ord, int cnt, int wg, int iters, bool use_graph) {
        BatchSummary sum;
        sum.count = cnt;
        sum.wg_per_cu = wg;
        sum.iters_per_launch = iters;
        sum.grid = wg * cus;
        lkb_args a = make_args(opt, dev, ord, cnt, iters);
        a.x0 = d_x0;
        a.y0 = d_y0;
        a.s0 = d_s0;
        a.warm_flag = d_wflag;
        a.warm_floor = opt.warm_floor;
        a.warm_mu = opt.warm_mu;
        a.warm_max_infeas = opt.warm_max_infeas;
        const int max_launches = (opt.max_iter + 1 + iters - 1) / iters + 1;
        void *e0 = nullptr, *e1 = nullptr;
        check(lkd_event_create(&e0), "event");
        check(lkd_event_create(&e1), "event");
        check(lkd_event_record(e0, stream), "event");
        check(lkd_batch_reset(dev.ctl, dev.state, ord, cnt, 1, stream), "reset");
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
                    check(lkd_batch_reset(dev.ctl, dev.state, ord, cnt, 0, stream), "reset");
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
                    check(lkd_batch_reset(dev.ctl, dev.state, ord, cnt, 0, stream), "reset");
                    check(lkd_batch_solve(&a, sum.grid, stream), "batch launch");
                    sum.launches++;
                }
                check(lkd_memcpy_dtoh(&done, dev.ctl + 1, sizeof(uint32_t), stream), "progress");
                if (static_cast<int>(done) >= cnt) break;
            }
        }
        check(lkd_event_record(e1, stream), "event");
        float ms = 0;
        check(lkd_event_elapsed_ms(&ms, e0, e1), "elapsed");
        sum.ms = ms;
        lkd_event_destroy(e0);
        lkd_event_destroy(e1);
        check(lkd_memcpy_dtoh(&done, dev.ctl + 1, sizeof(uint32_t), stream), "progress");
        if (static_cast<int>(done) < cnt)
            throw std::runtime_error("BatchedMehrotraSolver: " + std::to_string(cnt - static_cast<int>(done)) +
                                     " LPs did not finish within the launch budget");
        std::vector<lkb_state> st(count);
        check(lkd_memcpy_dtoh(st.data(), dev.state, sizeof(lkb_state) * count, stream), "states");
        std::vector<int32_t> ho(cnt);
        if (ord == dev.order) ho = img.order;
        else check(lkd_memcpy_dtoh(ho.data(), ord, sizeof(int32_t) * cnt, stream), "order");
        for (int t = 0; t < cnt; ++t) {
            const int i = ho[t];
            res[i] = to_result(st[i], clock_khz);
            if (st[i].status == 0) sum.optimal++;
            const double mm = img.lps[i].m;
            sum.flops_chol += st[i].factorizations * mm * mm * mm / 3.0;
        }
        have_solution = false;
        return sum;
    }

    int default_wg() const { return opt.wg_per_cu > 0 ? opt.wg_per_cu : 2; }
    int default_iters() const { return opt.iters_per_launch > 0 ? opt.iters_per_launch : opt.max_iter + 1; }

    BatchSummary solve() {
        BatchSummary sum = run(dev.order, count, default_wg(), default_iters(), opt.use_graph != 0);
        solved = true;
        return sum;
    }

    BatchSummary autotune(int sample) {
        sample = std::max(1, std::min(sample, count));
        // every (count / sample)-th LP of the size order: the sample spans the size range
        std::vector<int32_t> so(sample);
        for (int t = 0; t < sample; ++t)
            so[t] = img.order[static_cast<size_t>(static_cast<int64_t>(t) * count / sample)];
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

    void set_warm(const double* x0, const double* y0, const double* s0, const int32_t* flags) {
        if (!d_x0) {
            d_x0 = static_cast<double*>(arena->allocate_uninitialized(sizeof(double) * img.n_xn));
            d_s0 = static_cast<double*>(arena->allocate_uninitialized(sizeof(double) * img.n_xn));
            d_y0 = static_cast<double*>(arena->allocate_uninitialized(sizeof(double) * img.n_ym));
            d_wflag = static_cast<int32_t*>(arena->allocate(sizeof(int32_t) * count));
        }
        std::vector<int32_t> f(count, 0);
        if (flags) std::copy(flags, flags + count, f.begin());
        else if (x0) std::fill(f.begin(), f.end(), 1);
        if (x0) {
            if (!y0 || !s0) throw std::invalid_argument("set_warm_start: x, y and s together");
            check(lkd_memcpy_htod(d_x0, x0, sizeof(double) * img.n_xn, stream), "warm x");
            check(lkd_memcpy_htod(d_s0, s0, sizeof(double) * img.n_xn, stream), "warm s");
            check(lkd_memcpy_htod(d_y0, y0, sizeof(double) * img.n_ym, stream), "warm y");
        } else {
            std::fill(f.begin(), f.end(), 0);
        }
        check(lkd_memcpy_htod(d_wflag, f.data(), sizeof(int32_t) * count, stream), "warm flags");
        check(lkd_stream_sync(stream), "warm");
    }

    void crossover(int32_t* basis, double* x, int32_t* status, int32_t* pivots) {
        if (!solved) throw std::invalid_argument("crossover: no solution yet (call solve() first)");
        if (!d_basis) {
            d_basis = static_cast<int32_t*>(arena->allocate_uninitialized(sizeof(int32_t) * img.n_ym));
            d_xv = static_cast<double*>(arena->allocate_uninitialized(sizeof(double) * img.n_xn));
            d_xstat = static_cast<int32_t*>(arena->allocate_uninitialized(sizeof(int32_t) * 2 * count));
        }
        const lkb_args a = make_args(opt, dev, dev.order, count, 1);
        lkx_args xo;
        xo.x = dev.x;
        xo.basis = d_basis;
        xo.xv = d_xv;
        xo.status = d_xstat;
        xo.pivots = d_xstat + count;
        check(lkd_batch_crossover(&a, &xo, stream), "crossover launch");
        if (basis) check(lkd_memcpy_dtoh(basis, d_basis, sizeof(int32_t) * img.n_ym, stream), "basis");
        if (x) check(lkd_memcpy_dtoh(x, d_xv, sizeof(double) * img.n_xn, stream), "basic x");
        if (status) check(lkd_memcpy_dtoh(status, d_xstat, sizeof(int32_t) * count, stream), "status");
        if (pivots) check(lkd_memcpy_dtoh(pivots, d_xstat + count, sizeof(int32_t) * count, stream), "pivots");
        check(lkd_stream_sync(stream), "crossover");
    }

    void fetch_solution() {
        if (have_solution) return;
        hx.resize(static_cast<size_t>(img.n_xn));
        hs.resize(static_cast<size_t>(img.n_xn));
        hy.resize(static_cast<size_t>(img.n_ym));
        check(lkd_memcpy_dtoh(hx.data(), dev.x, sizeof(double) * img.n_xn, stream), "x");
        check(lkd_memcpy_dtoh(hs.data(), dev.s, sizeof(double) * img.n_xn, stream), "s");
        check(lkd_memcpy_dtoh(hy.data(), dev.y, sizeof(double) * img.n_ym, stream), "y");
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
void BatchedMehrotraSolver::set_warm_start(const double* x0, const double* y0, const double* s0,
                                           const int32_t* flags) {
    impl_->set_warm(x0, y0, s0, flags);
}
void BatchedMehrotraSolver::crossover(int32_t* basis, double* x, int32_t* status, int32_t* pivots) {
    impl_->crossover(basis, x, status, pivots);
}
void BatchedMehrotraSolver::solution(int i, double* x, double* y, double* s) const {
    if (i < 0 || i >= impl_->count) throw std::invalid_argument("solution: LP index out of range");
    impl_->fetch_solution();
    const lkb_lp& d = impl_->img.lps[i];
    if (x) std::copy(impl_->hx.begin() + d.on, impl_->hx.begin() + d.on + d.n, x);
    if (s) std::copy(impl_->hs.begin() + d.on, impl_->hs.begin() + d.on + d.n, s);
    if (y) std::copy(impl_->hy.begin() + d.om, impl_->hy.begin() + d.om + d.m, y);
}

void crossover_points(int count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                      const int32_t* rowidx, const double* vals, const double* b, const double* c,
                      const double* x, int32_t* basis, double* xv, int32_t* status,
                      int32_t* pivots) {
    if (!x) throw std::invalid_argument("crossover_points: null x");
    const BatchImage img = pack_batch(count, m, n, colptr, rowidx, vals, b, c, 1);
    gpu::VRAMArena arena(img.device_bytes() + img.crossover_bytes() + (1u << 20));
    void* stream = nullptr;
    check(lkd_stream_create(&stream), "stream");
    try {
        const BatchDevice dev =
            place_batch(img, [&](size_t nb) { return arena.allocate_uninitialized(nb); }, stream);
        check(lkd_memcpy_htod(dev.x, x, sizeof(double) * img.n_xn, stream), "points");
        int32_t* d_basis = static_cast<int32_t*>(arena.allocate_uninitialized(sizeof(int32_t) * img.n_ym));
        double* d_xv = static_cast<double*>(arena.allocate_uninitialized(sizeof(double) * img.n_xn));
        int32_t* d_st = static_cast<int32_t*>(arena.allocate_uninitialized(sizeof(int32_t) * 2 * count));
        BatchOptions opt;
        const lkb_args a = make_args(opt, dev, dev.order, count, 1);
        lkx_args xo;
        xo.x = dev.x;
        xo.basis = d_basis;
        xo.xv = d_xv;
        xo.status = d_st;
        xo.pivots = d_st + count;
        check(lkd_batch_crossover(&a, &xo, stream), "crossover launch");
        if (basis) check(lkd_memcpy_dtoh(basis, d_basis, sizeof(int32_t) * img.n_ym, stream), "basis");
        if (xv) check(lkd_memcpy_dtoh(xv, d_xv, sizeof(double) * img.n_xn, stream), "basic x");
        if (status) check(lkd_memcpy_dtoh(status, d_st, sizeof(int32_t) * count, stream), "status");
        if (pivots) check(lkd_memcpy_dtoh(pivots, d_st + count, sizeof(int32_t) * count, stream), "pivots");
        check(lkd_stream_sync(stream), "crossover");
    } catch (...) {
        lkd_stream_sync(stream);
        lkd_stream_destroy(stream);
        throw;
    }
    check(lkd_stream_destroy(stream), "stream");
}

BatchSummary solve_batch_noalloc(int count, const int32_t* m, const int32_t* n,
                                 const int32_t* colptr, const int32_t* rowidx, const double* vals,
                                 const double* b, const double* c, const BatchOptions& opt,
                                 double* x, double* y, double* s, BatchLPResult* results) {
    static std::mutex mu;   // one call at a time per process: they share the scratch
    check_options(opt);
    if (!x || !y || !s || !results) throw std::invalid_argument("solve_batch_noalloc: null output");
    const BatchImage img = pack_batch(count, m, n, colptr, rowidx, vals, b, c, opt.order);
    std::lock_guard<std::mutex> lock(mu);
    void* base = nullptr;
    size_t cap = 0;
    check(lkd_batch_scratch(&base, &cap), "scratch address");
    const size_t need = img.device_bytes();
    if (need > cap)
        throw std::invalid_argument("solve_batch_noalloc: the batch needs " + std::to_string(need) +
                                    " bytes of device scratch; the fixed scratch holds " +
                                    std::to_string(cap) + " bytes (use BatchedMehrotraSolver)");
    size_t off = 0;
    auto take = [&](size_t nb) {
        void* p = static_cast<char*>(base) + off;
        off += BatchImage::rnd(nb);
        if (off > cap) throw std::logic_error("solve_batch_noalloc: scratch layout overflow");
        return p;
    };
    void* stream = nullptr;   // the null stream: no stream or event is created
    const auto t0 = std::chrono::steady_clock::now();
    const BatchDevice dev = place_batch(img, take, stream);
    const int grid = (opt.wg_per_cu > 0 ? opt.wg_per_cu : 2) * device_cus();
    const lkb_args a = make_args(opt, dev, dev.order, count, opt.max_iter + 1);
    check(lkd_batch_reset(dev.ctl, dev.state, dev.order, count, 1, stream), "reset");
    check(lkd_batch_solve(&a, grid, stream), "batch launch");
    uint32_t done = 0;
    check(lkd_memcpy_dtoh(&done, dev.ctl + 1, sizeof(uint32_t), stream), "progress");
    if (static_cast<int>(done) < count)
        throw std::runtime_error("solve_batch_noalloc: " + std::to_string(count - static_cast<int>(done)) +
                                 " LPs did not finish");
    std::vector<lkb_state> st(count);
    check(lkd_memcpy_dtoh(st.data(), dev.state, sizeof(lkb_state) * count, stream), "states");
    check(lkd_memcpy_dtoh(x, dev.x, sizeof(double) * img.n_xn, stream), "x");
    check(lkd_memcpy_dtoh(s, dev.s, sizeof(double) * img.n_xn, stream), "s");
    check(lkd_memcpy_dtoh(y, dev.y, sizeof(double) * img.n_ym, stream), "y");
    BatchSummary sum;
    sum.count = count;
    sum.launches = 1;
    sum.grid = grid;
    sum.wg_per_cu = grid / device_cus();
    sum.iters_per_launch = opt.max_iter + 1;
    sum.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const double khz = device_clock_khz();
    for (int i = 0; i < count; ++i) {
        results[i] = to_result(st[i], khz);
        if (st[i].status == 0) sum.optimal++;
        const double mm = img.lps[i].m;
        sum.flops_chol += st[i].factorizations * mm * mm * mm / 3.0;
    }
    return sum;
}

}  // namespace ipm

namespace simplex {

void solve_batch(int count, const int32_t* m, const int32_t* n, const int32_t* colptr,
                 const int32_t* rowidx, const double* vals, const double* b, const double* c,
                 int32_t max_iter, int32_t* basis, double* x, double* y, int32_t* status,
                 int32_t* iterations, double* objective) {
    using namespace ipm;
    if (max_iter < 0) throw std::invalid_argument("simplex::solve_batch: negative iteration limit");
    const BatchImage img = pack_batch(count, m, n, colptr, rowidx, vals, b, c, 1);
    const size_t out = BatchImage::rnd(sizeof(int32_t) * img.n_ym) + BatchImage::rnd(sizeof(double) * img.n_xn) +
                       BatchImage::rnd(sizeof(double) * img.n_ym) + BatchImage::rnd(sizeof(int32_t) * 3 * count) +
                       BatchImage::rnd(sizeof(double) * count);
    gpu::VRAMArena arena(img.device_bytes() + out + (1u << 20));
    void* stream = nullptr;
    check(lkd_stream_create(&stream), "stream");
    try {
        auto take = [&](size_t nb) { return arena.allocate_uninitialized(nb); };
        const BatchDevice dev = place_batch(img, take, stream);
        lks_args sx;
        sx.basis = static_cast<int32_t*>(take(sizeof(int32_t) * img.n_ym));
        sx.x = static_cast<double*>(take(sizeof(double) * img.n_xn));
        sx.y = static_cast<double*>(take(sizeof(double) * img.n_ym));
        sx.status = static_cast<int32_t*>(take(sizeof(int32_t) * 3 * count));
        sx.iterations = sx.status + count;
        sx.objective = static_cast<double*>(take(sizeof(double) * count));
        sx.max_iter = max_iter;
        BatchOptions opt;
        const lkb_args a = make_args(opt, dev, dev.order, count, 1);
        check(lkd_batch_simplex(&a, &sx, stream), "simplex launch");
        if (basis) check(lkd_memcpy_dtoh(basis, sx.basis, sizeof(int32_t) * img.n_ym, stream), "basis");
        if (x) check(lkd_memcpy_dtoh(x, sx.x, sizeof(double) * img.n_xn, stream), "x");
        if (y) check(lkd_memcpy_dtoh(y, sx.y, sizeof(double) * img.n_ym, stream), "y");
        if (status) check(lkd_memcpy_dtoh(status, sx.status, sizeof(int32_t) * count, stream), "status");
        if (iterations)
            check(lkd_memcpy_dtoh(iterations, sx.iterations, sizeof(int32_t) * 2 * count, stream), "iterations");
        if (objective) check(lkd_memcpy_dtoh(objective, sx.objective, sizeof(double) * count, stream), "objective");
        check(lkd_stream_sync(stream), "simplex");
    } catch (...) {
        lkd_stream_sync(stream);
        lkd_stream_destroy(stream);
        throw;
    }
    check(lkd_stream_destroy(stream), "stream");
}

}  // namespace simplex
}  // namespace sankhya

// GPUKKTCholeskySolver: plan construction on the host, numeric work through the device C ABI.
//
// Follows src/cuda/kkt.cu (reference) for semantics:
//   factorize_device  kkt.cu:556-603  (M0 + 1e-9 I, unequilibrated, throws if not SPD)
//   solve_device      kkt.cu:620-870  (D = 1/sqrt(max|M0_ii|,1e-12), delta 1e-9 x10 up to 5
//                                      tries, refinement against M0 with ||.||_inf, rollback on
//                                      stagnation, no larger delta after a refinement failure)
// and replaces its mechanisms:
//   per-row merge assembly       -> pair map sorted by (destination, column), fixed fma order
//   level-by-level row Cholesky  -> supernodal tile schedule (POTRF/TRSM/SYRK/UPDATE)
//   cuSPARSE SpSV per solve      -> supernodal triangular-solve schedule
//   refactorization per solve    -> one cached factor per Theta
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <map>
#include <string>
#include "gpu/kkt.hpp"
#include "parallel.hpp"
#include "lakshya_device.h"

namespace sankhya {
namespace gpu {

namespace {

void check(int e, const char* what) {
    if (e != 0)
        throw std::runtime_error(std::string("HIP call failed (") + what + "): " +
                                 lkd_error_string(e));
}

constexpr int TILE = 64;           // block-column width (diagonal blocks, substitution blocks)
constexpr int GEMM_MT = 128;        // COLUPD / UPDATE tile rows
constexpr int GEMM_NC = 64;         // COLUPD / UPDATE tile columns
constexpr int GEMM_TARGET_WG = 256;   // UPDATE K splitting aims at this many workgroups
constexpr int KSPLIT_MIN = 128;     // minimum K columns per split part
constexpr int STEP_TARGET_WG = 512; // a STEP launch with fewer tiles splits K (PART launch) ...
constexpr int STEP_KMIN = 64;       // ... into parts of at least this many columns
constexpr int PSUM_MIN = 3;         // tiles with at least this many parts get them summed by PSUM
constexpr int STEPS_RECENT = 192;   // STEPS: K columns a chain tile computes itself
constexpr int STEPS_PART_LAG = 1;   // STEPS: steps between a PART's inputs and its queue place
constexpr int STEPS_SPLIT_CLS = 2;  // STEPS: tiles of class <= this (0 next-diagonal, 1 below it, 2 others) split
constexpr int SOLVE_WIDE_TILES = 8;   // level solves: 8 wavefronts per supernode above this
constexpr int SOLVE_MAX_G = 4;        // forward level solves: workgroups per wide supernode
constexpr int SOLVE_BWD_G = 2;        // backward level solves: workgroups per wide supernode
constexpr int FWD_ROWS = 256;
constexpr int WIDE_MIN_BLOCKS = 16;   // level solves: levels of at most WIDE_MAX_SN supernodes, one
constexpr int WIDE_MAX_SN = 8;        // of at least WIDE_MIN_BLOCKS blocks, take the wide form

// Order a launch's tasks for dispatch.  Workgroup b runs task (b % 8) * per + b / 8 with
// per = ceil(n / 8): segment x of `per` consecutive tasks runs on XCD x, in list order.  Inside
// every segment the critical tasks move to the front, so they start with the first workgroups.
void order_for_dispatch(std::vector<lkd_task>& v, const std::vector<char>& crit) {
    const size_t n = v.size(), per = (n + 7) / 8;
    std::vector<lkd_task> out;
    out.reserve(n);
    for (size_t a = 0; a < n; a += per) {
        const size_t b = std::min(n, a + per);
        for (size_t i = a; i < b; ++i)
            if (crit[i]) out.push_back(v[i]);
        for (size_t i = a; i < b; ++i)
            if (!crit[i]) out.push_back(v[i]);
    }
    v.swap(out);
}

struct Launch {
    int kind;        // 1 diag 2 update 3 diagwb 5 step 6 part 7 steps | 20 forward / 21 backward level
    int64_t begin;
    int32_t count;
    int32_t nw = 0;  // level solves: wavefronts per workgroup
    int32_t g = 1;   // level solves: workgroups per supernode
    int32_t q = -1;  // STEPS: launch number (its queue counters and offsets)
    int32_t grid = 0;  // STEPS: workgroups
};

template <typename T>
T* dev_upload(VRAMArena& arena, const std::vector<T>& v, std::vector<void*>& owned) {
    if (v.empty()) return nullptr;
    T* p = static_cast<T*>(arena.allocate_uninitialized(v.size() * sizeof(T)));
    owned.push_back(p);
    check(lkd_memcpy_htod(p, v.data(), v.size() * sizeof(T), nullptr), "upload");
    return p;
}

}  // namespace

struct GPUKKTCholeskySolver::Impl {
    VRAMArena* arena_p;            // nullptr: host-only plan (inspection / CPU emulation)
    std::map<std::string, std::vector<char>>* keep = nullptr;   // host copies of uploads

    template <typename T>
    T* up(const std::vector<T>& v, const char* name) {
        if (keep) {
            auto& dst = (*keep)[name];
            dst.resize(v.size() * sizeof(T));
            if (!v.empty()) std::memcpy(dst.data(), v.data(), dst.size());
        }
        if (!arena_p) return nullptr;
        return dev_upload(*arena_p, v, owned);
    }
    Index m = 0, n = 0;
    KKTOptions opt;
    KKTStats st;
    void* stream = nullptr;
    bool own_stream = false;
    std::vector<void*> owned;

    ipm::SupernodalStructure S;
    std::vector<Index> perm, pinv;   // perm[i] = original row at permuted position i

    // device structure
    int32_t *d_perm = nullptr, *d_colptr = nullptr, *d_prow = nullptr, *d_prp = nullptr,
            *d_pci = nullptr;
    double *d_cval = nullptr, *d_pval = nullptr;
    int64_t *d_segptr = nullptr, *d_segdst = nullptr, *d_dseg = nullptr;
    int32_t *d_segi = nullptr, *d_segj = nullptr, *d_pk = nullptr;
    double* d_pp = nullptr;
    uint8_t* d_empty = nullptr;
    // rows / columns / pair-map segments handled by whole blocks (LKD_LONG_ROW / LKD_LONG_SEG)
    lkd_long long_prows{0, nullptr, 0, nullptr}, long_cols{0, nullptr, 0, nullptr}, long_diag{0, nullptr, 0, nullptr}, long_segs{0, nullptr, 0, nullptr};
    lkd_long mid_diag{0, nullptr, 0, nullptr}, mid_segs{0, nullptr, 0, nullptr};   // 16-lane class of the pair map
    int64_t* d_chunk_seg = nullptr;   // first segment of each LKD_PANEL_CHUNK panel chunk
    double* d_mval = nullptr;         // per-segment M0 entries (assembly scratch)
    int32_t *d_sn_first = nullptr, *d_sn_rows = nullptr, *d_col_to_sn = nullptr;
    int64_t *d_sn_rowptr = nullptr, *d_sn_panel = nullptr;
    lkd_task *d_ftasks = nullptr, *d_stasks = nullptr;
    std::vector<int32_t> fdeps;            // STEPS dependencies, 4 per factor task (-1 = none)
    std::vector<int32_t> fsdep;            // STEPS: the task staging the tile's diagonal block
    int32_t *d_fdeps = nullptr, *d_fsdep = nullptr;
    int32_t n_dag = 0;                     // STEPS launches
    int32_t n_xcd = 8;                     // STEPS queues per launch: one per XCD
    std::vector<int32_t> fqueue, fqoff;    // STEPS queue order (task indices per launch), offsets
    int32_t *d_fqueue = nullptr, *d_fqoff = nullptr;
    uint32_t* d_dag_flags = nullptr;       // per factor task: epoch of its completion (rows)
    uint32_t* d_dag_flags2 = nullptr;      // per factor task: epoch of its staged factor
    uint32_t* d_dag_ctl = nullptr;         // [0] epoch, [1 + q n_xcd + x] queue counters
    int64_t nseg = 0, npairs = 0;
    std::vector<Launch> flaunch, slaunch;

    // device numeric buffers
    void* d_panel = nullptr;       // sized for FP64
    void* d_diag = nullptr;
    int64_t diag_elems = 0;
    void* d_piv = nullptr;         // pivots L_jj per permuted column (sized for FP64)
    void* d_stage = nullptr;       // staged diagonal factors, LKD_STAGE per supernode of a level
    int64_t stage_elems = 0;
    void* d_work = nullptr;        // PART products (sized for FP64)
    int64_t work_elems = 0;
    double* d_below = nullptr;     // level solves: rows below each supernode
    double* d_pbuf = nullptr;      // level solves: z exchange / backward partial sums
    uint32_t* d_flags = nullptr;   // level solves: hand-off flags, then the epoch
    int32_t* d_err = nullptr;
    int32_t* d_serr = nullptr;     // level-solve status (bit 0x10: a bounded wait timed out)
    double *d_D = nullptr, *d_m0d = nullptr, *d_theta = nullptr;
    double *d_r = nullptr, *d_req = nullptr, *d_z = nullptr, *d_dy = nullptr, *d_bak = nullptr,
           *d_rhsp = nullptr,
           *d_u = nullptr, *d_rinf = nullptr, *d_csr = nullptr;
    double* d_pcg = nullptr;       // PCG work vectors (4m)
    int64_t* d_csr_index = nullptr;
    int64_t nnz_csr = 0;

    // cache state
    bool have_eq = false;          // equilibrated factor valid for the current M0
    bool have_theta = false;
    FactorPrecision cached_prec = FactorPrecision::FP64;
    double cached_delta = 0;
    int factor_kind = 0;           // 0 none, 1 unequilibrated (reference factorize), 2 equilibrated

    Impl(VRAMArena* a, const core::CSRMatrix& A, const ipm::SymbolicFactorization* sym,
         ipm::SupernodalStructure* pre, const KKTOptions& o,
         std::map<std::string, std::vector<char>>* keep_host = nullptr)
        : arena_p(a), keep(keep_host), m(A.rows), n(A.cols), opt(o) {
        if (opt.use_graphs)
            throw std::invalid_argument("GPUKKTCholeskySolver: launch-graph capture is not supported");
        auto t0 = std::chrono::steady_clock::now();
        stream = opt.stream;
        if (!stream && arena_p) {
            check(lkd_stream_create(&stream), "stream create");
            own_stream = true;
        }
        if (pre) {
            S = std::move(*pre);
        } else {
            if (static_cast<Index>(sym->P.size()) != m)
                throw std::invalid_argument("GPUKKTCholeskySolver: permutation size mismatch");
            ipm::compute_supernodal_structure(A, sym->P, opt.supernodes, S);
        }
        if (S.m != m) throw std::invalid_argument("GPUKKTCholeskySolver: structure size mismatch");
        perm = S.perm;
        pinv.assign(static_cast<size_t>(m), 0);
        for (Index i = 0; i < m; ++i) pinv[perm[i]] = i;
        build_matrices(A);
        build_pairs();
        if (arena_p) {   // XCDs of the device: 32 compute units each (1 .. 8)
            int cus = 0;
            if (lkd_device_cus(&cus) == 0 && cus > 0) n_xcd = std::max(1, std::min(8, (cus + 16) / 32));
        }
        build_factor_plan();
        build_solve_plan();
        if (sym) build_csr_index(*sym);
        if (keep) {
            up(perm, "perm");
            up(S.sn_rowptr, "sn_rowptr");
            up(S.sn_rows, "sn_rows");
            up(S.sn_panel, "sn_panel");
            up(S.col_to_sn, "col_to_sn");
            std::vector<int32_t> f(S.sn_first.begin(), S.sn_first.end());
            up(f, "sn_first");
            std::vector<int64_t> fl, sl;
            for (const auto& L : flaunch) { fl.push_back(L.kind); fl.push_back(L.begin); fl.push_back(L.count); }
            for (const auto& L : slaunch) { sl.push_back(L.kind); sl.push_back(L.begin); sl.push_back(L.count); }
            up(fl, "flaunch");
            up(sl, "slaunch");
            std::vector<int64_t> de{diag_elems};
            up(de, "diag_elems");
            std::vector<int64_t> se{stage_elems};
            up(se, "stage_elems");
        }
        if (arena_p) allocate_numeric();
        st.m = m;
        st.n = n;
        st.nnz_A = A.row_ptrs.empty() ? 0 : A.row_ptrs[m];
        st.nsuper = S.nsuper;
        st.nnz_L = S.nnz_L;
        st.nnz_L_exact = S.nnz_L_exact;
        st.flops_exact = S.flops_exact;
        st.flops_panel = S.flops_panel;
        st.levels = S.max_level + 1;
        st.panel_elems = S.sn_panel[S.nsuper];
        st.npairs = npairs;
        st.nsegments = nseg;
        st.launches_factor = static_cast<int32_t>(flaunch.size());
        st.launches_solve = static_cast<int32_t>(slaunch.size());
        st.symbolic_ms = std::chrono::duration<double, std::milli>(
                             std::chrono::steady_clock::now() - t0).count();
    }

    ~Impl() {
        for (int p = 0; p < 2; ++p) {
        }
        if (stream) lkd_stream_sync(stream);
        for (void* e : timer.events) lkd_event_destroy(e);
        for (auto it = owned.rbegin(); it != owned.rend(); ++it) {
            try { arena_p->free(*it); } catch (...) {}
        }
        if (own_stream) lkd_stream_destroy(stream);
    }

    // -------------------------------------------------------------------- plan construction
    void build_matrices(const core::CSRMatrix& A) {
        const int64_t nz = A.row_ptrs[m];
        // CSC of A with permuted row indices (columns keep the original order)
        std::vector<Index> colptr(static_cast<size_t>(n) + 1, 0), prow(static_cast<size_t>(nz));
        std::vector<double> cval(static_cast<size_t>(nz));
        for (int64_t p = 0; p < nz; ++p) colptr[A.col_indices[p] + 1]++;
        for (Index j = 0; j < n; ++j) colptr[j + 1] += colptr[j];
        {
            std::vector<Index> nxt(colptr.begin(), colptr.end() - 1);
            for (Index i = 0; i < m; ++i)       // rows visited in original order
                for (Index p = A.row_ptrs[i]; p < A.row_ptrs[i + 1]; ++p) {
                    const Index q = nxt[A.col_indices[p]]++;
                    prow[q] = pinv[i];
                    cval[q] = A.values[p];
                }
        }
        // permuted CSR
        std::vector<Index> prp(static_cast<size_t>(m) + 1, 0), pci(static_cast<size_t>(nz));
        std::vector<double> pval(static_cast<size_t>(nz));
        for (Index k = 0; k < m; ++k) {
            const Index i = perm[k];
            prp[k + 1] = prp[k] + (A.row_ptrs[i + 1] - A.row_ptrs[i]);
            std::copy(A.col_indices.begin() + A.row_ptrs[i], A.col_indices.begin() + A.row_ptrs[i + 1],
                      pci.begin() + prp[k]);
            std::copy(A.values.begin() + A.row_ptrs[i], A.values.begin() + A.row_ptrs[i + 1],
                      pval.begin() + prp[k]);
        }
        h_colptr = colptr; h_prow = prow; h_cval = cval;
        long_prows = long_list(prp, m, LKD_LONG_ROW, "long_prows");
        long_cols = long_list(colptr, n, LKD_LONG_ROW, "long_cols");
        split_long(long_prows, prp, "long_prows_part");
        split_long(long_cols, colptr, "long_cols_part");
        d_colptr = up(colptr, "colptr");
        d_prow = up(prow, "prow");
        d_cval = up(cval, "cval");
        d_prp = up(prp, "prp");
        d_pci = up(pci, "pci");
        d_pval = up(pval, "pval");
        d_perm = up(perm, "perm");
        d_empty = up(S.empty_row, "empty_row");
    }
    std::vector<Index> h_colptr, h_prow;
    std::vector<double> h_cval;

    // indices k < count with ptr[k+1] - ptr[k] > thr, uploaded
    template <typename P>
    lkd_long long_list(const std::vector<P>& ptr, int64_t count, int64_t thr, const char* name) {
        std::vector<int32_t> idx;
        for (int64_t k = 0; k < count; ++k)
            if (static_cast<int64_t>(ptr[k + 1]) - static_cast<int64_t>(ptr[k]) > thr)
                idx.push_back(static_cast<int32_t>(k));
        return {static_cast<int32_t>(idx.size()), up(idx, name), 0, nullptr};
    }

    // row passes: the listed rows are summed in pieces of about LKD_LONG_PIECE entries by a
    // pre-pass; scratch for the piece sums
    template <typename P>
    void split_long(lkd_long& lg, const std::vector<P>& ptr, const char* name) {
        int64_t longest = 0;
        for (size_t k = 0; k + 1 < ptr.size(); ++k)
            longest = std::max(longest, static_cast<int64_t>(ptr[k + 1]) - static_cast<int64_t>(ptr[k]));
        lg.split = static_cast<int32_t>(std::min<int64_t>(LKD_LONG_MAX_SPLIT, (longest + LKD_LONG_PIECE - 1) / LKD_LONG_PIECE));
        lg.part = nullptr;
        if (lg.n > 0 && lg.split > 1)
            lg.part = up(std::vector<double>(static_cast<size_t>(lg.n) * lg.split, 0.0), name);
    }

    int64_t panel_index(Index i, Index j) const {   // i >= j, permuted
        const Index s = S.col_to_sn[j];
        const int64_t r0 = S.sn_rowptr[s], r1 = S.sn_rowptr[s + 1];
        const Index f = S.sn_first[s];
        auto b = S.sn_rows.begin() + r0 + (j - f), e = S.sn_rows.begin() + r1;
        auto it = std::lower_bound(b, e, i);
        if (it == e || *it != i) throw std::logic_error("KKT plan: entry outside the L structure");
        return S.sn_panel[s] + static_cast<int64_t>(j - f) * (r1 - r0) + (it - (S.sn_rows.begin() + r0));
    }

    void build_pairs() {
        struct Pair { int64_t dst; int32_t a, b, k; };
        // The pairs of every column k of A in (a, b) order, at off[k] (columns in parallel); then
        // a stable counting sort by destination column of L and, per column, a stable sort by
        // panel slot.  A column's slots form one contiguous range, increasing with the column,
        // so the result is the stable order by slot of the generation order.
        std::vector<int64_t> off(static_cast<size_t>(n) + 1, 0);
        for (Index k = 0; k < n; ++k) {
            const int64_t c = h_colptr[k + 1] - h_colptr[k];
            off[static_cast<size_t>(k) + 1] = off[static_cast<size_t>(k)] + c * (c + 1) / 2;
        }
        const int64_t total = off[static_cast<size_t>(n)];
        std::vector<Pair> gen(static_cast<size_t>(total));
        std::vector<Index> dcol(static_cast<size_t>(total));
        host::parallel_rows(n, [&] {
            return [&](int64_t kk) {
                const Index k = static_cast<Index>(kk);
                int64_t q = off[static_cast<size_t>(k)];
                for (Index a = h_colptr[k]; a < h_colptr[k + 1]; ++a)
                    for (Index b = h_colptr[k]; b < h_colptr[k + 1]; ++b) {
                        const Index i = h_prow[a], j = h_prow[b];
                        if (i < j) continue;
                        gen[static_cast<size_t>(q)] = {panel_index(i, j), a, b, k};
                        dcol[static_cast<size_t>(q)] = j;
                        ++q;
                    }
            };
        });
        std::vector<int64_t> cstart(static_cast<size_t>(m) + 1, 0);
        for (Index j : dcol) ++cstart[static_cast<size_t>(j) + 1];
        for (Index j = 0; j < m; ++j) cstart[static_cast<size_t>(j) + 1] += cstart[static_cast<size_t>(j)];
        std::vector<Pair> pairs(static_cast<size_t>(total));
        {
            std::vector<int64_t> nxt(cstart.begin(), cstart.end() - 1);
            for (int64_t q = 0; q < total; ++q)
                pairs[static_cast<size_t>(nxt[static_cast<size_t>(dcol[static_cast<size_t>(q)])]++)] =
                    gen[static_cast<size_t>(q)];
        }
        std::vector<Pair>().swap(gen);
        std::vector<Index>().swap(dcol);
        host::parallel_rows(m, [&] {
            return [&](int64_t j) {
                std::stable_sort(pairs.begin() + cstart[static_cast<size_t>(j)],
                                 pairs.begin() + cstart[static_cast<size_t>(j) + 1],
                                 [](const Pair& x, const Pair& y) { return x.dst < y.dst; });
            };
        });
        npairs = static_cast<int64_t>(pairs.size());
        std::vector<int64_t> segptr, segdst;
        std::vector<int32_t> segi, segj, pk(pairs.size());
        std::vector<double> pp(pairs.size());
        std::vector<int64_t> dseg(static_cast<size_t>(m), -1);
        for (size_t q = 0; q < pairs.size(); ++q) {
            if (q == 0 || pairs[q].dst != pairs[q - 1].dst) {
                segptr.push_back(static_cast<int64_t>(q));
                segdst.push_back(pairs[q].dst);
                const Index i = h_prow[pairs[q].a], j = h_prow[pairs[q].b];
                segi.push_back(i);
                segj.push_back(j);
                if (i == j) dseg[i] = static_cast<int64_t>(segdst.size()) - 1;
            }
            pp[q] = h_cval[pairs[q].a] * h_cval[pairs[q].b];
            pk[q] = pairs[q].k;
        }
        segptr.push_back(static_cast<int64_t>(pairs.size()));
        nseg = static_cast<int64_t>(segdst.size());
        pairs.clear();
        pairs.shrink_to_fit();
        if (nseg >= (int64_t(1) << 31)) throw std::length_error("KKT plan: more than 2^31 M0 entries");
        long_segs = long_list(segptr, nseg, LKD_LONG_SEG, "long_segs");
        {
            std::vector<int32_t> ms;
            for (int64_t t = 0; t < nseg; ++t) {
                const int64_t len = segptr[t + 1] - segptr[t];
                if (len > LKD_SHORT_SEG && len <= LKD_LONG_SEG) ms.push_back(static_cast<int32_t>(t));
            }
            mid_segs = {static_cast<int32_t>(ms.size()), up(ms, "mid_segs"), 0, nullptr};
            std::vector<int32_t> ld, md;
            for (Index i = 0; i < m; ++i) {
                if (dseg[i] < 0) continue;
                const int64_t len = segptr[dseg[i] + 1] - segptr[dseg[i]];
                if (len > LKD_LONG_SEG) ld.push_back(i);
                else if (len > LKD_SHORT_SEG) md.push_back(i);
            }
            long_diag = {static_cast<int32_t>(ld.size()), up(ld, "long_diag"), 0, nullptr};
            mid_diag = {static_cast<int32_t>(md.size()), up(md, "mid_diag"), 0, nullptr};
        }
        {
            const int64_t pe = S.sn_panel[S.nsuper];
            const int64_t nchunk = (pe + LKD_PANEL_CHUNK - 1) / LKD_PANEL_CHUNK;
            std::vector<int64_t> cs(static_cast<size_t>(nchunk) + 1);
            for (int64_t c = 0; c <= nchunk; ++c)
                cs[c] = std::lower_bound(segdst.begin(), segdst.end(), c * LKD_PANEL_CHUNK) - segdst.begin();
            d_chunk_seg = up(cs, "chunk_seg");
        }
        d_segptr = up(segptr, "segptr");
        d_segdst = up(segdst, "segdst");
        d_segi = up(segi, "segi");
        d_segj = up(segj, "segj");
        d_pp = up(pp, "pp");
        d_pk = up(pk, "pk");
        d_dseg = up(dseg, "dseg");
    }

    // diag-copy offsets per supernode block column
    std::vector<int64_t> diag_off_start;   // [nsuper+1] start offset for the supernode's blocks

    int64_t diag_off(Index s, int blk) const {
        const Index nc = S.sn_first[s + 1] - S.sn_first[s];
        int64_t off = diag_off_start[s];
        for (int b = 0; b < blk; ++b) {
            const int kb = std::min(TILE, nc - b * TILE);
            off += static_cast<int64_t>(kb) * kb;
        }
        return off;
    }

    // A STEPS launch keeps every supernode on one XCD (its tiles hand data to each other through
    // that XCD's L2).  A level whose work one or a few supernodes dominate -- a dense top
    // separator -- would then leave most XCDs idle; such a level takes one launch per step
    // instead, whose tiles spread over all XCDs and whose hand-offs are the launch boundaries.
    // Balanced: the busiest XCD of a largest-first assignment (dense-trapezoid flops per
    // supernode) holds at most 1.5 times the mean.
    bool xcd_balanced(const std::vector<Index>& lv) const {
        std::vector<double> w;
        double total = 0.0;
        for (Index s : lv) {
            const double nc = S.sn_first[s + 1] - S.sn_first[s];
            const double ld = static_cast<double>(S.sn_rowptr[s + 1] - S.sn_rowptr[s]);
            const double f = nc * ld * ld - nc * nc * ld + nc * nc * nc / 3.0;
            w.push_back(f);
            total += f;
        }
        std::sort(w.begin(), w.end(), std::greater<double>());
        std::vector<double> bin(static_cast<size_t>(n_xcd), 0.0);
        for (double f : w) *std::min_element(bin.begin(), bin.end()) += f;
        return *std::max_element(bin.begin(), bin.end()) <= 1.5 * total / n_xcd;
    }

    void build_factor_plan() {
        const Index ns = S.nsuper;
        // the UPDATE scatter places a row inside a target's columns at row - first without a
        // search: every supernode's row list must start with its own columns, in order
        for (Index s = 0; s < ns; ++s)
            for (Index k = 0; k < S.sn_first[s + 1] - S.sn_first[s]; ++k)
                if (S.sn_rows[S.sn_rowptr[s] + k] != S.sn_first[s] + k)
                    throw std::logic_error("KKT plan: a supernode's rows do not start with its columns");
        diag_off_start.assign(static_cast<size_t>(ns) + 1, 0);
        for (Index s = 0; s < ns; ++s) {
            const Index nc = S.sn_first[s + 1] - S.sn_first[s];
            int64_t sz = 0;
            for (Index k0 = 0; k0 < nc; k0 += TILE) {
                const int kb = std::min<Index>(TILE, nc - k0);
                sz += static_cast<int64_t>(kb) * kb;
            }
            diag_off_start[s + 1] = diag_off_start[s] + sz;
        }
        diag_elems = diag_off_start[ns];
        std::vector<std::vector<Index>> by_level(static_cast<size_t>(S.max_level) + 1);
        for (Index s = 0; s < ns; ++s) by_level[S.sn_level[s]].push_back(s);
        levels = by_level;

        std::vector<lkd_task> tasks;
        auto base = [&](Index s) {
            lkd_task t{};
            t.off = S.sn_panel[s];
            t.ld = static_cast<int32_t>(S.sn_rowptr[s + 1] - S.sn_rowptr[s]);
            t.ncol = S.sn_first[s + 1] - S.sn_first[s];
            t.sn = s;
            t.first = S.sn_first[s];
            t.aux2 = S.sn_rowptr[s];
            t.ws = -1;
            return t;
        };
        auto emit = [&](int kind, int64_t begin) {
            const int64_t cnt = static_cast<int64_t>(tasks.size()) - begin;
            if (cnt > 0) flaunch.push_back({kind, begin, static_cast<int32_t>(cnt)});
        };
        // K split: enough workgroups to cover the CUs, at least KSPLIT_MIN columns per part
        auto nsplit_for = [](int64_t tiles, int64_t K) {
            if (tiles <= 0 || K <= 0) return 1;
            const int64_t want = (GEMM_TARGET_WG + tiles - 1) / tiles;
            return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>({want, K / KSPLIT_MIN, 64})));
        };
        auto kpart = [](int K, int ns, int j) {   // split boundaries on multiples of 16
            return static_cast<int32_t>((static_cast<int64_t>(K) * j / ns) / 16 * 16);
        };
        std::vector<lkd_task> diag_tasks;   // one per (supernode, block column): DIAGWB
        std::vector<int32_t> sidx(static_cast<size_t>(ns), -1);   // supernode -> index in its level
        for (const auto& lv : by_level) {
            for (size_t si = 0; si < lv.size(); ++si) sidx[lv[si]] = static_cast<int32_t>(si);
            int maxblk = 0;
            for (Index s : lv) {
                const Index nc = S.sn_first[s + 1] - S.sn_first[s];
                maxblk = std::max<int>(maxblk, (nc + TILE - 1) / TILE);
            }
            // staging slots: with one launch per step, two per supernode of the level (block k
            // is read from slot k % 2 while block k + 1 is written to the other one); with the
            // STEPS launch, where steps overlap, one per block column of the level
            const bool dag = opt.step_dag == 2 || (opt.step_dag == 1 && xcd_balanced(lv));
            std::vector<int64_t> blkbase(lv.size() + 1, 0);
            for (size_t si = 0; si < lv.size(); ++si) {
                const Index nc = S.sn_first[lv[si] + 1] - S.sn_first[lv[si]];
                blkbase[si + 1] = blkbase[si] + (dag ? (nc + TILE - 1) / TILE : 2);
            }
            auto slot = [&](size_t si, int k) {
                return (blkbase[si] + (dag ? k : (k & 1))) * static_cast<int64_t>(LKD_STAGE);
            };
            stage_elems = std::max<int64_t>(stage_elems, blkbase[lv.size()] * LKD_STAGE);
            // STEPS: the tiles of all steps in step order, with, per tile, the step and rows
            // PART tasks: dstep = the step whose tiles complete their K range, after = the step
            // after whose tiles they are queued (dstep + STEPS_PART_LAG, before their tile's step)
            struct DagTile { size_t si; int k; int r0, r1; bool nsplit; int cls; bool part; int after, dstep; };
            std::vector<lkd_task> dag_tasks;
            std::vector<DagTile> dag_meta;
            for (Index s : lv) {
                lkd_task t = base(s);
                for (int k = 0; k * TILE < t.ncol; ++k) {
                    lkd_task u = t;
                    u.k0 = k * TILE;
                    u.kb = std::min(TILE, t.ncol - u.k0);
                    u.r0 = u.k0;
                    u.aux = diag_off(s, k);
                    diag_tasks.push_back(u);
                }
            }
            // DIAG: block column 0
            int64_t b0 = static_cast<int64_t>(tasks.size());
            for (size_t si = 0; si < lv.size(); ++si) {
                lkd_task t = base(lv[si]);
                t.kb = std::min(TILE, t.ncol);
                t.aux = diag_off(lv[si], 0);
                t.ws = slot(si, 0);
                tasks.push_back(t);
            }
            emit(1, b0);
            // STEP k: 128-row tiles below the diagonal block of block column k
            std::map<std::pair<Index, int32_t>, std::pair<int64_t, int>> carried;   // (sn, r0) -> summed products
            int64_t carried_end = 0;
            for (int k = 0; k < maxblk; ++k) {
                std::vector<lkd_task> tl;
                std::vector<lkd_task> dg;   // one launch per step: block k's factorization
                for (size_t si = 0; si < lv.size(); ++si) {
                    lkd_task t = base(lv[si]);
                    if (k * TILE >= t.ncol) continue;
                    t.k0 = k * TILE;
                    t.kb = std::min(TILE, t.ncol - t.k0);
                    t.ws = slot(si, k);
                    if (!dag && k > 0) {   // factored in the PART launch of step k (beside its
                        lkd_task d = t;    // partial products), not at the end of STEP k - 1
                        d.r0 = t.k0;
                        d.aux = diag_off(lv[si], k);
                        d.nsplit = 2;
                        dg.push_back(d);
                    }
                    const int k1 = t.k0 + t.kb;
                    int r0 = k1;
                    if (k1 < t.ncol) {   // the next diagonal block: its own tile, first in the list
                        lkd_task u = t;
                        u.r0 = k1;
                        u.c0 = std::min(TILE, t.ncol - k1);
                        if (dag) {       // STEPS: the tile factors the block once its rows are done
                            u.nsplit = 1;
                            u.aux = diag_off(lv[si], k + 1);
                            u.aux2 = slot(si, k + 1);
                        } else if (k1 % GEMM_MT == 0) {
                            // one launch per step: tiles on 128-row boundaries, the same for
                            // consecutive steps below their diagonal blocks (paired PART tasks)
                            u.c0 = static_cast<int>(std::min<int64_t>(GEMM_MT, t.ld - k1));
                        }
                        tl.push_back(u);
                        r0 += u.c0;
                    }
                    for (; r0 < t.ld; r0 += GEMM_MT) {
                        lkd_task u = t;
                        u.r0 = r0;
                        u.c0 = std::min(GEMM_MT, t.ld - r0);
                        tl.push_back(u);
                    }
                }
                if (tl.empty() && dg.empty()) continue;   // (the last block: only its factorization)
                if (dag) {   // whole K; in the queue, the step's next-diagonal tiles first, then
                             // the tiles right below them (the next one waits for those), then
                             // the others
                    std::vector<int> cls(tl.size(), 2);
                    for (size_t i = 0; i < tl.size(); ++i)
                        if (tl[i].nsplit) {
                            cls[i] = 0;
                            if (i + 1 < tl.size() && tl[i + 1].sn == tl[i].sn) cls[i + 1] = 1;
                        }
                    for (int pass = 0; pass < 3; ++pass)
                        for (size_t i = 0; i < tl.size(); ++i) {
                            if (cls[i] != pass) continue;
                            lkd_task u = tl[i];
                            u.kbeg = 0;
                            u.kend = u.k0;
                            u.np = 0;
                            u.px = 0;
                            dag_meta.push_back({static_cast<size_t>(sidx[u.sn]), k, u.r0, u.r0 + u.c0,
                                                u.nsplit != 0, cls[i], false, k, k});
                            dag_tasks.push_back(u);
                        }
                    continue;
                }
                // few tiles: split every tile's K = k0 columns into S parts computed by a PART
                // launch (the STEP then only adds them); the launch count per step is one more
                int kmax = 0;
                for (const auto& u : tl) kmax = std::max(kmax, u.k0);
                const int64_t nt = static_cast<int64_t>(tl.size());
                int S = 1;
                if (nt < STEP_TARGET_WG && kmax >= 2 * STEP_KMIN)
                    S = static_cast<int>(std::min<int64_t>(STEP_TARGET_WG / nt, kmax / STEP_KMIN));
                std::vector<lkd_task> parts;
                std::vector<char> pcrit, scrit;
                int64_t px = carried_end;   // slots below carried_end hold products of step k
                // Tiles of step k - 1's pairs: their products over [0, k0 - 64) are summed in the
                // work buffer; they compute block column k - 1 (the last 64 columns) themselves.
                // An even step k pairs its tiles below block k + 1 with step k + 1 (same rows):
                // PART tasks of two block columns read each A tile once (part = 2).
                std::vector<lkd_task> next_sums;
                std::map<std::pair<Index, int32_t>, std::pair<int64_t, int>> pairs;
                const bool pair_step = k % 2 == 0 && S > 1;
                for (auto& u : tl) {
                    const auto c = carried.find({u.sn, u.r0});
                    if (c != carried.end()) {
                        u.px = c->second.first;
                        u.np = c->second.second;
                        u.kbeg = u.k0 - TILE;
                        u.kend = u.k0;
                        scrit.push_back(0);
                        continue;
                    }
                    const int Su = std::min(S, std::max(1, u.k0 / STEP_KMIN));
                    const bool paired = pair_step && Su > 1 && u.k0 + TILE < u.ncol && u.r0 >= u.k0 + 2 * TILE;
                    if (paired) {
                        const int64_t p1 = px, p2 = px + static_cast<int64_t>(Su) * LKD_PART;
                        u.kbeg = u.kend = 0;
                        u.np = Su;
                        u.px = p1;
                        for (int j = 0; j < Su; ++j) {
                            lkd_task q = u;
                            q.kbeg = kpart(u.k0, Su, j);
                            q.kend = j + 1 == Su ? u.k0 : kpart(u.k0, Su, j + 1);
                            q.np = 0;
                            q.part = 2;
                            q.px = p1 + static_cast<int64_t>(j) * LKD_PART;
                            q.ws = p2 + static_cast<int64_t>(j) * LKD_PART;
                            parts.push_back(q);
                            pcrit.push_back(0);
                        }
                        pairs[{u.sn, u.r0}] = {p2, Su};
                        if (Su >= PSUM_MIN)
                            for (int sl = 0; sl < LKD_PART / 512; ++sl) {
                                lkd_task q = u;
                                q.px = p2;
                                q.kbeg = sl;
                                q.kend = 0;
                                next_sums.push_back(q);
                            }
                        px = p2 + static_cast<int64_t>(Su) * LKD_PART;
                        scrit.push_back(0);
                        continue;
                    }
                    if (Su > 1) {
                        u.kbeg = u.kend = 0;
                        u.np = Su;
                        u.px = px;
                        for (int j = 0; j < Su; ++j) {
                            lkd_task q = u;
                            q.kbeg = kpart(u.k0, Su, j);
                            q.kend = j + 1 == Su ? u.k0 : kpart(u.k0, Su, j + 1);
                            q.np = 0;
                            q.px = px + static_cast<int64_t>(j) * LKD_PART;
                            parts.push_back(q);
                            pcrit.push_back(u.nsplit ? 1 : 0);
                        }
                        px += static_cast<int64_t>(Su) * LKD_PART;
                    } else {
                        u.kbeg = 0;
                        u.kend = u.k0;
                        u.np = 0;
                        u.px = 0;
                    }
                    scrit.push_back(u.nsplit ? 1 : 0);
                }
                work_elems = std::max(work_elems, px);
                carried.clear();
                carried_end = 0;
                for (const auto& e : pairs) {
                    carried[e.first] = {e.second.first, e.second.second >= PSUM_MIN ? 1 : e.second.second};
                    carried_end = px;   // step k + 1 keeps these slots; its own parts go above
                }
                if (!parts.empty() || !dg.empty()) {
                    order_for_dispatch(parts, pcrit);
                    parts.insert(parts.begin(), dg.begin(), dg.end());   // first workgroups
                    b0 = static_cast<int64_t>(tasks.size());
                    tasks.insert(tasks.end(), parts.begin(), parts.end());
                    emit(6, b0);
                }
                // tiles with many parts: PSUM adds them (slices on many workgroups), the STEP
                // then reads the one sum
                std::vector<lkd_task> sums = next_sums;
                for (auto& u : tl) {
                    if (u.np < PSUM_MIN || u.kend > 0) continue;   // (carried tiles: summed before)
                    for (int sl = 0; sl < LKD_PART / 512; ++sl) {
                        lkd_task q = u;
                        q.kbeg = sl;
                        q.kend = 0;
                        sums.push_back(q);
                    }
                    u.np = 1;
                }
                if (!sums.empty()) {
                    b0 = static_cast<int64_t>(tasks.size());
                    tasks.insert(tasks.end(), sums.begin(), sums.end());
                    emit(8, b0);
                }
                order_for_dispatch(tl, scrit);
                b0 = static_cast<int64_t>(tasks.size());
                tasks.insert(tasks.end(), tl.begin(), tl.end());
                emit(5, b0);
            }
            if (!dag_tasks.empty()) {
                // Tiles (of class <= STEPS_SPLIT_CLS) leave all but their last STEPS_RECENT
                // columns of K to a PART task queued right after the step that completes those
                // columns, so that product is ready early and the chains of tiles through the
                // steps only compute the recent columns.
                {
                    const size_t nt = dag_tasks.size();
                    int64_t px = 0;
                    for (size_t i = 0; i < nt; ++i) {
                        lkd_task& u = dag_tasks[i];
                        if (dag_meta[i].cls > STEPS_SPLIT_CLS || u.k0 < STEPS_RECENT + 2 * TILE) continue;
                        const int kp = (u.k0 - STEPS_RECENT) / TILE * TILE;
                        lkd_task q = u;
                        q.kbeg = 0;
                        q.kend = kp;
                        q.np = 0;
                        q.px = px;
                        q.nsplit = 0;
                        q.prio = 0;
                        q.part = 1;
                        u.kbeg = kp;
                        u.np = 1;
                        u.px = px;
                        px += LKD_PART;
                        DagTile d = dag_meta[i];
                        d.nsplit = false;
                        d.part = true;
                        d.dstep = kp / TILE - 1;
                        d.after = std::min(d.k - 1, d.dstep + STEPS_PART_LAG);
                        dag_meta.push_back(d);
                        dag_tasks.push_back(q);
                    }
                    work_elems = std::max(work_elems, px);
                    // queue order: the tiles of step k, then the PART tasks queued after it
                    std::vector<size_t> ord;
                    for (int k = 0, kmax = dag_meta[nt - 1].k; k <= kmax; ++k) {
                        for (size_t i = 0; i < nt; ++i)
                            if (dag_meta[i].k == k) ord.push_back(i);
                        for (size_t i = nt; i < dag_tasks.size(); ++i)
                            if (dag_meta[i].after == k) ord.push_back(i);
                    }
                    if (ord.size() != dag_tasks.size()) throw std::logic_error("KKT plan: STEPS order");
                    std::vector<lkd_task> t2;
                    std::vector<DagTile> m2;
                    for (size_t i : ord) {
                        t2.push_back(dag_tasks[i]);
                        m2.push_back(dag_meta[i]);
                    }
                    dag_tasks.swap(t2);
                    dag_meta.swap(m2);
                }
                // the dependencies below are taken from the tiles listed before a task: every
                // tile of the previous step (for a PART: of its data step) that it reads must be
                // listed before it
                {
                    std::vector<std::vector<size_t>> by_sn(lv.size());
                    for (size_t i = 0; i < dag_tasks.size(); ++i)
                        if (!dag_meta[i].part) by_sn[dag_meta[i].si].push_back(i);
                    for (size_t i = 0; i < dag_tasks.size(); ++i) {
                        const DagTile& d = dag_meta[i];
                        const lkd_task& u = dag_tasks[i];
                        const int step = d.part ? d.dstep : d.k - 1;
                        for (size_t j : by_sn[d.si]) {
                            const DagTile& e = dag_meta[j];
                            if (e.k != step) continue;
                            const bool reads = (e.r0 < d.r1 && d.r0 < e.r1) || (!d.part && e.nsplit) ||
                                               (d.part && e.r0 < u.k0 + u.kb && u.k0 < e.r1);
                            if (reads && j >= i)
                                throw std::logic_error("KKT plan: STEPS task listed before a tile it reads");
                        }
                    }
                }
                // a tile waits for the tiles of the previous step of its supernode whose rows
                // overlap its rows, for the one that factored its diagonal block, and for its PART;
                // a PART waits for the tiles of the step it follows that overlap its rows or the
                // rows of its diagonal block
                b0 = static_cast<int64_t>(tasks.size());
                fdeps.resize(static_cast<size_t>(4 * b0), -1);
                fsdep.resize(static_cast<size_t>(b0), -1);
                std::vector<std::vector<int32_t>> tiles_of(lv.size());   // STEP tiles so far, per supernode
                std::map<std::pair<size_t, int64_t>, int32_t> part_of;   // (supernode, px) -> PART
                for (size_t i = 0; i < dag_tasks.size(); ++i)
                    if (dag_meta[i].part) part_of[{dag_meta[i].si, dag_tasks[i].px}] = static_cast<int32_t>(i);
                for (size_t i = 0; i < dag_tasks.size(); ++i) {
                    const DagTile& d = dag_meta[i];
                    const lkd_task& u = dag_tasks[i];
                    std::vector<int32_t> dep;
                    int32_t sdep = -1;
                    auto overl = [&](int step, int a0, int a1) {
                        for (int32_t j : tiles_of[d.si]) {
                            const DagTile& e = dag_meta[static_cast<size_t>(j)];
                            if (e.k == step && e.r0 < a1 && a0 < e.r1 &&
                                std::find(dep.begin(), dep.end(), j) == dep.end())
                                dep.push_back(j);
                        }
                    };
                    if (d.part) {
                        overl(d.dstep, d.r0, d.r1);
                        overl(d.dstep, u.k0, u.k0 + u.kb);
                    } else {
                        for (int32_t j : tiles_of[d.si]) {
                            const DagTile& e = dag_meta[static_cast<size_t>(j)];
                            if (e.k == d.k - 1 && e.nsplit) {
                                dep.push_back(j);
                                sdep = j;
                            }
                        }
                        overl(d.k - 1, d.r0, d.r1);
                        if (u.np > 0) dep.push_back(part_of.at({d.si, u.px}));
                        tiles_of[d.si].push_back(static_cast<int32_t>(i));
                    }
                    fsdep.push_back(sdep);
                    if (dep.size() > 4) throw std::logic_error("KKT plan: STEPS task with more than 4 dependencies");
                    for (size_t q = 0; q < 4; ++q) fdeps.push_back(q < dep.size() ? dep[q] : -1);
                    if (d.nsplit)   // the chain of next-diagonal tiles: what it waits for is critical
                        for (int32_t j : dep) dag_tasks[static_cast<size_t>(j)].prio = 1;
                }
                // queues: every supernode's tiles on one XCD, supernodes dealt by work to the
                // least loaded XCD; each queue keeps the list order
                std::vector<double> sw(lv.size(), 0.0), xw(static_cast<size_t>(n_xcd), 0.0);
                for (const auto& u : dag_tasks) sw[static_cast<size_t>(sidx[u.sn])] += task_flops(u, 5) + 1e6;
                std::vector<size_t> ord(lv.size());
                for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
                std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return sw[a] > sw[b]; });
                std::vector<int> sx(lv.size(), 0);
                for (size_t i : ord) {
                    const int x = static_cast<int>(std::min_element(xw.begin(), xw.end()) - xw.begin());
                    sx[i] = x;
                    xw[static_cast<size_t>(x)] += sw[i];
                }
                fqueue.resize(static_cast<size_t>(b0), -1);
                int32_t longest = 0;
                for (int x = 0; x < n_xcd; ++x) {
                    fqoff.push_back(static_cast<int32_t>(fqueue.size() - static_cast<size_t>(b0)));
                    int32_t len = 0;
                    for (size_t i = 0; i < dag_tasks.size(); ++i)
                        if (sx[dag_meta[i].si] == x) {
                            fqueue.push_back(static_cast<int32_t>(i));
                            ++len;
                        }
                    longest = std::max(longest, len);
                }
                fqoff.push_back(static_cast<int32_t>(dag_tasks.size()));
                tasks.insert(tasks.end(), dag_tasks.begin(), dag_tasks.end());
                emit(7, b0);
                flaunch.back().q = n_dag++;
                flaunch.back().grid = n_xcd * longest;
            }
            // UPDATE: L21 L21^T of every supernode into its ancestors, split K
            int64_t tiles = 0, maxK = 0;
            for (Index s : lv) {
                const lkd_task t = base(s);
                if (t.ld <= t.ncol) continue;
                for (int r0 = t.ncol; r0 < t.ld; r0 += GEMM_MT)
                    for (int c0 = t.ncol; c0 < std::min(t.ld, r0 + GEMM_MT); c0 += GEMM_NC) ++tiles;
                maxK = std::max<int64_t>(maxK, t.ncol);
            }
            b0 = static_cast<int64_t>(tasks.size());
            for (Index s : lv) {
                lkd_task t = base(s);
                if (t.ld <= t.ncol) continue;
                const int ns = nsplit_for(tiles, t.ncol);
                for (int r0 = t.ncol; r0 < t.ld; r0 += GEMM_MT)
                    for (int c0 = t.ncol; c0 < std::min(t.ld, r0 + GEMM_MT); c0 += GEMM_NC)
                        for (int j = 0; j < ns; ++j) {
                            lkd_task u = t;
                            u.r0 = r0;
                            u.c0 = c0;
                            u.kbeg = kpart(t.ncol, ns, j);
                            u.kend = j + 1 == ns ? t.ncol : kpart(t.ncol, ns, j + 1);
                            tasks.push_back(u);
                        }
            }
            (void)maxK;
            emit(2, b0);
        }
        // DIAGWB: every factored diagonal block into the panel (no kernel reads them before)
        {
            const int64_t b0 = static_cast<int64_t>(tasks.size());
            tasks.insert(tasks.end(), diag_tasks.begin(), diag_tasks.end());
            emit(3, b0);
        }
        fdeps.resize(4 * tasks.size(), -1);
        fsdep.resize(tasks.size(), -1);
        fqueue.resize(tasks.size(), -1);
        for (const auto& L : flaunch)
            for (int64_t q = L.begin; q < L.begin + L.count; ++q) {
                validate(tasks[q], L.kind == 7 ? (tasks[q].part ? 6 : 5) : L.kind);
                if (L.kind != 7) continue;
                const lkd_task& t = tasks[q];
                bool has_part = t.np == 0;   // a tile with a product waits for its PART
                for (int d = 0; d < 4 && !has_part; ++d) {
                    const int32_t j = fdeps[4 * q + d];
                    has_part = j >= 0 && tasks[L.begin + j].part && tasks[L.begin + j].px == t.px &&
                               tasks[L.begin + j].kend == t.kbeg;
                }
                if (t.part ? t.np != 0 || t.nsplit || t.kbeg != 0 : t.np > 1 || !has_part)
                    throw std::logic_error("KKT plan: STEPS tile with partial products");
                for (int d = 0; d < 5; ++d) {   // earlier tiles of the same supernode
                    const int32_t j = d < 4 ? fdeps[4 * q + d] : fsdep[q];
                    if (j < -1 || j >= q - L.begin || (j >= 0 && tasks[L.begin + j].sn != tasks[q].sn))
                        throw std::logic_error("KKT plan: invalid STEPS dependency");
                    if (d == 4 && j >= 0 && !tasks[L.begin + j].nsplit)
                        throw std::logic_error("KKT plan: STEPS staging dependency is not a factoring tile");
                }
            }
        for (const auto& L : flaunch) {   // STEPS queues: a permutation, dependencies first
            if (L.kind != 7) continue;
            const int32_t* off = fqoff.data() + static_cast<size_t>(L.q) * (n_xcd + 1);
            std::vector<int32_t> pos(static_cast<size_t>(L.count), -1), xq(static_cast<size_t>(L.count), -1);
            int32_t longest = 0;
            for (int x = 0; x < n_xcd; ++x) {
                longest = std::max(longest, off[x + 1] - off[x]);
                for (int32_t p = off[x]; p < off[x + 1]; ++p) {
                    const int32_t i = fqueue[static_cast<size_t>(L.begin + p)];
                    if (i < 0 || i >= L.count || pos[i] >= 0) throw std::logic_error("KKT plan: invalid STEPS queue");
                    pos[i] = p;
                    xq[i] = x;
                }
            }
            if (off[0] != 0 || off[n_xcd] != L.count || L.grid < n_xcd * longest)
                throw std::logic_error("KKT plan: invalid STEPS queue offsets");
            for (int32_t i = 0; i < L.count; ++i)
                for (int d = 0; d < 4; ++d) {
                    const int32_t j = fdeps[4 * (L.begin + i) + d];
                    if (j >= 0 && (xq[j] != xq[i] || pos[j] >= pos[i]))
                        throw std::logic_error("KKT plan: STEPS dependency outside its queue order");
                }
        }
        d_ftasks = up(tasks, "ftasks");
        d_fdeps = up(fdeps, "fdeps");
        d_fsdep = up(fsdep, "fsdep");
        d_fqueue = up(fqueue, "fqueue");
        d_fqoff = up(fqoff, "fqoff");
        if (keep) {
            std::vector<int32_t> nx{n_xcd};
            up(nx, "n_xcd");
        }
        h_ftasks = tasks;
        n_ftasks = static_cast<int64_t>(tasks.size());
    }
    std::vector<std::vector<Index>> levels;
    int64_t n_ftasks = 0, n_stasks = 0;

    // Every task's accesses must stay inside its supernode's panel, row list and diagonal copy.
    void validate(const lkd_task& t, int kind) const {
        auto bad = [&](const char* why) {
            throw std::logic_error(std::string("KKT plan: invalid task (") + why + ")");
        };
        const Index s = t.sn;
        if (s < 0 || s >= S.nsuper) bad("supernode");
        const int64_t ld = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
        const Index nc = S.sn_first[s + 1] - S.sn_first[s];
        if (t.ld != ld || t.ncol != nc || t.off != S.sn_panel[s] || t.first != S.sn_first[s] ||
            (kind != 5 && kind != 6 && kind != 8 && t.aux2 != S.sn_rowptr[s]))   // STEP: aux2 = next staging slot
            bad("geometry");
        if (S.sn_panel[s] + ld * nc > S.sn_panel[S.nsuper]) bad("panel extent");
        if (t.first + nc > m) bad("columns");
        if (t.k0 < 0 || t.kb < 0 || t.k0 + t.kb > nc || t.kb > TILE) bad("block column");
        auto diag_ok = [&](int64_t off, int kb) { return off >= 0 && off + int64_t(kb) * kb <= diag_elems; };
        auto slot_ok = [&](int64_t off) { return off >= 0 && off + LKD_STAGE <= stage_elems; };
        switch (kind) {
            case 1:   // DIAG: first diagonal block, its staging slot and copy
                if (t.k0 != 0 || t.kb < 1 || !diag_ok(t.aux, t.kb) || !slot_ok(t.ws)) bad("diag");
                break;
            case 3:   // DIAGWB: a diagonal block and its copy
                if (t.kb < 1 || t.r0 != t.k0 || !diag_ok(t.aux, t.kb)) bad("diagwb");
                break;
            case 8:   // PSUM: a slice of a STEP tile's partial products
                if (t.np < PSUM_MIN || t.px < 0 || t.px + int64_t(t.np) * LKD_PART > work_elems ||
                    t.kbeg < 0 || t.kbeg >= LKD_PART / 512)
                    bad("psum");
                break;
            case 6:   // PART launch: the step's diagonal block (nsplit = 2), like DIAG
                if (t.nsplit == 2) {
                    if (t.k0 % TILE || t.kb < 1 || t.r0 != t.k0 || !diag_ok(t.aux, t.kb) || !slot_ok(t.ws))
                        bad("part-launch diag");
                    break;
                }
                // PART: one K range of a STEP tile's GEMM into the work buffer; part = 2: of two
                // block columns, the second one's product at ws
                if (t.part == 2 && (t.k0 + t.kb >= nc || t.r0 < t.k0 + 2 * TILE || t.ws < 0 ||
                                    t.ws + LKD_PART > work_elems || (t.ws < t.px + LKD_PART && t.px < t.ws + LKD_PART)))
                    bad("paired part");
                if (t.kb < 1 || t.r0 < t.k0 + t.kb || t.c0 < 1 || t.c0 > GEMM_MT || t.r0 + t.c0 > ld) bad("part rows");
                if (t.kbeg < 0 || t.kbeg >= t.kend || t.kend > t.k0) bad("part K");
                if (t.px < 0 || t.px + LKD_PART > work_elems) bad("part work");
                break;
            case 5: { // STEP: a tile of c0 <= 128 rows below the diagonal block of block column k
                if (t.kb < 1 || t.r0 < t.k0 + t.kb || t.c0 < 1 || t.c0 > GEMM_MT || t.r0 + t.c0 > ld) bad("step rows");
                if (t.kbeg < 0 || t.kbeg > t.kend || t.kend > t.k0) bad("step K");
                if (t.np < 0 || (t.np > 0 && (t.px < 0 || t.px + int64_t(t.np) * LKD_PART > work_elems))) bad("step work");
                if (t.np == 0 && (t.kbeg != 0 || t.kend != t.k0)) bad("step K cover");
                if (t.r0 < nc && (t.r0 % TILE)) bad("step tile alignment");   // diagonal blocks whole
                if (!slot_ok(t.ws)) bad("step slot");
                if (t.nsplit) {
                    const int k1 = t.k0 + t.kb;
                    if (t.r0 != k1 || t.c0 > TILE || k1 + t.c0 > nc || !diag_ok(t.aux, t.c0) ||
                        !slot_ok(t.aux2) || t.aux2 == t.ws)
                        bad("step next block");
                }
                break;
            }
            case 2:   // UPDATE: tile of the L21 rows, lower triangle, K inside the columns
                if (t.c0 < nc || t.c0 >= ld || t.r0 < nc || t.r0 >= ld || t.c0 >= t.r0 + GEMM_MT) bad("update");
                if (t.kbeg < 0 || t.kend > nc || t.kbeg >= t.kend) bad("update K");
                break;
            case 20:
            case 21: {  // level solves: workgroups, wavefronts, scratch / exchange / flag offsets
                const int nb = (nc + TILE - 1) / TILE;
                if (t.nsplit != 4 && t.nsplit != 8) bad("solve wavefronts");
                if ((t.c0 != 1 && !(t.c0 == (kind == 20 ? SOLVE_MAX_G : SOLVE_BWD_G) && t.nsplit == 8)) ||
                    t.r0 < 0 || t.r0 >= t.c0)
                    bad("solve workgroups");
                if (t.ws < 0 || t.ws + int64_t(nb) * t.c0 * t.nsplit * TILE > solve_pbuf || t.kbeg < 0 ||
                    t.kbeg + (ld - nc) > solve_below || t.px < 0 || t.px + nb + int64_t(nb) * t.c0 > solve_flags)
                    bad("solve offsets");
                if (!diag_ok(t.aux, 0) || t.aux + int64_t(nb - 1) * TILE * TILE + int64_t(nc - (nb - 1) * TILE) * (nc - (nb - 1) * TILE) > diag_elems)
                    bad("solve diagonal copies");
                break;
            }
            case 22:
            case 23: {  // wide level solves: a tile (forward) / block column (backward) of a supernode
                const int nb = (nc + TILE - 1) / TILE;
                const int nt = nb + static_cast<int>((ld - nc + TILE - 1) / TILE);
                if (t.r0 < 0 || t.r0 >= (kind == 22 ? nt : nb)) bad("wide solve tile");
                if (t.px < 0 || t.px + 2 * int64_t(nb) > solve_flags) bad("wide solve flags");
                if (!diag_ok(t.aux, 0) || t.aux + int64_t(nb - 1) * TILE * TILE + int64_t(nc - (nb - 1) * TILE) * (nc - (nb - 1) * TILE) > diag_elems)
                    bad("solve diagonal copies");
                break;
            }
            default: bad("kind");
        }
    }

    // Solve plan: per etree level one launch per direction.  Wavefronts per workgroup: 8 when
    // some supernode of the level has more than SOLVE_WIDE_TILES row tiles, else 4.  Forward
    // launches of wide levels whose G = SOLVE_MAX_G workgroups per supernode all fit on the
    // compute units (one each) spread every supernode over G workgroups; they take their tasks
    // from per-XCD queues, so a supernode's workgroups share an L2 and hand off without L2
    // write-backs.  The backward direction uses SOLVE_BWD_G workgroups on those levels (its
    // hand-offs between workgroups are per block, so fewer of them pay).
    // task.nsplit = wavefronts, c0 = G, r0 = g, ws = exchange offset, px = flag offset,
    // kbeg = scratch offset.
    void build_solve_plan() {
        std::vector<lkd_task> tasks;
        std::vector<std::vector<lkd_task>> per_level, per_level_bwd;
        int cus = 0;
        if (lkd_device_cus(&cus) != 0 || cus < 1) cus = 0;   // no device: one workgroup each
        int64_t nbelow = 0, npbuf = 0, nflags = 0;
        std::vector<char> wide_f, wide_b;
        for (const auto& lv : levels) {
            int nw = 4, maxnt = 1;
            Index maxnc = 0;
            for (const Index s : lv) {
                const Index nc = S.sn_first[s + 1] - S.sn_first[s];
                const int64_t ld = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
                const int nt = static_cast<int>((nc + TILE - 1) / TILE + (ld - nc + TILE - 1) / TILE);
                maxnt = std::max(maxnt, nt);
                maxnc = std::max(maxnc, nc);
                if (nt > SOLVE_WIDE_TILES) nw = 8;
            }
            // wide form: a task per row tile (forward, by tile) / block column (backward, by
            // blocks from the last), in ticket order.  Backward for every level with a supernode
            // of WIDE_MIN_BLOCKS blocks; forward only where the level has at most WIDE_MAX_SN
            // supernodes (with more, the grouped forward below is faster)
            const bool big = maxnc >= WIDE_MIN_BLOCKS * TILE;
            const bool wf = big && static_cast<int>(lv.size()) <= WIDE_MAX_SN, wb = big;
            std::vector<lkd_task> wlt, wlb;
            if (wf || wb) {
                std::vector<lkd_task> heads;
                for (const Index s : lv) {
                    lkd_task t{};
                    t.off = S.sn_panel[s];
                    t.ld = static_cast<int32_t>(S.sn_rowptr[s + 1] - S.sn_rowptr[s]);
                    t.ncol = S.sn_first[s + 1] - S.sn_first[s];
                    t.sn = s;
                    t.first = S.sn_first[s];
                    t.aux = diag_off(s, 0);
                    t.aux2 = S.sn_rowptr[s];
                    t.nsplit = 0;
                    t.c0 = 1;
                    t.px = nflags;
                    nflags += 2 * ((t.ncol + TILE - 1) / TILE);
                    heads.push_back(t);
                }
                for (int k = 0; k < maxnt; ++k)
                    for (const auto& h : heads) {
                        const int nb = (h.ncol + TILE - 1) / TILE;
                        const int nt = nb + static_cast<int>((h.ld - h.ncol + TILE - 1) / TILE);
                        lkd_task u = h;
                        if (wf && k < nt) {
                            u.r0 = k;
                            wlt.push_back(u);
                        }
                        if (wb && k < nb) {
                            u.r0 = nb - 1 - k;
                            wlb.push_back(u);
                        }
                    }
            }
            wide_f.push_back(wf);
            wide_b.push_back(wb);
            if (wf && wb) {
                per_level.push_back(std::move(wlt));
                per_level_bwd.push_back(std::move(wlb));
                continue;
            }
            // forward: SOLVE_MAX_G workgroups per wide supernode when the level's workgroups all
            // fit on the compute units (one each); kernels exist for G = 1 and 4
            int G = 1;
            if (nw == 8 && cus > 0 && static_cast<int64_t>(lv.size()) * SOLVE_MAX_G <= cus &&
                maxnt >= 2 * nw)
                G = SOLVE_MAX_G;
            std::vector<lkd_task> lt, lb;
            for (const Index s : lv) {
                lkd_task t{};
                t.off = S.sn_panel[s];
                t.ld = static_cast<int32_t>(S.sn_rowptr[s + 1] - S.sn_rowptr[s]);
                t.ncol = S.sn_first[s + 1] - S.sn_first[s];
                t.sn = s;
                t.first = S.sn_first[s];
                t.aux = diag_off(s, 0);
                t.aux2 = S.sn_rowptr[s];
                t.nsplit = nw;
                t.c0 = G;
                t.ws = npbuf;
                t.px = nflags;
                t.kbeg = static_cast<int32_t>(nbelow);
                const int64_t nb = (t.ncol + TILE - 1) / TILE;
                npbuf += nb * G * nw * TILE;
                nflags += nb + nb * G;
                nbelow += t.ld - t.ncol;
                for (int g = 0; g < G; ++g) {
                    lkd_task u = t;
                    u.r0 = g;
                    lt.push_back(u);
                }
                const int Gb = G > 1 ? SOLVE_BWD_G : 1;
                for (int g = 0; g < Gb; ++g) {
                    lkd_task u = t;
                    u.c0 = Gb;
                    u.r0 = g;
                    lb.push_back(u);
                }
            }
            per_level.push_back(wf ? std::move(wlt) : std::move(lt));
            per_level_bwd.push_back(wb ? std::move(wlb) : std::move(lb));
        }
        auto emit = [&](int kind, const std::vector<lkd_task>& lt) {
            if (lt.empty()) return;
            slaunch.push_back({kind, static_cast<int64_t>(tasks.size()), static_cast<int32_t>(lt.size()),
                               lt.empty() ? 4 : lt[0].nsplit, lt.empty() ? 1 : lt[0].c0});
            const int G = slaunch.back().g;
            squeue.resize(tasks.size(), -1);
            if (kind >= 22) {   // wide: one ticket counter (queue 0), tasks in list order
                Launch& L = slaunch.back();
                L.q = n_sq++;
                for (int x = 0; x <= n_xcd; ++x) sqoff.push_back(x == 0 ? 0 : static_cast<int32_t>(lt.size()));
                for (size_t i = 0; i < lt.size(); ++i) squeue.push_back(static_cast<int32_t>(i));
                L.grid = static_cast<int32_t>(lt.size());
            } else if (G > 1) {   // XCD queues: supernodes dealt in turn, each one's G tasks together
                Launch& L = slaunch.back();
                L.q = n_sq++;
                const int nsn = static_cast<int>(lt.size()) / G;
                int32_t longest = 0;
                for (int x = 0; x < n_xcd; ++x) {
                    sqoff.push_back(static_cast<int32_t>(squeue.size() - tasks.size()));
                    int32_t len = 0;
                    for (int sn = x; sn < nsn; sn += n_xcd)
                        for (int g = 0; g < G; ++g, ++len) squeue.push_back(sn * G + g);
                    longest = std::max(longest, len);
                }
                sqoff.push_back(static_cast<int32_t>(lt.size()));
                L.grid = n_xcd * longest;
            }
            tasks.insert(tasks.end(), lt.begin(), lt.end());
            squeue.resize(tasks.size(), -1);
        };
        for (size_t l = 0; l < per_level.size(); ++l) emit(wide_f[l] ? 22 : 20, per_level[l]);   // forward
        for (size_t l = per_level_bwd.size(); l-- > 0;) emit(wide_b[l] ? 23 : 21, per_level_bwd[l]);   // backward
        solve_pbuf = npbuf;
        solve_flags = nflags;
        solve_below = nbelow;
        for (const auto& L : slaunch) {
            for (int64_t q = L.begin; q < L.begin + L.count; ++q) validate(tasks[q], L.kind);
            if (L.kind >= 22) {   // wide: every task once, each task's producers earlier
                std::map<std::pair<int32_t, int32_t>, int32_t> pos;
                for (int32_t i = 0; i < L.count; ++i) {
                    const lkd_task& t = tasks[L.begin + i];
                    if (squeue[static_cast<size_t>(L.begin + i)] != i || !pos.emplace(std::make_pair(t.sn, t.r0), i).second)
                        throw std::logic_error("KKT plan: invalid wide solve queue");
                }
                for (int32_t i = 0; i < L.count; ++i) {
                    const lkd_task& t = tasks[L.begin + i];
                    const int nb = (t.ncol + TILE - 1) / TILE;
                    const int lo = L.kind == 22 ? 0 : t.r0 + 1, hi = L.kind == 22 ? std::min(t.r0, nb) : nb;
                    for (int j = lo; j < hi; ++j) {
                        auto it = pos.find(std::make_pair(t.sn, static_cast<int32_t>(j)));
                        if (it == pos.end() || it->second >= i) throw std::logic_error("KKT plan: wide solve order");
                    }
                }
                continue;
            }
            if (L.g == 1) continue;
            // queues: every task once; a supernode's tasks consecutive, in one queue, g in order
            const int32_t* off = sqoff.data() + static_cast<size_t>(L.q) * (n_xcd + 1);
            std::vector<char> seen(static_cast<size_t>(L.count), 0);
            int32_t longest = 0;
            for (int x = 0; x < n_xcd; ++x) {
                longest = std::max(longest, off[x + 1] - off[x]);
                for (int32_t p = off[x]; p < off[x + 1]; ++p) {
                    const int32_t i = squeue[static_cast<size_t>(L.begin + p)];
                    if (i < 0 || i >= L.count || seen[i]) throw std::logic_error("KKT plan: invalid solve queue");
                    seen[i] = 1;
                    const lkd_task& t = tasks[L.begin + i];
                    if (t.c0 != L.g || t.r0 != (p - off[x]) % L.g ||
                        (t.r0 > 0 && tasks[L.begin + squeue[L.begin + p - 1]].sn != t.sn))
                        throw std::logic_error("KKT plan: solve queue splits a supernode");
                }
            }
            if (off[0] != 0 || off[n_xcd] != L.count || L.grid < n_xcd * longest)
                throw std::logic_error("KKT plan: invalid solve queue offsets");
        }
        d_stasks = up(tasks, "stasks");
        d_squeue = up(squeue, "squeue");
        d_sqoff = up(sqoff, "sqoff");
        n_stasks = static_cast<int64_t>(tasks.size());
    }
    int64_t solve_pbuf = 0, solve_below = 0, solve_flags = 0;
    std::vector<int32_t> squeue, sqoff;    // XCD queues of the multi-workgroup level solves
    int32_t *d_squeue = nullptr, *d_sqoff = nullptr;
    int32_t n_sq = 0;                      // level solves with XCD queues

    // reference CSR layout of L (sym.L_pattern rows + diagonal last) -> panel indices
    void build_csr_index(const ipm::SymbolicFactorization& sym) {
        const auto& Lp = sym.L_pattern;
        if (static_cast<Index>(Lp.row_ptrs.size()) != m + 1) return;
        // sym ordering index -> build (postordered) index
        std::vector<Index> q(static_cast<size_t>(m));
        for (Index i = 0; i < m; ++i) q[i] = pinv[sym.P[i]];
        std::vector<int64_t> idx;
        idx.reserve(Lp.col_indices.size() + static_cast<size_t>(m));
        for (Index i = 0; i < m; ++i) {
            for (Index p = Lp.row_ptrs[i]; p < Lp.row_ptrs[i + 1]; ++p) {
                const Index j = Lp.col_indices[p];
                Index a = q[i], b = q[j];
                if (a < b) std::swap(a, b);
                int64_t pi = -1;
                try { pi = panel_index(a, b); } catch (...) { pi = -1; }
                idx.push_back(pi < 0 ? 0 : pi);
            }
            idx.push_back(panel_index(q[i], q[i]));
        }
        nnz_csr = static_cast<int64_t>(idx.size());
        d_csr_index = up(idx, "csr_index");
    }

    void allocate_numeric() {
        const int64_t pe = std::max<int64_t>(S.sn_panel[S.nsuper], 1);
        d_panel = arena_p->allocate_uninitialized(static_cast<size_t>(pe) * sizeof(double));
        owned.push_back(d_panel);
        d_diag = arena_p->allocate_uninitialized(static_cast<size_t>(std::max<int64_t>(diag_elems, 1)) * sizeof(double));
        owned.push_back(d_diag);
        d_piv = arena_p->allocate_uninitialized(static_cast<size_t>(std::max<Index>(m, 1)) * sizeof(double));
        owned.push_back(d_piv);
        d_stage = arena_p->allocate_uninitialized(static_cast<size_t>(std::max<int64_t>(stage_elems, 1)) * sizeof(double));
        owned.push_back(d_stage);
        d_work = arena_p->allocate_uninitialized(static_cast<size_t>(std::max<int64_t>(work_elems, 1)) * sizeof(double));
        owned.push_back(d_work);
        if (n_dag > 0) {   // zero-filled: epoch 0 is never current
            d_dag_flags = static_cast<uint32_t*>(arena_p->allocate(static_cast<size_t>(n_ftasks) * sizeof(uint32_t)));
            owned.push_back(d_dag_flags);
            d_dag_flags2 = static_cast<uint32_t*>(arena_p->allocate(static_cast<size_t>(n_ftasks) * sizeof(uint32_t)));
            owned.push_back(d_dag_flags2);
            d_dag_ctl = static_cast<uint32_t*>(arena_p->allocate(static_cast<size_t>(1 + n_dag * n_xcd) * sizeof(uint32_t)));
            owned.push_back(d_dag_ctl);
        }
        d_below = static_cast<double*>(arena_p->allocate(static_cast<size_t>(std::max<int64_t>(solve_below, 1)) * sizeof(double)));
        owned.push_back(d_below);
        d_pbuf = static_cast<double*>(arena_p->allocate(static_cast<size_t>(std::max<int64_t>(solve_pbuf, 1)) * sizeof(double)));
        owned.push_back(d_pbuf);
        d_flags = static_cast<uint32_t*>(arena_p->allocate(static_cast<size_t>(solve_flags + 1 + n_sq * n_xcd) * sizeof(uint32_t)));
        owned.push_back(d_flags);   // flags [0, solve_flags), epoch [solve_flags], queue counters; zero-filled
        auto dv = [&](int64_t len) {
            double* p = static_cast<double*>(arena_p->allocate(static_cast<size_t>(std::max<int64_t>(len, 1)) * sizeof(double)));
            owned.push_back(p);
            return p;
        };
        d_err = static_cast<int32_t*>(arena_p->allocate(sizeof(int32_t) * 4));
        owned.push_back(d_err);
        d_D = dv(m);
        d_m0d = dv(m);
        d_mval = dv(nseg);
        d_theta = dv(n);
        d_r = dv(m);
        d_req = dv(m);
        d_z = dv(m);
        d_dy = dv(m);
        d_bak = dv(m);
        d_rhsp = dv(m);
        d_u = dv(n);
        // the level-solve status word sits just before the kS slots, so one copy reads both
        double* ks_block = dv(1 + LKD_KS_PART + 8 * LKD_MAX_NBLK);
        d_serr = reinterpret_cast<int32_t*>(ks_block);
        d_rinf = ks_block + 1;
        d_pcg = dv(4 * static_cast<int64_t>(m));
        if (nnz_csr > 0) d_csr = dv(nnz_csr);
        std::vector<int32_t> sn_first(S.sn_first.begin(), S.sn_first.end());
        d_sn_first = dev_upload(*arena_p, sn_first, owned);
        d_sn_rowptr = dev_upload(*arena_p, S.sn_rowptr, owned);
        d_sn_rows = dev_upload(*arena_p, S.sn_rows, owned);
        d_sn_panel = dev_upload(*arena_p, S.sn_panel, owned);
        d_col_to_sn = dev_upload(*arena_p, S.col_to_sn, owned);
        check(lkd_stream_sync(nullptr), "plan upload");
    }

    lkd_snplan plan() const {
        lkd_snplan p;
        p.nsuper = S.nsuper;
        p.sn_first = d_sn_first;
        p.sn_rowptr = d_sn_rowptr;
        p.sn_rows = d_sn_rows;
        p.sn_panel = d_sn_panel;
        p.col_to_sn = d_col_to_sn;
        return p;
    }

    // -------------------------------------------------------------------- numeric phases
    void run_factor_schedule(int prec) {
        const lkd_snplan pl = plan();
        if (n_dag > 0) check(lkd_chol_dag_begin(d_dag_ctl, n_dag * n_xcd, stream), "steps begin");
        if (!timer.on) {
            for (const auto& L : flaunch) check(launch_factor(prec, L, pl), "factor kernel");
            check_factor_queues();
            return;
        }
        // one event at every change of launch kind: a segment of consecutive launches of one
        // kind is timed as a whole (its launches follow each other on the stream)
        const size_t need = flaunch.size() + 1;
        if (timer.used + need > timer.events.size()) {
            check(lkd_stream_sync(stream), "timing sync");
            timer_collect();
            while (timer.events.size() < need) {
                void* e = nullptr;
                check(lkd_event_create(&e), "event");
                timer.events.push_back(e);
            }
        }
        const size_t first = timer.used;
        check(lkd_event_record(timer.events[timer.used++], stream), "event");
        for (size_t i = 0; i < flaunch.size(); ++i) {
            const int slot = slot_of(flaunch[i]);
            if (i == 0 || slot != slot_of(flaunch[i - 1]))
                timer.pending.push_back({slot, timer.used - 1, 0, 0, 0.0});
            check(launch_factor(prec, flaunch[i], pl), "factor kernel");
            auto& seg = timer.pending.back();
            seg.count += 1;
            seg.flops += launch_flops[i];
            if (i + 1 == flaunch.size() || slot_of(flaunch[i + 1]) != slot) {
                check(lkd_event_record(timer.events[timer.used++], stream), "event");
                seg.end = timer.used - 1;
            }
        }
        timer.pending.push_back({4, first, timer.used - 1, 1, S.flops_exact});
        check_factor_queues();
    }

    // Every ticket of every STEPS queue must have been taken (error bit 0x40 otherwise): a task
    // nobody waits on would otherwise be skipped without a trace.
    void check_factor_queues() {
        if (n_dag > 0)
            check(lkd_queue_check(d_dag_ctl + 1, d_fqoff, n_dag, n_xcd, 0x40, d_err, stream), "queue check");
    }

    // 0 STEP, 1 DIAG + DIAGWB, 2 UPDATE, 5 PART (3 assembly and 4 totals are not launch kinds)
    static int slot_of(const Launch& L) {
        return L.kind == 5 || L.kind == 7 ? 0 : L.kind == 6 || L.kind == 8 ? 5 : L.kind == 2 ? 2 : 1;
    }

    // Live launch timing: HIP events on the solver's stream at every change of launch kind in
    // the factorization schedule, resolved after the stream has drained.  Slots as
    // profile_factor (0 STEP, 1 DIAG + DIAGWB, 2 UPDATE, 5 PART; 4 = whole factorizations).
    struct Segment {
        int slot;
        size_t begin, end;
        int count;
        double flops;
    };
    struct LaunchTimer {
        bool on = false;
        std::vector<void*> events;
        size_t used = 0;
        std::vector<Segment> pending;
        double ms[6] = {}, flops[6] = {}, launches[6] = {};
    } timer;
    std::vector<double> launch_flops;   // per factorization launch, stored (panel) flops

    void timer_collect() {
        for (const auto& sg : timer.pending) {
            float t = 0;
            check(lkd_event_elapsed_ms(&t, timer.events[sg.begin], timer.events[sg.end]), "elapsed");
            timer.ms[sg.slot] += t;
            timer.launches[sg.slot] += sg.count;
            timer.flops[sg.slot] += sg.flops;
        }
        timer.pending.clear();
        timer.used = 0;
    }

    void set_timing(bool on) {
        if (launch_flops.size() != flaunch.size()) {
            launch_flops.assign(flaunch.size(), 0.0);
            for (size_t i = 0; i < flaunch.size(); ++i)
                for (int64_t q = flaunch[i].begin; q < flaunch[i].begin + flaunch[i].count; ++q)
                    launch_flops[i] += task_flops(h_ftasks[static_cast<size_t>(q)], flaunch[i].kind);
        }
        check(lkd_stream_sync(stream), "timing sync");
        timer_collect();
        timer.on = on;
        for (int k = 0; k < 6; ++k) timer.ms[k] = timer.flops[k] = timer.launches[k] = 0;
    }

    void read_timing(double ms[6], double flops[6], double launches[6]) {
        check(lkd_stream_sync(stream), "timing sync");
        timer_collect();
        for (int k = 0; k < 6; ++k) {
            ms[k] = timer.ms[k];
            flops[k] = timer.flops[k];
            launches[k] = timer.launches[k];
        }
    }

    int launch_factor(int prec, const Launch& L, const lkd_snplan& pl) {
        const lkd_task* t = d_ftasks + L.begin;
        switch (L.kind) {
            case 1: return lkd_chol_diag(prec, L.count, t, d_panel, d_diag, d_piv, d_stage, d_err, stream);
            case 2: return lkd_chol_update(prec, L.count, t, d_panel, pl, d_err, stream);
            case 3:   // only the reference-layout export reads the panel's diagonal blocks (the
                      // solves use the row-major copies)
                return d_csr ? lkd_chol_diagwb(prec, L.count, t, d_panel, d_diag, d_piv, stream) : 0;
            case 6: return lkd_chol_part(prec, L.count, t, d_panel, d_diag, d_piv, d_stage, d_work, d_err, stream);
            case 8: return lkd_chol_psum(prec, L.count, t, d_work, stream);
            case 7: return lkd_chol_step_dag(prec, L.count, t, d_fdeps + 4 * L.begin, d_fsdep + L.begin,
                                             d_fqueue + L.begin, d_fqoff + static_cast<int64_t>(L.q) * (n_xcd + 1),
                                             n_xcd, L.grid, d_dag_flags + L.begin, d_dag_flags2 + L.begin,
                                             d_dag_ctl, L.q, d_panel, d_diag, d_piv, d_stage, d_work, d_err,
                                             stream);
            default: return lkd_chol_step(prec, L.count, t, d_panel, d_diag, d_piv, d_stage, d_work, d_err, stream);
        }
    }

    // One factorization of the current M with HIP events around every launch: per launch kind
    // (0 STEP, 1 DIAG + DIAGWB, 2 UPDATE) the summed device time, the launch count and the
    // algorithmic flops of the tasks; slot 3 = assembly (prep, diagonal, zero fill, pair sums);
    // slot 4 = whole factorization incl. assembly (flops: sum over columns of L of c_j^2).
    void profile_factor(int prec, double ms[6], double flops[6], double launches[6]) {
        for (int k = 0; k < 6; ++k) ms[k] = flops[k] = launches[k] = 0;
        if (!have_theta) throw std::logic_error("profile_factor: no Theta");
        const lkd_snplan pl = plan();
        void *e0, *e1, *a0, *a1;
        check(lkd_event_create(&e0), "event");
        check(lkd_event_create(&e1), "event");
        check(lkd_event_create(&a0), "event");
        check(lkd_event_create(&a1), "event");
        check(lkd_event_record(a0, stream), "event");
        check(lkd_event_record(e0, stream), "event");
        check(lkd_assemble_diag(m, d_dseg, d_segptr, d_pp, d_pk, d_theta, 1, d_D, d_m0d, mid_diag, long_diag,
                                stream), "diag");
        check(lkd_memset(d_err, 0, sizeof(int32_t), stream), "zero");
        check(lkd_assemble_pairs(prec, nseg, d_segptr, d_segdst, d_segi, d_segj, d_pp, d_pk, d_theta,
                                 d_D, d_empty, 1e-9, mid_segs, long_segs, d_mval, S.sn_panel[S.nsuper],
                                 d_chunk_seg, d_panel, stream), "assemble");
        check(lkd_event_record(e1, stream), "event");
        {
            float t_ms = 0;
            check(lkd_event_elapsed_ms(&t_ms, e0, e1), "elapsed");
            ms[3] = t_ms;
            launches[3] = 4;
            flops[3] = 2.0 * static_cast<double>(npairs);
        }
        // a new STEPS epoch and fresh queue counters, as in every factorization
        if (n_dag > 0) check(lkd_chol_dag_begin(d_dag_ctl, n_dag * n_xcd, stream), "steps begin");
        for (const auto& L : flaunch) {
            check(lkd_event_record(e0, stream), "event");
            check(launch_factor(prec, L, pl), "factor kernel");
            check(lkd_event_record(e1, stream), "event");
            float t_ms = 0;
            check(lkd_event_elapsed_ms(&t_ms, e0, e1), "elapsed");
            const int slot = slot_of(L);
            ms[slot] += t_ms;
            launches[slot] += 1;
            for (int64_t q = L.begin; q < L.begin + L.count; ++q) flops[slot] += task_flops(h_ftasks[q], L.kind);
        }
        check(lkd_event_record(a1, stream), "event");
        float tot = 0;
        check(lkd_event_elapsed_ms(&tot, a0, a1), "elapsed");
        ms[4] = tot;
        flops[4] = S.flops_exact;
        launches[4] = static_cast<double>(flaunch.size());
        lkd_event_destroy(e0); lkd_event_destroy(e1); lkd_event_destroy(a0); lkd_event_destroy(a1);
        have_eq = false;   // the profiled factor uses delta 1e-9 only; force a normal refactor next
    }
    std::vector<lkd_task> h_ftasks;
    // algorithmic flops of one task (lower-triangle work only)
    static double task_flops(const lkd_task& t, int kind) {
        const double kb = t.kb;
        switch (kind) {
            case 6:
                if (t.nsplit == 2) return kb * kb * kb / 3.0;
                return 2.0 * t.c0 * (t.part == 2 ? std::min(2 * TILE, t.ncol - t.k0) : kb) * (t.kend - t.kbeg);
            case 8: return 0.0;
            case 7:
                if (t.part) return 2.0 * t.c0 * kb * (t.kend - t.kbeg);
                [[fallthrough]];
            case 5: {   // GEMM + solve of the tile; diagonal-block updates; next block's factor
                const double ra = t.c0;
                double f = 2.0 * ra * kb * (t.kend - t.kbeg) + ra * kb * kb;
                for (int h = 0; h < 2; ++h) {
                    const int d0 = t.r0 + 64 * h;
                    if (64 * h >= t.c0 || d0 >= t.ncol) break;
                    const double db = std::min({64, t.c0 - 64 * h, t.ncol - d0});
                    f += db * (db + 1) * kb;
                }
                if (t.nsplit) f += ra * ra * ra / 3.0;
                return f;
            }
            case 1: return kb * kb * kb / 3.0;
            case 3: return 0.0;
            default: {
                const double ra = std::min(GEMM_MT, t.ld - t.r0), rb = std::min(GEMM_NC, t.ld - t.c0);
                double useful = 0;   // entries with row >= column
                for (int c = 0; c < rb; ++c) useful += std::max(0.0, ra - std::max(0, t.c0 + c - t.r0));
                return 2.0 * useful * (t.kend - t.kbeg);
            }
        }
    }

    void run_solve_schedule(int prec, double* x) {
        uint32_t* epoch = d_flags + solve_flags;
        check(lkd_trsv_epoch(epoch, epoch + 1, n_sq * n_xcd, stream), "solve epoch");
        for (const auto& L : slaunch) {
            if (L.kind >= 22) {
                check(lkd_trsv_wide(prec, L.kind == 22 ? 0 : 1, L.count, d_stasks + L.begin, d_squeue + L.begin,
                                    epoch + 1 + static_cast<int64_t>(L.q) * n_xcd, d_panel, d_diag, d_sn_rows, x,
                                    d_flags, epoch, d_serr, stream), "solve kernel");
                continue;
            }
            const int dir = L.kind == 20 ? 0 : 1;
            const bool q = L.g > 1;
            check(lkd_trsv_level(prec, dir, L.nw, L.g, L.count, d_stasks + L.begin, d_panel, d_diag, d_sn_rows,
                                 x, d_below, d_pbuf, d_flags, epoch, q ? d_squeue + L.begin : nullptr,
                                 q ? d_sqoff + static_cast<int64_t>(L.q) * (n_xcd + 1) : nullptr, n_xcd,
                                 q ? epoch + 1 + static_cast<int64_t>(L.q) * n_xcd : nullptr, L.grid,
                                 d_serr, stream), "solve kernel");
        }
        // every ticket of the multi-workgroup level queues was taken (bit 0x40 otherwise)
        if (n_sq > 0)
            check(lkd_queue_check(epoch + 1, d_sqoff, n_sq, n_xcd, 0x40, d_serr, stream), "queue check");
    }

    // A level solve whose waits timed out produced an invalid vector: report it.  The status
    // comes with the kS read when no level solve ran after it, else it is read here.
    void check_solve_sync() {
        int32_t e = h_serr;
        if (solve_unread) check(lkd_memcpy_dtoh(&e, d_serr, sizeof(int32_t), stream), "solve status");
        if (e == 0) return;
        h_serr = 0;
        check(lkd_memset(d_serr, 0, sizeof(int32_t), stream), "reset");
        check(lkd_stream_sync(stream), "reset");
        if (e & 0x40) throw std::runtime_error("KKT solve: a level-solve queue was not drained");
        throw std::runtime_error("KKT solve: level synchronisation timed out");
    }

    int32_t h_serr = 0;            // status word as of the last read_ks
    bool solve_unread = false;     // a level solve ran after the last read_ks
    void trsv(int prec, double* x) {
        run_solve_schedule(prec, x);
        solve_unread = true;
    }

    void factor_numeric(int prec) {
        run_factor_schedule(prec);
    }

    // assemble (D, delta) and factor; returns true when every pivot was positive
    bool assemble_and_factor(const double* theta, bool equil, double delta, int prec,
                             bool need_diag) {
        if (need_diag) {
            check(lkd_assemble_diag(m, d_dseg, d_segptr, d_pp, d_pk, theta, equil ? 1 : 0, d_D, d_m0d,
                                    mid_diag, long_diag, stream), "diag");
        }
        check(lkd_memset(d_err, 0, sizeof(int32_t), stream), "err zero");
        check(lkd_assemble_pairs(prec, nseg, d_segptr, d_segdst, d_segi, d_segj, d_pp, d_pk, theta,
                                 d_D, d_empty, delta, mid_segs, long_segs, d_mval, S.sn_panel[S.nsuper],
                                 d_chunk_seg, d_panel, stream), "assemble");
        factor_numeric(prec);
        int32_t err = 0;
        check(lkd_memcpy_dtoh(&err, d_err, sizeof(int32_t), stream), "err read");
        st.factorizations++;
        if (err & 2) throw std::logic_error("KKT factorization: update outside the L structure");
        if (err & 0x20) throw std::runtime_error("KKT factorization: a STEPS wait exceeded its poll limit");
        if (err & 0x40) throw std::runtime_error("KKT factorization: a STEPS queue was not drained");
        return err == 0;
    }

    void keep_theta(const double* theta) {
        if (theta != d_theta)
            check(lkd_memcpy_dtod(d_theta, theta, static_cast<size_t>(n) * sizeof(double), stream), "theta copy");
        have_theta = true;
    }

    int factor_equilibrated(const double* theta, FactorPrecision p) {
        keep_theta(theta);
        const int prec = p == FactorPrecision::FP32 ? LKD_F32 : LKD_F64;
        double delta = p == FactorPrecision::FP32 ? opt.delta0_fp32 : opt.delta0_fp64;
        have_eq = false;
        factor_kind = 0;
        for (int r = 0; r < opt.delta_retries; ++r) {
            if (r > 0) delta *= 10.0;
            if (assemble_and_factor(d_theta, true, delta, prec, r == 0)) {
                have_eq = true;
                factor_kind = 2;
                cached_prec = p;
                cached_delta = delta;
                retries_used = r;
                return r;
            }
        }
        return -1;
    }
    int retries_used = 0;

    // r = rhs_perm - M0 dy (M0 dy = A_p (theta o (A_p^T dy))), req = D r;
    // kS[0..3] = max|r|, max|rhs|, max|dy|, max M0_ii
    void residual() {
        check(lkd_kkt_residual(m, n, d_colptr, d_prow, d_cval, d_prp, d_pci, d_pval, d_theta,
                               d_dy, d_rhsp, d_m0d, d_u, d_r, d_D, d_req, d_rinf, long_prows, long_cols,
                               stream), "residual");
    }
    double hk[LKD_KS_SLOTS];
    void read_ks() {   // the status word and the kS slots in one copy
        double blk[1 + LKD_KS_SLOTS];
        check(lkd_memcpy_dtoh(blk, d_rinf - 1, sizeof(blk), stream), "kkt scalars");
        std::memcpy(&h_serr, blk, sizeof(int32_t));
        std::copy(blk + 1, blk + 1 + LKD_KS_SLOTS, hk);
        solve_unread = false;
    }
    // backward-error acceptance (relative-gap mode): ||r|| <= 1e-12 (||rhs|| + max M0_ii ||dy||)
    bool backward_ok(double rinf) const { return rinf <= 1e-12 * (hk[1] + hk[3] * hk[2]); }

    // Direct solve with the cached factor, then preconditioned CG on M0 dy = rhs (preconditioner
    // D (L L^T)^{-1} D with the cached FP64 or FP32 factor).  Residual replacement every 10 steps
    // and before accepting; the best true-residual iterate is kept.
    bool solve_pcg(double* d_rhs, double tol, int max_it, bool backward, KKTSolveInfo* info) {
        if (!have_eq) throw std::logic_error("KKT solve: no factorization");
        const int prec = cached_prec == FactorPrecision::FP32 ? LKD_F32 : LKD_F64;
        const size_t vb = static_cast<size_t>(m) * sizeof(double);
        double* d_p = d_pcg;
        double* d_q = d_pcg + m;
        double* d_best = d_pcg + 2 * static_cast<int64_t>(m);
        st.solves++;
        check(lkd_kkt_gather(m, d_perm, d_rhs, d_D, d_rhsp, d_req, stream), "gather");
        check(lkd_memcpy_dtod(d_z, d_req, vb, stream), "copy");
        trsv(prec, d_z);
        check(lkd_kkt_update(m, d_D, d_z, d_dy, nullptr, 0, stream), "update");
        residual();
        read_ks();
        double best = hk[0];
        double bw_hk[LKD_KS_SLOTS];
        std::copy(hk, hk + LKD_KS_SLOTS, bw_hk);   // norms at the best iterate
        bool ok = best <= tol;
        int its = 0, stall = 0;
        if (!ok) {
            check(lkd_memcpy_dtod(d_best, d_dy, vb, stream), "copy");
            check(lkd_memcpy_dtod(d_z, d_req, vb, stream), "copy");
            trsv(prec, d_z);
            int sa = 5, sb = 7;   // rz slots (ping-pong), pq slot 6
            check(lkd_pcg_precdot(m, d_D, d_z, d_r, d_rinf, sa, stream), "precdot");
            check(lkd_memcpy_dtod(d_p, d_z, vb, stream), "copy");
            for (int it = 0; it < max_it; ++it) {
                its = it + 1;
                check(lkd_pcg_matvec(m, n, d_colptr, d_prow, d_cval, d_prp, d_pci, d_pval, d_theta,
                                     d_p, d_u, d_q, d_rinf, 6, long_prows, long_cols, stream), "matvec");
                check(lkd_pcg_update(m, d_rinf, sa, 6, d_p, d_q, d_dy, d_r, d_D, d_req, stream), "pcg update");
                read_ks();
                if (!(hk[6] > 0.0)) break;
                const double rr = hk[4];
                if (rr <= tol || it == max_it - 1 || (it + 1) % 5 == 0) {
                    residual();
                    read_ks();
                    const double rt = hk[0];
                    if (rt < 0.5 * best) stall = 0; else ++stall;
                    if (rt < best) {
                        best = rt;
                        std::copy(hk, hk + LKD_KS_SLOTS, bw_hk);
                        check(lkd_memcpy_dtod(d_best, d_dy, vb, stream), "copy");
                    }
                    if (rt <= tol) { ok = true; break; }
                    if (stall >= 2) break;     // stagnated: accept at the backward level
                }
                check(lkd_memcpy_dtod(d_z, d_req, vb, stream), "copy");
                trsv(prec, d_z);
                check(lkd_pcg_precdot(m, d_D, d_z, d_r, d_rinf, sb, stream), "precdot");
                check(lkd_pcg_dir(m, d_rinf, sb, sa, d_z, d_p, stream), "pcg dir");
                std::swap(sa, sb);
            }
            if (!ok) check(lkd_memcpy_dtod(d_dy, d_best, vb, stream), "copy");
        }
        if (!ok && backward) {
            std::copy(bw_hk, bw_hk + LKD_KS_SLOTS, hk);
            ok = backward_ok(best);
        }
        st.refinement_steps += its;
        check_solve_sync();
        check(lkd_kkt_scatter(m, d_perm, d_dy, d_rhs, stream), "scatter");
        if (info) {
            info->retries = retries_used;
            info->refinement_steps = its;
            info->pcg_iterations = its;
            info->delta = cached_delta;
            info->residual_inf = best;
        }
        return ok;
    }

    bool solve_cached(double* d_rhs, double tol, KKTSolveInfo* info, int max_refine) {
        if (!have_eq) throw std::logic_error("KKT solve: no factorization");
        const int prec = cached_prec == FactorPrecision::FP32 ? LKD_F32 : LKD_F64;
        st.solves++;
        check(lkd_kkt_gather(m, d_perm, d_rhs, d_D, d_rhsp, d_req, stream), "gather");
        check(lkd_memcpy_dtod(d_z, d_req, static_cast<size_t>(m) * sizeof(double), stream), "copy");
        trsv(prec, d_z);
        check(lkd_kkt_update(m, d_D, d_z, d_dy, nullptr, 0, stream), "update");
        residual();
        read_ks();
        double before = hk[0];
        bool ok = before <= tol;
        int steps = 0;
        double after = before;
        if (!ok) {
            for (int it = 0; it < max_refine; ++it) {
                check(lkd_memcpy_dtod(d_z, d_req, static_cast<size_t>(m) * sizeof(double), stream), "copy");
                trsv(prec, d_z);
                check(lkd_kkt_update(m, d_D, d_z, d_dy, d_bak, 1, stream), "update");
                residual();
                read_ks();
                after = hk[0];
                ++steps;
                if (after <= tol) { ok = true; break; }
                if (after >= before) {
                    check(lkd_memcpy_dtod(d_dy, d_bak, static_cast<size_t>(m) * sizeof(double), stream), "rollback");
                    after = before;
                    break;
                }
                before = after;
            }
        }
        st.refinement_steps += steps;
        check_solve_sync();
        check(lkd_kkt_scatter(m, d_perm, d_dy, d_rhs, stream), "scatter");
        if (info) {
            info->retries = retries_used;
            info->refinement_steps = steps;
            info->delta = cached_delta;
            info->residual_inf = after;
        }
        return ok;
    }
};

// ------------------------------------------------------------------------------------------
GPUKKTCholeskySolver::GPUKKTCholeskySolver(VRAMArena& arena, const core::CSRMatrix& A,
                                           const ipm::SymbolicFactorization& sym,
                                           const core::CSRMatrix& M_pattern)
    : impl_(nullptr) {
    if (M_pattern.rows != A.rows || M_pattern.cols != A.rows)
        throw std::invalid_argument("GPUKKTCholeskySolver: M_pattern dimension mismatch");
    impl_.reset(new Impl(&arena, A, &sym, nullptr, KKTOptions{}));
}

GPUKKTCholeskySolver::GPUKKTCholeskySolver(VRAMArena& arena, const core::CSRMatrix& A,
                                           const ipm::SymbolicFactorization& sym,
                                           const KKTOptions& opt)
    : impl_(new Impl(&arena, A, &sym, nullptr, opt)) {}

GPUKKTCholeskySolver::GPUKKTCholeskySolver(VRAMArena& arena, const core::CSRMatrix& A,
                                           ipm::SupernodalStructure&& structure,
                                           const KKTOptions& opt)
    : impl_(new Impl(&arena, A, nullptr, &structure, opt)) {}

std::map<std::string, std::vector<char>> GPUKKTCholeskySolver::export_plan(
    const core::CSRMatrix& A, const std::vector<Index>& P, const ipm::SupernodeOptions& so,
    int step_dag) {
    std::map<std::string, std::vector<char>> out;
    ipm::SymbolicFactorization sym;
    ipm::compute_symbolic_factorization(A, P, sym);
    KKTOptions ko;
    ko.supernodes = so;
    ko.step_dag = step_dag;
    Impl impl(nullptr, A, &sym, nullptr, ko, &out);
    return out;
}

std::size_t GPUKKTCholeskySolver::estimate_bytes(const core::CSRMatrix& A,
                                                 const ipm::SupernodalStructure& S) {
    const int64_t m = A.rows, n = A.cols, nz = A.row_ptrs.empty() ? 0 : A.row_ptrs[A.rows];
    std::vector<int64_t> cnt(static_cast<size_t>(n), 0);
    for (int64_t p = 0; p < nz; ++p) cnt[A.col_indices[p]]++;
    int64_t pairs = 0;
    for (int64_t c : cnt) pairs += c * (c + 1) / 2;
    int64_t diag = 0, tasks = 0, slots = 0, step_tiles = 0;
    for (Index s = 0; s < S.nsuper; ++s) {
        const int64_t nc = S.sn_first[s + 1] - S.sn_first[s];
        const int64_t ld = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
        diag += nc * 64;
        slots += std::max<int64_t>(2, (nc + 63) / 64);   // staging slots, all levels at once
        const int64_t bc = (nc + 63) / 64, br = (ld + 63) / 64;
        step_tiles += bc * (br + 2);
        tasks += 4 + bc * (br + 4) + bc * br * (br + 1) / 2 + br * br;
    }
    const int64_t bytes = 8 * S.sn_panel[S.nsuper] + 8 * diag + 20 * pairs + 36 * pairs +
                          2 * (16 * nz + 8 * (m + n)) + 8 * nz + 8 * 24 * (m + n) +
                          128 * tasks + 8 * 4096 + static_cast<int64_t>(S.sn_rows.size()) * 4 +
                          8 * LKD_STAGE * slots + 8 * int64_t(STEP_TARGET_WG) * LKD_PART +
                          8 * int64_t(LKD_PART) * step_tiles +   // STEPS: PART products
                          24 * static_cast<int64_t>(S.nsuper) + 8 * m + 256 * 96;
    return static_cast<std::size_t>(bytes);
}

GPUKKTCholeskySolver::~GPUKKTCholeskySolver() = default;

void GPUKKTCholeskySolver::gpu_cholesky_factorize(const core::CSRMatrix& A,
                                                  const std::vector<Float>& Theta) {
    (void)A;
    if (static_cast<Index>(Theta.size()) != impl_->n)
        throw std::invalid_argument("Theta dimension mismatch");
    check(lkd_memcpy_htod(impl_->d_theta, Theta.data(), Theta.size() * sizeof(double),
                          impl_->stream), "theta upload");
    gpu_cholesky_factorize_device(impl_->d_theta);
}

void GPUKKTCholeskySolver::gpu_cholesky_factorize_device(const Float* d_Theta) {
    auto& I = *impl_;
    I.keep_theta(d_Theta);
    I.have_eq = false;
    const bool ok = I.assemble_and_factor(I.d_theta, false, 1e-9, LKD_F64, true);
    I.factor_kind = 1;
    if (I.d_csr) check(lkd_panel_to_csr(LKD_F64, I.nnz_csr, I.d_csr_index, I.d_panel, I.d_csr, I.stream), "L gather");
    check(lkd_stream_sync(I.stream), "sync");
    if (!ok) throw std::runtime_error("Numerical factorization failed: matrix is not positive definite.");
}

bool GPUKKTCholeskySolver::gpu_cholesky_solve(std::vector<Float>& rhs, Float tol) {
    if (static_cast<Index>(rhs.size()) != impl_->m)
        throw std::invalid_argument("rhs dimension mismatch");
    auto& I = *impl_;
    double* d = I.d_pcg;   // 4m scratch; first m used as the rhs buffer
    check(lkd_memcpy_htod(d, rhs.data(), rhs.size() * sizeof(double), I.stream), "rhs upload");
    const bool ok = gpu_cholesky_solve_device(d, tol);
    check(lkd_memcpy_dtoh(rhs.data(), d, rhs.size() * sizeof(double), I.stream), "rhs download");
    return ok;
}

bool GPUKKTCholeskySolver::gpu_cholesky_solve_device(Float* d_rhs, Float tol) {
    auto& I = *impl_;
    if (!I.have_theta) throw std::logic_error("gpu_cholesky_solve_device: factorize first");
    if (!I.have_eq || I.cached_prec != FactorPrecision::FP64) {
        if (I.factor_equilibrated(I.d_theta, FactorPrecision::FP64) < 0) {
            // every delta failed: dy = 0 (reference returns the zeroed dy)
            check(lkd_memset(I.d_dy, 0, static_cast<size_t>(I.m) * sizeof(double), I.stream), "zero");
            check(lkd_kkt_scatter(I.m, I.d_perm, I.d_dy, d_rhs, I.stream), "scatter");
            check(lkd_stream_sync(I.stream), "sync");
            return false;
        }
    }
    const bool ok = I.solve_cached(d_rhs, tol, nullptr, 10);
    check(lkd_stream_sync(I.stream), "sync");
    return ok;
}

Float* GPUKKTCholeskySolver::get_device_L_values() const { return impl_->d_csr; }

void GPUKKTCholeskySolver::set_regularization(double delta0_fp64, double delta0_fp32, int retries) {
    if (!(delta0_fp64 > 0.0) || !(delta0_fp32 > 0.0) || retries < 1)
        throw std::invalid_argument("set_regularization: delta0 > 0 and retries >= 1 required");
    impl_->opt.delta0_fp64 = delta0_fp64;
    impl_->opt.delta0_fp32 = delta0_fp32;
    impl_->opt.delta_retries = retries;
}

int GPUKKTCholeskySolver::factor_equilibrated(const Float* d_Theta, FactorPrecision prec) {
    return impl_->factor_equilibrated(d_Theta, prec);
}

bool GPUKKTCholeskySolver::solve_cached(Float* d_rhs, Float tol, KKTSolveInfo* info,
                                        int max_refine) {
    return impl_->solve_cached(d_rhs, tol, info, max_refine);
}

bool GPUKKTCholeskySolver::solve_pcg(Float* d_rhs, Float tol, int max_iter, KKTSolveInfo* info,
                                     bool backward_error_acceptance) {
    return impl_->solve_pcg(d_rhs, tol, max_iter, backward_error_acceptance, info);
}

void GPUKKTCholeskySolver::triangular_solve_permuted(Float* d_x) {
    auto& I = *impl_;
    const int prec = I.cached_prec == FactorPrecision::FP32 && I.factor_kind == 2 ? LKD_F32 : LKD_F64;
    I.trsv(prec, d_x);
    check(lkd_stream_sync(I.stream), "sync");
    I.check_solve_sync();
}

const KKTStats& GPUKKTCholeskySolver::stats() const { return impl_->st; }
void GPUKKTCholeskySolver::set_launch_timing(bool on) { impl_->set_timing(on); }
void GPUKKTCholeskySolver::launch_timing(double ms[6], double flops[6], double launches[6]) {
    impl_->read_timing(ms, flops, launches);
}
void GPUKKTCholeskySolver::profile_factor(FactorPrecision prec, double ms[6], double flops[6],
                                          double launches[6]) {
    impl_->profile_factor(prec == FactorPrecision::FP32 ? LKD_F32 : LKD_F64, ms, flops, launches);
}
const std::vector<Index>& GPUKKTCholeskySolver::permutation() const { return impl_->perm; }
void* GPUKKTCholeskySolver::stream() const { return impl_->stream; }
Index GPUKKTCholeskySolver::rows() const { return impl_->m; }
Index GPUKKTCholeskySolver::cols() const { return impl_->n; }

}  // namespace gpu
}  // namespace sankhya

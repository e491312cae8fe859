// MehrotraSolver: the interior-point loop of src/ipm/mehrotra_device.cu:695-1047 (reference) on
// device-resident vectors.  Every vector operation of an iteration runs in a fused kernel of the
// device layer; scalars (norms, mu, step bounds, all backtracking trials) come back in one small
// read per decision point instead of one Thrust reduction + host sync each.
//
// Same algorithm and constants as the reference (see include/ipm/mehrotra.hpp).  Deliberate
// differences, all numerically equivalent in exact arithmetic:
//  * one factorization of D*M0*D + delta*I per iteration, reused by predictor, corrector and
//    centred solves (the reference refactorizes the same matrix in every solve call);
//  * the reference's extra unequilibrated factorization of M0 + 1e-9 I at the start of each
//    iteration (which only decides whether solve() throws) is not performed;
//  * backtracking trials evaluate rp + a*(A dx) and rd + a*(A^T dy + ds) instead of recomputing
//    A(x + a dx) - b and A^T(y + a dy) + s + a ds - c per trial;
//  * Ruiz scaling, the fill-reducing ordering and the symbolic phase run once in the constructor
//    as in the reference, with AMD instead of the O(m^2) explicit minimum degree by default.
// Relative-gap mode (extension; the reference's loop stops at an absolute 1e-6): the factor is
// regularized from delta = 1e-14 instead of 1e-9, KKT solves use PCG preconditioned by it (a
// backward-error acceptance only once PCG stagnates), the e1 re-check of the KKT residual is
// dropped, the stop rule adds absolute residual and per-column complementarity targets, and when
// both merit searches fail the longest finite, positive centred trial is taken (non-monotone).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include "gpu/kkt.hpp"
#include "gpu/vram_arena.hpp"
#include "ipm/mehrotra.hpp"
#include "ipm/symbolic.hpp"
#include "lakshya_device.h"

namespace sankhya {
namespace ipm {

namespace {
void check(int e, const char* what) {
    if (e != 0)
        throw std::runtime_error(std::string("HIP call failed (") + what + "): " +
                                 lkd_error_string(e));
}
constexpr double LANE_DIV = 4.0;   // lanes per row of a sparse pass: mean entries / LANE_DIV

int pow2_group(double avg) {
    int g = 1;
    while (g < 64 && g < avg) g <<= 1;
    return g;
}
using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t) {
    return std::chrono::duration<double, std::milli>(clk::now() - t).count();
}
}  // namespace

void ruiz_equilibrate(core::CSRMatrix& A, std::vector<Float>& b, std::vector<Float>& c,
                      std::vector<Float>& R, std::vector<Float>& C) {
    const Index m = A.rows, n = A.cols;
    if (static_cast<Index>(b.size()) != m || static_cast<Index>(c.size()) != n)
        throw std::invalid_argument("ruiz_equilibrate: dimension mismatch");
    R.assign(static_cast<size_t>(m), 1.0);
    C.assign(static_cast<size_t>(n), 1.0);
    // per-pass factors: 1/sqrt(max |a|) over the row / column, 1 for (near) empty ones
    std::vector<Float> fr(static_cast<size_t>(m)), fc(static_cast<size_t>(n));
    for (int pass = 0; pass < 10; ++pass) {
        std::fill(fr.begin(), fr.end(), 0.0);
        std::fill(fc.begin(), fc.end(), 0.0);
        for (Index i = 0; i < m; ++i)
            for (Index p = A.row_ptrs[i]; p < A.row_ptrs[i + 1]; ++p) {
                const Float a = std::abs(A.values[p]);
                const Index j = A.col_indices[p];
                if (a > fr[i]) fr[i] = a;
                if (a > fc[j]) fc[j] = a;
            }
        for (Float& f : fr) f = f > 1e-8 ? 1.0 / std::sqrt(f) : 1.0;
        for (Float& f : fc) f = f > 1e-8 ? 1.0 / std::sqrt(f) : 1.0;
        for (Index i = 0; i < m; ++i) {
            R[i] *= fr[i];
            b[i] *= fr[i];
            for (Index p = A.row_ptrs[i]; p < A.row_ptrs[i + 1]; ++p)
                A.values[p] *= fr[i] * fc[A.col_indices[p]];
        }
        for (Index j = 0; j < n; ++j) {
            C[j] *= fc[j];
            c[j] *= fc[j];
        }
    }
}

class MehrotraSolver::Impl {
public:
    const core::Model& model_;
    IpmOptions opt_;
    Index m_ = 0, n_ = 0;
    core::CSRMatrix A_;              // scaled
    std::vector<Float> R_, C_, b_, c_;
    double binf_ = 0, cinf_ = 0;
    std::unique_ptr<gpu::VRAMArena> arena_;
    std::unique_ptr<gpu::GPUKKTCholeskySolver> kkt_;
    void* stream_ = nullptr;
    lkd_ipm P_{};
    double setup_ms_ = 0;

    // device vectors
    double *x_, *y_, *s_, *rp_, *rd_, *th_, *rxs_, *v_, *dy_, *dxa_, *dsa_, *dx_, *ds_, *qd_, *qp_;
    double *S_, *part_;
    std::vector<double> hS_;

    Impl(const core::Model& model, const IpmOptions& opt) : model_(model), opt_(opt) {
        auto t0 = clk::now();
        m_ = model.A.rows;
        n_ = model.A.cols;
        if (m_ <= 0 || n_ <= 0) throw std::invalid_argument("MehrotraSolver: empty model");
        A_ = core::to_csr(model.A);
        b_ = model.rhs;
        c_ = model.obj;
        for (Float v : b_) binf_ = std::max(binf_, std::abs(v));
        for (Float v : c_) cinf_ = std::max(cinf_, std::abs(v));
        ruiz();
        // ordering + supernodal structure (once)
        std::vector<Index> P = opt_.ordering == IpmOptions::Ordering::ReferenceMinimumDegree
                                   ? compute_amd_ordering(A_)
                                   : compute_fill_reducing_ordering(A_);
        SupernodalStructure S;
        compute_supernodal_structure(A_, P, SupernodeOptions{}, S);
        const size_t kkt_bytes = gpu::GPUKKTCholeskySolver::estimate_bytes(A_, S);
        const int64_t nz = A_.row_ptrs[m_];
        const int nb_rows = LKD_MAX_NBLK, nb_cols = LKD_MAX_NBLK;
        const size_t ipm_bytes = sizeof(double) * (20 * static_cast<size_t>(m_ + n_) +
                                                   2 * static_cast<size_t>(nz)) +
                                 sizeof(int32_t) * (2 * static_cast<size_t>(nz) + m_ + n_ + 2) +
                                 sizeof(double) * (LKS_TOTAL + 256 * LKD_MAX_NBLK) + (1u << 20);
        arena_.reset(new gpu::VRAMArena(kkt_bytes + ipm_bytes + (64u << 20)));
        check(lkd_stream_create(&stream_), "stream");
        gpu::KKTOptions ko;
        ko.stream = stream_;
        ko.use_graphs = opt_.use_graphs;
        if (opt_.stop_rule == IpmOptions::StopRule::RelativeGap) {
            // 1e-14 keeps the preconditioner exact down to eigenvalues ~1e-13 of D M0 D, so PCG
            // converges in a step or two late in the solve.  FP32 panels start at 1e-8 (about
            // the FP32 rounding of a unit diagonal; a breakdown retries at 10x): against 1e-6 it
            // cut the PCG steps of a block-angular solve (96,050 rows) from 153 to 86 and of the
            // 700 x 700 transport LP from 539 to 393
            ko.delta0_fp64 = 1e-14;
            ko.delta0_fp32 = 1e-8;
            ko.delta_retries = 10;
        }
        kkt_.reset(new gpu::GPUKKTCholeskySolver(*arena_, A_, std::move(S), ko));

        // IPM device data
        auto up_i = [&](const std::vector<Index>& v) {
            auto* p = static_cast<int32_t*>(arena_->allocate_uninitialized(v.size() * sizeof(int32_t)));
            check(lkd_memcpy_htod(p, v.data(), v.size() * sizeof(int32_t), stream_), "upload");
            return p;
        };
        auto up_d = [&](const std::vector<Float>& v) {
            auto* p = static_cast<double*>(arena_->allocate_uninitialized(std::max<size_t>(v.size(), 1) * sizeof(double)));
            check(lkd_memcpy_htod(p, v.data(), v.size() * sizeof(double), stream_), "upload");
            return p;
        };
        // rows / columns longer than LKD_LONG_ROW go to whole blocks; the lane-group width of the
        // other rows follows their mean length
        // (listed rows are summed in pieces of about LKD_LONG_PIECE entries by a pre-pass)
        auto long_list = [&](const std::vector<Index>& ptr, Index count, lkd_long& lg, double& mean) {
            std::vector<Index> idx;
            int64_t short_nz = 0, longest = 0;
            for (Index k = 0; k < count; ++k) {
                const int64_t len = ptr[k + 1] - ptr[k];
                if (len > LKD_LONG_ROW) {
                    idx.push_back(k);
                    longest = std::max(longest, len);
                } else {
                    short_nz += len;
                }
            }
            lg.n = static_cast<int32_t>(idx.size());
            mean = static_cast<double>(short_nz) / std::max<int64_t>(1, count - lg.n);
            lg.idx = idx.empty() ? nullptr : up_i(idx);
            lg.split = static_cast<int32_t>(std::min<int64_t>(LKD_LONG_MAX_SPLIT, (longest + LKD_LONG_PIECE - 1) / LKD_LONG_PIECE));
            lg.part = nullptr;
            if (lg.n > 0 && lg.split > 1)
                lg.part = static_cast<double*>(arena_->allocate_uninitialized(
                    static_cast<size_t>(lg.n) * static_cast<size_t>(lg.split) * sizeof(double)));
        };
        double mean_row = 0, mean_col = 0;
        long_list(A_.row_ptrs, m_, P_.long_rows, mean_row);
        core::CSCMatrix Acsc(m_, n_, {}, {}, {});
        {   // A^T in CSR = CSC of the scaled A
            std::vector<Index> cp(static_cast<size_t>(n_) + 1, 0), ri(static_cast<size_t>(nz));
            std::vector<Float> va(static_cast<size_t>(nz));
            for (int64_t p = 0; p < nz; ++p) cp[A_.col_indices[p] + 1]++;
            for (Index j = 0; j < n_; ++j) cp[j + 1] += cp[j];
            std::vector<Index> nx(cp.begin(), cp.end() - 1);
            for (Index i = 0; i < m_; ++i)
                for (Index p = A_.row_ptrs[i]; p < A_.row_ptrs[i + 1]; ++p) {
                    const Index q = nx[A_.col_indices[p]]++;
                    ri[q] = i;
                    va[q] = A_.values[p];
                }
            long_list(cp, n_, P_.long_cols, mean_col);
            P_.trp = up_i(cp);
            P_.tci = up_i(ri);
            P_.tval = up_d(va);
        }
        P_.m = m_;
        P_.n = n_;
        P_.arp = up_i(A_.row_ptrs);
        P_.aci = up_i(A_.col_indices);
        P_.aval = up_d(A_.values);
        P_.b = up_d(b_);
        P_.c = up_d(c_);
        P_.R = up_d(R_);
        P_.C = up_d(C_);
        P_.grp_rows = pow2_group(mean_row / LANE_DIV);
        P_.grp_cols = pow2_group(mean_col / LANE_DIV);
        P_.nblk_rows = std::max(1, std::min<int>(nb_rows, static_cast<int>((static_cast<int64_t>(m_) * P_.grp_rows + 255) / 256)));
        P_.nblk_cols = std::max(1, std::min<int>(nb_cols, static_cast<int>((static_cast<int64_t>(n_) * P_.grp_cols + 255) / 256)));
        auto dv = [&](int64_t len) {
            return static_cast<double*>(arena_->allocate(static_cast<size_t>(std::max<int64_t>(len, 1)) * sizeof(double)));
        };
        part_ = dv(256 * LKD_MAX_NBLK);   // up to 6 LKD_MAX_TRIALS + 1 quantities per pass
        S_ = dv(LKS_TOTAL);
        P_.part = part_;
        P_.S = S_;
        x_ = dv(n_); y_ = dv(m_); s_ = dv(n_); rp_ = dv(m_); rd_ = dv(n_); th_ = dv(n_);
        rxs_ = dv(n_); v_ = dv(n_); dy_ = dv(m_); dxa_ = dv(n_); dsa_ = dv(n_); dx_ = dv(n_);
        ds_ = dv(n_); qd_ = dv(n_); qp_ = dv(m_);
        hS_.assign(LKS_TOTAL, 0.0);
        check(lkd_stream_sync(stream_), "setup");
        setup_ms_ = ms_since(t0);
    }

    ~Impl() {
        if (stream_) lkd_stream_sync(stream_);
        kkt_.reset();
        arena_.reset();
        if (stream_) lkd_stream_destroy(stream_);
    }

    void ruiz() { ruiz_equilibrate(A_, b_, c_, R_, C_); }

    void read_scalars(int count) {
        check(lkd_memcpy_dtoh(hS_.data(), S_, sizeof(double) * count, stream_), "scalars");
    }

    struct Norms { double rp, rd, mu, cx, by, rpinf, rdinf, xsmax; };

    Norms residuals() {
        check(lkd_ipm_residuals(&P_, x_, y_, s_, rp_, rd_, th_, stream_), "residuals");
        read_scalars(LKS_NSLOTS);
        Norms r;
        r.rp = std::sqrt(hS_[LKS_RP2]);
        r.rd = std::sqrt(hS_[LKS_RD2]);
        r.mu = hS_[LKS_XS] / n_;
        r.cx = hS_[LKS_CX];
        r.by = hS_[LKS_BY];
        r.rpinf = hS_[LKS_RPINF];
        r.rdinf = hS_[LKS_RDINF];
        r.xsmax = hS_[LKS_XSMAX];
        return r;
    }

    struct Dir { double e1, e2, e3, apmax, admax; };

    // rhs, solve, direction; returns false when the KKT solve failed
    bool direction(double sigma_mu, const double* dxa, const double* dsa, double* dx, double* ds,
                   double tol, const Norms& nr, Dir& d, MehrotraResult& res, bool& used_fp32) {
        check(lkd_ipm_rhs(&P_, x_, s_, th_, rd_, rp_, dxa, dsa, sigma_mu, rxs_, v_, dy_, stream_), "rhs");
        gpu::KKTSolveInfo info;
        const bool rel = opt_.stop_rule == IpmOptions::StopRule::RelativeGap;
        // reference mode: refinement against M0 as src/cuda/kkt.cu:771-842; relative-gap mode and
        // FP32 factors: PCG preconditioned by the factor (backward-error acceptance in rel mode)
        auto run = [&](bool fp32) {
            bool r;
            if (rel || fp32) r = kkt_->solve_pcg(dy_, tol, fp32 ? opt_.mixed_max_refine : 60, &info, rel);
            else r = kkt_->solve_cached(dy_, tol, &info, 10);
            res.refinement_steps += info.refinement_steps;
            res.solves++;
            return r;
        };
        bool ok;
        if (used_fp32) {
            // keep the right-hand side in case the FP64 factor has to take over
            check(lkd_memcpy_dtod(qp_, dy_, sizeof(double) * m_, stream_), "copy");
            ok = run(true);
            if (!ok) {
                factor(false, res);
                used_fp32 = false;
                if (fp64_fail_) return false;
                check(lkd_memcpy_dtod(dy_, qp_, sizeof(double) * m_, stream_), "copy");
                ok = run(false);
            }
        } else {
            ok = run(false);
        }
        if (opt_.verbose)
            std::printf("    solve: sigma_mu %.3e ok %d its %d res %.3e delta %.1e tol %.1e\n", sigma_mu, ok,
                        info.refinement_steps, info.residual_inf, info.delta, tol);
        if (!ok) return false;
        check(lkd_ipm_direction(&P_, x_, s_, y_, th_, rd_, rp_, rxs_, dy_, dx, ds, qd_, qp_, stream_), "direction");
        // predictor: mu_aff at the affine step lengths (taken from the direction's slots on the
        // device) comes with the same scalar read
        if (!dxa) check(lkd_ipm_mu_aff_step(&P_, x_, s_, dx, ds, stream_), "mu_aff");
        read_scalars(LKS_NSLOTS);
        if (opt_.verbose)
            std::printf("    dir: e1 %.3e e2 %.3e e3 %.3e apmax %.3e admax %.3e\n",
                        std::sqrt(hS_[LKS_R1]) / std::max(1.0, nr.rp), std::sqrt(hS_[LKS_R2]) / std::max(1.0, nr.rd),
                        std::sqrt(hS_[LKS_R3]) / std::max(1.0, std::sqrt(hS_[LKS_RXS2])), hS_[LKS_APMAX], hS_[LKS_ADMAX]);
        // e1 re-checks the KKT residual; in relative-gap mode the solve's own (backward-error)
        // acceptance replaces it
        d.e1 = opt_.stop_rule == IpmOptions::StopRule::RelativeGap
                   ? 0.0 : std::sqrt(hS_[LKS_R1]) / std::max(1.0, nr.rp);
        d.e2 = std::sqrt(hS_[LKS_R2]) / std::max(1.0, nr.rd);
        d.e3 = std::sqrt(hS_[LKS_R3]) / std::max(1.0, std::sqrt(hS_[LKS_RXS2]));
        d.apmax = hS_[LKS_APMAX];
        d.admax = hS_[LKS_ADMAX];
        return true;
    }

    // Relative-gap mode: dual-only trials (x stays; s, y move by b ds, b dy, b halving from
    // 0.995 of the dual step bound, 15 trials) of the current direction; the first that is finite,
    // positive, makes Armijo progress on merit0, keeps the met targets and has a merit below
    // `below` is returned (its index, b in *bd), else -1.
    template <typename Keeps>
    int dual_only_trial(double admax, double merit0, double below, const Keeps& keeps, double& bd) {
        const double ad0 = std::min(1.0, 0.995 * admax);
        const int ntd = 15;
        check(lkd_ipm_trials(&P_, x_, s_, y_, dx_, ds_, dy_, rp_, rd_, qp_, qd_, 0.0, ad0, ntd, stream_), "trials");
        read_scalars(LKS_TRIAL_BASE + LKS_TRIAL_STRIDE * ntd);
        double b = ad0;
        for (int t = 0; t < ntd; ++t, b *= 0.5) {
            const double* T = &hS_[LKS_TRIAL_BASE + t * LKS_TRIAL_STRIDE];
            const double nrp2 = std::sqrt(T[0]), nrd2 = std::sqrt(T[1]), mu2 = T[2] / n_;
            const double merit = nrp2 / rp0_ + nrd2 / rd0_ + mu2 / mu0_;
            const bool finite = std::isfinite(nrp2) && std::isfinite(nrd2) && std::isfinite(mu2) &&
                                std::isfinite(merit) && std::isfinite(merit0) && std::isfinite(T[4]) &&
                                std::isfinite(b) && std::isfinite(T[5]);
            const bool positive = T[4] > 0.0 && mu2 > 0.0 && b > 0.0;
            const bool progress = merit < merit0 && merit <= (1.0 - 1e-4 * b) * merit0 && keeps(T);
            if (finite && positive && progress && merit < below) {
                bd = b;
                return t;
            }
        }
        return -1;
    }

    bool fp64_fail_ = false;

    // factor for the current Theta; fp32 selects the FP32 panel (mixed precision)
    bool factor(bool fp32, MehrotraResult& res) {
        auto t = clk::now();
        const int r = kkt_->factor_equilibrated(th_, fp32 ? gpu::FactorPrecision::FP32
                                                          : gpu::FactorPrecision::FP64);
        res.factor_ms += ms_since(t);
        res.factorizations++;
        if (fp32) res.fp32_factorizations++;
        else res.fp64_factorizations++;
        if (!fp32) fp64_fail_ = r < 0;
        return r >= 0;
    }

    void finish(MehrotraResult& res, const Norms& nr) {
        res.primal_residual = nr.rp;
        res.dual_residual = nr.rd;
        res.duality_gap = nr.mu;
        res.objective_value = nr.cx;
        res.dual_objective = nr.by;
        res.rel_gap = std::abs(nr.cx - nr.by) / std::max({1.0, std::abs(nr.cx), std::abs(nr.by)});
        res.rel_primal_residual = nr.rpinf / (1.0 + binf_);
        res.rel_dual_residual = nr.rdinf / (1.0 + cinf_);
        std::vector<double> x(static_cast<size_t>(n_)), y(static_cast<size_t>(m_)), s(static_cast<size_t>(n_));
        check(lkd_memcpy_dtoh(x.data(), x_, sizeof(double) * n_, stream_), "x");
        check(lkd_memcpy_dtoh(y.data(), y_, sizeof(double) * m_, stream_), "y");
        check(lkd_memcpy_dtoh(s.data(), s_, sizeof(double) * n_, stream_), "s");
        res.x.resize(static_cast<size_t>(n_));
        res.pi.resize(static_cast<size_t>(m_));
        res.s.resize(static_cast<size_t>(n_));
        for (Index j = 0; j < n_; ++j) { res.x[j] = x[j] * C_[j]; res.s[j] = s[j] / C_[j]; }
        for (Index i = 0; i < m_; ++i) res.pi[i] = y[i] * R_[i];
        const auto& st = kkt_->stats();
        res.nnz_L = static_cast<double>(st.nnz_L);
        res.flops_per_factorization = st.flops_exact;
    }

    bool stop(const Norms& nr) const {
        if (opt_.stop_rule == IpmOptions::StopRule::Reference)
            return nr.rp < opt_.tol && nr.rd < opt_.tol && nr.mu < opt_.tol;
        const double gap = std::abs(nr.cx - nr.by) / std::max({1.0, std::abs(nr.cx), std::abs(nr.by)});
        // Absolute targets as well (the FP64 certificate checks |Ax-b| and the reduced costs
        // absolutely), floored where FP64 cannot resolve them.
        return gap <= opt_.gap_tol && nr.rpinf / (1.0 + binf_) <= opt_.gap_tol &&
               nr.rdinf / (1.0 + cinf_) <= opt_.gap_tol && nr.rpinf <= primal_target() &&
               nr.rdinf <= dual_target() &&
               nr.xsmax <= std::max(kRelAbsMargin * opt_.gap_tol,
                                    1e-14 * std::max({1.0, std::abs(nr.cx), std::abs(nr.by)}));
    }
    // relative-gap mode: absolute targets of ||Ax - b||_inf and ||A^T y + s - c||_inf (original
    // space), a fraction kRelAbsMargin of gap_tol (the FP64 certificate checks them absolutely)
    static constexpr double kRelAbsMargin = 0.9;
    // relative-gap mode: a corrector step below this competes with the dual-only trials
    static constexpr double kShortStep = 1e-2;
    double rp0_ = 1, rd0_ = 1, mu0_ = 1;   // merit normalisation of the current solve
    double primal_target() const { return std::max(kRelAbsMargin * opt_.gap_tol, 1e-12 * (1.0 + binf_)); }
    double dual_target() const { return std::max(kRelAbsMargin * opt_.gap_tol, 1e-12 * (1.0 + cinf_)); }

    void init_point() {
        if (opt_.warm_x && opt_.warm_y && opt_.warm_s) {
            std::vector<double> x(static_cast<size_t>(n_)), y(static_cast<size_t>(m_)), s(static_cast<size_t>(n_));
            for (Index j = 0; j < n_; ++j) {
                x[j] = std::max(opt_.warm_x[j] / C_[j], opt_.warm_floor);
                s[j] = std::max(opt_.warm_s[j] * C_[j], opt_.warm_floor);
            }
            for (Index i = 0; i < m_; ++i) y[i] = opt_.warm_y[i] / R_[i];
            check(lkd_memcpy_htod(x_, x.data(), sizeof(double) * n_, stream_), "warm x");
            check(lkd_memcpy_htod(y_, y.data(), sizeof(double) * m_, stream_), "warm y");
            check(lkd_memcpy_htod(s_, s.data(), sizeof(double) * n_, stream_), "warm s");
        } else {
            check(lkd_ipm_init(m_, n_, x_, y_, s_, stream_), "init");
        }
    }

    MehrotraResult solve() {
        auto t0 = clk::now();
        MehrotraResult res;
        res.status = simplex::SimplexStatus::Optimal;
        res.iterations = 0;
        res.setup_ms = setup_ms_;
        init_point();
        double rp0 = 1, rd0 = 1, mu0 = 1;
        const Index max_iter = opt_.max_iter;
        Norms nr{};
        bool converged = false, broke = false;
        for (Index iter = 0; iter < max_iter; ++iter) {
            nr = residuals();
            if (iter == 0) {
                rp0 = std::max(1.0, nr.rp);
                rd0 = std::max(1.0, nr.rd);
                mu0 = std::max(1.0, nr.mu);
                rp0_ = rp0;
                rd0_ = rd0;
                mu0_ = mu0;
            }
            if (opt_.verbose)
                std::printf("  Iter %d | rp: %.3e | rd: %.3e | mu: %.3e | obj: %.12e\n", iter,
                            nr.rp, nr.rd, nr.mu, nr.cx);
            if (stop(nr)) {
                res.status = simplex::SimplexStatus::Optimal;
                res.iterations = iter;
                converged = true;
                break;
            }
            // ---- factorization for this iteration ----
            bool use_fp32 = opt_.precision == IpmOptions::Precision::Mixed;
            if (use_fp32 && theta_spread() > opt_.mixed_fp64_spread) use_fp32 = false;
            bool fok = factor(use_fp32, res);
            if (!fok && use_fp32) { use_fp32 = false; fok = factor(false, res); }
            if (!fok) { res.status = simplex::SimplexStatus::IterationLimit; res.iterations = iter; res.fail_code = 1; broke = true; break; }
            const double tol = 1e-9 * std::max(1.0, nr.rp);

            // ---- predictor ----
            Dir d;
            const bool pok = direction(0.0, nullptr, nullptr, dxa_, dsa_, tol, nr, d, res, use_fp32);
            if (!pok || d.e1 > 1e-8 || d.e2 > 1e-8 || d.e3 > 1e-8) {
                res.status = simplex::SimplexStatus::IterationLimit;
                res.iterations = iter;
                res.fail_code = pok ? 3 : 2;
                if (opt_.verbose) std::printf("    predictor failed: solve %d e %.2e %.2e %.2e\n", pok, d.e1, d.e2, d.e3);
                broke = true;
                break;
            }
            // MUAFF = (x + ap dxa)^T (s + ad dsa) at ap = min(1, apmax), ad = min(1, admax), read
            // with the predictor's direction scalars
            const double mu_aff = hS_[LKS_MUAFF] / n_;
            const double sigma = std::min(1.0, std::pow(std::max(0.0, mu_aff) / std::max(1e-16, nr.mu), 3.0));

            // ---- corrector ----
            if (!direction(sigma * nr.mu, dxa_, dsa_, dx_, ds_, tol, nr, d, res, use_fp32)) {
                res.status = simplex::SimplexStatus::IterationLimit;
                res.iterations = iter;
                res.fail_code = 4;
                broke = true;
                break;
            }
            const bool corrector_valid = !(d.e1 > 1e-8 || d.e2 > 1e-8 || d.e3 > 1e-8);
            const bool rel = opt_.stop_rule == IpmOptions::StopRule::RelativeGap;
            double ap = std::min(1.0, 0.995 * d.apmax), ad = std::min(1.0, 0.995 * d.admax);
            const double merit0 = nr.rp / rp0 + nr.rd / rd0 + nr.mu / mu0;
            bool accepted = false;
            // relative-gap mode: a residual already inside its absolute target must stay inside
            const bool keep_p = rel && nr.rpinf <= primal_target(), keep_d = rel && nr.rdinf <= dual_target();
            auto keeps = [&](const double* T) {
                return !(keep_p && T[7] > primal_target()) && !(keep_d && T[6] > dual_target());
            };
            for (int dir = corrector_valid ? 0 : 1; dir < 2 && !accepted; ++dir) {
                if (dir == 1) {
                    Dir dc;
                    if (!direction(0.5 * nr.mu, nullptr, nullptr, dx_, ds_, tol, nr, dc, res, use_fp32)) break;
                    if (dc.e1 > 1e-8 || dc.e2 > 1e-8 || dc.e3 > 1e-8) break;
                    ap = ad = std::min(1.0, 0.995 * std::min(dc.apmax, dc.admax));
                }
                const int ntr = dir == 0 ? 15 : 30;
                check(lkd_ipm_trials(&P_, x_, s_, y_, dx_, ds_, dy_, rp_, rd_, qp_, qd_, ap, ad, ntr, stream_), "trials");
                read_scalars(LKS_TRIAL_BASE + LKS_TRIAL_STRIDE * ntr);
                double a = ap, b = ad;
                // relative mode: first finite, positive centred trial that keeps the met targets
                double fa = -1.0, fb = -1.0;
                for (int t = 0; t < ntr; ++t) {
                    const double* T = &hS_[LKS_TRIAL_BASE + t * LKS_TRIAL_STRIDE];
                    const double nrp2 = std::sqrt(T[0]), nrd2 = std::sqrt(T[1]), mu2 = T[2] / n_;
                    const double mx = T[3], ms = T[4], ny = std::sqrt(T[5]);
                    const double merit = nrp2 / rp0 + nrd2 / rd0 + mu2 / mu0;
                    const double step = std::min(a, b);
                    const bool finite = std::isfinite(nrp2) && std::isfinite(nrd2) && std::isfinite(mu2) &&
                                        std::isfinite(merit) && std::isfinite(merit0) && std::isfinite(mx) &&
                                        std::isfinite(ms) && std::isfinite(a) && std::isfinite(b) &&
                                        std::isfinite(ny);
                    const bool positive = mx > 0.0 && ms > 0.0 && mu2 > 0.0 && step > 0.0;
                    const bool progress = merit < merit0 && merit <= (1.0 - 1e-4 * step) * merit0 && keeps(T);
                    if (finite && positive && progress) {
                        // relative-gap mode: a short corrector step with both residuals inside
                        // their targets competes with the dual-only trials (better merit wins)
                        double ca = a, cb = b;
                        if (rel && dir == 0 && step < kShortStep && keep_p && keep_d) {
                            const int t2 = dual_only_trial(d.admax, merit0, merit, keeps, cb);
                            if (t2 >= 0) {
                                ca = 0.0;
                                if (opt_.verbose) std::printf("    step: dual-only trial %d ad %.3e beats the corrector\n", t2, cb);
                            } else {
                                cb = b;
                            }
                        }
                        check(lkd_ipm_commit(&P_, x_, s_, y_, dx_, ds_, dy_, ca, cb, stream_), "commit");
                        accepted = true;
                        if (opt_.verbose)
                            std::printf("    step: dir %d trial %d ap %.3e ad %.3e sigma %.3e pcg %d\n", dir, t, ca, cb, sigma, res.refinement_steps);
                        break;
                    }
                    if (rel && dir == 1 && finite && positive && keeps(T) && fa < 0) { fa = a; fb = b; }
                    a *= 0.5;
                    b *= 0.5;
                }
                if (!accepted && rel && dir == 0) {
                    // relative-gap mode: dual-only trials of the same direction -- the primal
                    // point (and its residual) stays; near the optimum the normal equations
                    // limit the primal accuracy of a direction, not its dual accuracy
                    double bd = 0.0;
                    const int t = dual_only_trial(d.admax, merit0, HUGE_VAL, keeps, bd);
                    if (t >= 0) {
                        check(lkd_ipm_commit(&P_, x_, s_, y_, dx_, ds_, dy_, 0.0, bd, stream_), "commit");
                        accepted = true;
                        if (opt_.verbose) std::printf("    step: dual-only trial %d ad %.3e\n", t, bd);
                    }
                }
                if (!accepted && rel && dir == 1 && fa > 0) {
                    // relative-gap mode: non-monotone fallback to the longest admissible centred step
                    check(lkd_ipm_commit(&P_, x_, s_, y_, dx_, ds_, dy_, fa, fb, stream_), "commit");
                    accepted = true;
                    if (opt_.verbose) std::printf("    step: centred fallback a %.3e\n", fa);
                }
            }
            if (!accepted) {
                res.status = simplex::SimplexStatus::IterationLimit;
                res.iterations = iter;
                res.fail_code = 5;
                broke = true;
                break;
            }
        }
        if (!converged) {
            nr = residuals();
            if (!broke) {
                res.status = simplex::SimplexStatus::IterationLimit;
                res.iterations = max_iter;
            }
        }
        finish(res, nr);
        res.solve_ms = ms_since(t0);
        return res;
    }

    void update_data(const std::vector<Float>& c, const std::vector<Float>& b) {
        if (static_cast<Index>(c.size()) != n_ || static_cast<Index>(b.size()) != m_)
            throw std::invalid_argument("update_objective_and_rhs: dimension mismatch");
        b_ = b;
        c_ = c;
        binf_ = cinf_ = 0;
        for (Index i = 0; i < m_; ++i) { binf_ = std::max(binf_, std::abs(b[i])); b_[i] *= R_[i]; }
        for (Index j = 0; j < n_; ++j) { cinf_ = std::max(cinf_, std::abs(c[j])); c_[j] *= C_[j]; }
        check(lkd_memcpy_htod(const_cast<double*>(P_.b), b_.data(), sizeof(double) * m_, stream_), "b");
        check(lkd_memcpy_htod(const_cast<double*>(P_.c), c_.data(), sizeof(double) * n_, stream_), "c");
        check(lkd_stream_sync(stream_), "sync");
    }

    std::vector<Float> wx_, wy_, ws_;
    void set_warm(const std::vector<Float>& x, const std::vector<Float>& y, const std::vector<Float>& s) {
        if (static_cast<Index>(x.size()) != n_ || static_cast<Index>(y.size()) != m_ ||
            static_cast<Index>(s.size()) != n_)
            throw std::invalid_argument("set_warm_start: dimension mismatch");
        wx_ = x; wy_ = y; ws_ = s;
        opt_.warm_x = wx_.data(); opt_.warm_y = wy_.data(); opt_.warm_s = ws_.data();
    }
    void clear_warm() { opt_.warm_x = opt_.warm_y = opt_.warm_s = nullptr; }

    // max(Theta) / min(Theta) of the current iterate: both extremes come with the residual
    // scalars read at the top of the iteration (the pass that forms Theta reduces them)
    double theta_spread() const {
        const double hi = hS_[LKS_THMAX], lo = hS_[LKS_THMIN];
        return lo > 0 ? hi / lo : 1e300;
    }
};

MehrotraSolver::MehrotraSolver(const core::Model& model) : impl_(new Impl(model, IpmOptions{})) {}
MehrotraSolver::MehrotraSolver(const core::Model& model, const IpmOptions& options)
    : impl_(new Impl(model, options)) {}
MehrotraSolver::~MehrotraSolver() { delete impl_; }
MehrotraResult MehrotraSolver::solve() { return impl_->solve(); }
void MehrotraSolver::update_objective_and_rhs(const std::vector<Float>& c,
                                              const std::vector<Float>& b) {
    impl_->update_data(c, b);
}
void MehrotraSolver::set_warm_start(const std::vector<Float>& x, const std::vector<Float>& y,
                                    const std::vector<Float>& s) {
    impl_->set_warm(x, y, s);
}
void MehrotraSolver::clear_warm_start() { impl_->clear_warm(); }
void MehrotraSolver::profile_factor(bool fp32, double ms[6], double flops[6], double launches[6]) {
    impl_->kkt_->profile_factor(fp32 ? gpu::FactorPrecision::FP32 : gpu::FactorPrecision::FP64,
                                ms, flops, launches);
    check(lkd_stream_sync(impl_->stream_), "sync");
}
void MehrotraSolver::set_launch_timing(bool on) { impl_->kkt_->set_launch_timing(on); }
void MehrotraSolver::launch_timing(double ms[6], double flops[6], double launches[6]) {
    impl_->kkt_->launch_timing(ms, flops, launches);
}
void MehrotraSolver::kkt_statistics(double out[12]) {
    const auto& st = impl_->kkt_->stats();
    const double v[12] = {double(st.m), double(st.n), double(st.nnz_A), double(st.nsuper),
                          double(st.nnz_L), double(st.nnz_L_exact), st.flops_exact, st.flops_panel,
                          double(st.levels), double(st.launches_factor), double(st.launches_solve),
                          st.symbolic_ms};
    for (int i = 0; i < 12; ++i) out[i] = v[i];
}

}  // namespace ipm
}  // namespace sankhya

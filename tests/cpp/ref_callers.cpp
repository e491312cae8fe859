// A reference-side caller compiled against this repository's include/ and linked to
// liblakshya.so: it drives the C++ classes exactly as the reference's own callers do.
//   kkt   -- GPUKKTCholeskySolver through the four-argument constructor with an explicit
//            M_pattern, host factorize, get_device_L_values() read back in the reference CSR
//            layout, host solve (the steps of tests/unit/test_gpu_kkt.cu:21-121);
//   ipm   -- MehrotraSolver(model).solve() on the 2 x 4 textbook LP built with add_variable /
//            add_constraint / finalize (tests/unit/test_mehrotra.cu:29-78), then the
//            reference-signature verify_optimal(model, x, pi, &res);
//   crossover -- generate_crossover_basis on a DeviceModel from upload_to_device, the two cases
//            of tests/unit/test_crossover.cu:15-133;
//   simplex -- simplex::solve_with_phase1(model, basis, x) (the device route of
//            src/simplex/phase1.cpp:124) on the models of tests/unit/test_phase1.cpp:123-430
//            built as that file's make_model does (direct CSC), then verify_optimal with
//            pi = B^-T c_B as the reference CLI forms it (sankhya_benchmark_cli.cpp:62-71);
//   model -- the Model layout: bilinear_terms is a member as in src/core/problem.hpp:49.
// Exit status 0 when every check holds; each failed check prints one line.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "core/problem.hpp"
#include "gpu/device_model.hpp"
#include "gpu/kkt.hpp"
#include "gpu/vram_arena.hpp"
#include "ipm/crossover.hpp"
#include "ipm/mehrotra.hpp"
#include "ipm/symbolic.hpp"
#include "simplex/primal.hpp"
#include "verifier/certificate.hpp"

using namespace sankhya;

static int g_fail = 0;
#define EXPECT(cond)                                                         \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);      \
            ++g_fail;                                                        \
        }                                                                    \
    } while (0)

static int run_kkt() {
    // A (3 x 5) whose normal matrix with Theta = 1 is [[6,1,1],[1,5,0],[1,0,5]]
    core::CSRMatrix A;
    A.rows = 3;
    A.cols = 5;
    A.row_ptrs = {0, 3, 5, 7};
    A.col_indices = {0, 3, 4, 1, 3, 2, 4};
    A.values = {2.0, 1.0, 1.0, 2.0, 1.0, 2.0, 1.0};
    const std::vector<Float> theta(5, 1.0);

    const std::vector<Index> P = ipm::compute_amd_ordering(A);
    EXPECT((P == std::vector<Index>{1, 0, 2}));
    ipm::SymbolicFactorization sym;
    ipm::compute_symbolic_factorization(A, P, sym);

    // the full permuted pattern of M, as the reference's callers pass it
    core::CSRMatrix Mp;
    Mp.rows = Mp.cols = 3;
    Mp.row_ptrs = {0, 2, 5, 7};
    Mp.col_indices = {0, 1, 0, 1, 2, 1, 2};
    Mp.values.assign(7, 0.0);

    gpu::VRAMArena arena(1 << 20);
    gpu::GPUKKTCholeskySolver kkt(arena, A, sym, Mp);
    kkt.gpu_cholesky_factorize(A, theta);

    const size_t nnzL = sym.L_pattern.values.size() + 3;   // strictly lower + one diagonal per row
    std::vector<Float> Lv(nnzL, 0.0);
    EXPECT(hipMemcpy(Lv.data(), kkt.get_device_L_values(), nnzL * sizeof(Float),
                     hipMemcpyDeviceToHost) == hipSuccess);
    // rows of L in CSR, the diagonal last in each row
    const int rp[4] = {0, 1, 3, 5}, ci[5] = {0, 0, 1, 1, 2};
    double L[3][3] = {};
    for (int i = 0; i < 3; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) L[i][ci[p]] = Lv[p];
    const double Mperm[3][3] = {{5, 1, 0}, {1, 6, 1}, {0, 1, 5}};
    double fro = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += L[i][k] * L[j][k];
            fro += (s - Mperm[i][j]) * (s - Mperm[i][j]);
        }
    EXPECT(std::sqrt(fro) < 1e-6);

    std::vector<Float> rhs = {11.0, 11.0, 16.0};
    kkt.gpu_cholesky_solve(rhs);
    EXPECT(std::fabs(rhs[0] - 1.0) < 1e-6 && std::fabs(rhs[1] - 2.0) < 1e-6 &&
           std::fabs(rhs[2] - 3.0) < 1e-6);

    // a Theta of the wrong length is std::invalid_argument
    bool threw = false;
    try {
        kkt.gpu_cholesky_factorize(A, std::vector<Float>(4, 1.0));
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    EXPECT(threw);
    std::printf("kkt: factor residual %.3e, x = (%.9f, %.9f, %.9f)\n", std::sqrt(fro), rhs[0],
                rhs[1], rhs[2]);
    return 0;
}

static int run_ipm() {
    // min -3 x1 - 5 x2  s.t.  x1 + x2 + x3 = 4,  2 x1 + 3 x2 + x4 = 12,  x >= 0  (optimum -20)
    core::Model model;
    model.sense = OptimizationSense::Minimize;
    for (double c : {-3.0, -5.0, 0.0, 0.0}) model.add_variable(c);
    model.add_constraint({0, 1, 2}, {1.0, 1.0, 1.0}, 4.0);
    model.add_constraint({0, 1, 3}, {2.0, 3.0, 1.0}, 12.0);
    model.finalize();

    ipm::MehrotraSolver solver(model);
    const ipm::MehrotraResult r = solver.solve();
    EXPECT(r.status == simplex::SimplexStatus::Optimal);
    EXPECT(r.iterations > 0);
    EXPECT(r.primal_residual < 1e-6 && r.dual_residual < 1e-6 && r.duality_gap < 1e-6);
    EXPECT(std::fabs(r.objective_value + 20.0) <= 1e-4);
    EXPECT(r.x.size() == 4 && r.pi.size() == 2);
    for (Float v : r.x) EXPECT(std::isfinite(v));
    for (Float v : r.pi) EXPECT(std::isfinite(v));
    Float res = -1;
    EXPECT(verifier::verify_optimal(model, r.x, r.pi, &res));
    EXPECT(res >= 0 && res < 1e-6);
    std::printf("ipm: status %d, %d iterations, objective %.9f, |Ax-b| %.2e\n",
                static_cast<int>(r.status), r.iterations, r.objective_value, res);
    return 0;
}

static int run_crossover() {
    // tests/unit/test_crossover.cu:15-114: min -x1 - x2  s.t.  x1 + 2 x2 + x3 = 4,
    // 2 x1 + x2 + x4 = 4 from the interior point (1, 1, 1, 1), through a DeviceModel uploaded
    // into an arena; the basis must be m valid, nonsingular columns whose basic solution is
    // feasible with every nonbasic column AtLower
    core::Model model;
    model.sense = OptimizationSense::Minimize;
    for (double c : {-1.0, -1.0, 0.0, 0.0}) model.add_variable(c);
    model.add_constraint({0, 1, 2}, {1.0, 2.0, 1.0}, 4.0);
    model.add_constraint({0, 1, 3}, {2.0, 1.0, 1.0}, 4.0);
    model.finalize();
    gpu::VRAMArena arena;
    gpu::DeviceModel d_model = gpu::upload_to_device(model, arena);
    const std::vector<Float> x_ipm = {1.0, 1.0, 1.0, 1.0};
    simplex::Basis basis = ipm::generate_crossover_basis(d_model, x_ipm);
    EXPECT(basis.basic_indices.size() == 2 && basis.col_status.size() == 4);
    EXPECT(basis.basic_indices[0] < 4 && basis.basic_indices[1] < 4);
    int num_basic = 0;
    for (int j = 0; j < 4; ++j) num_basic += basis.col_status[j] == simplex::BasisStatus::Basic;
    EXPECT(num_basic == 2);
    auto col = [&](Index j) {
        std::vector<Float> v(2, 0.0);
        for (Index p = model.A.col_ptrs[j]; p < model.A.col_ptrs[j + 1]; ++p)
            v[model.A.row_indices[p]] = model.A.values[p];
        return v;
    };
    const Index b0 = basis.basic_indices[0], b1 = basis.basic_indices[1];
    const std::vector<Float> A0 = col(b0), A1 = col(b1);
    const Float det = A0[0] * A1[1] - A0[1] * A1[0];
    EXPECT(std::fabs(det) > 1e-7);
    std::vector<Float> x(4, 0.0);
    x[b0] = (4.0 * A1[1] - 4.0 * A1[0]) / det;
    x[b1] = (A0[0] * 4.0 - A0[1] * 4.0) / det;
    EXPECT(std::fabs(x[0] + 2 * x[1] + x[2] - 4.0) <= 1e-6);
    EXPECT(std::fabs(2 * x[0] + x[1] + x[3] - 4.0) <= 1e-6);
    for (int j = 0; j < 4; ++j) {
        EXPECT(x[j] >= -1e-6);
        if (j != b0 && j != b1) EXPECT(basis.col_status[j] == simplex::BasisStatus::AtLower);
    }
    // the same call on the host model
    simplex::Basis hb = ipm::generate_crossover_basis(model, x_ipm);
    EXPECT(hb.basic_indices == basis.basic_indices && hb.col_status == basis.col_status);
    std::printf("crossover: basis (%d, %d), x = (%.6f, %.6f, %.6f, %.6f), objective %.6f\n", b0, b1,
                x[0], x[1], x[2], x[3], -x[0] - x[1]);

    // tests/unit/test_crossover.cu:117-133: dependent rows throw the rank-deficiency message
    core::Model dep;
    dep.sense = OptimizationSense::Minimize;
    dep.add_variable(-1.0);
    dep.add_variable(0.0);
    dep.add_constraint({0, 1}, {1.0, 1.0}, 2.0);
    dep.add_constraint({0, 1}, {2.0, 2.0}, 4.0);
    dep.finalize();
    gpu::DeviceModel d_dep = gpu::upload_to_device(dep, arena);
    bool threw = false;
    try {
        ipm::generate_crossover_basis(d_dep, std::vector<Float>{1.0, 1.0});
    } catch (const std::runtime_error& e) {
        threw = std::strstr(e.what(), "Crossover failed: Matrix is rank deficient") != nullptr;
    }
    EXPECT(threw);
    return 0;
}

// tests/unit/test_phase1.cpp:44-70: a Model from CSC arrays, lb = 0, ub = +inf unless given
static core::Model phase1_model(Index m, Index n, std::vector<Index> cp, std::vector<Index> ri,
                                std::vector<Float> v, std::vector<Float> b, std::vector<Float> c,
                                OptimizationSense sense = OptimizationSense::Minimize,
                                std::vector<Float> lb = {}, std::vector<Float> ub = {}) {
    core::Model mo;
    mo.sense = sense;
    mo.obj = c;
    mo.rhs = b;
    mo.lb = lb.empty() ? std::vector<Float>(n, 0.0) : lb;
    mo.ub = ub.empty() ? std::vector<Float>(n, math::kInfinity) : ub;
    mo.vtype.assign(n, VariableType::Continuous);
    mo.A.rows = m;
    mo.A.cols = n;
    mo.A.col_ptrs = cp;
    mo.A.row_indices = ri;
    mo.A.values = v;
    return mo;
}

static int run_simplex() {
    struct Case { const char* name; core::Model model; simplex::SimplexStatus status; double obj; };
    using S = simplex::SimplexStatus;
    std::vector<Case> cases;
    cases.push_back({"test1", phase1_model(2, 4, {0, 2, 3, 4, 5}, {0, 1, 0, 0, 1}, {1, 1, 1, 1, 1}, {4, 2}, {-1, -1, 0, 0}), S::Optimal, -4.0});
    cases.push_back({"test2", phase1_model(1, 2, {0, 1, 2}, {0, 0}, {1, 1}, {3}, {1, 1}), S::Optimal, 3.0});
    cases.push_back({"test3", phase1_model(1, 1, {0, 1}, {0}, {-1}, {-5}, {1}), S::Optimal, 5.0});
    cases.push_back({"test4", phase1_model(2, 2, {0, 2, 4}, {0, 1, 0, 1}, {1, 1, 1, -1}, {1, 0}, {1, 2}), S::Optimal, 1.5});
    cases.push_back({"test5", phase1_model(2, 2, {0, 2, 4}, {0, 1, 0, 1}, {1, 1, 1, 1}, {1, 2}, {0, 0}), S::Infeasible, 0.0});
    cases.push_back({"test6", phase1_model(2, 2, {0, 2, 4}, {0, 1, 0, 1}, {1, 2, 1, 2}, {3, 6}, {1, 1}), S::IterationLimit, 0.0});
    cases.push_back({"test9", phase1_model(2, 4, {0, 2, 3, 4, 5}, {0, 1, 0, 0, 1}, {1, 1, 1, 1, 1}, {4, 2}, {-3, -2, 0, 0}), S::Optimal, -10.0});
    // a Maximize model (the route minimizes -c) and a shifted lower bound
    cases.push_back({"max", phase1_model(2, 4, {0, 2, 3, 4, 5}, {0, 1, 0, 0, 1}, {1, 1, 1, 1, 1}, {4, 2}, {3, 2, 0, 0},
                                         OptimizationSense::Maximize), S::Optimal, 10.0});
    cases.push_back({"lb", phase1_model(1, 2, {0, 1, 2}, {0, 0}, {1, 1}, {5}, {1, 2}, OptimizationSense::Minimize,
                                        {0.0, 1.0}), S::Optimal, 6.0});
    // a finite upper bound: the reference's unsupported case (phase1.cpp:151-161)
    cases.push_back({"ub", phase1_model(1, 2, {0, 1, 2}, {0, 0}, {1, 1}, {5}, {1, 2}, OptimizationSense::Minimize,
                                        {}, {3.0, math::kInfinity}), S::IterationLimit, 0.0});
    for (auto& cs : cases) {
        simplex::Basis basis;
        std::vector<Float> x, pi;
        const S st = simplex::solve_with_phase1(cs.model, basis, x, &pi);
        EXPECT(st == cs.status);
        if (st != cs.status) std::printf("  case %s: status %d\n", cs.name, static_cast<int>(st));
        if (st != S::Optimal) continue;
        double obj = 0.0;
        for (size_t j = 0; j < x.size(); ++j) obj += cs.model.obj[j] * x[j];
        EXPECT(std::fabs(obj - cs.obj) <= 1e-9);
        for (Index bi : basis.basic_indices) EXPECT(bi >= 0 && bi < cs.model.A.cols);
        Float res = -1.0;
        EXPECT(verifier::verify_optimal(cs.model, x, pi, &res));
        std::printf("simplex %s: objective %.6f, residual %.2e\n", cs.name, obj, res);
    }
    return 0;
}

static int run_model() {
    core::Model model;
    model.add_variable(1.0);
    model.add_variable(1.0);
    model.add_variable(0.0);
    core::BilinearTerm t{0, 1, 2, {0, 1, 2, 3}};
    model.bilinear_terms.push_back(t);
    EXPECT(model.bilinear_terms.size() == 1 && model.bilinear_terms[0].w_col == 2 &&
           model.bilinear_terms[0].mccormick_rows[3] == 3);
    return 0;
}

int main(int argc, char** argv) {
    const char* what = argc > 1 ? argv[1] : "all";
    try {
        if (!std::strcmp(what, "model") || !std::strcmp(what, "all")) run_model();
        if (!std::strcmp(what, "kkt") || !std::strcmp(what, "all")) run_kkt();
        if (!std::strcmp(what, "ipm") || !std::strcmp(what, "all")) run_ipm();
        if (!std::strcmp(what, "crossover") || !std::strcmp(what, "all")) run_crossover();
        if (!std::strcmp(what, "simplex") || !std::strcmp(what, "all")) run_simplex();
    } catch (const std::exception& e) {
        std::printf("FAIL: exception %s\n", e.what());
        return 2;
    }
    std::printf("%s\n", g_fail ? "FAILED" : "OK");
    return g_fail ? 1 : 0;
}

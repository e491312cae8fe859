// Command-line solve of an MPS file with the JSON result contract of the reference's
// tests/benchmark/sankhya_benchmark_cli.cpp:15-100 (the object app.py and harness.py read: the
// last {...} on stdout with the keys solver, status, execution_path, fallback_triggered,
// solve_time_ms, iterations, objective, primal_residual, variables, constraints, nnz,
// certificate_pass).
//
//   lakshya_benchmark_cli <file.mps>             the device IPM in the reference stop rule
//   lakshya_benchmark_cli --relative <file.mps>  the 1e-8 relative-gap rule from the start
//
// Flow: read_mps (a parse error prints {"error": "..."} and exits 0, as the reference does);
// MehrotraSolver(model).solve(); an Optimal result is certified with verify_optimal.  On
// IterationLimit the reference falls back to its CPU Phase-I / Phase-II simplex
// (sankhya_benchmark_cli.cpp:48-75); here the same route runs on the device
// (simplex::solve_with_phase1, execution_path "GPU+GPU_SIMPLEX") for models of up to 1,600 rows,
// and its optimum is certified with pi = B^-T c_B as the reference does; a non-Optimal route
// ends as "FailedRecovery".  Larger models re-solve on the device in the relative-gap mode
// instead (its stricter end-game rules carry stalled solves to a certificate;
// "GPU+GPU_RECOVERY").  solve_time_ms covers the solver set-up, the solves and the
// certificate; iterations are those of the first solve, as in the reference.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "core/problem.hpp"
#include "ipm/mehrotra.hpp"
#include "parsers/mps_reader.hpp"
#include "simplex/primal.hpp"
#include "verifier/certificate.hpp"

using namespace sankhya;

namespace {

std::string escaped(const std::string& s) {
    std::string o;
    for (char ch : s) {
        if (ch == '"' || ch == '\\') o += '\\';
        if (ch == '\n') { o += "\\n"; continue; }
        o += ch;
    }
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    bool relative = false;
    std::string path;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "--relative") == 0) relative = true;
        else if (path.empty()) path = argv[i];
        else path.clear(), i = argc;   // more than one file: usage
    }
    if (path.empty()) {
        std::fprintf(stderr, "Usage: lakshya_benchmark_cli [--relative] <mps_file>\n");
        return 1;
    }

    core::Model model;
    try {
        model = parsers::read_mps(path);
    } catch (const std::exception& e) {
        std::printf("{\"error\": \"%s\"}\n", escaped(e.what()).c_str());
        return 0;
    }

    const auto t0 = std::chrono::steady_clock::now();
    std::string status = "Failed";
    double objective = 0.0, primal_residual = 0.0;
    bool cert = false, fallback = false;
    const char* path_name = "GPU_ONLY";
    long iterations = 0;
    try {
        ipm::IpmOptions opt;
        if (relative) opt.stop_rule = ipm::IpmOptions::StopRule::RelativeGap;
        ipm::MehrotraResult r;
        {
            ipm::MehrotraSolver solver(model, opt);
            r = solver.solve();
        }
        iterations = r.iterations;
        if (r.status == simplex::SimplexStatus::Optimal) {
            status = "Optimal";
            objective = r.objective_value;
            primal_residual = r.primal_residual;
            cert = verifier::verify_optimal(model, r.x, r.pi);
        } else if (r.status == simplex::SimplexStatus::IterationLimit) {
            fallback = true;
            status = "FailedRecovery";
            if (model.A.rows <= simplex::kDeviceSimplexMaxRows) {
                path_name = "GPU+GPU_SIMPLEX";
                simplex::Basis basis;
                std::vector<Float> x, pi;
                if (simplex::solve_with_phase1(model, basis, x, &pi) == simplex::SimplexStatus::Optimal) {
                    status = "Optimal";
                    objective = 0.0;
                    for (std::size_t j = 0; j < x.size(); ++j) objective += x[j] * model.obj[j];
                    Float res = 0.0;
                    cert = verifier::verify_optimal(model, x, pi, &res);
                    primal_residual = res;
                }
            } else if (!relative) {
                path_name = "GPU+GPU_RECOVERY";
                ipm::IpmOptions rec;
                rec.stop_rule = ipm::IpmOptions::StopRule::RelativeGap;
                ipm::MehrotraSolver solver(model, rec);
                const ipm::MehrotraResult q = solver.solve();
                if (q.status == simplex::SimplexStatus::Optimal) {
                    status = "Optimal";
                    objective = q.objective_value;
                    Float res = 0.0;
                    cert = verifier::verify_optimal(model, q.x, q.pi, &res);
                    primal_residual = res;
                }
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "solver error: %s\n", e.what());
        status = "Failed";
    }
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    std::printf("{\n");
    std::printf("  \"solver\": \"SANKHYA\",\n");
    std::printf("  \"status\": \"%s\",\n", status.c_str());
    std::printf("  \"execution_path\": \"%s\",\n", path_name);
    std::printf("  \"fallback_triggered\": \"%s\",\n", fallback ? "YES" : "NO");
    std::printf("  \"solve_time_ms\": %.17g,\n", ms);
    std::printf("  \"iterations\": %ld,\n", iterations);
    std::printf("  \"objective\": %.17g,\n", objective);
    std::printf("  \"primal_residual\": %.17g,\n", primal_residual);
    std::printf("  \"variables\": %zu,\n", model.obj.size());
    std::printf("  \"constraints\": %zu,\n", model.rhs.size());
    std::printf("  \"nnz\": %zu,\n", model.A.values.size());
    std::printf("  \"certificate_pass\": %s\n", cert ? "true" : "false");
    std::printf("}\n");
    return 0;
}

// Independent FP64 optimality certificate check.
// The four-argument form is the reference's function (src/verifier/certificate.hpp:26-31,
// same signature and mangled name, epsilon 1e-6 as src/verifier/certificate.cpp:37).  The
// five-argument overload takes the tolerance explicitly (1e-8 for this build's certificates).
#pragma once
#include <vector>
#include "core/problem.hpp"
#include "sankhya/types.hpp"

namespace sankhya {
namespace verifier {

bool verify_optimal(const core::Model& model, const std::vector<Float>& x,
                    const std::vector<Float>& pi, Float* out_primal_residual = nullptr);
bool verify_optimal(const core::Model& model, const std::vector<Float>& x,
                    const std::vector<Float>& pi, Float* out_primal_residual, Float tol);

}  // namespace verifier
}  // namespace sankhya

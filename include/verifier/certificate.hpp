// Independent FP64 optimality certificate check.
// Interface of src/verifier/certificate.hpp:26-31 (reference) plus a trailing tolerance
// argument whose default (1e-6) reproduces the reference's hard-coded epsilon
// (src/verifier/certificate.cpp:37).
#pragma once
#include <vector>
#include "core/problem.hpp"
#include "sankhya/types.hpp"

namespace sankhya {
namespace verifier {

bool verify_optimal(const core::Model& model, const std::vector<Float>& x,
                    const std::vector<Float>& pi, Float* out_primal_residual = nullptr,
                    Float tol = 1e-6);

}  // namespace verifier
}  // namespace sankhya

// Scalar/index types and global tolerances of the solver interface.
// Mirrors include/sankhya/types.hpp:8-18 of the reference (same names, same values).
#pragma once
#include <cstdint>
#include <limits>

namespace sankhya {

using Float = double;     // all host-side and FP64 device arithmetic
using Index = int32_t;    // row/column/nonzero indices

enum class VariableType { Continuous, Integer, Binary };
enum class OptimizationSense { Minimize, Maximize };

namespace math {
constexpr Float kInfinity = std::numeric_limits<Float>::infinity();
constexpr Float kDefaultFeasibilityTol = 1e-6;   // IPM stop rule in reference mode
constexpr Float kDefaultPivotTol = 1e-7;
}  // namespace math

}  // namespace sankhya

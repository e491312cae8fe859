// Termination status shared by the interior-point solver and the simplex route.
// Same enumerators as src/simplex/primal.hpp:26-31 (reference). The simplex routines themselves
// are not part of this library (the interior-point path only uses the status type).
#pragma once
namespace sankhya {
namespace simplex {
enum class SimplexStatus { Optimal, Infeasible, Unbounded, IterationLimit };
}  // namespace simplex
}  // namespace sankhya

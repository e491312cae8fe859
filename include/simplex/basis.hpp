// Basis of an LP in the equality form  A x = b, lb <= x <= ub: the same names and members as
// src/simplex/basis.hpp:23-44 (reference), so code written against that header compiles here.
#pragma once

#include <vector>

#include "sankhya/types.hpp"

namespace sankhya {
namespace simplex {

enum class BasisStatus { Basic, AtLower, AtUpper, Superbasic };

struct Basis {
    std::vector<BasisStatus> col_status;   // one per column
    std::vector<Index> basic_indices;      // one per row; >= cols marks a row left artificial
};

}  // namespace simplex
}  // namespace sankhya

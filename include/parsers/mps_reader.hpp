// MPS input (src/parsers/mps_reader.hpp:8-20, reference): free-format NAME / ROWS / COLUMNS /
// RHS / BOUNDS sections into a finalized core::Model in equality form.
//   * L rows get a slack (+1, bounds [0, inf)), G rows a surplus (-1, [0, inf)); the slack
//     columns follow the structural columns, in row order;
//   * the first N row is the objective (its entries are the cost vector; a later entry for the
//     same column replaces an earlier one); entries of further N rows are dropped;
//   * only the first RHS vector and the first BOUNDS vector are read;
//   * bounds LO UP FX FR MI PL BV LI UI; 'MARKER' 'INTORG' / 'INTEND' make columns Integer;
//   * RANGES, SOS and any other section, unknown row types, unknown bound types, references to
//     undeclared rows or columns, missing values and a nonzero objective constant throw
//     std::runtime_error ("MPS parser error: ...");
//   * the model's sense is Minimize.
#pragma once
#include <string>
#include "core/problem.hpp"

namespace sankhya {
namespace parsers {

core::Model read_mps(const std::string& filename);

}  // namespace parsers
}  // namespace sankhya

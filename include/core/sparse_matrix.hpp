// Compressed sparse column / row matrices and the CSC -> CSR conversion.
// Interface of src/core/sparse_matrix.hpp:28-75 (reference): same members and semantics.
#pragma once
#include <cstddef>
#include <vector>
#include "sankhya/types.hpp"

namespace sankhya {
namespace core {

// Column c owns entries [col_ptrs[c], col_ptrs[c+1]); row indices strictly increase inside a column.
struct CSCMatrix {
    Index rows;
    Index cols;
    std::vector<Index> col_ptrs;
    std::vector<Index> row_indices;
    std::vector<Float> values;

    CSCMatrix(Index rows, Index cols, std::vector<Float> vals,
              std::vector<Index> rows_idx, std::vector<Index> col_ptrs);
    Index nnz() const;
    Float get(Index r, Index c) const;   // 0 when absent or out of range
};

// Row i owns entries [row_ptrs[i], row_ptrs[i+1]); column indices strictly increase inside a row.
struct CSRMatrix {
    Index rows = 0;
    Index cols = 0;
    std::vector<Index> row_ptrs;
    std::vector<Index> col_indices;
    std::vector<Float> values;
};

// O(nnz) counting-sort transpose of the storage order; values are copied bit-exactly.
CSRMatrix to_csr(const CSCMatrix& csc);

}  // namespace core
}  // namespace sankhya

// Host symbolic phase for the normal equations M = A*Theta*A^T.
//
// Reference interface (kept): SymbolicFactorization, compute_amd_ordering,
// compute_symbolic_factorization -- src/ipm/symbolic.hpp:20-63.
//
// Extensions used by the GPU factorization: a quotient-graph approximate-minimum-degree
// ordering (compute_fill_reducing_ordering) and the supernodal structure
// (compute_supernodal_structure) that the device factorization is planned from.
#pragma once
#include <cstdint>
#include <vector>
#include "core/sparse_matrix.hpp"

namespace sankhya {
namespace ipm {

// Permutation P (P[i] = original row placed at position i), its inverse, the elimination tree
// of P*A*A^T*P^T and the strictly-lower pattern of its Cholesky factor in CSR (ascending columns).
struct SymbolicFactorization {
    std::vector<Index> P;
    std::vector<Index> P_inv;
    std::vector<Index> parent;       // -1 for roots
    core::CSRMatrix L_pattern;       // strictly lower, values are dummies (1.0)
};

// Reference ordering rule (src/ipm/symbolic.cpp:11-98): explicit-graph minimum degree on the
// A*A^T graph, lowest index wins ties, the eliminated vertex's neighbours become a clique.
// Exact but O(m^2)+; intended for small systems and for reproducing the reference's orderings.
std::vector<Index> compute_amd_ordering(const core::CSRMatrix& A);

// Elimination tree + L pattern for a given ordering (src/ipm/symbolic.cpp:100-218 semantics;
// computed here in O(nnz(L)) with Liu's algorithm and row-subtree traversal).
void compute_symbolic_factorization(const core::CSRMatrix& A, const std::vector<Index>& P,
                                    SymbolicFactorization& out_factorization);

// Approximate minimum degree (quotient graph, element absorption, supervariables, dense-row
// postponement) on the A*A^T graph.  Scales to m ~ 1e5 with dense linking rows.
std::vector<Index> compute_fill_reducing_ordering(const core::CSRMatrix& A);

// Supernodal structure of the permuted normal matrix, the static input of the device
// factorization plan.  All row indices are permuted positions.
struct SupernodalStructure {
    Index m = 0;
    Index nsuper = 0;
    std::vector<Index> perm;          // [m] ordering actually used (input ordering, postordered)
    std::vector<Index> sn_first;      // [nsuper+1] first column of each supernode (contiguous)
    std::vector<Index> sn_parent;     // parent supernode in the supernodal etree (-1 root)
    std::vector<Index> sn_level;      // height above the leaves (leaves = 0)
    std::vector<int64_t> sn_rowptr;   // [nsuper+1] offsets into sn_rows
    std::vector<Index> sn_rows;       // row structure; first ncols entries are the diagonal block
    std::vector<int64_t> sn_panel;    // [nsuper+1] offsets (in elements) of column-major panels
    std::vector<Index> col_to_sn;     // [m] supernode owning each column
    std::vector<uint8_t> empty_row;   // [m] permuted rows of A with no entries
    int64_t nnz_L = 0;                // structural nonzeros of L incl. diagonal (panel lower part)
    int64_t nnz_L_exact = 0;          // nonzeros of the exact (unrelaxed) factor incl. diagonal
    double flops_exact = 0;           // sum_j c_j^2 over the exact column counts
    double flops_panel = 0;           // same sum over the relaxed (stored) column counts
    Index max_level = 0;
};

struct SupernodeOptions {
    // A child is merged into its parent when the merged supernode has at most `relax_small`
    // columns, or when the explicit zeros added stay below `relax_frac` of the merged panel.
    Index relax_small = 16;
    double relax_frac = 0.30;
    Index max_cols = 1 << 30;
    // A child of more than `serial_child` columns is not merged into a parent when a sibling
    // subtree carries at least `serial_frac` of the child's subtree work: the merge would move
    // the child's whole factorization behind that sibling's (one etree level up), where it runs
    // alone instead of beside it.
    Index serial_child = 32;
    double serial_frac = 0.5;
};

void compute_supernodal_structure(const core::CSRMatrix& A, const std::vector<Index>& P,
                                  const SupernodeOptions& opt, SupernodalStructure& out);

}  // namespace ipm
}  // namespace sankhya

// Host symbolic phase of the normal equations M = A*Theta*A^T (permuted: P*M*P^T).
//
//  * compute_amd_ordering: the reference ordering rule (src/ipm/symbolic.cpp:11-98) --
//    explicit-graph minimum degree, lowest index wins ties, neighbours of the eliminated
//    vertex become a clique.  Restated with the same selection rule, so the same permutation.
//  * compute_symbolic_factorization: etree + strictly-lower L pattern for a permutation
//    (src/ipm/symbolic.cpp:100-218 semantics); computed with Liu's etree algorithm and the
//    row-subtree characterisation of L's rows instead of the reference's O(m^2) child scan.
//  * compute_supernodal_structure: postordered etree, exact column counts, fundamental
//    supernodes, relaxed amalgamation and supernode row structures (device factorization input).
#include <algorithm>
#include <numeric>
#include <stdexcept>
#include "ipm/symbolic.hpp"
#include "parallel.hpp"

namespace sankhya {
namespace ipm {

namespace {

// Rows of A sharing column k, listed per column (the A^T view the graph of A*A^T is built from).
std::vector<std::vector<Index>> rows_by_column(const core::CSRMatrix& A) {
    std::vector<std::vector<Index>> by_col(static_cast<size_t>(A.cols));
    for (Index i = 0; i < A.rows; ++i)
        for (Index p = A.row_ptrs[i]; p < A.row_ptrs[i + 1]; ++p)
            by_col[static_cast<size_t>(A.col_indices[p])].push_back(i);
    return by_col;
}

// Strictly-lower adjacency of the permuted graph: row i lists j < i (ascending, unique).
void permuted_lower_graph(const core::CSRMatrix& A, const std::vector<Index>& pinv,
                          std::vector<Index>& ptr, std::vector<Index>& idx) {
    const Index m = A.rows;
    // column lists (original rows), then per permuted row r (original row i = P[r]) the permuted
    // rows below r that share a column, deduplicated with a per-thread marker and sorted; rows
    // in parallel, concatenated in order
    std::vector<Index> cstart(static_cast<size_t>(A.cols) + 1, 0), crow, P(static_cast<size_t>(m));
    for (Index i = 0; i < m; ++i) P[static_cast<size_t>(pinv[static_cast<size_t>(i)])] = i;
    {
        for (Index p = 0; p < A.row_ptrs[m]; ++p) cstart[static_cast<size_t>(A.col_indices[p]) + 1]++;
        for (Index k = 0; k < A.cols; ++k) cstart[k + 1] += cstart[k];
        crow.resize(static_cast<size_t>(cstart[A.cols]));
        std::vector<Index> nxt(cstart.begin(), cstart.end() - 1);
        for (Index i = 0; i < m; ++i)
            for (Index p = A.row_ptrs[i]; p < A.row_ptrs[i + 1]; ++p)
                crow[static_cast<size_t>(nxt[static_cast<size_t>(A.col_indices[p])]++)] =
                    pinv[static_cast<size_t>(i)];
    }
    std::vector<std::vector<Index>> rows(static_cast<size_t>(m));
    host::parallel_rows(m, [&] {
        return [&, mark = std::vector<Index>(static_cast<size_t>(m), -1)](int64_t rr) mutable {
            const Index r = static_cast<Index>(rr), i = P[static_cast<size_t>(r)];
            auto& v = rows[static_cast<size_t>(r)];
            for (Index p = A.row_ptrs[i]; p < A.row_ptrs[i + 1]; ++p) {
                const Index k = A.col_indices[p];
                for (Index a = cstart[k]; a < cstart[k + 1]; ++a) {
                    const Index rb = crow[static_cast<size_t>(a)];
                    if (rb >= r || mark[static_cast<size_t>(rb)] == r) continue;
                    mark[static_cast<size_t>(rb)] = r;
                    v.push_back(rb);
                }
            }
            std::sort(v.begin(), v.end());
        };
    });
    ptr.assign(static_cast<size_t>(m) + 1, 0);
    for (Index r = 0; r < m; ++r) ptr[r + 1] = ptr[r] + static_cast<Index>(rows[static_cast<size_t>(r)].size());
    idx.resize(static_cast<size_t>(ptr[m]));
    host::parallel_rows(m, [&] {
        return [&](int64_t r) {
            std::copy(rows[static_cast<size_t>(r)].begin(), rows[static_cast<size_t>(r)].end(),
                      idx.begin() + ptr[static_cast<size_t>(r)]);
            std::vector<Index>().swap(rows[static_cast<size_t>(r)]);
        };
    });
}

// Liu's elimination-tree algorithm with path compression over the lower rows.
std::vector<Index> etree_from_lower(Index m, const std::vector<Index>& ptr,
                                    const std::vector<Index>& idx) {
    std::vector<Index> parent(static_cast<size_t>(m), -1), anc(static_cast<size_t>(m), -1);
    for (Index i = 0; i < m; ++i) {
        for (Index p = ptr[i]; p < ptr[i + 1]; ++p) {
            Index j = idx[p];
            while (j != -1 && j < i) {
                const Index next = anc[j];
                anc[j] = i;
                if (next == -1) { parent[j] = i; break; }
                j = next;
            }
        }
    }
    return parent;
}

// Depth-first postorder (children visited in ascending order): post[k] = node at position k.
std::vector<Index> postorder(const std::vector<Index>& parent) {
    const Index m = static_cast<Index>(parent.size());
    std::vector<Index> head(static_cast<size_t>(m), -1), next(static_cast<size_t>(m), -1);
    for (Index j = m - 1; j >= 0; --j) {
        const Index p = parent[j];
        if (p != -1) { next[j] = head[p]; head[p] = j; }
    }
    std::vector<Index> post;
    post.reserve(static_cast<size_t>(m));
    std::vector<Index> stack;
    for (Index r = 0; r < m; ++r) {
        if (parent[r] != -1) continue;
        stack.push_back(r);
        while (!stack.empty()) {
            const Index j = stack.back();
            if (head[j] != -1) {           // descend into the next unvisited child
                const Index c = head[j];
                head[j] = next[c];
                stack.push_back(c);
            } else {
                post.push_back(j);
                stack.pop_back();
            }
        }
    }
    return post;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Reference ordering rule (src/ipm/symbolic.cpp:11-98).
std::vector<Index> compute_amd_ordering(const core::CSRMatrix& A) {
    const Index m = A.rows;
    auto by_col = rows_by_column(A);
    std::vector<std::vector<Index>> adj(static_cast<size_t>(m));
    for (const auto& rows : by_col)
        for (size_t a = 0; a < rows.size(); ++a)
            for (size_t b = a + 1; b < rows.size(); ++b) {
                adj[static_cast<size_t>(rows[a])].push_back(rows[b]);
                adj[static_cast<size_t>(rows[b])].push_back(rows[a]);
            }
    for (Index i = 0; i < m; ++i) {
        auto& v = adj[static_cast<size_t>(i)];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        auto it = std::lower_bound(v.begin(), v.end(), i);
        if (it != v.end() && *it == i) v.erase(it);
    }
    std::vector<Index> order;
    order.reserve(static_cast<size_t>(m));
    std::vector<char> done(static_cast<size_t>(m), 0);
    for (Index step = 0; step < m; ++step) {
        Index best = -1;
        size_t best_deg = static_cast<size_t>(m) + 1;
        for (Index i = 0; i < m; ++i)      // strict '<' in ascending order: lowest index wins ties
            if (!done[i] && adj[i].size() < best_deg) { best_deg = adj[i].size(); best = i; }
        if (best < 0) throw std::runtime_error("compute_amd_ordering: invalid state");
        done[best] = 1;
        order.push_back(best);
        const std::vector<Index> nb = adj[static_cast<size_t>(best)];
        for (size_t a = 0; a < nb.size(); ++a) {
            auto& va = adj[static_cast<size_t>(nb[a])];
            va.erase(std::remove(va.begin(), va.end(), best), va.end());
            for (size_t b = 0; b < nb.size(); ++b) {
                if (a == b) continue;
                auto it = std::lower_bound(va.begin(), va.end(), nb[b]);
                if (it == va.end() || *it != nb[b]) va.insert(it, nb[b]);
            }
        }
    }
    return order;
}

// ---------------------------------------------------------------------------------------------
void compute_symbolic_factorization(const core::CSRMatrix& A, const std::vector<Index>& P,
                                    SymbolicFactorization& out) {
    const Index m = A.rows;
    if (static_cast<Index>(P.size()) != m)
        throw std::invalid_argument("compute_symbolic_factorization: permutation size mismatch");
    out.P = P;
    out.P_inv.assign(static_cast<size_t>(m), -1);
    for (Index i = 0; i < m; ++i) {
        if (P[i] < 0 || P[i] >= m || out.P_inv[P[i]] != -1)
            throw std::invalid_argument("compute_symbolic_factorization: P is not a permutation");
        out.P_inv[P[i]] = i;
    }
    std::vector<Index> lp, li;
    permuted_lower_graph(A, out.P_inv, lp, li);
    out.parent = etree_from_lower(m, lp, li);

    // Row i of L = union of etree paths from each j in M(i, <i) up to i (row subtree).
    auto& L = out.L_pattern;
    L.rows = m;
    L.cols = m;
    L.row_ptrs.assign(static_cast<size_t>(m) + 1, 0);
    L.col_indices.clear();
    std::vector<Index> mark(static_cast<size_t>(m), -1), row;
    for (Index i = 0; i < m; ++i) {
        row.clear();
        mark[i] = i;
        for (Index p = lp[i]; p < lp[i + 1]; ++p)
            for (Index j = li[p]; j != -1 && mark[j] != i; j = out.parent[j]) {
                mark[j] = i;
                row.push_back(j);
            }
        std::sort(row.begin(), row.end());
        L.col_indices.insert(L.col_indices.end(), row.begin(), row.end());
        L.row_ptrs[i + 1] = static_cast<Index>(L.col_indices.size());
    }
    L.values.assign(L.col_indices.size(), 1.0);
}

// ---------------------------------------------------------------------------------------------
void compute_supernodal_structure(const core::CSRMatrix& A, const std::vector<Index>& P_in,
                                  const SupernodeOptions& opt, SupernodalStructure& S) {
    const Index m = A.rows;
    if (static_cast<Index>(P_in.size()) != m)
        throw std::invalid_argument("compute_supernodal_structure: permutation size mismatch");
    std::vector<Index> pinv(static_cast<size_t>(m));
    for (Index i = 0; i < m; ++i) pinv[P_in[i]] = i;

    // 1. etree of the given ordering, then relabel by a postorder (same fill, contiguous subtrees).
    std::vector<Index> lp, li;
    permuted_lower_graph(A, pinv, lp, li);
    std::vector<Index> parent0 = etree_from_lower(m, lp, li);
    std::vector<Index> post = postorder(parent0);
    std::vector<Index> P(static_cast<size_t>(m));
    for (Index k = 0; k < m; ++k) P[k] = P_in[post[k]];
    for (Index i = 0; i < m; ++i) pinv[P[i]] = i;
    permuted_lower_graph(A, pinv, lp, li);
    std::vector<Index> parent = etree_from_lower(m, lp, li);

    // 2. exact column counts (incl. diagonal) by row-subtree traversal.
    std::vector<int64_t> cc(static_cast<size_t>(m), 1);
    {
        std::vector<Index> mark(static_cast<size_t>(m), -1);
        for (Index i = 0; i < m; ++i) {
            mark[i] = i;
            for (Index p = lp[i]; p < lp[i + 1]; ++p)
                for (Index j = li[p]; j != -1 && mark[j] != i; j = parent[j]) {
                    mark[j] = i;
                    cc[j]++;
                }
        }
    }
    std::vector<Index> nchild(static_cast<size_t>(m), 0);
    for (Index j = 0; j < m; ++j)
        if (parent[j] != -1) nchild[parent[j]]++;

    // 3. fundamental supernodes.
    std::vector<Index> first;
    for (Index j = 0; j < m; ++j) {
        const bool chain = j > 0 && parent[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1;
        if (!chain) first.push_back(j);
    }
    first.push_back(m);
    Index ns = static_cast<Index>(first.size()) - 1;
    std::vector<Index> colsn(static_cast<size_t>(m));
    for (Index s = 0; s < ns; ++s)
        for (Index j = first[s]; j < first[s + 1]; ++j) colsn[j] = s;
    // per-supernode: columns, panel rows (ld), exact nonzeros of the lower trapezoid.
    std::vector<int64_t> ncol(static_cast<size_t>(ns)), ld(static_cast<size_t>(ns)),
        exact(static_cast<size_t>(ns), 0);
    std::vector<Index> sp(static_cast<size_t>(ns), -1);
    for (Index s = 0; s < ns; ++s) {
        ncol[s] = first[s + 1] - first[s];
        ld[s] = cc[first[s]];
        for (Index j = first[s]; j < first[s + 1]; ++j) exact[s] += cc[j];
        const Index pj = parent[first[s + 1] - 1];
        sp[s] = pj == -1 ? -1 : colsn[pj];
    }

    // 4. relaxed amalgamation: top-down, merge the contiguous preceding child into its parent
    //    (never a wide child into a parent with other children: see serial_child).
    auto stored = [](int64_t n, int64_t l) { return n * l - n * (n - 1) / 2; };
    std::vector<char> alive(static_cast<size_t>(ns), 1);
    std::vector<std::vector<Index>> skids(static_cast<size_t>(ns));
    std::vector<double> subw(static_cast<size_t>(ns), 0.0);   // subtree work, sum of cc^2
    for (Index s = 0; s < ns; ++s) {                          // children precede parents
        for (Index j = first[s]; j < first[s + 1]; ++j) subw[s] += static_cast<double>(cc[j]) * cc[j];
        if (sp[s] != -1) {
            skids[sp[s]].push_back(s);
            subw[sp[s]] += subw[s];
        }
    }
    std::vector<Index> absorbed_into(static_cast<size_t>(ns));
    std::iota(absorbed_into.begin(), absorbed_into.end(), 0);
    std::vector<Index> sfirst(first.begin(), first.end() - 1);
    for (Index p = ns - 1; p >= 0; --p) {
        if (!alive[p]) continue;
        while (true) {
            const Index f = sfirst[p];
            if (f == 0) break;
            Index c = colsn[f - 1];
            while (absorbed_into[c] != c) c = absorbed_into[c];
            if (!alive[c] || c == p) break;
            Index cp = sp[c];
            while (cp != -1 && absorbed_into[cp] != cp) cp = absorbed_into[cp];
            if (cp != p) break;
            const int64_t n = ncol[c] + ncol[p];
            if (n > opt.max_cols) break;
            if (ncol[c] > opt.serial_child) {
                double other = 0.0;
                for (const Index k : skids[p])
                    if (k != c) other = std::max(other, subw[k]);
                if (other >= opt.serial_frac * subw[c]) break;
            }
            const int64_t l = ncol[c] + ld[p];
            const int64_t st = stored(n, l);
            const int64_t zeros = st - (exact[c] + exact[p]);
            const bool ok = n <= opt.relax_small ||
                            static_cast<double>(zeros) <= opt.relax_frac * static_cast<double>(st);
            if (!ok) break;
            alive[c] = 0;
            absorbed_into[c] = p;
            skids[p].erase(std::find(skids[p].begin(), skids[p].end(), c));
            skids[p].insert(skids[p].end(), skids[c].begin(), skids[c].end());
            sfirst[p] = sfirst[c];
            ncol[p] = n;
            ld[p] = l;
            exact[p] += exact[c];
        }
    }
    // compact the surviving supernodes in column order
    std::vector<Index> newid(static_cast<size_t>(ns), -1);
    S.sn_first.clear();
    Index nsup = 0;
    for (Index s = 0; s < ns; ++s)
        if (alive[s]) { newid[s] = nsup++; }
    std::vector<std::pair<Index, Index>> order;   // (first col, old id)
    for (Index s = 0; s < ns; ++s)
        if (alive[s]) order.emplace_back(sfirst[s], s);
    std::sort(order.begin(), order.end());
    std::vector<Index> old_of(static_cast<size_t>(nsup));
    for (Index k = 0; k < nsup; ++k) { old_of[k] = order[k].second; newid[order[k].second] = k; }
    S.m = m;
    S.nsuper = nsup;
    S.sn_first.resize(static_cast<size_t>(nsup) + 1);
    for (Index k = 0; k < nsup; ++k) S.sn_first[k] = order[k].first;
    S.sn_first[nsup] = m;
    S.col_to_sn.assign(static_cast<size_t>(m), -1);
    for (Index k = 0; k < nsup; ++k)
        for (Index j = S.sn_first[k]; j < S.sn_first[k + 1]; ++j) S.col_to_sn[j] = k;
    S.sn_parent.assign(static_cast<size_t>(nsup), -1);
    for (Index k = 0; k < nsup; ++k) {
        const Index pj = parent[S.sn_first[k + 1] - 1];
        S.sn_parent[k] = pj == -1 ? -1 : S.col_to_sn[pj];
    }

    // 5. supernode row structures: own columns, M's lower entries of its columns, children's rows.
    std::vector<Index> up_ptr(static_cast<size_t>(m) + 1, 0), up_idx;  // column view of lower M
    for (Index i = 0; i < m; ++i)
        for (Index p = lp[i]; p < lp[i + 1]; ++p) up_ptr[static_cast<size_t>(li[p]) + 1]++;
    for (Index j = 0; j < m; ++j) up_ptr[j + 1] += up_ptr[j];
    up_idx.resize(static_cast<size_t>(up_ptr[m]));
    {
        std::vector<Index> nxt(up_ptr.begin(), up_ptr.end() - 1);
        for (Index i = 0; i < m; ++i)
            for (Index p = lp[i]; p < lp[i + 1]; ++p) up_idx[static_cast<size_t>(nxt[li[p]]++)] = i;
    }
    std::vector<std::vector<Index>> kids(static_cast<size_t>(nsup));
    for (Index k = 0; k < nsup; ++k)
        if (S.sn_parent[k] != -1) kids[static_cast<size_t>(S.sn_parent[k])].push_back(k);
    S.sn_rowptr.assign(static_cast<size_t>(nsup) + 1, 0);
    S.sn_rows.clear();
    std::vector<std::vector<Index>> rows(static_cast<size_t>(nsup));
    std::vector<Index> mark(static_cast<size_t>(m), -1);
    for (Index k = 0; k < nsup; ++k) {
        const Index f = S.sn_first[k], l = S.sn_first[k + 1];
        std::vector<Index> below;
        for (Index j = f; j < l; ++j)
            for (Index p = up_ptr[j]; p < up_ptr[j + 1]; ++p) {
                const Index i = up_idx[p];
                if (i >= l && mark[i] != k) { mark[i] = k; below.push_back(i); }
            }
        for (Index c : kids[static_cast<size_t>(k)]) {
            const auto& rc = rows[static_cast<size_t>(c)];
            const Index nc = S.sn_first[c + 1] - S.sn_first[c];
            for (size_t t = static_cast<size_t>(nc); t < rc.size(); ++t) {
                const Index i = rc[t];
                if (i >= l && mark[i] != k) { mark[i] = k; below.push_back(i); }
            }
        }
        std::sort(below.begin(), below.end());
        auto& rk = rows[static_cast<size_t>(k)];
        rk.reserve(static_cast<size_t>(l - f) + below.size());
        for (Index j = f; j < l; ++j) rk.push_back(j);
        rk.insert(rk.end(), below.begin(), below.end());
    }
    S.sn_level.assign(static_cast<size_t>(nsup), 0);
    S.max_level = 0;
    for (Index k = 0; k < nsup; ++k) {   // children precede parents (postorder)
        const Index p = S.sn_parent[k];
        if (p != -1) S.sn_level[p] = std::max(S.sn_level[p], S.sn_level[k] + 1);
        S.max_level = std::max(S.max_level, S.sn_level[k]);
    }
    S.sn_panel.assign(static_cast<size_t>(nsup) + 1, 0);
    S.nnz_L = 0;
    S.flops_panel = 0;
    for (Index k = 0; k < nsup; ++k) {
        const auto& rk = rows[static_cast<size_t>(k)];
        S.sn_rowptr[k + 1] = S.sn_rowptr[k] + static_cast<int64_t>(rk.size());
        S.sn_rows.insert(S.sn_rows.end(), rk.begin(), rk.end());
        const int64_t n = S.sn_first[k + 1] - S.sn_first[k];
        const int64_t L = static_cast<int64_t>(rk.size());
        S.sn_panel[k + 1] = S.sn_panel[k] + n * L;
        S.nnz_L += n * L - n * (n - 1) / 2;
        for (int64_t t = 0; t < n; ++t) S.flops_panel += static_cast<double>(L - t) * (L - t);
    }
    S.nnz_L_exact = 0;
    S.flops_exact = 0;
    for (Index j = 0; j < m; ++j) {
        S.nnz_L_exact += cc[j];
        S.flops_exact += static_cast<double>(cc[j]) * cc[j];
    }
    S.empty_row.assign(static_cast<size_t>(m), 0);
    for (Index i = 0; i < m; ++i)
        if (A.row_ptrs[P[i] + 1] == A.row_ptrs[P[i]]) S.empty_row[i] = 1;
    S.perm = P;
}

}  // namespace ipm
}  // namespace sankhya

// Approximate minimum degree ordering of the A*A^T graph on a quotient graph.
//
// Written from the published algorithm (Amestoy, Davis & Duff, "An approximate minimum degree
// ordering algorithm", SIAM J. Matrix Anal. Appl. 17(4), 1996): variables and elements,
// element absorption (plain and aggressive), supervariable detection by hashing, mass
// elimination, approximate external degrees from the |Le \ Lp| counts, and postponement of
// dense rows (degree > max(16, 10*sqrt(m))) to the end of the order.
// The reference orders with an explicit-graph minimum degree (src/ipm/symbolic.cpp:11-98),
// which is O(m^2)+ and cannot order the block-angular or 20k-row systems.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>
#include "ipm/symbolic.hpp"
#include "parallel.hpp"

namespace sankhya {
namespace ipm {

std::vector<Index> compute_fill_reducing_ordering(const core::CSRMatrix& A) {
    const Index m = A.rows;
    std::vector<Index> order;
    order.reserve(static_cast<size_t>(m));
    if (m == 0) return order;

    // ---- graph of A*A^T (no self loops): row i's neighbours are the rows of the columns it
    //      touches, deduplicated with a per-thread marker and sorted (rows in parallel) ----
    std::vector<std::vector<Index>> avar(static_cast<size_t>(m));   // variable adjacency
    {
        const Index nc = A.cols;
        std::vector<Index> cptr(static_cast<size_t>(nc) + 1, 0);
        for (Index p = 0; p < A.row_ptrs[m]; ++p) ++cptr[static_cast<size_t>(A.col_indices[p]) + 1];
        for (Index j = 0; j < nc; ++j) cptr[static_cast<size_t>(j) + 1] += cptr[static_cast<size_t>(j)];
        std::vector<Index> crow(static_cast<size_t>(A.row_ptrs[m]));
        {
            std::vector<Index> fill(cptr.begin(), cptr.end() - 1);
            for (Index i = 0; i < m; ++i)
                for (Index p = A.row_ptrs[i]; p < A.row_ptrs[i + 1]; ++p)
                    crow[static_cast<size_t>(fill[static_cast<size_t>(A.col_indices[p])]++)] = i;
        }
        host::parallel_rows(m, [&] {
            return [&, mark = std::vector<Index>(static_cast<size_t>(m), -1)](int64_t ii) mutable {
                const Index i = static_cast<Index>(ii);
                auto& v = avar[static_cast<size_t>(i)];
                for (Index p = A.row_ptrs[i]; p < A.row_ptrs[i + 1]; ++p) {
                    const Index j = A.col_indices[p];
                    for (Index q = cptr[static_cast<size_t>(j)]; q < cptr[static_cast<size_t>(j) + 1]; ++q) {
                        const Index r = crow[static_cast<size_t>(q)];
                        if (r == i || mark[static_cast<size_t>(r)] == i) continue;
                        mark[static_cast<size_t>(r)] = i;
                        v.push_back(r);
                    }
                }
                std::sort(v.begin(), v.end());
                v.shrink_to_fit();
            };
        });
    }

    // ---- state ----
    enum : char { VAR = 0, ELEMENT = 1, DEAD = 2, MERGED = 3, DENSE = 4 };
    std::vector<char> kind(static_cast<size_t>(m), VAR);
    std::vector<std::vector<Index>> elist(static_cast<size_t>(m));   // elements adjacent to var
    std::vector<std::vector<Index>> lset(static_cast<size_t>(m));    // variables of an element
    std::vector<std::vector<Index>> members(static_cast<size_t>(m)); // merged vars (supervariable)
    std::vector<int64_t> wt(static_cast<size_t>(m), 1);              // supervariable weight
    std::vector<int64_t> esz(static_cast<size_t>(m), 0);             // weighted |L_e|
    std::vector<int64_t> deg(static_cast<size_t>(m), 0);             // approx external degree
    std::vector<int64_t> wcnt(static_cast<size_t>(m), -1);           // |L_e \ L_p| scratch
    std::vector<Index> mark(static_cast<size_t>(m), -1);
    int64_t remaining = m;

    // ---- degree buckets (doubly linked) ----
    std::vector<Index> bhead(static_cast<size_t>(m) + 1, -1), bnext(static_cast<size_t>(m), -1),
        bprev(static_cast<size_t>(m), -1);
    int64_t mindeg = 0;
    auto bucket_insert = [&](Index i) {
        int64_t d = std::min<int64_t>(std::max<int64_t>(deg[i], 0), m);
        deg[i] = d;
        bprev[i] = -1;
        bnext[i] = bhead[d];
        if (bhead[d] != -1) bprev[bhead[d]] = i;
        bhead[d] = i;
        if (d < mindeg) mindeg = d;
    };
    auto bucket_remove = [&](Index i) {
        const int64_t d = deg[i];
        if (bprev[i] != -1) bnext[bprev[i]] = bnext[i]; else bhead[d] = bnext[i];
        if (bnext[i] != -1) bprev[bnext[i]] = bprev[i];
        bnext[i] = bprev[i] = -1;
    };

    // ---- dense rows are postponed ----
    const int64_t dense = std::max<int64_t>(16, static_cast<int64_t>(10.0 * std::sqrt(double(m))));
    std::vector<Index> dense_rows;
    for (Index i = 0; i < m; ++i)
        if (static_cast<int64_t>(avar[i].size()) > dense) { kind[i] = DENSE; dense_rows.push_back(i); }
    remaining -= static_cast<int64_t>(dense_rows.size());
    for (Index i = 0; i < m; ++i) {
        if (kind[i] != VAR) continue;
        auto& v = avar[i];
        v.erase(std::remove_if(v.begin(), v.end(), [&](Index j) { return kind[j] == DENSE; }),
                v.end());
        deg[i] = static_cast<int64_t>(v.size());
        bucket_insert(i);
    }
    for (Index i : dense_rows) { avar[i].clear(); avar[i].shrink_to_fit(); }

    std::vector<Index> lp;          // pivot element's variables
    std::vector<std::pair<uint64_t, Index>> hashed;
    int64_t eliminated = 0;
    const int64_t target = m - static_cast<int64_t>(dense_rows.size());
    Index stamp = 0;

    auto emit = [&](Index v) {
        order.push_back(v);
        for (Index u : members[v]) order.push_back(u);
        members[v].clear();
        members[v].shrink_to_fit();
    };

    while (eliminated < target) {
        // -- select the pivot supervariable of minimum approximate degree --
        while (mindeg <= m && bhead[mindeg] == -1) ++mindeg;
        const Index p = bhead[mindeg];
        bucket_remove(p);
        kind[p] = ELEMENT;
        emit(p);
        eliminated += wt[p];
        remaining -= wt[p];

        // -- Lp = (A_p U union of L_e, e in E_p) \ {p} over principal variables --
        ++stamp;
        lp.clear();
        mark[p] = stamp;
        int64_t lpw = 0;
        for (Index v : avar[p])
            if (kind[v] == VAR && mark[v] != stamp) { mark[v] = stamp; lp.push_back(v); lpw += wt[v]; }
        for (Index e : elist[p]) {
            if (kind[e] != ELEMENT) continue;
            for (Index v : lset[e])
                if (kind[v] == VAR && mark[v] != stamp) { mark[v] = stamp; lp.push_back(v); lpw += wt[v]; }
            kind[e] = DEAD;                   // absorbed into p
            lset[e].clear();
            lset[e].shrink_to_fit();
        }
        avar[p].clear(); avar[p].shrink_to_fit();
        elist[p].clear(); elist[p].shrink_to_fit();
        esz[p] = lpw;

        // -- |L_e \ L_p| for elements adjacent to Lp --
        for (Index i : lp) {
            for (Index e : elist[i]) {
                if (kind[e] != ELEMENT || e == p) continue;
                if (wcnt[e] < 0) wcnt[e] = esz[e];
                wcnt[e] -= wt[i];
            }
        }
        // -- prune adjacency, aggressive absorption, approximate degrees --
        for (Index i : lp) {
            bucket_remove(i);
            auto& ei = elist[i];
            int64_t ext = 0;
            size_t w = 0;
            for (Index e : ei) {
                if (kind[e] != ELEMENT || e == p) continue;
                if (wcnt[e] == 0) {           // L_e subset of L_p: e absorbed by p
                    kind[e] = DEAD;
                    lset[e].clear();
                    lset[e].shrink_to_fit();
                    continue;
                }
                ext += wcnt[e];
                ei[w++] = e;
            }
            ei.resize(w);
            ei.push_back(p);
            auto& ai = avar[i];
            w = 0;
            int64_t aw = 0;
            for (Index v : ai) {
                if (kind[v] != VAR || mark[v] == stamp) continue;   // in Lp or gone
                ai[w++] = v;
                aw += wt[v];
            }
            ai.resize(w);
            const int64_t lpo = lpw - wt[i];
            int64_t d = std::min<int64_t>(remaining - wt[i], deg[i] + lpo);
            d = std::min<int64_t>(d, aw + lpo + ext);
            deg[i] = std::max<int64_t>(d, 0);
        }
        for (Index i : lp)
            for (Index e : elist[i])
                if (e != p && e >= 0 && e < m) wcnt[e] = -1;
        lset[p] = lp;

        // -- supervariable detection: identical (A_i, E_i) lists --
        hashed.clear();
        for (Index i : lp) {
            uint64_t h = 1469598103934665603ULL;
            auto& ai = avar[i];
            auto& ei = elist[i];
            std::sort(ai.begin(), ai.end());
            std::sort(ei.begin(), ei.end());
            for (Index v : ai) h = (h ^ static_cast<uint64_t>(v + 1)) * 1099511628211ULL;
            h ^= 0x9e3779b97f4a7c15ULL;
            for (Index e : ei) h = (h ^ static_cast<uint64_t>(e + 1)) * 1099511628211ULL;
            hashed.emplace_back(h, i);
        }
        std::sort(hashed.begin(), hashed.end());
        for (size_t a = 0; a < hashed.size();) {
            size_t b = a + 1;
            while (b < hashed.size() && hashed[b].first == hashed[a].first) ++b;
            for (size_t x = a; x < b; ++x) {
                const Index i = hashed[x].second;
                if (kind[i] != VAR) continue;
                for (size_t y = x + 1; y < b; ++y) {
                    const Index j = hashed[y].second;
                    if (kind[j] != VAR) continue;
                    if (avar[i] != avar[j] || elist[i] != elist[j]) continue;
                    // merge j into i
                    kind[j] = MERGED;
                    wt[i] += wt[j];
                    deg[i] -= wt[j];
                    members[i].push_back(j);
                    for (Index u : members[j]) members[i].push_back(u);
                    members[j].clear();
                    members[j].shrink_to_fit();
                    avar[j].clear(); avar[j].shrink_to_fit();
                    elist[j].clear(); elist[j].shrink_to_fit();
                }
            }
            a = b;
        }

        // -- mass elimination: variables whose only neighbour is the new element --
        size_t keep = 0;
        for (Index i : lp) {
            if (kind[i] != VAR) continue;
            if (avar[i].empty() && elist[i].size() == 1 && elist[i][0] == p) {
                kind[i] = DEAD;
                emit(i);
                eliminated += wt[i];
                remaining -= wt[i];
                esz[p] -= wt[i];
                elist[i].clear();
                continue;
            }
            lp[keep++] = i;
        }
        lp.resize(keep);
        lset[p] = lp;
        for (Index i : lp) {
            deg[i] = std::min<int64_t>(deg[i], remaining - wt[i]);
            bucket_insert(i);
        }
        if (lp.empty()) kind[p] = DEAD;
    }
    for (Index i : dense_rows) order.push_back(i);
    return order;
}

}  // namespace ipm
}  // namespace sankhya

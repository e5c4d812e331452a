// Symbolic factorization: elimination tree, row/column counts, and the new
// supernodal structure the device factorization runs on.
#include <algorithm>
#include <map>
#include <numeric>

#include "aphi_host.hpp"

namespace aphi_amd {

namespace {

// Entries of column k above the diagonal, ascending (src/symbolic.cpp:18-27).
template <typename F>
inline void for_earlier(const SymmetricAdjacency& adj, int k, F&& f) {
    for (int t = adj.offset[static_cast<size_t>(k)]; t < adj.offset[static_cast<size_t>(k) + 1]; ++t) {
        const int i = adj.neighbor[static_cast<size_t>(t)];
        if (i < k) f(i);
    }
}

}  // namespace

namespace {

std::vector<int> etree_of(const SymmetricAdjacency& adj) {
    const int n = adj.rows();
    std::vector<int> parent(static_cast<size_t>(n), -1), anc(static_cast<size_t>(n), -1);
    for (int k = 0; k < n; ++k) {
        for_earlier(adj, k, [&](int i) {
            while (i != -1 && i < k) {
                const int nx = anc[static_cast<size_t>(i)];
                anc[static_cast<size_t>(i)] = k;
                if (nx == -1) parent[static_cast<size_t>(i)] = k;
                i = nx;
            }
        });
    }
    return parent;
}

}  // namespace

std::vector<int> elimination_tree(const SparsityPattern& permuted) {
    if (permuted.rows != permuted.cols) throw std::invalid_argument("elimination_tree: the pattern must be square");
    return etree_of(SymmetricAdjacency::build(permuted));
}

namespace {

// Row k of L is the tree-reach of A(k, 0:k-1) (src/symbolic.cpp:29-55). Counts
// rows and, when `cols` is given, columns, without storing the pattern.
void reach_counts(const SymmetricAdjacency& adj, const std::vector<int>& parent, std::vector<int>& rows,
                  std::vector<int>* cols) {
    const int n = adj.rows();
    std::vector<int> mark(static_cast<size_t>(n), -1);
    rows.assign(static_cast<size_t>(n), 0);
    if (cols) cols->assign(static_cast<size_t>(n), 0);
    for (int k = 0; k < n; ++k) {
        mark[static_cast<size_t>(k)] = k;
        int count = 1;
        if (cols) ++(*cols)[static_cast<size_t>(k)];
        for_earlier(adj, k, [&](int i) {
            while (i >= 0 && mark[static_cast<size_t>(i)] != k) {
                mark[static_cast<size_t>(i)] = k;
                ++count;
                if (cols) ++(*cols)[static_cast<size_t>(i)];
                i = parent[static_cast<size_t>(i)];
            }
        });
        rows[static_cast<size_t>(k)] = count;
    }
}

// The same row and column counts in time nearly linear in nnz(A) (Gilbert, Ng and
// Peyton, SIAM J. Matrix Anal. Appl. 15(4), 1994, as set out in Davis, "Direct
// Methods for Sparse Linear Systems", SIAM 2006, section 4.5): the row subtree of
// row i is the union of the tree paths from its leaves up to i. Visiting the
// nodes in postorder, node j is a leaf of row subtree i (A(i, j) != 0, i > j)
// exactly when its first descendant comes after the previous leaf's; the new path
// from j stops below q, the least common ancestor of j and the previous leaf,
// found by a path-compressed union-find over the nodes already finished. Row i
// gains depth(j) - depth(q) nodes (q = i for the first leaf); column counts come
// from the same leaves as differences summed up the tree.
void skeleton_counts(const SymmetricAdjacency& adj, const std::vector<int>& parent, std::vector<int>& rows,
                     std::vector<int>& cols) {
    const int n = adj.rows();
    const auto at = [](std::vector<int>& v, int i) -> int& { return v[static_cast<size_t>(i)]; };
    // postorder, children ascending
    std::vector<int> head(static_cast<size_t>(n), -1), next(static_cast<size_t>(n), -1), post, stack;
    post.reserve(static_cast<size_t>(n));
    for (int j = n - 1; j >= 0; --j) {
        const int p = parent[static_cast<size_t>(j)];
        if (p < 0) continue;
        at(next, j) = at(head, p);
        at(head, p) = j;
    }
    for (int r = 0; r < n; ++r) {
        if (parent[static_cast<size_t>(r)] >= 0) continue;
        stack.push_back(r);
        while (!stack.empty()) {
            const int v = stack.back();
            const int c = at(head, v);
            if (c < 0) {
                stack.pop_back();
                post.push_back(v);
            } else {
                at(head, v) = at(next, c);
                stack.push_back(c);
            }
        }
    }
    // depth (roots 0; parents come after their children in postorder) and first
    // descendant (postorder position); a leaf of the tree starts its column at 1
    std::vector<int> depth(static_cast<size_t>(n), 0), first(static_cast<size_t>(n), -1);
    cols.assign(static_cast<size_t>(n), 0);
    for (int k = n - 1; k >= 0; --k) {
        const int j = post[static_cast<size_t>(k)], p = parent[static_cast<size_t>(j)];
        at(depth, j) = p < 0 ? 0 : at(depth, p) + 1;
    }
    for (int k = 0; k < n; ++k) {
        int j = post[static_cast<size_t>(k)];
        at(cols, j) = at(first, j) == -1 ? 1 : 0;
        for (; j != -1 && at(first, j) == -1; j = parent[static_cast<size_t>(j)]) at(first, j) = k;
    }
    std::vector<int> ancestor(static_cast<size_t>(n)), maxfirst(static_cast<size_t>(n), -1),
        prevleaf(static_cast<size_t>(n), -1);
    std::iota(ancestor.begin(), ancestor.end(), 0);
    rows.assign(static_cast<size_t>(n), 1);
    for (int k = 0; k < n; ++k) {
        const int j = post[static_cast<size_t>(k)], p = parent[static_cast<size_t>(j)];
        if (p != -1) --at(cols, p);
        for (int t = adj.offset[static_cast<size_t>(j)]; t < adj.offset[static_cast<size_t>(j) + 1]; ++t) {
            const int i = adj.neighbor[static_cast<size_t>(t)];
            if (i <= j || at(first, j) <= at(maxfirst, i)) continue;  // j is no leaf of row subtree i
            at(maxfirst, i) = at(first, j);
            const int jprev = at(prevleaf, i);
            at(prevleaf, i) = j;
            int q = i;
            if (jprev != -1) {
                for (q = jprev; q != at(ancestor, q);) q = at(ancestor, q);
                for (int u = jprev; u != q;) {
                    const int up = at(ancestor, u);
                    at(ancestor, u) = q;
                    u = up;
                }
                --at(cols, q);
            }
            ++at(cols, j);
            at(rows, i) += at(depth, j) - at(depth, q);
        }
        if (p != -1) at(ancestor, j) = p;
    }
    for (int j = 0; j < n; ++j) {  // parents have larger indices than their children
        const int p = parent[static_cast<size_t>(j)];
        if (p != -1) at(cols, p) += at(cols, j);
    }
}

}  // namespace

std::vector<int> factor_row_counts(const SparsityPattern& permuted, const std::vector<int>& parent) {
    if (static_cast<int>(parent.size()) != permuted.rows)
        throw std::invalid_argument("factor_row_counts: the elimination tree has " + std::to_string(parent.size()) +
                                    " entries for a pattern of " + std::to_string(permuted.rows) + " rows");
    std::vector<int> rows;
    reach_counts(SymmetricAdjacency::build(permuted), parent, rows, nullptr);
    return rows;
}

SolverAnalysis analyze(const SparsityPattern& pattern, Ordering ordering, size_t max_factor_nnz) {
    SolverAnalysis a;
    a.ordering = ordering;
    a.permutation = compute_ordering(pattern, ordering);
    a.permuted = permute_pattern(pattern, a.permutation);
    if (a.permuted.rows != a.permuted.cols) throw std::invalid_argument("analyze: the pattern must be square");
    const SymmetricAdjacency adj = SymmetricAdjacency::build(a.permuted);
    a.parent = etree_of(adj);
    skeleton_counts(adj, a.parent, a.row_count, a.col_count);
    for (int c : a.row_count) a.predicted_nnz += static_cast<size_t>(c);
    for (int p : a.parent)
        if (p < 0) ++a.num_roots;
    if (max_factor_nnz != 0 && a.predicted_nnz > max_factor_nnz)
        throw std::invalid_argument("analyze: the ordering predicts " + std::to_string(a.predicted_nnz) +
                                    " nonzeros in L, over the budget of " + std::to_string(max_factor_nnz) +
                                    ". Refusing before allocating it.");
    return a;
}

double ldlt_flops(const std::vector<int>& col_count) {
    double f = 0.0;
    for (int c : col_count) {
        const double m = static_cast<double>(c - 1);
        f += m * (m + 1.0) / 2.0 + 2.0 * m;
    }
    return 8.0 * f;
}

// ------------------------------------------------------------------ supernodes

SupernodePlan build_supernodes(const SolverAnalysis& an, const SupernodeOptions& opt) {
    const int n = an.permuted.rows;
    SupernodePlan sp;
    sp.n = n;
    sp.exact_factor_nnz = an.predicted_nnz;
    sp.flops = ldlt_flops(an.col_count);

    // 1. Postorder of the elimination tree (children ascending, roots ascending),
    // so every subtree -- and every amalgamated supernode -- is a contiguous range
    // of columns. Postordering is an equivalent reordering: same fill, same counts.
    std::vector<int> head(static_cast<size_t>(n), -1), next(static_cast<size_t>(n), -1);
    for (int j = n - 1; j >= 0; --j) {
        const int p = an.parent[static_cast<size_t>(j)];
        if (p < 0) continue;
        next[static_cast<size_t>(j)] = head[static_cast<size_t>(p)];
        head[static_cast<size_t>(p)] = j;
    }
    std::vector<int> post;
    post.reserve(static_cast<size_t>(n));
    {
        std::vector<int> stack;
        for (int r = 0; r < n; ++r) {
            if (an.parent[static_cast<size_t>(r)] >= 0) continue;
            stack.push_back(r);
            while (!stack.empty()) {
                const int v = stack.back();
                const int c = head[static_cast<size_t>(v)];
                if (c < 0) {
                    stack.pop_back();
                    post.push_back(v);
                } else {
                    head[static_cast<size_t>(v)] = next[static_cast<size_t>(c)];
                    stack.push_back(c);
                }
            }
        }
    }
    std::vector<int> ipost(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) ipost[static_cast<size_t>(post[static_cast<size_t>(k)])] = k;
    std::vector<int> parent(static_cast<size_t>(n)), cc(static_cast<size_t>(n)), nchild(static_cast<size_t>(n), 0);
    for (int k = 0; k < n; ++k) {
        const int p = an.parent[static_cast<size_t>(post[static_cast<size_t>(k)])];
        parent[static_cast<size_t>(k)] = p < 0 ? -1 : ipost[static_cast<size_t>(p)];
        cc[static_cast<size_t>(k)] = an.col_count[static_cast<size_t>(post[static_cast<size_t>(k)])];
        if (p >= 0) ++nchild[static_cast<size_t>(parent[static_cast<size_t>(k)])];
    }
    sp.perm.resize(static_cast<size_t>(n));
    sp.iperm.resize(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) {
        sp.perm[static_cast<size_t>(k)] = an.permutation.perm[static_cast<size_t>(post[static_cast<size_t>(k)])];
        sp.iperm[static_cast<size_t>(sp.perm[static_cast<size_t>(k)])] = k;
    }

    // 2. Fundamental supernodes: j+1 continues j's supernode when it is j's only
    // parent-child link and the column counts nest exactly.
    std::vector<int> fstart;
    for (int j = 0; j < n; ++j) {
        const bool joins = j > 0 && parent[static_cast<size_t>(j) - 1] == j &&
                           cc[static_cast<size_t>(j) - 1] == cc[static_cast<size_t>(j)] + 1 &&
                           nchild[static_cast<size_t>(j)] == 1;
        if (!joins) fstart.push_back(j);
    }
    const int nf = static_cast<int>(fstart.size());
    fstart.push_back(n);
    std::vector<int> fsuper(static_cast<size_t>(n));
    for (int s = 0; s < nf; ++s)
        for (int j = fstart[static_cast<size_t>(s)]; j < fstart[static_cast<size_t>(s) + 1]; ++j) fsuper[static_cast<size_t>(j)] = s;

    // 3. Relaxed amalgamation, bottom-up: a supernode merges into its parent when
    // it is the parent's last child (so the columns stay contiguous) and the
    // explicit zeros this adds stay under a threshold that loosens for narrow
    // supernodes, where wider fronts pay for themselves on the device.
    std::vector<int64_t> k(static_cast<size_t>(nf)), m(static_cast<size_t>(nf));
    std::vector<double> nz(static_cast<size_t>(nf), 0.0);
    std::vector<int> fpar(static_cast<size_t>(nf), -1), merged_into(static_cast<size_t>(nf), -1);
    std::vector<int> first_col(static_cast<size_t>(nf));
    for (int s = 0; s < nf; ++s) {
        const int f = fstart[static_cast<size_t>(s)], l = fstart[static_cast<size_t>(s) + 1] - 1;
        k[static_cast<size_t>(s)] = l - f + 1;
        m[static_cast<size_t>(s)] = cc[static_cast<size_t>(f)];
        first_col[static_cast<size_t>(s)] = f;
        for (int j = f; j <= l; ++j) nz[static_cast<size_t>(s)] += cc[static_cast<size_t>(j)];
        const int pj = parent[static_cast<size_t>(l)];
        fpar[static_cast<size_t>(s)] = pj < 0 ? -1 : fsuper[static_cast<size_t>(pj)];
    }
    std::vector<int> last_col(static_cast<size_t>(nf));
    for (int s = 0; s < nf; ++s) last_col[static_cast<size_t>(s)] = fstart[static_cast<size_t>(s) + 1] - 1;
    for (int s = 0; s < nf; ++s) {
        const int p = fpar[static_cast<size_t>(s)];
        if (p < 0) continue;
        // contiguity: s's last column is immediately before p's first column
        if (last_col[static_cast<size_t>(s)] + 1 != first_col[static_cast<size_t>(p)]) continue;
        const int64_t kn = k[static_cast<size_t>(s)] + k[static_cast<size_t>(p)];
        const int64_t mn = k[static_cast<size_t>(s)] + m[static_cast<size_t>(p)];
        if (kn > opt.max_cols) continue;
        const double stored = static_cast<double>(kn) * mn - static_cast<double>(kn) * (kn - 1) / 2.0;
        const double zf = (stored - nz[static_cast<size_t>(s)] - nz[static_cast<size_t>(p)]) / stored;
        bool merge;
        if (kn <= 4) merge = true;
        else if (kn <= opt.relax_small) merge = zf < 0.8;
        else if (kn <= 3 * opt.relax_small) merge = zf < opt.relax_fill;
        else merge = zf < opt.relax_fill * 0.5;
        if (!merge) continue;
        // p absorbs s: p now starts at s's first column
        k[static_cast<size_t>(p)] = kn;
        m[static_cast<size_t>(p)] = mn;
        nz[static_cast<size_t>(p)] += nz[static_cast<size_t>(s)];
        first_col[static_cast<size_t>(p)] = first_col[static_cast<size_t>(s)];
        merged_into[static_cast<size_t>(s)] = p;
    }

    // 4. Final partition.
    sp.first.clear();
    std::vector<int64_t> expect_m;
    for (int s = 0; s < nf; ++s) {
        if (merged_into[static_cast<size_t>(s)] >= 0) continue;
        sp.first.push_back(first_col[static_cast<size_t>(s)]);
        expect_m.push_back(m[static_cast<size_t>(s)]);
    }
    sp.num_super = static_cast<int>(sp.first.size());
    sp.first.push_back(n);
    const int ns = sp.num_super;
    sp.col_to_super.assign(static_cast<size_t>(n), -1);
    for (int s = 0; s < ns; ++s)
        for (int j = sp.first[static_cast<size_t>(s)]; j < sp.first[static_cast<size_t>(s) + 1]; ++j)
            sp.col_to_super[static_cast<size_t>(j)] = s;
    for (int s = 1; s <= ns; ++s)
        if (sp.first[static_cast<size_t>(s)] <= sp.first[static_cast<size_t>(s) - 1])
            throw std::logic_error("build_supernodes: supernodes are not contiguous and increasing");
    sp.sparent.assign(static_cast<size_t>(ns), -1);
    for (int s = 0; s < ns; ++s) {
        const int l = sp.first[static_cast<size_t>(s) + 1] - 1;
        const int pj = parent[static_cast<size_t>(l)];
        sp.sparent[static_cast<size_t>(s)] = pj < 0 ? -1 : sp.col_to_super[static_cast<size_t>(pj)];
        if (pj >= 0 && sp.sparent[static_cast<size_t>(s)] <= s)
            throw std::logic_error("build_supernodes: a supernode's parent precedes it");
    }

    // 5. Row structure: own columns, then the union of the lower pattern of its
    // columns and of its children's structures, beyond its last column.
    // The pattern in factorization order: an.permuted (AMD order) relabelled
    // through the postorder.
    Permutation relabel;
    relabel.perm = post;
    relabel.iperm = ipost;
    const SparsityPattern fperm = permute_pattern(an.permuted, relabel);
    const SymmetricAdjacency adj = SymmetricAdjacency::build(fperm);

    sp.child_ptr.assign(static_cast<size_t>(ns) + 1, 0);
    for (int s = 0; s < ns; ++s)
        if (sp.sparent[static_cast<size_t>(s)] >= 0) ++sp.child_ptr[static_cast<size_t>(sp.sparent[static_cast<size_t>(s)]) + 1];
    for (int s = 0; s < ns; ++s) sp.child_ptr[static_cast<size_t>(s) + 1] += sp.child_ptr[static_cast<size_t>(s)];
    sp.child.assign(static_cast<size_t>(sp.child_ptr[static_cast<size_t>(ns)]), 0);
    {
        std::vector<int> cur(sp.child_ptr.begin(), sp.child_ptr.end() - 1);
        for (int s = 0; s < ns; ++s)
            if (sp.sparent[static_cast<size_t>(s)] >= 0)
                sp.child[static_cast<size_t>(cur[static_cast<size_t>(sp.sparent[static_cast<size_t>(s)])]++)] = s;
    }

    sp.row_ptr.assign(static_cast<size_t>(ns) + 1, 0);
    sp.rows.clear();
    std::vector<int> mark(static_cast<size_t>(n), -1), tmp;
    for (int s = 0; s < ns; ++s) {
        const int f = sp.first[static_cast<size_t>(s)], l = sp.first[static_cast<size_t>(s) + 1] - 1;
        tmp.clear();
        for (int j = f; j <= l; ++j) mark[static_cast<size_t>(j)] = s;
        for (int j = f; j <= l; ++j) {
            for (int t = adj.offset[static_cast<size_t>(j)]; t < adj.offset[static_cast<size_t>(j) + 1]; ++t) {
                const int i = adj.neighbor[static_cast<size_t>(t)];
                if (i > l && mark[static_cast<size_t>(i)] != s) {
                    mark[static_cast<size_t>(i)] = s;
                    tmp.push_back(i);
                }
            }
        }
        for (int c = sp.child_ptr[static_cast<size_t>(s)]; c < sp.child_ptr[static_cast<size_t>(s) + 1]; ++c) {
            const int ch = sp.child[static_cast<size_t>(c)];
            const int kc = sp.first[static_cast<size_t>(ch) + 1] - sp.first[static_cast<size_t>(ch)];
            for (int64_t r = sp.row_ptr[static_cast<size_t>(ch)] + kc; r < sp.row_ptr[static_cast<size_t>(ch) + 1]; ++r) {
                const int i = sp.rows[static_cast<size_t>(r)];
                if (i > l && mark[static_cast<size_t>(i)] != s) {
                    mark[static_cast<size_t>(i)] = s;
                    tmp.push_back(i);
                }
            }
        }
        std::sort(tmp.begin(), tmp.end());
        for (int j = f; j <= l; ++j) sp.rows.push_back(j);
        sp.rows.insert(sp.rows.end(), tmp.begin(), tmp.end());
        sp.row_ptr[static_cast<size_t>(s) + 1] = static_cast<int64_t>(sp.rows.size());
        const int64_t mm = sp.row_ptr[static_cast<size_t>(s) + 1] - sp.row_ptr[static_cast<size_t>(s)];
        // The amalgamated structure is the top fundamental supernode's plus the
        // merged columns.
        if (mm != expect_m[static_cast<size_t>(s)])
            throw std::logic_error("build_supernodes: row structure disagrees with the column counts");
    }

    // 6. Relative indices into the parent, levels, flat panel layout.
    sp.rel.assign(sp.rows.size(), -1);
    sp.level.assign(static_cast<size_t>(ns), 0);
    sp.panel_offset.assign(static_cast<size_t>(ns) + 1, 0);
    for (int s = 0; s < ns; ++s) {
        const int64_t b = sp.row_ptr[static_cast<size_t>(s)], e = sp.row_ptr[static_cast<size_t>(s) + 1];
        const int64_t kk = sp.first[static_cast<size_t>(s) + 1] - sp.first[static_cast<size_t>(s)];
        sp.panel_offset[static_cast<size_t>(s) + 1] = sp.panel_offset[static_cast<size_t>(s)] + (e - b) * kk;
        const int p = sp.sparent[static_cast<size_t>(s)];
        if (p >= 0) {
            sp.level[static_cast<size_t>(p)] = std::max(sp.level[static_cast<size_t>(p)], sp.level[static_cast<size_t>(s)] + 1);
            const int64_t pb = sp.row_ptr[static_cast<size_t>(p)], pe = sp.row_ptr[static_cast<size_t>(p) + 1];
            int64_t q = pb;
            for (int64_t r = b + kk; r < e; ++r) {
                const int want = sp.rows[static_cast<size_t>(r)];
                while (q < pe && sp.rows[static_cast<size_t>(q)] < want) ++q;
                if (q == pe || sp.rows[static_cast<size_t>(q)] != want)
                    throw std::logic_error("build_supernodes: a child's row is missing from its parent");
                sp.rel[static_cast<size_t>(r)] = static_cast<int>(q - pb);
            }
        } else if (e - b != kk) {
            throw std::logic_error("build_supernodes: a root supernode has rows beyond its columns");
        }
    }
    sp.num_levels = 0;
    for (int s = 0; s < ns; ++s) sp.num_levels = std::max(sp.num_levels, sp.level[static_cast<size_t>(s)] + 1);
    sp.factor_entries = static_cast<size_t>(sp.panel_offset[static_cast<size_t>(ns)]);

    // 7. Contribution-block workspace. A block lives from its own level to its
    // parent's level, both inclusive; blocks whose lifetimes overlap get disjoint
    // ranges. A simulated first-fit allocator over the level schedule assigns
    // the offsets.
    sp.cb_offset.assign(static_cast<size_t>(ns), -1);
    {
        std::vector<std::vector<int>> born(static_cast<size_t>(sp.num_levels)), dies(static_cast<size_t>(sp.num_levels));
        for (int s = 0; s < ns; ++s) {
            const int p = sp.sparent[static_cast<size_t>(s)];
            if (p < 0) continue;
            born[static_cast<size_t>(sp.level[static_cast<size_t>(s)])].push_back(s);
            dies[static_cast<size_t>(sp.level[static_cast<size_t>(p)])].push_back(s);
        }
        std::map<int64_t, int64_t> free_blocks;  // offset -> size
        int64_t top = 0;
        const auto size_of = [&](int s) {
            const int64_t mm = sp.row_ptr[static_cast<size_t>(s) + 1] - sp.row_ptr[static_cast<size_t>(s)];
            const int64_t kk = sp.first[static_cast<size_t>(s) + 1] - sp.first[static_cast<size_t>(s)];
            const int64_t c = mm - kk;
            return ((c * c + 63) / 64) * 64;  // keep every block 1 KiB-aligned
        };
        for (int L = 0; L < sp.num_levels; ++L) {
            // Blocks written at level L are placed before the blocks level L reads
            // are released: within one level the writes of some fronts overlap the
            // reads of others, so a lifetime covers both of its end levels.
            for (int s : born[static_cast<size_t>(L)]) {
                const int64_t sz = size_of(s);
                int64_t off = -1;
                for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
                    if (it->second >= sz) {
                        off = it->first;
                        const int64_t rest = it->second - sz;
                        free_blocks.erase(it);
                        if (rest > 0) free_blocks.emplace(off + sz, rest);
                        break;
                    }
                }
                if (off < 0) {
                    // extend the pool; merge with a trailing free block if any
                    if (!free_blocks.empty()) {
                        auto last = std::prev(free_blocks.end());
                        if (last->first + last->second == top) {
                            off = last->first;
                            free_blocks.erase(last);
                            top = off + sz;
                        }
                    }
                    if (off < 0) {
                        off = top;
                        top += sz;
                    }
                }
                sp.cb_offset[static_cast<size_t>(s)] = off;
            }
            for (int s : dies[static_cast<size_t>(L)]) {
                int64_t off = sp.cb_offset[static_cast<size_t>(s)], sz = size_of(s);
                if (sz == 0) continue;
                auto it = free_blocks.emplace(off, sz).first;
                if (it != free_blocks.begin()) {
                    auto pv = std::prev(it);
                    if (pv->first + pv->second == it->first) {
                        pv->second += it->second;
                        free_blocks.erase(it);
                        it = pv;
                    }
                }
                auto nx = std::next(it);
                if (nx != free_blocks.end() && it->first + it->second == nx->first) {
                    it->second += nx->second;
                    free_blocks.erase(nx);
                }
            }
        }
        sp.cb_pool_entries = top;
    }
    return sp;
}

std::vector<int64_t> slot_to_panel(const SparsityPattern& pattern, const SolverAnalysis& a, const SupernodePlan& sp) {
    (void)a;
    std::vector<int64_t> out(pattern.nnz(), -1);
    for (int r = 0; r < pattern.rows; ++r) {
        for (int k = pattern.row_ptr[static_cast<size_t>(r)]; k < pattern.row_ptr[static_cast<size_t>(r) + 1]; ++k) {
            const int c = pattern.col_index[static_cast<size_t>(k)];
            int i = sp.iperm[static_cast<size_t>(r)], j = sp.iperm[static_cast<size_t>(c)];
            if (i < j) std::swap(i, j);  // (row i, column j), lower
            const int s = sp.col_to_super[static_cast<size_t>(j)];
            const int64_t b = sp.row_ptr[static_cast<size_t>(s)], e = sp.row_ptr[static_cast<size_t>(s) + 1];
            const auto it = std::lower_bound(sp.rows.begin() + b, sp.rows.begin() + e, i);
            if (it == sp.rows.begin() + e || *it != i)
                throw std::logic_error("slot_to_panel: an entry of A is outside its supernode's structure");
            const int64_t lr = it - (sp.rows.begin() + b);
            const int64_t lc = j - sp.first[static_cast<size_t>(s)];
            out[static_cast<size_t>(k)] = sp.panel_offset[static_cast<size_t>(s)] + lc * (e - b) + lr;
        }
    }
    return out;
}

}  // namespace aphi_amd

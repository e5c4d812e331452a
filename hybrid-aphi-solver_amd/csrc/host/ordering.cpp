// Fill-reducing orderings on the host: natural, reverse Cuthill-McKee and the
// reference's approximate minimum degree with supervariables.
//
// The factor's pattern depends on every tie-break of the ordering, so the AMD
// below follows src/ordering.cpp:220-510 step for step: LIFO degree buckets, the
// min-degree scan, L_p gathered from A_p then from E_p's members, absorption of
// E_p, pruning of each A_i/E_i, weighted |L_e \ L_p|, the three-way degree
// bound, and supervariable detection by a (sum of neighbours) mod n hash with
// head-inserted chains and a persistent `mark` array.
#include <algorithm>
#include <numeric>

#include "aphi_host.hpp"

namespace aphi_amd {

bool Permutation::is_valid() const {
    if (perm.size() != iperm.size()) return false;
    const int n = static_cast<int>(perm.size());
    std::vector<uint8_t> hit(static_cast<size_t>(n), 0);
    for (int i = 0; i < n; ++i) {
        const int o = perm[static_cast<size_t>(i)];
        if (o < 0 || o >= n || hit[static_cast<size_t>(o)]) return false;
        hit[static_cast<size_t>(o)] = 1;
        if (iperm[static_cast<size_t>(o)] != i) return false;
    }
    return true;
}

SymmetricAdjacency SymmetricAdjacency::build(const SparsityPattern& pattern) {
    const int n = pattern.rows;
    SymmetricAdjacency a;
    a.offset.assign(static_cast<size_t>(n) + 1, 0);
    for (int r = 0; r < n; ++r)
        for (int k = pattern.row_ptr[static_cast<size_t>(r)]; k < pattern.row_ptr[static_cast<size_t>(r) + 1]; ++k) {
            const int c = pattern.col_index[static_cast<size_t>(k)];
            if (c == r) continue;
            ++a.offset[static_cast<size_t>(r) + 1];
            ++a.offset[static_cast<size_t>(c) + 1];
        }
    for (int r = 0; r < n; ++r) a.offset[static_cast<size_t>(r) + 1] += a.offset[static_cast<size_t>(r)];
    std::vector<int> raw(static_cast<size_t>(a.offset.back()));
    std::vector<int> cur(a.offset.begin(), a.offset.end() - 1);
    for (int r = 0; r < n; ++r)
        for (int k = pattern.row_ptr[static_cast<size_t>(r)]; k < pattern.row_ptr[static_cast<size_t>(r) + 1]; ++k) {
            const int c = pattern.col_index[static_cast<size_t>(k)];
            if (c == r) continue;
            raw[static_cast<size_t>(cur[static_cast<size_t>(r)]++)] = c;
            raw[static_cast<size_t>(cur[static_cast<size_t>(c)]++)] = r;
        }
    std::vector<int> off(static_cast<size_t>(n) + 1, 0);
    a.neighbor.reserve(raw.size());
    for (int r = 0; r < n; ++r) {
        auto b = raw.begin() + a.offset[static_cast<size_t>(r)];
        auto e = raw.begin() + a.offset[static_cast<size_t>(r) + 1];
        std::sort(b, e);
        auto last = std::unique(b, e);
        off[static_cast<size_t>(r)] = static_cast<int>(a.neighbor.size());
        a.neighbor.insert(a.neighbor.end(), b, last);
    }
    off[static_cast<size_t>(n)] = static_cast<int>(a.neighbor.size());
    a.offset = std::move(off);
    return a;
}

namespace {

// ---------------------------------------------------------------- AMD

class MinimumDegree {
public:
    MinimumDegree(const SymmetricAdjacency& adj, int n)
        : n_(n), kind_(static_cast<size_t>(n), kVar), vars_(static_cast<size_t>(n)),
          elems_(static_cast<size_t>(n)), members_(static_cast<size_t>(n)), degree_(static_cast<size_t>(n)),
          weight_(static_cast<size_t>(n), 1), absorbed_(static_cast<size_t>(n)),
          elem_weight_(static_cast<size_t>(n), 0), bucket_head_(static_cast<size_t>(n) + 1, -1),
          bucket_next_(static_cast<size_t>(n), -1), bucket_prev_(static_cast<size_t>(n), -1),
          stamp_(static_cast<size_t>(n), -1), external_(static_cast<size_t>(n), 0),
          chain_head_(static_cast<size_t>(n), -1), chain_next_(static_cast<size_t>(n), -1),
          mark_(static_cast<size_t>(n), -1) {
        for (int i = 0; i < n; ++i) {
            auto& v = vars_[static_cast<size_t>(i)];
            v.assign(adj.neighbor.begin() + adj.offset[static_cast<size_t>(i)],
                     adj.neighbor.begin() + adj.offset[static_cast<size_t>(i) + 1]);
            degree_[static_cast<size_t>(i)] = static_cast<int>(v.size());
        }
        for (int i = 0; i < n; ++i) push(i);
    }

    std::vector<int> run() {
        std::vector<int> order;
        order.reserve(static_cast<size_t>(n_));
        int min_deg = 0, done = 0, step = 0;
        std::vector<int> lp, touched, used;
        while (done < n_) {
            while (min_deg <= n_ && bucket_head_[static_cast<size_t>(min_deg)] < 0) ++min_deg;
            if (min_deg > n_) break;
            const int p = bucket_head_[static_cast<size_t>(min_deg)];
            pop(p);
            order.push_back(p);
            for (int q : absorbed_[static_cast<size_t>(p)]) order.push_back(q);
            done += weight_[static_cast<size_t>(p)];
            const int k = step++;

            // L_p, principal variables only: A_p first, then E_p's members.
            lp.clear();
            for (int j : vars_[static_cast<size_t>(p)]) {
                if (kind_[static_cast<size_t>(j)] != kVar || stamp_[static_cast<size_t>(j)] == k) continue;
                stamp_[static_cast<size_t>(j)] = k;
                lp.push_back(j);
            }
            for (int e : elems_[static_cast<size_t>(p)]) {
                if (kind_[static_cast<size_t>(e)] != kElem) continue;
                for (int j : members_[static_cast<size_t>(e)]) {
                    if (kind_[static_cast<size_t>(j)] != kVar || j == p || stamp_[static_cast<size_t>(j)] == k) continue;
                    stamp_[static_cast<size_t>(j)] = k;
                    lp.push_back(j);
                }
            }
            // E_p is absorbed into the new element p.
            for (int e : elems_[static_cast<size_t>(p)]) {
                if (kind_[static_cast<size_t>(e)] != kElem) continue;
                kind_[static_cast<size_t>(e)] = kGone;
                release(members_[static_cast<size_t>(e)]);
            }
            // Prune A_i and E_i of every i in L_p; E_i gains p.
            for (int i : lp) {
                auto& ai = vars_[static_cast<size_t>(i)];
                size_t w = 0;
                for (int j : ai) {
                    if (j == p || kind_[static_cast<size_t>(j)] != kVar || stamp_[static_cast<size_t>(j)] == k) continue;
                    ai[w++] = j;
                }
                ai.resize(w);
                auto& ei = elems_[static_cast<size_t>(i)];
                w = 0;
                for (int e : ei)
                    if (kind_[static_cast<size_t>(e)] == kElem) ei[w++] = e;
                ei.resize(w);
                ei.push_back(p);
            }
            // |L_e \ L_p|, weighted, for each element adjacent to L_p.
            touched.clear();
            const int estamp = -2 - k;
            for (int i : lp) {
                for (int e : elems_[static_cast<size_t>(i)]) {
                    if (e == p) continue;
                    if (stamp_[static_cast<size_t>(e)] != estamp) {
                        stamp_[static_cast<size_t>(e)] = estamp;
                        external_[static_cast<size_t>(e)] = elem_weight_[static_cast<size_t>(e)];
                        touched.push_back(e);
                    }
                    external_[static_cast<size_t>(e)] -= weight_[static_cast<size_t>(i)];
                }
            }
            long long lp_weight = 0;
            for (int i : lp) lp_weight += weight_[static_cast<size_t>(i)];
            for (int i : lp) {
                long long ext = 0;
                for (int e : elems_[static_cast<size_t>(i)])
                    if (e != p) ext += std::max(0, external_[static_cast<size_t>(e)]);
                long long aw = 0;
                for (int j : vars_[static_cast<size_t>(i)])
                    if (kind_[static_cast<size_t>(j)] == kVar) aw += weight_[static_cast<size_t>(j)];
                const long long wi = weight_[static_cast<size_t>(i)];
                const long long approx = aw + (lp_weight - wi) + ext;
                const long long remaining = static_cast<long long>(n_) - done;
                const long long grew = static_cast<long long>(degree_[static_cast<size_t>(i)]) + lp_weight - wi;
                const long long d = std::min(remaining, std::min(grew, approx));
                pop(i);
                degree_[static_cast<size_t>(i)] = static_cast<int>(std::max(0LL, d));
                push(i);
                min_deg = std::min(min_deg, degree_[static_cast<size_t>(i)]);
            }
            for (int e : touched) stamp_[static_cast<size_t>(e)] = -1;

            kind_[static_cast<size_t>(p)] = kElem;
            members_[static_cast<size_t>(p)] = lp;
            elem_weight_[static_cast<size_t>(p)] = static_cast<int>(lp_weight);
            release(vars_[static_cast<size_t>(p)]);
            release(elems_[static_cast<size_t>(p)]);

            // Supervariables within L_p.
            used.clear();
            for (int i : lp) {
                if (kind_[static_cast<size_t>(i)] != kVar) continue;
                long long h = 0;
                for (int j : vars_[static_cast<size_t>(i)]) h += j;
                for (int e : elems_[static_cast<size_t>(i)]) h += e;
                h %= n_;
                const size_t hb = static_cast<size_t>(h);
                if (chain_head_[hb] < 0) used.push_back(static_cast<int>(h));
                chain_next_[static_cast<size_t>(i)] = chain_head_[hb];
                chain_head_[hb] = i;
            }
            for (int h : used) {
                for (int i = chain_head_[static_cast<size_t>(h)]; i >= 0; i = chain_next_[static_cast<size_t>(i)]) {
                    if (kind_[static_cast<size_t>(i)] != kVar) continue;
                    for (int j : vars_[static_cast<size_t>(i)]) mark_[static_cast<size_t>(j)] = i;
                    for (int e : elems_[static_cast<size_t>(i)]) mark_[static_cast<size_t>(e)] = i;
                    for (int j = chain_next_[static_cast<size_t>(i)]; j >= 0; j = chain_next_[static_cast<size_t>(j)]) {
                        if (kind_[static_cast<size_t>(j)] != kVar) continue;
                        if (vars_[static_cast<size_t>(j)].size() != vars_[static_cast<size_t>(i)].size() ||
                            elems_[static_cast<size_t>(j)].size() != elems_[static_cast<size_t>(i)].size())
                            continue;
                        bool same = true;
                        for (int q : vars_[static_cast<size_t>(j)])
                            if (mark_[static_cast<size_t>(q)] != i) { same = false; break; }
                        if (same)
                            for (int e : elems_[static_cast<size_t>(j)])
                                if (mark_[static_cast<size_t>(e)] != i) { same = false; break; }
                        if (!same) continue;
                        pop(j);
                        weight_[static_cast<size_t>(i)] += weight_[static_cast<size_t>(j)];
                        auto& ab = absorbed_[static_cast<size_t>(i)];
                        ab.push_back(j);
                        ab.insert(ab.end(), absorbed_[static_cast<size_t>(j)].begin(), absorbed_[static_cast<size_t>(j)].end());
                        release(absorbed_[static_cast<size_t>(j)]);
                        weight_[static_cast<size_t>(j)] = 0;
                        kind_[static_cast<size_t>(j)] = kGone;
                        release(vars_[static_cast<size_t>(j)]);
                        release(elems_[static_cast<size_t>(j)]);
                    }
                }
                chain_head_[static_cast<size_t>(h)] = -1;
            }
        }
        if (static_cast<int>(order.size()) != n_)
            throw std::logic_error("approximate_minimum_degree: ordered " + std::to_string(order.size()) + " of " +
                                   std::to_string(n_) + " unknowns");
        return order;
    }

private:
    enum : int8_t { kVar = 0, kElem = 1, kGone = 2 };

    static void release(std::vector<int>& v) { std::vector<int>().swap(v); }

    // LIFO doubly linked degree buckets.
    void push(int i) {
        const size_t d = static_cast<size_t>(std::min(degree_[static_cast<size_t>(i)], n_));
        bucket_next_[static_cast<size_t>(i)] = bucket_head_[d];
        bucket_prev_[static_cast<size_t>(i)] = -1;
        if (bucket_head_[d] >= 0) bucket_prev_[static_cast<size_t>(bucket_head_[d])] = i;
        bucket_head_[d] = i;
    }
    void pop(int i) {
        const size_t d = static_cast<size_t>(std::min(degree_[static_cast<size_t>(i)], n_));
        const int pv = bucket_prev_[static_cast<size_t>(i)], nx = bucket_next_[static_cast<size_t>(i)];
        if (pv >= 0) bucket_next_[static_cast<size_t>(pv)] = nx;
        else if (bucket_head_[d] == i) bucket_head_[d] = nx;
        if (nx >= 0) bucket_prev_[static_cast<size_t>(nx)] = pv;
        bucket_prev_[static_cast<size_t>(i)] = bucket_next_[static_cast<size_t>(i)] = -1;
    }

    int n_;
    std::vector<int8_t> kind_;
    std::vector<std::vector<int>> vars_, elems_, members_;
    std::vector<int> degree_, weight_;
    std::vector<std::vector<int>> absorbed_;
    std::vector<int> elem_weight_;
    std::vector<int> bucket_head_, bucket_next_, bucket_prev_;
    std::vector<int> stamp_, external_;
    std::vector<int> chain_head_, chain_next_, mark_;
};

// ---------------------------------------------------------------- RCM
// src/ordering.cpp:113-179, 512-535: BFS with each vertex's new neighbours
// sorted by (degree, index); George-Liu pseudo-peripheral start, at most 5 passes.
int level_bfs(const SymmetricAdjacency& adj, int root, std::vector<uint8_t>& seen, std::vector<int>& order,
              std::vector<int>& level) {
    const size_t first = order.size();
    seen[static_cast<size_t>(root)] = 1;
    order.push_back(root);
    level[static_cast<size_t>(root)] = 0;
    int depth = 0;
    for (size_t h = first; h < order.size(); ++h) {
        const int v = order[h];
        const int lv = level[static_cast<size_t>(v)];
        const size_t b = order.size();
        for (int k = adj.offset[static_cast<size_t>(v)]; k < adj.offset[static_cast<size_t>(v) + 1]; ++k) {
            const int w = adj.neighbor[static_cast<size_t>(k)];
            if (seen[static_cast<size_t>(w)]) continue;
            seen[static_cast<size_t>(w)] = 1;
            level[static_cast<size_t>(w)] = lv + 1;
            depth = std::max(depth, lv + 1);
            order.push_back(w);
        }
        std::sort(order.begin() + static_cast<std::ptrdiff_t>(b), order.end(), [&](int x, int y) {
            const int dx = adj.degree(x), dy = adj.degree(y);
            return dx != dy ? dx < dy : x < y;
        });
    }
    return depth;
}

std::vector<int> reverse_cuthill_mckee(const SymmetricAdjacency& adj, int n) {
    std::vector<uint8_t> seen(static_cast<size_t>(n), 0), sseen(static_cast<size_t>(n), 0);
    std::vector<int> level(static_cast<size_t>(n), 0), slevel(static_cast<size_t>(n), 0), order, sorder;
    order.reserve(static_cast<size_t>(n));
    for (int s = 0; s < n; ++s) {
        if (seen[static_cast<size_t>(s)]) continue;
        int best = s, best_depth = -1;
        for (int pass = 0; pass < 5; ++pass) {
            std::fill(sseen.begin(), sseen.end(), 0);
            sorder.clear();
            const int depth = level_bfs(adj, best, sseen, sorder, slevel);
            if (depth <= best_depth) break;
            best_depth = depth;
            int cand = best, cand_deg = -1;
            for (int v : sorder) {
                if (slevel[static_cast<size_t>(v)] != depth) continue;
                const int d = adj.degree(v);
                if (cand_deg < 0 || d < cand_deg) { cand = v; cand_deg = d; }
            }
            if (cand == best) break;
            best = cand;
        }
        level_bfs(adj, best, seen, order, level);
    }
    std::reverse(order.begin(), order.end());
    return order;
}

}  // namespace

Permutation compute_ordering(const SparsityPattern& pattern, Ordering ordering) {
    if (pattern.rows != pattern.cols) throw std::invalid_argument("compute_ordering: the pattern must be square");
    Permutation p;
    if (ordering == Ordering::Natural) {
        p.perm.resize(static_cast<size_t>(pattern.rows));
        std::iota(p.perm.begin(), p.perm.end(), 0);
    } else {
        const SymmetricAdjacency adj = SymmetricAdjacency::build(pattern);
        p.perm = ordering == Ordering::ApproximateMinimumDegree ? MinimumDegree(adj, pattern.rows).run()
                                                                : reverse_cuthill_mckee(adj, pattern.rows);
    }
    p.iperm.assign(p.perm.size(), 0);
    for (size_t i = 0; i < p.perm.size(); ++i) p.iperm[static_cast<size_t>(p.perm[i])] = static_cast<int>(i);
    return p;
}

SparsityPattern permute_pattern(const SparsityPattern& pattern, const Permutation& p) {
    if (p.size() != pattern.rows)
        throw std::invalid_argument("permute_pattern: the permutation is sized " + std::to_string(p.size()) +
                                    " but the pattern has " + std::to_string(pattern.rows) + " rows");
    SparsityPattern out;
    out.rows = pattern.rows;
    out.cols = pattern.cols;
    out.upper_only = pattern.upper_only;
    const size_t n = static_cast<size_t>(out.rows);
    out.row_ptr.assign(n + 1, 0);
    const auto place = [&](int r, int c, int& row, int& col) {
        row = p.iperm[static_cast<size_t>(r)];
        col = p.iperm[static_cast<size_t>(c)];
        if (out.upper_only && col < row) std::swap(row, col);
    };
    for (int r = 0; r < pattern.rows; ++r)
        for (int k = pattern.row_ptr[static_cast<size_t>(r)]; k < pattern.row_ptr[static_cast<size_t>(r) + 1]; ++k) {
            int row, col;
            place(r, pattern.col_index[static_cast<size_t>(k)], row, col);
            ++out.row_ptr[static_cast<size_t>(row) + 1];
        }
    for (size_t r = 0; r < n; ++r) out.row_ptr[r + 1] += out.row_ptr[r];
    std::vector<int> cur(out.row_ptr.begin(), out.row_ptr.end() - 1);
    out.col_index.assign(static_cast<size_t>(out.row_ptr.back()), 0);
    for (int r = 0; r < pattern.rows; ++r)
        for (int k = pattern.row_ptr[static_cast<size_t>(r)]; k < pattern.row_ptr[static_cast<size_t>(r) + 1]; ++k) {
            int row, col;
            place(r, pattern.col_index[static_cast<size_t>(k)], row, col);
            out.col_index[static_cast<size_t>(cur[static_cast<size_t>(row)]++)] = col;
        }
    for (size_t r = 0; r < n; ++r)
        std::sort(out.col_index.begin() + out.row_ptr[r], out.col_index.begin() + out.row_ptr[r + 1]);
    return out;
}

}  // namespace aphi_amd

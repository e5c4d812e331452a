// Supernodal multifrontal complex-symmetric LDL^T (no pivoting) on gfx950.
//
// Replaces the reference's scalar up-looking factorize_ldlt
// (src/factorization.cpp:109-218) with the same mathematics organised by
// fronts: P A P^T = L D L^T, L unit lower, D complex, nothing conjugated.
//
// Storage (per supernode s with k columns and m rows):
//   panel  P[panel_off[s] + c*m + r]       m x k column-major: L below the
//                                          diagonal, D on it
//   update CB[cb_off[s] + c*(m-k) + r]     (m-k)^2 column-major, lower half used
// A front is the m x m lower matrix whose first k columns live in the panel and
// the rest in the update block.
//
// Level-synchronous schedule: fronts of one elimination-tree height run
// together. Small fronts (m <= small_max) are assembled, factored and written
// back by one workgroup entirely in LDS. Large fronts use a blocked right-looking
// loop: extend-add, then per NB-column block a diagonal factorization, a panel
// multiply by (L_bb^-T D_bb^-1), and an MFMA trailing update of the remaining
// front (the only O(m^2 k) step).
#include <algorithm>
#include <climits>
#include <stdexcept>

#include "common.hpp"
#include "factor.hpp"

namespace aphi_dev {

namespace {

template <typename T>
using C = typename cplx<T>::type;

template <typename T>
__device__ __forceinline__ C<T> to_c(cd v) {
    return C<T>{static_cast<T>(v.re), static_cast<T>(v.im)};
}

// Address of front element (i, j), i >= j, of supernode s.
template <typename T>
struct FrontRef {
    C<T>* panel;
    C<T>* cb;
    int m, k;
    __device__ __forceinline__ C<T>* at(int i, int j) const {
        return j < k ? panel + static_cast<int64_t>(j) * m + i : cb + static_cast<int64_t>(j - k) * (m - k) + (i - k);
    }
};

template <typename T>
__device__ __forceinline__ FrontRef<T> front(const FrontPlan& fp, int s, C<T>* P, C<T>* CB) {
    FrontRef<T> f;
    f.m = static_cast<int>(fp.row_ptr[s + 1] - fp.row_ptr[s]);
    f.k = fp.first[s + 1] - fp.first[s];
    f.panel = P + fp.panel_off[s];
    f.cb = f.m > f.k ? CB + fp.cb_off[s] : nullptr;
    return f;
}

__device__ __forceinline__ void stat_min(double* a, double v) {
    atomicMin(reinterpret_cast<unsigned long long*>(a), static_cast<unsigned long long>(__double_as_longlong(v)));
}

// The statistics slot of the calling wave: neighbouring workgroups and the waves
// of one workgroup land on different slots.
__device__ __forceinline__ StatSlot* stat_slot(FactorStatsDev* st) {
    const unsigned k = blockIdx.x * 37u + blockIdx.y * 101u + threadIdx.x / 64u;
    return &st->slot[k % kStatSlots];
}

// One lane's running pivot statistics: |D| extremes and the smallest reference
// column whose pivot is not above the floor (src/factorization.cpp:191-210).
struct PivotAcc {
    double mn = __builtin_inf();
    double mx = 0.0;
    int fail = INT_MAX;
    // col: factor-order column; its reference column is looked up only on failure
    __device__ __forceinline__ void add(double mag, const int* col_ref, int col, double floor_) {
        mn = fmin(mn, mag);
        mx = fmax(mx, mag);
        if (!(mag > floor_)) fail = min(fail, col_ref[col]);
    }
};

template <typename T>
__device__ __forceinline__ double cmag(C<T> d) {
    return hypot(static_cast<double>(d.re), static_cast<double>(d.im));
}

// Reduces the lanes' pivot statistics across the (whole, converged) wave, then
// one atomic per statistic into the wave's slot.
__device__ __forceinline__ void flush_pivots(FactorStatsDev* st, PivotAcc a) {
    for (int off = 32; off > 0; off >>= 1) {
        a.mn = fmin(a.mn, __shfl_xor(a.mn, off));
        a.mx = fmax(a.mx, __shfl_xor(a.mx, off));
        a.fail = min(a.fail, __shfl_xor(a.fail, off));
    }
    if ((threadIdx.x & 63) == 0) {
        StatSlot* sl = stat_slot(st);
        if (a.mn < __builtin_inf()) stat_min(&sl->smallest_pivot, a.mn);
        if (a.mx > 0.0) atomic_max_nonneg(&sl->largest_pivot, a.mx);
        if (a.fail != INT_MAX) atomicMin(&sl->fail_col, a.fail);
    }
}

// Block-wide max of a non-negative value into the stats: a wave64 reduction, then
// one atomic per wave into its slot.
__device__ __forceinline__ void block_max_multiplier(FactorStatsDev* st, double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0 && v > 0.0) atomic_max_nonneg(&stat_slot(st)->largest_multiplier, v);
}

}  // namespace

// ------------------------------------------------------------------ setup

template <typename T>
__global__ void k_zero(int64_t n, C<T>* __restrict__ p) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = C<T>{0, 0};
}

__global__ void k_stats_init(FactorStatsDev* st) {
    const int i = threadIdx.x;
    if (i < kStatSlots) st->slot[i] = StatSlot{__builtin_inf(), 0.0, 0.0, INT_MAX, 0};
}

// The equilibrated A, scattered into the panels (each entry lands in the front
// of the supernode owning its lower column; no two entries share a position).
template <typename T>
__global__ void k_scatter_panel(int64_t nnz, const int64_t* __restrict__ s2p, const cd* __restrict__ a,
                                C<T>* __restrict__ P) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < nnz) P[s2p[k]] = to_c<T>(a[k]);
}

// Statistics of a factorization stopped by a pivot at or below the floor. The
// reference's up-looking loop (src/factorization.cpp:124-210) stops after row
// `fail` of its own order: its extremes cover the pivots of reference columns
// 0..fail and the multipliers l_kj of rows k <= fail. Every one of those entries
// depends only on columns before `fail` in that order (descendants come first),
// so the device's values of them are the finished ones even though it factored
// on past the failure. A flat grid over the panels (P[panel_off[s] + c*m + r]:
// r == c the pivot, r > c a multiplier of row rows[row_ptr[s] + r]) gathers
// the extremes over exactly those entries into freshly initialised slots.
template <typename T>
__global__ void k_failure_stats(int ns, FrontPlan fp, const C<T>* __restrict__ P, int fail,
                                FactorStatsDev* st) {
    PivotAcc acc;
    double mult = 0.0;
    const int64_t total = fp.panel_off[ns];
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += stride) {
        int lo = 0, hi = ns - 1;  // the supernode s with panel_off[s] <= e < panel_off[s + 1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (fp.panel_off[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        const int k = fp.first[lo + 1] - fp.first[lo];
        const int64_t m = fp.row_ptr[lo + 1] - fp.row_ptr[lo];
        const int64_t off = e - fp.panel_off[lo];
        if (off >= m * k) continue;
        const int c = static_cast<int>(off / m);
        const int64_t r = off - static_cast<int64_t>(c) * m;
        if (r < c) continue;  // above the diagonal: unused
        const int row = fp.rows[fp.row_ptr[lo] + r];
        if (fp.col_ref[row] > fail) continue;
        const double mag = cmag<T>(P[e]);
        if (r == c) {
            acc.mn = fmin(acc.mn, mag);
            acc.mx = fmax(acc.mx, mag);
        } else {
            mult = fmax(mult, mag);
        }
    }
    flush_pivots(st, acc);
    block_max_multiplier(st, mult);
}

// ------------------------------------------------------------------ small fronts

// One workgroup per front: assemble (panel + children's update blocks, children
// in order), factor all k columns, write the panel and the update block back.
template <typename T>
__global__ void __launch_bounds__(256) k_front_small(const int* __restrict__ nodes, FrontPlan fp, C<T>* P, C<T>* CB,
                                                     C<T>* __restrict__ xinv, double pivot_floor, FactorStatsDev* st) {
    extern __shared__ unsigned char smem_raw[];
    C<T>* F = reinterpret_cast<C<T>*>(smem_raw);
    const int s = nodes[blockIdx.x];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    const int m = fr.m, k = fr.k;
    C<T>* lcol = F + m * m;
    const int tid = threadIdx.x, nt = blockDim.x;

    for (int idx = tid; idx < m * m; idx += nt) {
        const int j = idx / m, i = idx % m;
        F[idx] = (j < k) ? fr.panel[idx] : C<T>{0, 0};
        (void)i;
    }
    for (int c = fp.child_ptr[s]; c < fp.child_ptr[s + 1]; ++c) {
        __syncthreads();
        const int ch = fp.child[c];
        const int cm = static_cast<int>(fp.row_ptr[ch + 1] - fp.row_ptr[ch]);
        const int ck = fp.first[ch + 1] - fp.first[ch];
        const int cu = cm - ck;
        const C<T>* cb = CB + fp.cb_off[ch];
        const int* rel = fp.rel + fp.row_ptr[ch] + ck;
        for (int idx = tid; idx < cu * cu; idx += nt) {
            const int j = idx / cu, i = idx % cu;
            if (i < j) continue;
            C<T>& dst = F[rel[j] * m + rel[i]];
            dst = gadd(dst, cb[idx]);
        }
    }
    __syncthreads();

    double lmax = 0.0;
    PivotAcc piv;
    for (int j = 0; j < k; ++j) {
        const C<T> d = F[j * m + j];
        if (tid == 0) piv.add(cmag<T>(d), fp.col_ref, fp.first[s] + j, pivot_floor);
        const C<T> rd = grecip(d);
        for (int i = j + 1 + tid; i < m; i += nt) {
            const C<T> l = gmul(F[j * m + i], rd);
            lcol[i] = l;
            lmax = fmax(lmax, static_cast<double>(hypot(l.re, l.im)));
        }
        __syncthreads();
        // F[i][c] -= F[i][j] * l_c for j < c <= i
        const int w = m - j - 1;
        for (int idx = tid; idx < w * w; idx += nt) {
            const int c = j + 1 + idx / w, i = j + 1 + idx % w;
            if (i < c) continue;
            F[c * m + i] = gfms(F[c * m + i], F[j * m + i], lcol[c]);
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < m; i += nt) F[j * m + i] = lcol[i];
        __syncthreads();
    }
    if (tid < 64) flush_pivots(st, piv);
    block_max_multiplier(st, lmax);

    // Inverses of the unit-lower diagonal blocks (kNB columns each), built in the
    // unused strict upper triangle of F: Linv[r][c] (r > c) lives at F[r*m + c].
    for (int bb = 0; bb < k; bb += kNB) {
        const int nbx = min(kNB, k - bb);
        if (tid < nbx) {
            const int c = tid;
            for (int r = c + 1; r < nbx; ++r) {
                C<T> acc = F[(bb + c) * m + (bb + r)];  // q = c term: L[r][c] * 1
                for (int q = c + 1; q < r; ++q)
                    acc = gfma(acc, F[(bb + q) * m + (bb + r)], F[(bb + q) * m + (bb + c)]);
                F[(bb + r) * m + (bb + c)] = C<T>{-acc.re, -acc.im};
            }
        }
        __syncthreads();
        C<T>* Xg = xinv + fp.xinv_off[s] + static_cast<int64_t>(bb / kNB) * kNB * kNB;
        for (int idx = tid; idx < nbx * nbx; idx += nt) {
            const int q = idx / nbx, c = idx % nbx;  // column q, row c
            Xg[idx] = c > q ? F[(bb + c) * m + (bb + q)] : C<T>{static_cast<T>(c == q ? 1 : 0), 0};
        }
    }

    for (int idx = tid; idx < m * k; idx += nt) fr.panel[idx] = (idx % m) < (idx / m) ? C<T>{0, 0} : F[idx];
    const int u = m - k;
    for (int idx = tid; idx < u * u; idx += nt) {
        const int j = idx / u, i = idx % u;
        fr.cb[idx] = F[(k + j) * m + (k + i)];
    }
}

// ------------------------------------------------------------------ large fronts

// Extend-add for large fronts. Workgroup = (front, block of 64 front columns,
// chunk of ea_rows<T>() front rows); it owns those entries exclusively. It maps
// the chunk's rows and columns to up to kEaChildren children at a time (LDS, -1
// where the child has none; the child index ranges come from the precomputed
// bounds), then takes the columns ea_entries<T>() / (ea_rows<T>() / 64) per wave
// at a time with the entries in registers: wave w holds columns c0 + w + 4t
// (t < 16), lane l rows r0 + l + 64u. In child order every entry adds its child
// entry if there is one. Each update-block
// entry is then written once (zero where no child contributes) and each touched
// pivot-column entry becomes P + (sum of its children's entries): one write per
// entry, deterministic without atomics.
template <typename T>
__global__ void __launch_bounds__(256) k_extend_add(const int* __restrict__ nodes, const int* __restrict__ blk_start,
                                                    int nnodes, FrontPlan fp, C<T>* P, C<T>* CB) {
    constexpr int R = ea_rows<T>();
    constexpr int CT = 16;  // columns per wave
    constexpr int NCH = kEaChildren;
    __shared__ int invc[NCH][64];
    __shared__ int invr[NCH][R];
    __shared__ int cinfo[NCH][5];     // child columns [a, e) and rows [ia, ie) in the chunk, update rows
    __shared__ int64_t coff[NCH][2];  // offsets of the child's update-row indices and update block
    const int tid = threadIdx.x, lane = tid % 64;
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int bx = static_cast<int>(blockIdx.x);
    // the front of this block: 64-way search, blk_start[lo] <= bx < blk_start[hi]
    int lo = 0, hi = nnodes;
    while (hi - lo > 1) {
        const int step = (hi - lo + 63) / 64;
        const int idx = lo + lane * step;
        const unsigned long long le = __ballot(idx < hi && blk_start[idx] <= bx);
        lo += (63 - __builtin_clzll(le)) * step;  // lane 0 always qualifies
        hi = min(hi, lo + step);
    }
    const int s = nodes[lo];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    const int cblk = bx - blk_start[lo];
    const int c0 = cblk * 64;
    const int c1 = min(c0 + 64, fr.m);
    const int r0 = max(c0, static_cast<int>(blockIdx.y) * R);
    const int r1 = min(fr.m, (static_cast<int>(blockIdx.y) + 1) * R);
    if (r0 >= r1) return;
    const int q0 = r0 / 64, q1 = (r1 + 63) / 64;  // row range in 64-blocks (r0, r1 multiples of 64 or m)
    const int cfirst = fp.child_ptr[s], clast = fp.child_ptr[s + 1];

    // maps of children [cc, cc + nc) for the whole chunk
    auto build_maps = [&](int cc, int nc) {
        __syncthreads();  // the previous maps are no longer read
        for (int i = tid; i < NCH * 64; i += 256) (&invc[0][0])[i] = -1;
        for (int i = tid; i < NCH * R; i += 256) (&invr[0][0])[i] = -1;
        if (tid < nc) {
            const int ch = fp.child[cc + tid];
            const int* bnd = fp.ea_bound + fp.ea_bound_off[ch];
            const int ck = fp.first[ch + 1] - fp.first[ch];
            cinfo[tid][0] = bnd[cblk];
            cinfo[tid][1] = bnd[cblk + 1];
            cinfo[tid][2] = bnd[q0];
            cinfo[tid][3] = bnd[q1];
            cinfo[tid][4] = static_cast<int>(fp.row_ptr[ch + 1] - fp.row_ptr[ch]) - ck;
            coff[tid][0] = fp.row_ptr[ch] + ck;
            coff[tid][1] = fp.cb_off[ch];
        }
        __syncthreads();
        for (int i = tid; i < nc * (64 + R); i += 256) {
            const int c = i / (64 + R), r = i % (64 + R);
            const int* rel = fp.rel + coff[c][0];
            if (r < 64) {
                const int jc = cinfo[c][0] + r;
                if (jc < cinfo[c][1]) invc[c][rel[jc] - c0] = jc;
            } else {
                const int ic = cinfo[c][2] + (r - 64);
                if (ic < cinfo[c][3]) invr[c][rel[ic] - r0] = ic;
            }
        }
        __syncthreads();
    };
    const bool one_map = clast - cfirst <= NCH;
    if (one_map && clast > cfirst) build_maps(cfirst, clast - cfirst);

    // passes of CP columns per wave; a pass reads and writes whole R-row column
    // segments (RL rows per lane), EA entries per lane
    constexpr int EA = ea_entries<T>();
    constexpr int RL = R / 64, CP = EA / RL;
    static_assert(RL * CP == EA && CT % CP == 0 && EA <= 32, "EA entries per lane and pass");
    for (int q = 0; q < CT / CP; ++q) {
        C<T> acc[CP][RL];
        uint32_t touched = 0;  // bit g * RL + u: a child contributed to (column g, row u)
#pragma unroll
        for (int g = 0; g < CP; ++g)
#pragma unroll
            for (int u = 0; u < RL; ++u) acc[g][u] = C<T>{0, 0};
        for (int cc = cfirst; cc < clast; cc += NCH) {
            const int nc = min(NCH, clast - cc);
            if (!one_map) build_maps(cc, nc);
            for (int c = 0; c < nc; ++c) {
                if (cinfo[c][0] >= cinfo[c][1] || cinfo[c][2] >= cinfo[c][3]) continue;  // nothing in this chunk
                const int cu = __builtin_amdgcn_readfirstlane(cinfo[c][4]);
                const C<T>* cb = CB + coff[c][1];
                int cr[RL];
#pragma unroll
                for (int u = 0; u < RL; ++u) cr[u] = invr[c][lane + 64 * u];
                C<T> v[CP][RL];
                bool ok[CP][RL];
#pragma unroll
                for (int g = 0; g < CP; ++g) {
                    const int cj = __builtin_amdgcn_readfirstlane(invc[c][wave + 4 * (CP * q + g)]);
                    const C<T>* col = cb + static_cast<int64_t>(cj) * cu;
#pragma unroll
                    for (int u = 0; u < RL; ++u) {
                        // cr >= cj >= 0 puts the entry in the child's lower part (and the
                        // parent's: the relative indices increase)
                        ok[g][u] = cj >= 0 && cr[u] >= cj;
                        v[g][u] = C<T>{0, 0};
                        if (ok[g][u]) v[g][u] = col[cr[u]];
                    }
                }
#pragma unroll
                for (int g = 0; g < CP; ++g)
#pragma unroll
                    for (int u = 0; u < RL; ++u)
                        if (ok[g][u]) {
                            acc[g][u] = gadd(acc[g][u], v[g][u]);
                            touched |= 1u << (g * RL + u);
                        }
            }
        }
        // update-block entries (j >= k) always, pivot-column entries where touched
#pragma unroll
        for (int g = 0; g < CP; ++g) {
            const int j = c0 + wave + 4 * (CP * q + g);
            if (j >= c1) break;
#pragma unroll
            for (int u = 0; u < RL; ++u) {
                const int i = r0 + lane + 64 * u;
                if (i >= r1 || i < j) continue;
                C<T>* dst = fr.at(i, j);
                if (j >= fr.k) *dst = acc[g][u];
                else if (touched >> (g * RL + u) & 1u) *dst = gadd(*dst, acc[g][u]);
            }
        }
    }
}

// Diagonal block of step b for each front with k > b, one 1024-thread workgroup
// per front. Items: the 2080 lower-triangle entries of the 64 x 64 block (rows
// and columns beyond nb padded with the identity) and the 2016 strict-lower
// entries of X = L_bb^-1, four per thread, kept in registers. Step j does both
// eliminations with one barrier: A[i][c] -= (A[i][j]/A[j][j]) A[c][j] for
// j < c <= i, and X[i][c] -= L[i][j] X[j][c] for c < j < i (with X[i][j] set to
// -L[i][j] at step j), reading the pivot column of A and row j of X from LDS;
// their owners publish column j+1 and row j+1 into the other buffer of each
// pair. A column of L is final (and written out) at its own step. Then the
// inverse (kept for the solves) and M = L_bb^-T D_bb^-1 (for the panel step)
// are written out.
template <typename T>
__global__ void __launch_bounds__(1024) k_diag(const int* __restrict__ nodes, int b, FrontPlan fp, C<T>* P, C<T>* CB,
                                               C<T>* __restrict__ mws, C<T>* __restrict__ xinv, double pivot_floor,
                                               FactorStatsDev* st) {
    extern __shared__ unsigned char smem_raw[];
    constexpr int LD = kNB + 1;
    constexpr int NA = kNB * (kNB + 1) / 2;      // A items (lower triangle with the diagonal)
    constexpr int NX = kNB * (kNB - 1) / 2;      // X items (strict lower triangle)
    constexpr int NT = 1024;
    // items per thread: A items in the first NEA slots, X items in the next NEX, so
    // that every slot's kind is known at compile time (slots padded with empty items)
    constexpr int NEA = (NA + NT - 1) / NT, NEX = (NX + NT - 1) / NT, NE = NEA + NEX;
    C<T>* Xs = reinterpret_cast<C<T>*>(smem_raw); // Xs[c*LD + r] = X[r][c]
    C<T>* colA = Xs + kNB * LD;                   // two pivot-column buffers
    C<T>* rowX = colA + 2 * kNB;                  // two pivot-row buffers
    C<T>* Dl = rowX + 2 * kNB;
    const int s = nodes[blockIdx.x];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    if (fr.k <= b) return;
    const int nb = min(kNB, fr.k - b);
    const int tid = threadIdx.x;
    C<T>* blk = fr.panel + static_cast<int64_t>(b) * fr.m + b;  // element (b, b) of the front

    // item -> (row, column, is_x); columns of A start at c*kNB - c(c-1)/2, of X at c*(kNB-1) - c(c-1)/2
    int ir[NE], ic[NE];
    bool isx[NE];
    T vre[NE], vim[NE];
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        const bool x = u >= NEA;
        const int e = tid + NT * (x ? u - NEA : u);
        const int len = x ? kNB - 1 : kNB;       // entries of column 0
        const int total = x ? NX : NA;
        const float q = 2.0f * len + 1.0f;
        int c = static_cast<int>((q - sqrtf(q * q - 8.0f * static_cast<float>(min(e, total - 1)))) * 0.5f);
        c = max(0, min(c, len - 1));
        while (c > 0 && c * len - c * (c - 1) / 2 > e) --c;
        while (c + 1 < len && (c + 1) * len - (c + 1) * c / 2 <= e) ++c;
        const bool valid = e < total;
        const int r = c + (x ? 1 : 0) + (e - (c * len - c * (c - 1) / 2));
        ir[u] = valid ? r : 0;
        ic[u] = valid ? c : kNB + 1;  // past every step, never published
        isx[u] = x;
    }
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        const int r = ir[u], c = ic[u];
        const bool load = !isx[u] && c < nb && r < nb;
        const C<T> v = blk[static_cast<int64_t>(load ? c : 0) * fr.m + (load ? r : 0)];
        vre[u] = load ? v.re : static_cast<T>(!isx[u] && r == c ? 1 : 0);
        vim[u] = load ? v.im : static_cast<T>(0);
        if (!isx[u] && c == 0) colA[r] = C<T>{vre[u], vim[u]};
    }
    __syncthreads();
    // Steps at which each item changes: A[r][c] at j <= c (updated for j < c,
    // turned into L[r][c] at j = c), X[r][c] at c <= j < r. Each wave skips an item
    // slot outside the union of its lanes' ranges (a wave-uniform branch).
    int lo[NE], hi[NE];
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        int l0 = isx[u] ? ic[u] : 0, h0 = isx[u] ? ir[u] - 1 : ic[u];
        if (ic[u] > kNB) {  // an empty slot: no step
            l0 = kNB;
            h0 = -1;
        }
        for (int off = 32; off > 0; off >>= 1) {
            l0 = min(l0, __shfl_xor(l0, off));
            h0 = max(h0, __shfl_xor(h0, off));
        }
        lo[u] = __builtin_amdgcn_readfirstlane(l0);
        hi[u] = __builtin_amdgcn_readfirstlane(h0);
    }
    // steps j >= nb only meet the identity padding: every entry read afterwards
    // (L, D for r, c < nb; X for r < nb) is final after step nb - 1
    for (int j = 0; j < nb; ++j) {
        const C<T>* cj = colA + (j & 1) * kNB;
        C<T>* cn = colA + ((j + 1) & 1) * kNB;
        const C<T>* rj = rowX + (j & 1) * kNB;
        C<T>* rn = rowX + ((j + 1) & 1) * kNB;
        const C<T> rd = grecip(cj[j]);
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            if (j < lo[u] || j > hi[u]) continue;
            const int r = ir[u], c = min(ic[u], kNB - 1);
            const C<T> ar = cj[r];
            const C<T> ac = isx[u] ? rj[c] : cj[c];
            const C<T> lr = gmul(ar, rd);  // L[r][j]
            const C<T> cur = C<T>{vre[u], vim[u]};
            const C<T> upd = gfms(cur, lr, ac);
            // A: update for c > j; column j becomes L below the diagonal (D stays on it).
            // X: X[r][j] = -L[r][j]; update for c < j < r.
            const bool do_upd = isx[u] ? (c < j && r > j) : (c > j);
            const bool do_set = isx[u] ? (c == j) : (c == j && r > j);
            const T sgn = isx[u] ? T(-1) : T(1);
            const T nre = do_upd ? upd.re : (do_set ? sgn * lr.re : cur.re);
            const T nim = do_upd ? upd.im : (do_set ? sgn * lr.im : cur.im);
            vre[u] = nre;
            vim[u] = nim;
            const bool pub = isx[u] ? (r == j + 1) : (ic[u] == j + 1);
            if (pub) (isx[u] ? rn[c] : cn[r]) = C<T>{nre, nim};
        }
        __syncthreads();
    }
    // L and D out to the front (stores kept out of the loop so no step waits on
    // them), X into LDS (column c, row r at Xs[c*LD + r]; diagonal 1, upper unused)
    double lmax = 0.0;
    for (int idx = tid; idx < kNB; idx += NT) Xs[idx * LD + idx] = C<T>{1, 0};
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        const int r = ir[u], c = ic[u];
        const C<T> v = C<T>{vre[u], vim[u]};
        if (isx[u]) {
            if (c < kNB) Xs[c * LD + r] = v;
        } else {
            if (r == c) Dl[r] = v;
            if (r > c && r < nb) lmax = fmax(lmax, static_cast<double>(hypot(v.re, v.im)));
            if (r < nb && c < nb) blk[static_cast<int64_t>(c) * fr.m + r] = v;
        }
    }
    __syncthreads();
    if (tid < 64) {
        PivotAcc piv;
        if (tid < nb) piv.add(cmag<T>(Dl[tid]), fp.col_ref, fp.first[s] + b + tid, pivot_floor);
        flush_pivots(st, piv);
    }
    block_max_multiplier(st, lmax);
    // M[q][c] = Linv[c][q] / D[c] (c >= q) at mws[q*kNB + c]; the inverse, column q
    // and row c at q*nb + c, is kept for the solves
    C<T>* M = mws + static_cast<int64_t>(blockIdx.x) * kNB * kNB;
    C<T>* Xg = xinv + fp.xinv_off[s] + static_cast<int64_t>(b / kNB) * kNB * kNB;
    for (int idx = tid; idx < nb * nb; idx += NT) {
        const int q = idx / nb, c = idx % nb;
        const C<T> v = c >= q ? Xs[q * LD + c] : C<T>{0, 0};
        Xg[idx] = v;
        M[q * kNB + c] = c >= q ? gmul(v, grecip(Dl[c])) : C<T>{0, 0};
    }
}

// Panel step: rows i in [b+nb, m) of block column b become L = F * M (M upper
// triangular, nb x nb, zero-padded to 64 x 64 here). One workgroup per (64-row
// block, front), wave w the 16 rows r0 + 16 w, on MFMA: 16 values of F per lane
// in registers, then the four 16-column tiles of M read from L2 one after the
// other. Each wave reads all of its rows before writing them, so the in-place
// write needs no staging, and the kernel uses no LDS: it fits beside the
// trailing-update workgroups that fill the device meanwhile.
template <typename T>
__global__ void __launch_bounds__(256) k_panel(const int* __restrict__ nodes, int b, FrontPlan fp, C<T>* P, C<T>* CB,
                                               const C<T>* __restrict__ mws, FactorStatsDev* st, C<T>* W) {
    using V4 = typename Mfma<T>::v4;
    const int s = nodes[blockIdx.y];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    if (fr.k <= b) return;
    const int nb = min(kNB, fr.k - b);
    const int rw = b + nb + static_cast<int>(blockIdx.x) * 64 + 16 * (static_cast<int>(threadIdx.x) / 64);
    const C<T>* M = mws + static_cast<int64_t>(blockIdx.y) * kNB * kNB;
    const int lane = threadIdx.x % 64, p = lane % 16, kq = lane / 16;
    double lmax = 0.0;
    if (rw < fr.m) {
        C<T> a[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = 4 * u + kq;
            a[u] = q < nb ? fr.panel[static_cast<int64_t>(b + q) * fr.m + min(rw + p, fr.m - 1)] : C<T>{0, 0};
        }

#pragma unroll 1
        for (int ct = 0; ct < kNB; ct += 16) {
            if (ct >= nb) break;
            // complex float: -D of the lane's column, loaded with this tile's M (the W
            // rows are -L D)
            C<T> dn = C<T>{0, 0};
            if constexpr (std::is_same<T, float>::value) {
                if (W != nullptr && ct + p < nb) {
                    const C<T> d = fr.panel[static_cast<int64_t>(b + ct + p) * fr.m + b + ct + p];
                    dn = C<T>{-d.re, -d.im};
                }
            }
            C<T> g[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int q = 4 * u + kq, c = ct + p;
                g[u] = (q < nb && c < nb) ? M[q * kNB + c] : C<T>{0, 0};
            }
            V4 accr = V4{0, 0, 0, 0}, acci = V4{0, 0, 0, 0};
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                accr = Mfma<T>::mac(a[u].re, g[u].re, accr);
                accr = Mfma<T>::mac(-a[u].im, g[u].im, accr);
                acci = Mfma<T>::mac(a[u].re, g[u].im, acci);
                acci = Mfma<T>::mac(a[u].im, g[u].re, acci);
            }
            // D[row][col]: row = Mfma<T>::row(lane, v) (front row), col = lane % 16 (column);
            // complex float also writes the update's W operand into the front's W window
            C<T>* wcol = nullptr;
            if constexpr (std::is_same<T, float>::value)
                if (W != nullptr) wcol = W + fp.woff[s] + static_cast<int64_t>((b + ct + p) % kWCols) * fr.m;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = rw + Mfma<T>::row(lane, v), c = ct + p;
                if (i < fr.m && c < nb) {
                    const C<T> l = C<T>{accr[v], acci[v]};
                    fr.panel[static_cast<int64_t>(b + c) * fr.m + i] = l;
                    if (wcol != nullptr) wcol[i] = gmul(l, dn);
                    lmax = fmax(lmax, static_cast<double>(hypot(l.re, l.im)));
                }
            }
        }
    }
    block_max_multiplier(st, lmax);
}

// ------------------------------------------------------------------ trailing update (MFMA)
//
// F[i][j] -= sum_c L[i][c] D[c] L[j][c] over a range of columns c, for the lower
// triangle of a region of the trailing front. One workgroup per square tile
// (ti >= tj) of one front; each wave owns a block of 16 x 16 MFMA tiles. The MFMA
// "row" operand is L[j] and the "column" operand W[i] = -L[i] D, so each lane's
// output column is a front row i: stores are contiguous.

// Tiling of the update. A wave moves one column segment of the L panel into LDS
// with a single 16-byte-per-lane DMA load (global_load_lds_dwordx4): 128 rows of
// complex float or 64 rows of complex double. That segment length is the tile
// side: 128 x 128 tiles for complex float with eight waves of 32 x 64 (three
// accumulator sets, see k_update), 64 x 64 for complex double with four waves of
// 32 x 32 (two sets). The per-wave accumulators set the register budget: 96 or 64
// registers of accumulators per lane, two waves per SIMD.
//
// Small fronts waste most of a 128 x 128 tile, so complex float also has a 64 x 64
// variant (Small) with four waves of 32 x 32 and 16-column stages (36 KB of LDS,
// four workgroups per compute unit); its wave DMA covers 64 rows of two columns.
// Either tiling accumulates each entry's products in the same order (k ascending,
// two per MFMA), so the two give identical results.
template <typename T, bool Small = false> struct UpdTile;
template <> struct UpdTile<float, false> {
    static constexpr int TS = 128, WX = 2, WY = 4, GX = 4, GY = 2, KC = 32, NW = GX * GY, MINB = 1, WPC = 1;
};
template <> struct UpdTile<float, true> {
    static constexpr int TS = 64, WX = 2, WY = 2, GX = 2, GY = 2, KC = 16, NW = GX * GY, MINB = 4, WPC = 4;
};
template <> struct UpdTile<double, false> {
    static constexpr int TS = 64, WX = 2, WY = 2, GX = 2, GY = 2, KC = 8, NW = GX * GY, MINB = 2, WPC = 3;
};

__device__ __forceinline__ void lds_dma16(const void* src, void* lds_dst) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)lds_dst, 16, 0, 0);
}

// F[i][j] -= sum_{c in [c0, c1)} L[i][c] D[c] L[j][c] for i >= j over the target
// region: columns [t0, t1) and rows [t0, m), with c1 = min(c0 + width, k),
// t0 = max(c1, min(tlo, k)) and t1 = min(thi, k) (thi < 0: t1 = m). The factor
// loop uses it for the narrow update inside an outer block (t1 = the block's
// end), the next block's columns of an outer block's update (t1 = the next
// block's end), and the rest of the trailing front (t0 = the next block's end,
// t1 = m), the last two on separate streams so that the next block's
// factorization overlaps the bulk of the update.
//
// Square tiles of the region's lower part, column-major enumeration; a
// persistent grid of gridDim.x workgroups per front walks them. Workgroups are
// placed on the XCDs round-robin by their linear index, so the tiles are split
// into eight contiguous ranges, one per XCD (neighbouring tiles share L panel
// rows in that XCD's L2). Per stage of KC columns, L rows j0.. (MFMA "row"
// operand) and i0.. (the "column" operand, multiplied by -D at read time) arrive
// by DMA into one of two LDS buffers while the other is being consumed; one
// barrier per stage. The stage numbering runs on across tiles, so the first
// stage of the next tile is in flight during the last stage of this one; the F
// tile is read in the epilogue and the sums stored. Each lane's output column is a front
// row i, so stores are contiguous. Reads past a column's end (rows >= m) land in
// rows of the tile that are never stored, so the factor storage carries a small
// zeroed tail for them.
template <typename T, bool Small = false>
__global__ void __launch_bounds__((64 * UpdTile<T, Small>::NW), (UpdTile<T, Small>::MINB))
    k_update(const int* __restrict__ nodes, int c0, int width, int tlo, int thi, FrontPlan fp, C<T>* P, C<T>* CB,
             const C<T>* __restrict__ W) {
    using V4 = typename Mfma<T>::v4;
    using UT = UpdTile<T, Small>;
    constexpr bool k3 = std::is_same<T, float>::value;  // complex float: W operand from the W windows
    constexpr int TS = UT::TS, WX = UT::WX, WY = UT::WY, KC = UT::KC, NW = UT::NW;
    static_assert(UT::GX * WX * 16 == TS && UT::GY * WY * 16 == TS, "waves tile the update tile");
    constexpr int RPL = 16 / static_cast<int>(sizeof(C<T>));  // rows per lane per DMA
    constexpr int CPD = 64 * RPL / TS;                       // tile column segments per wave DMA
    static_assert(TS * CPD == 64 * RPL && (CPD == 1 || CPD == 2), "a wave DMA covers whole column segments");
    static_assert((2 * KC / CPD) % NW == 0 && KC % CPD == 0, "stage DMAs divide among the waves");
    // LDS row stride (complex): padded, except where one DMA fills two adjacent rows
    constexpr int LDR = CPD == 1 ? TS + 16 : TS;
    __shared__ C<T> Lc[2][KC][LDR], Wc[2][KC][LDR];
    __shared__ C<T> Dsh[k3 ? 1 : kOuter];
    const int s = nodes[blockIdx.y];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    if (fr.k <= c0) return;
    const int c1 = min(c0 + width, fr.k);
    const int t0 = max(c1, min(tlo, fr.k));
    const int tend = thi < 0 ? fr.m : min(thi, fr.k);
    if (tend <= t0) return;
    const int K = c1 - c0;
    const int nrt = (fr.m - t0 + TS - 1) / TS;
    const int nct = (tend - t0 + TS - 1) / TS;
    const int ntiles = nct * nrt - nct * (nct - 1) / 2;
    // this workgroup's tiles: the workgroups fall into ng = min(8, G) groups by
    // blockIdx.x % ng (the XCDs, when G is a multiple of 8); group x takes a
    // contiguous range of tiles in proportion to its size, shared round-robin
    const int G = static_cast<int>(gridDim.x);
    const int ng = min(8, G);
    const int xcd = static_cast<int>(blockIdx.x) % ng, lidx = static_cast<int>(blockIdx.x) / ng;
    const int nloc = (G - xcd + ng - 1) / ng;
    int before = 0;
    for (int x = 0; x < xcd; ++x) before += (G - x + ng - 1) / ng;
    const int tfirst = static_cast<int>(static_cast<int64_t>(ntiles) * before / G);
    const int tlast = static_cast<int>(static_cast<int64_t>(ntiles) * (before + nloc) / G);
    int t = tfirst + lidx;
    if (t >= tlast) return;
    // tile index -> (i0, j0); column tile tj holds nrt - tj tiles, first(tj) = tj*nrt - tj(tj-1)/2
    const auto tile_origin = [&](int tt, int& i0, int& j0) {
        const double q2 = 2.0 * nrt + 1.0;
        int tj = static_cast<int>((q2 - sqrt(q2 * q2 - 8.0 * tt)) * 0.5);
        tj = max(0, min(tj, nct - 1));
        while (tj > 0 && tj * nrt - tj * (tj - 1) / 2 > tt) --tj;
        while (tj + 1 < nct && (tj + 1) * nrt - (tj + 1) * tj / 2 <= tt) ++tj;
        const int ti = tj + (tt - (tj * nrt - tj * (tj - 1) / 2));
        i0 = t0 + ti * TS;
        j0 = t0 + tj * TS;
    };

    const int tid = threadIdx.x;
    // the wave index is uniform: address arithmetic derived from it stays scalar
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64;
    const int wi = (wave % UT::GX) * (16 * WX), wj = (wave / UT::GX) * (16 * WY);  // wave offsets (i, j)
    const C<T>* L0 = fr.panel + static_cast<int64_t>(c0) * fr.m;
    const C<T>* W0 = k3 ? W + fp.woff[s] : nullptr;
    const int nst = (K + KC - 1) / KC;

    // stage st of the tile at (i0, j0): 2*KC column segments (L rows j0.. then rows i0..), 2*KC/NW per wave
    // (CPD = 2: lanes 0-31 fill row q of the LDS buffer, lanes 32-63 row q + 1)
    const auto issue_stage = [&](int buf, int st, int i0, int j0) {
        const int sub = lane / (64 / CPD), lr = lane % (64 / CPD);
#pragma unroll
        for (int u = 0; u < 2 * KC / CPD / NW; ++u) {
            const int seg = (wave + NW * u) * CPD;
            const int mat = seg / KC, q = seg % KC;
            const int cu = st * KC + q + sub, c = min(cu, K - 1);
            const int row = min((mat ? i0 : j0) + RPL * lr, fr.m - 1);
            const C<T>* src = L0 + static_cast<int64_t>(c) * fr.m + row;
            // complex float: W of column c; past the update's last column the zero
            // column at the start of the W buffer, so those steps add nothing
            if (k3 && mat) src = cu < K ? W0 + static_cast<int64_t>((c0 + c) % kWCols) * fr.m + row : W + row;
            lds_dma16(src, mat ? &Wc[buf][q][0] : &Lc[buf][q][0]);
        }
    };
    int i0, j0;
    tile_origin(t, i0, j0);
    int gs = 0;  // stage counter across tiles: LDS buffer gs & 1
    issue_stage(0, 0, i0, j0);
    if constexpr (!k3) {
        for (int q = tid; q < nst * KC; q += 64 * NW) {
            const int qq = min(q, K - 1);
            const C<T> d = L0[static_cast<int64_t>(qq) * fr.m + c0 + qq];
            Dsh[q] = q < K ? C<T>{-d.re, -d.im} : C<T>{0, 0};
        }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    // FP32 tiles use a three-product form of the complex product a w (a = L[j] =
    // ar + i ai, w = W[i] = wr + i wi, W = -L D prepared by the panel kernel):
    // t1 = ar wr, t2 = ai wi, t3 = (ar + ai)(wr + wi), re = t1 - t2,
    // im = (t3 - t1) - t2, summed in S1, S2 and S3: three MFMAs per product instead
    // of four, and two VALU additions per step for the operand sums. FP64 tiles use
    // the four-product form (accr, acci) with D applied at read time.
    if constexpr (k3) {
        // FP32: each wave's (16 WX) x (16 WY) block as XB x YB tiles of 32 x 32 on
        // v_mfma_f32_32x32x2f32 (16 passes: half the MFMA instructions of 16x16x4
        // for the same flops). Operands: lane l supplies A[l%32][l/32] and
        // B[l/32][l%32]; register e of the accumulator holds
        // D[8(e/4) + 4(l/32) + e%4][l%32].
        typedef float v16 __attribute__((ext_vector_type(16)));
        constexpr int XB = WX / 2, YB = WY / 2;
        for (;;) {
            v16 s1[XB][YB], s2[XB][YB], s3[XB][YB];
#pragma unroll
            for (int x = 0; x < XB; ++x)
#pragma unroll
                for (int y = 0; y < YB; ++y)
#pragma unroll
                    for (int e = 0; e < 16; ++e) s1[x][y][e] = s2[x][y][e] = s3[x][y][e] = 0.0f;
            const int tn = t + nloc;  // this workgroup's next tile
            int ni0 = 0, nj0 = 0;
            if (tn < tlast) tile_origin(tn, ni0, nj0);
            const int r32 = lane % 32, k2q = lane / 32;
            for (int st = 0; st < nst; ++st, ++gs) {
                const int buf = gs & 1;
                if (st + 1 < nst) issue_stage(buf ^ 1, st + 1, i0, j0);
                else if (tn < tlast) issue_stage(buf ^ 1, 0, ni0, nj0);
                // the LDS operands of a step are read one step ahead, so a step's
                // MFMAs do not wait on its own LDS reads
                C<T> ra[2][YB], rw[2][XB];
                const auto fetch = [&](int slot, int k2) {
                    const int kq = k2 + k2q;
#pragma unroll
                    for (int y = 0; y < YB; ++y) ra[slot][y] = Lc[buf][kq][wj + 32 * y + r32];
#pragma unroll
                    for (int x = 0; x < XB; ++x) rw[slot][x] = Wc[buf][kq][wi + 32 * x + r32];
                };
                fetch(0, 0);
#pragma unroll
                for (int k2 = 0; k2 < KC; k2 += 2) {
                    const int cur = (k2 / 2) & 1;
                    if (k2 + 2 < KC) fetch(cur ^ 1, k2 + 2);
                    __builtin_amdgcn_sched_barrier(0);  // keep the next step's reads ahead of this step's MFMAs
                    float bs[XB];
#pragma unroll
                    for (int x = 0; x < XB; ++x) bs[x] = rw[cur][x].re + rw[cur][x].im;
#pragma unroll
                    for (int y = 0; y < YB; ++y) {
                        const float ar = ra[cur][y].re, ai = ra[cur][y].im, as = ar + ai;
#pragma unroll
                        for (int x = 0; x < XB; ++x) {
                            s1[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, rw[cur][x].re, s1[x][y], 0, 0, 0);
                            s2[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, rw[cur][x].im, s2[x][y], 0, 0, 0);
                            s3[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(as, bs[x], s3[x][y], 0, 0, 0);
                        }
                    }
                }
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __syncthreads();
            }
            // D[p][q]: p (j) = 8(e/4) + 4(lane/32) + e%4, q (i) = lane%32; per tile all
            // the F loads first, then the sums and stores
#pragma unroll
            for (int x = 0; x < XB; ++x) {
                const int i = i0 + wi + 32 * x + r32;
#pragma unroll
                for (int y = 0; y < YB; ++y) {
                    C<T> f[16];
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int j = j0 + wj + 32 * y + 8 * (e / 4) + 4 * k2q + e % 4;
                        const int jj = min(j, fr.m - 1), ii = max(min(i, fr.m - 1), jj);  // a stored entry
                        f[e] = *fr.at(ii, jj);
                    }
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int j = j0 + wj + 32 * y + 8 * (e / 4) + 4 * k2q + e % 4;
                        const C<T> out =
                            C<T>{f[e].re + (s1[x][y][e] - s2[x][y][e]), f[e].im + ((s3[x][y][e] - s1[x][y][e]) - s2[x][y][e])};
                        if (i < fr.m && j < tend && i >= j) *fr.at(i, j) = out;
                    }
                }
            }
            if (tn >= tlast) break;
            t = tn;
            i0 = ni0;
            j0 = nj0;
        }
        return;
    }
    for (;;) {
        V4 accr[WX][WY], acci[WX][WY];
#pragma unroll
        for (int x = 0; x < WX; ++x)
#pragma unroll
            for (int y = 0; y < WY; ++y) {
                accr[x][y] = V4{0, 0, 0, 0};
                acci[x][y] = V4{0, 0, 0, 0};
            }
        const int tn = t + nloc;  // this workgroup's next tile
        int ni0 = 0, nj0 = 0;
        if (tn < tlast) tile_origin(tn, ni0, nj0);
        for (int st = 0; st < nst; ++st, ++gs) {
            const int buf = gs & 1;
            if (st + 1 < nst) issue_stage(buf ^ 1, st + 1, i0, j0);
            else if (tn < tlast) issue_stage(buf ^ 1, 0, ni0, nj0);
            // the LDS operands of a step are read one step ahead, so a step's MFMAs
            // do not wait on its own LDS reads
            C<T> ra[2][WY], rw[2][WX], rd[2];
            const auto fetch = [&](int slot, int k4) {
                const int kq = k4 + lane / 16;
                rd[slot] = Dsh[st * KC + kq];
#pragma unroll
                for (int y = 0; y < WY; ++y) ra[slot][y] = Lc[buf][kq][wj + y * 16 + (lane % 16)];
#pragma unroll
                for (int x = 0; x < WX; ++x) rw[slot][x] = Wc[buf][kq][wi + x * 16 + (lane % 16)];
            };
            fetch(0, 0);
#pragma unroll
            for (int k4 = 0; k4 < KC; k4 += 4) {
                const int cur = (k4 / 4) & 1;
                if (k4 + 4 < KC) fetch(cur ^ 1, k4 + 4);
                __builtin_amdgcn_sched_barrier(0);  // keep the next step's reads ahead of this step's MFMAs
                T ar[WY], ai[WY], br[WX], bi[WX];
#pragma unroll
                for (int y = 0; y < WY; ++y) {
                    ar[y] = ra[cur][y].re;
                    ai[y] = ra[cur][y].im;
                }
#pragma unroll
                for (int x = 0; x < WX; ++x) {
                    const C<T> w = gmul(rw[cur][x], rd[cur]);
                    br[x] = w.re;
                    bi[x] = w.im;
                }
#pragma unroll
                for (int x = 0; x < WX; ++x)
#pragma unroll
                    for (int y = 0; y < WY; ++y) {
                        accr[x][y] = Mfma<T>::mac(ar[y], br[x], accr[x][y]);
                        accr[x][y] = Mfma<T>::mac(-ai[y], bi[x], accr[x][y]);
                        acci[x][y] = Mfma<T>::mac(ar[y], bi[x], acci[x][y]);
                        acci[x][y] = Mfma<T>::mac(ai[y], br[x], acci[x][y]);
                    }
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        }
        // D[p][q]: lane holds p = Mfma<T>::row(lane, v) (j), q = lane%16 (i); all
        // the F loads first, then the sums and stores
        C<T> f[WX][WY][4];
#pragma unroll
        for (int x = 0; x < WX; ++x)
#pragma unroll
            for (int y = 0; y < WY; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = i0 + wi + x * 16 + (lane % 16);
                    const int j = j0 + wj + y * 16 + Mfma<T>::row(lane, v);
                    const int jj = min(j, fr.m - 1), ii = max(min(i, fr.m - 1), jj);  // a stored entry
                    f[x][y][v] = *fr.at(ii, jj);
                }
#pragma unroll
        for (int x = 0; x < WX; ++x)
#pragma unroll
            for (int y = 0; y < WY; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = i0 + wi + x * 16 + (lane % 16);
                    const int j = j0 + wj + y * 16 + Mfma<T>::row(lane, v);
                    const C<T> out = C<T>{f[x][y][v].re + accr[x][y][v], f[x][y][v].im + acci[x][y][v]};
                    if (i < fr.m && j < tend && i >= j) *fr.at(i, j) = out;
                }
        if (tn >= tlast) break;
        t = tn;
        i0 = ni0;
        j0 = nj0;
    }
}

// ------------------------------------------------------------------ launchers

template <typename T>
void launch_zero(hipStream_t st, int64_t n, void* p) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_zero<T>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, n,
                       static_cast<C<T>*>(p));
}

template <typename T>
void launch_scatter_panel(hipStream_t st, int64_t nnz, const int64_t* s2p, const cd* a, void* P) {
    if (nnz <= 0) return;
    hipLaunchKernelGGL(k_scatter_panel<T>, dim3(static_cast<unsigned>((nnz + 255) / 256)), dim3(256), 0, st, nnz, s2p, a,
                       static_cast<C<T>*>(P));
}

void launch_stats_init(hipStream_t st, FactorStatsDev* stats) {
    hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(kStatSlots), 0, st, stats);
}

template <typename T>
void launch_failure_stats(hipStream_t st, int ns, int64_t entries, const FrontPlan& fp, const void* P, int fail,
                          FactorStatsDev* stats) {
    hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(kStatSlots), 0, st, stats);
    if (ns <= 0 || entries <= 0) return;
    const int64_t blocks = std::min<int64_t>((entries + 255) / 256, 8192);
    hipLaunchKernelGGL(k_failure_stats<T>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, ns, fp,
                       static_cast<const C<T>*>(P), fail, stats);
}

template <typename T>
size_t small_lds_bytes(int max_m) {
    return static_cast<size_t>(max_m) * (max_m + 1) * sizeof(C<T>);
}

template <typename T>
void launch_front_small(hipStream_t st, const int* nodes, int count, int max_m, const FrontPlan& fp, void* P, void* CB,
                        void* xinv, double pivot_floor, FactorStatsDev* stats) {
    if (count <= 0) return;
    allow_large_lds(k_front_small<T>, small_lds_bytes<T>(max_m));
    hipLaunchKernelGGL(k_front_small<T>, dim3(count), dim3(256), small_lds_bytes<T>(max_m), st, nodes, fp,
                       static_cast<C<T>*>(P), static_cast<C<T>*>(CB), static_cast<C<T>*>(xinv), pivot_floor, stats);
}

template <typename T>
void launch_extend_add(hipStream_t st, const int* nodes, const int* blk_start, int count, int total_blocks,
                       int max_m, const FrontPlan& fp, void* P, void* CB) {
    const int row_chunks = (max_m + ea_rows<T>() - 1) / ea_rows<T>();
    if (count <= 0 || total_blocks <= 0 || row_chunks <= 0) return;
    hipLaunchKernelGGL(k_extend_add<T>, dim3(total_blocks, row_chunks), dim3(256), 0, st, nodes, blk_start, count, fp,
                       static_cast<C<T>*>(P), static_cast<C<T>*>(CB));
}

template <typename T>
void launch_block_step(hipStream_t st, const int* nodes, int count, int b, int max_row_blocks, int max_tiles,
                       const FrontPlan& fp, void* P, void* CB, void* mws, void* xinv, double pivot_floor,
                       FactorStatsDev* stats, void* W) {
    const size_t diag_lds = static_cast<size_t>(kNB * (kNB + 1) + 5 * kNB) * sizeof(C<T>);
    allow_large_lds(k_diag<T>, diag_lds);
    hipLaunchKernelGGL(k_diag<T>, dim3(count), dim3(1024), diag_lds, st, nodes, b, fp, static_cast<C<T>*>(P),
                       static_cast<C<T>*>(CB), static_cast<C<T>*>(mws), static_cast<C<T>*>(xinv), pivot_floor, stats);
    if (max_row_blocks > 0) {
        hipLaunchKernelGGL(k_panel<T>, dim3(max_row_blocks, count), dim3(256), 0, st, nodes, b, fp,
                           static_cast<C<T>*>(P), static_cast<C<T>*>(CB), static_cast<const C<T>*>(mws), stats,
                           static_cast<C<T>*>(W));
    }
    (void)max_tiles;
}

namespace {
// compute units of the current device (cached per device)
int device_cus() {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cus[dev] = v;
    }
    return cus[dev];
}
}  // namespace

template <typename T>
void launch_update(hipStream_t st, const int* nodes, int count, int c0, int width, int tlo, int thi, int max_tiles,
                   int max_wgs, const FrontPlan& fp, void* P, void* CB, const void* W, bool small) {
    if (count <= 0 || max_tiles <= 0) return;
    if constexpr (std::is_same<T, float>::value) {
        if (W == nullptr) throw std::invalid_argument("launch_update: complex float needs the W windows");
        if (small) {
            hipLaunchKernelGGL((k_update<float, true>), dim3(max_tiles, count), dim3(64 * UpdTile<float, true>::NW), 0,
                               st, nodes, c0, width, tlo, thi, fp, static_cast<C<T>*>(P), static_cast<C<T>*>(CB),
                               static_cast<const C<T>*>(W));
            return;
        }
    }
    // workgroups per front: one per tile, or (max_wgs > 0) a persistent grid of
    // about max_wgs workgroups over all fronts, a multiple of 8 (one share per
    // XCD). A persistent grid holds its compute units until it ends, so the
    // update that runs beside the factorization of the next block takes one
    // tile per workgroup and lets the other stream's work in between tiles.
    int g = max_tiles;
    if (max_wgs > 0) g = std::min(max_tiles, std::max(8, (max_wgs / count) / 8 * 8));
    hipLaunchKernelGGL(k_update<T>, dim3(g, count), dim3(64 * UpdTile<T>::NW), 0, st, nodes, c0, width, tlo, thi, fp,
                       static_cast<C<T>*>(P), static_cast<C<T>*>(CB), static_cast<const C<T>*>(W));
}

template <typename T>
int update_tile(bool small) {
    if constexpr (std::is_same<T, float>::value) return small ? UpdTile<float, true>::TS : UpdTile<float>::TS;
    return UpdTile<T>::TS;
}

template void launch_failure_stats<double>(hipStream_t, int, int64_t, const FrontPlan&, const void*, int,
                                           FactorStatsDev*);
template void launch_failure_stats<float>(hipStream_t, int, int64_t, const FrontPlan&, const void*, int,
                                          FactorStatsDev*);
template void launch_zero<double>(hipStream_t, int64_t, void*);
template void launch_zero<float>(hipStream_t, int64_t, void*);
template void launch_scatter_panel<double>(hipStream_t, int64_t, const int64_t*, const cd*, void*);
template void launch_scatter_panel<float>(hipStream_t, int64_t, const int64_t*, const cd*, void*);
template size_t small_lds_bytes<double>(int);
template size_t small_lds_bytes<float>(int);
template void launch_front_small<double>(hipStream_t, const int*, int, int, const FrontPlan&, void*, void*, void*,
                                         double, FactorStatsDev*);
template void launch_front_small<float>(hipStream_t, const int*, int, int, const FrontPlan&, void*, void*, void*,
                                        double, FactorStatsDev*);
template void launch_extend_add<double>(hipStream_t, const int*, const int*, int, int, int, const FrontPlan&, void*, void*);
template void launch_extend_add<float>(hipStream_t, const int*, const int*, int, int, int, const FrontPlan&, void*, void*);
template void launch_update<double>(hipStream_t, const int*, int, int, int, int, int, int, int, const FrontPlan&, void*, void*,
                                   const void*, bool);
template void launch_update<float>(hipStream_t, const int*, int, int, int, int, int, int, int, const FrontPlan&, void*, void*,
                                  const void*, bool);
template <typename T> int update_resident_wgs() { return device_cus() * UpdTile<T>::WPC; }
template int update_resident_wgs<double>();
template int update_resident_wgs<float>();
template int update_tile<double>(bool);
template int update_tile<float>(bool);
template void launch_block_step<double>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, void*, void*,
                                        void*, void*, double, FactorStatsDev*, void*);
template void launch_block_step<float>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, void*, void*,
                                       void*, void*, double, FactorStatsDev*, void*);

}  // namespace aphi_dev

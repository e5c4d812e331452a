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

// Records |d| and checks it against the floor (src/factorization.cpp:191-210).
template <typename T>
__device__ __forceinline__ void record_pivot(FactorStatsDev* st, const FrontPlan& fp, int col, C<T> d, double floor_) {
    const double mag = hypot(static_cast<double>(d.re), static_cast<double>(d.im));
    stat_min(&st->smallest_pivot, mag);
    atomic_max_nonneg(&st->largest_pivot, mag);
    if (!(mag > floor_)) atomicMin(&st->fail_col, fp.col_ref[col]);
}

// Block-wide max of a non-negative value into the stats.
__device__ __forceinline__ void block_max_multiplier(FactorStatsDev* st, double v, double* red) {
    // wave64 reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&st->largest_multiplier, v);
    (void)red;
}

}  // namespace

// ------------------------------------------------------------------ setup

template <typename T>
__global__ void k_zero(int64_t n, C<T>* __restrict__ p) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = C<T>{0, 0};
}

// The equilibrated A, scattered into the panels (each entry lands in the front
// of the supernode owning its lower column; no two entries share a position).
template <typename T>
__global__ void k_scatter_panel(int64_t nnz, const int64_t* __restrict__ s2p, const cd* __restrict__ a,
                                C<T>* __restrict__ P) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < nnz) P[s2p[k]] = to_c<T>(a[k]);
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
    for (int j = 0; j < k; ++j) {
        const C<T> d = F[j * m + j];
        if (tid == 0) record_pivot<T>(st, fp, fp.first[s] + j, d, pivot_floor);
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
    block_max_multiplier(st, lmax, nullptr);

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
// chunk of kEaRows front rows); it owns those entries exclusively, clears their
// update-block part, then adds each child's contributions in child order --
// deterministic without atomics. The child index ranges come from the
// precomputed bounds; each wave takes two child columns at a time with all of
// their rows (kEaRows / 64 per lane) loaded before any store.
template <typename T>
__global__ void __launch_bounds__(256) k_extend_add(const int* __restrict__ nodes, const int* __restrict__ blk_start,
                                                    int nnodes, FrontPlan fp, C<T>* P, C<T>* CB) {
    constexpr int RL = kEaRows / 64;  // rows per lane per column
    // locate the front this block belongs to
    int lo = 0, hi = nnodes;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (blk_start[mid] <= static_cast<int>(blockIdx.x)) lo = mid;
        else hi = mid;
    }
    const int s = nodes[lo];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    const int cblk = static_cast<int>(blockIdx.x) - blk_start[lo];
    const int c0 = cblk * 64;
    const int c1 = min(c0 + 64, fr.m);
    const int r0 = max(c0, static_cast<int>(blockIdx.y) * kEaRows);
    const int r1 = min(fr.m, (static_cast<int>(blockIdx.y) + 1) * kEaRows);
    if (r0 >= r1) return;
    const int q0 = r0 / 64, q1 = (r1 + 63) / 64;  // row range in 64-blocks (r0, r1 multiples of 64 or m)
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;

    // clear owned update-block entries (lower part)
    for (int j = max(c0, fr.k) + wave; j < c1; j += 4)
        for (int i = max(j, r0) + lane; i < r1; i += 64) *fr.at(i, j) = C<T>{0, 0};

    for (int c = fp.child_ptr[s]; c < fp.child_ptr[s + 1]; ++c) {
        __syncthreads();
        const int ch = fp.child[c];
        const int cm = static_cast<int>(fp.row_ptr[ch + 1] - fp.row_ptr[ch]);
        const int ck = fp.first[ch + 1] - fp.first[ch];
        const int cu = cm - ck;
        const int* rel = fp.rel + fp.row_ptr[ch] + ck;
        const int* bnd = fp.ea_bound + fp.ea_bound_off[ch];
        const int a = bnd[cblk], e = bnd[cblk + 1];   // child columns landing in [c0, c1)
        const int ia = bnd[q0], ie = bnd[q1];         // child rows landing in [r0, r1)
        const C<T>* cb = CB + fp.cb_off[ch];
        for (int jb = a + 2 * wave; jb < e; jb += 8) {
            int dr[2][RL];
            C<T> v[2][RL];
            int pj[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int jc = jb + h;
                pj[h] = jc < e ? rel[jc] : 0;
                const int base = max(jc, ia);
#pragma unroll
                for (int u = 0; u < RL; ++u) {
                    const int ic = base + lane + 64 * u;
                    const bool ok = jc < e && ic < ie;
                    dr[h][u] = ok ? rel[ic] : -1;
                    v[h][u] = ok ? cb[static_cast<int64_t>(jc) * cu + ic] : C<T>{0, 0};
                }
            }
            C<T> cur[2][RL];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int u = 0; u < RL; ++u)
                    if (dr[h][u] >= 0) cur[h][u] = *fr.at(dr[h][u], pj[h]);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int u = 0; u < RL; ++u)
                    if (dr[h][u] >= 0) *fr.at(dr[h][u], pj[h]) = gadd(cur[h][u], v[h][u]);
        }
    }
}

// Diagonal block of step b for each front with k > b, one wavefront per front.
// Lane r holds row r of the nb x nb block in registers (rows and columns beyond
// nb padded with the identity); the right-looking LDL^T takes each column value
// it needs from the owning lane with a scalar lane read, so the elimination runs
// without barriers or LDS traffic. Then lane c forms column c of L_bb^-1 by
// forward substitution (L read as LDS broadcasts), and the block inverse (kept
// for the solves) and M = L_bb^-T D_bb^-1 (for the panel step) are written out.
template <typename T>
__global__ void __launch_bounds__(64) k_diag(const int* __restrict__ nodes, int b, FrontPlan fp, C<T>* P, C<T>* CB,
                                             C<T>* __restrict__ mws, C<T>* __restrict__ xinv, double pivot_floor,
                                             FactorStatsDev* st) {
    extern __shared__ unsigned char smem_raw[];
    constexpr int LD = kNB + 1;
    C<T>* Ls = reinterpret_cast<C<T>*>(smem_raw);  // Ls[c*LD + r]: factored block, then L_bb^-1 by columns
    C<T>* Dl = Ls + kNB * LD;                     // pivots
    const int s = nodes[blockIdx.x];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    if (fr.k <= b) return;
    const int nb = min(kNB, fr.k - b);
    const int lane = threadIdx.x;
    C<T>* blk = fr.panel + static_cast<int64_t>(b) * fr.m + b;  // element (b, b) of the front
    double lmax = 0.0;
    if constexpr (sizeof(T) == 4) {
        C<T> a[kNB];
#pragma unroll
        for (int c = 0; c < kNB; ++c) {
            const bool in = lane < nb && c < nb;
            a[c] = in ? (c <= lane ? blk[static_cast<int64_t>(c) * fr.m + lane] : C<T>{0, 0})
                      : C<T>{static_cast<T>(lane == c ? 1 : 0), 0};
        }
        // a[c] of lane r is A[r][c]; entries above the diagonal are scratch.
#pragma unroll
        for (int j = 0; j < kNB; ++j) {
            const C<T> aj = a[j];
            const C<T> l = gmul(aj, grecip(lane_read(aj, j)));
            if (lane > j) a[j] = l;
#pragma unroll
            for (int c = j + 1; c < kNB; ++c) a[c] = gfms(a[c], l, lane_read(aj, c));  // A[r][c] -= l_r A[c][j]
        }
#pragma unroll
        for (int c = 0; c < kNB; ++c) Ls[c * LD + lane] = a[c];
    } else {
        // complex double: the same elimination with row `lane` in LDS (64 complex
        // doubles per lane would not fit the register file)
        C<T>* U = Dl + kNB;  // the pivot column before scaling
#pragma unroll 8
        for (int c = 0; c < kNB; ++c) {
            const bool in = lane < nb && c < nb;
            Ls[c * LD + lane] = in ? (c <= lane ? blk[static_cast<int64_t>(c) * fr.m + lane] : C<T>{0, 0})
                                   : C<T>{static_cast<T>(lane == c ? 1 : 0), 0};
        }
        __syncthreads();
        for (int j = 0; j < kNB; ++j) {
            const C<T> aj = Ls[j * LD + lane];
            const C<T> l = gmul(aj, grecip(Ls[j * LD + j]));
            U[lane] = aj;
            __syncthreads();
            if (lane > j) Ls[j * LD + lane] = l;
#pragma unroll 4
            for (int c = j + 1; c < kNB; ++c) Ls[c * LD + lane] = gfms(Ls[c * LD + lane], l, U[c]);
            __syncthreads();
        }
    }
    __syncthreads();
    for (int c = 0; c < nb; ++c) {
        const C<T> v = Ls[c * LD + lane];
        if (c < lane && lane < nb) lmax = fmax(lmax, static_cast<double>(hypot(v.re, v.im)));
        if (c <= lane && lane < nb) blk[static_cast<int64_t>(c) * fr.m + lane] = v;
    }
    const C<T> dd = Ls[lane * LD + lane];
    Dl[lane] = dd;
    if (lane < nb) record_pivot<T>(st, fp, fp.first[s] + b + lane, dd, pivot_floor);
    block_max_multiplier(st, lmax, nullptr);
    // column `lane` of L_bb^-1: x[r] = delta(r, lane) - sum_{q<r} L[r][q] x[q];
    // x in registers for complex float, in a per-lane LDS column for complex
    // double (whose 64 entries would not fit the register file next to the rest)
    if constexpr (sizeof(T) == 4) {
        C<T> x[kNB];
#pragma unroll
        for (int r = 0; r < kNB; ++r) {
            C<T> acc = C<T>{static_cast<T>(lane == r ? 1 : 0), 0};
#pragma unroll
            for (int q = 0; q < r; ++q) acc = gfms(acc, Ls[q * LD + r], x[q]);
            x[r] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kNB; ++r) Ls[lane * LD + r] = x[r];  // Ls[q*LD + c] = Linv[c][q]
    } else {
        C<T>* Xs = Dl + 2 * kNB;  // Xs[r*kNB + lane] = x[r] of lane
#pragma unroll
        for (int r = 0; r < kNB; ++r) {
            C<T> acc = C<T>{static_cast<T>(lane == r ? 1 : 0), 0};
#pragma unroll 16
            for (int q = 0; q < r; ++q) acc = gfms(acc, Ls[q * LD + r], Xs[q * kNB + lane]);
            Xs[r * kNB + lane] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kNB; ++r) Ls[lane * LD + r] = Xs[r * kNB + lane];
    }
    __syncthreads();
    // M[q][c] = Linv[c][q] / D[c] (c >= q) at mws[q*kNB + c]; the inverse, column q
    // and row c at q*nb + c, is kept for the solves
    C<T>* M = mws + static_cast<int64_t>(blockIdx.x) * kNB * kNB;
    C<T>* Xg = xinv + fp.xinv_off[s] + static_cast<int64_t>(b / kNB) * kNB * kNB;
    for (int idx = lane; idx < nb * nb; idx += 64) {
        const int q = idx / nb, c = idx % nb;
        const C<T> v = c >= q ? Ls[q * LD + c] : C<T>{0, 0};
        Xg[idx] = v;
        M[q * kNB + c] = c >= q ? gmul(v, grecip(Dl[c])) : C<T>{0, 0};
    }
}

// Panel step: rows i in [b+nb, m) of block column b become L = F * M. One
// workgroup per (64-row block, front); its rows are staged in LDS first, so the
// in-place write is safe.
template <typename T>
__global__ void __launch_bounds__(256) k_panel(const int* __restrict__ nodes, int b, FrontPlan fp, C<T>* P, C<T>* CB,
                                               const C<T>* __restrict__ mws, FactorStatsDev* st) {
    extern __shared__ unsigned char smem_raw[];
    const int s = nodes[blockIdx.y];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    if (fr.k <= b) return;
    const int nb = min(kNB, fr.k - b);
    const int r0 = b + nb + static_cast<int>(blockIdx.x) * 64;
    if (r0 >= fr.m) return;
    const int nr = min(64, fr.m - r0);
    C<T>* Fs = reinterpret_cast<C<T>*>(smem_raw);  // Fs[q*64 + row]
    C<T>* Ms = Fs + 64 * kNB;                     // Ms[q*kNB + c]
    const C<T>* M = mws + static_cast<int64_t>(blockIdx.y) * kNB * kNB;
    const int tid = threadIdx.x, nt = blockDim.x;
    {
        // all loads of the row block and of M in flight before the LDS stores
        C<T> f[64 * kNB / 256], g[kNB * kNB / 256];
#pragma unroll
        for (int u = 0; u < 64 * kNB / 256; ++u) {
            const int idx = tid + 256 * u, q = idx / 64, r = idx % 64;
            f[u] = (q < nb && r < nr) ? fr.panel[static_cast<int64_t>(b + q) * fr.m + r0 + r] : C<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < kNB * kNB / 256; ++u) {
            const int idx = tid + 256 * u;
            g[u] = M[idx];
        }
#pragma unroll
        for (int u = 0; u < 64 * kNB / 256; ++u) Fs[tid + 256 * u] = f[u];
#pragma unroll
        for (int u = 0; u < kNB * kNB / 256; ++u) Ms[tid + 256 * u] = g[u];
    }
    __syncthreads();
    const int r = tid % 64;
    double lmax = 0.0;
    for (int c = tid / 64; c < nb; c += nt / 64) {
        C<T> acc = C<T>{0, 0};
        for (int q = 0; q <= c; ++q) acc = gfma(acc, Fs[q * 64 + r], Ms[q * kNB + c]);
        if (r < nr) {
            fr.panel[static_cast<int64_t>(b + c) * fr.m + r0 + r] = acc;
            lmax = fmax(lmax, static_cast<double>(hypot(acc.re, acc.im)));
        }
    }
    block_max_multiplier(st, lmax, nullptr);
}

// ------------------------------------------------------------------ trailing update (MFMA)
//
// F[i][j] -= sum_c L[i][c] D[c] L[j][c] over the block column c in [b, b+nb),
// for the lower triangle of rows/columns [b+nb, m). One workgroup per 64x64 tile
// (ti >= tj) of one front; four waves, each a 32x32 quadrant of 2x2 MFMA 16x16
// blocks with separate real/imaginary accumulators (complex product as four real
// products). The MFMA "row" operand is L[j] and the "column" operand W[i] =
// L[i] D, so each lane's output column is a front row i: stores are contiguous.

// Tiling of the update. A wave moves one column segment of the L panel into LDS
// with a single 16-byte-per-lane DMA load (global_load_lds_dwordx4): 128 rows of
// complex float or 64 rows of complex double. That segment length is the tile
// side: 128x128 tiles for complex float (each wave a 64x64 quadrant of 4x4 MFMA
// blocks), 64x64 for complex double (2x2 blocks per wave).
template <typename T> struct UpdTile;
template <> struct UpdTile<float> { static constexpr int WB = 4, KC = 16; };
template <> struct UpdTile<double> { static constexpr int WB = 2, KC = 8; };

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
// factorization overlaps the bulk of the update. One workgroup per square tile
// of the region's lower part (column-major enumeration). The tile of F is loaded into the MFMA accumulators first and
// written back at the end (no read-modify-write epilogue). Per stage of KC
// columns, L rows j0.. (MFMA "row" operand) and i0.. (the "column" operand,
// multiplied by -D at read time) arrive by DMA into one of two LDS buffers
// while the other is being consumed; one barrier per stage. Each lane's output
// column is a front row i, so stores are contiguous. Reads past a column's end
// (rows >= m) land in rows of the tile that are never stored, so the factor
// storage carries a small zeroed tail for them.
template <typename T>
__global__ void __launch_bounds__(256, 2) k_update(const int* __restrict__ nodes, int c0, int width, int tlo, int thi,
                                                   FrontPlan fp, C<T>* P, C<T>* CB) {
    using V4 = typename Mfma<T>::v4;
    constexpr int WB = UpdTile<T>::WB, KC = UpdTile<T>::KC;
    constexpr int TS = 32 * WB;                          // tile side
    constexpr int RPL = 16 / static_cast<int>(sizeof(C<T>));  // rows per lane per DMA
    static_assert(TS == 64 * RPL, "one wave DMA covers one tile column segment");
    constexpr int LDR = TS + 16;                         // LDS row stride (complex)
    __shared__ C<T> Lc[2][KC][LDR], Wc[2][KC][LDR];
    __shared__ C<T> Dsh[kOuter];
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
    const int t = static_cast<int>(blockIdx.x);
    if (t >= nct * nrt - nct * (nct - 1) / 2) return;
    // column tile tj holds nrt - tj tiles; first(tj) = tj*nrt - tj(tj-1)/2
    const double q2 = 2.0 * nrt + 1.0;
    int tj = static_cast<int>((q2 - sqrt(q2 * q2 - 8.0 * t)) * 0.5);
    tj = max(0, min(tj, nct - 1));
    while (tj > 0 && tj * nrt - tj * (tj - 1) / 2 > t) --tj;
    while (tj + 1 < nct && (tj + 1) * nrt - (tj + 1) * tj / 2 <= t) ++tj;
    const int ti = tj + (t - (tj * nrt - tj * (tj - 1) / 2));
    const int i0 = t0 + ti * TS, j0 = t0 + tj * TS;

    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wi = (wave & 1) * (TS / 2), wj = (wave >> 1) * (TS / 2);  // quadrant offsets (i, j)
    const C<T>* L0 = fr.panel + static_cast<int64_t>(c0) * fr.m;
    const int nst = (K + KC - 1) / KC;

    // stage st: 2*KC column segments (L rows j0.. then rows i0..), KC/2 per wave
    const auto issue_stage = [&](int buf, int st) {
#pragma unroll
        for (int u = 0; u < 2 * KC / 4; ++u) {
            const int jdx = wave + 4 * u;
            const int mat = jdx / KC, q = jdx % KC;
            const int c = min(st * KC + q, K - 1);
            const int row = min((mat ? i0 : j0) + RPL * lane, fr.m - 1);
            lds_dma16(L0 + static_cast<int64_t>(c) * fr.m + row, mat ? &Wc[buf][q][0] : &Lc[buf][q][0]);
        }
    };
    issue_stage(0, 0);

    V4 accr[WB][WB], acci[WB][WB];
#pragma unroll
    for (int x = 0; x < WB; ++x)
#pragma unroll
        for (int y = 0; y < WB; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wi + x * 16 + (lane % 16);
                const int j = j0 + wj + y * 16 + Mfma<T>::row(lane, v);
                const int jj = min(j, fr.m - 1), ii = max(min(i, fr.m - 1), jj);  // a stored entry
                const C<T> c = *fr.at(ii, jj);
                const bool ok = i < fr.m && j < tend && i >= j;
                accr[x][y][v] = ok ? c.re : static_cast<T>(0);
                acci[x][y][v] = ok ? c.im : static_cast<T>(0);
            }
    for (int q = tid; q < nst * KC; q += 256) {
        const int qq = min(q, K - 1);
        const C<T> d = L0[static_cast<int64_t>(qq) * fr.m + c0 + qq];
        Dsh[q] = q < K ? C<T>{-d.re, -d.im} : C<T>{0, 0};
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) issue_stage(buf ^ 1, st + 1);
#pragma unroll
        for (int k4 = 0; k4 < KC; k4 += 4) {
            const int kq = k4 + lane / 16;
            const C<T> d = Dsh[st * KC + kq];
            T ar[WB], ai[WB], br[WB], bi[WB];
#pragma unroll
            for (int y = 0; y < WB; ++y) {
                const C<T> a = Lc[buf][kq][wj + y * 16 + (lane % 16)];
                ar[y] = a.re;
                ai[y] = a.im;
            }
#pragma unroll
            for (int x = 0; x < WB; ++x) {
                const C<T> w = gmul(Wc[buf][kq][wi + x * 16 + (lane % 16)], d);
                br[x] = w.re;
                bi[x] = w.im;
            }
#pragma unroll
            for (int x = 0; x < WB; ++x)
#pragma unroll
                for (int y = 0; y < WB; ++y) {
                    accr[x][y] = Mfma<T>::mac(ar[y], br[x], accr[x][y]);
                    accr[x][y] = Mfma<T>::mac(-ai[y], bi[x], accr[x][y]);
                    acci[x][y] = Mfma<T>::mac(ar[y], bi[x], acci[x][y]);
                    acci[x][y] = Mfma<T>::mac(ai[y], br[x], acci[x][y]);
                }
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }
    // D[p][q]: lane holds p = Mfma<T>::row(lane, v) (j), q = lane%16 (i)
#pragma unroll
    for (int x = 0; x < WB; ++x)
#pragma unroll
        for (int y = 0; y < WB; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wi + x * 16 + (lane % 16);
                const int j = j0 + wj + y * 16 + Mfma<T>::row(lane, v);
                if (i < fr.m && j < tend && i >= j) *fr.at(i, j) = C<T>{accr[x][y][v], acci[x][y][v]};
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
                       int row_chunks, const FrontPlan& fp, void* P, void* CB) {
    if (count <= 0 || total_blocks <= 0 || row_chunks <= 0) return;
    hipLaunchKernelGGL(k_extend_add<T>, dim3(total_blocks, row_chunks), dim3(256), 0, st, nodes, blk_start, count, fp,
                       static_cast<C<T>*>(P), static_cast<C<T>*>(CB));
}

template <typename T>
void launch_block_step(hipStream_t st, const int* nodes, int count, int b, int max_row_blocks, int max_tiles,
                       const FrontPlan& fp, void* P, void* CB, void* mws, void* xinv, double pivot_floor,
                       FactorStatsDev* stats) {
    const size_t diag_lds = static_cast<size_t>(kNB * (kNB + 1) + 2 * kNB + (sizeof(T) == 8 ? kNB * kNB : 0)) * sizeof(C<T>);
    allow_large_lds(k_diag<T>, diag_lds);
    hipLaunchKernelGGL(k_diag<T>, dim3(count), dim3(64), diag_lds, st, nodes, b, fp, static_cast<C<T>*>(P),
                       static_cast<C<T>*>(CB), static_cast<C<T>*>(mws), static_cast<C<T>*>(xinv), pivot_floor, stats);
    if (max_row_blocks > 0) {
        const size_t panel_lds = static_cast<size_t>(64 * kNB + kNB * kNB) * sizeof(C<T>);
        allow_large_lds(k_panel<T>, panel_lds);
        hipLaunchKernelGGL(k_panel<T>, dim3(max_row_blocks, count), dim3(256), panel_lds, st, nodes, b, fp,
                           static_cast<C<T>*>(P), static_cast<C<T>*>(CB), static_cast<const C<T>*>(mws), stats);
    }
    (void)max_tiles;
}

template <typename T>
void launch_update(hipStream_t st, const int* nodes, int count, int c0, int width, int tlo, int thi, int max_tiles,
                   const FrontPlan& fp, void* P, void* CB) {
    if (count <= 0 || max_tiles <= 0) return;
    hipLaunchKernelGGL(k_update<T>, dim3(max_tiles, count), dim3(256), 0, st, nodes, c0, width, tlo, thi, fp,
                       static_cast<C<T>*>(P), static_cast<C<T>*>(CB));
}

template <typename T>
int update_tile() {
    return 32 * UpdTile<T>::WB;
}

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
template void launch_update<double>(hipStream_t, const int*, int, int, int, int, int, int, const FrontPlan&, void*, void*);
template void launch_update<float>(hipStream_t, const int*, int, int, int, int, int, int, const FrontPlan&, void*, void*);
template int update_tile<double>();
template int update_tile<float>();
template void launch_block_step<double>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, void*, void*,
                                        void*, void*, double, FactorStatsDev*);
template void launch_block_step<float>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, void*, void*,
                                       void*, void*, double, FactorStatsDev*);

}  // namespace aphi_dev

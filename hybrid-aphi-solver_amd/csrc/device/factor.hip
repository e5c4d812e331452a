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

// Extend-add for large fronts. Workgroup = (front, block of 64 front columns);
// it owns those columns exclusively, clears their update-block part, then adds
// each child's contributions in child order -- deterministic without atomics.
template <typename T>
__global__ void __launch_bounds__(256) k_extend_add(const int* __restrict__ nodes, const int* __restrict__ blk_start,
                                                    int nnodes, FrontPlan fp, C<T>* P, C<T>* CB) {
    // locate the front this block belongs to
    int lo = 0, hi = nnodes;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (blk_start[mid] <= static_cast<int>(blockIdx.x)) lo = mid;
        else hi = mid;
    }
    const int s = nodes[lo];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    const int c0 = (static_cast<int>(blockIdx.x) - blk_start[lo]) * 64;
    const int c1 = min(c0 + 64, fr.m);
    const int tid = threadIdx.x, nt = blockDim.x;

    // clear owned update-block columns (lower part)
    for (int j = max(c0, fr.k); j < c1; ++j)
        for (int i = j + tid; i < fr.m; i += nt) *fr.at(i, j) = C<T>{0, 0};

    for (int c = fp.child_ptr[s]; c < fp.child_ptr[s + 1]; ++c) {
        __syncthreads();
        const int ch = fp.child[c];
        const int cm = static_cast<int>(fp.row_ptr[ch + 1] - fp.row_ptr[ch]);
        const int ck = fp.first[ch + 1] - fp.first[ch];
        const int cu = cm - ck;
        const int* rel = fp.rel + fp.row_ptr[ch] + ck;
        // child columns mapping into [c0, c1): rel is increasing
        int a = 0, b = cu;
        while (a < b) { const int mid = (a + b) / 2; if (rel[mid] < c0) a = mid + 1; else b = mid; }
        int e = a, f = cu;
        while (e < f) { const int mid = (e + f) / 2; if (rel[mid] < c1) e = mid + 1; else f = mid; }
        const C<T>* cb = CB + fp.cb_off[ch];
        for (int jc = a; jc < e; ++jc) {
            const int pj = rel[jc];
            for (int ic = jc + tid; ic < cu; ic += nt) {
                C<T>* dst = fr.at(rel[ic], pj);
                *dst = gadd(*dst, cb[static_cast<int64_t>(jc) * cu + ic]);
            }
        }
    }
}

// Diagonal block of step b for each front with k > b: LDL^T of the nbb x nbb
// block in LDS, written back; then M = L_bb^-T D_bb^-1 (upper triangular) into
// the per-front workspace for the panel step.
template <typename T>
__global__ void __launch_bounds__(256) k_diag(const int* __restrict__ nodes, int b, FrontPlan fp, C<T>* P, C<T>* CB,
                                              C<T>* __restrict__ mws, C<T>* __restrict__ xinv, double pivot_floor,
                                              FactorStatsDev* st) {
    extern __shared__ unsigned char smem_raw[];
    const int s = nodes[blockIdx.x];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    if (fr.k <= b) return;
    const int nb = min(kNB, fr.k - b);
    C<T>* A = reinterpret_cast<C<T>*>(smem_raw);  // A[c*kNB + r]
    C<T>* X = A + kNB * kNB;                      // L^-1, X[c*kNB + r]
    C<T>* lcol = X + kNB * kNB;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < nb * nb; idx += nt) {
        const int c = idx / nb, r = idx % nb;
        A[c * kNB + r] = (r >= c) ? *fr.at(b + r, b + c) : C<T>{0, 0};
    }
    __syncthreads();
    double lmax = 0.0;
    for (int j = 0; j < nb; ++j) {
        const C<T> d = A[j * kNB + j];
        if (tid == 0) record_pivot<T>(st, fp, fp.first[s] + b + j, d, pivot_floor);
        const C<T> rd = grecip(d);
        for (int i = j + 1 + tid; i < nb; i += nt) {
            const C<T> l = gmul(A[j * kNB + i], rd);
            lcol[i] = l;
            lmax = fmax(lmax, static_cast<double>(hypot(l.re, l.im)));
        }
        __syncthreads();
        const int w = nb - j - 1;
        for (int idx = tid; idx < w * w; idx += nt) {
            const int c = j + 1 + idx / w, i = j + 1 + idx % w;
            if (i < c) continue;
            A[c * kNB + i] = gfms(A[c * kNB + i], A[j * kNB + i], lcol[c]);
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < nb; i += nt) A[j * kNB + i] = lcol[i];
        __syncthreads();
    }
    block_max_multiplier(st, lmax, nullptr);
    for (int idx = tid; idx < nb * nb; idx += nt) {
        const int c = idx / nb, r = idx % nb;
        if (r >= c) *fr.at(b + r, b + c) = A[c * kNB + r];
    }
    // X = L^-1 (unit lower), one column per thread
    if (tid < nb) {
        const int c = tid;
        for (int r = 0; r < nb; ++r) X[c * kNB + r] = C<T>{static_cast<T>(r == c ? 1 : 0), 0};
        for (int r = c + 1; r < nb; ++r) {
            C<T> acc = C<T>{0, 0};
            for (int q = c; q < r; ++q) acc = gfma(acc, A[q * kNB + r], X[c * kNB + q]);
            X[c * kNB + r] = C<T>{-acc.re, -acc.im};
        }
    }
    __syncthreads();
    // M[q][c] = X[c][q] / D[c]  (c >= q), stored mws[q*kNB + c]; X kept for the solves
    C<T>* M = mws + static_cast<int64_t>(blockIdx.x) * kNB * kNB;
    C<T>* Xg = xinv + fp.xinv_off[s] + static_cast<int64_t>(b / kNB) * kNB * kNB;
    for (int idx = tid; idx < nb * nb; idx += nt) {
        const int q = idx / nb, c = idx % nb;
        M[q * kNB + c] = (c >= q) ? gmul(X[q * kNB + c], grecip(A[c * kNB + c])) : C<T>{0, 0};
        Xg[q * nb + c] = (c >= q) ? X[q * kNB + c] : C<T>{0, 0};  // column q, row c
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
    for (int idx = tid; idx < nb * 64; idx += nt) {
        const int q = idx / 64, r = idx % 64;
        Fs[q * 64 + r] = r < nr ? fr.panel[static_cast<int64_t>(b + q) * fr.m + r0 + r] : C<T>{0, 0};
    }
    for (int idx = tid; idx < nb * nb; idx += nt) {
        const int q = idx / nb, c = idx % nb;
        Ms[q * kNB + c] = M[q * kNB + c];
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

template <typename T> struct Mfma;

// Operands: lane l supplies A[l%16][l/16] and B[l/16][l%16] for both types. The
// accumulator layouts differ (measured on gfx950): register v of lane l holds
// D[l/16 + 4v][l%16] for f64 and D[4(l/16) + v][l%16] for f32.
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return lane / 16 + 4 * v; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane / 16) + v; }
};

template <typename T>
__global__ void __launch_bounds__(256) k_update(const int* __restrict__ nodes, int b, FrontPlan fp, C<T>* P, C<T>* CB) {
    using V4 = typename Mfma<T>::v4;
    constexpr int KC = 16;  // block columns staged per iteration
    __shared__ T Lr[KC][64], Li[KC][64], Wr[KC][64], Wi[KC][64];
    __shared__ C<T> Dsh[kNB];
    const int s = nodes[blockIdx.y];
    const FrontRef<T> fr = front<T>(fp, s, P, CB);
    if (fr.k <= b) return;
    const int nb = min(kNB, fr.k - b);
    const int t0 = b + nb;
    const int nrem = fr.m - t0;
    if (nrem <= 0) return;
    const int ntile = (nrem + 63) / 64;
    const int t = static_cast<int>(blockIdx.x);
    if (t >= ntile * (ntile + 1) / 2) return;
    // lower-triangular tile index -> (ti, tj), ti >= tj
    int ti = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    while (ti * (ti + 1) / 2 > t) --ti;
    const int tj = t - ti * (ti + 1) / 2;
    const int i0 = t0 + ti * 64, j0 = t0 + tj * 64;

    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;  // quadrant offsets (i, j)

    for (int q = tid; q < nb; q += 256) Dsh[q] = fr.panel[static_cast<int64_t>(b + q) * fr.m + (b + q)];

    V4 accr[2][2], acci[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            accr[x][y] = V4{0, 0, 0, 0};
            acci[x][y] = V4{0, 0, 0, 0};
        }

    for (int kk = 0; kk < nb; kk += KC) {
        __syncthreads();
        // stage L rows j0.. and W rows i0.. for columns b+kk .. b+kk+KC
        for (int idx = tid; idx < KC * 64; idx += 256) {
            const int q = idx / 64, r = idx % 64;
            const int c = kk + q;
            C<T> lj = C<T>{0, 0}, li = C<T>{0, 0};
            if (c < nb) {
                const C<T>* col = fr.panel + static_cast<int64_t>(b + c) * fr.m;
                if (j0 + r < fr.m) lj = col[j0 + r];
                if (i0 + r < fr.m) li = gmul(col[i0 + r], Dsh[c]);
            }
            Lr[q][r] = lj.re;
            Li[q][r] = lj.im;
            Wr[q][r] = li.re;
            Wi[q][r] = li.im;
        }
        __syncthreads();
#pragma unroll
        for (int k4 = 0; k4 < KC; k4 += 4) {
            const int kq = k4 + lane / 16;
            T ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                ar[y] = Lr[kq][wj + y * 16 + (lane % 16)];
                ai[y] = Li[kq][wj + y * 16 + (lane % 16)];
            }
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                br[x] = Wr[kq][wi + x * 16 + (lane % 16)];
                bi[x] = Wi[kq][wi + x * 16 + (lane % 16)];
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    accr[x][y] = Mfma<T>::mac(ar[y], br[x], accr[x][y]);
                    accr[x][y] = Mfma<T>::mac(-ai[y], bi[x], accr[x][y]);
                    acci[x][y] = Mfma<T>::mac(ar[y], bi[x], acci[x][y]);
                    acci[x][y] = Mfma<T>::mac(ai[y], br[x], acci[x][y]);
                }
        }
    }
    // D[p][q]: lane holds p = Mfma<T>::row(lane, v) (j), q = lane%16 (i)
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wi + x * 16 + (lane % 16);
                const int j = j0 + wj + y * 16 + Mfma<T>::row(lane, v);
                if (i < fr.m && j < fr.m && i >= j) {
                    C<T>* dst = fr.at(i, j);
                    C<T> cur = *dst;
                    cur.re -= accr[x][y][v];
                    cur.im -= acci[x][y][v];
                    *dst = cur;
                }
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
                       const FrontPlan& fp, void* P, void* CB) {
    if (count <= 0 || total_blocks <= 0) return;
    hipLaunchKernelGGL(k_extend_add<T>, dim3(total_blocks), dim3(256), 0, st, nodes, blk_start, count, fp,
                       static_cast<C<T>*>(P), static_cast<C<T>*>(CB));
}

template <typename T>
void launch_block_step(hipStream_t st, const int* nodes, int count, int b, int max_row_blocks, int max_tiles,
                       const FrontPlan& fp, void* P, void* CB, void* mws, void* xinv, double pivot_floor,
                       FactorStatsDev* stats) {
    const size_t diag_lds = static_cast<size_t>(2 * kNB * kNB + kNB) * sizeof(C<T>);
    allow_large_lds(k_diag<T>, diag_lds);
    hipLaunchKernelGGL(k_diag<T>, dim3(count), dim3(256), diag_lds, st, nodes, b, fp, static_cast<C<T>*>(P),
                       static_cast<C<T>*>(CB), static_cast<C<T>*>(mws), static_cast<C<T>*>(xinv), pivot_floor, stats);
    if (max_row_blocks > 0) {
        const size_t panel_lds = static_cast<size_t>(64 * kNB + kNB * kNB) * sizeof(C<T>);
        allow_large_lds(k_panel<T>, panel_lds);
        hipLaunchKernelGGL(k_panel<T>, dim3(max_row_blocks, count), dim3(256), panel_lds, st, nodes, b, fp,
                           static_cast<C<T>*>(P), static_cast<C<T>*>(CB), static_cast<const C<T>*>(mws), stats);
    }
    if (max_tiles > 0) launch_update<T>(st, nodes, count, b, max_tiles, fp, P, CB);
}

template <typename T>
void launch_update(hipStream_t st, const int* nodes, int count, int b, int max_tiles, const FrontPlan& fp, void* P,
                   void* CB) {
    hipLaunchKernelGGL(k_update<T>, dim3(max_tiles, count), dim3(256), 0, st, nodes, b, fp, static_cast<C<T>*>(P),
                       static_cast<C<T>*>(CB));
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
template void launch_extend_add<double>(hipStream_t, const int*, const int*, int, int, const FrontPlan&, void*, void*);
template void launch_extend_add<float>(hipStream_t, const int*, const int*, int, int, const FrontPlan&, void*, void*);
template void launch_update<double>(hipStream_t, const int*, int, int, int, const FrontPlan&, void*, void*);
template void launch_update<float>(hipStream_t, const int*, int, int, int, const FrontPlan&, void*, void*);
template void launch_block_step<double>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, void*, void*,
                                        void*, void*, double, FactorStatsDev*);
template void launch_block_step<float>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, void*, void*,
                                       void*, void*, double, FactorStatsDev*);

}  // namespace aphi_dev

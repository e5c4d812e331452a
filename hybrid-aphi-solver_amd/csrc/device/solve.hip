// Multi-RHS supernodal triangular solves (src/factorization.cpp:254-313, organised
// by fronts): forward through L (unit), divide by D, backward through L^T -- the
// transpose, never the conjugate transpose.
//
// Forward, leaves to root: each supernode's front vector W_s (m x nrhs) takes its
// own right-hand side rows plus its children's update vectors (children in order,
// so the sums are deterministic), solves its unit-lower diagonal part and leaves
// W_s[k:m] - L21 y as its own update vector. Backward, root to leaves: x_s =
// L11^-T (D^-1 y_s - L21^T x[rows beyond]), where the rows beyond belong to
// ancestors that are already solved. Each supernode goes in blocks of kSB
// columns whose diagonal-block inverses are formed once after the
// factorization, so a block costs two (forward) or three (backward) parallel
// launches instead of a chain of substitutions.
//
// Vectors are kept in groups of kQ = 16 right-hand sides, interleaved: entry
// (row i, right-hand side q) of group g at [(g * rows + i) * 16 + q], with rows = n
// for the permuted right-hand side, y and x, and the total front row count for
// the front vectors W. A 16-row by 16-column tile of any of them is then one
// contiguous 16-complex segment per row, which is the MFMA operand shape: every
// step kernel is a set of 16 x 16 complex tiles, one wave per (tile, 64-wide
// slice of the inner dimension), the slices summed in LDS in a fixed order.
#include <algorithm>

#include "common.hpp"
#include "factor.hpp"

namespace aphi_dev {

namespace {

template <typename T>
using C = typename cplx<T>::type;

struct Sn {
    int m, k, first;
    int64_t rp;
};

__device__ __forceinline__ Sn sn(const FrontPlan& fp, int s) {
    Sn x;
    x.rp = fp.row_ptr[s];
    x.m = static_cast<int>(fp.row_ptr[s + 1] - x.rp);
    x.first = fp.first[s];
    x.k = fp.first[s + 1] - x.first;
    return x;
}

template <typename T>
using V4 = typename Mfma<T>::v4;

// One wave: acc += A (16 x 64) B (64 x 16) for one 64-wide slice of the inner
// dimension, all 32 operand loads issued before the first MFMA. Lane l supplies
// A[l%16][l/16 + 4u] and B[l/16 + 4u][l%16], u = 0..15; a complex product is four
// real MFMAs.
template <typename T, typename FA, typename FB>
__device__ __forceinline__ void cmma64(V4<T>& accr, V4<T>& acci, FA a_at, FB b_at) {
    const int lane = static_cast<int>(threadIdx.x % 64);
    const int p = lane % 16, kq = lane / 16;
    C<T> a[16], b[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) a[u] = a_at(p, 4 * u + kq);
#pragma unroll
    for (int u = 0; u < 16; ++u) b[u] = b_at(4 * u + kq, p);
    __builtin_amdgcn_sched_barrier(0);  // every operand load in flight before the first MFMA
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        accr = Mfma<T>::mac(a[u].re, b[u].re, accr);
        accr = Mfma<T>::mac(-a[u].im, b[u].im, accr);
        acci = Mfma<T>::mac(a[u].re, b[u].im, acci);
        acci = Mfma<T>::mac(a[u].im, b[u].re, acci);
    }
}

// The four waves' partial 16 x 16 tiles summed in wave order; out(p, q, v) is
// called by thread (p = tid / 16, q = tid % 16) with the sum.
template <typename T, typename FO>
__device__ __forceinline__ void reduce4(C<T> (*part)[256], const V4<T>& accr, const V4<T>& acci, FO out) {
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
#pragma unroll
    for (int v = 0; v < 4; ++v) part[wave][Mfma<T>::row(lane, v) * 16 + lane % 16] = C<T>{accr[v], acci[v]};
    __syncthreads();
    C<T> t = part[0][tid];
    t = gadd(t, part[1][tid]);
    t = gadd(t, part[2][tid]);
    t = gadd(t, part[3][tid]);
    out(tid / 16, tid % 16, t);
}

// reduce4 for NW = 4 waves; a one-wave workgroup (the levels of many small
// fronts, where one wave covers the whole inner dimension) emits each lane's own
// four entries instead.
template <typename T, int NW, typename FO>
__device__ __forceinline__ void reduce_waves(C<T> (*part)[256], const V4<T>& accr, const V4<T>& acci, FO out) {
    static_assert(NW == 1 || NW == 4, "one or four waves");
    if constexpr (NW == 1) {
        const int lane = static_cast<int>(threadIdx.x % 64);
#pragma unroll
        for (int v = 0; v < 4; ++v) out(Mfma<T>::row(lane, v), lane % 16, C<T>{accr[v], acci[v]});
    } else {
        reduce4<T>(part, accr, acci, out);
    }
}

}  // namespace

// b in the caller's layout (column r at b + r * ldb) -> permuted, equilibrated,
// column-scaled and interleaved; one thread per (row, right-hand side)
template <typename T>
__global__ void __launch_bounds__(256) k_perm_in(int n, int nrhs, const int* __restrict__ perm,
                                                 const double* __restrict__ d, const cd* __restrict__ b, int64_t ldb,
                                                 const double* __restrict__ colscale, C<T>* __restrict__ out) {
    const int i = static_cast<int>(blockIdx.x) * 16 + static_cast<int>(threadIdx.x) / kQ;
    const int q = static_cast<int>(threadIdx.x) % kQ, g = static_cast<int>(blockIdx.y);
    const int r = g * kQ + q;
    if (i >= n) return;
    C<T> o = C<T>{0, 0};
    if (r < nrhs) {
        const int p = perm[i];
        cd v = cmul(b[r * ldb + p], cd{d[p], 0.0});
        const double s = colscale[r];
        if (s != 1.0) v = cd{v.re * s, v.im * s};
        o = C<T>{static_cast<T>(v.re), static_cast<T>(v.im)};
    }
    out[(static_cast<int64_t>(g) * n + i) * kQ + q] = o;
}

template <typename T>
__global__ void __launch_bounds__(256) k_perm_out(int n, int nrhs, const int* __restrict__ perm,
                                                  const double* __restrict__ d, const C<T>* __restrict__ xt,
                                                  const double* __restrict__ colscale, cd* __restrict__ x, int64_t ldx,
                                                  int accumulate) {
    const int i = static_cast<int>(blockIdx.x) * 16 + static_cast<int>(threadIdx.x) / kQ;
    const int q = static_cast<int>(threadIdx.x) % kQ, g = static_cast<int>(blockIdx.y);
    const int r = g * kQ + q;
    if (i >= n || r >= nrhs) return;
    const int p = perm[i];
    const C<T> t = xt[(static_cast<int64_t>(g) * n + i) * kQ + q];
    cd v = cd{static_cast<double>(t.re), static_cast<double>(t.im)};
    const double s = colscale[r];
    if (s != 1.0) v = cd{v.re / s, v.im / s};
    v = cmul(v, cd{d[p], 0.0});
    x[r * ldx + p] = accumulate ? cadd(x[r * ldx + p], v) : v;
}

// Rows [r0, r0 + kGatherRows) of W_s := [b rows of s ; 0] + the children's update
// vectors, children in order, accumulated in LDS. A child's update rows land at
// increasing parent positions (rel), so the ones in this chunk are a contiguous
// range, read from the extend-add bounds for fronts of more than one chunk.
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_gather(const int* __restrict__ nodes, FrontPlan fp,
                                                    const C<T>* __restrict__ bt, int n, int64_t R, C<T>* W) {
    extern __shared__ unsigned char smem_raw[];
    C<T>* acc = reinterpret_cast<C<T>*>(smem_raw);  // [kGatherRows][kQ]
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z);
    const Sn x = sn(fp, s);
    const int r0 = static_cast<int>(blockIdx.x) * kGatherRows;
    if (r0 >= x.m) return;
    const int rows = min(kGatherRows, x.m - r0);
    const int tid = threadIdx.x;
    for (int idx = tid; idx < rows * kQ; idx += 256) {
        const int i = r0 + idx / kQ;
        acc[idx] = i < x.k ? bt[(static_cast<int64_t>(g) * n + x.first + i) * kQ + idx % kQ] : C<T>{0, 0};
    }
    for (int c = fp.child_ptr[s]; c < fp.child_ptr[s + 1]; ++c) {
        const int ch = fp.child[c];
        const Sn y = sn(fp, ch);
        int lo = 0, hi = y.m - y.k;
        if (x.m > kGatherRows) {
            const int* bnd = fp.ea_bound + fp.ea_bound_off[ch];
            lo = bnd[r0 / 64];
            hi = bnd[min((r0 + kGatherRows) / 64, (x.m + 63) / 64)];
        }
        const int* rel = fp.rel + y.rp + y.k;
        const C<T>* Wc = W + (static_cast<int64_t>(g) * R + y.rp + y.k) * kQ;
        __syncthreads();
        for (int idx = tid; idx < (hi - lo) * kQ; idx += 256) {
            const int i = lo + idx / kQ, q = idx % kQ;
            C<T>& a = acc[(rel[i] - r0) * kQ + q];
            a = gadd(a, Wc[static_cast<int64_t>(i) * kQ + q]);
        }
    }
    __syncthreads();
    C<T>* Ws = W + (static_cast<int64_t>(g) * R + x.rp + r0) * kQ;
    for (int idx = tid; idx < rows * kQ; idx += 256) Ws[idx] = acc[idx];
}

// ------------------------------------------------------------------ diagonal-block inverses
//
// X = L_BB^-1 for every kSB-column block B of every supernode (the last one nb x
// nb), built after the factorization from the kNB-block inverses the diagonal
// steps kept: X[ib][jb] = -Xinv[ib] * sum_{kb=jb}^{ib-1} L[ib][kb] X[kb][jb] over
// kNB x kNB sub-blocks. Workgroup = (block, sub-block column jb); it walks down
// its column, reading back what it wrote (the barrier orders global memory
// within the workgroup). Stored twice: column-major (forward solve) and
// row-major (backward solve), upper blocks zero.
template <typename T>
__global__ void __launch_bounds__(256) k_inv_blocks(const int2* __restrict__ pairs, FrontPlan fp,
                                                    const C<T>* __restrict__ P, const C<T>* __restrict__ xinv,
                                                    C<T>* X, C<T>* XT) {
    extern __shared__ unsigned char smem_raw[];
    C<T>* As = reinterpret_cast<C<T>*>(smem_raw);  // 64 x 64, column-major (LD 65)
    C<T>* Bs = As + kNB * (kNB + 1);               // 64 x 64, column-major (LD 65)
    constexpr int LD = kNB + 1;
    const int2 pr = pairs[blockIdx.x];
    const int s = pr.x, B = pr.y;
    const Sn x = sn(fp, s);
    const int nb = min(kSB, x.k - B);
    const int nblk = (nb + kNB - 1) / kNB;
    const int jb = static_cast<int>(blockIdx.y);
    if (jb >= nblk) return;
    const C<T>* Ps = P + fp.panel_off[s];
    C<T>* Xs = X + fp.xsb_off[s] + static_cast<int64_t>(B / kSB) * kSB * kSB;
    C<T>* XTs = XT + fp.xsb_off[s] + static_cast<int64_t>(B / kSB) * kSB * kSB;
    const int tid = threadIdx.x;
    const int row = tid % 64, g = tid / 64;  // output row and a group of 16 output columns
    const int c0 = jb * kNB, ncj = min(kNB, nb - c0);
    // zero the blocks above the diagonal of this sub-block column
    for (int idx = tid; idx < c0 * ncj; idx += 256) {
        const int rr = idx % c0, cc = c0 + idx / c0;
        Xs[static_cast<int64_t>(cc) * nb + rr] = C<T>{0, 0};
        XTs[static_cast<int64_t>(rr) * nb + cc] = C<T>{0, 0};
    }
    for (int ib = jb; ib < nblk; ++ib) {
        const int r0 = ib * kNB, nri = min(kNB, nb - r0);
        const C<T>* Xi = xinv + fp.xinv_off[s] + static_cast<int64_t>((B + r0) / kNB) * kNB * kNB;  // nri x nri
        C<T> acc[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u] = C<T>{0, 0};
        if (ib == jb) {
            // X[jb][jb] = Xinv[jb]
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int cc = g * 16 + u;
                if (row < nri && cc < ncj) acc[u] = Xi[cc * nri + row];
            }
        } else {
            // S = sum_kb L[ib][kb] X[kb][jb]
            for (int kb = jb; kb < ib; ++kb) {
                const int k0 = kb * kNB, nk = min(kNB, nb - k0);
                __syncthreads();
                for (int idx = tid; idx < kNB * kNB; idx += 256) {
                    const int rr = idx % kNB, cc = idx / kNB;
                    As[cc * LD + rr] = (rr < nri && cc < nk) ? Ps[static_cast<int64_t>(B + k0 + cc) * x.m + B + r0 + rr]
                                                             : C<T>{0, 0};
                    Bs[cc * LD + rr] = (rr < nk && cc < ncj) ? Xs[static_cast<int64_t>(c0 + cc) * nb + k0 + rr]
                                                             : C<T>{0, 0};
                }
                __syncthreads();
                for (int q = 0; q < kNB; ++q) {
                    const C<T> a = As[q * LD + row];
#pragma unroll
                    for (int u = 0; u < 16; ++u) acc[u] = gfma(acc[u], a, Bs[(g * 16 + u) * LD + q]);
                }
            }
            // X[ib][jb] = -Xinv[ib] S
            __syncthreads();
            for (int idx = tid; idx < kNB * kNB; idx += 256) {
                const int rr = idx % kNB, cc = idx / kNB;
                As[cc * LD + rr] = (rr < nri && cc < nri) ? Xi[cc * nri + rr] : C<T>{0, 0};
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) Bs[(g * 16 + u) * LD + row] = acc[u];
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 16; ++u) acc[u] = C<T>{0, 0};
            for (int q = 0; q < kNB; ++q) {
                const C<T> a = As[q * LD + row];
#pragma unroll
                for (int u = 0; u < 16; ++u) acc[u] = gfms(acc[u], a, Bs[(g * 16 + u) * LD + q]);
            }
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int cc = g * 16 + u;
            if (row < nri && cc < ncj) {
                Xs[static_cast<int64_t>(c0 + cc) * nb + r0 + row] = acc[u];
                XTs[static_cast<int64_t>(r0 + row) * nb + c0 + cc] = acc[u];
            }
        }
        __syncthreads();  // orders this sub-block's global writes before the next reads
    }
}

// dinv[first(s) + j] = 1 / D[j] of every supernode s (one workgroup per supernode)
template <typename T>
__global__ void __launch_bounds__(64) k_dinv(FrontPlan fp, const C<T>* __restrict__ P, C<T>* __restrict__ dinv) {
    const int s = static_cast<int>(blockIdx.x);
    const Sn x = sn(fp, s);
    const C<T>* Ps = P + fp.panel_off[s];
    for (int j = threadIdx.x; j < x.k; j += 64) dinv[x.first + j] = grecip(Ps[static_cast<int64_t>(j) * x.m + j]);
}

// ------------------------------------------------------------------ forward steps
//
// Step B of every supernode with k > B: y_B = X_B w_B (X = L_BB^-1), then the
// rows below the block take W[i] -= L[i, B:B+nb] y_B.

// y rows [16 bx, 16 bx + 16) of the block: y = X w (X lower triangular, stored
// with zeros above the diagonal); wave w takes the columns 64 w + 256 j, 64 at a
// time (NW = 1: blocks of at most 64 columns)
template <typename T, int NW = 4>
__global__ void __launch_bounds__(64 * NW) k_fwd_diag(const int* __restrict__ nodes, int B, FrontPlan fp,
                                                      const C<T>* __restrict__ Xsb, const C<T>* __restrict__ W,
                                                      C<T>* __restrict__ Y, int n, int64_t R) {
    __shared__ C<T> part[NW][256];
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z);
    const Sn x = sn(fp, s);
    if (x.k <= B) return;
    const int nb = min(kSB, x.k - B);
    const int c0 = static_cast<int>(blockIdx.x) * 16;
    if (c0 >= nb) return;
    const int kmax = min(nb, c0 + 16);  // X[c][k] = 0 for k > c
    const int k0 = 64 * (static_cast<int>(threadIdx.x) / 64);
    const C<T>* Xs = Xsb + fp.xsb_off[s] + static_cast<int64_t>(B / kSB) * kSB * kSB;
    const C<T>* Ws = W + (static_cast<int64_t>(g) * R + x.rp + B) * kQ;
    V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
    for (int kc = k0; kc < kmax; kc += 64 * NW)
        cmma64<T>(accr, acci,
                  [&](int p, int k) { return Xs[static_cast<int64_t>(min(kc + k, nb - 1)) * nb + min(c0 + p, nb - 1)]; },
                  [&](int k, int q) {
                      const C<T> v = Ws[min(kc + k, kmax - 1) * kQ + q];  // clamped read, then select
                      return kc + k < kmax ? v : C<T>{0, 0};
                  });
    C<T>* Yg = Y + (static_cast<int64_t>(g) * n + x.first + B) * kQ;
    reduce_waves<T, NW>(part, accr, acci, [&](int p, int q, C<T> t) {
        if (c0 + p < nb) Yg[(c0 + p) * kQ + q] = t;
    });
}

// Rows of a forward update: the triangle rows [B + nb, k) below block B (RECT
// false: the block chain), or the rows [k, m) below the front's pivots with the
// blocks [ks bps, (ks + 1) bps) of the front (RECT true: one launch per level after
// its chain, so that the bulk of L streams in large launches instead of one small
// launch per block; the front's blocks split ksplit ways so that a level of few
// fronts still has enough workgroups)
struct FwdRows {
    int lo, hi, b0, b1;  // rows [lo, hi), columns [b0, b1) in blocks of kSB
};

template <bool RECT>
__device__ __forceinline__ bool fwd_rows(const Sn& x, int B, int ks, int bps, FwdRows& r) {
    if constexpr (RECT) {
        r = FwdRows{x.k, x.m, ks * bps * kSB, min(x.k, (ks + 1) * bps * kSB)};
        if (r.b0 >= r.b1) return false;
    } else {
        if (x.k <= B) return false;
        r = FwdRows{B + min(kSB, x.k - B), x.k, B, B + 1};
    }
    return r.lo < r.hi;
}

// Where a forward update's sums go: W -= sum directly (one split), or the split's
// partial sums part[((f * ksplit + ks) * G + g) * prow + (i - k)][q], subtracted by
// k_fwd_rect_reduce in split order
template <typename T>
struct FwdOut {
    C<T>* part;
    int ksplit, bps, prow;
};

// W[i] -= L[i, blocks] y_blocks for the 16 rows i0.. of the update rows; wave w
// takes the columns 64 w + 256 j of each block (64 at a time)
template <typename T, bool RECT>
__global__ void __launch_bounds__(256) k_fwd_update(const int* __restrict__ nodes, int B, FrontPlan fp,
                                                    const C<T>* __restrict__ P, C<T>* W, const C<T>* __restrict__ Y,
                                                    int n, int64_t R, FwdOut<T> fo) {
    __shared__ C<T> part[4][256];
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z), G = static_cast<int>(gridDim.z);
    const Sn x = sn(fp, s);
    const int tiles = static_cast<int>(gridDim.x) / fo.ksplit;
    const int ks = static_cast<int>(blockIdx.x) / tiles;
    FwdRows rr;
    if (!fwd_rows<RECT>(x, B, ks, fo.bps, rr)) return;
    const int i0 = rr.lo + (static_cast<int>(blockIdx.x) % tiles) * 16;
    if (i0 >= rr.hi) return;
    const int k0 = 64 * (static_cast<int>(threadIdx.x) / 64);
    const C<T>* Ps = P + fp.panel_off[s];
    V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
    for (int Bb = rr.b0; Bb < rr.b1; Bb += kSB) {
        const int nb = min(kSB, x.k - Bb);
        const C<T>* Yg = Y + (static_cast<int64_t>(g) * n + x.first + Bb) * kQ;
        for (int kc = k0; kc < nb; kc += 256)
            cmma64<T>(accr, acci,
                      [&](int p, int k) {
                          return Ps[static_cast<int64_t>(Bb + min(kc + k, nb - 1)) * x.m + min(i0 + p, x.m - 1)];
                      },
                      [&](int k, int q) {
                          const C<T> v = Yg[min(kc + k, nb - 1) * kQ + q];  // clamped read, then select
                          return kc + k < nb ? v : C<T>{0, 0};
                      });
    }
    C<T>* Wg = W + (static_cast<int64_t>(g) * R + x.rp) * kQ;
    C<T>* Pg = fo.ksplit > 1 ? fo.part + ((static_cast<int64_t>(blockIdx.y) * fo.ksplit + ks) * G + g) * fo.prow * kQ
                             : nullptr;  // rows i - k
    reduce4<T>(part, accr, acci, [&](int p, int q, C<T> t) {
        if (i0 + p < rr.hi) {
            if (Pg) {
                Pg[static_cast<int64_t>(i0 + p - x.k) * kQ + q] = t;
            } else {
                C<T>& dst = Wg[static_cast<int64_t>(i0 + p) * kQ + q];
                dst = gsub(dst, t);
            }
        }
    });
}

// One wave's sums of the complex-float forward update for the 32 rows i0.. over
// the columns [c0, c1) of a front (whole blocks of kSB, the last possibly short):
// wave w takes the columns [w kSB / 4, (w + 1) kSB / 4) of each block. Lane (p, kq) loads
// rows i0 + 2p and i0 + 2p + 1 of a column with one 16-byte load (each load
// instruction covers 256 contiguous bytes of four columns), the two rows going to
// two 16-row MFMA tiles (even rows r0/i0a, odd rows r1/i1a; tile row r of tile h
// is front row i0 + 2 r + h). A wave's operands for its next 32 columns are loaded
// before it multiplies the current ones.
template <int NW>
__device__ __forceinline__ void fwd2_sums(const Sn& x, const cf* __restrict__ Ps, const cf* __restrict__ Yg, int i0,
                                          int c0, int c1, V4<float>& r0, V4<float>& i0a, V4<float>& r1,
                                          V4<float>& i1a) {
    using VF = V4<float>;
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int tid = static_cast<int>(threadIdx.x), wave = tid / 64, lane = tid % 64;
    // a block's columns split into four wave slices of kSB / 4, each in stages of 32
    // (one-wave workgroups: blocks of at most 64 columns, in the first slice)
    constexpr int SPB = kSB / 128;  // stages per wave slice
    const int p = lane % 16, kq = lane / 16, k0 = (kSB / 4) * wave;
    // three-product complex form (k1 = (ar + ai) br, k2 = ar (bi - br), k3 = ai (br + bi);
    // re = k1 - k3, im = k1 + k2): three MFMAs per complex product instead of four
    VF s1a = VF{0, 0, 0, 0}, s2a = VF{0, 0, 0, 0}, s3a = VF{0, 0, 0, 0};
    VF s1b = VF{0, 0, 0, 0}, s2b = VF{0, 0, 0, 0}, s3b = VF{0, 0, 0, 0};
    const int row = min(i0 + 2 * p, x.m - 1);  // rows row and row + 1 (the second may lie past m: never stored)
    // The wave's columns in stages of 32: stage j is the k steps 8 (j % SPB) ..
    // 8 (j % SPB) + 7 of its slice of block c0 + (j / SPB) kSB; steps(j) is how
    // many of them lie inside the block (uniform over the wave)
    const int nstage = SPB * ((c1 - c0 + kSB - 1) / kSB);
    const auto steps = [&](int j) {
        if (j >= nstage) return 0;
        const int Bb = c0 + (j / SPB) * kSB, u0 = 8 * (j % SPB);
        return max(0, min(8, (min(kSB, x.k - Bb) - k0 - 4 * u0 + 3) / 4));
    };
    // Four-wave workgroups load every step of a stage (clamped columns past the
    // block or past the last stage), so that the loads in flight are known at
    // compile time and a stage waits only for its own; one-wave workgroups (blocks
    // of at most 64 columns) load only the steps they use.
    const auto load = [&](int j, int nu, f4 (&a)[8], cf (&b)[8]) {
        if constexpr (NW > 1) j = min(j, nstage - 1);
        const int Bb = c0 + (j / SPB) * kSB, kb = k0 + 32 * (j % SPB);
        const int last = min(kSB, x.k - Bb) - 1;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (NW > 1 || u < nu)
                a[u] = *reinterpret_cast<const f4*>(Ps + static_cast<int64_t>(Bb + min(kb + 4 * u + kq, last)) * x.m + row);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (NW > 1 || u < nu) {
                const int k = kb + 4 * u + kq;
                const cf v = Yg[(Bb + min(k, last)) * kQ + p];  // clamped read, then select
                b[u] = k <= last ? v : cf{0, 0};
            }
    };
    const auto mma = [&](int nu, const f4 (&a)[8], const cf (&b)[8]) {
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (u < nu) {
                const float bd = b[u].im - b[u].re, bs = b[u].re + b[u].im;
                s1a = Mfma<float>::mac(a[u].x + a[u].y, b[u].re, s1a);
                s2a = Mfma<float>::mac(a[u].x, bd, s2a);
                s3a = Mfma<float>::mac(a[u].y, bs, s3a);
                s1b = Mfma<float>::mac(a[u].z + a[u].w, b[u].re, s1b);
                s2b = Mfma<float>::mac(a[u].z, bd, s2b);
                s3b = Mfma<float>::mac(a[u].w, bs, s3b);
            }
    };
    f4 a0[8], a1[8];
    cf b0[8], b1[8];
    int nu0 = steps(0);
    load(0, nu0, a0, b0);
    for (int j = 0; j < nstage; j += 2) {
        const int nu1 = steps(j + 1);
        load(j + 1, nu1, a1, b1);
        __builtin_amdgcn_sched_barrier(0);  // the next stage's loads in flight before these MFMAs
        mma(nu0, a0, b0);
        nu0 = steps(j + 2);
        load(j + 2, nu0, a0, b0);
        __builtin_amdgcn_sched_barrier(0);
        mma(nu1, a1, b1);
    }
    r0 = s1a - s3a;
    i0a = s1a + s2a;
    r1 = s1b - s3b;
    i1a = s1b + s2b;
}

// The same update for complex64 with 32 rows per workgroup: lane (p, kq) loads
// rows i0 + 2p and i0 + 2p + 1 of a column with one 16-byte load (each load
// instruction covers 256 contiguous bytes of four columns, twice the run of the
// one-row-per-lane mapping), and the two rows go to two 16-row MFMA tiles (even
// rows, odd rows). A wave's operands for its slice of block b + 1 are loaded
// before it multiplies those of block b, and a slice narrower than 64 columns
// issues only the steps its columns need. NW = 1: blocks of at most 64 columns,
// each lane updating its own eight entries.
template <int NW, bool RECT>
__global__ void __launch_bounds__(64 * NW) k_fwd_update2(const int* __restrict__ nodes, int B, FrontPlan fp,
                                                         const cf* __restrict__ P, cf* W, const cf* __restrict__ Y, int n,
                                                         int64_t R, FwdOut<float> fo) {
    using VF = V4<float>;
    typedef float f4 __attribute__((ext_vector_type(4)));
    __shared__ cf part[2][NW][256];
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z), G = static_cast<int>(gridDim.z);
    const Sn x = sn(fp, s);
    const int tiles = static_cast<int>(gridDim.x) / fo.ksplit;
    const int ks = static_cast<int>(blockIdx.x) / tiles;
    FwdRows rr;
    if (!fwd_rows<RECT>(x, B, ks, fo.bps, rr)) return;
    const int i0 = rr.lo + (static_cast<int>(blockIdx.x) % tiles) * 32;
    if (i0 >= rr.hi) return;
    const int tid = static_cast<int>(threadIdx.x), wave = tid / 64, lane = tid % 64, p = lane % 16;
    const cf* Ps = P + fp.panel_off[s];
    const cf* Yg = Y + (static_cast<int64_t>(g) * n + x.first) * kQ;
    VF r0, i0a, r1, i1a;
    fwd2_sums<NW>(x, Ps, Yg, i0, rr.b0, rr.b1, r0, i0a, r1, i1a);
    cf* Wg = W + (static_cast<int64_t>(g) * R + x.rp) * kQ;
    cf* Pg = fo.ksplit > 1 ? fo.part + ((static_cast<int64_t>(blockIdx.y) * fo.ksplit + ks) * G + g) * fo.prow * kQ
                           : nullptr;  // rows i - k
    if constexpr (NW == 1) {
        // tile row r of tile h is front row i0 + 2 r + h
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = i0 + 2 * Mfma<float>::row(lane, v) + h;
                if (i < rr.hi) {
                    const cf t = h ? cf{r1[v], i1a[v]} : cf{r0[v], i0a[v]};
                    if (Pg) {
                        Pg[static_cast<int64_t>(i - x.k) * kQ + p] = t;
                    } else {
                        cf& dst = Wg[static_cast<int64_t>(i) * kQ + p];
                        dst = gsub(dst, t);
                    }
                }
            }
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            part[0][wave][Mfma<float>::row(lane, v) * 16 + p] = cf{r0[v], i0a[v]};
            part[1][wave][Mfma<float>::row(lane, v) * 16 + p] = cf{r1[v], i1a[v]};
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            cf t = part[h][0][tid];
            t = gadd(t, part[h][1][tid]);
            t = gadd(t, part[h][2][tid]);
            t = gadd(t, part[h][3][tid]);
            const int i = i0 + 2 * (tid / 16) + h;  // tile row tid / 16 of tile h is front row i0 + 2 (tid / 16) + h
            if (i < rr.hi) {
                if (Pg) {
                    Pg[static_cast<int64_t>(i - x.k) * kQ + tid % 16] = t;
                } else {
                    cf& dst = Wg[static_cast<int64_t>(i) * kQ + tid % 16];
                    dst = gsub(dst, t);
                }
            }
        }
    }
}

// W[k + r] -= the forward splits' partial sums of row r, in split order; one thread
// per (row, right-hand side)
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_rect_reduce(const int* __restrict__ nodes, FrontPlan fp, C<T>* W,
                                                         int64_t R, FwdOut<T> fo) {
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z), G = static_cast<int>(gridDim.z);
    const Sn x = sn(fp, s);
    const int idx = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    const int r = idx / kQ, q = idx % kQ;
    if (r >= x.m - x.k) return;
    const int nsplit = ((x.k + kSB - 1) / kSB + fo.bps - 1) / fo.bps;  // splits that hold blocks of this front
    const C<T>* pp = fo.part + (static_cast<int64_t>(blockIdx.y) * fo.ksplit * G + g) * fo.prow * kQ +
                     static_cast<int64_t>(r) * kQ + q;
    const int64_t sstride = static_cast<int64_t>(G) * fo.prow * kQ;
    C<T> t = pp[0];
    for (int ks = 1; ks < nsplit; ++ks) t = gadd(t, pp[ks * sstride]);
    C<T>& dst = W[(static_cast<int64_t>(g) * R + x.rp + x.k + r) * kQ + q];
    dst = gsub(dst, t);
}

// ------------------------------------------------------------------ backward steps
//
// Per level, first the sums over the rows below the pivots, t = L[k:m, 0:k]^T
// x[rows k..m), which need only the ancestors' x: partial sums over kRectRows rows
// per workgroup, reduced in chunk order into the level's t (column stride rk).
// Then the block chain, blocks in decreasing order: x_B = X_B^T (D_B^-1 y_B - t_B),
// and right after it the columns before the block take the rows of block B:
// t[0:B] += L[B:B+nb, 0:B]^T x_B. Each entry of t therefore sums its rows' terms in
// a fixed order: the rows below the pivots by chunk, then the blocks from the
// last to its own.

// Rows and columns of a backward product: the rows [k, m) below the pivots with
// all k columns, in chunks of kRectRows rows (RECT), or the rows of block B with
// the columns [0, B) before it, in one chunk (the chain's update)
struct BwdRows {
    int lo, hi, chunk_rows, ncol;
};

template <bool RECT>
__device__ __forceinline__ bool bwd_rows(const Sn& x, int B, BwdRows& r) {
    if constexpr (RECT) {
        r = BwdRows{x.k, x.m, kRectRows, x.k};
    } else {
        if (x.k <= B) return false;
        r = BwdRows{B, B + min(kSB, x.k - B), kSB, B};
    }
    return r.lo < r.hi && r.ncol > 0;
}

// out[c][q] for the 16 columns c0.. over the rows [i0, i0 + chunk_rows) of the
// product's rows (wave w takes rows i0 + 64 w + 256 j): RECT writes
// partial[((f * stride + chunk) * G + g)][c][q], the chain's update adds to
// t[(f * G + g)][c][q] (column stride cstride both)
template <typename T, bool RECT>
__global__ void __launch_bounds__(256) k_bwd_partial(const int* __restrict__ nodes, int B, int stride, int ctiles,
                                                     int cstride, FrontPlan fp, const C<T>* __restrict__ P,
                                                     const C<T>* __restrict__ X, int n, C<T>* __restrict__ out) {
    __shared__ C<T> part[4][256];
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z), G = static_cast<int>(gridDim.z);
    const Sn x = sn(fp, s);
    BwdRows rr;
    if (!bwd_rows<RECT>(x, B, rr)) return;
    const int chunk = static_cast<int>(blockIdx.x) / ctiles, c0 = (static_cast<int>(blockIdx.x) % ctiles) * 16;
    const int i0 = rr.lo + chunk * rr.chunk_rows;
    if (i0 >= rr.hi || c0 >= rr.ncol) return;
    const int iend = min(i0 + rr.chunk_rows, rr.hi);
    const C<T>* Ps = P + fp.panel_off[s];
    const int* rows = fp.rows + x.rp;
    const C<T>* Xg = X + static_cast<int64_t>(g) * n * kQ;
    const int lane = static_cast<int>(threadIdx.x % 64), p = lane % 16, kq = lane / 16;
    const C<T>* col = Ps + static_cast<int64_t>(min(c0 + p, rr.ncol - 1)) * x.m;
    V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
    for (int iw = i0 + 64 * (static_cast<int>(threadIdx.x) / 64); iw < iend; iw += 256) {
        // the row indices first (one wait), then every operand load before the MFMAs
        int gr[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) gr[u] = rows[min(iw + 4 * u + kq, x.m - 1)];
        C<T> a[16], b[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) a[u] = col[min(iw + 4 * u + kq, x.m - 1)];
#pragma unroll
        for (int u = 0; u < 16; ++u) b[u] = Xg[static_cast<int64_t>(gr[u]) * kQ + p];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const C<T> bb = iw + 4 * u + kq < iend ? b[u] : C<T>{0, 0};
            accr = Mfma<T>::mac(a[u].re, bb.re, accr);
            accr = Mfma<T>::mac(-a[u].im, bb.im, accr);
            acci = Mfma<T>::mac(a[u].re, bb.im, acci);
            acci = Mfma<T>::mac(a[u].im, bb.re, acci);
        }
    }
    if constexpr (RECT) {
        C<T>* o = out + ((static_cast<int64_t>(blockIdx.y) * stride + chunk) * G + g) * cstride * kQ;
        reduce4<T>(part, accr, acci, [&](int pp, int q, C<T> t) {
            if (c0 + pp < rr.ncol) o[(c0 + pp) * kQ + q] = t;
        });
    } else {
        C<T>* o = out + (static_cast<int64_t>(blockIdx.y) * G + g) * cstride * kQ;
        reduce4<T>(part, accr, acci, [&](int pp, int q, C<T> t) {
            if (c0 + pp < rr.ncol) o[(c0 + pp) * kQ + q] = gadd(o[(c0 + pp) * kQ + q], t);
        });
    }
}

// The same products for complex64 with one 16-byte load per lane for two
// adjacent rows of a column (eight contiguous rows of each of 16 columns per load
// instruction instead of four): within each group of eight rows, the even rows
// form one MFMA k-group and the odd rows the next. The chunk's row indices are
// staged in LDS first, so that the x loads they feed do not wait on the memory
// counter, and each wave loads its next 32 rows' operands before it multiplies the
// current ones. NW = 1: one wave walks the chunk 64 rows at a time (levels whose
// fronts have at most 64 such rows).
template <int NW, bool RECT>
__global__ void __launch_bounds__(64 * NW) k_bwd_partial2(const int* __restrict__ nodes, int B, int stride, int ctiles,
                                                          int cstride, FrontPlan fp, const cf* __restrict__ P,
                                                          const cf* __restrict__ X, int n, cf* __restrict__ out) {
    using VF = V4<float>;
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int kIdx = RECT ? kRectRows : kSB;  // rows of a chunk
    __shared__ cf part[NW][256];
    __shared__ int sidx[kIdx];
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z), G = static_cast<int>(gridDim.z);
    const Sn x = sn(fp, s);
    BwdRows rr;
    if (!bwd_rows<RECT>(x, B, rr)) return;
    const int chunk = static_cast<int>(blockIdx.x) / ctiles, c0 = (static_cast<int>(blockIdx.x) % ctiles) * 16;
    const int i0 = rr.lo + chunk * rr.chunk_rows;
    if (i0 >= rr.hi || c0 >= rr.ncol) return;
    const int iend = min(i0 + rr.chunk_rows, rr.hi);
    const cf* Ps = P + fp.panel_off[s];
    const int* rows = fp.rows + x.rp;
    const cf* Xg = X + static_cast<int64_t>(g) * n * kQ;
    const int lane = static_cast<int>(threadIdx.x % 64), p = lane % 16, kq = lane / 16;
    const cf* col = Ps + static_cast<int64_t>(min(c0 + p, rr.ncol - 1)) * x.m;
    // the chunk's row indices (positions past its end repeat the last one)
    for (int i = static_cast<int>(threadIdx.x); i < kIdx; i += 64 * NW) sidx[i] = rows[min(i0 + i, iend - 1)];
    __syncthreads();
    // three-product complex form, as in k_fwd_update2
    VF s1 = VF{0, 0, 0, 0}, s2 = VF{0, 0, 0, 0}, s3 = VF{0, 0, 0, 0};
    // The wave's rows in stages of 32: rows iw .. iw + 31 (four groups of eight);
    // steps(iw) is how many groups hold rows of the product (uniform over the
    // wave). Four-wave workgroups load every group (clamped rows), so that the loads
    // in flight are known at compile time; one-wave workgroups load only the
    // groups they use.
    const auto steps = [&](int iw) { return iw < iend ? min(4, (iend - iw + 7) / 8) : 0; };
    const auto load = [&](int iw, int nu, f4 (&a)[4], cf (&b)[8]) {
#pragma unroll
        for (int u = 0; u < 4; ++u)  // the second row of the pair may lie past the rows: its product is masked
            if (NW > 1 || u < nu) a[u] = *reinterpret_cast<const f4*>(col + min(iw + 8 * u + 2 * kq, iend - 1));
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (NW > 1 || u < nu) {
                const int r = min(iw + 8 * u + 2 * kq - i0, kIdx - 2);
                b[2 * u] = Xg[static_cast<int64_t>(sidx[r]) * kQ + p];
                b[2 * u + 1] = Xg[static_cast<int64_t>(sidx[r + 1]) * kQ + p];
            }
    };
    const auto mma = [&](int iw, int nu, const f4 (&a)[4], const cf (&b)[8]) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < nu) {
                const int r = iw + 8 * u + 2 * kq;
                const cf b0 = r < iend ? b[2 * u] : cf{0, 0};
                const cf b1 = r + 1 < iend ? b[2 * u + 1] : cf{0, 0};
                s1 = Mfma<float>::mac(a[u].x + a[u].y, b0.re, s1);
                s2 = Mfma<float>::mac(a[u].x, b0.im - b0.re, s2);
                s3 = Mfma<float>::mac(a[u].y, b0.re + b0.im, s3);
                s1 = Mfma<float>::mac(a[u].z + a[u].w, b1.re, s1);
                s2 = Mfma<float>::mac(a[u].z, b1.im - b1.re, s2);
                s3 = Mfma<float>::mac(a[u].w, b1.re + b1.im, s3);
            }
    };
    // stage s of a wave: rows i0 + 64 w + 256 (s / 2) * NW / 4 ... in order: the
    // wave's 64 rows of each pass, first half then second half
    const int w = static_cast<int>(threadIdx.x) / 64;
    const auto stage_row = [&](int st) { return i0 + 64 * w + 64 * NW * (st / 2) + 32 * (st % 2); };
    f4 a0[4], a1[4];
    cf b0[8], b1[8];
    int st = 0;
    int nu0 = steps(stage_row(0));
    load(stage_row(0), nu0, a0, b0);
    for (; stage_row(st) < iend; st += 2) {
        const int r0 = stage_row(st), r1 = stage_row(st + 1), r2 = stage_row(st + 2);
        const int nu1 = steps(r1);
        load(r1, nu1, a1, b1);
        __builtin_amdgcn_sched_barrier(0);  // the next rows' loads in flight before these MFMAs
        mma(r0, nu0, a0, b0);
        nu0 = steps(r2);
        load(r2, nu0, a0, b0);
        __builtin_amdgcn_sched_barrier(0);
        mma(r1, nu1, a1, b1);
    }
    const VF accr = s1 - s3, acci = s1 + s2;
    if constexpr (RECT) {
        cf* o = out + ((static_cast<int64_t>(blockIdx.y) * stride + chunk) * G + g) * cstride * kQ;
        reduce_waves<float, NW>(part, accr, acci, [&](int pp, int q, cf t) {
            if (c0 + pp < rr.ncol) o[(c0 + pp) * kQ + q] = t;
        });
    } else {
        cf* o = out + (static_cast<int64_t>(blockIdx.y) * G + g) * cstride * kQ;
        reduce_waves<float, NW>(part, accr, acci, [&](int pp, int q, cf t) {
            if (c0 + pp < rr.ncol) o[(c0 + pp) * kQ + q] = gadd(o[(c0 + pp) * kQ + q], t);
        });
    }
}

// t[(f * G + g)][c][q] (column stride cstride) = sum over the chunks of the rows
// below the pivots, for every column c < k (zero for a front without such rows),
// one column per workgroup: 16 summers per right-hand side, combined in a fixed
// order
template <typename T>
__global__ void __launch_bounds__(256) k_rect_reduce(const int* __restrict__ nodes, int stride, int cstride,
                                                     FrontPlan fp, const C<T>* __restrict__ partial,
                                                     C<T>* __restrict__ t) {
    __shared__ C<T> red[16][17];
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z), G = static_cast<int>(gridDim.z);
    const Sn x = sn(fp, s);
    const int c = static_cast<int>(blockIdx.x);
    if (c >= x.k) return;
    const int nch = x.m > x.k ? (x.m - x.k + kRectRows - 1) / kRectRows : 0;
    const int tid = threadIdx.x, q = tid % 16, p = tid / 16;
    const int64_t chstride = static_cast<int64_t>(G) * cstride * kQ;
    const C<T>* pp = partial + (static_cast<int64_t>(blockIdx.y) * stride * G + g) * cstride * kQ + c * kQ + q;
    C<T> v = C<T>{0, 0};
#pragma unroll 8
    for (int ch = p; ch < nch; ch += 16) v = gadd(v, pp[ch * chstride]);
    red[p][q] = v;
    __syncthreads();
    if (p == 0) {
        C<T> w = red[0][q];
        for (int j = 1; j < 16; ++j) w = gadd(w, red[j][q]);
        t[((static_cast<int64_t>(blockIdx.y) * G + g) * cstride + c) * kQ + q] = w;
    }
}

// x rows [16 bx, 16 bx + 16) of the block: x[c] = sum_{k >= c} X[k][c] z[k],
// z = y D^-1 - t[B + k] (t column stride rk; read for every front when all is set,
// else only for fronts with rows below their pivots, zero for the others or when t
// is null); D^-1 from k_dinv; wave w takes k in [c0 + 64 w + 256 j, + 64) (NW = 1:
// blocks of at most 64 columns)
template <typename T, int NW = 4>
__global__ void __launch_bounds__(64 * NW) k_bwd_diag(const int* __restrict__ nodes, int B, FrontPlan fp,
                                                      const C<T>* __restrict__ dinv, const C<T>* __restrict__ XTsb,
                                                      const C<T>* __restrict__ Y, const C<T>* __restrict__ tl, int rk,
                                                      int all, C<T>* __restrict__ X, int n) {
    __shared__ C<T> part[NW][256];
    const int s = nodes[blockIdx.y];
    const int g = static_cast<int>(blockIdx.z), G = static_cast<int>(gridDim.z);
    const Sn x = sn(fp, s);
    if (x.k <= B) return;
    const int nb = min(kSB, x.k - B);
    const int c0 = static_cast<int>(blockIdx.x) * 16;
    if (c0 >= nb) return;
    const int k0 = c0 + 64 * (static_cast<int>(threadIdx.x) / 64);
    const bool has_t = tl != nullptr && (all || x.m > x.k);
    const C<T>* dv = dinv + x.first + B;
    const C<T>* XTs = XTsb + fp.xsb_off[s] + static_cast<int64_t>(B / kSB) * kSB * kSB;
    const C<T>* Yg = Y + (static_cast<int64_t>(g) * n + x.first + B) * kQ;
    const C<T>* tg = has_t ? tl + ((static_cast<int64_t>(blockIdx.y) * G + g) * rk + B) * kQ : nullptr;
    V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
    for (int kc = k0; kc < nb; kc += 64 * NW) {
        // every load (clamped indices) issued before the first wait, then z, then the MFMAs
        const int lane = static_cast<int>(threadIdx.x % 64), p = lane % 16, kq = lane / 16;
        C<T> a[16], yv[16], dw[16], tv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int kk = min(kc + 4 * u + kq, nb - 1);
            a[u] = XTs[static_cast<int64_t>(kk) * nb + min(c0 + p, nb - 1)];
            yv[u] = Yg[kk * kQ + p];
            dw[u] = dv[kk];
            tv[u] = has_t ? tg[kk * kQ + p] : C<T>{0, 0};
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const C<T> z = gsub(gmul(yv[u], dw[u]), tv[u]);
            const C<T> b = kc + 4 * u + kq < nb ? z : C<T>{0, 0};
            accr = Mfma<T>::mac(a[u].re, b.re, accr);
            accr = Mfma<T>::mac(-a[u].im, b.im, accr);
            acci = Mfma<T>::mac(a[u].re, b.im, acci);
            acci = Mfma<T>::mac(a[u].im, b.re, acci);
        }
    }
    C<T>* Xg = X + (static_cast<int64_t>(g) * n + x.first + B) * kQ;
    reduce_waves<T, NW>(part, accr, acci, [&](int p, int q, C<T> t) {
        if (c0 + p < nb) Xg[(c0 + p) * kQ + q] = t;
    });
}

// ------------------------------------------------------------------ levels of small fronts
//
// A level whose fronts have at most kSmallK pivots and whose front vectors fit in
// kSmallLds runs its forward step in one launch, one workgroup per (front, group
// of 16 right-hand sides): the front vector w (m x 16) is gathered in LDS (own
// right-hand side rows, then the children's update vectors in order), y = X w[0:k]
// (X the inverse of the unit-lower diagonal block) goes to Y and to LDS, and the
// update vector w[k:m] - L21 y goes to W. Each output tile is one wave's 16 x 16
// complex MFMA tile; the inner dimension runs in chunks of 32, every operand of a
// chunk loaded before its MFMAs. (The backward step of such levels stays with the
// general kernels, whose row-chunked partial sums measured faster.)

// acc += A (16 x K) B (K x 16), K = 4 steps; a_at(p, j) must be finite and b_at(j, q)
// zero for j >= K (both are called for j up to K rounded up to 32)
template <typename T, typename FA, typename FB>
__device__ __forceinline__ void cmma_k(V4<T>& accr, V4<T>& acci, int steps, FA a_at, FB b_at) {
    const int lane = static_cast<int>(threadIdx.x % 64), p = lane % 16, kq = lane / 16;
    for (int s0 = 0; s0 < steps; s0 += 8) {
        C<T> a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = a_at(p, 4 * (s0 + u) + kq);
#pragma unroll
        for (int u = 0; u < 8; ++u) b[u] = b_at(4 * (s0 + u) + kq, p);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            accr = Mfma<T>::mac(a[u].re, b[u].re, accr);
            accr = Mfma<T>::mac(-a[u].im, b[u].im, accr);
            acci = Mfma<T>::mac(a[u].re, b[u].im, acci);
            acci = Mfma<T>::mac(a[u].im, b[u].re, acci);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fwd_small(const int* __restrict__ nodes, FrontPlan fp,
                                                   const C<T>* __restrict__ P, const C<T>* __restrict__ Xsb,
                                                   const C<T>* __restrict__ bt, int n, int64_t R, C<T>* W,
                                                   C<T>* __restrict__ Y) {
    extern __shared__ unsigned char smem_raw[];
    C<T>* w = reinterpret_cast<C<T>*>(smem_raw);  // m x kQ
    const int s = nodes[blockIdx.x];
    const int g = static_cast<int>(blockIdx.y);
    const Sn x = sn(fp, s);
    C<T>* yl = w + static_cast<int64_t>(x.m) * kQ;  // k x kQ
    const int tid = static_cast<int>(threadIdx.x), wave = tid / 64, lane = tid % 64;
    // w = [b rows of s ; 0] + the children's update vectors, children in order
    for (int idx = tid; idx < x.m * kQ; idx += 256) {
        const int i = idx / kQ;
        w[idx] = i < x.k ? bt[(static_cast<int64_t>(g) * n + x.first + i) * kQ + idx % kQ] : C<T>{0, 0};
    }
    for (int c = fp.child_ptr[s]; c < fp.child_ptr[s + 1]; ++c) {
        const int ch = fp.child[c];
        const Sn y = sn(fp, ch);
        const int* rel = fp.rel + y.rp + y.k;
        const C<T>* Wc = W + (static_cast<int64_t>(g) * R + y.rp + y.k) * kQ;
        __syncthreads();
        for (int idx = tid; idx < (y.m - y.k) * kQ; idx += 256) {
            const int i = idx / kQ, q = idx % kQ;
            C<T>& a = w[rel[i] * kQ + q];
            a = gadd(a, Wc[static_cast<int64_t>(i) * kQ + q]);
        }
    }
    __syncthreads();
    // y = X w[0:k] (X[r][j] is stored as zero for j > r)
    const int nb = x.k;
    const C<T>* Xs = Xsb + fp.xsb_off[s];
    C<T>* Yg = Y + (static_cast<int64_t>(g) * n + x.first) * kQ;
    for (int t = wave; 16 * t < nb; t += 4) {
        const int r0 = 16 * t;
        V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
        cmma_k<T>(accr, acci, (min(nb, r0 + 16) + 3) / 4,
                  [&](int pp, int j) { return Xs[static_cast<int64_t>(min(j, nb - 1)) * nb + min(r0 + pp, nb - 1)]; },
                  [&](int j, int q) { return j < nb ? w[j * kQ + q] : C<T>{0, 0}; });
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = r0 + Mfma<T>::row(lane, v);
            if (r < nb) {
                const C<T> val = C<T>{accr[v], acci[v]};
                yl[r * kQ + lane % 16] = val;
                Yg[r * kQ + lane % 16] = val;
            }
        }
    }
    __syncthreads();
    // the update vector: W[k:m] = w[k:m] - L21 y
    const C<T>* Ps = P + fp.panel_off[s];
    C<T>* Wg = W + (static_cast<int64_t>(g) * R + x.rp) * kQ;
    for (int t = wave; 16 * t < x.m - x.k; t += 4) {
        const int i0 = x.k + 16 * t;
        V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
        cmma_k<T>(accr, acci, (nb + 3) / 4,
                  [&](int pp, int j) { return Ps[static_cast<int64_t>(min(j, nb - 1)) * x.m + min(i0 + pp, x.m - 1)]; },
                  [&](int j, int q) { return j < nb ? yl[j * kQ + q] : C<T>{0, 0}; });
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = i0 + Mfma<T>::row(lane, v);
            if (r < x.m) Wg[static_cast<int64_t>(r) * kQ + lane % 16] = gsub(w[r * kQ + lane % 16], C<T>{accr[v], acci[v]});
        }
    }
}

// colscale[r] = 1 (mode 0) or 1/||b_r|| (mode 1, 1 for a zero column)
__global__ void k_colscale_final(int nrhs, int nparts, const double* __restrict__ partials, double* __restrict__ colscale,
                                 int mode) {
    const int r = threadIdx.x;
    if (r >= nrhs) return;
    if (mode == 0) {
        colscale[r] = 1.0;
        return;
    }
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += partials[r * nparts + p];
    colscale[r] = s > 0.0 ? 1.0 / sqrt(s) : 1.0;
}

// ------------------------------------------------------------------ launchers

int rhs_groups(int nrhs) { return (nrhs + kQ - 1) / kQ; }

template <typename T>
void launch_perm_in(hipStream_t st, int n, int nrhs, const int* perm, const double* d, const cd* b, int64_t ldb,
                    const double* colscale, void* out) {
    hipLaunchKernelGGL(k_perm_in<T>, dim3((n + 15) / 16, rhs_groups(nrhs)), dim3(256), 0, st, n, nrhs, perm, d, b, ldb,
                       colscale, static_cast<C<T>*>(out));
}

template <typename T>
void launch_perm_out(hipStream_t st, int n, int nrhs, const int* perm, const double* d, const void* xt,
                     const double* colscale, cd* x, int64_t ldx, int accumulate) {
    hipLaunchKernelGGL(k_perm_out<T>, dim3((n + 15) / 16, rhs_groups(nrhs)), dim3(256), 0, st, n, nrhs, perm, d,
                       static_cast<const C<T>*>(xt), colscale, x, ldx, accumulate);
}

template <typename T>
void launch_fwd_gather(hipStream_t st, const int* nodes, int count, int max_m, int nrhs, const FrontPlan& fp,
                       const void* b, int n, int64_t R, void* W) {
    if (count <= 0 || max_m <= 0) return;
    // fronts of at most kGatherRows rows use only their own rows of the chunk
    const size_t lds = static_cast<size_t>(std::min(kGatherRows, max_m)) * kQ * sizeof(C<T>);
    allow_large_lds(k_fwd_gather<T>, static_cast<size_t>(kGatherRows) * kQ * sizeof(C<T>));
    hipLaunchKernelGGL(k_fwd_gather<T>, dim3((max_m + kGatherRows - 1) / kGatherRows, count, rhs_groups(nrhs)),
                       dim3(256), lds, st, nodes, fp, static_cast<const C<T>*>(b), n, R, static_cast<C<T>*>(W));
}

template <typename T>
void launch_inv_blocks(hipStream_t st, const int2* pairs, int npairs, const FrontPlan& fp, const void* P,
                       const void* xinv, void* X, void* XT) {
    if (npairs <= 0) return;
    const size_t lds = static_cast<size_t>(2 * kNB * (kNB + 1)) * sizeof(C<T>);
    allow_large_lds(k_inv_blocks<T>, lds);
    hipLaunchKernelGGL(k_inv_blocks<T>, dim3(npairs, kSB / kNB), dim3(256), lds, st, pairs, fp,
                       static_cast<const C<T>*>(P), static_cast<const C<T>*>(xinv), static_cast<C<T>*>(X),
                       static_cast<C<T>*>(XT));
}

template <typename T>
void launch_fwd_step(hipStream_t st, const int* nodes, int count, int B, int max_nb, int max_tri_rows, int nrhs,
                     const FrontPlan& fp, const void* P, const void* Xsb, void* W, void* Y, int n, int64_t R) {
    if (count <= 0 || max_nb <= 0) return;
    const int G = rhs_groups(nrhs);
    // blocks of at most 64 columns (the levels of many small fronts): one-wave workgroups
    const bool one = max_nb <= 64;
    if (one)
        hipLaunchKernelGGL((k_fwd_diag<T, 1>), dim3((max_nb + 15) / 16, count, G), dim3(64), 0, st, nodes, B, fp,
                           static_cast<const C<T>*>(Xsb), static_cast<const C<T>*>(W), static_cast<C<T>*>(Y), n, R);
    else
        hipLaunchKernelGGL((k_fwd_diag<T, 4>), dim3((max_nb + 15) / 16, count, G), dim3(256), 0, st, nodes, B, fp,
                           static_cast<const C<T>*>(Xsb), static_cast<const C<T>*>(W), static_cast<C<T>*>(Y), n, R);
    if (max_tri_rows > 0) {
        if constexpr (std::is_same<T, float>::value) {
            if (one)
                hipLaunchKernelGGL((k_fwd_update2<1, false>), dim3((max_tri_rows + 31) / 32, count, G), dim3(64), 0, st,
                                   nodes, B, fp, static_cast<const cf*>(P), static_cast<cf*>(W), static_cast<const cf*>(Y),
                                   n, R, FwdOut<float>{nullptr, 1, 1, 0});
            else
                hipLaunchKernelGGL((k_fwd_update2<4, false>), dim3((max_tri_rows + 31) / 32, count, G), dim3(256), 0, st,
                                   nodes, B, fp, static_cast<const cf*>(P), static_cast<cf*>(W), static_cast<const cf*>(Y),
                                   n, R, FwdOut<float>{nullptr, 1, 1, 0});
        } else
            hipLaunchKernelGGL((k_fwd_update<T, false>), dim3((max_tri_rows + 15) / 16, count, G), dim3(256), 0, st, nodes,
                               B, fp, static_cast<const C<T>*>(P), static_cast<C<T>*>(W), static_cast<const C<T>*>(Y), n,
                               R, FwdOut<T>{nullptr, 1, 1, 0});
    }
}

// Splits of a level's forward rectangle: enough workgroups to fill the device
// (about kFwdRectWgs) from row tiles x column splits, each split whole blocks
int fwd_rect_split(int count, int max_k, int max_rect_rows, bool fp32, int* bps) {
    const int nbk = (std::max(max_k, 1) + kSB - 1) / kSB;
    const int rows_per_wg = fp32 ? 32 : 16;
    const int64_t wgs = static_cast<int64_t>(count) * ((std::max(max_rect_rows, 1) + rows_per_wg - 1) / rows_per_wg);
    int ks = static_cast<int>(std::min<int64_t>(nbk, std::max<int64_t>(1, (kFwdRectWgs + wgs - 1) / wgs)));
    const int b = (nbk + ks - 1) / ks;
    if (bps) *bps = b;
    return (nbk + b - 1) / b;
}

template <typename T>
void launch_fwd_rect(hipStream_t st, const int* nodes, int count, int max_k, int max_rect_rows, int nrhs,
                     const FrontPlan& fp, const void* P, void* W, const void* Y, int n, int64_t R, void* part) {
    if (count <= 0 || max_k <= 0 || max_rect_rows <= 0) return;
    const int G = rhs_groups(nrhs);
    constexpr bool f32 = std::is_same<T, float>::value;
    int bps = 1;
    const int ks = fwd_rect_split(count, max_k, max_rect_rows, f32, &bps);
    const FwdOut<T> fo{static_cast<C<T>*>(part), ks, bps, max_rect_rows};
    const int tiles = (max_rect_rows + (f32 ? 31 : 15)) / (f32 ? 32 : 16);
    const dim3 grid(tiles * ks, count, G);
    if constexpr (f32) {
        if (max_k <= 64)
            hipLaunchKernelGGL((k_fwd_update2<1, true>), grid, dim3(64), 0, st, nodes, 0, fp, static_cast<const cf*>(P),
                               static_cast<cf*>(W), static_cast<const cf*>(Y), n, R, fo);
        else
            hipLaunchKernelGGL((k_fwd_update2<4, true>), grid, dim3(256), 0, st, nodes, 0, fp, static_cast<const cf*>(P),
                               static_cast<cf*>(W), static_cast<const cf*>(Y), n, R, fo);
    } else {
        hipLaunchKernelGGL((k_fwd_update<T, true>), grid, dim3(256), 0, st, nodes, 0, fp, static_cast<const C<T>*>(P),
                           static_cast<C<T>*>(W), static_cast<const C<T>*>(Y), n, R, fo);
    }
    if (ks > 1)
        hipLaunchKernelGGL(k_fwd_rect_reduce<T>, dim3((max_rect_rows * kQ + 255) / 256, count, G), dim3(256), 0, st,
                           nodes, fp, static_cast<C<T>*>(W), R, fo);
}

int rect_chunks(int rect_rows) { return std::max(1, (rect_rows + kRectRows - 1) / kRectRows); }
int rect_cols(int max_k) { return ((std::max(max_k, 1) + 15) / 16) * 16; }

template <typename T>
const void* launch_bwd_rect(hipStream_t st, const int* nodes, int count, int max_k, int max_rect_rows, int nrhs,
                            bool all, const FrontPlan& fp, const void* P, const void* X, int n, void* rpart,
                            void* t) {
    if (count <= 0 || max_k <= 0) return nullptr;
    if (max_rect_rows <= 0 && !all) return nullptr;
    const int G = rhs_groups(nrhs);
    const int ctiles = (max_k + 15) / 16, rk = rect_cols(max_k);
    const int stride = rect_chunks(max_rect_rows);
    if (max_rect_rows > 0) {
        const dim3 grid(stride * ctiles, count, G);
        if constexpr (std::is_same<T, float>::value) {
            if (max_rect_rows <= 64)
                hipLaunchKernelGGL((k_bwd_partial2<1, true>), grid, dim3(64), 0, st, nodes, 0, stride, ctiles, rk, fp,
                                   static_cast<const cf*>(P), static_cast<const cf*>(X), n, static_cast<cf*>(rpart));
            else
                hipLaunchKernelGGL((k_bwd_partial2<4, true>), grid, dim3(256), 0, st, nodes, 0, stride, ctiles, rk, fp,
                                   static_cast<const cf*>(P), static_cast<const cf*>(X), n, static_cast<cf*>(rpart));
        } else {
            hipLaunchKernelGGL((k_bwd_partial<T, true>), grid, dim3(256), 0, st, nodes, 0, stride, ctiles, rk, fp,
                               static_cast<const C<T>*>(P), static_cast<const C<T>*>(X), n, static_cast<C<T>*>(rpart));
        }
        // with one chunk of rows the partial sums are already t (same layout) for
        // the fronts that have such rows, which is all a diagonal step reads when
        // the chain has no updates
        if (stride == 1 && !all) return rpart;
    }
    hipLaunchKernelGGL(k_rect_reduce<T>, dim3(max_k, count, G), dim3(256), 0, st, nodes, stride, rk, fp,
                       static_cast<const C<T>*>(rpart), static_cast<C<T>*>(t));
    return t;
}

template <typename T>
void launch_bwd_step(hipStream_t st, const int* nodes, int count, int B, int max_nb, int nrhs, const FrontPlan& fp,
                     const void* P, const void* dinv, const void* XTsb, const void* Y, void* X, int n, void* t, int rk,
                     bool all) {
    if (count <= 0 || max_nb <= 0) return;
    const int G = rhs_groups(nrhs);
    const int ctiles = (max_nb + 15) / 16;
    if (max_nb <= 64)
        hipLaunchKernelGGL((k_bwd_diag<T, 1>), dim3(ctiles, count, G), dim3(64), 0, st, nodes, B, fp,
                           static_cast<const C<T>*>(dinv), static_cast<const C<T>*>(XTsb), static_cast<const C<T>*>(Y),
                           static_cast<const C<T>*>(t), rk, all ? 1 : 0, static_cast<C<T>*>(X), n);
    else
        hipLaunchKernelGGL((k_bwd_diag<T, 4>), dim3(ctiles, count, G), dim3(256), 0, st, nodes, B, fp,
                           static_cast<const C<T>*>(dinv), static_cast<const C<T>*>(XTsb), static_cast<const C<T>*>(Y),
                           static_cast<const C<T>*>(t), rk, all ? 1 : 0, static_cast<C<T>*>(X), n);
    if (B == 0) return;
    // the columns before the block take its rows: t[0:B] += L[B:B+nb, 0:B]^T x_B
    const dim3 grid((B + 15) / 16, count, G);
    if constexpr (std::is_same<T, float>::value) {
        if (max_nb <= 64)
            hipLaunchKernelGGL((k_bwd_partial2<1, false>), grid, dim3(64), 0, st, nodes, B, 1, (B + 15) / 16, rk, fp,
                               static_cast<const cf*>(P), static_cast<const cf*>(X), n, static_cast<cf*>(t));
        else
            hipLaunchKernelGGL((k_bwd_partial2<4, false>), grid, dim3(256), 0, st, nodes, B, 1, (B + 15) / 16, rk, fp,
                               static_cast<const cf*>(P), static_cast<const cf*>(X), n, static_cast<cf*>(t));
    } else {
        hipLaunchKernelGGL((k_bwd_partial<T, false>), grid, dim3(256), 0, st, nodes, B, 1, (B + 15) / 16, rk, fp,
                           static_cast<const C<T>*>(P), static_cast<const C<T>*>(X), n, static_cast<C<T>*>(t));
    }
}

template <typename T>
bool small_level(int max_k, int max_m) {
    return max_k <= kSmallK &&
           static_cast<size_t>(max_m + kSmallK) * kQ * sizeof(C<T>) <= static_cast<size_t>(kSmallLds);
}

template <typename T>
void launch_fwd_small(hipStream_t st, const int* nodes, int count, int max_m, int nrhs, const FrontPlan& fp,
                      const void* P, const void* Xsb, const void* b, int n, int64_t R, void* W, void* Y) {
    if (count <= 0) return;
    const size_t lds = static_cast<size_t>(max_m + kSmallK) * kQ * sizeof(C<T>);
    allow_large_lds(k_fwd_small<T>, lds);
    hipLaunchKernelGGL(k_fwd_small<T>, dim3(count, rhs_groups(nrhs)), dim3(256), lds, st, nodes, fp,
                       static_cast<const C<T>*>(P), static_cast<const C<T>*>(Xsb), static_cast<const C<T>*>(b), n, R,
                       static_cast<C<T>*>(W), static_cast<C<T>*>(Y));
}

template bool small_level<double>(int, int);
template bool small_level<float>(int, int);
template void launch_fwd_small<double>(hipStream_t, const int*, int, int, int, const FrontPlan&, const void*, const void*, const void*, int, int64_t, void*, void*);
template void launch_fwd_small<float>(hipStream_t, const int*, int, int, int, const FrontPlan&, const void*, const void*, const void*, int, int64_t, void*, void*);

void launch_column_scale(hipStream_t st, int n, int nrhs, const cd* b, int64_t ldb, double* partials, double* colscale,
                         int mode) {
    const int nparts = 64;
    if (mode != 0)
        hipLaunchKernelGGL(k_sumsq_partial, dim3(nparts, nrhs), dim3(256), 0, st, n, b, ldb, partials);
    hipLaunchKernelGGL(k_colscale_final, dim3(1), dim3(64), 0, st, nrhs, nparts, partials, colscale, mode);
}

template void launch_perm_in<double>(hipStream_t, int, int, const int*, const double*, const cd*, int64_t, const double*, void*);
template void launch_perm_in<float>(hipStream_t, int, int, const int*, const double*, const cd*, int64_t, const double*, void*);
template void launch_perm_out<double>(hipStream_t, int, int, const int*, const double*, const void*, const double*, cd*, int64_t, int);
template void launch_perm_out<float>(hipStream_t, int, int, const int*, const double*, const void*, const double*, cd*, int64_t, int);
template void launch_fwd_gather<double>(hipStream_t, const int*, int, int, int, const FrontPlan&, const void*, int, int64_t, void*);
template void launch_fwd_gather<float>(hipStream_t, const int*, int, int, int, const FrontPlan&, const void*, int, int64_t, void*);
template void launch_fwd_step<double>(hipStream_t, const int*, int, int, int, int, int, const FrontPlan&, const void*, const void*, void*, void*, int, int64_t);
template void launch_fwd_step<float>(hipStream_t, const int*, int, int, int, int, int, const FrontPlan&, const void*, const void*, void*, void*, int, int64_t);
template void launch_fwd_rect<double>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, const void*, void*, const void*, int, int64_t, void*);
template void launch_fwd_rect<float>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, const void*, void*, const void*, int, int64_t, void*);
template const void* launch_bwd_rect<double>(hipStream_t, const int*, int, int, int, int, bool, const FrontPlan&, const void*, const void*, int, void*, void*);
template const void* launch_bwd_rect<float>(hipStream_t, const int*, int, int, int, int, bool, const FrontPlan&, const void*, const void*, int, void*, void*);
template void launch_bwd_step<double>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, const void*, const void*, const void*, const void*, void*, int, void*, int, bool);
template void launch_bwd_step<float>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, const void*, const void*, const void*, const void*, void*, int, void*, int, bool);
template <typename T>
void launch_dinv(hipStream_t st, int ns, const FrontPlan& fp, const void* P, void* dinv) {
    if (ns <= 0) return;
    hipLaunchKernelGGL(k_dinv<T>, dim3(ns), dim3(64), 0, st, fp, static_cast<const C<T>*>(P), static_cast<C<T>*>(dinv));
}
template void launch_dinv<double>(hipStream_t, int, const FrontPlan&, const void*, void*);
template void launch_dinv<float>(hipStream_t, int, const FrontPlan&, const void*, void*);
template void launch_inv_blocks<double>(hipStream_t, const int2*, int, const FrontPlan&, const void*, const void*, void*, void*);
template void launch_inv_blocks<float>(hipStream_t, const int2*, int, const FrontPlan&, const void*, const void*, void*, void*);

}  // namespace aphi_dev

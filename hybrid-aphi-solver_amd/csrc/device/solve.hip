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

}  // namespace

template <typename T>
__global__ void k_perm_in(int n, int nrhs, const int* __restrict__ perm, const double* __restrict__ d,
                          const cd* __restrict__ b, int64_t ldb, const double* __restrict__ colscale, C<T>* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n || r >= nrhs) return;
    const int p = perm[i];
    cd v = cmul(b[r * ldb + p], cd{d[p], 0.0});
    const double s = colscale[r];
    if (s != 1.0) v = cd{v.re * s, v.im * s};
    out[static_cast<int64_t>(r) * n + i] = C<T>{static_cast<T>(v.re), static_cast<T>(v.im)};
}

template <typename T>
__global__ void k_perm_out(int n, int nrhs, const int* __restrict__ perm, const double* __restrict__ d,
                           const C<T>* __restrict__ xt, const double* __restrict__ colscale, cd* __restrict__ x,
                           int64_t ldx, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n || r >= nrhs) return;
    const int p = perm[i];
    const C<T> t = xt[static_cast<int64_t>(r) * n + i];
    cd v = cd{static_cast<double>(t.re), static_cast<double>(t.im)};
    const double s = colscale[r];
    if (s != 1.0) v = cd{v.re / s, v.im / s};
    v = cmul(v, cd{d[p], 0.0});
    x[r * ldx + p] = accumulate ? cadd(x[r * ldx + p], v) : v;
}

// W_s := [b rows of s ; 0] + extend-add of the children's update vectors.
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_gather(const int* __restrict__ nodes, int nrhs, FrontPlan fp,
                                                    const C<T>* __restrict__ b, int n, C<T>* W) {
    const int s = nodes[blockIdx.x];
    const int r = blockIdx.y;
    const Sn x = sn(fp, s);
    C<T>* Ws = W + x.rp * nrhs + static_cast<int64_t>(r) * x.m;
    for (int i = threadIdx.x; i < x.m; i += blockDim.x)
        Ws[i] = i < x.k ? b[static_cast<int64_t>(r) * n + x.first + i] : C<T>{0, 0};
    for (int c = fp.child_ptr[s]; c < fp.child_ptr[s + 1]; ++c) {
        __syncthreads();
        const int ch = fp.child[c];
        const Sn y = sn(fp, ch);
        const int cu = y.m - y.k;
        const C<T>* Wc = W + y.rp * nrhs + static_cast<int64_t>(r) * y.m + y.k;
        const int* rel = fp.rel + y.rp + y.k;
        for (int i = threadIdx.x; i < cu; i += blockDim.x) Ws[rel[i]] = gadd(Ws[rel[i]], Wc[i]);
    }
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
    C<T>* Xs = X + fp.x256_off[s] + static_cast<int64_t>(B / kSB) * kSB * kSB;
    C<T>* XTs = XT + fp.x256_off[s] + static_cast<int64_t>(B / kSB) * kSB * kSB;
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

// ------------------------------------------------------------------ forward steps
//
// Step B of every supernode with k > B: y_B = X_B w_B (X = L_BB^-1), then the
// rows below the block take W[i] -= L[i, B:B+nb] y_B. Right-hand sides go in
// groups of 16.

// The step kernels are small complex GEMMs: each wave owns 16x16 output tiles
// (16 rows of the block or front x 16 right-hand sides) accumulated on MFMA,
// with the right-hand-side operand staged in LDS as [row][rhs] (zero-padded).
template <typename T>
using V4 = typename Mfma<T>::v4;

// y rows [64 bx, 64 bx + 64) of the block: y = X w; wave = 16 rows
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_diag(const int* __restrict__ nodes, int B, int nrhs, FrontPlan fp,
                                                  const C<T>* __restrict__ X256, const C<T>* __restrict__ W,
                                                  C<T>* __restrict__ Y, int n) {
    __shared__ C<T> wl[kSB * 16];  // wl[q*16 + r]
    const int s = nodes[blockIdx.y];
    const Sn x = sn(fp, s);
    if (x.k <= B) return;
    const int nb = min(kSB, x.k - B);
    const int c0 = static_cast<int>(blockIdx.x) * 64;
    if (c0 >= nb) return;
    const int qmax = min(nb, c0 + 64);  // X[c][q] = 0 for q > c
    const int K4 = (qmax + 3) & ~3;
    const C<T>* Xs = X256 + fp.x256_off[s] + static_cast<int64_t>(B / kSB) * kSB * kSB;
    const C<T>* Ws = W + x.rp * nrhs;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int cw = c0 + 16 * wave;
    for (int r0 = 0; r0 < nrhs; r0 += 16) {
        const int nr = min(16, nrhs - r0);
        __syncthreads();
        {
            C<T> t[kSB * 16 / 256];  // all loads in flight before the LDS stores
#pragma unroll
            for (int u = 0; u < kSB * 16 / 256; ++u) {
                const int idx = tid + 256 * u, q = idx / 16, r = idx % 16;
                t[u] = (q < qmax && r < nr) ? Ws[static_cast<int64_t>(r0 + r) * x.m + B + q] : C<T>{0, 0};
            }
#pragma unroll
            for (int u = 0; u < kSB * 16 / 256; ++u) wl[tid + 256 * u] = t[u];
        }
        __syncthreads();
        if (cw < nb) {
            V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
            cmma16<T>(accr, acci, K4,
                      [&](int pp, int k) { return Xs[static_cast<int64_t>(min(k, nb - 1)) * nb + min(cw + pp, nb - 1)]; },
                      [&](int k, int q) { return wl[k * 16 + q]; });
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = cw + Mfma<T>::row(lane, v), r = lane % 16;
                if (c < nb && r < nr) Y[static_cast<int64_t>(r0 + r) * n + x.first + B + c] = C<T>{accr[v], acci[v]};
            }
        }
    }
}

// W[i] -= L[i, B:B+nb] y_B for 64 rows per workgroup; wave = 16 rows
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_update(const int* __restrict__ nodes, int B, int nrhs, FrontPlan fp,
                                                    const C<T>* __restrict__ P, C<T>* W, const C<T>* __restrict__ Y,
                                                    int n) {
    __shared__ C<T> yl[kSB * 16];  // yl[c*16 + r]
    const int s = nodes[blockIdx.y];
    const Sn x = sn(fp, s);
    if (x.k <= B) return;
    const int nb = min(kSB, x.k - B);
    const int i0 = B + nb + static_cast<int>(blockIdx.x) * 64;
    if (i0 >= x.m) return;
    const int K4 = (nb + 3) & ~3;
    const C<T>* Ps = P + fp.panel_off[s];
    C<T>* Ws = W + x.rp * nrhs;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int iw = i0 + 16 * wave;
    for (int r0 = 0; r0 < nrhs; r0 += 16) {
        const int nr = min(16, nrhs - r0);
        __syncthreads();
        {
            C<T> t[kSB * 16 / 256];
#pragma unroll
            for (int u = 0; u < kSB * 16 / 256; ++u) {
                const int idx = tid + 256 * u, c = idx / 16, r = idx % 16;
                t[u] = (c < nb && r < nr) ? Y[static_cast<int64_t>(r0 + r) * n + x.first + B + c] : C<T>{0, 0};
            }
#pragma unroll
            for (int u = 0; u < kSB * 16 / 256; ++u) yl[tid + 256 * u] = t[u];
        }
        __syncthreads();
        if (iw < x.m) {
            V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
            cmma16<T>(accr, acci, K4,
                      [&](int pp, int k) {
                          return Ps[static_cast<int64_t>(B + min(k, nb - 1)) * x.m + min(iw + pp, x.m - 1)];
                      },
                      [&](int k, int q) { return yl[k * 16 + q]; });
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = iw + Mfma<T>::row(lane, v), r = lane % 16;
                if (i < x.m && r < nr) {
                    C<T>& dst = Ws[static_cast<int64_t>(r0 + r) * x.m + i];
                    dst = C<T>{dst.re - accr[v], dst.im - acci[v]};
                }
            }
        }
    }
}

// ------------------------------------------------------------------ backward steps
//
// Step B (blocks in decreasing order): t_B = L[rows below, B:B+nb]^T x[rows below]
// in 64-row partial sums, reduced in chunk order, then x_B = X_B^T (D_B^-1 y_B - t_B).

// partial[(slot*stride + bx)][r][c] over 64 rows; wave = 16-column tiles w, w+4, ...
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_partial(const int* __restrict__ nodes, int B, int nrhs, int stride,
                                                     FrontPlan fp, const C<T>* __restrict__ P,
                                                     const C<T>* __restrict__ X, int n, C<T>* __restrict__ partial) {
    __shared__ C<T> xt[64 * 16];  // xt[row*16 + r]
    const int s = nodes[blockIdx.y];
    const Sn x = sn(fp, s);
    if (x.k <= B) return;
    const int nb = min(kSB, x.k - B);
    const int i0 = B + nb + static_cast<int>(blockIdx.x) * 64;
    if (i0 >= x.m) return;
    const int nrows = min(64, x.m - i0);
    const C<T>* Ps = P + fp.panel_off[s];
    const int* rows = fp.rows + x.rp;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    C<T>* out = partial + (static_cast<int64_t>(blockIdx.y) * stride + blockIdx.x) * kSB * nrhs;
    for (int r0 = 0; r0 < nrhs; r0 += 16) {
        const int nr = min(16, nrhs - r0);
        __syncthreads();
        {
            int g[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int row = (tid + 256 * u) / 16;
                g[u] = row < nrows ? rows[i0 + row] : 0;
            }
            C<T> t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int idx = tid + 256 * u, row = idx / 16, r = idx % 16;
                t[u] = (row < nrows && r < nr) ? X[static_cast<int64_t>(r0 + r) * n + g[u]] : C<T>{0, 0};
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) xt[tid + 256 * u] = t[u];
        }
        __syncthreads();
        for (int ct = 16 * wave; ct < nb; ct += 64) {
            V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
            cmma16<T>(accr, acci, 64,
                      [&](int pp, int k) {
                          return Ps[static_cast<int64_t>(B + min(ct + pp, nb - 1)) * x.m + min(i0 + k, x.m - 1)];
                      },
                      [&](int k, int q) { return xt[k * 16 + q]; });
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = ct + Mfma<T>::row(lane, v), r = lane % 16;
                if (c < nb && r < nr) out[static_cast<int64_t>(r0 + r) * kSB + c] = C<T>{accr[v], acci[v]};
            }
        }
    }
}

// tsum[slot][r][c] = sum over chunks of partial, for 16 columns and one RHS per
// workgroup: 16 summers per output, combined in a fixed order
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_reduce(const int* __restrict__ nodes, int B, int nrhs, int stride,
                                                    int cblocks, FrontPlan fp, const C<T>* __restrict__ partial,
                                                    C<T>* __restrict__ tsum) {
    __shared__ C<T> red[16][17];
    const int s = nodes[blockIdx.y];
    const Sn x = sn(fp, s);
    if (x.k <= B) return;
    const int nb = min(kSB, x.k - B);
    const int nch = x.m > B + nb ? (x.m - B - nb + 63) / 64 : 0;
    const int cblk = static_cast<int>(blockIdx.x) % cblocks, r = static_cast<int>(blockIdx.x) / cblocks;
    if (cblk * 16 >= nb || r >= nrhs) return;
    const int tid = threadIdx.x, o = tid % 16, p = tid / 16;
    const int c = cblk * 16 + o;
    const C<T>* pp = partial + static_cast<int64_t>(blockIdx.y) * stride * kSB * nrhs + static_cast<int64_t>(r) * kSB + c;
    C<T> t = C<T>{0, 0};
#pragma unroll 8
    for (int ch = p; ch < nch; ch += 16) t = gadd(t, pp[static_cast<int64_t>(ch) * kSB * nrhs]);
    red[p][o] = t;
    __syncthreads();
    if (p == 0 && c < nb) {
        C<T> v = red[0][o];
        for (int q = 1; q < 16; ++q) v = gadd(v, red[q][o]);
        tsum[(static_cast<int64_t>(blockIdx.y) * nrhs + r) * kSB + c] = v;
    }
}

// x rows [64 bx, 64 bx + 64) of the block: x[c] = sum_{q >= c} X[q][c] z[q],
// z = y D^-1 - t; wave = 16 rows
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_diag(const int* __restrict__ nodes, int B, int nrhs, FrontPlan fp,
                                                  const C<T>* __restrict__ P, const C<T>* __restrict__ XT256,
                                                  const C<T>* __restrict__ Y, const C<T>* __restrict__ tsum,
                                                  C<T>* __restrict__ X, int n) {
    __shared__ C<T> zl[kSB * 16];  // zl[q*16 + r], q counted from c0
    const int s = nodes[blockIdx.y];
    const Sn x = sn(fp, s);
    if (x.k <= B) return;
    const int nb = min(kSB, x.k - B);
    const int c0 = static_cast<int>(blockIdx.x) * 64;
    if (c0 >= nb) return;
    const int K4 = (nb - c0 + 3) & ~3;
    const bool below = x.m > B + nb;
    const C<T>* Ps = P + fp.panel_off[s];
    const C<T>* XTs = XT256 + fp.x256_off[s] + static_cast<int64_t>(B / kSB) * kSB * kSB;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int cw = c0 + 16 * wave;
    for (int r0 = 0; r0 < nrhs; r0 += 16) {
        const int nr = min(16, nrhs - r0);
        __syncthreads();
        {
            C<T> yv[kSB * 16 / 256], dv[kSB * 16 / 256], tv[kSB * 16 / 256];
#pragma unroll
            for (int u = 0; u < kSB * 16 / 256; ++u) {
                const int idx = tid + 256 * u, q = c0 + idx / 16, r = idx % 16;
                const bool ok = q < nb && r < nr;
                yv[u] = ok ? Y[static_cast<int64_t>(r0 + r) * n + x.first + B + q] : C<T>{0, 0};
                dv[u] = ok ? Ps[static_cast<int64_t>(B + q) * x.m + B + q] : C<T>{1, 0};
                tv[u] = (ok && below) ? tsum[(static_cast<int64_t>(blockIdx.y) * nrhs + r0 + r) * kSB + q] : C<T>{0, 0};
            }
#pragma unroll
            for (int u = 0; u < kSB * 16 / 256; ++u) zl[tid + 256 * u] = gsub(gmul(yv[u], grecip(dv[u])), tv[u]);
        }
        __syncthreads();
        if (cw < nb) {
            V4<T> accr = V4<T>{0, 0, 0, 0}, acci = V4<T>{0, 0, 0, 0};
            cmma16<T>(accr, acci, K4,
                      [&](int pp, int k) {
                          return XTs[static_cast<int64_t>(min(c0 + k, nb - 1)) * nb + min(cw + pp, nb - 1)];
                      },
                      [&](int k, int q) { return zl[k * 16 + q]; });
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = cw + Mfma<T>::row(lane, v), r = lane % 16;
                if (c < nb && r < nr) X[static_cast<int64_t>(r0 + r) * n + x.first + B + c] = C<T>{accr[v], acci[v]};
            }
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

template <typename T>
void launch_perm_in(hipStream_t st, int n, int nrhs, const int* perm, const double* d, const cd* b, int64_t ldb,
                    const double* colscale, void* out) {
    hipLaunchKernelGGL(k_perm_in<T>, dim3((n + 255) / 256, nrhs), dim3(256), 0, st, n, nrhs, perm, d, b, ldb, colscale,
                       static_cast<C<T>*>(out));
}

template <typename T>
void launch_perm_out(hipStream_t st, int n, int nrhs, const int* perm, const double* d, const void* xt,
                     const double* colscale, cd* x, int64_t ldx, int accumulate) {
    hipLaunchKernelGGL(k_perm_out<T>, dim3((n + 255) / 256, nrhs), dim3(256), 0, st, n, nrhs, perm, d,
                       static_cast<const C<T>*>(xt), colscale, x, ldx, accumulate);
}

template <typename T>
void launch_fwd_gather(hipStream_t st, const int* nodes, int count, int nrhs, const FrontPlan& fp, const void* b, int n,
                       void* W) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fwd_gather<T>, dim3(count, nrhs), dim3(256), 0, st, nodes, nrhs, fp,
                       static_cast<const C<T>*>(b), n, static_cast<C<T>*>(W));
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
void launch_fwd_step(hipStream_t st, const int* nodes, int count, int B, int max_nb, int max_row_blocks, int nrhs,
                     const FrontPlan& fp, const void* P, const void* X256, void* W, void* Y, int n) {
    if (count <= 0 || max_nb <= 0) return;
    hipLaunchKernelGGL(k_fwd_diag<T>, dim3((max_nb + 63) / 64, count), dim3(256), 0, st, nodes, B, nrhs, fp,
                       static_cast<const C<T>*>(X256), static_cast<const C<T>*>(W), static_cast<C<T>*>(Y), n);
    if (max_row_blocks > 0) {
        hipLaunchKernelGGL(k_fwd_update<T>, dim3(max_row_blocks, count), dim3(256), 0, st, nodes, B, nrhs, fp,
                           static_cast<const C<T>*>(P), static_cast<C<T>*>(W), static_cast<const C<T>*>(Y), n);
    }
}

template <typename T>
void launch_bwd_step(hipStream_t st, const int* nodes, int count, int B, int max_nb, int max_row_blocks, int nrhs,
                     const FrontPlan& fp, const void* P, const void* XT256, const void* Y, void* X, int n,
                     void* partial, void* tsum) {
    if (count <= 0 || max_nb <= 0) return;
    const int stride = max_row_blocks > 0 ? max_row_blocks : 1;
    if (max_row_blocks > 0) {
        hipLaunchKernelGGL(k_bwd_partial<T>, dim3(max_row_blocks, count), dim3(256), 0, st, nodes, B, nrhs, stride,
                           fp, static_cast<const C<T>*>(P), static_cast<const C<T>*>(X), n, static_cast<C<T>*>(partial));
        hipLaunchKernelGGL(k_bwd_reduce<T>, dim3(((max_nb + 15) / 16) * nrhs, count), dim3(256), 0, st, nodes, B, nrhs,
                           stride, (max_nb + 15) / 16, fp, static_cast<const C<T>*>(partial), static_cast<C<T>*>(tsum));
    }
    hipLaunchKernelGGL(k_bwd_diag<T>, dim3((max_nb + 63) / 64, count), dim3(256), 0, st, nodes, B, nrhs, fp,
                       static_cast<const C<T>*>(P), static_cast<const C<T>*>(XT256), static_cast<const C<T>*>(Y),
                       static_cast<const C<T>*>(tsum), static_cast<C<T>*>(X), n);
}

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
template void launch_fwd_gather<double>(hipStream_t, const int*, int, int, const FrontPlan&, const void*, int, void*);
template void launch_fwd_gather<float>(hipStream_t, const int*, int, int, const FrontPlan&, const void*, int, void*);
template void launch_fwd_step<double>(hipStream_t, const int*, int, int, int, int, int, const FrontPlan&, const void*, const void*, void*, void*, int);
template void launch_fwd_step<float>(hipStream_t, const int*, int, int, int, int, int, const FrontPlan&, const void*, const void*, void*, void*, int);
template void launch_bwd_step<double>(hipStream_t, const int*, int, int, int, int, int, const FrontPlan&, const void*, const void*, const void*, void*, int, void*, void*);
template void launch_bwd_step<float>(hipStream_t, const int*, int, int, int, int, int, const FrontPlan&, const void*, const void*, const void*, void*, int, void*, void*);
template void launch_inv_blocks<double>(hipStream_t, const int2*, int, const FrontPlan&, const void*, const void*, void*, void*);
template void launch_inv_blocks<float>(hipStream_t, const int2*, int, const FrontPlan&, const void*, const void*, void*, void*);

}  // namespace aphi_dev

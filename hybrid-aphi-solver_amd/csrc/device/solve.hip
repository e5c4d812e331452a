// Multi-RHS supernodal triangular solves (src/factorization.cpp:254-313, organised
// by fronts): forward through L (unit), divide by D, backward through L^T -- the
// transpose, never the conjugate transpose.
//
// Forward, leaves to root: each supernode's front vector W_s (m x nrhs) takes its
// own right-hand side rows plus its children's update vectors (children in order,
// so the sums are deterministic), solves its unit-lower diagonal part and leaves
// W_s[k:m] - L21 y as its own update vector. Backward, root to leaves: x_s =
// L11^-T (D^-1 y_s - L21^T x[rows beyond]), where the rows beyond belong to
// ancestors that are already solved. Blocks of NB columns bound the work of one
// workgroup; row-parallel kernels carry the rest.
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

// Forward step b of every supernode with k > b, one launch: each workgroup forms
// y_b = L_bb^-1 w_b from the stored block inverse (a 64x64 GEMV, recomputed per
// workgroup instead of a serial substitution), workgroup 0 writes y_b to Y, and
// every workgroup updates its 256 rows below the block: W[i] -= L[i, b:b+nb] y_b.
// Rows b..b+nb of W are only read in this launch, so there is no race. Global
// loads are issued in unrolled batches so that each thread has many in flight:
// the step is latency-bound otherwise.
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_block(const int* __restrict__ nodes, int b, int nrhs, FrontPlan fp,
                                                   const C<T>* __restrict__ P, const C<T>* __restrict__ xinv, C<T>* W,
                                                   C<T>* __restrict__ Y, int n) {
    constexpr int CH = sizeof(T) == 4 ? 32 : 16;  // L columns held in registers per batch
    __shared__ C<T> Xl[kNB * kNB];
    __shared__ C<T> Wb[16 * kNB];
    __shared__ C<T> Yb[16 * kNB];
    const int s = nodes[blockIdx.y];
    const Sn x = sn(fp, s);
    if (x.k <= b) return;
    const int nb = min(kNB, x.k - b);
    const int i0 = b + nb + static_cast<int>(blockIdx.x) * 256;
    if (blockIdx.x > 0 && i0 >= x.m) return;
    const C<T>* Ps = P + fp.panel_off[s];
    const C<T>* Xg = xinv + fp.xinv_off[s] + static_cast<int64_t>(b / kNB) * kNB * kNB;
    C<T>* Ws = W + x.rp * nrhs;
    const int tid = threadIdx.x;
    {
        C<T> t[kNB * kNB / 256];
#pragma unroll
        for (int u = 0; u < kNB * kNB / 256; ++u) {
            const int idx = tid + 256 * u;
            t[u] = idx < nb * nb ? Xg[idx] : C<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < kNB * kNB / 256; ++u) Xl[tid + 256 * u] = t[u];  // column q, row c at q*nb + c
    }
    const int i = i0 + tid;
    for (int r0 = 0; r0 < nrhs; r0 += 16) {
        const int nr = min(16, nrhs - r0);
        __syncthreads();
        {
            C<T> t[16 * kNB / 256];
#pragma unroll
            for (int u = 0; u < 16 * kNB / 256; ++u) {
                const int idx = tid + 256 * u, r = idx / kNB, c = idx % kNB;
                t[u] = (r < nr && c < nb) ? Ws[static_cast<int64_t>(r0 + r) * x.m + b + c] : C<T>{0, 0};
            }
#pragma unroll
            for (int u = 0; u < 16 * kNB / 256; ++u) Wb[tid + 256 * u] = t[u];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 16 * kNB / 256; ++u) {
            const int idx = tid + 256 * u, r = idx / kNB, c = idx % kNB;
            C<T> acc = C<T>{0, 0};
            if (r < nr && c < nb) {
                for (int q = 0; q <= c; ++q) acc = gfma(acc, Xl[q * nb + c], Wb[r * kNB + q]);
                if (blockIdx.x == 0) Y[static_cast<int64_t>(r0 + r) * n + x.first + b + c] = acc;
            }
            Yb[idx] = acc;  // zero outside nb x nr
        }
        __syncthreads();
        if (i < x.m) {
            C<T> acc[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = C<T>{0, 0};
            for (int c0 = 0; c0 < nb; c0 += CH) {
                C<T> l[CH];
#pragma unroll
                for (int u = 0; u < CH; ++u)
                    l[u] = c0 + u < nb ? Ps[static_cast<int64_t>(b + c0 + u) * x.m + i] : C<T>{0, 0};
#pragma unroll
                for (int u = 0; u < CH; ++u)
#pragma unroll
                    for (int r = 0; r < 16; ++r) acc[r] = gfma(acc[r], l[u], Yb[r * kNB + c0 + u]);
            }
            C<T> w[16];
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if (r < nr) w[r] = Ws[static_cast<int64_t>(r0 + r) * x.m + i];
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if (r < nr) Ws[static_cast<int64_t>(r0 + r) * x.m + i] = gsub(w[r], acc[r]);
        }
    }
}

// partial[slot][blk] (nb x nrhs, c-major) = sum over 64 rows of L[row, b+c] * x[row]
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_partial(const int* __restrict__ nodes, int b, int nrhs, int stride,
                                                     FrontPlan fp, const C<T>* __restrict__ P,
                                                     const C<T>* __restrict__ X, int n, C<T>* __restrict__ partial) {
    extern __shared__ unsigned char smem_raw[];
    const int s = nodes[blockIdx.y];
    const Sn x = sn(fp, s);
    if (x.k <= b) return;
    const int nb = min(kNB, x.k - b);
    const int i0 = b + nb + static_cast<int>(blockIdx.x) * 64;
    if (i0 >= x.m) return;
    const int nrows = min(64, x.m - i0);
    C<T>* Lt = reinterpret_cast<C<T>*>(smem_raw);  // Lt[c*64 + row]
    C<T>* Xt = Lt + kNB * 64;                      // Xt[row*nrhs + r]
    const C<T>* Ps = P + fp.panel_off[s];
    const int* rows = fp.rows + x.rp;
    const int tid = threadIdx.x;
    {
        C<T> t[kNB * 64 / 256];
#pragma unroll
        for (int u = 0; u < kNB * 64 / 256; ++u) {
            const int idx = tid + 256 * u, c = idx / 64, row = idx % 64;
            t[u] = (c < nb && row < nrows) ? Ps[static_cast<int64_t>(b + c) * x.m + i0 + row] : C<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < kNB * 64 / 256; ++u) Lt[tid + 256 * u] = t[u];
    }
    {
        // rows x nrhs gather of the solved unknowns (nrhs <= 64)
        int g[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = tid + 256 * u, row = idx % 64;
            g[u] = (idx < nrhs * 64 && row < nrows) ? rows[i0 + row] : -1;
        }
        C<T> t[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = tid + 256 * u, r = idx / 64;
            t[u] = g[u] >= 0 ? X[static_cast<int64_t>(r) * n + g[u]] : C<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = tid + 256 * u, r = idx / 64, row = idx % 64;
            if (idx < nrhs * 64) Xt[row * nrhs + r] = t[u];
        }
    }
    __syncthreads();
    C<T>* out = partial + (static_cast<int64_t>(blockIdx.y) * stride + blockIdx.x) * kNB * nrhs;
    for (int idx = tid; idx < nb * nrhs; idx += 256) {
        const int c = idx / nrhs, r = idx % nrhs;
        C<T> acc = C<T>{0, 0};
#pragma unroll 8
        for (int row = 0; row < 64; ++row) acc = gfma(acc, Lt[c * 64 + row], Xt[row * nrhs + r]);
        out[c * nrhs + r] = acc;
    }
}

// x_b = L_bb^-T (D_b^-1 y_b - sum of partials), written to X: a GEMV with the
// stored block inverse (L_bb^-T)[c][q] = Linv[q][c], q >= c. The partials are
// summed in block order (deterministic).
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_final(const int* __restrict__ nodes, int b, int nrhs, int stride,
                                                   FrontPlan fp, const C<T>* __restrict__ P,
                                                   const C<T>* __restrict__ xinv, const C<T>* __restrict__ Y,
                                                   C<T>* __restrict__ X, int n, const C<T>* __restrict__ partial) {
    constexpr int LD = kNB + 1;  // padded rows: lanes read along c
    __shared__ C<T> Xl[kNB * LD];
    __shared__ C<T> Z[16 * kNB];
    const int s = nodes[blockIdx.x];
    const Sn x = sn(fp, s);
    if (x.k <= b) return;
    const int nb = min(kNB, x.k - b);
    const int nblk = x.m > b + nb ? (x.m - b - nb + 63) / 64 : 0;
    const C<T>* Ps = P + fp.panel_off[s];
    const C<T>* Xg = xinv + fp.xinv_off[s] + static_cast<int64_t>(b / kNB) * kNB * kNB;
    const C<T>* part = partial + static_cast<int64_t>(blockIdx.x) * stride * kNB * nrhs;
    const int tid = threadIdx.x;
    {
        C<T> t[kNB * kNB / 256];
#pragma unroll
        for (int u = 0; u < kNB * kNB / 256; ++u) {
            const int idx = tid + 256 * u;
            t[u] = idx < nb * nb ? Xg[idx] : C<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < kNB * kNB / 256; ++u) {
            const int idx = tid + 256 * u;
            if (idx < nb * nb) Xl[(idx / nb) * LD + idx % nb] = t[u];  // Xl[j*LD + i] = Linv[i][j]
        }
    }
    for (int r0 = 0; r0 < nrhs; r0 += 16) {
        const int nr = min(16, nrhs - r0);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 16 * kNB / 256; ++u) {
            const int idx = tid + 256 * u, c = idx / 16, r = idx % 16;
            if (c < nb && r < nr) {
                C<T> t = C<T>{0, 0};
                const C<T>* pp = part + c * nrhs + (r0 + r);
#pragma unroll 8
                for (int blk = 0; blk < nblk; ++blk) t = gadd(t, pp[static_cast<int64_t>(blk) * kNB * nrhs]);
                const C<T> y = Y[static_cast<int64_t>(r0 + r) * n + x.first + b + c];
                const C<T> d = Ps[static_cast<int64_t>(b + c) * x.m + b + c];
                Z[r * kNB + c] = gsub(gmul(y, grecip(d)), t);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 16 * kNB / 256; ++u) {
            const int idx = tid + 256 * u, r = idx / kNB, c = idx % kNB;
            if (c < nb && r < nr) {
                C<T> acc = C<T>{0, 0};
                for (int q = c; q < nb; ++q) acc = gfma(acc, Xl[c * LD + q], Z[r * kNB + q]);
                X[static_cast<int64_t>(r0 + r) * n + x.first + b + c] = acc;
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
void launch_fwd_step(hipStream_t st, const int* nodes, int count, int b, int max_row_blocks, int nrhs,
                     const FrontPlan& fp, const void* P, const void* xinv, void* W, void* Y, int n) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fwd_block<T>, dim3(max_row_blocks > 0 ? max_row_blocks : 1, count), dim3(256), 0, st, nodes,
                       b, nrhs, fp, static_cast<const C<T>*>(P), static_cast<const C<T>*>(xinv), static_cast<C<T>*>(W),
                       static_cast<C<T>*>(Y), n);
}

template <typename T>
void launch_bwd_step(hipStream_t st, const int* nodes, int count, int b, int max_row_blocks, int nrhs,
                     const FrontPlan& fp, const void* P, const void* xinv, const void* Y, void* X, int n, void* partial) {
    if (count <= 0) return;
    const int stride = max_row_blocks > 0 ? max_row_blocks : 1;
    if (max_row_blocks > 0) {
        const size_t lds = static_cast<size_t>(kNB * 64 + nrhs * 64) * sizeof(C<T>);
        allow_large_lds(k_bwd_partial<T>, lds);
        hipLaunchKernelGGL(k_bwd_partial<T>, dim3(max_row_blocks, count), dim3(256), lds, st, nodes, b, nrhs, stride, fp,
                           static_cast<const C<T>*>(P), static_cast<const C<T>*>(X), n, static_cast<C<T>*>(partial));
    }
    hipLaunchKernelGGL(k_bwd_final<T>, dim3(count), dim3(256), 0, st, nodes, b, nrhs, stride, fp,
                       static_cast<const C<T>*>(P), static_cast<const C<T>*>(xinv), static_cast<const C<T>*>(Y),
                       static_cast<C<T>*>(X), n, static_cast<const C<T>*>(partial));
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
template void launch_fwd_step<double>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, const void*, const void*, void*, void*, int);
template void launch_fwd_step<float>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, const void*, const void*, void*, void*, int);
template void launch_bwd_step<double>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, const void*, const void*, const void*, void*, int, void*);
template void launch_bwd_step<float>(hipStream_t, const int*, int, int, int, int, const FrontPlan&, const void*, const void*, const void*, void*, int, void*);

}  // namespace aphi_dev

// Device refill (element kernels + deterministic slot reduction), symmetric Ruiz
// equilibration, and the FP64 operator kernels the solve needs (A*x residual,
// norms, permutation with scaling).
//
// Floating-point contraction is disabled in this file: the element entries and
// their reduction reproduce the reference's rounding step for step
// (src/basis_functions.cpp:15-139, src/element_matrix.cpp:27-117,
// src/assembly.cpp:255-422), so a device refill can be compared bit for bit.
#pragma clang fp contract(off)

#include "common.hpp"
#include "kernels.hpp"

namespace aphi_dev {

namespace {

// |a + j b| as the reference's std::abs(std::complex<double>) computes it on its
// host (glibc's hypot): Borges' method ("An improved algorithm for hypot(a, b)",
// 2019) without fused multiply-adds -- sqrt(ax^2 + ay^2) followed by one
// correction step -- with the inputs scaled by 2^-600 / 2^600 away from overflow
// and underflow, and ax + ay once ay is below ax's last bit. Equilibration row
// maxima and max |A_ij| then carry the reference's bits (a correctly rounded or
// differently rounded modulus moves them by an ulp now and then).
__device__ __forceinline__ double hypot_kernel(double ax, double ay) {
    double h = sqrt(ax * ax + ay * ay), t1, t2;
    if (h <= 2.0 * ay) {
        const double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        const double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}
__device__ __forceinline__ double abs_like_host(double a, double b) {
    if (!isfinite(a) || !isfinite(b)) return (isinf(a) || isinf(b)) ? INFINITY : a + b;
    a = fabs(a);
    b = fabs(b);
    const double ax = a < b ? b : a, ay = a < b ? a : b;
    constexpr double kScale = 0x1p-600, kLarge = 0x1p+511, kTiny = 0x1p-459, kEps = 0x1p-54;
    if (ax > kLarge) return ay <= ax * kEps ? ax + ay : hypot_kernel(ax * kScale, ay * kScale) / kScale;
    if (ay < kTiny) return ax >= ay / kEps ? ax + ay : hypot_kernel(ax / kScale, ay / kScale) * kScale;
    if (ay <= ax * kEps) return ax + ay;
    return hypot_kernel(ax, ay);
}

__constant__ int kEV[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {2, 3}, {1, 3}};

constexpr int kTetsPerBlock = 16;
constexpr int kLanesPerTet = 16;

// Gauss-Jordan with partial pivoting on [1 x y z] rows (src/basis_functions.cpp:15-57),
// fully unrolled so everything stays in registers. Returns false for a pivot
// below 1e-14 (a degenerate tet).
__device__ bool invert4(double m[4][4], double inv[4][4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) inv[i][j] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int col = 0; col < 4; ++col) {
        int prow = col;
        double pval = fabs(m[col][col]);
#pragma unroll
        for (int row = col + 1; row < 4; ++row) {
            if (fabs(m[row][col]) > pval) {
                prow = row;
                pval = fabs(m[row][col]);
            }
        }
        if (pval < 1e-14) return false;
#pragma unroll
        for (int row = col + 1; row < 4; ++row) {
            if (row == prow) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    double t = m[row][k]; m[row][k] = m[col][k]; m[col][k] = t;
                    t = inv[row][k]; inv[row][k] = inv[col][k]; inv[col][k] = t;
                }
            }
        }
        const double piv = m[col][col];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            m[col][k] /= piv;
            inv[col][k] /= piv;
        }
#pragma unroll
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            const double f = m[row][col];
            if (f == 0.0) continue;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                m[row][k] -= f * m[col][k];
                inv[row][k] -= f * inv[col][k];
            }
        }
    }
    return true;
}

struct Geo {
    double g[4][3];  // grad L_i
    double vol;
};

__device__ __forceinline__ void whitney(const Geo& G, int e, const double* L, double out[3]) {
    const int vi = kEV[e][0], vj = kEV[e][1];
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = G.g[vj][c] * L[vi] - G.g[vi][c] * L[vj];
}

__device__ __forceinline__ void p2grad(const Geo& G, int a, const double* L, double out[3]) {
    if (a < 4) {
        const double s = 4.0 * L[a] - 1.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c] = G.g[a][c] * s;
    } else {
        const int v0 = kEV[a - 4][0], v1 = kEV[a - 4][1];
        const double s1 = 4.0 * L[v1], s0 = 4.0 * L[v0];
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c] = G.g[v0][c] * s1 + G.g[v1][c] * s0;
    }
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// compute_tet_geometry (src/basis_functions.cpp:15-85) of tet t: barycentric
// gradients and |signed_tet_volume| (src/mesh.cpp:151-161). False for a
// degenerate tet.
__device__ bool tet_geometry(const double* __restrict__ coords, const int* __restrict__ tets, int t, Geo& G) {
    double m[4][4], inv[4][4];
    double p[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int v = tets[4 * t + j];
        p[j][0] = coords[3 * v + 0];
        p[j][1] = coords[3 * v + 1];
        p[j][2] = coords[3 * v + 2];
        m[j][0] = 1.0;
        m[j][1] = p[j][0];
        m[j][2] = p[j][1];
        m[j][3] = p[j][2];
    }
    if (!invert4(m, inv)) {
#pragma unroll
        for (int i = 0; i < 4; ++i) G.g[i][0] = G.g[i][1] = G.g[i][2] = 0.0;
        G.vol = 0.0;
        return false;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        G.g[i][0] = inv[1][i];
        G.g[i][1] = inv[2][i];
        G.g[i][2] = inv[3][i];
    }
    double a[3], b[3], c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        a[k] = p[1][k] - p[0][k];
        b[k] = p[2][k] - p[0][k];
        c[k] = p[3][k] - p[0][k];
    }
    const double x = a[1] * b[2] - a[2] * b[1];
    const double y = a[2] * b[0] - a[0] * b[2];
    const double z = a[0] * b[1] - a[1] * b[0];
    G.vol = fabs((x * c[0] + y * c[1] + z * c[2]) / 6.0);
    return true;
}

}  // namespace

// One 16-lane group per tet: lane 0 builds the geometry into LDS, then the 16
// lanes split the 196 element entries (36 AA, 60 APhi, 100 PhiPhi).
__global__ void __launch_bounds__(256) k_element(int num_tets, const double* __restrict__ coords,
                                                 const int* __restrict__ tets, const int* __restrict__ tet_body,
                                                 const BodyCoef* __restrict__ coef, cd* __restrict__ E,
                                                 int* __restrict__ error_flag) {
    __shared__ Geo geo[kTetsPerBlock];
    __shared__ double qp[4][4];
    const int local = threadIdx.x / kLanesPerTet;
    const int lane = threadIdx.x % kLanesPerTet;
    const int t = blockIdx.x * kTetsPerBlock + local;
    if (threadIdx.x == 0) {
        // src/quadrature.cpp:16-35
        const double s5 = sqrt(5.0);
        const double a = (5.0 + 3.0 * s5) / 20.0;
        const double b = (5.0 - s5) / 20.0;
        for (int k = 0; k < 4; ++k)
            for (int i = 0; i < 4; ++i) qp[k][i] = (i == k) ? a : b;
    }
    if (lane == 0 && t < num_tets) {
        Geo G;
        if (!tet_geometry(coords, tets, t, G)) atomicOr(error_flag, 1);
        geo[local] = G;
    }
    __syncthreads();
    if (t >= num_tets) return;
    const Geo& G = geo[local];
    const BodyCoef cf = coef[tet_body[t]];
    cd* out = E + static_cast<int64_t>(t) * 196;
    const double wvol = 0.25 * G.vol;

    for (int e = lane; e < 196; e += kLanesPerTet) {
        cd acc;
        if (e < 36) {
            // kernel_AA: nu*V*curl.curl, then alpha * (w * W_i.W_j) per point
            const int i = e / 6, j = e % 6;
            double ci[3], cj[3];
            {
                const int a0 = kEV[i][0], a1 = kEV[i][1], b0 = kEV[j][0], b1 = kEV[j][1];
                ci[0] = (G.g[a0][1] * G.g[a1][2] - G.g[a0][2] * G.g[a1][1]) * 2.0;
                ci[1] = (G.g[a0][2] * G.g[a1][0] - G.g[a0][0] * G.g[a1][2]) * 2.0;
                ci[2] = (G.g[a0][0] * G.g[a1][1] - G.g[a0][1] * G.g[a1][0]) * 2.0;
                cj[0] = (G.g[b0][1] * G.g[b1][2] - G.g[b0][2] * G.g[b1][1]) * 2.0;
                cj[1] = (G.g[b0][2] * G.g[b1][0] - G.g[b0][0] * G.g[b1][2]) * 2.0;
                cj[2] = (G.g[b0][0] * G.g[b1][1] - G.g[b0][1] * G.g[b1][0]) * 2.0;
            }
            acc.re = cf.nu * G.vol * dot3(ci, cj);
            acc.im = 0.0;
            if (!(cf.alpha.re == 0.0 && cf.alpha.im == 0.0)) {
                for (int k = 0; k < 4; ++k) {
                    double wi[3], wj[3];
                    whitney(G, i, qp[k], wi);
                    whitney(G, j, qp[k], wj);
                    const double s = wvol * dot3(wi, wj);
                    acc = cadd(acc, cscale(cf.alpha, s));
                }
            }
        } else if (e < 96) {
            // kernel_APhi: beta * (w * W_i . grad S_b)
            const int i = (e - 36) / 10, b = (e - 36) % 10;
            acc = {0.0, 0.0};
            if (!(cf.beta.re == 0.0 && cf.beta.im == 0.0)) {
                for (int k = 0; k < 4; ++k) {
                    double wi[3], gb[3];
                    whitney(G, i, qp[k], wi);
                    p2grad(G, b, qp[k], gb);
                    const double s = wvol * dot3(wi, gb);
                    acc = cadd(acc, cscale(cf.beta, s));
                }
            }
        } else {
            // kernel_PhiPhi: beta * (w * grad S_a . grad S_b)
            const int a = (e - 96) / 10, b = (e - 96) % 10;
            acc = {0.0, 0.0};
            if (!(cf.beta.re == 0.0 && cf.beta.im == 0.0)) {
                for (int k = 0; k < 4; ++k) {
                    double ga[3], gb[3];
                    p2grad(G, a, qp[k], ga);
                    p2grad(G, b, qp[k], gb);
                    const double s = wvol * dot3(ga, gb);
                    acc = cadd(acc, cscale(cf.beta, s));
                }
            }
        }
        out[e] = acc;
    }
}

// One thread per target (a matrix slot, or an rhs entry at num_slots + i). The
// contributions are summed in the recorded order, each formed exactly as the
// reference's scatter forms it.
__global__ void __launch_bounds__(256) k_assemble_reduce(int64_t num_targets, int64_t num_slots,
                                                         const int64_t* __restrict__ tptr,
                                                         const uint32_t* __restrict__ entry,
                                                         const uint8_t* __restrict__ kind,
                                                         const cd* __restrict__ E, const FreqScales* __restrict__ scales,
                                                         int nseg, const int64_t* __restrict__ seg_slots,
                                                         const int64_t* __restrict__ seg_rows,
                                                         cd* __restrict__ values, cd* __restrict__ rhs) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= num_targets) return;
    // the problem this target belongs to (a batch is several problems side by
    // side: their slots, then their rhs rows), by bisection of the offsets
    int seg = 0;
    if (nseg > 1) {
        const int64_t* off = t < num_slots ? seg_slots : seg_rows;
        const int64_t key = t < num_slots ? t : t - num_slots;
        int lo = 0, hi = nseg;  // off[lo] <= key < off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (off[mid] <= key) lo = mid;
            else hi = mid;
        }
        seg = lo;
    }
    const FreqScales& sc = scales[seg];
    cd acc = {0.0, 0.0};
    for (int64_t c = tptr[t]; c < tptr[t + 1]; ++c) {
        const int k = kind[c];
        const double sgn = (k & 0x10) ? -1.0 : 1.0;
        const int vid = k >> 5;
        switch (k & 0x0f) {
            case 0:  // (cp*cq) * aa
                acc = cadd(acc, cscale(E[entry[c]], sgn));
                break;
            case 1:  // (column * (ce*cn)) * ap
                acc = cadd(acc, cmul(cscale(sc.column, sgn), E[entry[c]]));
                break;
            case 2:  // ((row*jw) * (cn*ce)) * ap
                acc = cadd(acc, cmul(cscale(sc.row_jw, sgn), E[entry[c]]));
                break;
            case 3:  // ((row*column) * (cp*cq)) * pp
                acc = cadd(acc, cmul(cscale(sc.row_column, sgn), E[entry[c]]));
                break;
            case 4:  // rhs -= ((column*ce) * ap) * fixed
                acc = csub(acc, cmul(cmul(cscale(sc.column, sgn), E[entry[c]]), sc.over_column[vid]));
                break;
            case 5:  // rhs -= (((row*column)*cp) * pp) * fixed
                acc = csub(acc, cmul(cmul(cscale(sc.row_column, sgn), E[entry[c]]), sc.over_column[vid]));
                break;
            case 6:  // voltage port diagonal
                acc = cadd(acc, cd{1.0, 0.0});
                break;
            case 7:  // voltage port rhs: V / column
                acc = cadd(acc, sc.over_column[vid]);
                break;
            case 8:  // current port rhs: row * I
                acc = cadd(acc, sc.times_row[vid]);
                break;
            default:
                break;
        }
    }
    if (t < num_slots) values[t] = acc;
    else rhs[t - num_slots] = acc;
}

// ------------------------------------------------------------------ equilibration
// src/equilibration.cpp:94-139: every stored entry counts toward both its row and
// its column. Max is exact, so the atomic max is order independent.

__global__ void k_copy_cd(int64_t n, const cd* __restrict__ src, cd* __restrict__ dst) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_fill_flag(int* __restrict__ flag) {
    if (threadIdx.x == 0) *flag = 0;
}

__global__ void k_fill_d(int64_t n, double* __restrict__ p, double v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// compute_symmetric_equilibration's row maxima (src/equilibration.cpp:94-139),
// gathered: 16 lanes per row i take
// the stored entries of row i (CSR) and of column i (entries (r, i), r < i,
// through the transpose index), then a max over the 16 lanes: no atomics, and
// the maximum does not depend on the order of the comparisons.
__global__ void __launch_bounds__(256) k_eq_rowmax16(int n, const int* __restrict__ row_ptr, const int* __restrict__ tptr,
                                                     const int* __restrict__ tslot, const cd* __restrict__ w,
                                                     double* __restrict__ rowmax) {
    const int i = static_cast<int>(blockIdx.x) * 16 + static_cast<int>(threadIdx.x) / 16;
    const int q = static_cast<int>(threadIdx.x) % 16;
    double mx = 0.0;
    if (i < n) {
        for (int k = row_ptr[i] + q; k < row_ptr[i + 1]; k += 16) mx = fmax(mx, abs_like_host(w[k].re, w[k].im));
        for (int t = tptr[i] + q; t < tptr[i + 1]; t += 16) {
            const cd v = w[tslot[t]];
            mx = fmax(mx, abs_like_host(v.re, v.im));
        }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 16));
    if (i < n && q == 0) rowmax[i] = mx;
}

__global__ void k_eq_factor(int n, double* __restrict__ rowmax, double* __restrict__ factor, double* __restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double rm = rowmax[i];
    const double f = rm > 0.0 ? 1.0 / sqrt(rm) : 1.0;
    factor[i] = f;
    d[i] *= f;
    rowmax[i] = 0.0;
}

__global__ void k_eq_scale(int64_t nnz, const int* __restrict__ row_of, const int* __restrict__ col,
                           const double* __restrict__ factor, cd* __restrict__ w) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    w[k] = cscale(w[k], factor[row_of[k]] * factor[col[k]]);
}

// apply_symmetric_equilibration (src/equilibration.cpp:141-162): values * (d_i d_j, 0)
__global__ void k_eq_apply(int64_t nnz, const int* __restrict__ row_of, const int* __restrict__ col,
                           const double* __restrict__ d, const cd* __restrict__ a, cd* __restrict__ out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    out[k] = cmul(a[k], cd{d[row_of[k]] * d[col[k]], 0.0});
}

// ------------------------------------------------------------------ operator

// x (column r at x + r * ldx) -> groups of 16 right-hand sides interleaved,
// xi[(g * n + i) * 16 + q] = x column 16 g + q (zero past nrhs); one thread per
// (row, right-hand side)
__global__ void __launch_bounds__(256) k_interleave16(int n, int nrhs, const cd* __restrict__ x, int64_t ldx,
                                                      cd* __restrict__ xi) {
    const int i = static_cast<int>(blockIdx.x) * 16 + static_cast<int>(threadIdx.x) / 16;
    const int q = static_cast<int>(threadIdx.x) % 16, g = static_cast<int>(blockIdx.y);
    const int r = 16 * g + q;
    if (i >= n) return;
    xi[(static_cast<int64_t>(g) * n + i) * 16 + q] = r < nrhs ? x[r * ldx + i] : cd{0.0, 0.0};
}

// y = A x for the symmetric upper-stored A (include/aphi_solver/sparse_symmetric.hpp:116-131),
// accumulated per row in the reference's order: the mirrored entries (r, i), r < i
// ascending, then row i's own entries. x interleaved by k_interleave16, y
// column-major with leading dim ldy. Sixteen consecutive threads take one row
// and the sixteen right-hand sides of a group: each matrix entry is read once
// per group and the x values it multiplies are one contiguous 256-byte segment.
__global__ void __launch_bounds__(256) k_symv16(int n, int nrhs, const int* __restrict__ row_ptr,
                                                const int* __restrict__ col, const int* __restrict__ tptr,
                                                const int* __restrict__ tslot, const int* __restrict__ trow,
                                                const cd* __restrict__ a, const cd* __restrict__ xi,
                                                cd* __restrict__ y, int64_t ldy) {
    const int i = static_cast<int>(blockIdx.x) * 16 + static_cast<int>(threadIdx.x) / 16;
    const int q = static_cast<int>(threadIdx.x) % 16, g = static_cast<int>(blockIdx.y);
    const int r = 16 * g + q;
    if (i >= n) return;
    const cd* xg = xi + static_cast<int64_t>(g) * n * 16 + q;
    cd acc = {0.0, 0.0};
    // eight terms at a time: their indices, then their operands, are all loaded
    // before the first product (one memory latency per eight terms); the sum
    // still runs term by term in the reference's order
    constexpr int U = 8;
    const int t0 = tptr[i], t1 = tptr[i + 1];
    for (int t = t0; t < t1; t += U) {
        int sl[U], rw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int tt = min(t + u, t1 - 1);
            sl[u] = tslot[tt];
            rw[u] = trow[tt];
        }
        cd av[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            av[u] = a[sl[u]];
            xv[u] = xg[static_cast<int64_t>(rw[u]) * 16];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (t + u < t1) acc = cadd(acc, cmul(av[u], xv[u]));
    }
    const int k0 = row_ptr[i], k1 = row_ptr[i + 1];
    for (int k = k0; k < k1; k += U) {
        int cl[U];
        cd av[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k + u, k1 - 1);
            cl[u] = col[kk];
            av[u] = a[kk];
        }
        cd xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = xg[static_cast<int64_t>(cl[u]) * 16];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k + u < k1) acc = cadd(acc, cmul(av[u], xv[u]));
    }
    if (r < nrhs) y[r * ldy + i] = acc;
}

// r = b - A x, in place into r (which holds A x on entry)
__global__ void k_residual(int n, int nrhs, const cd* __restrict__ b, int64_t ldb, cd* __restrict__ r, int64_t ldr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nrhs) return;
    r[c * ldr + i] = csub(b[c * ldb + i], r[c * ldr + i]);
}

// Per-RHS sums of |v|^2 over rows, into partials[c * nblocks + blockIdx.x]; a
// fixed-shape tree so the result does not depend on scheduling.
__global__ void __launch_bounds__(256) k_sumsq_partial(int n, const cd* __restrict__ v, int64_t ldv,
                                                       double* __restrict__ partials) {
    __shared__ double red[256];
    const int c = blockIdx.y;
    const cd* col = v + c * ldv;
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += col[i].re * col[i].re + col[i].im * col[i].im;
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[c * gridDim.x + blockIdx.x] = red[0];
}

// b~[i] = d[perm[i]] * b[perm[i]]  (scale_rhs then P b; the order of the product
// (b * (d, 0)) follows src/equilibration.cpp:73-82)
__global__ void k_scale_permute_in(int n, int nrhs, const int* __restrict__ perm, const double* __restrict__ d,
                                   const cd* __restrict__ b, int64_t ldb, cd* __restrict__ out, int64_t ldo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nrhs) return;
    const int p = perm[i];
    out[c * ldo + i] = cmul(b[c * ldb + p], cd{d[p], 0.0});
}

// x[perm[i]] = d[perm[i]] * x~[i]  (P^T then recover_equilibrated_solution)
__global__ void k_permute_scale_out(int n, int nrhs, const int* __restrict__ perm, const double* __restrict__ d,
                                    const cd* __restrict__ xt, int64_t ldt, cd* __restrict__ x, int64_t ldx, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nrhs) return;
    const int p = perm[i];
    const cd v = cmul(xt[c * ldt + i], cd{d[p], 0.0});
    x[c * ldx + p] = accumulate ? cadd(x[c * ldx + p], v) : v;
}

// ------------------------------------------------------------------ GMRES refinement
//
// Vectors are complex128, n per right-hand side; basis vector v of RHS r lives at
// V + v * vstride + r * n. Reductions go through fixed-shape partial sums.

// out[r][i] = in[r][i] * d[i] (power = 1) or in[r][i] / d[i] (power = -1)
__global__ void k_scale_d(int n, int nrhs, const double* __restrict__ d, int power, const cd* __restrict__ in,
                          int64_t ldi, cd* __restrict__ out, int64_t ldo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n || r >= nrhs) return;
    const cd v = in[r * ldi + i];
    out[r * ldo + i] = power > 0 ? cscale(v, d[i]) : cd{v.re / d[i], v.im / d[i]};
}

// A batch holds several problems side by side: problem s owns rows
// [off[s], off[s + 1]) of every vector, and each (problem, right-hand side) pair
// is a column of its own, c = r * nseg + s, with its own inner products and
// coefficients. One problem: nseg = 1, off = {0, n}, row_seg = nullptr.
__device__ __forceinline__ int seg_of(const int* __restrict__ row_seg, int i) { return row_seg ? row_seg[i] : 0; }

// partials[(v * ncol + c) * gridDim.x + blk] = sum over the block's rows of
// column c of conj(V_v) w; grid (parts, nvec, ncol)
__global__ void __launch_bounds__(256) k_dot_partial(int n, int nrhs, int nseg, const int64_t* __restrict__ off,
                                                     const cd* __restrict__ V, int64_t vstride,
                                                     const cd* __restrict__ w, cd* __restrict__ partials) {
    __shared__ double rre[256], rim[256];
    const int v = blockIdx.y, c = blockIdx.z;
    const int r = c / nseg, sg = c % nseg;
    const cd* a = V + v * vstride + static_cast<int64_t>(r) * n;
    const cd* b = w + static_cast<int64_t>(r) * n;
    const int lo = static_cast<int>(off[sg]), hi = static_cast<int>(off[sg + 1]);
    double sr = 0.0, si = 0.0;
    for (int i = lo + static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x); i < hi; i += gridDim.x * 256) {
        const cd x = a[i], y = b[i];
        sr += x.re * y.re + x.im * y.im;
        si += x.re * y.im - x.im * y.re;
    }
    rre[threadIdx.x] = sr;
    rim[threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            rre[threadIdx.x] += rre[threadIdx.x + s];
            rim[threadIdx.x] += rim[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partials[(static_cast<int64_t>(v) * nrhs * nseg + c) * gridDim.x + blockIdx.x] = cd{rre[0], rim[0]};
}

// w[r][i] -= sum_v h[v * ncol + c] V_v[r][i], c = r * nseg + (problem of row i)
__global__ void k_sub_basis(int n, int nrhs, int nseg, const int* __restrict__ row_seg, int nvec,
                            const cd* __restrict__ V, int64_t vstride, const cd* __restrict__ h, cd* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n || r >= nrhs) return;
    const int c = r * nseg + seg_of(row_seg, i), ncol = nrhs * nseg;
    cd acc = w[static_cast<int64_t>(r) * n + i];
    for (int v = 0; v < nvec; ++v) acc = csub(acc, cmul(h[v * ncol + c], V[v * vstride + static_cast<int64_t>(r) * n + i]));
    w[static_cast<int64_t>(r) * n + i] = acc;
}

// out[r][i] = w[r][i] * s[c]
__global__ void k_scale_cols(int n, int nrhs, int nseg, const int* __restrict__ row_seg, const cd* __restrict__ w,
                             const double* __restrict__ s, cd* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n || r >= nrhs) return;
    out[static_cast<int64_t>(r) * n + i] = cscale(w[static_cast<int64_t>(r) * n + i], s[r * nseg + seg_of(row_seg, i)]);
}

// x[r][i] += sum_v y[v * ncol + c] Z_v[r][i]
__global__ void k_add_basis(int n, int nrhs, int nseg, const int* __restrict__ row_seg, int nvec,
                            const cd* __restrict__ Z, int64_t zstride, const cd* __restrict__ y, cd* __restrict__ x,
                            int64_t ldx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n || r >= nrhs) return;
    const int c = r * nseg + seg_of(row_seg, i), ncol = nrhs * nseg;
    cd acc = x[r * ldx + i];
    for (int v = 0; v < nvec; ++v) acc = cadd(acc, cmul(y[v * ncol + c], Z[v * zstride + static_cast<int64_t>(r) * n + i]));
    x[r * ldx + i] = acc;
}

// Per column c = r * nseg + s: the sum over the problem's rows of |v_i|^2
// (power 0), (d_i |v_i|)^2 (power 1) or (|v_i| / d_i)^2 (power -1), as partials
// partials[c * gridDim.x + blk] of a fixed-shape tree; grid (parts, ncol)
__global__ void __launch_bounds__(256) k_seg_sumsq(int nseg, const int64_t* __restrict__ off, const cd* __restrict__ v,
                                                   int64_t ldv, const double* __restrict__ d, int power,
                                                   double* __restrict__ partials) {
    __shared__ double red[256];
    const int c = blockIdx.y;
    const int r = c / nseg, sg = c % nseg;
    const cd* col = v + r * ldv;
    const int64_t lo = off[sg], hi = off[sg + 1];
    double s = 0.0;
    for (int64_t i = lo + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < hi; i += static_cast<int64_t>(gridDim.x) * 256) {
        if (power == 0) {
            s += col[i].re * col[i].re + col[i].im * col[i].im;
        } else {
            const double w = power > 0 ? d[i] : 1.0 / d[i];
            const double re = col[i].re * w, im = col[i].im * w;
            s += re * re + im * im;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[c * gridDim.x + blockIdx.x] = red[0];
}

// Per problem s: max |v_k| over its stored entries [off[s], off[s + 1]), as
// partials[s * gridDim.x + blk]; grid (parts, nseg)
__global__ void __launch_bounds__(256) k_seg_absmax(const int64_t* __restrict__ off, const cd* __restrict__ v,
                                                    double* __restrict__ partials) {
    __shared__ double red[256];
    const int sg = blockIdx.y;
    double s = 0.0;
    for (int64_t i = off[sg] + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < off[sg + 1];
         i += static_cast<int64_t>(gridDim.x) * 256)
        s = fmax(s, abs_like_host(v[i].re, v[i].im));
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[sg * gridDim.x + blockIdx.x] = red[0];
}

// ------------------------------------------------------------------ port current

// Per tet and right-hand side: the centroid current density of compute_fields
// (src/postprocess.cpp:311-400) -- A from the edge unknowns with the global
// edge signs, Phi from the Phi unknowns (scaled back under ScaledPhi) or the
// prescribed values, E = -j w A - grad Phi, J = sigma E -- and its flux term of
// terminal_current (src/postprocess.cpp:538-544): (J . axis) |V|. Written in the
// reference's operation order (contraction off), so each term is the
// reference's bit for bit. Non-conducting tets contribute nothing.
namespace {
// E = -j w A - grad Phi at the centroid of tet t (compute_fields' evaluate at
// kCentroid, src/postprocess.cpp:357-391), in the reference's operation order.
__device__ __forceinline__ void centroid_e(const TetFieldPlan& tp, const cd* __restrict__ xr, int t, const Geo& G,
                                           cd minus_jw, cd scale, cd e[3]) {
    const int* d = tp.dofs + 16 * static_cast<int64_t>(t);
    const double L[4] = {0.25, 0.25, 0.25, 0.25};
    cd av[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const cd ae = d[k] >= 0 ? xr[d[k]] : cd{0.0, 0.0};
        double w[3];
        whitney(G, k, L, w);
        const double sg = static_cast<double>(tp.signs[6 * static_cast<int64_t>(t) + k]);
#pragma unroll
        for (int c = 0; c < 3; ++c) av[c] = cadd(av[c], cscale(ae, w[c] * sg));
    }
    cd gp[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int b = 0; b < 10; ++b) {
        const int ib = d[6 + b];
        cd pn;
        if (ib >= 0) pn = cmul(cscale(scale, tp.phi_coeff[10 * static_cast<int64_t>(t) + b]), xr[ib]);
        else pn = tp.phi_fixed[10 * static_cast<int64_t>(t) + b];
        double gr[3];
        p2grad(G, b, L, gr);
#pragma unroll
        for (int c = 0; c < 3; ++c) gp[c] = cadd(gp[c], cscale(pn, gr[c]));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] = csub(cmul(minus_jw, av[c]), gp[c]);
}
}  // namespace

__global__ void __launch_bounds__(256) k_tet_current(int num_tets, TetFieldPlan tp, int nrhs, const cd* __restrict__ x,
                                                     int64_t ldx, cd minus_jw, cd scale, double ax, double ay, double az,
                                                     int nseg, const int64_t* __restrict__ seg_tets,
                                                     const cd* __restrict__ seg_factors, cd* __restrict__ contrib) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (t >= num_tets || r >= nrhs) return;
    if (nseg > 1) {
        // a batch: the problem of tet t brings its own -j w and Phi scale
        int lo = 0, hi = nseg;
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (seg_tets[mid] <= t) lo = mid;
            else hi = mid;
        }
        minus_jw = seg_factors[2 * lo];
        scale = seg_factors[2 * lo + 1];
    }
    const double sigma = tp.body_sigma[tp.tet_body[t]];
    cd out = cd{0.0, 0.0};
    if (sigma > 0.0) {
        Geo G;
        tet_geometry(tp.coords, tp.tets, t, G);
        cd e[3];
        centroid_e(tp, x + static_cast<int64_t>(r) * ldx, t, G, minus_jw, scale, e);
        cd j[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) j[c] = cscale(e[c], sigma);
        out = cscale(cadd(cadd(cscale(j[0], ax), cscale(j[1], ay)), cscale(j[2], az)), G.vol);
    }
    contrib[static_cast<int64_t>(r) * num_tets + t] = out;
}

// e_tet of compute_fields (src/postprocess.cpp:389-391) for every tet, one per thread.
__global__ void __launch_bounds__(256) k_tet_efield(int num_tets, TetFieldPlan tp, const cd* __restrict__ x, cd minus_jw,
                                                    cd scale, cd* __restrict__ e_tet) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tets) return;
    Geo G;
    tet_geometry(tp.coords, tp.tets, t, G);
    cd e[3];
    centroid_e(tp, x, t, G, minus_jw, scale, e);
#pragma unroll
    for (int c = 0; c < 3; ++c) e_tet[3 * static_cast<int64_t>(t) + c] = e[c];
}

// ------------------------------------------------------------------ compute_fields
//
// FieldOutput of compute_fields (src/postprocess.cpp:223-481). First one thread
// per tet: the edge unknowns and resolved Phi coefficients, B = curl A, A and E at
// the centroid (J = sigma E) and A and E at each of the ten P2 nodes, every term
// in the reference's operation order (contraction off in this file). Then one
// thread per P2 node: the volume-weighted sums over its (tet, local node)
// entries in tet order -- the order the reference's tet loop adds them -- with
// the side selection of the most conducting / most permeable tets, and the
// normalisation and constitutive relations at the node.
namespace {
constexpr double kMu0 = 4.0e-7 * 3.14159265358979323846;

// compute_fields' evaluate (src/postprocess.cpp:364-386) at barycentric point L
__device__ __forceinline__ void eval_ae(const Geo& G, const double* L, const cd a[6], const double sg[6],
                                        const cd ph[10], cd minus_jw, cd av[3], cd ev[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) av[c] = cd{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double w[3];
        whitney(G, k, L, w);
#pragma unroll
        for (int c = 0; c < 3; ++c) av[c] = cadd(av[c], cscale(a[k], w[c] * sg[k]));
    }
    cd gp[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int b = 0; b < 10; ++b) {
        double gr[3];
        p2grad(G, b, L, gr);
#pragma unroll
        for (int c = 0; c < 3; ++c) gp[c] = cadd(gp[c], cscale(ph[b], gr[c]));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) ev[c] = csub(cmul(minus_jw, av[c]), gp[c]);
}
}  // namespace

__global__ void __launch_bounds__(256) k_fields_tet(int num_tets, TetFieldPlan tp, const double* __restrict__ body_mu_r,
                                                    const cd* __restrict__ x, cd minus_jw, cd scale, cd* __restrict__ a_tet,
                                                    cd* __restrict__ b_tet, cd* __restrict__ e_tet, cd* __restrict__ j_tet,
                                                    double* __restrict__ sigma_tet, double* __restrict__ tet_vol,
                                                    double* __restrict__ tet_mu, cd* __restrict__ node_val) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tets) return;
    Geo G;
    tet_geometry(tp.coords, tp.tets, t, G);
    const int* d = tp.dofs + 16 * static_cast<int64_t>(t);
    cd a[6], ph[10];
    double sg[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        a[k] = d[k] >= 0 ? x[d[k]] : cd{0.0, 0.0};
        sg[k] = static_cast<double>(tp.signs[6 * static_cast<int64_t>(t) + k]);
    }
#pragma unroll
    for (int b = 0; b < 10; ++b) {
        const int ib = d[6 + b];
        ph[b] = ib >= 0 ? cmul(cscale(scale, tp.phi_coeff[10 * static_cast<int64_t>(t) + b]), x[ib])
                        : tp.phi_fixed[10 * static_cast<int64_t>(t) + b];
    }
    // B = curl A: (grad L_vi x grad L_vj) * 2, then the edge sign (src/basis_functions.cpp:101-116)
    cd bc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double* gi = G.g[kEV[k][0]];
        const double* gj = G.g[kEV[k][1]];
        const double cr[3] = {(gi[1] * gj[2] - gi[2] * gj[1]) * 2.0, (gi[2] * gj[0] - gi[0] * gj[2]) * 2.0,
                              (gi[0] * gj[1] - gi[1] * gj[0]) * 2.0};
#pragma unroll
        for (int c = 0; c < 3; ++c) bc[c] = cadd(bc[c], cscale(a[k], cr[c] * sg[k]));
    }
    const int body = tp.tet_body[t];
    const double sigma = tp.body_sigma[body];
    sigma_tet[t] = sigma;
    tet_vol[t] = G.vol;
    tet_mu[t] = kMu0 * body_mu_r[body];
    const double Lc[4] = {0.25, 0.25, 0.25, 0.25};
    cd av[3], ev[3];
    eval_ae(G, Lc, a, sg, ph, minus_jw, av, ev);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        b_tet[3 * static_cast<int64_t>(t) + c] = bc[c];
        a_tet[3 * static_cast<int64_t>(t) + c] = av[c];
        e_tet[3 * static_cast<int64_t>(t) + c] = ev[c];
        j_tet[3 * static_cast<int64_t>(t) + c] = cscale(ev[c], sigma);
    }
    for (int k = 0; k < 10; ++k) {
        double L[4] = {0.0, 0.0, 0.0, 0.0};
        if (k < 4) {
            L[k] = 1.0;
        } else {
            L[kEV[k - 4][0]] = 0.5;
            L[kEV[k - 4][1]] = 0.5;
        }
        eval_ae(G, L, a, sg, ph, minus_jw, av, ev);
        cd* o = node_val + 6 * (10 * static_cast<int64_t>(t) + k);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            o[c] = av[c];
            o[3 + c] = ev[c];
        }
    }
}

__global__ void __launch_bounds__(256) k_fields_node(int num_p2, const int64_t* __restrict__ node_ptr,
                                                     const int* __restrict__ node_entry,
                                                     const double* __restrict__ sigma_tet,
                                                     const double* __restrict__ tet_mu, const double* __restrict__ tet_vol,
                                                     const cd* __restrict__ b_tet, const cd* __restrict__ node_val,
                                                     cd* __restrict__ a_node, cd* __restrict__ b_node,
                                                     cd* __restrict__ h_node, cd* __restrict__ e_node,
                                                     cd* __restrict__ j_node, double* __restrict__ vertex_weight,
                                                     int8_t* __restrict__ on_interface) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= num_p2) return;
    const int64_t e0 = node_ptr[u], e1 = node_ptr[u + 1];
    // std::min(a, b) = b < a ? b : a; std::max(a, b) = a < b ? b : a
    double sig_min = __builtin_inf(), sig_max = -__builtin_inf(), mu_min = __builtin_inf(), mu_max = -__builtin_inf();
    for (int64_t e = e0; e < e1; ++e) {
        const int t = node_entry[e] / 10;
        const double s = sigma_tet[t], m = tet_mu[t];
        sig_min = s < sig_min ? s : sig_min;
        sig_max = sig_max < s ? s : sig_max;
        mu_min = m < mu_min ? m : mu_min;
        mu_max = mu_max < m ? m : mu_max;
    }
    const bool sigma_uniform = sig_min == sig_max, mu_uniform = mu_min == mu_max;
    double vw = 0.0, ew = 0.0, bw = 0.0;
    cd a[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}}, ef[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}},
       bf[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    int8_t interface_flag = 0;
    for (int64_t e = e0; e < e1; ++e) {
        const int ent = node_entry[e], t = ent / 10;
        const double vol = tet_vol[t];
        const cd* o = node_val + 6 * static_cast<int64_t>(ent);
        vw = vw + vol;
#pragma unroll
        for (int c = 0; c < 3; ++c) a[c] = cadd(a[c], cscale(o[c], vol));
        if (sigma_uniform || sigma_tet[t] == sig_max) {
            ew = ew + vol;
#pragma unroll
            for (int c = 0; c < 3; ++c) ef[c] = cadd(ef[c], cscale(o[3 + c], vol));
        }
        if (mu_uniform || tet_mu[t] == mu_max) {
            bw = bw + vol;
#pragma unroll
            for (int c = 0; c < 3; ++c) bf[c] = cadd(bf[c], cscale(b_tet[3 * static_cast<int64_t>(t) + c], vol));
        }
        if (!sigma_uniform || !mu_uniform) interface_flag = 1;
    }
    vertex_weight[u] = vw;
    on_interface[u] = interface_flag;
    cd h[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}}, j[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (vw > 0.0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a[c] = cd{a[c].re / vw, a[c].im / vw};
        if (ew > 0.0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) ef[c] = cd{ef[c].re / ew, ef[c].im / ew};
        }
        if (bw > 0.0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) bf[c] = cd{bf[c].re / bw, bf[c].im / bw};
        }
        const double nu_node = 1.0 / mu_max;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            j[c] = cscale(ef[c], sig_max);
            h[c] = cscale(bf[c], nu_node);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a_node[3 * static_cast<int64_t>(u) + c] = a[c];
        e_node[3 * static_cast<int64_t>(u) + c] = ef[c];
        b_node[3 * static_cast<int64_t>(u) + c] = bf[c];
        j_node[3 * static_cast<int64_t>(u) + c] = j[c];
        h_node[3 * static_cast<int64_t>(u) + c] = h[c];
    }
}

// ------------------------------------------------------------------ gauge-free potential
//
// Per tet of a participating body (src/gauge_free_potential.cpp:112-151), the
// sixteen stiffness terms k_ij = (beta |V|) (grad L_i . grad L_j) and the four
// source terms (beta |V|) (E . grad L_i), in the reference's operation order.
__global__ void __launch_bounds__(256) k_gf_tet(int num_tets, const double* __restrict__ coords,
                                                const int* __restrict__ tets, const int* __restrict__ tet_body,
                                                const int* __restrict__ tet_on, const cd* __restrict__ beta,
                                                const cd* __restrict__ e_tet, cd* __restrict__ terms) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tets || !tet_on[t]) return;
    Geo G;
    tet_geometry(coords, tets, t, G);
    const cd bv = cscale(beta[tet_body[t]], G.vol);
    cd* out = terms + 20 * static_cast<int64_t>(t);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) out[4 * i + j] = cscale(bv, dot3(G.g[i], G.g[j]));
    const cd* E = e_tet + 3 * static_cast<int64_t>(t);
    const cd e0 = E[0], e1 = E[1], e2 = E[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const cd edotg = cadd(cadd(cscale(e0, G.g[i][0]), cscale(e1, G.g[i][1])), cscale(e2, G.g[i][2]));
        out[16 + i] = cmul(bv, edotg);
    }
}

// One thread per slot / rhs row: its terms summed from zero in plan order (the
// reference's tet, i, j order): slot += k_ij; rhs -= source_i; rhs -= k_ij * value.
__global__ void __launch_bounds__(256) k_gf_reduce(int64_t num_targets, int64_t nnz, const int64_t* __restrict__ tptr,
                                                   const uint32_t* __restrict__ entry, const uint8_t* __restrict__ kind,
                                                   const cd* __restrict__ values, const cd* __restrict__ terms,
                                                   cd* __restrict__ A, cd* __restrict__ rhs) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= num_targets) return;
    cd acc{0.0, 0.0};
    for (int64_t c = tptr[t]; c < tptr[t + 1]; ++c) {
        const int k = kind[c];
        const cd v = terms[entry[c]];
        if ((k & 3) == 0) acc = cadd(acc, v);
        else if ((k & 3) == 1) acc = csub(acc, v);
        else acc = csub(acc, cmul(v, values[k >> 2]));
    }
    if (t < nnz) A[t] = acc;
    else rhs[t - nnz] = acc;
}

// phi_vertex[v] = x[column[v]], or the prescribed value (src/gauge_free_potential.cpp:167-171)
__global__ void __launch_bounds__(256) k_gf_scatter(int num_nodes, const int* __restrict__ column,
                                                    const cd* __restrict__ prescribed, const cd* __restrict__ x,
                                                    cd* __restrict__ phi) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= num_nodes) return;
    const int c = column[v];
    phi[v] = c >= 0 ? x[c] : prescribed[v];
}

// current[r] = -(sum over conducting tets, in tet order, of the flux terms) / length:
// one wavefront per right-hand side adds the terms of 64 tets at a time lane by
// lane (the conducting lanes only, in ascending order), so the sum is the
// reference's sequential one. The terms of the next kCurBatch chunks of 64 are
// loaded while the current ones are added, so the memory latency overlaps the
// additions.
constexpr int kCurBatch = 4;
__global__ void __launch_bounds__(64) k_current_sum(int num_tets, int nseg, const int64_t* __restrict__ seg_tets,
                                                    TetFieldPlan tp, const cd* __restrict__ contrib,
                                                    const double* __restrict__ lengths, double length,
                                                    cd* __restrict__ current) {
    // block c = r * nseg + s sums problem s's tets in order for right-hand side r
    const int c = blockIdx.x, r = c / nseg, sg = c % nseg;
    const int lane = threadIdx.x;
    const cd* cr = contrib + static_cast<int64_t>(r) * num_tets;
    const int lo = static_cast<int>(seg_tets[sg]), hi = static_cast<int>(seg_tets[sg + 1]);
    double fre = 0.0, fim = 0.0;
    cd v[kCurBatch], vn[kCurBatch];
    bool on[kCurBatch], onn[kCurBatch];
    const auto load = [&](int tb, cd (&w)[kCurBatch], bool (&f)[kCurBatch]) {
#pragma unroll
        for (int u = 0; u < kCurBatch; ++u) {
            const int t = tb + 64 * u + lane, tc = max(lo, min(t, hi - 1));  // an index inside the problem
            f[u] = t < hi && tp.body_sigma[tp.tet_body[tc]] > 0.0;
            w[u] = cr[tc];
        }
    };
    if (lo < hi) load(lo, v, on);
    for (int tb = lo; tb < hi; tb += 64 * kCurBatch) {
        if (tb + 64 * kCurBatch < hi) load(tb + 64 * kCurBatch, vn, onn);
#pragma unroll
        for (int u = 0; u < kCurBatch; ++u) {
            // the conducting lanes in ascending order (most chunks have none)
            unsigned long long mask = __ballot(on[u]);
            while (mask != 0ull) {
                const int l = __builtin_ctzll(mask);
                fre = fre + lane_read(v[u].re, l);
                fim = fim + lane_read(v[u].im, l);
                mask &= mask - 1ull;
            }
        }
#pragma unroll
        for (int u = 0; u < kCurBatch; ++u) {
            v[u] = vn[u];
            on[u] = onn[u];
        }
    }
    const double len = lengths ? lengths[sg] : length;
    if (lane == 0) current[c] = cd{-fre / len, -fim / len};
}

}  // namespace aphi_dev

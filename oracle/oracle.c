#define _POSIX_C_SOURCE 199309L
/* CPU oracle of the A-Phi hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's algorithms for this path, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * Nothing on the product path links or calls it. Each function cites the
 * reference routine it follows; the arithmetic is written in the reference's
 * evaluation order and built with -ffp-contract=off, so that on x86-64 it rounds
 * as the reference does.
 *
 * Pinned by (see tests/test_oracle.py): the reference's own golden vectors
 * (the .npz fixtures under tests/golden: assembled values and rhs bit for bit, equilibration d,
 * D, x, backward error), the closed forms of tests/test_element_matrix.cpp
 * (unit-tet curl-curl table, partition of unity), LDL^T reconstruction against
 * the matrix, agreement with a dense LU solve, and the 2x2 hand factor
 * [[2,3],[3,5]] -> L = 3/2, D = (2, 1/2).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef double _Complex cplx;

/* ------------------------------------------------------------------ element kernels */

static const int kEV[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {2, 3}, {1, 3}};

/* compute_tet_geometry + invert4x4, src/basis_functions.cpp:15-85. Returns 0 for
 * a degenerate tet. g[i][c] = d L_i / d x_c. */
static int tet_geometry(const double p[4][3], double g[4][3], double* vol) {
    double m[4][4], inv[4][4];
    for (int j = 0; j < 4; ++j) {
        m[j][0] = 1.0;
        m[j][1] = p[j][0];
        m[j][2] = p[j][1];
        m[j][3] = p[j][2];
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) inv[i][j] = i == j ? 1.0 : 0.0;
    for (int col = 0; col < 4; ++col) {
        int pr = col;
        double pv = fabs(m[col][col]);
        for (int row = col + 1; row < 4; ++row)
            if (fabs(m[row][col]) > pv) {
                pr = row;
                pv = fabs(m[row][col]);
            }
        if (pv < 1e-14) return 0;
        if (pr != col)
            for (int k = 0; k < 4; ++k) {
                double t = m[pr][k]; m[pr][k] = m[col][k]; m[col][k] = t;
                t = inv[pr][k]; inv[pr][k] = inv[col][k]; inv[col][k] = t;
            }
        const double piv = m[col][col];
        for (int k = 0; k < 4; ++k) {
            m[col][k] /= piv;
            inv[col][k] /= piv;
        }
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            const double f = m[row][col];
            if (f == 0.0) continue;
            for (int k = 0; k < 4; ++k) {
                m[row][k] -= f * m[col][k];
                inv[row][k] -= f * inv[col][k];
            }
        }
    }
    for (int i = 0; i < 4; ++i) {
        g[i][0] = inv[1][i];
        g[i][1] = inv[2][i];
        g[i][2] = inv[3][i];
    }
    /* signed_tet_volume, src/mesh.cpp:151-161 */
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = p[1][k] - p[0][k];
        b[k] = p[2][k] - p[0][k];
        c[k] = p[3][k] - p[0][k];
    }
    const double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
    *vol = fabs((x * c[0] + y * c[1] + z * c[2]) / 6.0);
    return 1;
}

static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* whitney_edge_value, src/basis_functions.cpp:95-101 */
static void whitney(double g[4][3], int e, const double* L, double* out) {
    const int vi = kEV[e][0], vj = kEV[e][1];
    for (int c = 0; c < 3; ++c) out[c] = g[vj][c] * L[vi] - g[vi][c] * L[vj];
}

/* p2_nodal_gradient, src/basis_functions.cpp:127-139 */
static void p2grad(double g[4][3], int a, const double* L, double* out) {
    if (a < 4) {
        const double s = 4.0 * L[a] - 1.0;
        for (int c = 0; c < 3; ++c) out[c] = g[a][c] * s;
    } else {
        const int v0 = kEV[a - 4][0], v1 = kEV[a - 4][1];
        const double s1 = 4.0 * L[v1], s0 = 4.0 * L[v0];
        for (int c = 0; c < 3; ++c) out[c] = g[v0][c] * s1 + g[v1][c] * s0;
    }
}

static cplx cxd(cplx a, double s) { return CMPLX(creal(a) * s, cimag(a) * s); }

/* kernel_AA / kernel_APhi / kernel_PhiPhi, src/element_matrix.cpp:27-117, with
 * the degree-2 rule of src/quadrature.cpp:16-35. out: 36 AA, 60 APhi, 100 PhiPhi
 * (complex, row-major within each block). Returns 0 for a degenerate tet. */
int oracle_element(const double* coords12, double nu, const double* alpha, const double* beta, double* out392) {
    double p[4][3], g[4][3], vol;
    for (int j = 0; j < 4; ++j)
        for (int c = 0; c < 3; ++c) p[j][c] = coords12[3 * j + c];
    if (!tet_geometry(p, g, &vol)) return 0;
    const double s5 = sqrt(5.0);
    const double qa = (5.0 + 3.0 * s5) / 20.0, qb = (5.0 - s5) / 20.0;
    double qp[4][4];
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 4; ++i) qp[k][i] = i == k ? qa : qb;
    const cplx al = CMPLX(alpha[0], alpha[1]), be = CMPLX(beta[0], beta[1]);
    cplx* o = (cplx*)out392;
    double curl[6][3];
    for (int i = 0; i < 6; ++i) {
        const int a0 = kEV[i][0], a1 = kEV[i][1];
        curl[i][0] = (g[a0][1] * g[a1][2] - g[a0][2] * g[a1][1]) * 2.0;
        curl[i][1] = (g[a0][2] * g[a1][0] - g[a0][0] * g[a1][2]) * 2.0;
        curl[i][2] = (g[a0][0] * g[a1][1] - g[a0][1] * g[a1][0]) * 2.0;
    }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) o[i * 6 + j] = CMPLX(nu * vol * dot3(curl[i], curl[j]), 0.0);
    const int have_alpha = !(creal(al) == 0.0 && cimag(al) == 0.0);
    const int have_beta = !(creal(be) == 0.0 && cimag(be) == 0.0);
    for (int i = 0; i < 60; ++i) o[36 + i] = 0.0;
    for (int i = 0; i < 100; ++i) o[96 + i] = 0.0;
    for (int k = 0; k < 4; ++k) {
        const double w = 0.25 * vol;
        double wv[6][3], gr[10][3];
        for (int i = 0; i < 6; ++i) whitney(g, i, qp[k], wv[i]);
        for (int a = 0; a < 10; ++a) p2grad(g, a, qp[k], gr[a]);
        if (have_alpha)
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) {
                    const cplx t = cxd(al, w * dot3(wv[i], wv[j]));
                    o[i * 6 + j] = CMPLX(creal(o[i * 6 + j]) + creal(t), cimag(o[i * 6 + j]) + cimag(t));
                }
        if (have_beta) {
            for (int i = 0; i < 6; ++i)
                for (int b = 0; b < 10; ++b) {
                    const cplx t = cxd(be, w * dot3(wv[i], gr[b]));
                    cplx* d = &o[36 + i * 10 + b];
                    *d = CMPLX(creal(*d) + creal(t), cimag(*d) + cimag(t));
                }
            for (int a = 0; a < 10; ++a)
                for (int b = 0; b < 10; ++b) {
                    const cplx t = cxd(be, w * dot3(gr[a], gr[b]));
                    cplx* d = &o[96 + a * 10 + b];
                    *d = CMPLX(creal(*d) + creal(t), cimag(*d) + cimag(t));
                }
        }
    }
    return 1;
}

/* ------------------------------------------------------------------ refill */

static int find_slot(const int32_t* row_ptr, const int32_t* col, int r, int c) {
    int lo = row_ptr[r], hi = row_ptr[r + 1];
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (col[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return (lo < row_ptr[r + 1] && col[lo] == c) ? lo : -1;
}

static int tet_block_index(int n, int r, int c) { return c < r ? -1 : r * n - r * (r - 1) / 2 + (c - r); }

/* refill for the upper-triangle SymmetricSystem: scatter_into,
 * src/assembly.cpp:255-422, driven by the ScatterMap (src/assembly.cpp:196-232).
 *   body_tets/body_ptr: bodies in order, each body's tets ascending
 *   body_param: per body sigma, eps_r, mu_r (element_coefficients,
 *               src/element_matrix.cpp:10-25, is formed here per omega)
 *   tet_dofs: per tet 6 edge then 10 phi global indices (-1 eliminated)
 *   tet_coeff: per tet 6 edge signs (phi coefficients are 1)
 *   tet_fixed: per tet 10 prescribed phi values (complex) for eliminated nodes
 *   omega, conditioning (1 row_scaled, 2 scaled_phi): formulation_scales,
 *               src/assembly.cpp:12-45
 *   ports: per port row, is_current, value(re,im)
 * Returns -1 for a degenerate tet, -2 for a missing slot, -3 for an unsupported
 * conditioning or omega == 0, 0 on success. */
int oracle_refill(int n, int64_t nnz, const int32_t* row_ptr, const int32_t* col, const double* coords,
                  const int32_t* tet_nodes, int num_bodies, const int32_t* body_ptr, const int32_t* body_tets,
                  const double* body_param, const int32_t* tet_dofs, const double* tet_coeff, const double* tet_fixed,
                  const int32_t* live, const int32_t* pos, const int64_t* offset, const int32_t* slot,
                  double omega, int conditioning, int num_ports, const int32_t* port_row, const int32_t* port_current,
                  const double* port_value, double* values_out, double* rhs_out) {
    cplx* values = (cplx*)values_out;
    cplx* rhs = (cplx*)rhs_out;
    for (int64_t k = 0; k < nnz; ++k) values[k] = 0.0;
    for (int i = 0; i < n; ++i) rhs[i] = 0.0;
    const cplx jw = CMPLX(0.0, omega);
    cplx column = CMPLX(1.0, 0.0), row = CMPLX(1.0, 0.0);
    if (omega == 0.0 || (conditioning != 1 && conditioning != 2)) return -3;
    if (conditioning == 1) row = CMPLX(1.0, 0.0) / jw;
    else column = jw;
    const double mu0 = 4.0e-7 * 3.14159265358979323846, eps0 = 8.8541878128e-12;
    double el[392];
    for (int b = 0; b < num_bodies; ++b) {
        const double sigma = body_param[3 * b], mu = body_param[3 * b + 2] * mu0, eps = body_param[3 * b + 1] * eps0;
        const double cf[5] = {1.0 / mu, -omega * omega * eps, omega * sigma, sigma, omega * eps};
        for (int q = body_ptr[b]; q < body_ptr[b + 1]; ++q) {
            const int t = body_tets[q];
            double p12[12];
            for (int j = 0; j < 4; ++j)
                for (int c = 0; c < 3; ++c) p12[3 * j + c] = coords[3 * tet_nodes[4 * t + j] + c];
            if (!oracle_element(p12, cf[0], cf + 1, cf + 3, el)) return -1;
            const cplx* aa = (const cplx*)el;
            const cplx* ap = aa + 36;
            const cplx* pp = aa + 96;
            const int32_t* ed = tet_dofs + 16 * t;
            const int32_t* ph = ed + 6;
            const double* ec = tet_coeff + 6 * t;
            const cplx* fixed = (const cplx*)(tet_fixed + 20 * t);
            const int nl = live[t];
            const int32_t* ep = pos + 16 * t;
            const int32_t* pq = ep + 6;
            const int32_t* sl = slot + offset[t];
#define SLOT(r, c) (tet_block_index(nl, (r), (c)) < 0 ? -1 : sl[tet_block_index(nl, (r), (c))])
            for (int a = 0; a < 6; ++a) {
                if (ed[a] < 0) continue;
                for (int c = 0; c < 6; ++c) {
                    if (ed[c] < 0) continue;
                    const int s = SLOT(ep[a], ep[c]);
                    if (s < 0) continue;
                    values[s] += (ec[a] * ec[c]) * aa[a * 6 + c];
                }
            }
            for (int a = 0; a < 6; ++a) {
                if (ed[a] < 0) continue;
                for (int c = 0; c < 10; ++c) {
                    const cplx block = ap[a * 10 + c];
                    if (ph[c] >= 0) {
                        const int s = SLOT(ep[a], pq[c]);
                        if (s >= 0) values[s] += column * (ec[a] * 1.0) * block;
                    } else if (fixed[c] != 0.0) {
                        const cplx f = fixed[c] / column;
                        rhs[ed[a]] -= column * ec[a] * block * f;
                    }
                    if (ph[c] >= 0) {
                        const int s = SLOT(pq[c], ep[a]);
                        if (s >= 0) values[s] += row * jw * (1.0 * ec[a]) * block;
                    }
                }
            }
            for (int a = 0; a < 10; ++a) {
                if (ph[a] < 0) continue;
                for (int c = 0; c < 10; ++c) {
                    const cplx block = pp[a * 10 + c];
                    if (ph[c] >= 0) {
                        const int s = SLOT(pq[a], pq[c]);
                        if (s >= 0) values[s] += row * column * (1.0 * 1.0) * block;
                    } else if (fixed[c] != 0.0) {
                        const cplx f = fixed[c] / column;
                        rhs[ph[a]] -= row * column * 1.0 * block * f;
                    }
                }
            }
#undef SLOT
        }
    }
    for (int k = 0; k < num_ports; ++k) {
        const int r = port_row[k];
        const cplx v = CMPLX(port_value[2 * k], port_value[2 * k + 1]);
        if (port_current[k]) {
            rhs[r] += row * v;
        } else {
            const int s = find_slot(row_ptr, col, r, r);
            if (s < 0) return -2;
            values[s] += CMPLX(1.0, 0.0);
            rhs[r] += v / column;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ equilibration */

/* compute_symmetric_equilibration (upper triangle), src/equilibration.cpp:94-139 */
void oracle_equilibrate(int n, const int32_t* row_ptr, const int32_t* col, const double* values, int iterations,
                        double* d) {
    const int64_t nnz = row_ptr[n];
    cplx* w = malloc((size_t)nnz * sizeof(cplx));
    double* rm = malloc((size_t)n * sizeof(double));
    double* f = malloc((size_t)n * sizeof(double));
    memcpy(w, values, (size_t)nnz * sizeof(cplx));
    for (int i = 0; i < n; ++i) d[i] = 1.0;
    for (int it = 0; it < iterations; ++it) {
        for (int i = 0; i < n; ++i) rm[i] = 0.0;
        for (int i = 0; i < n; ++i)
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
                const double mag = cabs(w[k]);
                if (mag > rm[i]) rm[i] = mag;
                if (mag > rm[col[k]]) rm[col[k]] = mag;
            }
        for (int i = 0; i < n; ++i) {
            f[i] = rm[i] > 0.0 ? 1.0 / sqrt(rm[i]) : 1.0;
            d[i] *= f[i];
        }
        for (int i = 0; i < n; ++i)
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) w[k] = cxd(w[k], f[i] * f[col[k]]);
    }
    free(w);
    free(rm);
    free(f);
}

/* ------------------------------------------------------------------ LDL^T */

/* Up-looking complex-symmetric LDL^T of P A P^T (src/factorization.cpp:13-218):
 * permuted_lower, elimination tree and column counts (src/symbolic.cpp:58-196),
 * ereach, then per row k the column updates in topological order. No pivoting,
 * nothing conjugated. */
typedef struct {
    int n;
    int32_t *col_ptr, *row_index;
    cplx *value, *diag;
    int32_t* perm;
} oracle_factor;

static void lower_of_permuted(int n, const int32_t* row_ptr, const int32_t* col, const cplx* a, const int32_t* iperm,
                              int32_t** lp, int32_t** lc, cplx** lv) {
    int32_t* rp = calloc((size_t)n + 1, sizeof(int32_t));
    for (int r = 0; r < n; ++r)
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            int i = iperm[r], j = iperm[col[k]];
            if (j > i) { int t = i; i = j; j = t; }
            rp[i + 1]++;
        }
    for (int r = 0; r < n; ++r) rp[r + 1] += rp[r];
    int32_t* cur = malloc((size_t)n * sizeof(int32_t));
    memcpy(cur, rp, (size_t)n * sizeof(int32_t));
    int32_t* ci = malloc((size_t)rp[n] * sizeof(int32_t));
    cplx* v = malloc((size_t)rp[n] * sizeof(cplx));
    for (int r = 0; r < n; ++r)
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            int i = iperm[r], j = iperm[col[k]];
            if (j > i) { int t = i; i = j; j = t; }
            ci[cur[i]] = j;
            v[cur[i]++] = a[k];
        }
    /* rows sorted by column (insertion sort: rows are short) */
    for (int r = 0; r < n; ++r)
        for (int p = rp[r] + 1; p < rp[r + 1]; ++p) {
            const int32_t c = ci[p];
            const cplx x = v[p];
            int q = p - 1;
            while (q >= rp[r] && ci[q] > c) {
                ci[q + 1] = ci[q];
                v[q + 1] = v[q];
                --q;
            }
            ci[q + 1] = c;
            v[q + 1] = x;
        }
    free(cur);
    *lp = rp;
    *lc = ci;
    *lv = v;
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* stats: smallest_pivot, largest_pivot, largest_multiplier, columns_done, ok.
 * budget_s > 0 stops after that many seconds (rows done in stats[3]);
 * work[0] = real flops performed, work[1] = the flops of the whole factorization
 * (8 per complex multiply-add of the column updates, 16 per (row, column) pair). */
static oracle_factor* ldlt_budget(int n, const int32_t* row_ptr, const int32_t* col, const cplx* a, const int32_t* perm,
                                  double pivot_floor, double* stats, double budget_s, double* work);

static oracle_factor* ldlt(int n, const int32_t* row_ptr, const int32_t* col, const cplx* a, const int32_t* perm,
                           double pivot_floor, double* stats) {
    return ldlt_budget(n, row_ptr, col, a, perm, pivot_floor, stats, 0.0, NULL);
}

static oracle_factor* ldlt_budget(int n, const int32_t* row_ptr, const int32_t* col, const cplx* a, const int32_t* perm,
                                  double pivot_floor, double* stats, double budget_s, double* work) {
    int32_t* iperm = malloc((size_t)n * sizeof(int32_t));
    for (int i = 0; i < n; ++i) iperm[perm[i]] = i;
    int32_t *lp, *lc;
    cplx* lv;
    lower_of_permuted(n, row_ptr, col, a, iperm, &lp, &lc, &lv);
    /* elimination tree of the permuted pattern (row k's entries j < k are exactly
     * the earlier neighbours of k) */
    int32_t* parent = malloc((size_t)n * sizeof(int32_t));
    int32_t* anc = malloc((size_t)n * sizeof(int32_t));
    for (int i = 0; i < n; ++i) parent[i] = anc[i] = -1;
    for (int k = 0; k < n; ++k)
        for (int t = lp[k]; t < lp[k + 1]; ++t) {
            int i = lc[t];
            while (i != -1 && i < k) {
                const int nx = anc[i];
                anc[i] = k;
                if (nx == -1) parent[i] = k;
                i = nx;
            }
        }
    /* column counts via row reaches */
    int32_t* flag = malloc((size_t)n * sizeof(int32_t));
    int32_t* cc = calloc((size_t)n, sizeof(int32_t));
    for (int i = 0; i < n; ++i) flag[i] = -1;
    for (int k = 0; k < n; ++k) {
        flag[k] = k;
        for (int t = lp[k]; t < lp[k + 1]; ++t) {
            int i = lc[t];
            while (i >= 0 && i < k && flag[i] != k) {
                flag[i] = k;
                cc[i]++;
                i = parent[i];
            }
        }
    }
    oracle_factor* f = calloc(1, sizeof(oracle_factor));
    f->n = n;
    f->col_ptr = calloc((size_t)n + 1, sizeof(int32_t));
    for (int j = 0; j < n; ++j) f->col_ptr[j + 1] = f->col_ptr[j] + cc[j];
    f->row_index = malloc(((size_t)f->col_ptr[n] + 1) * sizeof(int32_t));
    f->value = malloc(((size_t)f->col_ptr[n] + 1) * sizeof(cplx));
    f->diag = calloc((size_t)n, sizeof(cplx));
    f->perm = malloc((size_t)n * sizeof(int32_t));
    memcpy(f->perm, perm, (size_t)n * sizeof(int32_t));
    cplx* y = calloc((size_t)n, sizeof(cplx));
    int32_t* stack = malloc((size_t)n * sizeof(int32_t));
    int32_t* filled = calloc((size_t)n, sizeof(int32_t));
    for (int i = 0; i < n; ++i) flag[i] = -1;
    double smallest = 0.0, largest = 0.0, lmax = 0.0;
    int first = 1, done = 0, ok = 1;
    double macs = 0.0, pairs = 0.0;
    if (work) {
        double total = 0.0;
        for (int j = 0; j < n; ++j) total += (double)cc[j] * (cc[j] - 1) / 2.0 + 2.0 * cc[j];
        work[1] = 8.0 * total;
    }
    const double t0 = budget_s > 0.0 ? now_s() : 0.0;
    for (int k = 0; k < n; ++k) {
        if (budget_s > 0.0 && (k & 63) == 0 && now_s() - t0 > budget_s) break;
        /* ereach, src/factorization.cpp:86-105 */
        int top = n;
        flag[k] = k;
        for (int t = lp[k]; t < lp[k + 1]; ++t) {
            int i = lc[t];
            if (i >= k) continue;
            int len = 0;
            while (flag[i] != k) {
                stack[len++] = i;
                flag[i] = k;
                i = parent[i];
                if (i < 0) break;
            }
            while (len > 0) stack[--top] = stack[--len];
        }
        cplx d = 0.0;
        for (int t = lp[k]; t < lp[k + 1]; ++t) {
            const int j = lc[t];
            if (j == k) d += lv[t];
            else if (j < k) y[j] += lv[t];
        }
        for (int s = top; s < n; ++s) {
            const int j = stack[s];
            const cplx yj = y[j];
            y[j] = 0.0;
            const int b = f->col_ptr[j], e = b + filled[j];
            for (int t = b; t < e; ++t) y[f->row_index[t]] -= f->value[t] * yj;
            macs += (double)filled[j];
            pairs += 1.0;
            const cplx lkj = yj / f->diag[j];
            d -= lkj * yj;
            f->row_index[e] = k;
            f->value[e] = lkj;
            if (cabs(lkj) > lmax) lmax = cabs(lkj);
            filled[j]++;
        }
        f->diag[k] = d;
        const double mag = cabs(d);
        if (mag > largest) largest = mag;
        if (first || mag < smallest) {
            smallest = mag;
            first = 0;
        }
        done = k + 1;
        if (!(mag > pivot_floor)) {
            ok = 0;
            break;
        }
    }
    stats[0] = smallest;
    stats[1] = largest;
    stats[2] = lmax;
    stats[3] = done;
    stats[4] = ok;
    if (work) work[0] = 8.0 * macs + 16.0 * pairs;
    free(iperm); free(lp); free(lc); free(lv); free(parent); free(anc); free(flag); free(cc);
    free(y); free(stack); free(filled);
    return f;
}

static void factor_free(oracle_factor* f) {
    if (!f) return;
    free(f->col_ptr); free(f->row_index); free(f->value); free(f->diag); free(f->perm);
    free(f);
}

/* solve_in_place + solve, src/factorization.cpp:254-313 */
static void factor_solve(const oracle_factor* f, const cplx* b, cplx* x) {
    const int n = f->n;
    cplx* y = malloc((size_t)n * sizeof(cplx));
    for (int i = 0; i < n; ++i) y[i] = b[f->perm[i]];
    for (int j = 0; j < n; ++j) {
        const cplx yj = y[j];
        for (int t = f->col_ptr[j]; t < f->col_ptr[j + 1]; ++t) y[f->row_index[t]] -= f->value[t] * yj;
    }
    for (int j = 0; j < n; ++j) y[j] /= f->diag[j];
    for (int j = n - 1; j >= 0; --j) {
        cplx acc = y[j];
        for (int t = f->col_ptr[j]; t < f->col_ptr[j + 1]; ++t) acc -= f->value[t] * y[f->row_index[t]];
        y[j] = acc;
    }
    for (int i = 0; i < n; ++i) x[f->perm[i]] = y[i];
    free(y);
}

/* SparseSymmetric::matvec, include/aphi_solver/sparse_symmetric.hpp:116-131 */
static void symv(int n, const int32_t* row_ptr, const int32_t* col, const cplx* a, const cplx* x, cplx* y) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int r = 0; r < n; ++r)
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            const int c = col[k];
            y[r] += a[k] * x[c];
            if (c != r) y[c] += a[k] * x[r];
        }
}

/* backward_error and relative_residual, src/factorization.cpp:315-329, 401-416.
 * out[0] = backward error, out[1] = ||Ax-b||/||b|| */
void oracle_backward_error(int n, const int32_t* row_ptr, const int32_t* col, const double* values, const double* xv,
                           const double* bv, double* out) {
    const cplx* a = (const cplx*)values;
    const cplx* x = (const cplx*)xv;
    const cplx* b = (const cplx*)bv;
    cplx* ax = malloc((size_t)n * sizeof(cplx));
    symv(n, row_ptr, col, a, x, ax);
    double num = 0.0, bn = 0.0, xn = 0.0, amax = 0.0;
    for (int i = 0; i < n; ++i) {
        const cplx r = ax[i] - b[i];
        num += creal(r) * creal(r) + cimag(r) * cimag(r);
        bn += creal(b[i]) * creal(b[i]) + cimag(b[i]) * cimag(b[i]);
        xn += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
    }
    for (int64_t k = 0; k < row_ptr[n]; ++k)
        if (cabs(a[k]) > amax) amax = cabs(a[k]);
    const double den = amax * sqrt(xn) + sqrt(bn);
    out[0] = den > 0.0 ? sqrt(num) / den : sqrt(num);
    out[1] = bn > 0.0 ? sqrt(num) / sqrt(bn) : sqrt(num);
    free(ax);
}

/* The residual r = b - A x itself, with A x accumulated as backward_error's
 * matvec does (include/aphi_solver/sparse_symmetric.hpp:116-131). */
void oracle_residual(int n, const int32_t* row_ptr, const int32_t* col, const double* values, const double* xv,
                     const double* bv, double* rv) {
    const cplx* a = (const cplx*)values;
    const cplx* b = (const cplx*)bv;
    cplx* r = (cplx*)rv;
    symv(n, row_ptr, col, a, (const cplx*)xv, r);
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
}

/* solve_symmetric, src/factorization.cpp:332-398, for nrhs right-hand sides
 * (column-major, ld = n): equilibrate (iterations, 0 disables), factor P A~ P^T
 * with the given permutation, solve, x = D y. stats[0..4] as ldlt();
 * stats[5] = max backward error, stats[6] = equilibration min, stats[7] = max.
 * diag_out (n complex, may be NULL) receives D. Returns 1 on success, 0 when a
 * pivot stopped the factorization. */
int oracle_solve_symmetric(int n, const int32_t* row_ptr, const int32_t* col, const double* values, int nrhs,
                           const double* bv, const int32_t* perm, int iterations, double pivot_floor, double* xv,
                           double* diag_out, double* stats) {
    const int64_t nnz = row_ptr[n];
    const cplx* a = (const cplx*)values;
    double* d = malloc((size_t)n * sizeof(double));
    if (iterations > 0) oracle_equilibrate(n, row_ptr, col, values, iterations, d);
    else for (int i = 0; i < n; ++i) d[i] = 1.0;
    cplx* as = malloc((size_t)nnz * sizeof(cplx));
    for (int r = 0; r < n; ++r)
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) as[k] = a[k] * CMPLX(d[r] * d[col[k]], 0.0);
    oracle_factor* f = ldlt(n, row_ptr, col, as, perm, pivot_floor, stats);
    double dmin = d[0], dmax = d[0];
    for (int i = 0; i < n; ++i) {
        if (d[i] < dmin) dmin = d[i];
        if (d[i] > dmax) dmax = d[i];
    }
    stats[6] = dmin;
    stats[7] = dmax;
    stats[5] = 0.0;
    const int ok = stats[4] != 0.0;
    if (ok) {
        cplx* bs = malloc((size_t)n * sizeof(cplx));
        cplx* y = malloc((size_t)n * sizeof(cplx));
        for (int r = 0; r < nrhs; ++r) {
            const cplx* b = (const cplx*)bv + (size_t)r * n;
            cplx* x = (cplx*)xv + (size_t)r * n;
            for (int i = 0; i < n; ++i) bs[i] = b[i] * CMPLX(d[i], 0.0);
            factor_solve(f, bs, y);
            for (int i = 0; i < n; ++i) x[i] = y[i] * CMPLX(d[i], 0.0);
            double be[2];
            oracle_backward_error(n, row_ptr, col, values, (const double*)x, (const double*)b, be);
            if (be[0] > stats[5]) stats[5] = be[0];
        }
        free(bs);
        free(y);
    }
    if (diag_out) memcpy(diag_out, f->diag, (size_t)n * sizeof(cplx));
    factor_free(f);
    free(as);
    free(d);
    return ok;
}

/* L D L^T reconstruction error max|(LDL^T - P A P^T)_ij| / max|A| over the
 * stored pattern of A and the factor's own pattern, for tests on small systems. */
double oracle_reconstruction_error(int n, const int32_t* row_ptr, const int32_t* col, const double* values,
                                   const int32_t* perm) {
    double st[5];
    const cplx* a = (const cplx*)values;
    oracle_factor* f = ldlt(n, row_ptr, col, a, perm, 0.0, st);
    if (n > 4000 || st[4] == 0.0) {
        factor_free(f);
        return -1.0;
    }
    cplx* dense = calloc((size_t)n * n, sizeof(cplx));
    int32_t* iperm = malloc((size_t)n * sizeof(int32_t));
    for (int i = 0; i < n; ++i) iperm[perm[i]] = i;
    double amax = 0.0;
    for (int r = 0; r < n; ++r)
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            const int i = iperm[r], j = iperm[col[k]];
            dense[(size_t)i * n + j] = a[k];
            dense[(size_t)j * n + i] = a[k];
            if (cabs(a[k]) > amax) amax = cabs(a[k]);
        }
    /* subtract sum_j D_j l_j l_j^T column by column */
    for (int j = 0; j < n; ++j) {
        const int b = f->col_ptr[j], e = f->col_ptr[j + 1];
        const cplx dj = f->diag[j];
        dense[(size_t)j * n + j] -= dj;
        for (int t = b; t < e; ++t) {
            const int i = f->row_index[t];
            const cplx v = f->value[t] * dj;
            dense[(size_t)i * n + j] -= v;
            dense[(size_t)j * n + i] -= v;
            for (int u = b; u < e; ++u) dense[(size_t)i * n + f->row_index[u]] -= v * f->value[u];
        }
    }
    double err = 0.0;
    for (size_t q = 0; q < (size_t)n * n; ++q)
        if (cabs(dense[q]) > err) err = cabs(dense[q]);
    free(dense);
    free(iperm);
    factor_free(f);
    return err / amax;
}

/* CPU baseline sample: equilibrate (10 Ruiz passes) and run the reference's
 * up-looking factorization of P A~ P^T for at most budget_s seconds.
 * out[0] = seconds spent factoring, out[1] = rows completed, out[2] = flops
 * performed, out[3] = flops of the whole factorization, out[4] = equilibration
 * seconds. */
void oracle_factor_sample(int n, const int32_t* row_ptr, const int32_t* col, const double* values, const int32_t* perm,
                          double budget_s, double* out) {
    const int64_t nnz = row_ptr[n];
    const cplx* a = (const cplx*)values;
    double* d = malloc((size_t)n * sizeof(double));
    const double e0 = now_s();
    oracle_equilibrate(n, row_ptr, col, values, 10, d);
    cplx* as = malloc((size_t)nnz * sizeof(cplx));
    for (int r = 0; r < n; ++r)
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) as[k] = a[k] * CMPLX(d[r] * d[col[k]], 0.0);
    const double e1 = now_s();
    double st[5], work[2] = {0.0, 0.0};
    const double t0 = now_s();
    oracle_factor* f = ldlt_budget(n, row_ptr, col, as, perm, 0.0, st, budget_s, work);
    const double t1 = now_s();
    out[0] = t1 - t0;
    out[1] = st[3];
    out[2] = work[0];
    out[3] = work[1];
    out[4] = e1 - e0;
    factor_free(f);
    free(as);
    free(d);
}

/* ------------------------------------------------------------------ port current */

/* compute_fields' centroid current density (src/postprocess.cpp:311-400: A from
 * the edge unknowns with the global edge signs, Phi from the Phi unknowns scaled
 * by Solution::phi_scale (:35-40) or the prescribed values, E = -jw A - grad Phi,
 * J = sigma E) and terminal_current (:526-547): I = -(1/length) sum over
 * conducting tets, in tet order, of (J . axis) |V|; impedance Z = V / I
 * (include/aphi_solver/postprocess.hpp:356-358), R = Re Z, L = Im Z / omega.
 * Complex products are written out as the reference's std::complex evaluates
 * them for finite values. out: I_re, I_im, R, L; terms (optional): the flux term
 * of each tet (2 doubles, zero for non-conducting tets). Returns 0 for a bad
 * length, like the reference's refusal. */
/* E = -j w A - grad Phi at the centroid of tet t: compute_fields' evaluate at
 * kCentroid (src/postprocess.cpp:311-391), in its operation order. Returns 0
 * for a degenerate tet; g and vol receive the tet's geometry. */
static int tet_centroid_e(int t, const double* coords, const int32_t* tet_nodes, const int32_t* tet_dofs,
                          const double* tet_signs, const double* tet_phi_coeff, const double* tet_phi_fixed,
                          const double* x, double omega, int conditioning, double e_re[3], double e_im[3],
                          double g[4][3], double* vol) {
    const double mjw_re = -0.0, mjw_im = -omega;                    /* -jw */
    const double sc_re = conditioning == 2 ? 0.0 : 1.0, sc_im = conditioning == 2 ? omega : 0.0;
    const double L[4] = {0.25, 0.25, 0.25, 0.25};
    double p[4][3];
    for (int j = 0; j < 4; ++j)
        for (int c = 0; c < 3; ++c) p[j][c] = coords[3 * tet_nodes[4 * t + j] + c];
    if (!tet_geometry(p, g, vol)) return 0;
    const int32_t* d = tet_dofs + 16 * (int64_t)t;
    double av_re[3] = {0.0, 0.0, 0.0}, av_im[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 6; ++k) {
        double ae_re = 0.0, ae_im = 0.0;
        if (d[k] >= 0) { ae_re = x[2 * d[k]]; ae_im = x[2 * d[k] + 1]; }
        double w[3];
        whitney(g, k, L, w);
        const double sg = tet_signs[6 * (int64_t)t + k];
        for (int c = 0; c < 3; ++c) {
            const double wc = w[c] * sg;
            av_re[c] = av_re[c] + ae_re * wc;
            av_im[c] = av_im[c] + ae_im * wc;
        }
    }
    double gp_re[3] = {0.0, 0.0, 0.0}, gp_im[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < 10; ++b) {
        const int ib = d[6 + b];
        double pn_re, pn_im;
        if (ib >= 0) {
            const double cf = tet_phi_coeff[10 * (int64_t)t + b];
            const double s_re = sc_re * cf, s_im = sc_im * cf;          /* scale * coeff */
            const double xr = x[2 * ib], xi = x[2 * ib + 1];
            pn_re = s_re * xr - s_im * xi;                              /* (scale * coeff) * x */
            pn_im = s_re * xi + s_im * xr;
        } else {
            pn_re = tet_phi_fixed[20 * (int64_t)t + 2 * b];
            pn_im = tet_phi_fixed[20 * (int64_t)t + 2 * b + 1];
        }
        double gr[3];
        p2grad(g, b, L, gr);
        for (int c = 0; c < 3; ++c) {
            gp_re[c] = gp_re[c] + pn_re * gr[c];
            gp_im[c] = gp_im[c] + pn_im * gr[c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        e_re[c] = (mjw_re * av_re[c] - mjw_im * av_im[c]) - gp_re[c];
        e_im[c] = (mjw_re * av_im[c] + mjw_im * av_re[c]) - gp_im[c];
    }
    return 1;
}

/* compute_fields' e_tet (src/postprocess.cpp:311-391) for every tet: 3 complex per tet. */
int oracle_tet_fields(int num_tets, const double* coords, const int32_t* tet_nodes, const int32_t* tet_dofs,
                      const double* tet_signs, const double* tet_phi_coeff, const double* tet_phi_fixed,
                      const double* x, double omega, int conditioning, double* e_out) {
    for (int t = 0; t < num_tets; ++t) {
        double e_re[3], e_im[3], g[4][3], vol;
        if (!tet_centroid_e(t, coords, tet_nodes, tet_dofs, tet_signs, tet_phi_coeff, tet_phi_fixed, x, omega,
                            conditioning, e_re, e_im, g, &vol))
            return 0;
        for (int c = 0; c < 3; ++c) {
            e_out[6 * (int64_t)t + 2 * c] = e_re[c];
            e_out[6 * (int64_t)t + 2 * c + 1] = e_im[c];
        }
    }
    return 1;
}

/* A and E = -jw A - grad Phi at the barycentric point L of tet t (compute_fields'
 * evaluate, src/postprocess.cpp:364-386), in its operation order; a and ph are
 * the tet's raw edge unknowns and resolved Phi coefficients. */
static void tet_eval(double g[4][3], const double L[4], const double* sg, const double a_re[6],
                     const double a_im[6], const double ph_re[10], const double ph_im[10], double omega,
                     double av_re[3], double av_im[3], double e_re[3], double e_im[3]) {
    const double mjw_re = -0.0, mjw_im = -omega;
    for (int c = 0; c < 3; ++c) { av_re[c] = 0.0; av_im[c] = 0.0; }
    for (int k = 0; k < 6; ++k) {
        double w[3];
        whitney(g, k, L, w);
        for (int c = 0; c < 3; ++c) {
            const double wc = w[c] * sg[k];
            av_re[c] = av_re[c] + a_re[k] * wc;
            av_im[c] = av_im[c] + a_im[k] * wc;
        }
    }
    double gp_re[3] = {0.0, 0.0, 0.0}, gp_im[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < 10; ++b) {
        double gr[3];
        p2grad(g, b, L, gr);
        for (int c = 0; c < 3; ++c) {
            gp_re[c] = gp_re[c] + ph_re[b] * gr[c];
            gp_im[c] = gp_im[c] + ph_im[b] * gr[c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        e_re[c] = (mjw_re * av_re[c] - mjw_im * av_im[c]) - gp_re[c];
        e_im[c] = (mjw_re * av_im[c] + mjw_im * av_re[c]) - gp_im[c];
    }
}

/* compute_fields (src/postprocess.cpp:223-481) without phi_node: per tet A, E, J at
 * the centroid, B = curl A and sigma; per P2 node the volume-weighted averages of A
 * (every incident tet), E (the most conducting side) and B (the most permeable
 * side), then J = sigma_max E and H = B / mu_max, with the interface flags and
 * counts. tet_p2 gives each tet's ten P2 nodes (vertices, then local edges in
 * kTetLocalEdgeVerts order). Complex outputs interleaved (re, im); counts[0] =
 * interface nodes, counts[1] = orphan nodes. Returns 0 for a degenerate tet. */
int oracle_compute_fields(int num_tets, int num_p2, const double* coords, const int32_t* tet_nodes,
                          const int32_t* tet_p2, const int32_t* tet_body, const double* body_sigma,
                          const double* body_mu_r, const int32_t* tet_dofs, const double* tet_signs,
                          const double* tet_phi_coeff, const double* tet_phi_fixed, const double* x, double omega,
                          int conditioning, double* a_tet, double* b_tet, double* e_tet, double* j_tet,
                          double* sigma_tet, double* a_node, double* b_node, double* h_node, double* e_node,
                          double* j_node, double* vertex_weight, int8_t* on_interface, int32_t* counts) {
    const double kMu0 = 4.0e-7 * 3.14159265358979323846;
    const size_t np = (size_t)num_p2;
    double* sig_min = malloc(np * sizeof(double));
    double* sig_max = malloc(np * sizeof(double));
    double* mu_min = malloc(np * sizeof(double));
    double* mu_max = malloc(np * sizeof(double));
    double* e_weight = calloc(np, sizeof(double));
    double* b_weight = calloc(np, sizeof(double));
    int ok = 1;
    for (size_t u = 0; u < np; ++u) {
        sig_min[u] = INFINITY; sig_max[u] = -INFINITY; mu_min[u] = INFINITY; mu_max[u] = -INFINITY;
        vertex_weight[u] = 0.0; on_interface[u] = 0;
    }
    memset(a_node, 0, 6 * np * sizeof(double));
    memset(b_node, 0, 6 * np * sizeof(double));
    memset(h_node, 0, 6 * np * sizeof(double));
    memset(e_node, 0, 6 * np * sizeof(double));
    memset(j_node, 0, 6 * np * sizeof(double));
    for (int t = 0; t < num_tets; ++t) {
        const double sgm = body_sigma[tet_body[t]], m = kMu0 * body_mu_r[tet_body[t]];
        for (int k = 0; k < 10; ++k) {
            const int u = tet_p2[10 * (int64_t)t + k];
            /* std::min(a, b) = b < a ? b : a; std::max(a, b) = a < b ? b : a */
            sig_min[u] = sgm < sig_min[u] ? sgm : sig_min[u];
            sig_max[u] = sig_max[u] < sgm ? sgm : sig_max[u];
            mu_min[u] = m < mu_min[u] ? m : mu_min[u];
            mu_max[u] = mu_max[u] < m ? m : mu_max[u];
        }
    }
    const double sc_re = conditioning == 2 ? 0.0 : 1.0, sc_im = conditioning == 2 ? omega : 0.0;
    for (int t = 0; t < num_tets && ok; ++t) {
        double p[4][3], g[4][3], vol;
        for (int j = 0; j < 4; ++j)
            for (int c = 0; c < 3; ++c) p[j][c] = coords[3 * tet_nodes[4 * t + j] + c];
        if (!tet_geometry(p, g, &vol)) { ok = 0; break; }
        const int32_t* d = tet_dofs + 16 * (int64_t)t;
        const double* sg = tet_signs + 6 * (int64_t)t;
        double a_re[6], a_im[6], ph_re[10], ph_im[10];
        for (int k = 0; k < 6; ++k) {
            a_re[k] = d[k] >= 0 ? x[2 * d[k]] : 0.0;
            a_im[k] = d[k] >= 0 ? x[2 * d[k] + 1] : 0.0;
        }
        for (int b = 0; b < 10; ++b) {
            const int ib = d[6 + b];
            if (ib >= 0) {
                const double cf = tet_phi_coeff[10 * (int64_t)t + b];
                const double s_re = sc_re * cf, s_im = sc_im * cf;
                const double xr = x[2 * ib], xi = x[2 * ib + 1];
                ph_re[b] = s_re * xr - s_im * xi;
                ph_im[b] = s_re * xi + s_im * xr;
            } else {
                ph_re[b] = tet_phi_fixed[20 * (int64_t)t + 2 * b];
                ph_im[b] = tet_phi_fixed[20 * (int64_t)t + 2 * b + 1];
            }
        }
        /* B = curl A: whitney_edge_curl_global = (grad L_vi x grad L_vj) * 2 * sign */
        double bc_re[3] = {0.0, 0.0, 0.0}, bc_im[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 6; ++k) {
            const double* gi = g[kEV[k][0]];
            const double* gj = g[kEV[k][1]];
            const double cr[3] = {(gi[1] * gj[2] - gi[2] * gj[1]) * 2.0, (gi[2] * gj[0] - gi[0] * gj[2]) * 2.0,
                                  (gi[0] * gj[1] - gi[1] * gj[0]) * 2.0};
            for (int c = 0; c < 3; ++c) {
                const double cc = cr[c] * sg[k];
                bc_re[c] = bc_re[c] + a_re[k] * cc;
                bc_im[c] = bc_im[c] + a_im[k] * cc;
            }
        }
        const int body = tet_body[t];
        const double mu = kMu0 * body_mu_r[body];
        const double sigma = body_sigma[body];
        sigma_tet[t] = sigma;
        const double Lc[4] = {0.25, 0.25, 0.25, 0.25};
        double av_re[3], av_im[3], ev_re[3], ev_im[3];
        tet_eval(g, Lc, sg, a_re, a_im, ph_re, ph_im, omega, av_re, av_im, ev_re, ev_im);
        for (int c = 0; c < 3; ++c) {
            b_tet[6 * (int64_t)t + 2 * c] = bc_re[c];
            b_tet[6 * (int64_t)t + 2 * c + 1] = bc_im[c];
            a_tet[6 * (int64_t)t + 2 * c] = av_re[c];
            a_tet[6 * (int64_t)t + 2 * c + 1] = av_im[c];
            j_tet[6 * (int64_t)t + 2 * c] = sigma * ev_re[c];
            j_tet[6 * (int64_t)t + 2 * c + 1] = sigma * ev_im[c];
            e_tet[6 * (int64_t)t + 2 * c] = ev_re[c];
            e_tet[6 * (int64_t)t + 2 * c + 1] = ev_im[c];
        }
        for (int k = 0; k < 10; ++k) {
            double L[4] = {0.0, 0.0, 0.0, 0.0};
            if (k < 4) {
                L[k] = 1.0;
            } else {
                L[kEV[k - 4][0]] = 0.5;
                L[kEV[k - 4][1]] = 0.5;
            }
            const int u = tet_p2[10 * (int64_t)t + k];
            tet_eval(g, L, sg, a_re, a_im, ph_re, ph_im, omega, av_re, av_im, ev_re, ev_im);
            vertex_weight[u] = vertex_weight[u] + vol;
            for (int c = 0; c < 3; ++c) {
                a_node[6 * (size_t)u + 2 * c] = a_node[6 * (size_t)u + 2 * c] + vol * av_re[c];
                a_node[6 * (size_t)u + 2 * c + 1] = a_node[6 * (size_t)u + 2 * c + 1] + vol * av_im[c];
            }
            const int sigma_uniform = sig_min[u] == sig_max[u];
            const int mu_uniform = mu_min[u] == mu_max[u];
            if (sigma_uniform || sigma == sig_max[u]) {
                e_weight[u] = e_weight[u] + vol;
                for (int c = 0; c < 3; ++c) {
                    e_node[6 * (size_t)u + 2 * c] = e_node[6 * (size_t)u + 2 * c] + vol * ev_re[c];
                    e_node[6 * (size_t)u + 2 * c + 1] = e_node[6 * (size_t)u + 2 * c + 1] + vol * ev_im[c];
                }
            }
            if (mu_uniform || mu == mu_max[u]) {
                b_weight[u] = b_weight[u] + vol;
                for (int c = 0; c < 3; ++c) {
                    b_node[6 * (size_t)u + 2 * c] = b_node[6 * (size_t)u + 2 * c] + vol * bc_re[c];
                    b_node[6 * (size_t)u + 2 * c + 1] = b_node[6 * (size_t)u + 2 * c + 1] + vol * bc_im[c];
                }
            }
            if (!sigma_uniform || !mu_uniform) on_interface[u] = 1;
        }
    }
    counts[0] = 0;
    counts[1] = 0;
    for (size_t u = 0; u < np && ok; ++u) {
        if (on_interface[u]) ++counts[0];
        if (vertex_weight[u] > 0.0) {
            for (int c = 0; c < 6; ++c) a_node[6 * u + c] = a_node[6 * u + c] / vertex_weight[u];
        } else {
            ++counts[1];
            continue;
        }
        if (e_weight[u] > 0.0)
            for (int c = 0; c < 6; ++c) e_node[6 * u + c] = e_node[6 * u + c] / e_weight[u];
        if (b_weight[u] > 0.0)
            for (int c = 0; c < 6; ++c) b_node[6 * u + c] = b_node[6 * u + c] / b_weight[u];
        const double nu_node = 1.0 / mu_max[u];
        for (int c = 0; c < 6; ++c) {
            j_node[6 * u + c] = sig_max[u] * e_node[6 * u + c];
            h_node[6 * u + c] = nu_node * b_node[6 * u + c];
        }
    }
    free(sig_min); free(sig_max); free(mu_min); free(mu_max); free(e_weight); free(b_weight);
    return ok;
}

int oracle_terminal_current(int num_tets, const double* coords, const int32_t* tet_nodes, const int32_t* tet_body,
                            const double* body_sigma, const int32_t* tet_dofs, const double* tet_signs,
                            const double* tet_phi_coeff, const double* tet_phi_fixed, const double* x, double omega,
                            int conditioning, const double* axis, double length, double voltage, double* out,
                            double* terms) {
    if (!(length > 0.0)) return 0;
    double f_re = 0.0, f_im = 0.0;
    for (int t = 0; t < num_tets; ++t) {
        const double sigma = body_sigma[tet_body[t]];
        if (terms) { terms[2 * t] = 0.0; terms[2 * t + 1] = 0.0; }
        if (!(sigma > 0.0)) continue;
        double e_re[3], e_im[3], g[4][3], vol;
        if (!tet_centroid_e(t, coords, tet_nodes, tet_dofs, tet_signs, tet_phi_coeff, tet_phi_fixed, x, omega,
                            conditioning, e_re, e_im, g, &vol))
            return 0;
        double j_re[3], j_im[3];
        for (int c = 0; c < 3; ++c) {
            j_re[c] = sigma * e_re[c];
            j_im[c] = sigma * e_im[c];
        }
        const double t_re = ((j_re[0] * axis[0] + j_re[1] * axis[1]) + j_re[2] * axis[2]) * vol;
        const double t_im = ((j_im[0] * axis[0] + j_im[1] * axis[1]) + j_im[2] * axis[2]) * vol;
        if (terms) { terms[2 * t] = t_re; terms[2 * t + 1] = t_im; }
        f_re = f_re + t_re;
        f_im = f_im + t_im;
    }
    const cplx cur = CMPLX(-f_re / length, -f_im / length);
    const cplx Z = CMPLX(voltage, 0.0) / cur;
    out[0] = creal(cur);
    out[1] = cimag(cur);
    out[2] = creal(Z);
    out[3] = cimag(Z) / omega;
    return 1;
}

/* ------------------------------------------------------------------ gauge-free potential */

/* The vertex map and P1 system of recover_gauge_free_potential
 * (src/gauge_free_potential.cpp:47-164): build_vertex_map, then per tet of a body
 * with beta = sigma + j w eps0 eps_r != 0, k_ij = (beta |V|) (grad L_i . grad L_j)
 * into the upper triangle and -(beta |V|) (E . grad L_i) into the rhs, eliminated
 * neighbours' -k_ij value too. Rows are kept sorted by column as the reference's
 * std::map keeps them, each entry starting from zero. phi_state per vertex:
 * 0 Free, 1 Absent, 2 Fixed, 3 Port, 4 Cut. Returns the number of stored
 * entries, or -1 when a row needs more than 64 entries or cap_nnz is too small. */
#define GF_ROW_CAP 64
int64_t oracle_gauge_free_system(int nv, int nt, const double* coords, const int32_t* tet_nodes, const int32_t* tet_body,
                                 const double* body_sigma, const double* body_eps_r, const int32_t* phi_state,
                                 const int32_t* phi_port, int num_ports, const int32_t* port_is_fixed,
                                 const double* port_value, const double* e_tet, double omega, int64_t cap_nnz,
                                 int32_t* unknowns_out, int32_t* column, double* prescribed, int32_t* port_column,
                                 int32_t* row_ptr, int32_t* col, double* values, double* rhs) {
    const double kEps0 = 8.8541878128e-12;
    int unknowns = 0;
    for (int v = 0; v < nv; ++v) { column[v] = -1; prescribed[2 * v] = 0.0; prescribed[2 * v + 1] = 0.0; }
    for (int p = 0; p < num_ports; ++p) {
        port_column[p] = -1;
        if (port_is_fixed[p]) continue;
        port_column[p] = unknowns++;
    }
    for (int v = 0; v < nv; ++v) {
        if (phi_state[v] == 3) {
            const int p = phi_port[v];
            if (p < 0) continue;
            if (port_is_fixed[p]) { prescribed[2 * v] = port_value[2 * p]; prescribed[2 * v + 1] = port_value[2 * p + 1]; }
            else column[v] = port_column[p];
        } else if (phi_state[v] == 0) {
            column[v] = unknowns++;
        }
    }
    *unknowns_out = unknowns;
    int* rn = (int*)calloc((size_t)(unknowns > 0 ? unknowns : 1), sizeof(int));
    int* rc = (int*)malloc(sizeof(int) * (size_t)(unknowns > 0 ? unknowns : 1) * GF_ROW_CAP);
    double* rv = (double*)malloc(sizeof(double) * 2 * (size_t)(unknowns > 0 ? unknowns : 1) * GF_ROW_CAP);
    for (int r = 0; r < unknowns; ++r) { rhs[2 * r] = 0.0; rhs[2 * r + 1] = 0.0; }
    int64_t status = 0;
    for (int t = 0; t < nt && status == 0; ++t) {
        const int body = tet_body[t];
        if (body < 0) continue;
        const double b_re = body_sigma[body], b_im = omega * kEps0 * body_eps_r[body];
        if (b_re == 0.0 && b_im == 0.0) continue;
        double p[4][3], g[4][3], vol;
        for (int j = 0; j < 4; ++j)
            for (int c = 0; c < 3; ++c) p[j][c] = coords[3 * tet_nodes[4 * t + j] + c];
        tet_geometry(p, g, &vol);
        const double bv_re = b_re * vol, bv_im = b_im * vol;             /* beta * volume */
        const double* E = e_tet + 6 * (int64_t)t;
        for (int i = 0; i < 4 && status == 0; ++i) {
            const int ri = column[tet_nodes[4 * t + i]];
            if (ri < 0) continue;
            /* E . grad L_i, complex times real, left to right */
            const double ed_re = (E[0] * g[i][0] + E[2] * g[i][1]) + E[4] * g[i][2];
            const double ed_im = (E[1] * g[i][0] + E[3] * g[i][1]) + E[5] * g[i][2];
            rhs[2 * ri] = rhs[2 * ri] - (bv_re * ed_re - bv_im * ed_im);
            rhs[2 * ri + 1] = rhs[2 * ri + 1] - (bv_re * ed_im + bv_im * ed_re);
            for (int j = 0; j < 4; ++j) {
                const double dg = (g[i][0] * g[j][0] + g[i][1] * g[j][1]) + g[i][2] * g[j][2];
                const double k_re = bv_re * dg, k_im = bv_im * dg;
                const int vj = tet_nodes[4 * t + j];
                const int cj = column[vj];
                if (cj < 0) {
                    const double v_re = prescribed[2 * vj], v_im = prescribed[2 * vj + 1];
                    if (v_re != 0.0 || v_im != 0.0) {
                        rhs[2 * ri] = rhs[2 * ri] - (k_re * v_re - k_im * v_im);
                        rhs[2 * ri + 1] = rhs[2 * ri + 1] - (k_re * v_im + k_im * v_re);
                    }
                    continue;
                }
                if (cj < ri) continue;
                int* cols = rc + (int64_t)ri * GF_ROW_CAP;
                double* vals = rv + 2 * (int64_t)ri * GF_ROW_CAP;
                int q = 0;
                while (q < rn[ri] && cols[q] < cj) ++q;
                if (q == rn[ri] || cols[q] != cj) {
                    if (rn[ri] == GF_ROW_CAP) { status = -1; break; }
                    for (int z = rn[ri]; z > q; --z) {
                        cols[z] = cols[z - 1];
                        vals[2 * z] = vals[2 * z - 2];
                        vals[2 * z + 1] = vals[2 * z - 1];
                    }
                    cols[q] = cj;
                    vals[2 * q] = 0.0;
                    vals[2 * q + 1] = 0.0;
                    ++rn[ri];
                }
                vals[2 * q] = vals[2 * q] + k_re;
                vals[2 * q + 1] = vals[2 * q + 1] + k_im;
            }
        }
    }
    int64_t nnz = 0;
    if (status == 0) {
        row_ptr[0] = 0;
        for (int r = 0; r < unknowns; ++r) {
            if (nnz + rn[r] > cap_nnz) { status = -1; break; }
            for (int q = 0; q < rn[r]; ++q) {
                col[nnz] = rc[(int64_t)r * GF_ROW_CAP + q];
                values[2 * nnz] = rv[2 * ((int64_t)r * GF_ROW_CAP + q)];
                values[2 * nnz + 1] = rv[2 * ((int64_t)r * GF_ROW_CAP + q) + 1];
                ++nnz;
            }
            row_ptr[r + 1] = (int32_t)nnz;
        }
    }
    free(rn);
    free(rc);
    free(rv);
    return status == 0 ? nnz : -1;
}

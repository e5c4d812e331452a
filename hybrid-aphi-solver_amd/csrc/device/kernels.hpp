// Kernel declarations and the small POD structs shared between the HIP kernels
// and their host-side launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace aphi_dev {

// Per body, per frequency: nu, alpha, beta of src/element_matrix.cpp:10-25.
struct BodyCoef {
    double nu;
    cd alpha;
    cd beta;
};

// Per frequency scalars of the scatter (src/assembly.cpp:12-45), formed on the
// host with std::complex so they are the reference's values exactly.
struct FreqScales {
    cd column;        // c
    cd row_jw;        // r * j*omega
    cd row_column;    // r * c
    cd over_column[8];// prescribed value id -> value / c
    cd times_row[8];  // port amplitude id -> r * I
};

// Everything the numeric factorization reads about the supernodal structure.
struct FrontPlan {
    const int* first;          // ns + 1
    const int64_t* row_ptr;    // ns + 1
    const int* rows;           // per supernode row lists
    const int* rel;            // relative indices into the parent (aligned with rows)
    const int* child_ptr;      // ns + 1
    const int* child;
    const int64_t* panel_off;  // ns + 1
    const int64_t* cb_off;     // ns
    const int* col_ref;        // factor-order column -> reference (analysis) column
    // Inverses of the unit-lower diagonal blocks, one per kNB columns of each
    // supernode (the last one nb x nb), column-major, kept for the solves.
    const int64_t* xinv_off;   // ns
    // For a child of a blocked front: bound[ea_bound_off[child] + q] is the first
    // index of the child's update rows whose parent position is >= 64 q
    // (q = 0 .. ceil(m_parent / 64)); -1 for other supernodes.
    const int64_t* ea_bound_off;  // ns
    const int* ea_bound;
    // Inverses of the kSB-column diagonal blocks (the last one nb x nb), for the solves.
    const int64_t* xsb_off;   // ns
    // Complex-float factorization: W = -L D of a front's columns, column c at
    // W + woff[s] + (c mod kWCols) * m (two outer blocks of columns, so the update of
    // one outer block can read its window while the next block's panels fill the
    // other); offsets restart at every level, -1 for fronts factored whole in LDS.
    // The buffer starts with a zero column at least as long as any front (the
    // operand of the steps past an update's last column).
    const int64_t* woff;  // ns
};

// Factorization statistics, spread over kStatSlots cache-line slots so that the
// atomics of many workgroups do not all meet on one address (same-address
// device-scope atomics serialise); the host reduces the slots afterwards.
constexpr int kStatSlots = 256;
struct alignas(128) StatSlot {
    double smallest_pivot;   // atomic min of |D|
    double largest_pivot;    // atomic max of |D|
    double largest_multiplier;
    int fail_col;            // smallest reference column whose pivot failed, or INT_MAX
    int pad;
};
struct FactorStatsDev {
    StatSlot slot[kStatSlots];
};

// assembly.hip
__global__ void k_element(int num_tets, const double* coords, const int* tets, const int* tet_body,
                          const BodyCoef* coef, cd* E, int* error_flag);
__global__ void k_assemble_reduce(int64_t num_targets, int64_t num_slots, const int64_t* tptr, const uint32_t* entry,
                                  const uint8_t* kind, const cd* E, const FreqScales* scales, int nseg,
                                  const int64_t* seg_slots, const int64_t* seg_rows, cd* values, cd* rhs);
__global__ void k_copy_cd(int64_t n, const cd* src, cd* dst);
__global__ void k_fill_d(int64_t n, double* p, double v);
__global__ void k_fill_flag(int* flag);
__global__ void k_eq_rowmax16(int n, const int* row_ptr, const int* tptr, const int* tslot, const cd* w, double* rowmax);
__global__ void k_eq_factor(int n, double* rowmax, double* factor, double* d);
__global__ void k_eq_scale(int64_t nnz, const int* row_of, const int* col, const double* factor, cd* w);
__global__ void k_eq_apply(int64_t nnz, const int* row_of, const int* col, const double* d, const cd* a, cd* out);
__global__ void k_interleave16(int n, int nrhs, const cd* x, int64_t ldx, cd* xi);
__global__ void k_symv16(int n, int nrhs, const int* row_ptr, const int* col, const int* tptr, const int* tslot,
                         const int* trow, const cd* a, const cd* xi, cd* y, int64_t ldy);
__global__ void k_residual(int n, int nrhs, const cd* b, int64_t ldb, cd* r, int64_t ldr);
__global__ void k_sumsq_partial(int n, const cd* v, int64_t ldv, double* partials);
__global__ void k_seg_sumsq(int nseg, const int64_t* off, const cd* v, int64_t ldv, const double* d, int power,
                            double* partials);
__global__ void k_seg_absmax(const int64_t* off, const cd* v, double* partials);
__global__ void k_scale_permute_in(int n, int nrhs, const int* perm, const double* d, const cd* b, int64_t ldb,
                                   cd* out, int64_t ldo);
__global__ void k_permute_scale_out(int n, int nrhs, const int* perm, const double* d, const cd* xt, int64_t ldt,
                                    cd* x, int64_t ldx, int accumulate);

// GMRES refinement (FP64)
__global__ void k_scale_d(int n, int nrhs, const double* d, int power, const cd* in, int64_t ldi, cd* out, int64_t ldo);
// GMRES over columns c = r * nseg + s (problem s of a batch, right-hand side r)
__global__ void k_dot_partial(int n, int nrhs, int nseg, const int64_t* off, const cd* V, int64_t vstride, const cd* w,
                              cd* partials);
__global__ void k_sub_basis(int n, int nrhs, int nseg, const int* row_seg, int nvec, const cd* V, int64_t vstride,
                            const cd* h, cd* w);
__global__ void k_scale_cols(int n, int nrhs, int nseg, const int* row_seg, const cd* w, const double* s, cd* out);
__global__ void k_add_basis(int n, int nrhs, int nseg, const int* row_seg, int nvec, const cd* Z, int64_t zstride,
                            const cd* y, cd* x, int64_t ldx);

// Per-tet data of the field recovery (compute_fields, src/postprocess.cpp:312-340).
struct TetFieldPlan {
    const double* coords;
    const int* tets;
    const int* tet_body;
    const double* body_sigma;
    const int* dofs;          // 16 per tet: 6 edge unknowns, 10 Phi unknowns
    const int8_t* signs;      // 6 per tet
    const double* phi_coeff;  // 10 per tet
    const cd* phi_fixed;      // 10 per tet
};
// nseg > 1 (a batch): tet t of problem s (seg_tets bounds) takes -j w and the Phi
// scale from seg_factors[2 s], [2 s + 1] instead of minus_jw, scale
__global__ void k_tet_current(int num_tets, TetFieldPlan tp, int nrhs, const cd* x, int64_t ldx, cd minus_jw, cd scale,
                              double ax, double ay, double az, int nseg, const int64_t* seg_tets, const cd* seg_factors,
                              cd* contrib);
// grid nrhs * nseg: problem s of right-hand side r at current[r * nseg + s];
// lengths per problem, or `length` for all when null
__global__ void k_current_sum(int num_tets, int nseg, const int64_t* seg_tets, TetFieldPlan tp, const cd* contrib,
                              const double* lengths, double length, cd* current);
__global__ void k_tet_efield(int num_tets, TetFieldPlan tp, const cd* x, cd minus_jw, cd scale, cd* e_tet);
// compute_fields: per tet (node_val: A then E at each of the ten P2 nodes, 6
// complex per (tet, local node)), then per P2 node over its entries in tet order
__global__ void k_fields_tet(int num_tets, TetFieldPlan tp, const double* body_mu_r, const cd* x, cd minus_jw, cd scale,
                             cd* a_tet, cd* b_tet, cd* e_tet, cd* j_tet, double* sigma_tet, double* tet_vol,
                             double* tet_mu, cd* node_val);
__global__ void k_fields_node(int num_p2, const int64_t* node_ptr, const int* node_entry, const double* sigma_tet,
                              const double* tet_mu, const double* tet_vol, const cd* b_tet, const cd* node_val,
                              cd* a_node, cd* b_node, cd* h_node, cd* e_node, cd* j_node, double* vertex_weight,
                              int8_t* on_interface);

// random.hip: normal variates of the counter-based generator (philox.hpp)
__global__ void k_random_normal(uint64_t seed, uint64_t offset, int64_t count, double* out);

// gauge-free potential (src/gauge_free_potential.cpp:87-177)
__global__ void k_gf_tet(int num_tets, const double* coords, const int* tets, const int* tet_body, const int* tet_on,
                         const cd* beta, const cd* e_tet, cd* terms);
__global__ void k_gf_reduce(int64_t num_targets, int64_t nnz, const int64_t* tptr, const uint32_t* entry,
                            const uint8_t* kind, const cd* values, const cd* terms, cd* A, cd* rhs);
__global__ void k_gf_scatter(int num_nodes, const int* column, const cd* prescribed, const cd* x, cd* phi);

}  // namespace aphi_dev

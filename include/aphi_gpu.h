/* C-ABI of the MI355X A-Phi path: element assembly into the symmetric CSR, and
 * complex-symmetric LDL^T factorization and solve, on one gfx950 GPU.
 *
 * Plain C types only: pointers, sizes, status codes. Complex arrays are
 * interleaved (re, im) doubles, the layout of std::complex<double>. Arrays
 * documented as "device" are HIP device pointers; everything else is host
 * memory. Every call returns an aphi_status; on failure aphi_last_error()
 * returns the message (thread-local, valid until the next call on the thread).
 *
 * Which reference entry point each call replaces is noted beside it
 * (paths relative to the reference repository).
 */
#ifndef APHI_GPU_H
#define APHI_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The host mirror maps them back onto the reference's error
 * conventions: INVALID_ARGUMENT -> std::invalid_argument, LOGIC_ERROR ->
 * std::logic_error, RUNTIME_ERROR -> std::runtime_error, PIVOT_FAILED -> the
 * `false` return of factorize_ldlt / solve_symmetric with the stats filled. */
typedef enum aphi_status {
    APHI_OK = 0,
    APHI_INVALID_ARGUMENT = 1,
    APHI_LOGIC_ERROR = 2,
    APHI_RUNTIME_ERROR = 3,
    APHI_DEVICE_ERROR = 4,
    APHI_OUT_OF_MEMORY = 5,
    APHI_PIVOT_FAILED = 6
} aphi_status;

/* include/aphi_solver/problem.hpp: enum class Conditioning */
typedef enum aphi_conditioning { APHI_NATURAL = 0, APHI_ROW_SCALED = 1, APHI_SCALED_PHI = 2 } aphi_conditioning;
/* include/aphi_solver/ordering.hpp: enum class Ordering */
typedef enum aphi_ordering { APHI_ORDER_NATURAL = 0, APHI_ORDER_RCM = 1, APHI_ORDER_AMD = 2 } aphi_ordering;
/* Arithmetic of the numeric factorization: complex128, or complex64 (MFMA
 * FP32) followed by FP64 iterative refinement. */
typedef enum aphi_precision { APHI_FP64 = 0, APHI_FP32 = 1 } aphi_precision;

const char* aphi_last_error(void);
const char* aphi_version(void);

/* ------------------------------------------------------------------ host symbolic problem
 * The integer half of the path, kept on the host and bit-exact with the
 * reference: mesh topology (src/mesh.cpp:26-142), binding and tree-cotree gauge
 * (src/problem_binding.cpp:208-700, src/tree_cotree.cpp:89-251), DOF map
 * (src/dof_map.cpp:8-194), upper-triangle sparsity (src/sparsity.cpp:19-136),
 * ScatterMap (src/assembly.cpp:196-232), ordering (src/ordering.cpp:54-608) and
 * the elimination tree / counts (src/symbolic.cpp:58-196), plus the supernodal
 * plan the device factorization uses. */
typedef struct aphi_problem aphi_problem;

/* A Kuhn-split cube (tools/generate_cube_mesh.py:20-48) of n^3 hexes, side
 * `side` metres, with a conductor bar of hex columns [bar_i0,bar_i1) x
 * [bar_j0,bar_j1) spanning z; voltage ports on the bar's z=side and z=0 faces;
 * n x A = 0 on the outer boundary; full-wave Phi. */
typedef struct aphi_cube_spec {
    int n;
    double side;
    int bar_i0, bar_i1, bar_j0, bar_j1;
    double sigma, eps_r, mu_r;
    double air_sigma, air_eps_r, air_mu_r;
    double v_top_re, v_top_im, v_bottom_re, v_bottom_im;
    int ordering;          /* aphi_ordering */
    int relax_small;       /* supernode amalgamation: <= 0 selects the default */
    double relax_fill;     /* <= 0 selects the default */
    /* 1: the z=side port is current-driven (a boundary-current port of amplitude
     * i_top amperes) instead of held at v_top */
    int top_current;
    double i_top_re, i_top_im;
} aphi_cube_spec;

typedef struct aphi_problem_counts {
    int64_t nodes, edges, faces, tets;
    int64_t num_a, num_phi, num_ports, num_total; /* DofMap counts */
    int64_t nnz;                                  /* stored upper-triangle entries */
    int64_t predicted_nnz;                        /* nnz(L) incl. diagonal (SolverAnalysis) */
    int64_t num_roots;
    int64_t tree_edges, dirichlet_edges, free_edges;
    int64_t num_super, num_levels, factor_entries, cb_pool_entries, contributions;
    double flops;                                 /* 8*sum_j [m_j(m_j+1)/2 + 2 m_j] */
    double analyze_ms, build_ms;
} aphi_problem_counts;

int aphi_problem_create_cube(const aphi_cube_spec* spec, aphi_problem** out);
/* count cube problems at once, built on up to `threads` host threads (0: one per
 * hardware thread): the batched and serving workloads' host symbolic phase, each
 * problem exactly what aphi_problem_create_cube gives for its spec. On a refusal
 * every out[i] is NULL and the message names the first failing index. */
int aphi_problem_create_cubes(const aphi_cube_spec* specs, int count, int threads, aphi_problem** out);
/* A problem from a caller's matrix pattern alone: the upper-triangle CSR of an
 * n x n complex-symmetric matrix (row_ptr n+1 entries from 0, columns strictly
 * ascending per row, none below the diagonal), refused with the messages of
 * SparseSymmetric::from_pattern (include/aphi_solver/sparse_symmetric.hpp:38-83).
 * Runs analyze(pattern, ordering, max_factor_nnz) (include/aphi_solver/symbolic.hpp:104,
 * src/symbolic.cpp:126-196; max_factor_nnz 0 = no budget) and builds the
 * supernodal plan. The problem has no mesh and no assembly plan: a device mesh
 * made from it takes its values through aphi_system_upload, and aphi_refill /
 * aphi_terminal_current refuse it. This is the path of any other caller of
 * solve_symmetric, e.g. the gauge-free potential (src/gauge_free_potential.cpp:177). */
int aphi_problem_create_csr(int32_t n, const int32_t* row_ptr, const int32_t* col_index, int ordering,
                            int64_t max_factor_nnz, int relax_small, double relax_fill, aphi_problem** out);
/* A problem from the reference host's own objects, flattened: what a caller of
 * refill(SymmetricSystem&, const BoundProblem&, const Mesh&, const DofMap&,
 * const SparsityPattern&, double omega, Conditioning, const ScatterMap*)
 * (include/aphi_solver/assembly.hpp:197-199) holds, for any mesh it has bound
 * (a .msh file, a generated cube). From it the library builds the device
 * assembly plan (every contribution to every slot and rhs entry in the order
 * src/assembly.cpp:255-422 accumulates them), then analyze(pattern, ordering)
 * (include/aphi_solver/symbolic.hpp:104) and the supernodal plan; describe() gives
 * the device mesh. The ScatterMap (ScatterMap::build, assembly.hpp:82-83) is taken
 * as given, or built from the pattern when its arrays are NULL. Refused (with the
 * reference's messages where it has them): indices out of range, a tet in no body
 * or in two, a pattern that is not upper-triangle CSR, a scatter slot the pattern
 * lacks. Such a problem has no mesh topology beyond its tets: the gauge-free
 * potential and the result writers refuse it. */
typedef struct aphi_bound_desc {
    int32_t num_nodes, num_tets, num_bodies, num_total, num_ports;
    const double* coords;          /* 3 * num_nodes: Mesh::nodes, metres */
    const int32_t* tet_nodes;      /* 4 * num_tets: Mesh::tets */
    const int8_t* tet_edge_signs;  /* 6 * num_tets: Mesh::tet_edge_signs */
    const double* body_sigma;      /* num_bodies: BoundProblem::bodies, in order */
    const double* body_eps_r;
    const double* body_mu_r;
    const int32_t* body_ptr;       /* num_bodies + 1 */
    const int32_t* body_tets;      /* body_ptr[num_bodies] = num_tets: each body's tets, ascending */
    /* DofMap::local_dofs(t, mesh, bound) of every tet: TetDofs::edge, ::phi, ::phi_fixed */
    const int32_t* edge_index;     /* 6 * num_tets, -1 eliminated */
    const double* edge_coeff;      /* 6 * num_tets */
    const int32_t* phi_index;      /* 10 * num_tets, -1 eliminated */
    const double* phi_coeff;       /* 10 * num_tets */
    const double* phi_fixed;       /* 20 * num_tets: complex prescribed values */
    /* ports, in BoundProblem::ports order */
    const int32_t* port_index;     /* num_ports: DofMap::port_index */
    const int32_t* port_current;   /* num_ports: 1 current-driven, 0 voltage-driven */
    const double* port_amplitude;  /* 2 * num_ports: Port::amplitude (current ports) or
                                      DofMap::port_value (voltage ports), complex */
    /* SparsityPattern, upper triangle */
    const int32_t* row_ptr;        /* num_total + 1 */
    const int32_t* col_index;      /* row_ptr[num_total] */
    /* ScatterMap (all four, or all NULL) */
    const int32_t* scatter_live;   /* num_tets: ScatterMap::live_count */
    const int32_t* scatter_pos;    /* 16 * num_tets: edge_at then phi_at */
    const int64_t* scatter_offset; /* num_tets + 1: start of each tet's packed block */
    const int32_t* scatter_slot;   /* scatter_offset[num_tets]: ScatterMap::slots */
} aphi_bound_desc;
int aphi_problem_create_bound(const aphi_bound_desc* desc, int ordering, int relax_small, double relax_fill,
                              aphi_problem** out);
void aphi_problem_destroy(aphi_problem* p);
int aphi_problem_counts_get(const aphi_problem* p, aphi_problem_counts* out);
/* Named integer arrays, copied into out[0..capacity); *count receives the full
 * length (call with capacity 0 to size). Names: tet_nodes, edges, tet_edges,
 * tet_edge_signs, tet_body, dirichlet_edge, tree_edge, edge_index, phi_index,
 * port_index, row_ptr, col_index, perm, iperm, parent, row_count, col_count,
 * scatter_live, scatter_pos, scatter_offset, scatter_slot, super_first,
 * super_parent, super_row_ptr, super_rows, super_rel, super_level, factor_perm, col_ref,
 * slot_to_panel, tet_dofs (6 edge + 10 phi per tet), body_ptr, body_tets, phi_state
 * (per P2 node: 0 free, 1 absent, 2 fixed, 3 port, 4 cut), phi_port, port_is_fixed;
 * of a gauge-free problem also gf_column, gf_port_column, gf_tet_on,
 * gf_target_ptr, gf_entry, gf_kind. */
int aphi_problem_get_int(const aphi_problem* p, const char* name, int64_t* out, int64_t capacity, int64_t* count);
/* Named double arrays: coords (3 per node), tet_phi_fixed (10 complex per tet),
 * body_param (sigma, eps_r, mu_r per body), port_value (complex per port: the
 * prescribed value of a voltage port, 0 for a current port), port_amplitude
 * (complex per port: Port::amplitude), tet_edge_coeff (6 per tet); of a
 * gauge-free problem also gf_prescribed (complex per vertex), gf_values. */
int aphi_problem_get_double(const aphi_problem* p, const char* name, double* out, int64_t capacity, int64_t* count);
/* Result files of a solution x (host, num_total complex, e.g. downloaded from
 * the device), in the reference's formats and number formatting: format 0 is
 * potential_at_nodes + write_potential (src/postprocess.cpp:42-123, 549-606),
 * format 1 is write_vtk without field data (src/postprocess.cpp:1055-1276), both
 * without run information. *bytes (may be NULL) receives the file size. */
int aphi_write_potential(const aphi_problem* p, const char* path, const double* x, double omega, int conditioning,
                         int format, int64_t* bytes);
/* The log sweep of src/input_file.cpp:19-50. */
int aphi_frequency_sweep(double start, double stop, int points, int logarithmic, double* out);

/* ------------------------------------------------------------------ field recovery
 * compute_fields (src/postprocess.cpp:223-481) and write_vtk with field data
 * (src/postprocess.cpp:1055-1276). The host plan says where each tet's ten P2
 * nodes (4 vertices, then the 6 local edges in kTetLocalEdgeVerts order) land and,
 * per P2 node, which (tet, local node) pairs contribute to its averages, tets
 * ascending -- the reference's accumulation order. */
typedef struct aphi_fields_desc {
    int32_t num_tets, num_p2;      /* P2 nodes: vertices, then edge midpoints */
    const int32_t* tet_p2;         /* 10 * num_tets */
    const int64_t* node_ptr;       /* num_p2 + 1 */
    const int32_t* node_entry;     /* 10 * num_tets: tet * 10 + local node, grouped by node */
} aphi_fields_desc;
/* Pointers into the problem's arrays (built on first call); a problem with no
 * mesh topology (pattern-only, bound or gauge-free) is refused. */
int aphi_problem_describe_fields(const aphi_problem* p, aphi_fields_desc* out);

/* The FieldOutput of compute_fields in caller-owned host arrays, complex values
 * interleaved (re, im): per tet A, B, E, J (3 complex each) and sigma; per P2
 * node the averages A, B, H, E, J (3 complex each), the A weight (the node's
 * incident volume) and the material-interface flag. */
typedef struct aphi_fields {
    double *a_tet, *b_tet, *e_tet, *j_tet;   /* 6 * num_tets each */
    double* sigma_tet;                       /* num_tets */
    double *a_node, *b_node, *h_node, *e_node, *j_node;  /* 6 * num_p2 each */
    double* vertex_weight;                   /* num_p2 */
    int8_t* on_material_interface;           /* num_p2 */
    int32_t num_interface_nodes, num_orphan_vertices;  /* filled on return */
} aphi_fields;
/* write_vtk(path, mesh, potential_at_nodes(x), solution, &fields, FieldSet::All)
 * without run information: the potential file of format 1 plus the cell data
 * (sigma, B, E, J per tet) and the point data (A, B, H, E, J and their phasor
 * magnitudes, the material-interface flag), in the reference's order and number
 * formatting. x: host, num_total complex. */
int aphi_write_vtk_fields(const aphi_problem* p, const char* path, const double* x, double omega, int conditioning,
                          const aphi_fields* fields, int64_t* bytes);

/* ------------------------------------------------------------------ device boundary
 * What the reference host hands the GPU once per mesh: geometry, the assembly
 * plan (every contribution to every slot / rhs entry, in the reference's
 * accumulation order), the pattern, and the supernodal factor plan. */
typedef struct aphi_mesh_desc {
    int32_t num_nodes, num_tets, num_bodies, n;
    int64_t nnz;
    const double* coords;          /* 3 * num_nodes, metres */
    const int32_t* tet_nodes;      /* 4 * num_tets */
    const int32_t* tet_body;       /* num_tets, index into the body arrays */
    const double* body_sigma;      /* num_bodies */
    const double* body_eps_r;
    const double* body_mu_r;
    /* assembly plan: contributions for target t (slot t < nnz, rhs row t - nnz)
     * are [target_ptr[t], target_ptr[t+1]) */
    const int64_t* target_ptr;     /* nnz + n + 1 */
    const uint32_t* entry;         /* tet * 196 + element entry */
    const uint8_t* kind;           /* contribution kind | sign << 4 | value id << 5 */
    int32_t num_values;            /* <= 8 prescribed values */
    const double* values;          /* 2 * num_values (complex) */
    /* upper-triangle CSR of the original ordering */
    const int32_t* row_ptr;        /* n + 1 */
    const int32_t* col_index;      /* nnz */
    /* supernodal plan (factorization order) */
    const int32_t* factor_perm;    /* n: factor position -> original unknown */
    const int32_t* col_ref;        /* n: factor position -> analysis (reference) position */
    int32_t num_super, num_levels;
    const int32_t* super_first;    /* num_super + 1 */
    const int32_t* super_parent;   /* num_super */
    const int64_t* super_row_ptr;  /* num_super + 1 */
    const int32_t* super_rows;     /* super_row_ptr[num_super] */
    const int32_t* super_rel;      /* aligned with super_rows */
    const int32_t* super_level;    /* num_super */
    const int64_t* panel_offset;   /* num_super + 1 */
    const int64_t* cb_offset;      /* num_super */
    int64_t cb_pool_entries;
    const int64_t* slot_to_panel;  /* nnz */
    double flops;                  /* algorithmic flops of one factorization */
    int64_t predicted_nnz;
    /* per-tet unknowns for the field recovery of compute_fields
     * (src/postprocess.cpp:312-340): DofMap::local_dofs of every tet */
    const int32_t* tet_dofs;       /* 16 * num_tets: 6 edge unknowns, 10 Phi unknowns (-1: eliminated) */
    const int8_t* tet_edge_signs;  /* 6 * num_tets: Mesh::tet_edge_signs */
    const double* tet_phi_coeff;   /* 10 * num_tets */
    const double* tet_phi_fixed;   /* 20 * num_tets: prescribed Phi (complex) where the unknown is -1 */
} aphi_mesh_desc;

/* Pointers into the problem's own arrays; valid while the problem lives. */
int aphi_problem_describe(const aphi_problem* p, aphi_mesh_desc* out);

/* ------------------------------------------------------------------ gauge-free potential
 * recover_gauge_free_potential (src/gauge_free_potential.cpp:87-177): the P1
 * system <beta grad Phi', grad lam> = -<beta E, grad lam> over the mesh
 * vertices, beta = sigma + j w eps0 eps_r, solved through solve_symmetric.
 *
 * aphi_problem_create_gauge_free builds, from a cube problem, the vertex map
 * (build_vertex_map, :47-85), the system's upper-triangle pattern (the
 * reference's per-row std::map, :108-153) with its ordering / elimination tree /
 * supernodal plan, and the ordered list of per-tet terms behind every slot and
 * rhs entry. The result is a pattern-only problem (like aphi_problem_create_csr)
 * whose describe() gives the device mesh; aphi_gauge_free_attach then hands that
 * mesh the plan below, aphi_gauge_free_refill assembles values and rhs on the
 * device from the field E of each tet (compute_fields' e_tet, see
 * aphi_tet_fields), aphi_solve_symmetric solves, and aphi_gauge_free_potential
 * maps the solution back to vertices and ports (:167-176). The plan assumes
 * w != 0 (so every body with sigma or eps_r nonzero takes part); the refill
 * refuses w == 0. */
int aphi_problem_create_gauge_free(const aphi_problem* cube, int ordering, aphi_problem** out);

typedef struct aphi_gauge_free_desc {
    int32_t num_nodes, num_tets, num_bodies, unknowns;
    int64_t nnz;
    const double* coords;          /* 3 * num_nodes */
    const int32_t* tet_nodes;      /* 4 * num_tets */
    const int32_t* tet_body;       /* num_tets */
    const double* body_sigma;      /* num_bodies */
    const double* body_eps_r;      /* num_bodies */
    const int32_t* tet_on;         /* num_tets: 1 if the tet's body takes part */
    const int32_t* column;         /* num_nodes: unknown of each vertex, -1 if eliminated */
    const double* prescribed;      /* 2 * num_nodes: value of an eliminated vertex (complex) */
    int32_t num_ports;
    const int32_t* port_column;    /* num_ports: unknown of a current-driven port, -1 if fixed */
    const double* port_value;      /* 2 * num_ports: DofMap::port_value (complex) */
    /* contributions to target t (slot t < nnz, rhs row t - nnz): [target_ptr[t],
     * target_ptr[t+1]); entry = tet * 20 + term (0..15: k_ij at i * 4 + j,
     * 16..19: source term of vertex i); kind = 0 (slot += k_ij), 1 (rhs -=
     * source_i), 2 | id << 2 (rhs -= k_ij * values[id]) */
    const int64_t* target_ptr;     /* nnz + unknowns + 1 */
    const uint32_t* entry;
    const uint8_t* kind;
    int32_t num_values;            /* <= 64 */
    const double* values;          /* 2 * num_values */
} aphi_gauge_free_desc;

int aphi_problem_describe_gauge_free(const aphi_problem* p, aphi_gauge_free_desc* out);

typedef struct aphi_context aphi_context;
typedef struct aphi_mesh aphi_mesh;

int aphi_context_create(int device, aphi_context** out);
/* Meshes created on a context must be destroyed before it. */
void aphi_context_destroy(aphi_context* ctx);
int aphi_context_synchronize(aphi_context* ctx);

/* Uploads one mesh's plan (once per mesh; reused at every frequency). */
int aphi_mesh_create(aphi_context* ctx, const aphi_mesh_desc* desc, aphi_mesh** out);
void aphi_mesh_destroy(aphi_mesh* mesh);
/* Per-launch event timing of the factorization's trailing-update kernel (adds
 * synchronisation-free event records around each launch; off by default).
 * on = 1: the usual schedule, whose look-ahead runs part of the updates on a second
 * stream beside the next block's (overlapping launches share the compute units, so
 * each is timed slower than it runs alone); on = 2: every launch on the library
 * stream, one after the other (the kernel's rate alone; the factor is the same). */
int aphi_mesh_set_profiling(aphi_mesh* mesh, int on);
/* The solve's per-level launches, and the factorization's, are captured into
 * HIP graphs and replayed (on by default; APHI_SOLVE_GRAPHS=0 and
 * APHI_FACTOR_GRAPHS=0 in the environment turn them off for all meshes). A capture fails while another host thread issues work on the legacy
 * default stream (a synchronous copy, a library running there), so a mesh used
 * beside other threads' GPU work -- a serving slot -- turns it off here and
 * launches the same kernels one by one (bitwise the same results). */
int aphi_mesh_set_solve_graphs(aphi_mesh* mesh, int on);
/* Tiling of the complex-float factorization's trailing update: 0 (default) picks
 * 64 x 64 tiles for launches whose tallest region has at most 8192 rows or whose
 * widest region has at most 512 columns, and 128 x 128 tiles otherwise; 1 forces
 * the small tiles, 2 the large ones. Both
 * give bitwise the same factor. No counterpart in the reference (tuning and
 * testing switch); other modes are refused (-1). */
int aphi_mesh_set_update_tiles(aphi_mesh* mesh, int mode);
/* Autotuner of the complex-float trailing update's tiles (no counterpart in the
 * reference). Each update launch falls in a class by its tallest region (rows:
 * <= 1024, <= 2048, ... , 8 buckets) and its widest region (columns: <= 64,
 * <= 128, ..., 8 buckets); the table holds per class 0 (the default rule of
 * aphi_mesh_set_update_tiles), 1 (64 x 64 tiles) or 2 (128 x 128 tiles), row-major
 * [height class][width class], 64 entries. The tuner factors the mesh's assembled
 * system in FP32 with every launch timed, once with each tiling (`repeats` times),
 * picks the faster tiling per class, then times whole factorizations with the
 * default rule and with that table and keeps the table only if it is faster.
 * Every tiling gives the same factor bits, so tuning never changes results; the
 * mesh holds a valid FP32 factor of its system afterwards. A table learnt on one
 * mesh can be handed to others (aphi_mesh_set_update_tile_table). A mode set by
 * aphi_mesh_set_update_tiles other than 0 overrides the table. */
int aphi_mesh_tune_update_tiles(aphi_mesh* mesh, int repeats, int8_t* table_out, double* ms_default, double* ms_tuned);
int aphi_mesh_get_update_tile_table(const aphi_mesh* mesh, int8_t* table);
int aphi_mesh_set_update_tile_table(aphi_mesh* mesh, const int8_t* table);
/* Device bytes currently held by this mesh (plan + numeric buffers). */
int64_t aphi_mesh_device_bytes(const aphi_mesh* mesh);
/* Frees the mesh's per-frequency state -- assembled values, equilibration, factor,
 * contribution blocks, solve and refinement work arrays, captured graphs -- and
 * keeps the plan (what ScatterMap::build and analyze produced: the symbolic factor),
 * so the mesh holds only its plan until the next aphi_refill / aphi_system_upload.
 * For callers that keep many meshes resident under a memory budget (a serving
 * loop): the next frequency on the mesh reuses the plan and allocates its numeric
 * state again. Calls that need an assembled system or a factor are refused until
 * then. Waits for the mesh's streams first. */
int aphi_mesh_release_workspace(aphi_mesh* mesh);

/* refill(SymmetricSystem&, bound, mesh, dofs, pattern, omega, conditioning, map)
 * -- include/aphi_solver/assembly.hpp:197, src/assembly.cpp:468-494. Clears and
 * assembles the upper-triangle values and the rhs on the device. Refuses a
 * conditioning that is not symmetric and omega == 0 under a scaled one. */
int aphi_refill(aphi_mesh* mesh, double omega, int conditioning);
/* The assembled system, copied to host: values 2*nnz doubles, rhs 2*n. */
int aphi_system_download(aphi_mesh* mesh, double* values, double* rhs);
/* Replaces the device system with a caller-assembled one (host arrays). */
int aphi_system_upload(aphi_mesh* mesh, const double* values, const double* rhs);
/* Device pointer to the assembled rhs (n complex), for solving the port system. */
const void* aphi_rhs_device(const aphi_mesh* mesh);
/* The assembled rhs copied into dst (device, n complex), device to device. */
int aphi_rhs_copy(aphi_mesh* mesh, void* dst);

/* Port current and impedance of each of nrhs solutions x (device, n complex per
 * column, leading dimension ldx): terminal_current (src/postprocess.cpp:526-547)
 * over the centroid current density J = sigma E of compute_fields
 * (src/postprocess.cpp:311-400, E = -j omega A - grad Phi), and impedance
 * (include/aphi_solver/postprocess.hpp:356-358). current: 2 * nrhs doubles
 * (complex I); resistance / inductance (may be null): Re(V/I) and Im(V/I)/omega
 * for the drive voltage V. The sum over tets runs in tet order, as the
 * reference's does, so a given x yields the reference's I bit for bit. Refuses
 * length <= 0 like the reference. */
int aphi_terminal_current(aphi_mesh* mesh, int nrhs, const void* x, int64_t ldx, double omega, int conditioning,
                          const double* axis, double length, double voltage, double* current, double* resistance,
                          double* inductance);

/* aphi_terminal_current without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * per-tet terms and the sums go to a scratch array compiled into the library,
 * one per device (32 MiB, held until the process ends), so num_tets * nrhs must
 * be at most 2,097,152 (e.g. up to 10 right-hand sides on 196,608 tets); larger
 * calls are refused (APHI_INVALID_ARGUMENT). Results are bitwise those of
 * aphi_terminal_current. Calls that overlap on one device (from several
 * threads or contexts) share the array and race: serialise them. */
int aphi_terminal_current_noalloc(aphi_mesh* mesh, int nrhs, const void* x, int64_t ldx, double omega,
                                  int conditioning, const double* axis, double length, double voltage,
                                  double* current, double* resistance, double* inductance);

/* ------------------------------------------------------------------ seeded draws
 * One counter-based generator for every seeded input (right-hand sides, problem
 * draws), the same bits on the host and on the device: Philox-4x32-10 (Salmon et
 * al., SC 2011) and a Box-Muller transform whose log, sqrt, sin and cos are
 * written in IEEE additions, multiplications and divisions. Normal number k of
 * stream `seed` uses counter block k / 2. No counterpart in the reference, whose
 * inputs come from files; this replaces its callers' ad-hoc draws for the
 * synthetic workloads. */
/* One Philox-4x32-10 block: out = philox(ctr[4], key[2]). */
int aphi_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
/* out[i] = normal number offset + i of stream seed (host memory). */
int aphi_random_normal_host(uint64_t seed, uint64_t offset, int64_t count, double* out);
/* Uniform number offset + i of stream seed in [0, 1), host memory: the 53-bit
 * words (w1:w0 >> 11) and (w3:w2 >> 11) of counter block k / 2, times 2^-53. */
int aphi_random_uniform_host(uint64_t seed, uint64_t offset, int64_t count, double* out);
/* The normals into device memory, on the context's stream. */
int aphi_random_normal(aphi_context* ctx, uint64_t seed, uint64_t offset, int64_t count, void* out);

/* The field E = -j w A - grad Phi of every tet at its centroid, for one solution
 * x (device, n complex): the e_tet of compute_fields (src/postprocess.cpp:311-391),
 * in its operation order, so bit for bit the reference's for the same x. e_tet:
 * device, 3 complex per tet. */
int aphi_tet_fields(aphi_mesh* mesh, const void* x, double omega, int conditioning, void* e_tet);

/* compute_fields (src/postprocess.cpp:223-481) on the device for a solution x
 * (device, n complex): per tet A and E at the centroid, J = sigma E and B = curl A;
 * per P2 node the volume-weighted averages of A (all incident tets), E (the most
 * conducting side) and B (the most permeable side), each node's sum in the
 * reference's tet order and every term in its operation order, so bit for bit
 * the reference's FieldOutput for the same x; then J = sigma_max E and
 * H = B / mu_max at the nodes. The mesh takes the field plan once
 * (aphi_fields_attach, from aphi_problem_describe_fields); the results land in
 * the caller's host arrays. */
int aphi_fields_attach(aphi_mesh* mesh, const aphi_fields_desc* desc);
int aphi_compute_fields(aphi_mesh* mesh, const void* x, double omega, int conditioning, aphi_fields* out);

/* Hands a mesh made from a gauge-free problem its assembly plan (once). */
int aphi_gauge_free_attach(aphi_mesh* mesh, const aphi_gauge_free_desc* desc);
/* The assembly of recover_gauge_free_potential (src/gauge_free_potential.cpp:112-164)
 * on the device: values and rhs of the P1 system from e_tet (device, 3 complex
 * per tet of the source mesh), each sum in the reference's order. */
int aphi_gauge_free_refill(aphi_mesh* mesh, const void* e_tet, double omega);
/* phi_vertex (device, num_nodes complex) and port_potential (host, 2 * num_ports
 * doubles) from the solution x (device, unknowns complex), as
 * src/gauge_free_potential.cpp:167-176 does. */
int aphi_gauge_free_potential(aphi_mesh* mesh, const void* x, void* phi_vertex, double* port_potential);

/* compute_symmetric_equilibration(a, iterations) + apply (src/equilibration.cpp:94-162).
 * iterations 0 disables scaling. d_out (host, n) may be NULL. */
int aphi_equilibrate(aphi_mesh* mesh, int iterations, double* d_out);

/* factorize_ldlt (include/aphi_solver/factorization.hpp:82, src/factorization.cpp:109-218)
 * of the equilibrated system. Returns APHI_PIVOT_FAILED (stats->ok = 0) when a
 * pivot is at or below pivot_floor. */
typedef struct aphi_factor_stats {
    int ok;
    int columns_done;
    double smallest_pivot, largest_pivot, largest_multiplier;
    int64_t nnz;
    double milliseconds;           /* whole factorization, HIP events on the library stream */
    double flops;                  /* algorithmic: 8*sum_j [m_j(m_j+1)/2 + 2 m_j] */
    /* Filled when profiling is on (aphi_mesh_set_profiling): the MFMA trailing
     * update, timed launch by launch with HIP events, and the real flops those
     * launches perform (8 per complex multiply-add over the lower triangles). */
    int64_t update_launches;
    double update_ms;
    double update_flops;
} aphi_factor_stats;
int aphi_factorize(aphi_mesh* mesh, int precision, double pivot_floor, aphi_factor_stats* stats);

/* D of the current factor, in factorization order (2*n doubles). Column j of the
 * factorization order is column col_ref[j] of the reference's: LDL^T without
 * pivoting is unique for a given elimination tree postorder, so this is the
 * reference's SymmetricFactor::diagonal (include/aphi_solver/factorization.hpp:41)
 * permuted by col_ref. */
int aphi_factor_diagonal(aphi_mesh* mesh, double* out);

/* solve (include/aphi_solver/factorization.hpp:110) for nrhs right-hand sides
 * with the current factor, undoing the equilibration: x = D * solve(D * b).
 * b and x are device arrays, column-major with leading dimensions ldb, ldx. */
int aphi_solve(aphi_mesh* mesh, int nrhs, const void* b, int64_t ldb, void* x, int64_t ldx);

/* solve_symmetric (include/aphi_solver/factorization.hpp:206) extended with FP64
 * iterative refinement: equilibrate, factor (FP64 or FP32), solve, then refine
 * with the FP64 residual of the ORIGINAL system until, for every right-hand side,
 * both the reference's normwise backward error and the same measure on the
 * equilibrated system are <= tolerance. An FP32 factor whose refinement has not
 * converged after max_refinements steps is replaced by an FP64 factor and the
 * refinement continues (report->precision_used says which finished). */
typedef struct aphi_solve_options {
    int precision;                 /* aphi_precision */
    int equilibration_iterations;  /* 10 in the reference; 0 disables */
    double pivot_floor;
    int max_refinements;           /* 0: no refinement (the reference's behaviour) */
    double tolerance;              /* backward-error target, e.g. 1e-12 */
} aphi_solve_options;

typedef struct aphi_solve_report {
    int ok;
    int refinement_steps;
    int precision_used;            /* FP32 falls back to FP64 when refinement stalls */
    double backward_error;         /* max over right-hand sides (reference definition) */
    double scaled_backward_error;  /* the same measure on the equilibrated system D A D */
    double residual;               /* max ||Ax-b||/||b|| */
    double equilibration_min, equilibration_max;
    double smallest_pivot, largest_pivot, largest_multiplier;
    int64_t factor_nnz;
    int columns_done;
    double factorize_ms, solve_ms, total_ms;
} aphi_solve_report;

int aphi_solve_symmetric(aphi_mesh* mesh, int nrhs, const void* b, int64_t ldb, void* x, int64_t ldx,
                         const aphi_solve_options* options, aphi_solve_report* report, double* backward_errors);

/* backward_error (include/aphi_solver/factorization.hpp:140, src/factorization.cpp:401-416)
 * and relative_residual (:315-329) per right-hand side, on the ORIGINAL system. */
int aphi_backward_error(aphi_mesh* mesh, int nrhs, const void* x, int64_t ldx, const void* b, int64_t ldb,
                        double* backward_errors, double* residuals);

/* ------------------------------------------------------------------ batches
 * Many independent problems in one device plan: the reference solves one problem
 * at a time (tools/solve_mesh.cpp:207-236 calls refill, src/assembly.cpp:255-422,
 * then solve_symmetric, src/factorization.cpp:332-398); a batch does the same for
 * every problem with shared launches. It is the block-diagonal union of the
 * problems' meshes: their unknowns side by side (problem p's at
 * [row_offsets[p], row_offsets[p+1]) of every vector), one element and one
 * scatter launch over all tets and slots, the fronts of equal elimination-tree
 * height of every problem factored in the same launches, and the multi-RHS solves
 * over all of them at once. Measures (backward error, refinement's inner
 * products) stay per problem: each problem's result is what it would be alone.
 * The generic mesh calls (aphi_equilibrate, aphi_factorize, aphi_solve,
 * aphi_refill with one frequency for all, aphi_system_download) act on the whole
 * batch; the calls whose outputs are per problem refuse a batch and have the
 * aphi_batch_ forms below. */
int aphi_batch_create(aphi_context* ctx, const aphi_mesh_desc* descs, int count, aphi_mesh** out);
/* count (may be NULL); row_offsets, tet_offsets: count + 1 entries each (may be NULL). */
int aphi_batch_info(const aphi_mesh* batch, int* count, int64_t* row_offsets, int64_t* tet_offsets);
/* refill of every problem, problem p at angular frequency omegas[p]. */
int aphi_batch_refill(aphi_mesh* batch, const double* omegas, int conditioning);
/* solve_symmetric of every problem with the options of aphi_solve_symmetric; b, x
 * span all problems' unknowns (n of the batch per right-hand side). reports:
 * count entries, each with its problem's backward errors (the factorization's
 * fields are the batch's). backward_errors (may be NULL): count * nrhs, problem-major. */
int aphi_batch_solve_symmetric(aphi_mesh* batch, int nrhs, const void* b, int64_t ldb, void* x, int64_t ldx,
                               const aphi_solve_options* options, aphi_solve_report* reports, double* backward_errors);
/* backward_error and relative_residual per problem and right-hand side
 * (count * nrhs each, problem-major). */
int aphi_batch_backward_error(aphi_mesh* batch, int nrhs, const void* x, int64_t ldx, const void* b, int64_t ldb,
                              double* backward_errors, double* residuals);
/* terminal_current + impedance per problem: problem p at omegas[p] with length
 * lengths[p] and drive voltage voltages[p], one axis for all; current 2 * count *
 * nrhs doubles, resistance / inductance count * nrhs (problem-major). Each
 * problem's sum runs over its own tets in order, as aphi_terminal_current does. */
int aphi_batch_terminal_current(aphi_mesh* batch, int nrhs, const void* x, int64_t ldx, const double* omegas,
                                int conditioning, const double* axis, const double* lengths, const double* voltages,
                                double* current, double* resistance, double* inductance);

#ifdef __cplusplus
}
#endif

#endif /* APHI_GPU_H */

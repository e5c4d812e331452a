// Host half of the A-Phi drop-in: everything integer (mesh numbering, binding,
// gauge, DOF map, sparsity, scatter map, ordering, elimination tree, supernodes)
// plus the per-mesh plans the device kernels consume.
//
// This is the part of the path the north-star keeps on the host, and it has to
// be bit-exact with the reference: the matrix pattern, and therefore every slot
// index and the factor's structure, is a function of these integers only.
// Each routine cites the reference routine whose observable behaviour (output
// order included) it reproduces.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace aphi_amd {

using Complex = std::complex<double>;

// include/aphi_solver/constants.hpp:12-13
constexpr double kMu0 = 4.0e-7 * 3.14159265358979323846;
constexpr double kEps0 = 8.8541878128e-12;

// include/aphi_solver/problem.hpp (Conditioning), ordering.hpp (Ordering)
enum class Conditioning : int { Natural = 0, RowScaled = 1, ScaledPhi = 2 };
enum class Ordering : int { Natural = 0, ReverseCuthillMcKee = 1, ApproximateMinimumDegree = 2 };
enum class PortType : int { BoundaryCurrent = 0, BoundaryVoltage = 1, InternalCurrent = 2, InternalVoltage = 3 };

inline bool is_current_driven(PortType t) {
    return t == PortType::BoundaryCurrent || t == PortType::InternalCurrent;
}

struct Vec3 {
    double x = 0.0, y = 0.0, z = 0.0;
    Vec3() = default;
    Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
    double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    Vec3 cross(const Vec3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
};

// Local edge (vertex pair) and face tables: include/aphi_solver/mesh.hpp:44-51, 62-67.
inline constexpr int kEdgeVerts[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {2, 3}, {1, 3}};
inline constexpr int kFaceVerts[4][3] = {{1, 2, 3}, {2, 3, 0}, {3, 0, 1}, {0, 1, 2}};

// ------------------------------------------------------------------ mesh

struct Mesh {
    std::vector<Vec3> nodes;
    std::vector<std::array<int, 4>> tets;
    std::vector<int> tet_tags;
    std::vector<std::array<int, 3>> tagged_faces;  // boundary triangles carrying a tag
    std::vector<int> tagged_face_tag;

    // Derived by build_topology (src/mesh.cpp:26-142).
    std::vector<std::pair<int, int>> edges;          // ascending (lo, hi)
    std::vector<std::array<int, 3>> faces;           // ascending (a, b, c)
    std::vector<std::array<int, 6>> tet_edges;
    std::vector<std::array<int8_t, 6>> tet_edge_signs;
    std::vector<std::array<int, 4>> tet_faces;
    std::vector<std::array<int, 2>> face_tets;
    std::vector<int> face_tet_count;

    int num_nodes() const { return static_cast<int>(nodes.size()); }
    int num_tets() const { return static_cast<int>(tets.size()); }
    int num_edges() const { return static_cast<int>(edges.size()); }
    int num_faces() const { return static_cast<int>(faces.size()); }
    bool is_boundary_face(int f) const { return face_tet_count[static_cast<size_t>(f)] == 1; }

    void build_topology();
    int find_edge(int i, int j) const;
    int find_face(int a, int b, int c) const;
    double signed_tet_volume(int t) const;
};

// The structured Kuhn cube of tools/generate_cube_mesh.py:20-48 (node order k,j,i;
// six tets per hex, all sharing the hex's main diagonal), scaled to `side` metres,
// with a conductor bar of hex columns [bar_i0,bar_i1) x [bar_j0,bar_j1) spanning z.
// Tags: air tets 1, conductor tets 2; the bar's z = side face is tagged 10 and its
// z = 0 face 11 (the generator writes none; these are the voltage-port terminals).
struct CubeSpec {
    int n = 6;
    double side = 0.1;
    int bar_i0 = 2, bar_i1 = 4, bar_j0 = 2, bar_j1 = 4;
    double sigma = 5.8e7, eps_r = 1.0, mu_r = 1.0;          // conductor
    double air_sigma = 0.0, air_eps_r = 1.0, air_mu_r = 1.0;
    Complex v_top{1.0, 0.0};
    Complex v_bottom{0.0, 0.0};
    bool top_current = false;  // the z = side port current-driven with amplitude i_top
    Complex i_top{1.0, 0.0};
};

Mesh make_kuhn_cube(const CubeSpec& spec);
double node_coordinate(const CubeSpec& spec, int index);  // side * index / n

// ------------------------------------------------------------------ binding

struct Body {
    std::string name;
    int tag = -1;
    double sigma = 0.0, eps_r = 1.0, mu_r = 1.0;
    std::vector<int> tets;  // ascending
    bool is_conductor() const { return sigma > 0.0; }
};

struct Port {
    std::string name;
    PortType type = PortType::BoundaryVoltage;
    Complex amplitude{0.0, 0.0};
    std::vector<int> faces, nodes, edges;  // sorted, unique
};

struct TreeCotree {
    std::vector<uint8_t> is_tree_edge;
    int tree_edge_count = 0, surface_tree_edge_count = 0, interior_tree_edge_count = 0;
    int num_groups = 0, num_reference_groups = 0, num_dirichlet_components = 0;
};

struct BoundProblem {
    std::vector<Body> bodies;
    std::vector<int> body_of_tet;
    std::vector<Port> ports;
    std::vector<uint8_t> phi_tet;
    std::vector<uint8_t> dirichlet_edge;
    TreeCotree gauge;
    bool full_wave = true;
};

// src/tree_cotree.cpp:253-263
std::vector<uint8_t> boundary_edge_mask(const Mesh& mesh);
// src/tree_cotree.cpp:106-251
TreeCotree build_tree_cotree(const Mesh& mesh, const std::vector<uint8_t>& dirichlet_edge);
// The subset of src/problem_binding.cpp:208-700 the cube configurations exercise:
// bodies (conductor first, then air, as the reference examples list them), two
// boundary-voltage ports, full-wave Phi support, n x A = 0 on the outer boundary.
BoundProblem bind_cube(const Mesh& mesh, const CubeSpec& spec);

// ------------------------------------------------------------------ DOF map

enum class EdgeDof : int8_t { Free = 0, Dirichlet = 1, Tree = 2 };
enum class PhiDof : int8_t { Free = 0, Absent = 1, Fixed = 2, Port = 3, Cut = 4 };

struct TetDofs {
    int edge_index[6];
    double edge_coeff[6];
    int phi_index[10];
    double phi_coeff[10];
    Complex phi_fixed[10];
    int phi_fixed_port[10];  // port whose value phi_fixed holds, or -1
};

struct DofMap {
    int num_nodes = 0, num_p2_nodes = 0;
    int num_a = 0, num_phi = 0, num_ports = 0, num_total = 0;
    std::vector<EdgeDof> edge_state;
    std::vector<int> edge_index;
    std::vector<PhiDof> phi_state;
    std::vector<int> phi_index, phi_port;
    std::vector<Complex> phi_fixed_value;
    std::vector<int> port_index;
    std::vector<uint8_t> port_is_fixed;
    std::vector<Complex> port_value;

    int vertex_p2(int v) const { return v; }
    int edge_p2(int e) const { return num_nodes + e; }
    TetDofs local_dofs(int tet, const Mesh& mesh, const BoundProblem& bound) const;
};

// src/dof_map.cpp:8-119
DofMap build_dof_map(const BoundProblem& bound, const Mesh& mesh);

// ------------------------------------------------------------------ sparsity

struct SparsityPattern {
    int rows = 0, cols = 0;
    bool upper_only = false;
    std::vector<int> row_ptr, col_index;
    size_t nnz() const { return col_index.size(); }
    int find_slot(int row, int col) const;
};

// src/sparsity.cpp:19-136
SparsityPattern build_sparsity(const DofMap& dofs, const BoundProblem& bound, const Mesh& mesh,
                               bool upper_only);

// include/aphi_solver/assembly.hpp:67-71
inline int tet_block_index(int n, bool packed, int row_at, int col_at) {
    if (!packed) return row_at * n + col_at;
    if (col_at < row_at) return -1;
    return row_at * n - row_at * (row_at - 1) / 2 + (col_at - row_at);
}
inline int tet_block_size(int n, bool packed) { return packed ? n * (n + 1) / 2 : n * n; }

// src/assembly.cpp:196-232: per tet, the live DOF count, the position of each local
// DOF among the live ones, and the packed block of value-array slots.
struct ScatterMap {
    bool packed = false;
    std::vector<int> live;         // per tet
    std::vector<int> pos;          // per tet: 6 edge then 10 phi positions (-1 eliminated)
    std::vector<int64_t> offset;   // per tet + 1
    std::vector<int> slot;         // -1 below the diagonal in a packed block
    int slot_of(int tet, int row_at, int col_at) const {
        const int at = tet_block_index(live[static_cast<size_t>(tet)], packed, row_at, col_at);
        return at < 0 ? -1 : slot[static_cast<size_t>(offset[static_cast<size_t>(tet)] + at)];
    }
};
ScatterMap build_scatter_map(const SparsityPattern& pattern, const DofMap& dofs, const Mesh& mesh,
                             const BoundProblem& bound);
// The same from each tet's local DOFs as a caller hands them over (the reference's
// DofMap::local_dofs for tet 0 .. num_tets - 1).
using TetDofsOf = std::function<TetDofs(int)>;
ScatterMap build_scatter_map(const SparsityPattern& pattern, int num_tets, const TetDofsOf& local);

// ------------------------------------------------------------------ assembly plan
//
// The device refill writes every tet's 196 element entries (36 AA, 60 APhi,
// 100 PhiPhi, src/element_matrix.cpp:27-117) to a scratch array, then reduces
// them per matrix slot / rhs entry over a list recorded here in EXACTLY the order
// src/assembly.cpp:255-422 accumulates them (bodies, tets ascending, AA, APhi/PhiA,
// PhiPhi, then port rows). Same terms, same order: a deterministic segmented
// reduction instead of atomics.
enum ContribKind : uint8_t {
    kAA = 0,          // values += (cp*cq) * aa
    kAPhi = 1,        // values += (column * (ce*cn)) * ap
    kPhiA = 2,        // values += ((row*jw) * (cn*ce)) * ap
    kPhiPhi = 3,      // values += ((row*column) * (cp*cq)) * pp
    kRhsAPhi = 4,     // rhs -= ((column*ce) * ap) * fixed
    kRhsPhiPhi = 5,   // rhs -= (((row*column)*cp) * pp) * fixed
    kPortDiag = 6,    // values += 1
    kPortVoltage = 7, // rhs += V / column
    kPortCurrent = 8, // rhs += row * I
};

struct AssemblyPlan {
    int num_slots = 0;               // nnz of the pattern; rhs targets follow at num_slots + i
    int num_rows = 0;
    std::vector<int64_t> target_ptr; // num_slots + num_rows + 1
    std::vector<uint32_t> entry;     // tet * 196 + local entry (unused for port kinds)
    std::vector<uint8_t> kind;       // ContribKind | (sign < 0 ? 0x10 : 0) | (value id << 5)
    std::vector<Complex> fixed_value;// table of prescribed physical values (value id)
    std::vector<int> tet_body;       // per tet: index into bodies
};
AssemblyPlan build_assembly_plan(const SparsityPattern& pattern, const ScatterMap& map,
                                 const DofMap& dofs, const Mesh& mesh, const BoundProblem& bound);
// The same with the tets' local DOFs from `local` (bodies, ports and the DofMap's
// port rows and values as in the overload above).
AssemblyPlan build_assembly_plan(const SparsityPattern& pattern, const ScatterMap& map, int num_tets,
                                 const TetDofsOf& local, const DofMap& dofs, const BoundProblem& bound);

// ------------------------------------------------------------------ ordering

struct SymmetricAdjacency {
    std::vector<int> offset, neighbor;
    int rows() const { return static_cast<int>(offset.size()) - 1; }
    int degree(int i) const { return offset[static_cast<size_t>(i) + 1] - offset[static_cast<size_t>(i)]; }
    static SymmetricAdjacency build(const SparsityPattern& pattern);  // src/ordering.cpp:54-102
};

struct Permutation {
    std::vector<int> perm;   // perm[new] = old
    std::vector<int> iperm;  // iperm[old] = new
    int size() const { return static_cast<int>(perm.size()); }
    bool is_valid() const;
};

Permutation compute_ordering(const SparsityPattern& pattern, Ordering ordering);    // :539-556
SparsityPattern permute_pattern(const SparsityPattern& pattern, const Permutation& p); // :558-608

// ------------------------------------------------------------------ symbolic

// src/symbolic.cpp:58-87 / 89-107 / 126-196 (counts only: the L pattern itself is
// never materialised on this path; the supernodal structure below replaces it).
struct SolverAnalysis {
    Ordering ordering = Ordering::ApproximateMinimumDegree;
    Permutation permutation;
    SparsityPattern permuted;
    std::vector<int> parent;
    std::vector<int> row_count;
    std::vector<int> col_count;
    size_t predicted_nnz = 0;
    int num_roots = 0;
};

std::vector<int> elimination_tree(const SparsityPattern& permuted);
std::vector<int> factor_row_counts(const SparsityPattern& permuted, const std::vector<int>& parent);
SolverAnalysis analyze(const SparsityPattern& pattern, Ordering ordering, size_t max_factor_nnz = 0);

// ------------------------------------------------------------------ supernodes
//
// New (the reference defers supernodes, docs/SOLVER_PLAN.md:331-363). A pure
// function of (parent, col_count) plus the relaxation parameters, so it is
// deterministic by construction.
struct SupernodePlan {
    int n = 0;
    int num_super = 0;
    // Factorization order: the analysis permutation followed by a postorder of its
    // elimination tree (an equivalent ordering with identical fill).
    std::vector<int> perm, iperm;
    std::vector<int> first;        // num_super + 1: columns [first[s], first[s+1])
    std::vector<int> col_to_super; // n
    std::vector<int> sparent;      // supernode parent, -1 for a root
    std::vector<int64_t> row_ptr;  // num_super + 1 into rows
    std::vector<int> rows;         // per supernode: its own columns, then the rest ascending
    std::vector<int> level;        // height above the leaves (leaf = 0)
    int num_levels = 0;
    // children CSR (in ascending supernode order)
    std::vector<int> child_ptr, child;
    // Flat storage layout for the L panels (m x k column-major, complex entries).
    std::vector<int64_t> panel_offset;  // num_super + 1
    // Contribution-block (update matrix, (m-k)^2) workspace: a static offset per
    // supernode from a lifetime-aware first-fit plan over the level schedule.
    std::vector<int64_t> cb_offset;
    int64_t cb_pool_entries = 0;
    // Per supernode: position of each of its rows in the parent's row list
    // (relative indices for extend-add), for rows[k..m).
    std::vector<int> rel;           // aligned with rows (entries for the first k rows unused)
    double flops = 0.0;             // algorithmic real flops of the numeric factorization
    size_t factor_entries = 0;      // sum m*k
    size_t exact_factor_nnz = 0;    // sum of col_count (the reference's nnz(L))
};

struct SupernodeOptions {
    int relax_small = 16;     // always merge a child into its parent when both are this small
    double relax_fill = 0.10; // otherwise allow this fraction of explicit zeros
    int max_cols = 1 << 30;   // cap on a supernode's width
};

SupernodePlan build_supernodes(const SolverAnalysis& analysis, const SupernodeOptions& opt);

// For each stored upper-triangle slot of the ORIGINAL pattern, where its value
// lands in the supernodal storage after permutation: the panel entry (panel
// offset + column-major index), for the front of the supernode owning column
// min(iperm[r], iperm[c]).
std::vector<int64_t> slot_to_panel(const SparsityPattern& pattern, const SolverAnalysis& a,
                                   const SupernodePlan& sp);

// Algorithmic real-flop count of the scalar factorization of the reference
// (8 * sum_j [m_j (m_j + 1) / 2 + 2 m_j], m_j = col_count[j] - 1).
double ldlt_flops(const std::vector<int>& col_count);

// ------------------------------------------------------------------ gauge-free potential
//
// The P1 system of recover_gauge_free_potential (src/gauge_free_potential.cpp:40-177):
// <beta grad(Phi'), grad(lam)> = -<beta E, grad(lam)> over the mesh vertices,
// beta = sigma + j w eps0 eps_r. Integer half only: which vertex is which unknown
// (build_vertex_map, :47-85), the upper-triangle pattern the reference's per-row
// std::map ends up with, and, for every stored slot and every rhs row, the
// per-tet terms that land there in the reference's accumulation order (tets
// ascending, then local vertex i, then local vertex j). The plan assumes
// w != 0, where a tet takes part unless its body has sigma = 0 and eps_r = 0.
struct GaugeFreePlan {
    int unknowns = 0;
    std::vector<int> column;             // per vertex: unknown, or -1 if eliminated
    std::vector<Complex> prescribed;     // per vertex: value when eliminated, else 0
    std::vector<int> port_column;        // per port: unknown, or -1 if the port is fixed
    std::vector<uint8_t> tet_on;         // per tet: 1 if its body takes part
    SparsityPattern pattern;             // upper CSR over the unknowns
    // contributions to target t (slot t < nnz, rhs row t - nnz) are
    // [target_ptr[t], target_ptr[t+1]); entry = tet * 20 + term, term 0..15 the
    // stiffness k_ij (i * 4 + j), 16..19 the source term of vertex i;
    // kind = 0 (slot += k_ij), 1 (rhs -= source_i), 2 | id << 2 (rhs -= k_ij * values[id])
    std::vector<int64_t> target_ptr;
    std::vector<uint32_t> entry;
    std::vector<uint8_t> kind;
    std::vector<Complex> values;         // distinct nonzero prescribed values
};

GaugeFreePlan build_gauge_free_plan(const Mesh& mesh, const BoundProblem& bound, const DofMap& dofs);

// ------------------------------------------------------------------ result files
//
// The reference's version string, as its writers print it (include/aphi_solver/version.hpp).
constexpr const char* kReferenceVersion = "0.0.1-dev";

// potential_at_nodes (src/postprocess.cpp:42-123): Phi at every P2 node (vertices,
// then edge midpoints), with a status letter F / T / C / P / A.
struct NodalPotential {
    int num_vertices = 0, num_edges = 0;
    std::vector<Vec3> position;
    std::vector<Complex> value;
    std::vector<char> status;
};
NodalPotential potential_at_nodes(const Mesh& mesh, const DofMap& dofs, const Complex* x, int n, double omega,
                                  int conditioning);
// write_potential (src/postprocess.cpp:549-606) without run information; returns bytes written
size_t write_potential(const std::string& path, const NodalPotential& p, double omega, int conditioning);
// write_vtk (src/postprocess.cpp:1055-1276) without field data or run information
size_t write_vtk_potential(const std::string& path, const Mesh& mesh, const NodalPotential& p, double omega,
                           int conditioning);
// compute_fields' FieldOutput as flat arrays (complex interleaved re, im; 3 complex
// per tet or node), owned by the caller
struct FieldsView {
    const double *a_tet, *b_tet, *e_tet, *j_tet, *sigma_tet;
    const double *a_node, *b_node, *h_node, *e_node, *j_node;
    const int8_t* on_material_interface;
};
// write_vtk with field data (FieldSet::All), no run information
size_t write_vtk_fields(const std::string& path, const Mesh& mesh, const NodalPotential& p, double omega,
                        int conditioning, const FieldsView& f);

// log-spaced sweep, src/input_file.cpp:19-50
std::vector<double> expand_frequency_sweep(double start, double stop, int points, bool logarithmic);

}  // namespace aphi_amd

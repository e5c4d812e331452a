// Reference-side binding of the MI355X path: the reference host's refill
// (symmetric overload, include/aphi_solver/assembly.hpp:197-199) and
// solve_symmetric (include/aphi_solver/factorization.hpp:206-210) with their
// exact signatures, implemented over the C-ABI of include/aphi_gpu.h. A
// maintainer builds the reference with this file in place of the CPU
// definitions of those two functions (INTEGRATION.md of this repository shows
// the two lines); every caller keeps its code.
//
// Per pattern (content hash of n, row_ptr and col_index, with the ordering) the
// binding keeps one device mesh: refill builds it from the bound objects it is
// given (aphi_problem_create_bound: the assembly plan in the reference's
// accumulation order, the ordering, the supernodal plan) the first time and
// reuses it at every later frequency; solve_symmetric on a matrix of that
// pattern solves on the same mesh, and on any other pattern builds one from the
// pattern alone (aphi_problem_create_csr). Errors come back as the reference's
// exception types and messages; a pivot at or below the floor returns false with
// the report filled, as the reference does.
#include "aphi_gpu_backend.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>

#include "aphi_gpu.h"
#include "aphi_solver/assembly.hpp"
#include "aphi_solver/dof_map.hpp"
#include "aphi_solver/mesh.hpp"
#include "aphi_solver/problem.hpp"
#include "aphi_solver/problem_binding.hpp"
#include "aphi_solver/sparsity.hpp"

namespace aphi_solver {
namespace {

using Complex = std::complex<double>;

void check(int status) {
    if (status == APHI_OK) return;
    const std::string msg = aphi_last_error();
    if (status == APHI_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    if (status == APHI_LOGIC_ERROR) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

void hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device memory owned by the binding (the C-ABI's solves take device pointers).
struct DeviceArray {
    void* p = nullptr;
    size_t bytes = 0;
    void reserve(size_t b) {
        if (b <= bytes) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hip(hipMalloc(&p, b), "hipMalloc");
        bytes = b;
    }
    ~DeviceArray() {
        if (p) (void)hipFree(p);
    }
};

struct Entry {
    aphi_problem* problem = nullptr;
    aphi_mesh* mesh = nullptr;
    bool has_plan = false;      // built from bound objects: refill works
    bool first_solve = true;    // analyze_ms is reported once, like a fresh analysis
    double analyze_ms = 0.0;
    DeviceArray b, x;
};

struct State {
    std::mutex mu;
    gpu::Settings settings;
    aphi_context* ctx = nullptr;
    std::map<uint64_t, Entry> meshes;
    ~State() { clear(); }
    void clear() {
        for (auto& kv : meshes) {
            if (kv.second.mesh) aphi_mesh_destroy(kv.second.mesh);
            if (kv.second.problem) aphi_problem_destroy(kv.second.problem);
        }
        meshes.clear();
        if (ctx) aphi_context_destroy(ctx);
        ctx = nullptr;
    }
};

State& state() {
    static State s;
    return s;
}

// FNV-1a over the pattern's integers and the ordering
uint64_t pattern_key(int n, const std::vector<int>& row_ptr, const std::vector<int>& col_index, int ordering) {
    uint64_t h = 1469598103934665603ull;
    const auto mix = [&](uint64_t v) {
        for (int i = 0; i < 8; ++i) {
            h ^= (v >> (8 * i)) & 0xffu;
            h *= 1099511628211ull;
        }
    };
    mix(static_cast<uint64_t>(n));
    mix(static_cast<uint64_t>(ordering));
    for (int v : row_ptr) mix(static_cast<uint32_t>(v));
    for (int v : col_index) mix(static_cast<uint32_t>(v));
    return h;
}

aphi_context* context(State& st) {
    if (!st.ctx) check(aphi_context_create(st.settings.device, &st.ctx));
    return st.ctx;
}

void attach(State& st, Entry& e, aphi_problem* p, bool has_plan) {
    if (e.mesh) aphi_mesh_destroy(e.mesh);
    if (e.problem) aphi_problem_destroy(e.problem);
    e.mesh = nullptr;
    e.problem = p;
    aphi_mesh_desc d{};
    check(aphi_problem_describe(p, &d));
    check(aphi_mesh_create(context(st), &d, &e.mesh));
    aphi_problem_counts c{};
    check(aphi_problem_counts_get(p, &c));
    e.analyze_ms = c.analyze_ms;
    e.first_solve = true;
    e.has_plan = has_plan;
}

// The bound objects flattened into aphi_bound_desc (arrays kept alive here).
struct Flat {
    std::vector<double> coords, sigma, eps_r, mu_r, edge_coeff, phi_coeff, phi_fixed, amplitude;
    std::vector<int32_t> tets, body_ptr, body_tets, edge_index, phi_index, port_index, port_current, live, pos, slot;
    std::vector<int8_t> signs;
    std::vector<int64_t> offset;
    aphi_bound_desc d{};
};

void flatten(const BoundProblem& bound, const Mesh& mesh, const DofMap& dofs, const SparsityPattern& pattern,
             const ScatterMap* scatter, Flat& f) {
    const int nv = static_cast<int>(mesh.nodes.size()), nt = mesh.num_tets();
    for (const Vec3& v : mesh.nodes) {
        f.coords.push_back(v.x);
        f.coords.push_back(v.y);
        f.coords.push_back(v.z);
    }
    for (int t = 0; t < nt; ++t) {
        for (int c = 0; c < 4; ++c) f.tets.push_back(mesh.tets[static_cast<size_t>(t)][static_cast<size_t>(c)]);
        for (int e = 0; e < 6; ++e) f.signs.push_back(mesh.tet_edge_signs[static_cast<size_t>(t)][static_cast<size_t>(e)]);
    }
    f.body_ptr.push_back(0);
    for (const BoundBody& b : bound.bodies) {
        f.sigma.push_back(b.sigma);
        f.eps_r.push_back(b.eps_r);
        f.mu_r.push_back(b.mu_r);
        f.body_tets.insert(f.body_tets.end(), b.tets.begin(), b.tets.end());
        f.body_ptr.push_back(static_cast<int32_t>(f.body_tets.size()));
    }
    for (int t = 0; t < nt; ++t) {
        const TetDofs td = dofs.local_dofs(t, mesh, bound);
        for (int i = 0; i < 6; ++i) {
            f.edge_index.push_back(td.edge[static_cast<size_t>(i)].index);
            f.edge_coeff.push_back(td.edge[static_cast<size_t>(i)].coeff);
        }
        for (int i = 0; i < 10; ++i) {
            f.phi_index.push_back(td.phi[static_cast<size_t>(i)].index);
            f.phi_coeff.push_back(td.phi[static_cast<size_t>(i)].coeff);
            f.phi_fixed.push_back(td.phi_fixed[static_cast<size_t>(i)].real());
            f.phi_fixed.push_back(td.phi_fixed[static_cast<size_t>(i)].imag());
        }
    }
    for (size_t k = 0; k < bound.ports.size(); ++k) {
        const bool current = is_current_driven(bound.ports[k].type);
        const Complex a = current ? bound.ports[k].amplitude : dofs.port_value[k];
        f.port_index.push_back(dofs.port_index[k]);
        f.port_current.push_back(current ? 1 : 0);
        f.amplitude.push_back(a.real());
        f.amplitude.push_back(a.imag());
    }
    if (scatter) {
        f.offset.push_back(0);
        for (int t = 0; t < nt; ++t) {
            const int live = scatter->live_count(t);
            f.live.push_back(live);
            for (int i = 0; i < 6; ++i) f.pos.push_back(scatter->edge_at(t)[i]);
            for (int i = 0; i < 10; ++i) f.pos.push_back(scatter->phi_at(t)[i]);
            const int size = tet_block_size(live, scatter->packed());
            f.slot.insert(f.slot.end(), scatter->slots(t), scatter->slots(t) + size);
            f.offset.push_back(static_cast<int64_t>(f.slot.size()));
        }
    }
    aphi_bound_desc& d = f.d;
    d.num_nodes = nv;
    d.num_tets = nt;
    d.num_bodies = static_cast<int32_t>(bound.bodies.size());
    d.num_total = dofs.num_total;
    d.num_ports = static_cast<int32_t>(bound.ports.size());
    d.coords = f.coords.data();
    d.tet_nodes = f.tets.data();
    d.tet_edge_signs = f.signs.data();
    d.body_sigma = f.sigma.data();
    d.body_eps_r = f.eps_r.data();
    d.body_mu_r = f.mu_r.data();
    d.body_ptr = f.body_ptr.data();
    d.body_tets = f.body_tets.data();
    d.edge_index = f.edge_index.data();
    d.edge_coeff = f.edge_coeff.data();
    d.phi_index = f.phi_index.data();
    d.phi_coeff = f.phi_coeff.data();
    d.phi_fixed = f.phi_fixed.data();
    d.port_index = f.port_index.data();
    d.port_current = f.port_current.data();
    d.port_amplitude = f.amplitude.data();
    d.row_ptr = pattern.row_ptr.data();
    d.col_index = pattern.col_index.data();
    if (scatter) {
        d.scatter_live = f.live.data();
        d.scatter_pos = f.pos.data();
        d.scatter_offset = f.offset.data();
        d.scatter_slot = f.slot.data();
    }
}

// solve_symmetric of nrhs right-hand sides on the pattern's device mesh
bool solve_on_device(const SparseSymmetricZ& a, int nrhs, const Complex* b, Complex* x, SolveReport& report,
                     const SolveOptions& options) {
    State& st = state();
    const int n = a.rows();
    if (n <= 0) throw std::invalid_argument("solve_symmetric: the matrix has no rows");
    const int ordering = static_cast<int>(options.ordering);
    Entry& e = st.meshes[pattern_key(n, a.row_ptr(), a.col_index(), ordering)];
    if (!e.mesh) {
        aphi_problem* p = nullptr;
        check(aphi_problem_create_csr(n, a.row_ptr().data(), a.col_index().data(), ordering, 0, 0, 0.0, &p));
        attach(st, e, p, false);
    }
    check(aphi_system_upload(e.mesh, reinterpret_cast<const double*>(a.values().data()),
                             reinterpret_cast<const double*>(b)));
    const size_t bytes = static_cast<size_t>(n) * static_cast<size_t>(nrhs) * sizeof(Complex);
    e.b.reserve(bytes);
    e.x.reserve(bytes);
    hip(hipMemcpy(e.b.p, b, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    aphi_solve_options opt{};
    opt.precision = st.settings.precision;
    opt.equilibration_iterations = options.equilibration_iterations;
    opt.pivot_floor = options.pivot_floor;
    opt.max_refinements = st.settings.max_refinements;
    opt.tolerance = st.settings.tolerance;
    aphi_solve_report r{};
    const int status = aphi_solve_symmetric(e.mesh, nrhs, e.b.p, n, e.x.p, n, &opt, &r, nullptr);
    report = SolveReport{};
    report.equilibration_min = r.equilibration_min;
    report.equilibration_max = r.equilibration_max;
    report.smallest_pivot = r.smallest_pivot;
    report.largest_pivot = r.largest_pivot;
    report.largest_multiplier = r.largest_multiplier;
    report.factor_nnz = static_cast<std::size_t>(r.factor_nnz);
    report.columns_done = r.columns_done;
    report.analyze_ms = e.first_solve ? e.analyze_ms : 0.0;
    report.factorize_ms = r.factorize_ms;
    e.first_solve = false;
    if (status == APHI_PIVOT_FAILED) return false;  // report filled as far as it got
    check(status);
    report.solve_ms = r.solve_ms;
    report.backward_error = r.backward_error;
    report.residual = r.residual;
    report.ok = r.ok != 0;
    hip(hipMemcpy(x, e.x.p, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    return true;
}

}  // namespace

void refill(SymmetricSystem& system, const BoundProblem& bound, const Mesh& mesh, const DofMap& dofs,
            const SparsityPattern& pattern, double omega, Conditioning formulation, const ScatterMap* scatter) {
    // the reference's refusals, with its messages (src/assembly.cpp:427-491)
    if (system.matrix.rows() != pattern.rows || system.matrix.nnz() != pattern.nnz())
        throw std::invalid_argument(
            "refill: this system was not made from this pattern. Their shapes must agree, or "
            "the slots the scatter writes to mean something else.");
    if (system.rhs.size() != static_cast<std::size_t>(dofs.num_total))
        throw std::invalid_argument("refill: the right-hand side is not one entry per unknown.");
    if (scatter != nullptr && scatter->num_tets() != mesh.num_tets())
        throw std::invalid_argument("refill: the ScatterMap was built for a different mesh (" +
                                    std::to_string(scatter->num_tets()) + " tets against " +
                                    std::to_string(mesh.num_tets()) + ").");
    if (scatter != nullptr && scatter->packed() != pattern.upper_only)
        throw std::invalid_argument(
            "refill: the ScatterMap and the pattern disagree about which half of the matrix is "
            "stored. Their blocks are laid out differently, so every entry would go to the "
            "wrong place.");
    if (!pattern.upper_only)
        throw std::invalid_argument(
            "refill: a SparseSymmetric holds only the upper triangle, so it needs a pattern "
            "built with SparsityStorage::UpperTriangle. A full pattern would reserve slots "
            "below the diagonal that nothing ever writes.");
    const FormulationScales scales = formulation_scales(formulation, omega);
    if (!scales.symmetric)
        throw std::invalid_argument(
            std::string("refill: conditioning = ") + conditioning_keyword(formulation) +
            " does not give a symmetric matrix, so only half of it cannot be stored. Its "
            "(Phi,A) block is j*omega times its (A,Phi) block. Use row_scaled or scaled_phi "
            "for symmetric storage, or a full SparseMatrixZ for this one.");
    State& st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    const int ordering = static_cast<int>(Ordering::ApproximateMinimumDegree);
    Entry& e = st.meshes[pattern_key(pattern.rows, pattern.row_ptr, pattern.col_index, ordering)];
    if (!e.has_plan) {
        Flat f;
        flatten(bound, mesh, dofs, pattern, scatter, f);
        aphi_problem* p = nullptr;
        check(aphi_problem_create_bound(&f.d, ordering, 0, 0.0, &p));
        attach(st, e, p, true);
    }
    check(aphi_refill(e.mesh, omega, static_cast<int>(formulation)));
    check(aphi_system_download(e.mesh, reinterpret_cast<double*>(system.matrix.mutable_values().data()),
                               reinterpret_cast<double*>(system.rhs.data())));
}

bool solve_symmetric(const SparseSymmetricZ& a, const std::vector<Complex>& b, std::vector<Complex>& x,
                     SolveReport& report, const SolveOptions& options, const SolverAnalysis* precomputed) {
    // The device derives its own analysis of the pattern (the same ordering, so
    // the same elimination tree); `precomputed` saves nothing here.
    (void)precomputed;
    if (static_cast<int>(b.size()) != a.rows())
        throw std::invalid_argument("solve_symmetric: b has the wrong length");
    State& st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    x.assign(b.size(), Complex(0.0, 0.0));
    return solve_on_device(a, 1, b.data(), x.data(), report, options);
}

namespace gpu {

void configure(const Settings& s) {
    if (s.precision != 0 && s.precision != 1) throw std::invalid_argument("gpu::configure: precision must be 0 or 1");
    if (s.max_refinements < 0) throw std::invalid_argument("gpu::configure: max_refinements must be >= 0");
    State& st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    if (s.device != st.settings.device) st.clear();
    st.settings = s;
}

Settings settings() {
    State& st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    return st.settings;
}

bool solve_symmetric_many(const SparseSymmetricZ& a, int nrhs, const std::vector<Complex>& B, std::vector<Complex>& X,
                          SolveReport& report, const SolveOptions& options) {
    if (nrhs < 1 || nrhs > 64) throw std::invalid_argument("solve_symmetric_many: nrhs must be 1..64");
    if (B.size() != static_cast<size_t>(a.rows()) * static_cast<size_t>(nrhs))
        throw std::invalid_argument("solve_symmetric_many: B must hold n * nrhs values");
    State& st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    X.assign(B.size(), Complex(0.0, 0.0));
    return solve_on_device(a, nrhs, B.data(), X.data(), report, options);
}

void release() {
    State& st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    st.clear();
}

}  // namespace gpu
}  // namespace aphi_solver

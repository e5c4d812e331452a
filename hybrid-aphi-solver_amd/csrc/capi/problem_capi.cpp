// C-ABI over the host symbolic problem: build a cube configuration end to end
// (mesh -> binding -> DOF map -> pattern -> scatter map -> assembly plan ->
// ordering -> elimination tree -> supernodes) and expose its integer arrays.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <future>
#include <thread>
#include <memory>
#include <cstring>
#include <string>
#include <vector>

#include "../host/aphi_host.hpp"
#include "status.hpp"

using namespace aphi_amd;

struct aphi_problem {
    bool from_csr = false;  // a bare pattern: no mesh, no assembly plan
    bool gauge_free = false;  // the P1 system of the gauge-free potential (plan in gf)
    GaugeFreePlan gf;
    std::vector<double> gf_prescribed, gf_port_value, gf_values, gf_eps_r;
    std::vector<int32_t> gf_column, gf_port_column, gf_tet_on;
    CubeSpec spec;
    Mesh mesh;
    BoundProblem bound;
    DofMap dofs;
    SparsityPattern pattern;
    ScatterMap map;
    AssemblyPlan plan;
    SolverAnalysis analysis;
    SupernodePlan sp;
    std::vector<int64_t> s2p;
    // flat copies handed out by aphi_problem_describe
    std::vector<double> coords;
    std::vector<int32_t> tet_nodes, tet_body, col_ref, super_level;
    std::vector<double> body_sigma, body_eps_r, body_mu_r, values;
    std::vector<int32_t> tet_dofs;
    std::vector<int8_t> tet_signs;
    std::vector<double> tet_phi_coeff, tet_phi_fixed;
    double analyze_ms = 0.0, build_ms = 0.0;
    // a problem made from a caller's bound objects (aphi_problem_create_bound):
    // the tets' local DOFs as handed over, no mesh topology beyond the tets
    bool from_bound = false;
    std::vector<TetDofs> given_dofs;
    // compute_fields' node plan, built on first use: each tet's ten P2 nodes and,
    // per P2 node, its (tet * 10 + local node) entries in tet order
    std::vector<int32_t> tet_p2, field_entry;
    std::vector<int64_t> field_ptr;
};

namespace {
TetDofs tet_dofs_of(const aphi_problem* p, int t) {
    return p->from_bound ? p->given_dofs[static_cast<size_t>(t)] : p->dofs.local_dofs(t, p->mesh, p->bound);
}

// The P2 node of local node k of tet t is tets[t][k] (k < 4) or the midpoint node
// of local edge k - 4 (src/postprocess.cpp:283-287, 405-414); a node's entries
// come out in tet order because the tets are visited in order.
void build_field_plan(aphi_problem* p) {
    if (!p->field_ptr.empty()) return;
    if (p->from_csr || p->from_bound || p->gauge_free)
        throw std::invalid_argument("compute_fields: the problem has no mesh topology");
    const Mesh& m = p->mesh;
    const int nt = m.num_tets(), np = p->dofs.num_p2_nodes;
    p->tet_p2.resize(10 * static_cast<size_t>(nt));
    std::vector<int64_t> ptr(static_cast<size_t>(np) + 1, 0);
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 10; ++k) {
            const int u = k < 4 ? p->dofs.vertex_p2(m.tets[static_cast<size_t>(t)][static_cast<size_t>(k)])
                                : p->dofs.edge_p2(m.tet_edges[static_cast<size_t>(t)][static_cast<size_t>(k - 4)]);
            p->tet_p2[10 * static_cast<size_t>(t) + static_cast<size_t>(k)] = u;
            ++ptr[static_cast<size_t>(u) + 1];
        }
    for (int u = 0; u < np; ++u) ptr[static_cast<size_t>(u) + 1] += ptr[static_cast<size_t>(u)];
    p->field_entry.assign(10 * static_cast<size_t>(nt), 0);
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < 10 * nt; ++e) {
        const int u = p->tet_p2[static_cast<size_t>(e)];
        p->field_entry[static_cast<size_t>(fill[static_cast<size_t>(u)]++)] = e;
    }
    p->field_ptr = std::move(ptr);
}
}  // namespace

namespace aphi_capi {
namespace {
thread_local std::string g_error;
}
void set_error(const std::string& m) { g_error = m; }
void clear_error() { g_error.clear(); }
}  // namespace aphi_capi

extern "C" const char* aphi_last_error(void) { return aphi_capi::g_error.c_str(); }
extern "C" const char* aphi_version(void) { return "aphi-mi355x 0.1"; }

namespace {

// analyze + supernodes + slot map, shared by both ways of making a problem
void finish_symbolic(aphi_problem* p, int ordering, size_t max_factor_nnz, int relax_small, double relax_fill) {
    const auto t1 = std::chrono::steady_clock::now();
    p->analysis = analyze(p->pattern, static_cast<Ordering>(ordering), max_factor_nnz);
    const auto t2 = std::chrono::steady_clock::now();
    SupernodeOptions opt;
    if (relax_small > 0) opt.relax_small = relax_small;
    if (relax_fill > 0.0) opt.relax_fill = relax_fill;
    p->sp = build_supernodes(p->analysis, opt);
    p->s2p = slot_to_panel(p->pattern, p->analysis, p->sp);
    const auto t3 = std::chrono::steady_clock::now();
    p->analyze_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    p->build_ms += std::chrono::duration<double, std::milli>(t3 - t2).count();
    const int n = p->sp.n;
    p->col_ref.resize(static_cast<size_t>(n));
    for (int j = 0; j < n; ++j)
        p->col_ref[static_cast<size_t>(j)] = p->analysis.permutation.iperm[static_cast<size_t>(p->sp.perm[static_cast<size_t>(j)])];
    p->super_level.assign(p->sp.level.begin(), p->sp.level.end());
}

}  // namespace

namespace {
// The flat copies aphi_problem_describe hands out: geometry, bodies, prescribed
// values and every tet's local unknowns.
void fill_flat(aphi_problem* p) {
    const Mesh& m = p->mesh;
    p->coords.resize(static_cast<size_t>(m.num_nodes()) * 3);
    for (int i = 0; i < m.num_nodes(); ++i) {
        p->coords[3 * static_cast<size_t>(i)] = m.nodes[static_cast<size_t>(i)].x;
        p->coords[3 * static_cast<size_t>(i) + 1] = m.nodes[static_cast<size_t>(i)].y;
        p->coords[3 * static_cast<size_t>(i) + 2] = m.nodes[static_cast<size_t>(i)].z;
    }
    p->tet_nodes.resize(static_cast<size_t>(m.num_tets()) * 4);
    for (int t = 0; t < m.num_tets(); ++t)
        for (int c = 0; c < 4; ++c) p->tet_nodes[4 * static_cast<size_t>(t) + c] = m.tets[static_cast<size_t>(t)][static_cast<size_t>(c)];
    p->tet_body.assign(p->plan.tet_body.begin(), p->plan.tet_body.end());
    for (const Body& b : p->bound.bodies) {
        p->body_sigma.push_back(b.sigma);
        p->body_eps_r.push_back(b.eps_r);
        p->body_mu_r.push_back(b.mu_r);
    }
    for (const Complex& v : p->plan.fixed_value) {
        p->values.push_back(v.real());
        p->values.push_back(v.imag());
    }
    {
        const size_t nt = static_cast<size_t>(m.num_tets());
        p->tet_dofs.resize(nt * 16);
        p->tet_signs.resize(nt * 6);
        p->tet_phi_coeff.resize(nt * 10);
        p->tet_phi_fixed.resize(nt * 20);
        for (size_t t = 0; t < nt; ++t) {
            const TetDofs d = tet_dofs_of(p, static_cast<int>(t));
            for (int i = 0; i < 6; ++i) {
                p->tet_dofs[16 * t + i] = d.edge_index[i];
                p->tet_signs[6 * t + i] = m.tet_edge_signs[t][static_cast<size_t>(i)];
            }
            for (int i = 0; i < 10; ++i) {
                p->tet_dofs[16 * t + 6 + i] = d.phi_index[i];
                p->tet_phi_coeff[10 * t + i] = d.phi_coeff[i];
                p->tet_phi_fixed[20 * t + 2 * i] = d.phi_fixed[i].real();
                p->tet_phi_fixed[20 * t + 2 * i + 1] = d.phi_fixed[i].imag();
            }
        }
    }
}
}  // namespace

extern "C" int aphi_problem_create_cube(const aphi_cube_spec* spec, aphi_problem** out) {
    return aphi_capi::guarded([&] {
        if (!spec || !out) throw std::invalid_argument("aphi_problem_create_cube: null argument");
        *out = nullptr;
        auto p = std::make_unique<aphi_problem>();
        const auto t0 = std::chrono::steady_clock::now();
        CubeSpec& s = p->spec;
        s.n = spec->n;
        s.side = spec->side;
        s.bar_i0 = spec->bar_i0;
        s.bar_i1 = spec->bar_i1;
        s.bar_j0 = spec->bar_j0;
        s.bar_j1 = spec->bar_j1;
        s.sigma = spec->sigma;
        s.eps_r = spec->eps_r;
        s.mu_r = spec->mu_r;
        s.air_sigma = spec->air_sigma;
        s.air_eps_r = spec->air_eps_r;
        s.air_mu_r = spec->air_mu_r;
        s.v_top = Complex(spec->v_top_re, spec->v_top_im);
        s.v_bottom = Complex(spec->v_bottom_re, spec->v_bottom_im);
        s.top_current = spec->top_current != 0;
        s.i_top = Complex(spec->i_top_re, spec->i_top_im);
        if (!(s.side > 0.0)) throw std::invalid_argument("aphi_problem_create_cube: side must be positive");
        if (spec->ordering < 0 || spec->ordering > 2) throw std::invalid_argument("aphi_problem_create_cube: unknown ordering");
        p->mesh = make_kuhn_cube(s);
        p->bound = bind_cube(p->mesh, s);
        p->dofs = build_dof_map(p->bound, p->mesh);
        p->pattern = build_sparsity(p->dofs, p->bound, p->mesh, true);
        p->map = build_scatter_map(p->pattern, p->dofs, p->mesh, p->bound);
        // the assembly plan needs only the pattern and the scatter map: it is built
        // on a second thread while the ordering and symbolic analysis run
        auto plan = std::async(std::launch::async, [&] {
            return build_assembly_plan(p->pattern, p->map, p->dofs, p->mesh, p->bound);
        });
        try {
            finish_symbolic(p.get(), spec->ordering, 0, spec->relax_small, spec->relax_fill);
        } catch (...) {
            plan.wait();
            throw;
        }
        p->plan = plan.get();
        p->build_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() -
                       p->analyze_ms - p->build_ms;

        fill_flat(p.get());
        const int n = p->sp.n;
        p->col_ref.resize(static_cast<size_t>(n));
        for (int j = 0; j < n; ++j)
            p->col_ref[static_cast<size_t>(j)] = p->analysis.permutation.iperm[static_cast<size_t>(p->sp.perm[static_cast<size_t>(j)])];
        p->super_level.assign(p->sp.level.begin(), p->sp.level.end());
        *out = p.release();
    });
}

extern "C" int aphi_problem_create_cubes(const aphi_cube_spec* specs, int count, int threads, aphi_problem** out) {
    return aphi_capi::guarded([&] {
        if (count < 0 || (count > 0 && (!specs || !out))) throw std::invalid_argument("aphi_problem_create_cubes: bad arguments");
        for (int i = 0; i < count; ++i) out[i] = nullptr;
        if (count == 0) return;
        int nt = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
        nt = std::max(1, std::min(nt, count));
        // problems are independent and the build keeps no shared state: the
        // workers take the next index until none is left; the first failure (by
        // index) is reported and every problem built is released
        std::atomic<int> next{0};
        std::vector<int> status(static_cast<size_t>(count), APHI_OK);
        std::vector<std::string> message(static_cast<size_t>(count));
        auto work = [&] {
            for (int i = next++; i < count; i = next++) {
                aphi_problem* p = nullptr;
                status[static_cast<size_t>(i)] = aphi_problem_create_cube(&specs[i], &p);
                if (status[static_cast<size_t>(i)] == APHI_OK) out[i] = p;
                else message[static_cast<size_t>(i)] = aphi_last_error();
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        for (int i = 0; i < count; ++i)
            if (status[static_cast<size_t>(i)] != APHI_OK) {
                for (int j = 0; j < count; ++j) {
                    delete out[j];
                    out[j] = nullptr;
                }
                const std::string msg = "problem " + std::to_string(i) + ": " + message[static_cast<size_t>(i)];
                if (status[static_cast<size_t>(i)] == APHI_INVALID_ARGUMENT) throw std::invalid_argument(msg);
                if (status[static_cast<size_t>(i)] == APHI_LOGIC_ERROR) throw std::logic_error(msg);
                throw std::runtime_error(msg);
            }
    });
}

extern "C" int aphi_problem_create_csr(int32_t n, const int32_t* row_ptr, const int32_t* col_index, int ordering,
                                       int64_t max_factor_nnz, int relax_small, double relax_fill,
                                       aphi_problem** out) {
    return aphi_capi::guarded([&] {
        if (!out || !row_ptr || (n > 0 && !col_index && row_ptr[n] > 0))
            throw std::invalid_argument("aphi_problem_create_csr: null argument");
        *out = nullptr;
        // the checks of SparseSymmetric::from_pattern (include/aphi_solver/sparse_symmetric.hpp:38-83)
        if (n < 0) throw std::invalid_argument("SparseSymmetric::from_pattern: negative rows");
        if (n == 0) throw std::invalid_argument("aphi_problem_create_csr: the matrix has no rows");
        if (row_ptr[0] != 0) throw std::invalid_argument("SparseSymmetric::from_pattern: row_ptr must run from 0 to col_index.size()");
        for (int r = 0; r < n; ++r) {
            const int begin = row_ptr[r], end = row_ptr[r + 1];
            if (end < begin) throw std::invalid_argument("SparseSymmetric::from_pattern: row_ptr is not non-decreasing");
            for (int k = begin; k < end; ++k) {
                const int c = col_index[k];
                if (c < 0 || c >= n) throw std::invalid_argument("SparseSymmetric::from_pattern: a column index is out of range");
                if (c < r)
                    throw std::invalid_argument("SparseSymmetric::from_pattern: entry (" + std::to_string(r) + ", " +
                                                std::to_string(c) +
                                                ") is below the diagonal. Only the upper triangle is stored; a "
                                                "lower-triangle entry would never be read, since matvec mirrors.");
                if (k > begin && c <= col_index[k - 1])
                    throw std::invalid_argument("SparseSymmetric::from_pattern: each row's columns must be strictly ascending");
            }
        }
        if (ordering < 0 || ordering > 2) throw std::invalid_argument("aphi_problem_create_csr: unknown ordering");
        if (max_factor_nnz < 0) throw std::invalid_argument("aphi_problem_create_csr: negative factor budget");
        auto p = std::make_unique<aphi_problem>();
        p->from_csr = true;
        p->pattern.rows = p->pattern.cols = n;
        p->pattern.upper_only = true;
        p->pattern.row_ptr.assign(row_ptr, row_ptr + n + 1);
        p->pattern.col_index.assign(col_index, col_index + row_ptr[n]);
        p->dofs.num_total = n;
        finish_symbolic(p.get(), ordering, static_cast<size_t>(max_factor_nnz), relax_small, relax_fill);
        *out = p.release();
    });
}

extern "C" int aphi_problem_create_bound(const aphi_bound_desc* d, int ordering, int relax_small, double relax_fill,
                                         aphi_problem** out) {
    return aphi_capi::guarded([&] {
        if (!d || !out) throw std::invalid_argument("aphi_problem_create_bound: null argument");
        *out = nullptr;
        const int nv = d->num_nodes, nt = d->num_tets, nb = d->num_bodies, n = d->num_total, np = d->num_ports;
        if (nv <= 0 || nt <= 0 || nb <= 0 || n <= 0 || np < 0)
            throw std::invalid_argument("aphi_problem_create_bound: empty mesh, bodies or unknowns");
        if (!d->coords || !d->tet_nodes || !d->tet_edge_signs || !d->body_sigma || !d->body_eps_r || !d->body_mu_r ||
            !d->body_ptr || !d->body_tets || !d->edge_index || !d->edge_coeff || !d->phi_index || !d->phi_coeff ||
            !d->phi_fixed || !d->row_ptr || !d->col_index || (np > 0 && (!d->port_index || !d->port_current ||
                                                                         !d->port_amplitude)))
            throw std::invalid_argument("aphi_problem_create_bound: null array");
        const bool given_map = d->scatter_live != nullptr;
        if (given_map != (d->scatter_pos != nullptr) || given_map != (d->scatter_offset != nullptr) ||
            given_map != (d->scatter_slot != nullptr))
            throw std::invalid_argument("aphi_problem_create_bound: pass all four ScatterMap arrays or none");
        if (ordering < 0 || ordering > 2) throw std::invalid_argument("aphi_problem_create_bound: unknown ordering");
        auto p = std::make_unique<aphi_problem>();
        p->from_bound = true;
        const auto t0 = std::chrono::steady_clock::now();
        // the pattern, with the checks of SparseSymmetric::from_pattern
        if (d->row_ptr[0] != 0) throw std::invalid_argument("SparseSymmetric::from_pattern: row_ptr must run from 0 to col_index.size()");
        for (int r = 0; r < n; ++r) {
            if (d->row_ptr[r + 1] < d->row_ptr[r])
                throw std::invalid_argument("SparseSymmetric::from_pattern: row_ptr is not non-decreasing");
            for (int k = d->row_ptr[r]; k < d->row_ptr[r + 1]; ++k) {
                const int c = d->col_index[k];
                if (c < r || c >= n)
                    throw std::invalid_argument("aphi_problem_create_bound: the pattern is not an upper-triangle CSR of "
                                                "num_total rows");
                if (k > d->row_ptr[r] && c <= d->col_index[k - 1])
                    throw std::invalid_argument("SparseSymmetric::from_pattern: each row's columns must be strictly ascending");
            }
        }
        p->pattern.rows = p->pattern.cols = n;
        p->pattern.upper_only = true;
        p->pattern.row_ptr.assign(d->row_ptr, d->row_ptr + n + 1);
        p->pattern.col_index.assign(d->col_index, d->col_index + d->row_ptr[n]);
        // the mesh: nodes, tets and their edge signs
        Mesh& m = p->mesh;
        m.nodes.resize(static_cast<size_t>(nv));
        for (int i = 0; i < nv; ++i) m.nodes[static_cast<size_t>(i)] = Vec3{d->coords[3 * i], d->coords[3 * i + 1], d->coords[3 * i + 2]};
        m.tets.resize(static_cast<size_t>(nt));
        m.tet_edge_signs.resize(static_cast<size_t>(nt));
        for (int t = 0; t < nt; ++t) {
            for (int c = 0; c < 4; ++c) {
                const int v = d->tet_nodes[4 * static_cast<int64_t>(t) + c];
                if (v < 0 || v >= nv) throw std::invalid_argument("aphi_problem_create_bound: a tet names a missing node");
                m.tets[static_cast<size_t>(t)][static_cast<size_t>(c)] = v;
            }
            for (int e = 0; e < 6; ++e) m.tet_edge_signs[static_cast<size_t>(t)][static_cast<size_t>(e)] = d->tet_edge_signs[6 * static_cast<int64_t>(t) + e];
        }
        // the bodies partition the tets (each body's tets ascending)
        if (d->body_ptr[0] != 0 || d->body_ptr[nb] != nt)
            throw std::invalid_argument("aphi_problem_create_bound: the bodies must cover every tet exactly once");
        std::vector<uint8_t> seen(static_cast<size_t>(nt), 0);
        for (int b = 0; b < nb; ++b) {
            Body body;
            body.sigma = d->body_sigma[b];
            body.eps_r = d->body_eps_r[b];
            body.mu_r = d->body_mu_r[b];
            if (d->body_ptr[b + 1] < d->body_ptr[b]) throw std::invalid_argument("aphi_problem_create_bound: body_ptr decreases");
            for (int q = d->body_ptr[b]; q < d->body_ptr[b + 1]; ++q) {
                const int t = d->body_tets[q];
                if (t < 0 || t >= nt || seen[static_cast<size_t>(t)] || (q > d->body_ptr[b] && t <= d->body_tets[q - 1]))
                    throw std::invalid_argument("aphi_problem_create_bound: the bodies must cover every tet exactly once, "
                                                "each body's tets ascending");
                seen[static_cast<size_t>(t)] = 1;
                body.tets.push_back(t);
            }
            p->bound.bodies.push_back(std::move(body));
        }
        // every tet's local unknowns (DofMap::local_dofs)
        p->given_dofs.resize(static_cast<size_t>(nt));
        for (int t = 0; t < nt; ++t) {
            TetDofs& o = p->given_dofs[static_cast<size_t>(t)];
            for (int i = 0; i < 6; ++i) {
                o.edge_index[i] = d->edge_index[6 * static_cast<int64_t>(t) + i];
                o.edge_coeff[i] = d->edge_coeff[6 * static_cast<int64_t>(t) + i];
                if (o.edge_index[i] < -1 || o.edge_index[i] >= n)
                    throw std::invalid_argument("aphi_problem_create_bound: an edge unknown is out of range");
            }
            for (int i = 0; i < 10; ++i) {
                o.phi_index[i] = d->phi_index[10 * static_cast<int64_t>(t) + i];
                o.phi_coeff[i] = d->phi_coeff[10 * static_cast<int64_t>(t) + i];
                o.phi_fixed[i] = Complex(d->phi_fixed[20 * static_cast<int64_t>(t) + 2 * i],
                                         d->phi_fixed[20 * static_cast<int64_t>(t) + 2 * i + 1]);
                o.phi_fixed_port[i] = -1;
                if (o.phi_index[i] < -1 || o.phi_index[i] >= n)
                    throw std::invalid_argument("aphi_problem_create_bound: a Phi unknown is out of range");
            }
        }
        // ports and the DofMap fields the scatter reads
        p->dofs.num_total = n;
        p->dofs.num_ports = np;
        for (int k = 0; k < np; ++k) {
            Port port;
            const bool current = d->port_current[k] != 0;
            port.type = current ? PortType::BoundaryCurrent : PortType::BoundaryVoltage;
            const Complex a(d->port_amplitude[2 * k], d->port_amplitude[2 * k + 1]);
            port.amplitude = a;
            p->bound.ports.push_back(port);
            if (d->port_index[k] < 0 || d->port_index[k] >= n)
                throw std::invalid_argument("aphi_problem_create_bound: a port row is out of range");
            p->dofs.port_index.push_back(d->port_index[k]);
            p->dofs.port_is_fixed.push_back(current ? 0 : 1);
            p->dofs.port_value.push_back(current ? Complex(0.0, 0.0) : a);
        }
        const TetDofsOf local = [&](int t) { return p->given_dofs[static_cast<size_t>(t)]; };
        if (given_map) {
            ScatterMap& sm = p->map;
            sm.packed = true;
            sm.live.assign(d->scatter_live, d->scatter_live + nt);
            sm.pos.assign(d->scatter_pos, d->scatter_pos + 16 * static_cast<int64_t>(nt));
            sm.offset.assign(d->scatter_offset, d->scatter_offset + nt + 1);
            if (sm.offset[0] != 0) throw std::invalid_argument("aphi_problem_create_bound: scatter offsets must start at 0");
            for (int t = 0; t < nt; ++t) {
                const int live = sm.live[static_cast<size_t>(t)];
                if (live < 0 || live > 16 ||
                    sm.offset[static_cast<size_t>(t) + 1] - sm.offset[static_cast<size_t>(t)] != tet_block_size(live, true))
                    throw std::invalid_argument("aphi_problem_create_bound: the ScatterMap blocks do not match their "
                                                "live counts (built for a different pattern?)");
                for (int i = 0; i < 16; ++i) {
                    const int q = sm.pos[16 * static_cast<size_t>(t) + i];
                    if (q < -1 || q >= live) throw std::invalid_argument("aphi_problem_create_bound: a scatter position is out of range");
                }
            }
            sm.slot.assign(d->scatter_slot, d->scatter_slot + sm.offset[static_cast<size_t>(nt)]);
            for (int v : sm.slot)
                if (v < -1 || v >= static_cast<int>(p->pattern.nnz()))
                    throw std::invalid_argument("aphi_problem_create_bound: a scatter slot lies outside the pattern");
        } else {
            p->map = build_scatter_map(p->pattern, nt, local);
        }
        p->plan = build_assembly_plan(p->pattern, p->map, nt, local, p->dofs, p->bound);
        finish_symbolic(p.get(), ordering, 0, relax_small, relax_fill);
        fill_flat(p.get());
        p->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() -
                      p->analyze_ms;
        *out = p.release();
    });
}

extern "C" int aphi_problem_create_gauge_free(const aphi_problem* cube, int ordering, aphi_problem** out) {
    return aphi_capi::guarded([&] {
        if (!cube || !out) throw std::invalid_argument("aphi_problem_create_gauge_free: null argument");
        if (cube->from_csr || cube->from_bound)
            throw std::invalid_argument("aphi_problem_create_gauge_free: the problem has no mesh topology");
        if (ordering < 0 || ordering > 2) throw std::invalid_argument("aphi_problem_create_gauge_free: unknown ordering");
        *out = nullptr;
        auto p = std::make_unique<aphi_problem>();
        p->from_csr = true;
        p->gauge_free = true;
        const auto t0 = std::chrono::steady_clock::now();
        p->gf = build_gauge_free_plan(cube->mesh, cube->bound, cube->dofs);
        if (p->gf.unknowns == 0) throw std::invalid_argument("aphi_problem_create_gauge_free: the potential has no unknowns");
        p->pattern = p->gf.pattern;
        p->dofs.num_total = p->gf.unknowns;
        // geometry and materials for the device terms
        p->coords = cube->coords;
        p->tet_nodes = cube->tet_nodes;
        p->tet_body = cube->tet_body;
        p->body_sigma = cube->body_sigma;
        p->gf_eps_r = cube->body_eps_r;
        const GaugeFreePlan& g = p->gf;
        p->gf_column.assign(g.column.begin(), g.column.end());
        p->gf_port_column.assign(g.port_column.begin(), g.port_column.end());
        p->gf_tet_on.assign(g.tet_on.begin(), g.tet_on.end());
        for (const Complex& v : g.prescribed) { p->gf_prescribed.push_back(v.real()); p->gf_prescribed.push_back(v.imag()); }
        for (const Complex& v : g.values) { p->gf_values.push_back(v.real()); p->gf_values.push_back(v.imag()); }
        for (const Complex& v : cube->dofs.port_value) { p->gf_port_value.push_back(v.real()); p->gf_port_value.push_back(v.imag()); }
        p->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        finish_symbolic(p.get(), ordering, 0, 0, 0.0);
        *out = p.release();
    });
}

extern "C" int aphi_problem_describe_gauge_free(const aphi_problem* p, aphi_gauge_free_desc* d) {
    return aphi_capi::guarded([&] {
        if (!p || !d) throw std::invalid_argument("aphi_problem_describe_gauge_free: null argument");
        if (!p->gauge_free) throw std::invalid_argument("aphi_problem_describe_gauge_free: not a gauge-free problem");
        const GaugeFreePlan& g = p->gf;
        *d = aphi_gauge_free_desc{};
        d->num_nodes = static_cast<int32_t>(g.column.size());
        d->num_tets = static_cast<int32_t>(g.tet_on.size());
        d->num_bodies = static_cast<int32_t>(p->body_sigma.size());
        d->unknowns = g.unknowns;
        d->nnz = static_cast<int64_t>(g.pattern.nnz());
        d->coords = p->coords.data();
        d->tet_nodes = p->tet_nodes.data();
        d->tet_body = p->tet_body.data();
        d->body_sigma = p->body_sigma.data();
        d->body_eps_r = p->gf_eps_r.data();
        d->tet_on = p->gf_tet_on.data();
        d->column = p->gf_column.data();
        d->prescribed = p->gf_prescribed.data();
        d->num_ports = static_cast<int32_t>(g.port_column.size());
        d->port_column = p->gf_port_column.data();
        d->port_value = p->gf_port_value.data();
        d->target_ptr = g.target_ptr.data();
        d->entry = g.entry.data();
        d->kind = g.kind.data();
        d->num_values = static_cast<int32_t>(g.values.size());
        d->values = p->gf_values.data();
    });
}

extern "C" int aphi_write_potential(const aphi_problem* p, const char* path, const double* x, double omega,
                                    int conditioning, int format, int64_t* bytes) {
    return aphi_capi::guarded([&] {
        if (!p || !path || !x) throw std::invalid_argument("aphi_write_potential: null argument");
        if (p->from_csr || p->from_bound)
            throw std::invalid_argument("aphi_write_potential: the problem has no mesh topology");
        if (conditioning < 0 || conditioning > 2) throw std::invalid_argument("aphi_write_potential: unknown conditioning");
        if (format != 0 && format != 1) throw std::invalid_argument("aphi_write_potential: format must be 0 (text) or 1 (VTK)");
        const NodalPotential np = potential_at_nodes(p->mesh, p->dofs, reinterpret_cast<const Complex*>(x),
                                                     p->dofs.num_total, omega, conditioning);
        const size_t n = format == 0 ? write_potential(path, np, omega, conditioning)
                                     : write_vtk_potential(path, p->mesh, np, omega, conditioning);
        if (bytes) *bytes = static_cast<int64_t>(n);
    });
}

extern "C" int aphi_problem_describe_fields(const aphi_problem* p, aphi_fields_desc* out) {
    return aphi_capi::guarded([&] {
        if (!p || !out) throw std::invalid_argument("aphi_problem_describe_fields: null argument");
        aphi_problem* q = const_cast<aphi_problem*>(p);  // the plan is a cache built on first use
        build_field_plan(q);
        out->num_tets = p->mesh.num_tets();
        out->num_p2 = p->dofs.num_p2_nodes;
        out->tet_p2 = q->tet_p2.data();
        out->node_ptr = q->field_ptr.data();
        out->node_entry = q->field_entry.data();
    });
}

extern "C" int aphi_write_vtk_fields(const aphi_problem* p, const char* path, const double* x, double omega,
                                     int conditioning, const aphi_fields* fields, int64_t* bytes) {
    return aphi_capi::guarded([&] {
        if (!p || !path || !x || !fields) throw std::invalid_argument("aphi_write_vtk_fields: null argument");
        if (p->from_csr || p->from_bound || p->gauge_free)
            throw std::invalid_argument("aphi_write_vtk_fields: the problem has no mesh topology");
        if (conditioning < 0 || conditioning > 2) throw std::invalid_argument("aphi_write_vtk_fields: unknown conditioning");
        const double* arrays[] = {fields->a_tet, fields->b_tet, fields->e_tet, fields->j_tet, fields->sigma_tet,
                                  fields->a_node, fields->b_node, fields->h_node, fields->e_node, fields->j_node};
        for (const double* a : arrays)
            if (!a) throw std::invalid_argument("aphi_write_vtk_fields: a field array is null");
        if (!fields->on_material_interface) throw std::invalid_argument("aphi_write_vtk_fields: a field array is null");
        const NodalPotential np = potential_at_nodes(p->mesh, p->dofs, reinterpret_cast<const Complex*>(x),
                                                     p->dofs.num_total, omega, conditioning);
        FieldsView f{fields->a_tet,  fields->b_tet,  fields->e_tet,  fields->j_tet,  fields->sigma_tet,
                     fields->a_node, fields->b_node, fields->h_node, fields->e_node, fields->j_node,
                     fields->on_material_interface};
        const size_t n = write_vtk_fields(path, p->mesh, np, omega, conditioning, f);
        if (bytes) *bytes = static_cast<int64_t>(n);
    });
}

extern "C" void aphi_problem_destroy(aphi_problem* p) { delete p; }

extern "C" int aphi_problem_counts_get(const aphi_problem* p, aphi_problem_counts* out) {
    return aphi_capi::guarded([&] {
        if (!p || !out) throw std::invalid_argument("aphi_problem_counts_get: null argument");
        aphi_problem_counts c{};
        c.nodes = p->mesh.num_nodes();
        c.edges = p->mesh.num_edges();
        c.faces = p->mesh.num_faces();
        c.tets = p->mesh.num_tets();
        c.num_a = p->dofs.num_a;
        c.num_phi = p->dofs.num_phi;
        c.num_ports = p->dofs.num_ports;
        c.num_total = p->dofs.num_total;
        c.nnz = static_cast<int64_t>(p->pattern.nnz());
        c.predicted_nnz = static_cast<int64_t>(p->analysis.predicted_nnz);
        c.num_roots = p->analysis.num_roots;
        for (EdgeDof e : p->dofs.edge_state) {
            if (e == EdgeDof::Tree) ++c.tree_edges;
            else if (e == EdgeDof::Dirichlet) ++c.dirichlet_edges;
            else ++c.free_edges;
        }
        c.num_super = p->sp.num_super;
        c.num_levels = p->sp.num_levels;
        c.factor_entries = static_cast<int64_t>(p->sp.factor_entries);
        c.cb_pool_entries = p->sp.cb_pool_entries;
        c.contributions = static_cast<int64_t>(p->plan.entry.size());
        c.flops = p->sp.flops;
        c.analyze_ms = p->analyze_ms;
        c.build_ms = p->build_ms;
        *out = c;
    });
}

namespace {

template <typename V>
void copy_out(const V& v, int64_t* out, int64_t capacity, int64_t* count) {
    *count = static_cast<int64_t>(v.size());
    const int64_t n = std::min<int64_t>(capacity, *count);
    for (int64_t i = 0; i < n; ++i) out[i] = static_cast<int64_t>(v[static_cast<size_t>(i)]);
}

}  // namespace

extern "C" int aphi_problem_get_int(const aphi_problem* p, const char* name, int64_t* out, int64_t capacity,
                                    int64_t* count) {
    return aphi_capi::guarded([&] {
        if (!p || !name || !count || (capacity > 0 && !out)) throw std::invalid_argument("aphi_problem_get_int: null argument");
        const std::string k(name);
        const Mesh& m = p->mesh;
        std::vector<int64_t> tmp;
        if (k == "tet_nodes") return copy_out(p->tet_nodes, out, capacity, count);
        if (k == "edges") {
            for (const auto& e : m.edges) { tmp.push_back(e.first); tmp.push_back(e.second); }
        } else if (k == "tet_edges") {
            for (const auto& te : m.tet_edges) for (int e : te) tmp.push_back(e);
        } else if (k == "tet_edge_signs") {
            for (const auto& ts : m.tet_edge_signs) for (int8_t s : ts) tmp.push_back(s);
        } else if (k == "tet_body") {
            return copy_out(p->tet_body, out, capacity, count);
        } else if (k == "dirichlet_edge") {
            return copy_out(p->bound.dirichlet_edge, out, capacity, count);
        } else if (k == "tree_edge") {
            return copy_out(p->bound.gauge.is_tree_edge, out, capacity, count);
        } else if (k == "edge_index") {
            return copy_out(p->dofs.edge_index, out, capacity, count);
        } else if (k == "phi_index") {
            return copy_out(p->dofs.phi_index, out, capacity, count);
        } else if (k == "port_index") {
            return copy_out(p->dofs.port_index, out, capacity, count);
        } else if (k == "row_ptr") {
            return copy_out(p->pattern.row_ptr, out, capacity, count);
        } else if (k == "col_index") {
            return copy_out(p->pattern.col_index, out, capacity, count);
        } else if (k == "perm") {
            return copy_out(p->analysis.permutation.perm, out, capacity, count);
        } else if (k == "iperm") {
            return copy_out(p->analysis.permutation.iperm, out, capacity, count);
        } else if (k == "parent") {
            return copy_out(p->analysis.parent, out, capacity, count);
        } else if (k == "row_count") {
            return copy_out(p->analysis.row_count, out, capacity, count);
        } else if (k == "col_count") {
            return copy_out(p->analysis.col_count, out, capacity, count);
        } else if (k == "scatter_live") {
            return copy_out(p->map.live, out, capacity, count);
        } else if (k == "scatter_pos") {
            return copy_out(p->map.pos, out, capacity, count);
        } else if (k == "scatter_offset") {
            return copy_out(p->map.offset, out, capacity, count);
        } else if (k == "scatter_slot") {
            return copy_out(p->map.slot, out, capacity, count);
        } else if (k == "super_first") {
            return copy_out(p->sp.first, out, capacity, count);
        } else if (k == "super_parent") {
            return copy_out(p->sp.sparent, out, capacity, count);
        } else if (k == "super_row_ptr") {
            return copy_out(p->sp.row_ptr, out, capacity, count);
        } else if (k == "super_rows") {
            return copy_out(p->sp.rows, out, capacity, count);
        } else if (k == "super_level") {
            return copy_out(p->sp.level, out, capacity, count);
        } else if (k == "factor_perm") {
            return copy_out(p->sp.perm, out, capacity, count);
        } else if (k == "tet_dofs") {
            for (int t = 0; t < m.num_tets(); ++t) {
                const TetDofs d = tet_dofs_of(p, t);
                for (int i = 0; i < 6; ++i) tmp.push_back(d.edge_index[i]);
                for (int i = 0; i < 10; ++i) tmp.push_back(d.phi_index[i]);
            }
        } else if (k == "body_ptr") {
            tmp.push_back(0);
            for (const Body& b : p->bound.bodies) tmp.push_back(tmp.back() + static_cast<int64_t>(b.tets.size()));
        } else if (k == "body_tets") {
            for (const Body& b : p->bound.bodies) tmp.insert(tmp.end(), b.tets.begin(), b.tets.end());
        } else if (k == "col_ref") {
            return copy_out(p->col_ref, out, capacity, count);
        } else if (k == "super_rel") {
            return copy_out(p->sp.rel, out, capacity, count);
        } else if (k == "slot_to_panel") {
            return copy_out(p->s2p, out, capacity, count);
        } else if (k == "phi_state") {
            for (PhiDof st : p->dofs.phi_state) tmp.push_back(static_cast<int64_t>(st));
        } else if (k == "phi_port") {
            return copy_out(p->dofs.phi_port, out, capacity, count);
        } else if (k == "port_is_fixed") {
            return copy_out(p->dofs.port_is_fixed, out, capacity, count);
        } else if (k == "gf_column") {
            return copy_out(p->gf.column, out, capacity, count);
        } else if (k == "gf_port_column") {
            return copy_out(p->gf.port_column, out, capacity, count);
        } else if (k == "gf_tet_on") {
            return copy_out(p->gf.tet_on, out, capacity, count);
        } else if (k == "gf_target_ptr") {
            return copy_out(p->gf.target_ptr, out, capacity, count);
        } else if (k == "gf_entry") {
            return copy_out(p->gf.entry, out, capacity, count);
        } else if (k == "gf_kind") {
            return copy_out(p->gf.kind, out, capacity, count);
        } else if (k == "tet_p2") {
            build_field_plan(const_cast<aphi_problem*>(p));
            return copy_out(p->tet_p2, out, capacity, count);
        } else if (k == "field_node_ptr") {
            build_field_plan(const_cast<aphi_problem*>(p));
            return copy_out(p->field_ptr, out, capacity, count);
        } else if (k == "field_node_entry") {
            build_field_plan(const_cast<aphi_problem*>(p));
            return copy_out(p->field_entry, out, capacity, count);
        } else {
            throw std::invalid_argument("aphi_problem_get_int: unknown array '" + k + "'");
        }
        copy_out(tmp, out, capacity, count);
    });
}

extern "C" int aphi_problem_get_double(const aphi_problem* p, const char* name, double* out, int64_t capacity,
                                       int64_t* count) {
    return aphi_capi::guarded([&] {
        if (!p || !name || !count || (capacity > 0 && !out)) throw std::invalid_argument("aphi_problem_get_double: null argument");
        const std::string k(name);
        std::vector<double> tmp;
        const std::vector<double>* src = &tmp;
        if (k == "coords") {
            src = &p->coords;
        } else if (k == "tet_phi_coeff") {
            src = &p->tet_phi_coeff;
        } else if (k == "tet_phi_fixed") {
            for (int t = 0; t < p->mesh.num_tets(); ++t) {
                const TetDofs d = tet_dofs_of(p, t);
                for (int i = 0; i < 10; ++i) { tmp.push_back(d.phi_fixed[i].real()); tmp.push_back(d.phi_fixed[i].imag()); }
            }
        } else if (k == "body_param") {
            for (const Body& b : p->bound.bodies) { tmp.push_back(b.sigma); tmp.push_back(b.eps_r); tmp.push_back(b.mu_r); }
        } else if (k == "port_value") {
            for (const Complex& v : p->dofs.port_value) { tmp.push_back(v.real()); tmp.push_back(v.imag()); }
        } else if (k == "port_amplitude") {
            for (const Port& q : p->bound.ports) { tmp.push_back(q.amplitude.real()); tmp.push_back(q.amplitude.imag()); }
        } else if (k == "tet_edge_coeff") {
            for (int t = 0; t < p->mesh.num_tets(); ++t) {
                const TetDofs d = tet_dofs_of(p, t);
                for (int i = 0; i < 6; ++i) tmp.push_back(d.edge_coeff[i]);
            }
        } else if (k == "gf_prescribed") {
            src = &p->gf_prescribed;
        } else if (k == "gf_values") {
            src = &p->gf_values;
        } else {
            throw std::invalid_argument("aphi_problem_get_double: unknown array '" + k + "'");
        }
        *count = static_cast<int64_t>(src->size());
        const int64_t n = std::min<int64_t>(capacity, *count);
        if (n > 0) std::memcpy(out, src->data(), static_cast<size_t>(n) * sizeof(double));
    });
}

extern "C" int aphi_frequency_sweep(double start, double stop, int points, int logarithmic, double* out) {
    return aphi_capi::guarded([&] {
        if (!out) throw std::invalid_argument("aphi_frequency_sweep: null output");
        const std::vector<double> f = expand_frequency_sweep(start, stop, points, logarithmic != 0);
        std::memcpy(out, f.data(), f.size() * sizeof(double));
    });
}

extern "C" int aphi_problem_describe(const aphi_problem* p, aphi_mesh_desc* d) {
    return aphi_capi::guarded([&] {
        if (!p || !d) throw std::invalid_argument("aphi_problem_describe: null argument");
        *d = aphi_mesh_desc{};
        d->num_nodes = p->mesh.num_nodes();
        d->num_tets = p->mesh.num_tets();
        d->num_bodies = static_cast<int32_t>(p->bound.bodies.size());
        d->n = p->dofs.num_total;
        d->nnz = static_cast<int64_t>(p->pattern.nnz());
        d->coords = p->coords.data();
        d->tet_nodes = p->tet_nodes.data();
        d->tet_body = p->tet_body.data();
        d->body_sigma = p->body_sigma.data();
        d->body_eps_r = p->body_eps_r.data();
        d->body_mu_r = p->body_mu_r.data();
        // a problem made from a bare pattern has no mesh and no assembly plan
        d->target_ptr = p->from_csr ? nullptr : p->plan.target_ptr.data();
        d->entry = p->from_csr ? nullptr : p->plan.entry.data();
        d->kind = p->from_csr ? nullptr : p->plan.kind.data();
        d->num_values = static_cast<int32_t>(p->plan.fixed_value.size());
        d->values = p->values.data();
        d->row_ptr = p->pattern.row_ptr.data();
        d->col_index = p->pattern.col_index.data();
        d->factor_perm = p->sp.perm.data();
        d->col_ref = p->col_ref.data();
        d->num_super = p->sp.num_super;
        d->num_levels = p->sp.num_levels;
        d->super_first = p->sp.first.data();
        d->super_parent = p->sp.sparent.data();
        d->super_row_ptr = p->sp.row_ptr.data();
        d->super_rows = p->sp.rows.data();
        d->super_rel = p->sp.rel.data();
        d->super_level = p->super_level.data();
        d->panel_offset = p->sp.panel_offset.data();
        d->cb_offset = p->sp.cb_offset.data();
        d->cb_pool_entries = p->sp.cb_pool_entries;
        d->slot_to_panel = p->s2p.data();
        d->flops = p->sp.flops;
        d->predicted_nnz = static_cast<int64_t>(p->analysis.predicted_nnz);
        if (!p->from_csr) {
            d->tet_dofs = p->tet_dofs.data();
            d->tet_edge_signs = p->tet_signs.data();
            d->tet_phi_coeff = p->tet_phi_coeff.data();
            d->tet_phi_fixed = p->tet_phi_fixed.data();
        }
    });
}

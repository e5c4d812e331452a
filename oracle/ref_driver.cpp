// Golden-vector driver over the REFERENCE implementation (checker only).
//
// Built by oracle/Makefile.ref against the reference's own sources where they
// lie, into oracle/_ref/. It constructs the Kuhn-cube configuration through the
// reference's public API (Mesh, Problem, bind_to_mesh, build_dof_map,
// build_sparsity, ScatterMap, analyze, refill, solve_symmetric, factorize_ldlt,
// compute_fields, terminal_current) and dumps what that produced, so the
// repository's host restatement and device path can be compared with it.
//
// usage: ref_driver OUTDIR n side bar_i0 bar_i1 bar_j0 bar_j1 sigma freq_hz mode
//   mode 0: symbolic only (topology, gauge, DOF map, pattern, scatter map, ordering)
//   mode 2: + refill at freq_hz and one solve_symmetric, with I, R and L (large meshes)
//   mode 3: mode 1 with the z = side port driven by a 1 A boundary current
//   mode 1: + refill at freq_hz, equilibration, factorization, solve, R and L,
//           the fields' e_tet, the gauge-free potential recovered from them, and
//           the potential.out / VTK files of the solution (no run information), and the
//           VTK with compute_fields' data (fields.vtk)
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "aphi_solver/assembly.hpp"
#include "aphi_solver/dof_map.hpp"
#include "aphi_solver/equilibration.hpp"
#include "aphi_solver/factorization.hpp"
#include "aphi_solver/gauge_free_potential.hpp"
#include "aphi_solver/mesh.hpp"
#include "aphi_solver/postprocess.hpp"
#include "aphi_solver/problem.hpp"
#include "aphi_solver/problem_binding.hpp"
#include "aphi_solver/sparsity.hpp"
#include "aphi_solver/symbolic.hpp"

using namespace aphi_solver;
using Z = std::complex<double>;

static std::string g_out;

template <typename T>
static void dump(const std::string& name, const std::vector<T>& v) {
    std::ofstream f(g_out + "/" + name + ".bin", std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

static std::vector<long long> ints(const std::vector<int>& v) { return {v.begin(), v.end()}; }

int main(int argc, char** argv) {
    if (argc != 11) {
        std::fprintf(stderr, "usage: ref_driver OUTDIR n side i0 i1 j0 j1 sigma freq mode\n");
        return 2;
    }
    g_out = argv[1];
    const int n = std::atoi(argv[2]);
    const double side = std::atof(argv[3]);
    const int i0 = std::atoi(argv[4]), i1 = std::atoi(argv[5]), j0 = std::atoi(argv[6]), j1 = std::atoi(argv[7]);
    const double sigma = std::atof(argv[8]);
    const double freq = std::atof(argv[9]);
    const int mode = std::atoi(argv[10]);

    // Kuhn cube, node order of tools/generate_cube_mesh.py, coordinates side*i/n.
    Mesh mesh;
    const int np = n + 1;
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i)
                mesh.nodes.emplace_back(side * static_cast<double>(i) / static_cast<double>(n),
                                        side * static_cast<double>(j) / static_cast<double>(n),
                                        side * static_cast<double>(k) / static_cast<double>(n));
    static const int kKuhn[6][4][3] = {
        {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}}, {{0, 0, 0}, {1, 1, 0}, {0, 1, 0}, {1, 1, 1}},
        {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}}, {{0, 0, 0}, {0, 1, 1}, {0, 0, 1}, {1, 1, 1}},
        {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}}, {{0, 0, 0}, {1, 0, 1}, {1, 0, 0}, {1, 1, 1}},
    };
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const bool cond = i >= i0 && i < i1 && j >= j0 && j < j1;
                for (int t = 0; t < 6; ++t) {
                    TetVerts v{};
                    for (int c = 0; c < 4; ++c)
                        v[static_cast<size_t>(c)] =
                            ((k + kKuhn[t][c][2]) * np + (j + kKuhn[t][c][1])) * np + (i + kKuhn[t][c][0]);
                    mesh.tets.push_back(v);
                    mesh.tet_tags.push_back(cond ? 2 : 1);
                }
            }
    // terminals: conductor triangles lying in z = side (tag 10) and z = 0 (tag 11)
    for (size_t t = 0; t < mesh.tets.size(); ++t) {
        if (mesh.tet_tags[t] != 2) continue;
        for (const auto& fv : kTetLocalFaceVerts) {
            std::array<int, 3> f{mesh.tets[t][static_cast<size_t>(fv[0])], mesh.tets[t][static_cast<size_t>(fv[1])],
                                 mesh.tets[t][static_cast<size_t>(fv[2])]};
            const int k0 = f[0] / (np * np);
            if (k0 != f[1] / (np * np) || k0 != f[2] / (np * np) || (k0 != 0 && k0 != n)) continue;
            std::sort(f.begin(), f.end());
            mesh.tagged_boundary_faces.push_back(TaggedFace{f, k0 == n ? 10 : 11});
        }
    }
    mesh.build_topology();

    Problem p;
    p.type = AnalysisType::Frequency;
    p.frequencies = {freq};
    p.formulation = Formulation::FullWave;
    p.conditioning = Conditioning::RowScaled;
    p.bodies.push_back(Body{"B1", 0, "2", sigma, 1.0, 1.0});
    p.bodies.push_back(Body{"B2", 0, "1", 0.0, 1.0, 1.0});
    Port top;
    top.name = "P1";
    top.type = PortType::BoundaryVoltage;
    top.surface = {"10"};
    top.amplitude = 1.0;
    Port bottom = top;
    if (mode == 3) {
        // the z = side terminal driven by a 1 A boundary current instead of 1 V
        top.type = PortType::BoundaryCurrent;
        top.amplitude = 1.0;
    }
    bottom.name = "P2";
    bottom.surface = {"11"};
    bottom.amplitude = 0.0;
    p.ports = {top, bottom};

    const BoundProblem bound = bind_to_mesh(p, mesh);
    const DofMap dofs = build_dof_map(bound, mesh);
    const SparsityPattern pattern = build_sparsity(dofs, bound, mesh, SparsityStorage::UpperTriangle);
    const ScatterMap map = ScatterMap::build(pattern, dofs, mesh, bound);
    const SolverAnalysis an = analyze(pattern, Ordering::ApproximateMinimumDegree);

    {
        std::vector<long long> e, te, ts, tree, dir, live, pos, off, slot;
        for (const auto& x : mesh.edges) { e.push_back(x.first); e.push_back(x.second); }
        for (const auto& x : mesh.tet_edges) for (int v : x) te.push_back(v);
        for (const auto& x : mesh.tet_edge_signs) for (signed char v : x) ts.push_back(v);
        for (bool b : bound.gauge.is_tree_edge) tree.push_back(b ? 1 : 0);
        for (bool b : bound.dirichlet_edge) dir.push_back(b ? 1 : 0);
        for (int t = 0; t < mesh.num_tets(); ++t) {
            const int nl = map.live_count(t);
            live.push_back(nl);
            for (int i = 0; i < 6; ++i) pos.push_back(map.edge_at(t)[i]);
            for (int i = 0; i < 10; ++i) pos.push_back(map.phi_at(t)[i]);
            off.push_back(static_cast<long long>(slot.size()));
            const int sz = tet_block_size(nl, map.packed());
            for (int a = 0; a < sz; ++a) slot.push_back(map.slots(t)[a]);
        }
        off.push_back(static_cast<long long>(slot.size()));
        dump("edges", e);
        dump("tet_edges", te);
        dump("tet_edge_signs", ts);
        dump("tree_edge", tree);
        dump("dirichlet_edge", dir);
        dump("edge_index", ints(dofs.edge_index));
        dump("phi_index", ints(dofs.phi_index));
        dump("row_ptr", ints(pattern.row_ptr));
        dump("col_index", ints(pattern.col_index));
        dump("perm", ints(an.permutation.perm));
        dump("parent", ints(an.parent));
        dump("row_count", ints(an.row_count));
        dump("col_count", ints(an.col_count));
        dump("scatter_live", live);
        dump("scatter_pos", pos);
        dump("scatter_offset", off);
        dump("scatter_slot", slot);
    }

    std::FILE* meta = std::fopen((g_out + "/meta.json").c_str(), "w");
    std::fprintf(meta, "{\"n\": %d, \"nodes\": %d, \"edges\": %d, \"tets\": %d, \"num_a\": %d, \"num_phi\": %d, "
                       "\"num_total\": %d, \"nnz\": %zu, \"predicted_nnz\": %zu, \"num_roots\": %d, "
                       "\"tree_edges\": %d",
                 n, mesh.num_nodes(), mesh.num_edges(), mesh.num_tets(), dofs.num_a, dofs.num_phi, dofs.num_total,
                 pattern.nnz(), an.predicted_nnz, an.num_roots, bound.gauge.tree_edge_count);
    if (mode == 2) {
        // one solve_symmetric (a single factorization), then I, R and L
        const double omega = 2.0 * 3.14159265358979323846 * freq;
        SymmetricSystem sys = make_symmetric_system(pattern, dofs.num_total);
        refill(sys, bound, mesh, dofs, pattern, omega, Conditioning::RowScaled, &map);
        dump("values", sys.matrix.values());
        dump("rhs", sys.rhs);
        std::vector<Z> x;
        SolveReport rep;
        SolveOptions opt;
        opt.equilibration_iterations = 10;
        const bool ok = solve_symmetric(sys.matrix, sys.rhs, x, rep, opt, &an);
        dump("x", x);
        Solution sol;
        sol.x = x;
        sol.conditioning = Conditioning::RowScaled;
        sol.omega = omega;
        const FieldOutput fields = compute_fields(mesh, bound, dofs, sol);
        const Z I = terminal_current(mesh, fields, Vec3(0.0, 0.0, 1.0), side);
        const Z Zin = impedance(1.0, I);
        std::fprintf(meta,
                     ", \"freq\": %.17g, \"omega\": %.17g, \"solve_ok\": %s, \"smallest_pivot\": %.17g, "
                     "\"largest_pivot\": %.17g, \"largest_multiplier\": %.17g, \"factor_nnz\": %zu, "
                     "\"factorize_ms\": %.6g, \"backward_error\": %.17g, \"residual\": %.17g, "
                     "\"equilibration_min\": %.17g, \"equilibration_max\": %.17g, \"I_re\": %.17g, \"I_im\": %.17g, "
                     "\"R\": %.17g, \"L\": %.17g",
                     freq, omega, ok ? "true" : "false", rep.smallest_pivot, rep.largest_pivot,
                     rep.largest_multiplier, rep.factor_nnz, rep.factorize_ms, rep.backward_error, rep.residual,
                     rep.equilibration_min, rep.equilibration_max, I.real(), I.imag(), Zin.real(),
                     Zin.imag() / omega);
    }
    if (mode == 1 || mode == 3) {
        const double omega = 2.0 * 3.14159265358979323846 * freq;
        SymmetricSystem sys = make_symmetric_system(pattern, dofs.num_total);
        refill(sys, bound, mesh, dofs, pattern, omega, Conditioning::RowScaled);
        SymmetricSystem sys2 = make_symmetric_system(pattern, dofs.num_total);
        refill(sys2, bound, mesh, dofs, pattern, omega, Conditioning::RowScaled, &map);
        const bool map_equal = sys.matrix.values() == sys2.matrix.values() && sys.rhs == sys2.rhs;
        dump("values", sys.matrix.values());
        dump("rhs", sys.rhs);

        const std::vector<double> d = compute_symmetric_equilibration(sys.matrix, 10);
        dump("equil_d", d);
        const SparseSymmetricZ scaled = apply_symmetric_equilibration(sys.matrix, d);
        SymmetricFactor factor;
        FactorStats fs;
        const bool fok = factorize_ldlt(scaled, an, factor, fs);
        dump("diag", factor.diagonal);

        std::vector<Z> x;
        SolveReport rep;
        SolveOptions opt;
        opt.equilibration_iterations = 10;
        const bool ok = solve_symmetric(sys.matrix, sys.rhs, x, rep, opt, &an);
        dump("x", x);

        Solution sol;
        sol.x = x;
        sol.conditioning = Conditioning::RowScaled;
        sol.omega = omega;
        const FieldOutput fields = compute_fields(mesh, bound, dofs, sol);
        const Z I = terminal_current(mesh, fields, Vec3(0.0, 0.0, 1.0), side);
        std::vector<Z> e_flat;
        for (const auto& e : fields.e_tet) e_flat.insert(e_flat.end(), e.begin(), e.end());
        dump("e_tet", e_flat);
        const GaugeFreePotential gfp = recover_gauge_free_potential(mesh, bound, dofs, fields, omega);
        dump("gf_phi_vertex", gfp.phi_vertex);
        dump("gf_port_potential", gfp.port_potential);
        const NodalPotential pot = potential_at_nodes(mesh, bound, dofs, sol);
        write_potential(g_out + "/potential.out", pot, sol, nullptr);
        write_vtk(g_out + "/solution.vtk", mesh, pot, sol);
        write_vtk(g_out + "/fields.vtk", mesh, pot, sol, &fields, FieldSet::All);
        const Z Zin = impedance(1.0, I);
        std::fprintf(meta,
                     ", \"freq\": %.17g, \"omega\": %.17g, \"scatter_map_bitwise_equal\": %s, \"factor_ok\": %s, "
                     "\"smallest_pivot\": %.17g, \"largest_pivot\": %.17g, \"largest_multiplier\": %.17g, "
                     "\"factor_nnz\": %zu, \"factorize_ms\": %.6g, \"solve_ok\": %s, \"backward_error\": %.17g, "
                     "\"residual\": %.17g, \"equilibration_min\": %.17g, \"equilibration_max\": %.17g, "
                     "\"I_re\": %.17g, \"I_im\": %.17g, \"R\": %.17g, \"L\": %.17g, \"gf_unknowns\": %d, "
                     "\"gf_converged\": %s, \"gf_residual\": %.17g",
                     freq, omega, map_equal ? "true" : "false", fok ? "true" : "false", fs.smallest_pivot,
                     fs.largest_pivot, fs.largest_multiplier, fs.nnz, fs.milliseconds, ok ? "true" : "false",
                     rep.backward_error, rep.residual, rep.equilibration_min, rep.equilibration_max, I.real(), I.imag(),
                     Zin.real(), Zin.imag() / omega, gfp.unknowns, gfp.converged ? "true" : "false", gfp.residual);
    }
    std::fprintf(meta, "}\n");
    std::fclose(meta);
    return 0;
}

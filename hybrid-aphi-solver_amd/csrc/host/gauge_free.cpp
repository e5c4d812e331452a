// Integer plan of the gauge-free potential's P1 system
// (src/gauge_free_potential.cpp:40-177): vertex map, pattern, and the ordered
// list of per-tet terms behind every slot and rhs entry.
#include <map>
#include <stdexcept>

#include "aphi_host.hpp"

namespace aphi_amd {

GaugeFreePlan build_gauge_free_plan(const Mesh& mesh, const BoundProblem& bound, const DofMap& dofs) {
    GaugeFreePlan g;
    const int nv = mesh.num_nodes();
    const int nt = mesh.num_tets();
    // build_vertex_map (src/gauge_free_potential.cpp:47-85): ports first, so
    // every vertex of a current-driven terminal shares its port's column; then
    // the free vertices in ascending order; Cut / Fixed / Absent stay prescribed
    // at zero, a voltage port's vertices at the port value.
    g.column.assign(static_cast<size_t>(nv), -1);
    g.prescribed.assign(static_cast<size_t>(nv), Complex{});
    g.port_column.assign(bound.ports.size(), -1);
    for (size_t p = 0; p < bound.ports.size(); ++p) {
        if (dofs.port_is_fixed[p]) continue;
        g.port_column[p] = g.unknowns++;
    }
    for (int v = 0; v < nv; ++v) {
        const size_t uv = static_cast<size_t>(v);
        const PhiDof st = dofs.phi_state[static_cast<size_t>(dofs.vertex_p2(v))];
        if (st == PhiDof::Port) {
            const int p = dofs.phi_port[static_cast<size_t>(dofs.vertex_p2(v))];
            if (p < 0) continue;
            const size_t up = static_cast<size_t>(p);
            if (dofs.port_is_fixed[up]) g.prescribed[uv] = dofs.port_value[up];
            else g.column[uv] = g.port_column[up];
        } else if (st == PhiDof::Free) {
            g.column[uv] = g.unknowns++;
        }
    }

    // which tets take part: a body, and beta = (sigma, w eps0 eps_r) != 0 for w != 0
    g.tet_on.assign(static_cast<size_t>(nt), 0);
    for (int t = 0; t < nt; ++t) {
        const int b = bound.body_of_tet[static_cast<size_t>(t)];
        if (b < 0) continue;
        const Body& body = bound.bodies[static_cast<size_t>(b)];
        g.tet_on[static_cast<size_t>(t)] = (body.sigma != 0.0 || body.eps_r != 0.0) ? 1 : 0;
    }

    // pattern: the (row, column >= row) pairs the reference's rows[ri][cj] += k touches
    std::vector<std::map<int, std::vector<uint32_t>>> rows(static_cast<size_t>(g.unknowns));
    std::vector<std::vector<std::pair<uint32_t, uint8_t>>> rhs(static_cast<size_t>(g.unknowns));
    auto value_id = [&](const Complex& v) {
        for (size_t k = 0; k < g.values.size(); ++k)
            if (g.values[k] == v) return static_cast<int>(k);
        if (g.values.size() >= 64) throw std::invalid_argument("gauge-free potential: more than 64 distinct prescribed values");
        g.values.push_back(v);
        return static_cast<int>(g.values.size() - 1);
    };
    for (int t = 0; t < nt; ++t) {
        if (!g.tet_on[static_cast<size_t>(t)]) continue;
        const std::array<int, 4>& verts = mesh.tets[static_cast<size_t>(t)];
        for (int i = 0; i < 4; ++i) {
            const int ri = g.column[static_cast<size_t>(verts[static_cast<size_t>(i)])];
            if (ri < 0) continue;
            rhs[static_cast<size_t>(ri)].emplace_back(static_cast<uint32_t>(t) * 20u + 16u + static_cast<uint32_t>(i), 1);
            for (int j = 0; j < 4; ++j) {
                const size_t vj = static_cast<size_t>(verts[static_cast<size_t>(j)]);
                const int cj = g.column[vj];
                const uint32_t e = static_cast<uint32_t>(t) * 20u + static_cast<uint32_t>(i * 4 + j);
                if (cj < 0) {
                    if (g.prescribed[vj] != Complex{})
                        rhs[static_cast<size_t>(ri)].emplace_back(e, static_cast<uint8_t>(2 | (value_id(g.prescribed[vj]) << 2)));
                    continue;
                }
                if (cj < ri) continue;
                rows[static_cast<size_t>(ri)][cj].push_back(e);
            }
        }
    }
    const int n = g.unknowns;
    g.pattern.rows = g.pattern.cols = n;
    g.pattern.upper_only = true;
    g.pattern.row_ptr.assign(static_cast<size_t>(n) + 1, 0);
    for (int r = 0; r < n; ++r)
        g.pattern.row_ptr[static_cast<size_t>(r) + 1] =
            g.pattern.row_ptr[static_cast<size_t>(r)] + static_cast<int>(rows[static_cast<size_t>(r)].size());
    g.pattern.col_index.reserve(static_cast<size_t>(g.pattern.row_ptr[static_cast<size_t>(n)]));
    g.target_ptr.assign(1, 0);
    for (int r = 0; r < n; ++r)
        for (const auto& kv : rows[static_cast<size_t>(r)]) {
            g.pattern.col_index.push_back(kv.first);
            for (uint32_t e : kv.second) {
                g.entry.push_back(e);
                g.kind.push_back(0);
            }
            g.target_ptr.push_back(static_cast<int64_t>(g.entry.size()));
        }
    for (int r = 0; r < n; ++r) {
        for (const auto& ek : rhs[static_cast<size_t>(r)]) {
            g.entry.push_back(ek.first);
            g.kind.push_back(ek.second);
        }
        g.target_ptr.push_back(static_cast<int64_t>(g.entry.size()));
    }
    return g;
}

}  // namespace aphi_amd

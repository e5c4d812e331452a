// Binding, tree-cotree gauge and DOF map for the cube configurations.
#include <algorithm>

#include "aphi_host.hpp"

namespace aphi_amd {

namespace {

std::vector<int> sorted_unique(std::vector<int> v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

// Union-find by rank with path halving (src/tree_cotree.cpp:58-89). The rank rule
// decides which root survives a union, and therefore which node a component is
// known by; the component numbering below depends on it only through find().
class DisjointSets {
public:
    explicit DisjointSets(int n) : up_(static_cast<size_t>(n)), rank_(static_cast<size_t>(n), 0) {
        for (int i = 0; i < n; ++i) up_[static_cast<size_t>(i)] = i;
    }
    int find(int x) {
        while (up_[static_cast<size_t>(x)] != x) {
            up_[static_cast<size_t>(x)] = up_[static_cast<size_t>(up_[static_cast<size_t>(x)])];
            x = up_[static_cast<size_t>(x)];
        }
        return x;
    }
    void unite(int a, int b) {
        a = find(a);
        b = find(b);
        if (a == b) return;
        if (rank_[static_cast<size_t>(a)] < rank_[static_cast<size_t>(b)]) std::swap(a, b);
        up_[static_cast<size_t>(b)] = a;
        if (rank_[static_cast<size_t>(a)] == rank_[static_cast<size_t>(b)]) ++rank_[static_cast<size_t>(a)];
    }

private:
    std::vector<int> up_, rank_;
};

}  // namespace

std::vector<uint8_t> boundary_edge_mask(const Mesh& mesh) {
    std::vector<uint8_t> mask(static_cast<size_t>(mesh.num_edges()), 0);
    for (int f = 0; f < mesh.num_faces(); ++f) {
        if (!mesh.is_boundary_face(f)) continue;
        const auto& v = mesh.faces[static_cast<size_t>(f)];
        mask[static_cast<size_t>(mesh.find_edge(v[0], v[1]))] = 1;
        mask[static_cast<size_t>(mesh.find_edge(v[1], v[2]))] = 1;
        mask[static_cast<size_t>(mesh.find_edge(v[0], v[2]))] = 1;
    }
    return mask;
}

TreeCotree build_tree_cotree(const Mesh& mesh, const std::vector<uint8_t>& dirichlet_edge) {
    const int n = mesh.num_nodes();
    const int m = mesh.num_edges();
    if (static_cast<int>(dirichlet_edge.size()) != m)
        throw std::invalid_argument("build_tree_cotree: dirichlet_edge must have one entry per mesh edge");
    TreeCotree out;
    out.is_tree_edge.assign(static_cast<size_t>(m), 0);
    if (n == 0) return out;

    // Node -> (neighbour, edge) lists filled in edge order: the BFS visiting order
    // below is defined by this layout (src/tree_cotree.cpp:10-52).
    std::vector<int> off(static_cast<size_t>(n) + 1, 0);
    for (const auto& e : mesh.edges) {
        ++off[static_cast<size_t>(e.first) + 1];
        ++off[static_cast<size_t>(e.second) + 1];
    }
    for (int i = 0; i < n; ++i) off[static_cast<size_t>(i) + 1] += off[static_cast<size_t>(i)];
    std::vector<int> nbr(static_cast<size_t>(2 * m)), via(static_cast<size_t>(2 * m));
    {
        std::vector<int> cur(off.begin(), off.end() - 1);
        for (int e = 0; e < m; ++e) {
            const int a = mesh.edges[static_cast<size_t>(e)].first, b = mesh.edges[static_cast<size_t>(e)].second;
            nbr[static_cast<size_t>(cur[static_cast<size_t>(a)])] = b;
            via[static_cast<size_t>(cur[static_cast<size_t>(a)]++)] = e;
            nbr[static_cast<size_t>(cur[static_cast<size_t>(b)])] = a;
            via[static_cast<size_t>(cur[static_cast<size_t>(b)]++)] = e;
        }
    }

    // Dirichlet surfaces = components of the Dirichlet-edge graph, numbered by
    // their lowest node.
    DisjointSets ds(n);
    std::vector<uint8_t> on_surface(static_cast<size_t>(n), 0);
    for (int e = 0; e < m; ++e) {
        if (!dirichlet_edge[static_cast<size_t>(e)]) continue;
        const int a = mesh.edges[static_cast<size_t>(e)].first, b = mesh.edges[static_cast<size_t>(e)].second;
        ds.unite(a, b);
        on_surface[static_cast<size_t>(a)] = on_surface[static_cast<size_t>(b)] = 1;
    }
    std::vector<int> comp_of_root(static_cast<size_t>(n), -1), comp(static_cast<size_t>(n), -1);
    int ncomp = 0, surface_nodes = 0;
    for (int i = 0; i < n; ++i) {
        if (!on_surface[static_cast<size_t>(i)]) continue;
        ++surface_nodes;
        int& c = comp_of_root[static_cast<size_t>(ds.find(i))];
        if (c < 0) c = ncomp++;
        comp[static_cast<size_t>(i)] = c;
    }
    std::vector<std::vector<int>> members(static_cast<size_t>(ncomp));
    for (int i = 0; i < n; ++i)
        if (comp[static_cast<size_t>(i)] >= 0) members[static_cast<size_t>(comp[static_cast<size_t>(i)])].push_back(i);

    // Surface spanning trees from Dirichlet edges only.
    int surface_tree = 0;
    {
        std::vector<uint8_t> seen(static_cast<size_t>(n), 0);
        std::vector<int> q;
        for (int c = 0; c < ncomp; ++c) {
            const int s = members[static_cast<size_t>(c)].front();
            seen[static_cast<size_t>(s)] = 1;
            q.assign(1, s);
            for (size_t h = 0; h < q.size(); ++h) {
                const int u = q[h];
                for (int k = off[static_cast<size_t>(u)]; k < off[static_cast<size_t>(u) + 1]; ++k) {
                    const int e = via[static_cast<size_t>(k)], v = nbr[static_cast<size_t>(k)];
                    if (!dirichlet_edge[static_cast<size_t>(e)] || seen[static_cast<size_t>(v)]) continue;
                    seen[static_cast<size_t>(v)] = 1;
                    out.is_tree_edge[static_cast<size_t>(e)] = 1;
                    ++surface_tree;
                    q.push_back(v);
                }
            }
        }
    }

    // Interior BFS: a surface is entered as a whole group through one edge.
    std::vector<uint8_t> visited(static_cast<size_t>(n), 0);
    std::vector<int> q;
    q.reserve(static_cast<size_t>(n));
    int groups = 0, roots = 0, interior_tree = 0;
    const auto take_group = [&](int node) {
        ++groups;
        const int c = comp[static_cast<size_t>(node)];
        if (c >= 0) {
            for (int w : members[static_cast<size_t>(c)]) {
                visited[static_cast<size_t>(w)] = 1;
                q.push_back(w);
            }
        } else {
            visited[static_cast<size_t>(node)] = 1;
            q.push_back(node);
        }
    };
    size_t head = 0;
    const auto drain = [&]() {
        while (head < q.size()) {
            const int u = q[head++];
            for (int k = off[static_cast<size_t>(u)]; k < off[static_cast<size_t>(u) + 1]; ++k) {
                const int v = nbr[static_cast<size_t>(k)];
                if (visited[static_cast<size_t>(v)]) continue;
                const int e = via[static_cast<size_t>(k)];
                if (dirichlet_edge[static_cast<size_t>(e)])
                    throw std::logic_error("build_tree_cotree: interior tree tried to use a Dirichlet edge");
                out.is_tree_edge[static_cast<size_t>(e)] = 1;
                ++interior_tree;
                take_group(v);
            }
        }
    };
    take_group(ncomp > 0 ? members.front().front() : 0);
    ++roots;
    drain();
    for (int s = 0; s < n; ++s) {
        if (visited[static_cast<size_t>(s)]) continue;
        take_group(s);
        ++roots;
        drain();
    }
    out.num_groups = groups;
    out.num_reference_groups = roots;
    out.num_dirichlet_components = ncomp;
    out.surface_tree_edge_count = surface_tree;
    out.interior_tree_edge_count = interior_tree;
    out.tree_edge_count = surface_tree + interior_tree;
    if (surface_tree != surface_nodes - ncomp)
        throw std::logic_error("build_tree_cotree: surface tree does not span each surface");
    if (out.tree_edge_count != n - roots)
        throw std::logic_error("build_tree_cotree: tree is not a spanning tree over all nodes");
    return out;
}

BoundProblem bind_cube(const Mesh& mesh, const CubeSpec& spec) {
    BoundProblem b;
    const int nt = mesh.num_tets();
    Body cond{"B1", 2, spec.sigma, spec.eps_r, spec.mu_r, {}};
    Body air{"B2", 1, spec.air_sigma, spec.air_eps_r, spec.air_mu_r, {}};
    b.bodies = {cond, air};
    b.body_of_tet.assign(static_cast<size_t>(nt), -1);
    for (int t = 0; t < nt; ++t) {
        const int tag = mesh.tet_tags[static_cast<size_t>(t)];
        const int body = tag == 2 ? 0 : (tag == 1 ? 1 : -1);
        if (body < 0) throw std::invalid_argument("bind_cube: a tet carries no body tag");
        b.body_of_tet[static_cast<size_t>(t)] = body;
        b.bodies[static_cast<size_t>(body)].tets.push_back(t);
    }
    for (const Body& body : b.bodies)
        if (body.tets.empty()) throw std::invalid_argument("bind_cube: body '" + body.name + "' has no tetrahedra");
    b.full_wave = true;
    b.phi_tet.assign(static_cast<size_t>(nt), 1);

    const auto port_on = [&](const char* name, int tag, Complex v) {
        Port p;
        p.name = name;
        p.type = PortType::BoundaryVoltage;
        p.amplitude = v;
        std::vector<int> faces, nodes, edges;
        for (size_t i = 0; i < mesh.tagged_faces.size(); ++i) {
            if (mesh.tagged_face_tag[i] != tag) continue;
            const auto& f = mesh.tagged_faces[i];
            const int fi = mesh.find_face(f[0], f[1], f[2]);
            if (fi < 0) throw std::invalid_argument("bind_cube: a tagged triangle is not a face of any tet");
            faces.push_back(fi);
            nodes.insert(nodes.end(), f.begin(), f.end());
            edges.push_back(mesh.find_edge(f[0], f[1]));
            edges.push_back(mesh.find_edge(f[1], f[2]));
            edges.push_back(mesh.find_edge(f[0], f[2]));
        }
        if (faces.empty()) throw std::invalid_argument(std::string("bind_cube: port '") + name + "' has no faces");
        p.faces = sorted_unique(std::move(faces));
        p.nodes = sorted_unique(std::move(nodes));
        p.edges = sorted_unique(std::move(edges));
        for (int f : p.faces)
            if (!mesh.is_boundary_face(f)) throw std::invalid_argument("bind_cube: a port face is interior");
        return p;
    };
    b.ports.push_back(port_on("P1", 10, spec.v_top));
    if (spec.top_current) {
        // a boundary current port: its terminal potential is an unknown and the
        // current drives its row (src/dof_map.cpp:91-103, src/assembly.cpp:396-403)
        b.ports.back().type = PortType::BoundaryCurrent;
        b.ports.back().amplitude = spec.i_top;
    }
    b.ports.push_back(port_on("P2", 11, spec.v_bottom));
    std::vector<int> shared;
    std::set_intersection(b.ports[0].nodes.begin(), b.ports[0].nodes.end(), b.ports[1].nodes.begin(),
                          b.ports[1].nodes.end(), std::back_inserter(shared));
    if (!shared.empty()) throw std::invalid_argument("bind_cube: the two terminals share nodes (short circuit)");

    b.dirichlet_edge = boundary_edge_mask(mesh);
    b.gauge = build_tree_cotree(mesh, b.dirichlet_edge);
    return b;
}

DofMap build_dof_map(const BoundProblem& bound, const Mesh& mesh) {
    DofMap d;
    d.num_nodes = mesh.num_nodes();
    d.num_p2_nodes = mesh.num_nodes() + mesh.num_edges();
    const size_t ne = static_cast<size_t>(mesh.num_edges());
    d.edge_state.assign(ne, EdgeDof::Free);
    d.edge_index.assign(ne, -1);
    for (size_t e = 0; e < ne; ++e) {
        if (bound.dirichlet_edge[e]) d.edge_state[e] = EdgeDof::Dirichlet;
        else if (bound.gauge.is_tree_edge[e]) d.edge_state[e] = EdgeDof::Tree;
        else d.edge_index[e] = d.num_a++;
    }

    const size_t np = static_cast<size_t>(d.num_p2_nodes);
    d.phi_state.assign(np, PhiDof::Absent);
    d.phi_index.assign(np, -1);
    d.phi_port.assign(np, -1);
    d.phi_fixed_value.assign(np, Complex(0.0, 0.0));
    std::vector<uint8_t> present(np, 0);
    for (int t = 0; t < mesh.num_tets(); ++t) {
        if (!bound.phi_tet[static_cast<size_t>(t)]) continue;
        for (int v : mesh.tets[static_cast<size_t>(t)]) present[static_cast<size_t>(d.vertex_p2(v))] = 1;
        for (int e : mesh.tet_edges[static_cast<size_t>(t)]) present[static_cast<size_t>(d.edge_p2(e))] = 1;
    }
    for (size_t k = 0; k < bound.ports.size(); ++k) {
        const Port& p = bound.ports[k];
        const bool internal = p.type == PortType::InternalCurrent || p.type == PortType::InternalVoltage;
        const PhiDof st = internal ? PhiDof::Cut : PhiDof::Port;
        for (int v : p.nodes) {
            d.phi_state[static_cast<size_t>(d.vertex_p2(v))] = st;
            d.phi_port[static_cast<size_t>(d.vertex_p2(v))] = static_cast<int>(k);
        }
        for (int e : p.edges) {
            d.phi_state[static_cast<size_t>(d.edge_p2(e))] = st;
            d.phi_port[static_cast<size_t>(d.edge_p2(e))] = static_cast<int>(k);
        }
    }
    for (size_t i = 0; i < np; ++i) {
        if (!present[i]) {
            d.phi_state[i] = PhiDof::Absent;
            continue;
        }
        if (d.phi_state[i] == PhiDof::Absent) d.phi_state[i] = PhiDof::Free;
        if (d.phi_state[i] == PhiDof::Free) d.phi_index[i] = d.num_phi++;
    }
    for (const Port& p : bound.ports) {
        d.port_index.push_back(-1);
        const bool fixed = !is_current_driven(p.type);
        d.port_is_fixed.push_back(fixed ? 1 : 0);
        d.port_value.push_back(fixed ? p.amplitude : Complex(0.0, 0.0));
    }
    for (size_t i = 0; i < np; ++i)
        if (d.phi_index[i] >= 0) d.phi_index[i] += d.num_a;
    d.num_ports = static_cast<int>(bound.ports.size());
    for (int k = 0; k < d.num_ports; ++k) d.port_index[static_cast<size_t>(k)] = d.num_a + d.num_phi + k;
    d.num_total = d.num_a + d.num_phi + d.num_ports;
    return d;
}

TetDofs DofMap::local_dofs(int tet, const Mesh& mesh, const BoundProblem& bound) const {
    (void)bound;  // only internal cuts consult the tet's side; the cube has none
    TetDofs o;
    const size_t t = static_cast<size_t>(tet);
    for (int le = 0; le < 6; ++le) {
        const int e = mesh.tet_edges[t][static_cast<size_t>(le)];
        o.edge_index[le] = edge_index[static_cast<size_t>(e)];
        o.edge_coeff[le] = static_cast<double>(mesh.tet_edge_signs[t][static_cast<size_t>(le)]);
    }
    int p2[10];
    for (int v = 0; v < 4; ++v) p2[v] = vertex_p2(mesh.tets[t][static_cast<size_t>(v)]);
    for (int le = 0; le < 6; ++le) p2[4 + le] = edge_p2(mesh.tet_edges[t][static_cast<size_t>(le)]);
    for (int ln = 0; ln < 10; ++ln) {
        o.phi_index[ln] = -1;
        o.phi_coeff[ln] = 0.0;
        o.phi_fixed[ln] = Complex(0.0, 0.0);
        o.phi_fixed_port[ln] = -1;
        const size_t i = static_cast<size_t>(p2[ln]);
        switch (phi_state[i]) {
            case PhiDof::Free:
                o.phi_index[ln] = phi_index[i];
                o.phi_coeff[ln] = 1.0;
                break;
            case PhiDof::Port: {
                const size_t k = static_cast<size_t>(phi_port[i]);
                if (port_is_fixed[k]) {
                    o.phi_fixed[ln] = port_value[k];
                    o.phi_fixed_port[ln] = static_cast<int>(k);
                } else {
                    o.phi_index[ln] = port_index[k];
                    o.phi_coeff[ln] = 1.0;
                }
                break;
            }
            case PhiDof::Fixed:
                o.phi_fixed[ln] = phi_fixed_value[i];
                break;
            case PhiDof::Cut:
                throw std::logic_error("local_dofs: internal cuts are not part of this path");
            case PhiDof::Absent:
                break;
        }
    }
    return o;
}

}  // namespace aphi_amd

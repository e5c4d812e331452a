// Result files in the reference's formats, from a solution handed over by the
// device path: the nodal potential (potential_at_nodes + write_potential,
// src/postprocess.cpp:42-123, 549-606) and the legacy-VTK quadratic-tet mesh with
// the potential and, when given, compute_fields' data computed on the device
// (write_vtk, src/postprocess.cpp:1055-1276).
// Numbers are written as the reference writes them (std::to_chars: the shortest
// text that reads back to the same double), so the same inputs give the same
// bytes.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "aphi_host.hpp"

namespace aphi_amd {
namespace {

struct Text {
    std::string s;
    void put(const char* t) { s += t; }
    void put(char c) { s += c; }
    void put(long long v) {
        char b[32];
        const auto r = std::to_chars(b, b + sizeof(b), v);
        s.append(b, r.ptr);
    }
    void put(int v) { put(static_cast<long long>(v)); }
    void put(double v) {
        char b[64];
        const auto r = std::to_chars(b, b + sizeof(b), v);
        if (r.ec != std::errc()) throw std::runtime_error("writer: double formatting failed");
        s.append(b, r.ptr);
    }
};

void save(const std::string& path, const std::string& text, const char* who) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error(std::string(who) + ": could not open '" + path + "' for writing");
    const size_t done = std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
    if (done != text.size()) throw std::runtime_error(std::string(who) + ": short write");
}

const char* conditioning_word(int c) {
    return c == 1 ? "row_scaled" : (c == 2 ? "scaled_phi" : "natural");
}

// complex product in the (ac - bd, ad + bc) form the reference's std::complex uses
Complex cmul(const Complex& a, const Complex& b) {
    return Complex(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}

}  // namespace

NodalPotential potential_at_nodes(const Mesh& mesh, const DofMap& dofs, const Complex* x, int n, double omega,
                                  int conditioning) {
    if (n != dofs.num_total)
        throw std::invalid_argument("potential_at_nodes: the solution has " + std::to_string(n) +
                                    " entries for a DOF map of " + std::to_string(dofs.num_total) +
                                    ". They must describe the same problem.");
    NodalPotential out;
    out.num_vertices = mesh.num_nodes();
    out.num_edges = mesh.num_edges();
    const int np = dofs.num_p2_nodes;
    out.position.resize(static_cast<size_t>(np));
    out.value.assign(static_cast<size_t>(np), Complex(0.0, 0.0));
    out.status.assign(static_cast<size_t>(np), 'A');
    for (int v = 0; v < mesh.num_nodes(); ++v) out.position[static_cast<size_t>(dofs.vertex_p2(v))] = mesh.nodes[static_cast<size_t>(v)];
    for (int e = 0; e < mesh.num_edges(); ++e) {
        const Vec3& p = mesh.nodes[static_cast<size_t>(mesh.edges[static_cast<size_t>(e)].first)];
        const Vec3& q = mesh.nodes[static_cast<size_t>(mesh.edges[static_cast<size_t>(e)].second)];
        out.position[static_cast<size_t>(dofs.edge_p2(e))] = Vec3{0.5 * (p.x + q.x), 0.5 * (p.y + q.y), 0.5 * (p.z + q.z)};
    }
    // Solution::phi_scale (src/postprocess.cpp:35-40): j w under ScaledPhi, else 1
    const Complex scale = conditioning == 2 ? Complex(0.0, omega) : Complex(1.0, 0.0);
    for (int i = 0; i < np; ++i) {
        const size_t u = static_cast<size_t>(i);
        switch (dofs.phi_state[u]) {
            case PhiDof::Free:
                out.value[u] = cmul(scale, x[dofs.phi_index[u]]);
                out.status[u] = 'F';
                break;
            case PhiDof::Port:
            case PhiDof::Cut: {
                const size_t k = static_cast<size_t>(dofs.phi_port[u]);
                out.value[u] = dofs.port_is_fixed[k] ? dofs.port_value[k] : cmul(scale, x[dofs.port_index[k]]);
                out.status[u] = dofs.phi_state[u] == PhiDof::Port ? 'T' : 'C';
                break;
            }
            case PhiDof::Fixed:
                out.value[u] = dofs.phi_fixed_value[u];
                out.status[u] = 'P';
                break;
            case PhiDof::Absent:
                out.value[u] = Complex(0.0, 0.0);
                out.status[u] = 'A';
                break;
        }
    }
    return out;
}

size_t write_potential(const std::string& path, const NodalPotential& p, double omega, int conditioning) {
    Text t;
    t.put("# potential.out -- nodal electric scalar potential Phi, in volts\n");
    t.put("# A-Phi solver ");
    t.put(kReferenceVersion);
    t.put("\n# conditioning ");
    t.put(conditioning_word(conditioning));
    t.put("   omega ");
    t.put(omega);
    t.put(" rad/s   frequency ");
    t.put(omega / 6.283185307179586476925286766559);
    t.put(" Hz\n# nodes ");
    t.put(static_cast<int>(p.position.size()));
    t.put("  (vertices ");
    t.put(p.num_vertices);
    t.put(", edge midpoints ");
    t.put(p.num_edges);
    t.put(")\n");
    t.put("# positions in metres. Mid-edge values are EXACT P2 unknowns, not interpolated.\n");
    t.put("# status  F free   T terminal   C cut (value is the PLUS side; Phi jumps here)"
          "   P prescribed   A absent\n");
    t.put("# index x y z Re(Phi) Im(Phi) status\n");
    for (size_t i = 0; i < p.position.size(); ++i) {
        t.put(static_cast<int>(i));
        t.put(' ');
        t.put(p.position[i].x);
        t.put(' ');
        t.put(p.position[i].y);
        t.put(' ');
        t.put(p.position[i].z);
        t.put(' ');
        t.put(p.value[i].real());
        t.put(' ');
        t.put(p.value[i].imag());
        t.put(' ');
        t.put(p.status[i]);
        t.put('\n');
    }
    save(path, t.s, "write_potential");
    return t.s.size();
}

namespace {
// put_vectors / put_magnitude of the reference's writer: the real or imaginary
// parts of 3 complex per item, and sqrt(sum |c|^2) per item
void put_vectors(Text& t, const char* name, const double* v, int n, bool real_part) {
    t.put("\nVECTORS ");
    t.put(name);
    t.put(" double\n");
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            if (k != 0) t.put(' ');
            t.put(v[6 * static_cast<size_t>(i) + 2 * k + (real_part ? 0 : 1)]);
        }
        t.put('\n');
    }
}

void put_magnitude(Text& t, const char* name, const double* v, int n) {
    t.put("\nSCALARS ");
    t.put(name);
    t.put(" double 1\nLOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += std::norm(Complex(v[6 * static_cast<size_t>(i) + 2 * k],
                                                           v[6 * static_cast<size_t>(i) + 2 * k + 1]));
        t.put(std::sqrt(s));
        t.put('\n');
    }
}

size_t write_vtk_impl(const std::string& path, const Mesh& mesh, const NodalPotential& p, double omega,
                      int conditioning, const FieldsView* f);
}  // namespace

size_t write_vtk_potential(const std::string& path, const Mesh& mesh, const NodalPotential& p, double omega,
                           int conditioning) {
    return write_vtk_impl(path, mesh, p, omega, conditioning, nullptr);
}

size_t write_vtk_fields(const std::string& path, const Mesh& mesh, const NodalPotential& p, double omega,
                        int conditioning, const FieldsView& f) {
    return write_vtk_impl(path, mesh, p, omega, conditioning, &f);
}

namespace {
size_t write_vtk_impl(const std::string& path, const Mesh& mesh, const NodalPotential& p, double omega,
                      int conditioning, const FieldsView* f) {
    if (static_cast<int>(p.position.size()) != mesh.num_nodes() + mesh.num_edges())
        throw std::invalid_argument("write_vtk: the potential has " + std::to_string(p.position.size()) +
                                    " nodes for a mesh of " + std::to_string(mesh.num_nodes()) + " vertices and " +
                                    std::to_string(mesh.num_edges()) + " edges");
    // VTK's quadratic tetrahedron lists its mid-edge nodes in this order of our local edges
    static const int kVtkEdge[6] = {0, 3, 1, 2, 5, 4};
    const int np = static_cast<int>(p.position.size());
    const int nc = mesh.num_tets();
    Text t;
    t.put("# vtk DataFile Version 3.0\nA-Phi fields, conditioning ");
    t.put(conditioning_word(conditioning));
    t.put(", omega ");
    t.put(omega);
    t.put(" rad/s\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS ");
    t.put(np);
    t.put(" double\n");
    for (const Vec3& q : p.position) {
        t.put(q.x);
        t.put(' ');
        t.put(q.y);
        t.put(' ');
        t.put(q.z);
        t.put('\n');
    }
    t.put("\nCELLS ");
    t.put(nc);
    t.put(' ');
    t.put(nc * 11);
    t.put('\n');
    for (int c = 0; c < nc; ++c) {
        t.put("10");
        for (int v = 0; v < 4; ++v) {
            t.put(' ');
            t.put(mesh.tets[static_cast<size_t>(c)][static_cast<size_t>(v)]);
        }
        for (int slot = 0; slot < 6; ++slot) {
            t.put(' ');
            t.put(mesh.num_nodes() + mesh.tet_edges[static_cast<size_t>(c)][static_cast<size_t>(kVtkEdge[slot])]);
        }
        t.put('\n');
    }
    t.put("\nCELL_TYPES ");
    t.put(nc);
    t.put('\n');
    for (int c = 0; c < nc; ++c) t.put("24\n");  // VTK_QUADRATIC_TETRA
    t.put("\nCELL_DATA ");
    t.put(nc);
    t.put("\nSCALARS body_tag int 1\nLOOKUP_TABLE default\n");
    for (int c = 0; c < nc; ++c) {
        t.put(c < static_cast<int>(mesh.tet_tags.size()) ? mesh.tet_tags[static_cast<size_t>(c)] : -1);
        t.put('\n');
    }
    if (f) {
        t.put("\nSCALARS sigma double 1\nLOOKUP_TABLE default\n");
        for (int c = 0; c < nc; ++c) {
            t.put(f->sigma_tet[c]);
            t.put('\n');
        }
        put_vectors(t, "B_cell_real", f->b_tet, nc, true);
        put_vectors(t, "B_cell_imag", f->b_tet, nc, false);
        put_vectors(t, "E_cell_real", f->e_tet, nc, true);
        put_vectors(t, "E_cell_imag", f->e_tet, nc, false);
        put_vectors(t, "J_cell_real", f->j_tet, nc, true);
        put_vectors(t, "J_cell_imag", f->j_tet, nc, false);
    }
    t.put("\nPOINT_DATA ");
    t.put(np);
    t.put("\nSCALARS phi_real double 1\nLOOKUP_TABLE default\n");
    for (int i = 0; i < np; ++i) {
        t.put(p.value[static_cast<size_t>(i)].real());
        t.put('\n');
    }
    t.put("\nSCALARS phi_imag double 1\nLOOKUP_TABLE default\n");
    for (int i = 0; i < np; ++i) {
        t.put(p.value[static_cast<size_t>(i)].imag());
        t.put('\n');
    }
    t.put("\nSCALARS phi_magnitude double 1\nLOOKUP_TABLE default\n");
    for (int i = 0; i < np; ++i) {
        t.put(std::abs(p.value[static_cast<size_t>(i)]));
        t.put('\n');
    }
    t.put("\nSCALARS phi_present int 1\nLOOKUP_TABLE default\n");
    for (int i = 0; i < np; ++i) {
        t.put(p.status[static_cast<size_t>(i)] == 'A' ? 0 : 1);
        t.put('\n');
    }
    if (f) {
        const struct {
            const char *re, *im, *mag;
            const double* v;
        } node_fields[] = {{"A_real", "A_imag", "A_magnitude", f->a_node},
                           {"B_real", "B_imag", "B_magnitude", f->b_node},
                           {"H_real", "H_imag", "H_magnitude", f->h_node},
                           {"E_real", "E_imag", "E_magnitude", f->e_node},
                           {"J_real", "J_imag", "J_magnitude", f->j_node}};
        for (const auto& nf : node_fields) {
            put_vectors(t, nf.re, nf.v, np, true);
            put_vectors(t, nf.im, nf.v, np, false);
            put_magnitude(t, nf.mag, nf.v, np);
        }
        t.put("\nSCALARS material_interface int 1\nLOOKUP_TABLE default\n");
        for (int i = 0; i < np; ++i) {
            t.put(static_cast<int>(f->on_material_interface[i]));
            t.put('\n');
        }
    }
    save(path, t.s, "write_vtk");
    return t.s.size();
}
}  // namespace

}  // namespace aphi_amd

// Kuhn cube generation and mesh topology.
//
// Topology follows src/mesh.cpp:26-142: edges and faces are numbered by sorting
// (lo, hi) / (a, b, c) keys, so their indices are the rank of the key, and each
// tet's local edge carries +1 when its local direction runs low -> high.
#include <algorithm>
#include <cmath>

#include "aphi_host.hpp"

namespace aphi_amd {

double node_coordinate(const CubeSpec& spec, int index) {
    return spec.side * static_cast<double>(index) / static_cast<double>(spec.n);
}

Mesh make_kuhn_cube(const CubeSpec& spec) {
    if (spec.n < 1) throw std::invalid_argument("make_kuhn_cube: n must be >= 1");
    if (!(spec.bar_i0 >= 0 && spec.bar_i0 < spec.bar_i1 && spec.bar_i1 <= spec.n &&
          spec.bar_j0 >= 0 && spec.bar_j0 < spec.bar_j1 && spec.bar_j1 <= spec.n)) {
        throw std::invalid_argument("make_kuhn_cube: the conductor block must lie inside the cube");
    }
    const int n = spec.n;
    const int np = n + 1;
    Mesh m;
    m.nodes.resize(static_cast<size_t>(np) * np * np);
    // Node order of tools/generate_cube_mesh.py:23-28: k outermost, i innermost.
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                m.nodes[static_cast<size_t>((k * np + j) * np + i)] =
                    Vec3(node_coordinate(spec, i), node_coordinate(spec, j), node_coordinate(spec, k));
            }
    // The six Kuhn tets of a unit hex, as corner offsets (generate_cube_mesh.py:32-39).
    static const int kKuhn[6][4][3] = {
        {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}}, {{0, 0, 0}, {1, 1, 0}, {0, 1, 0}, {1, 1, 1}},
        {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}}, {{0, 0, 0}, {0, 1, 1}, {0, 0, 1}, {1, 1, 1}},
        {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}}, {{0, 0, 0}, {1, 0, 1}, {1, 0, 0}, {1, 1, 1}},
    };
    const size_t nt = static_cast<size_t>(n) * n * n * 6;
    m.tets.reserve(nt);
    m.tet_tags.reserve(nt);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const bool conductor =
                    i >= spec.bar_i0 && i < spec.bar_i1 && j >= spec.bar_j0 && j < spec.bar_j1;
                for (int t = 0; t < 6; ++t) {
                    std::array<int, 4> v{};
                    for (int c = 0; c < 4; ++c) {
                        const int ii = i + kKuhn[t][c][0], jj = j + kKuhn[t][c][1], kk = k + kKuhn[t][c][2];
                        v[static_cast<size_t>(c)] = (kk * np + jj) * np + ii;
                    }
                    m.tets.push_back(v);
                    m.tet_tags.push_back(conductor ? 2 : 1);
                }
            }
    // Port terminals: the conductor's triangles on z = side (tag 10) and z = 0 (tag 11).
    const auto k_of = [np](int node) { return node / (np * np); };
    for (size_t t = 0; t < m.tets.size(); ++t) {
        if (m.tet_tags[t] != 2) continue;
        for (int lf = 0; lf < 4; ++lf) {
            std::array<int, 3> f{m.tets[t][static_cast<size_t>(kFaceVerts[lf][0])],
                                 m.tets[t][static_cast<size_t>(kFaceVerts[lf][1])],
                                 m.tets[t][static_cast<size_t>(kFaceVerts[lf][2])]};
            const int k0 = k_of(f[0]);
            if (k0 != k_of(f[1]) || k0 != k_of(f[2])) continue;
            if (k0 != n && k0 != 0) continue;
            std::sort(f.begin(), f.end());
            m.tagged_faces.push_back(f);
            m.tagged_face_tag.push_back(k0 == n ? 10 : 11);
        }
    }
    m.build_topology();
    return m;
}

void Mesh::build_topology() {
    const size_t nt = tets.size();
    if (tet_tags.size() != nt) tet_tags.assign(nt, -1);
    edges.clear();
    faces.clear();
    tet_edges.assign(nt, {});
    tet_edge_signs.assign(nt, {});
    tet_faces.assign(nt, {});

    // Edges: one (lo, hi, slot) record per local edge, sorted by key; the rank of
    // a distinct key is its global index.
    struct EdgeRec { int lo, hi, slot; };
    std::vector<EdgeRec> er(nt * 6);
    for (size_t t = 0; t < nt; ++t) {
        for (int le = 0; le < 6; ++le) {
            const int a = tets[t][static_cast<size_t>(kEdgeVerts[le][0])];
            const int b = tets[t][static_cast<size_t>(kEdgeVerts[le][1])];
            er[t * 6 + static_cast<size_t>(le)] = {std::min(a, b), std::max(a, b), static_cast<int>(t * 6) + le};
            tet_edge_signs[t][static_cast<size_t>(le)] = a < b ? int8_t(1) : int8_t(-1);
        }
    }
    std::sort(er.begin(), er.end(), [](const EdgeRec& x, const EdgeRec& y) {
        return x.lo != y.lo ? x.lo < y.lo : x.hi < y.hi;
    });
    for (size_t i = 0; i < er.size(); ++i) {
        if (i == 0 || er[i].lo != er[i - 1].lo || er[i].hi != er[i - 1].hi) edges.emplace_back(er[i].lo, er[i].hi);
        tet_edges[static_cast<size_t>(er[i].slot / 6)][static_cast<size_t>(er[i].slot % 6)] =
            static_cast<int>(edges.size()) - 1;
    }

    struct FaceRec { int a, b, c, slot; };
    std::vector<FaceRec> fr(nt * 4);
    for (size_t t = 0; t < nt; ++t) {
        for (int lf = 0; lf < 4; ++lf) {
            std::array<int, 3> k{tets[t][static_cast<size_t>(kFaceVerts[lf][0])],
                                 tets[t][static_cast<size_t>(kFaceVerts[lf][1])],
                                 tets[t][static_cast<size_t>(kFaceVerts[lf][2])]};
            std::sort(k.begin(), k.end());
            fr[t * 4 + static_cast<size_t>(lf)] = {k[0], k[1], k[2], static_cast<int>(t * 4) + lf};
        }
    }
    std::sort(fr.begin(), fr.end(), [](const FaceRec& x, const FaceRec& y) {
        if (x.a != y.a) return x.a < y.a;
        if (x.b != y.b) return x.b < y.b;
        return x.c < y.c;
    });
    for (size_t i = 0; i < fr.size(); ++i) {
        if (i == 0 || fr[i].a != fr[i - 1].a || fr[i].b != fr[i - 1].b || fr[i].c != fr[i - 1].c)
            faces.push_back({fr[i].a, fr[i].b, fr[i].c});
        tet_faces[static_cast<size_t>(fr[i].slot / 4)][static_cast<size_t>(fr[i].slot % 4)] =
            static_cast<int>(faces.size()) - 1;
    }

    face_tets.assign(faces.size(), {-1, -1});
    face_tet_count.assign(faces.size(), 0);
    for (size_t t = 0; t < nt; ++t) {
        for (int lf = 0; lf < 4; ++lf) {
            const size_t f = static_cast<size_t>(tet_faces[t][static_cast<size_t>(lf)]);
            if (face_tet_count[f] < 2) face_tets[f][static_cast<size_t>(face_tet_count[f])] = static_cast<int>(t);
            ++face_tet_count[f];
        }
    }
}

int Mesh::find_edge(int i, int j) const {
    const std::pair<int, int> key(std::min(i, j), std::max(i, j));
    const auto it = std::lower_bound(edges.begin(), edges.end(), key);
    return (it == edges.end() || *it != key) ? -1 : static_cast<int>(it - edges.begin());
}

int Mesh::find_face(int a, int b, int c) const {
    std::array<int, 3> key{a, b, c};
    std::sort(key.begin(), key.end());
    const auto it = std::lower_bound(faces.begin(), faces.end(), key);
    return (it == faces.end() || *it != key) ? -1 : static_cast<int>(it - faces.begin());
}

double Mesh::signed_tet_volume(int t) const {
    const auto& v = tets[static_cast<size_t>(t)];
    const Vec3& p0 = nodes[static_cast<size_t>(v[0])];
    const Vec3 a = nodes[static_cast<size_t>(v[1])] - p0;
    const Vec3 b = nodes[static_cast<size_t>(v[2])] - p0;
    const Vec3 c = nodes[static_cast<size_t>(v[3])] - p0;
    return a.cross(b).dot(c) / 6.0;
}

std::vector<double> expand_frequency_sweep(double start, double stop, int points, bool logarithmic) {
    if (points < 2) throw std::invalid_argument("expand_frequency_sweep: points must be >= 2");
    if (stop <= start) throw std::invalid_argument("expand_frequency_sweep: stop must exceed start");
    if (logarithmic && start <= 0.0)
        throw std::invalid_argument("expand_frequency_sweep: log spacing needs a positive start");
    std::vector<double> out(static_cast<size_t>(points));
    out.front() = start;
    out.back() = stop;
    const double steps = static_cast<double>(points - 1);
    if (logarithmic) {
        const double l0 = std::log(start);
        const double step = (std::log(stop) - l0) / steps;
        for (int i = 1; i < points - 1; ++i) out[static_cast<size_t>(i)] = std::exp(l0 + step * i);
    } else {
        const double step = (stop - start) / steps;
        for (int i = 1; i < points - 1; ++i) out[static_cast<size_t>(i)] = start + step * i;
    }
    return out;
}

}  // namespace aphi_amd

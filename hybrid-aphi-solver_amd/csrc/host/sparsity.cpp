// Symbolic assembly: the pattern, the scatter map, and the device assembly plan.
#include <algorithm>

#include "aphi_host.hpp"

namespace aphi_amd {

int SparsityPattern::find_slot(int row, int col) const {
    if (row < 0 || row >= rows) return -1;
    const auto b = col_index.begin() + row_ptr[static_cast<size_t>(row)];
    const auto e = col_index.begin() + row_ptr[static_cast<size_t>(row) + 1];
    const auto it = std::lower_bound(b, e, col);
    return (it == e || *it != col) ? -1 : static_cast<int>(it - col_index.begin());
}

namespace {

// Distinct live global DOFs of a tet, ascending (src/assembly.cpp:80-112,
// src/sparsity.cpp:31-46).
int live_dofs(const TetDofs& d, int* live) {
    int n = 0;
    for (int i = 0; i < 6; ++i)
        if (d.edge_index[i] >= 0) live[n++] = d.edge_index[i];
    for (int i = 0; i < 10; ++i)
        if (d.phi_index[i] >= 0) live[n++] = d.phi_index[i];
    std::sort(live, live + n);
    return static_cast<int>(std::unique(live, live + n) - live);
}

int position_of(const int* live, int n, int global) {
    return global < 0 ? -1 : static_cast<int>(std::lower_bound(live, live + n, global) - live);
}

}  // namespace

SparsityPattern build_sparsity(const DofMap& dofs, const BoundProblem& bound, const Mesh& mesh,
                               bool upper_only) {
    SparsityPattern p;
    p.rows = p.cols = dofs.num_total;
    p.upper_only = upper_only;
    const size_t n = static_cast<size_t>(p.rows);
    // Each row collects its candidate columns, then is sorted and uniqued; the
    // result is the ascending set of distinct columns, which is all the reference's
    // count/fill/compact passes define.
    std::vector<int> count(n + 1, 0);
    int live[16];
    const int nt = mesh.num_tets();
    for (int t = 0; t < nt; ++t) {
        const int m = live_dofs(dofs.local_dofs(t, mesh, bound), live);
        for (int i = 0; i < m; ++i) count[static_cast<size_t>(live[i]) + 1] += upper_only ? m - i : m;
    }
    for (size_t k = 0; k < dofs.port_is_fixed.size(); ++k)
        if (dofs.port_is_fixed[k]) count[static_cast<size_t>(dofs.port_index[k]) + 1] += 1;
    for (size_t r = 0; r < n; ++r) count[r + 1] += count[r];
    std::vector<int> col(static_cast<size_t>(count[n]));
    std::vector<int> cur(count.begin(), count.end() - 1);
    for (int t = 0; t < nt; ++t) {
        const int m = live_dofs(dofs.local_dofs(t, mesh, bound), live);
        for (int i = 0; i < m; ++i)
            for (int j = upper_only ? i : 0; j < m; ++j) col[static_cast<size_t>(cur[static_cast<size_t>(live[i])]++)] = live[j];
    }
    for (size_t k = 0; k < dofs.port_is_fixed.size(); ++k) {
        if (!dofs.port_is_fixed[k]) continue;
        const int r = dofs.port_index[k];
        col[static_cast<size_t>(cur[static_cast<size_t>(r)]++)] = r;
    }
    p.row_ptr.assign(n + 1, 0);
    int w = 0;
    for (size_t r = 0; r < n; ++r) {
        auto b = col.begin() + count[r];
        auto e = col.begin() + cur[r];
        std::sort(b, e);
        auto last = std::unique(b, e);
        p.row_ptr[r] = w;
        for (auto it = b; it != last; ++it) col[static_cast<size_t>(w++)] = *it;
    }
    p.row_ptr[n] = w;
    col.resize(static_cast<size_t>(w));
    p.col_index = std::move(col);
    return p;
}

ScatterMap build_scatter_map(const SparsityPattern& pattern, const DofMap& dofs, const Mesh& mesh,
                             const BoundProblem& bound) {
    return build_scatter_map(pattern, mesh.num_tets(), [&](int t) { return dofs.local_dofs(t, mesh, bound); });
}

ScatterMap build_scatter_map(const SparsityPattern& pattern, int nt, const TetDofsOf& local) {
    ScatterMap m;
    m.packed = pattern.upper_only;
    m.live.assign(static_cast<size_t>(nt), 0);
    m.pos.assign(static_cast<size_t>(nt) * 16, -1);
    m.offset.assign(static_cast<size_t>(nt) + 1, 0);
    std::vector<int> lives(static_cast<size_t>(nt) * 16);
    for (int t = 0; t < nt; ++t) {
        const TetDofs d = local(t);
        int* live = &lives[static_cast<size_t>(t) * 16];
        const int n = live_dofs(d, live);
        m.live[static_cast<size_t>(t)] = n;
        m.offset[static_cast<size_t>(t) + 1] = m.offset[static_cast<size_t>(t)] + tet_block_size(n, m.packed);
        for (int i = 0; i < 6; ++i) m.pos[static_cast<size_t>(t) * 16 + i] = position_of(live, n, d.edge_index[i]);
        for (int i = 0; i < 10; ++i) m.pos[static_cast<size_t>(t) * 16 + 6 + i] = position_of(live, n, d.phi_index[i]);
    }
    m.slot.assign(static_cast<size_t>(m.offset[static_cast<size_t>(nt)]), -1);
    for (int t = 0; t < nt; ++t) {
        const int* live = &lives[static_cast<size_t>(t) * 16];
        const int n = m.live[static_cast<size_t>(t)];
        int* out = &m.slot[static_cast<size_t>(m.offset[static_cast<size_t>(t)])];
        for (int r = 0; r < n; ++r) {
            const int row = live[r];
            int k = pattern.row_ptr[static_cast<size_t>(row)];
            const int end = pattern.row_ptr[static_cast<size_t>(row) + 1];
            for (int c = m.packed ? r : 0; c < n; ++c) {
                const int want = live[c];
                while (k < end && pattern.col_index[static_cast<size_t>(k)] < want) ++k;
                if (k == end || pattern.col_index[static_cast<size_t>(k)] != want)
                    throw std::logic_error("assemble: no slot for entry (" + std::to_string(row) + ", " +
                                           std::to_string(want) + "). The sparsity pattern and the scatter "
                                           "disagree about the matrix's shape.");
                out[tet_block_index(n, m.packed, r, c)] = k;
            }
        }
    }
    return m;
}

AssemblyPlan build_assembly_plan(const SparsityPattern& pattern, const ScatterMap& map,
                                 const DofMap& dofs, const Mesh& mesh, const BoundProblem& bound) {
    return build_assembly_plan(pattern, map, mesh.num_tets(), [&](int t) { return dofs.local_dofs(t, mesh, bound); },
                               dofs, bound);
}

AssemblyPlan build_assembly_plan(const SparsityPattern& pattern, const ScatterMap& map, int num_tets,
                                 const TetDofsOf& local, const DofMap& dofs, const BoundProblem& bound) {
    if (!pattern.upper_only || !map.packed)
        throw std::invalid_argument("build_assembly_plan: the device path stores the upper triangle only");
    if (static_cast<int>(map.live.size()) != num_tets)
        throw std::invalid_argument("build_assembly_plan: the ScatterMap was built for a different mesh");
    AssemblyPlan plan;
    plan.num_slots = static_cast<int>(pattern.nnz());
    plan.num_rows = dofs.num_total;
    const size_t ntarget = static_cast<size_t>(plan.num_slots) + static_cast<size_t>(plan.num_rows);
    plan.tet_body.assign(static_cast<size_t>(num_tets), -1);
    for (size_t b = 0; b < bound.bodies.size(); ++b)
        for (int t : bound.bodies[b].tets) plan.tet_body[static_cast<size_t>(t)] = static_cast<int>(b);

    // Prescribed values are few (one per voltage port); each gets a small id.
    const auto value_id = [&](const Complex& v) {
        for (size_t i = 0; i < plan.fixed_value.size(); ++i)
            if (plan.fixed_value[i] == v) return static_cast<int>(i);
        if (plan.fixed_value.size() >= 8)
            throw std::invalid_argument("build_assembly_plan: more than 8 distinct prescribed values");
        plan.fixed_value.push_back(v);
        return static_cast<int>(plan.fixed_value.size()) - 1;
    };

    std::vector<int64_t> count(ntarget + 1, 0);
    bool filling = false;
    std::vector<int64_t> cur;
    const auto emit = [&](int64_t target, uint32_t entry, int kind, double sign, int vid) {
        if (!filling) {
            ++count[static_cast<size_t>(target) + 1];
            return;
        }
        const size_t at = static_cast<size_t>(cur[static_cast<size_t>(target)]++);
        plan.entry[at] = entry;
        plan.kind[at] = static_cast<uint8_t>(kind | (sign < 0.0 ? 0x10 : 0) | (vid << 5));
    };

    const auto walk = [&]() {
        // Bodies, then each body's tets ascending: src/assembly.cpp:277-280.
        for (const Body& body : bound.bodies) {
            for (int t : body.tets) {
                const TetDofs d = local(t);
                const int* pos = &map.pos[static_cast<size_t>(t) * 16];
                const uint32_t base = static_cast<uint32_t>(t) * 196u;
                // (A, A): :300-315
                for (int p = 0; p < 6; ++p) {
                    if (d.edge_index[p] < 0) continue;
                    for (int q = 0; q < 6; ++q) {
                        if (d.edge_index[q] < 0) continue;
                        const int slot = map.slot_of(t, pos[p], pos[q]);
                        if (slot < 0) continue;
                        emit(slot, base + static_cast<uint32_t>(p * 6 + q), kAA, d.edge_coeff[p] * d.edge_coeff[q], 0);
                    }
                }
                // (A, Phi) and (Phi, A): :322-364
                for (int p = 0; p < 6; ++p) {
                    if (d.edge_index[p] < 0) continue;
                    for (int q = 0; q < 10; ++q) {
                        const uint32_t ent = base + 36u + static_cast<uint32_t>(p * 10 + q);
                        if (d.phi_index[q] >= 0) {
                            const int slot = map.slot_of(t, pos[p], pos[6 + q]);
                            if (slot >= 0) emit(slot, ent, kAPhi, d.edge_coeff[p] * d.phi_coeff[q], 0);
                        } else if (d.phi_fixed[q] != Complex(0.0, 0.0)) {
                            emit(static_cast<int64_t>(plan.num_slots) + d.edge_index[p], ent, kRhsAPhi,
                                 d.edge_coeff[p], value_id(d.phi_fixed[q]));
                        }
                        if (d.phi_index[q] >= 0) {
                            const int slot = map.slot_of(t, pos[6 + q], pos[p]);
                            if (slot >= 0) emit(slot, ent, kPhiA, d.phi_coeff[q] * d.edge_coeff[p], 0);
                        }
                    }
                }
                // (Phi, Phi): :367-388
                for (int p = 0; p < 10; ++p) {
                    if (d.phi_index[p] < 0) continue;
                    for (int q = 0; q < 10; ++q) {
                        const uint32_t ent = base + 96u + static_cast<uint32_t>(p * 10 + q);
                        if (d.phi_index[q] >= 0) {
                            const int slot = map.slot_of(t, pos[6 + p], pos[6 + q]);
                            if (slot >= 0) emit(slot, ent, kPhiPhi, d.phi_coeff[p] * d.phi_coeff[q], 0);
                        } else if (d.phi_fixed[q] != Complex(0.0, 0.0)) {
                            emit(static_cast<int64_t>(plan.num_slots) + d.phi_index[p], ent, kRhsPhiPhi,
                                 d.phi_coeff[p], value_id(d.phi_fixed[q]));
                        }
                    }
                }
            }
        }
        // Port rows: :396-421.
        for (size_t k = 0; k < bound.ports.size(); ++k) {
            const int row = dofs.port_index[k];
            if (is_current_driven(bound.ports[k].type)) {
                emit(static_cast<int64_t>(plan.num_slots) + row, 0, kPortCurrent, 1.0, value_id(bound.ports[k].amplitude));
            } else {
                const int slot = pattern.find_slot(row, row);
                if (slot < 0) throw std::logic_error("assemble: no slot for a voltage port's constraint diagonal");
                emit(slot, 0, kPortDiag, 1.0, 0);
                emit(static_cast<int64_t>(plan.num_slots) + row, 0, kPortVoltage, 1.0, value_id(dofs.port_value[k]));
            }
        }
    };

    walk();
    for (size_t i = 0; i < ntarget; ++i) count[i + 1] += count[i];
    plan.target_ptr = count;
    plan.entry.assign(static_cast<size_t>(count[ntarget]), 0);
    plan.kind.assign(static_cast<size_t>(count[ntarget]), 0);
    cur.assign(count.begin(), count.end() - 1);
    filling = true;
    plan.fixed_value.clear();
    walk();
    return plan;
}

}  // namespace aphi_amd

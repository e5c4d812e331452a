// C-ABI over the device path: context and stream, per-mesh plan upload, refill,
// equilibration, level-scheduled supernodal factorization, multi-RHS solves and
// FP64 iterative refinement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

#include "../device/factor.hpp"
#include "../device/philox.hpp"
#include "status.hpp"

using aphi_capi::DeviceError;
using aphi_capi::PivotFailure;
using aphi_dev::cd;
using Complex = std::complex<double>;

namespace {

constexpr double kMu0 = 4.0e-7 * 3.14159265358979323846;
constexpr double kEps0 = 8.8541878128e-12;
constexpr int kSmallMaxLimit = 96;  // LDS bound: fronts up to this many rows fit one workgroup
// the solve's forward gather reads extend-add bounds for fronts of more than one
// chunk of rows; those exist for every front that is not small
static_assert(kSmallMaxLimit < aphi_dev::kGatherRows, "gather chunks need extend-add bounds");

// Fronts up to this many rows are factored in LDS by one workgroup; the rest take
// the blocked path. APHI_SMALL_FRONT_MAX (0..96) overrides it, for tuning.
int small_front_max() {
    const char* e = std::getenv("APHI_SMALL_FRONT_MAX");
    if (!e) return kSmallMaxLimit;
    const int v = std::atoi(e);
    return v < 0 ? 0 : (v > kSmallMaxLimit ? kSmallMaxLimit : v);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw DeviceError(std::string(what) + ": " + hipGetErrorString(e), e == hipErrorOutOfMemory ? 2 : 1);
}
#define HIPCHK(x) hip_check((x), #x)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void reserve(size_t b) {
        if (b <= bytes && p) return;
        release();
        if (b == 0) b = 256;
        HIPCHK(hipMalloc(&p, b));
        bytes = b;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// host -> device on the library's own stream (never the legacy default stream,
// whose implicit ordering would tie it to other threads' streams), complete on return
template <typename T>
void upload(hipStream_t s, DevBuf& b, const T* src, size_t n) {
    b.reserve(n * sizeof(T));
    if (!n) return;
    HIPCHK(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
}

size_t csize(int precision) { return precision == APHI_FP64 ? sizeof(cd) : sizeof(aphi_dev::cf); }

struct Level {
    int small_count = 0, small_max_m = 0, large_count = 0, ea_blocks = 0, large_max_m = 0;
    int64_t off = 0;       // first node of this level in the node list: small ones, then large ones
    int64_t ea_off = 0;    // this level's extend-add block starts
    std::vector<int> m, k; // per node (small then large)
    int count() const { return small_count + large_count; }
};

}  // namespace

// Scratch of aphi_terminal_current_noalloc: the per-tet, per-right-hand-side
// current terms (num_tets * nrhs entries) and, after them, the nrhs sums. One copy
// per device, held until the process ends; every call on a device shares it.
namespace {
constexpr int64_t kNoAllocTerms = int64_t(1) << 21;  // 32 MiB of complex128 terms
constexpr int kNoAllocSums = 64;                      // nrhs <= 64
}  // namespace
static __device__ __align__(256) aphi_dev::cd scratch[kNoAllocTerms + kNoAllocSums];

struct aphi_context {
    int device = 0;
    aphi_dev::cd* scratch = nullptr;  // this device's address of `scratch` (hipGetSymbolAddress)
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // second stream of the factorization (default priority)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t main_done = nullptr, side_done = nullptr;  // ordering only (no timing)
    hipEvent_t caller = nullptr;  // the caller's legacy-stream work, joined at entry
};

// The solve's per-level launches (about 1,200 at 465k unknowns) captured once
// and replayed: they touch only the mesh's own buffers, whose addresses form the
// key, and depend on the right-hand sides only through the group count.
struct SolveGraph {
    hipGraphExec_t exec = nullptr;
    int precision = -1, groups = 0;
    std::vector<const void*> key;
};

struct aphi_mesh {
    aphi_context* ctx = nullptr;
    int n = 0, num_nodes = 0, num_tets = 0, num_bodies = 0, num_super = 0, num_levels = 0, num_values = 0;
    int64_t nnz = 0, num_targets = 0, rows_total = 0, factor_entries = 0, cb_entries = 0, predicted_nnz = 0;
    double flops = 0.0;
    std::vector<double> sigma, eps_r, mu_r;
    std::vector<Level> levels;
    int max_level_count = 0;
    // count * rect_chunks * rect_cols and count * rect_cols over levels (the
    // backward sums of the rows below the pivots, per padded nrhs)
    int64_t max_rect_partials = 0, max_rect_sums = 0;
    int64_t max_fwd_partials = 0;  // count * forward splits * rows below the pivots over levels, per padded nrhs
    // plan on the device
    DevBuf coords, tets, tet_body, tptr, entry, kind, row_ptr, col, row_of, ttptr, tslot, trow, perm, col_ref;
    DevBuf first, srow_ptr, srows, rel, child_ptr, child, panel_off, cb_off, s2p, nodes, ea_start, xinv_off;
    DevBuf ea_bound_off, ea_bound, xsb_off, inv_pairs;
    DevBuf woff;            // per supernode: its W window's offset within the level (-1: none)
    int64_t lw_entries = 0;  // the zero column and the largest level's W windows (complex float entries)
    int64_t lw_zero = 0;     // entries of the zero column at the start of the W buffer
    DevBuf tdofs, tsign, tcoef, tfixed, bsigma, tcontrib, tcurrent;  // port-current recovery
    // compute_fields: the node plan (aphi_fields_attach) and its work arrays
    int fields_np = 0;
    DevBuf fptr, fentry, bmu, fval, fvol, fmu, fsig, fa_tet, fb_tet, fe_tet, fj_tet;
    DevBuf fa_node, fb_node, fh_node, fe_node, fj_node, fweight, fflag;
    int64_t xsb_entries = 0;
    int num_inv_pairs = 0;
    int64_t xinv_entries = 0;
    // numeric state
    DevBuf coef, E, flag, A, rhs, Aeq, work, rowmax, fac, d, partials, stats;
    DevBuf P, CB, mws, Xinv, Xsb, XTsb, dinv;
    DevBuf LW;  // complex-float factorization: W = -L D windows of the current level (FrontPlan::woff)
    DevBuf bt, W, X, Yv, rpart, trect, fpart, colscale, res, norms, xil;
    DevBuf gV, gZ, gW, gU, gDots, gSmall;  // GMRES refinement
    int solve_nrhs_cap = 0;
    SolveGraph sgraph;
    // The factorization's launches (both streams, from the panel scatter to the
    // inverse diagonal) captured once and replayed at every frequency: they
    // depend on the values only through buffers whose addresses form the key.
    struct FactorGraph {
        hipGraphExec_t exec = nullptr;
        int precision = -1;
        double pivot_floor = 0.0;
        int64_t tiles_version = -1;
        std::vector<const void*> key;
        int64_t launches = 0;
        double update_flops = 0.0;
    } fgraph;
    // Tiling of the complex-float trailing update per launch class (tallest
    // region by bucket x widest region by bucket): 0 = the default rule, 1 = 64 x 64
    // tiles, 2 = 128 x 128 tiles (both give the same bits); set by the autotuner.
    static constexpr int kHeightClasses = 8, kWidthClasses = 8;
    int8_t tile_table[kHeightClasses][kWidthClasses] = {};
    int64_t tiles_version = 0;
    bool assembled = false, equilibrated = false, factored = false;
    bool has_plan = true;  // false for a problem made from a bare pattern (no refill)
    bool solve_graphs_on = true;  // aphi_mesh_set_solve_graphs: graph replay of the solve
    bool graphs_on = true;        // and of the factorization
    int update_tiles = 0;         // aphi_mesh_set_update_tiles: 0 by region height, 1 small, 2 large
    // gauge-free potential plan (aphi_gauge_free_attach), on a bare-pattern mesh
    struct GaugeFree {
        bool attached = false;
        int num_nodes = 0, num_tets = 0, num_bodies = 0;
        int64_t num_targets = 0;
        std::vector<double> sigma, eps_r;
        std::vector<int> port_column;
        std::vector<Complex> port_value;
        DevBuf coords, tets, tet_body, tet_on, column, prescribed, tptr, entry, kind, values, beta, terms;
    } gf;
    // The problems a mesh holds side by side: one for a mesh of one problem, more
    // for a batch (aphi_batch_create). Problem s owns unknowns [rows[s], rows[s+1]),
    // stored entries [slots[s], slots[s+1]), tets [tets[s], tets[s+1]) and bodies
    // [bodies[s], bodies[s+1]); its matrix is its own diagonal block.
    struct Segments {
        int count = 1;
        std::vector<int64_t> rows, slots, tets, bodies;
        std::vector<std::vector<Complex>> values;  // prescribed values of each problem
        DevBuf d_rows, d_slots, d_tets, row_seg, scales;
    } seg;
    int factor_precision = -1;
    std::vector<double> amax, amax_eq;  // max |A_ij| and max |(D A D)_ij| per problem
    double eq_min = 1.0, eq_max = 1.0;
    aphi_factor_stats last_stats{};
    bool profiling = false;
    bool serial_updates = false;  // profiling with the look-ahead's updates on the main stream
    std::vector<hipEvent_t> prof_events;

    aphi_dev::FrontPlan plan() const {
        aphi_dev::FrontPlan fp;
        fp.first = first.as<int>();
        fp.row_ptr = srow_ptr.as<int64_t>();
        fp.rows = srows.as<int>();
        fp.rel = rel.as<int>();
        fp.child_ptr = child_ptr.as<int>();
        fp.child = child.as<int>();
        fp.panel_off = panel_off.as<int64_t>();
        fp.cb_off = cb_off.as<int64_t>();
        fp.col_ref = col_ref.as<int>();
        fp.xinv_off = xinv_off.as<int64_t>();
        fp.ea_bound_off = ea_bound_off.as<int64_t>();
        fp.ea_bound = ea_bound.as<int>();
        fp.xsb_off = xsb_off.as<int64_t>();
        fp.woff = woff.as<int64_t>();
        return fp;
    }
    hipStream_t st() const { return ctx->stream; }
};

namespace {

unsigned blocks_for(int64_t n, int per = 256) { return static_cast<unsigned>((n + per - 1) / per); }

void require(bool ok, const char* msg) {
    if (!ok) throw std::invalid_argument(msg);
}

// ------------------------------------------------------------------ numerics helpers

// max |v_k| over each problem's stored entries of v (nnz entries)
std::vector<double> segment_absmax(aphi_mesh* m, const cd* v) {
    const int parts = m->seg.count == 1 ? 256 : 8;
    const int ns = m->seg.count;
    m->partials.reserve(std::max<size_t>(m->partials.bytes, static_cast<size_t>(parts) * ns * sizeof(double)));
    hipLaunchKernelGGL(aphi_dev::k_seg_absmax, dim3(parts, ns), dim3(256), 0, m->st(), m->seg.d_slots.as<int64_t>(), v,
                       m->partials.as<double>());
    std::vector<double> h(static_cast<size_t>(parts) * ns), out(static_cast<size_t>(ns), 0.0);
    HIPCHK(hipMemcpyAsync(h.data(), m->partials.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, m->st()));
    HIPCHK(hipStreamSynchronize(m->st()));
    for (int sg = 0; sg < ns; ++sg)
        for (int p = 0; p < parts; ++p) out[static_cast<size_t>(sg)] = std::max(out[static_cast<size_t>(sg)], h[static_cast<size_t>(sg) * parts + p]);
    return out;
}

void compute_amax(aphi_mesh* m) { m->amax = segment_absmax(m, m->A.as<cd>()); }

void equilibrate(aphi_mesh* m, int iterations, double* d_out) {
    require(m->assembled, "equilibrate: the system has not been assembled (call aphi_refill or aphi_system_upload)");
    require(iterations >= 0, "equilibrate: iterations must be >= 0");
    hipStream_t s = m->st();
    const int n = m->n;
    m->d.reserve(n * sizeof(double));
    m->Aeq.reserve(m->nnz * sizeof(cd));
    hipLaunchKernelGGL(aphi_dev::k_fill_d, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<int64_t>(n), m->d.as<double>(), 1.0);
    if (iterations > 0) {
        m->work.reserve(m->nnz * sizeof(cd));
        m->rowmax.reserve(n * sizeof(double));
        m->fac.reserve(n * sizeof(double));
        hipLaunchKernelGGL(aphi_dev::k_copy_cd, dim3(blocks_for(m->nnz)), dim3(256), 0, s, m->nnz, m->A.as<cd>(), m->work.as<cd>());
        hipLaunchKernelGGL(aphi_dev::k_fill_d, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<int64_t>(n), m->rowmax.as<double>(), 0.0);
        for (int it = 0; it < iterations; ++it) {
            hipLaunchKernelGGL(aphi_dev::k_eq_rowmax16, dim3(blocks_for(n, 16)), dim3(256), 0, s, n, m->row_ptr.as<int>(),
                               m->ttptr.as<int>(), m->tslot.as<int>(), m->work.as<cd>(), m->rowmax.as<double>());
            hipLaunchKernelGGL(aphi_dev::k_eq_factor, dim3(blocks_for(n)), dim3(256), 0, s, n, m->rowmax.as<double>(),
                               m->fac.as<double>(), m->d.as<double>());
            hipLaunchKernelGGL(aphi_dev::k_eq_scale, dim3(blocks_for(m->nnz)), dim3(256), 0, s, m->nnz, m->row_of.as<int>(),
                               m->col.as<int>(), m->fac.as<double>(), m->work.as<cd>());
        }
        hipLaunchKernelGGL(aphi_dev::k_eq_apply, dim3(blocks_for(m->nnz)), dim3(256), 0, s, m->nnz, m->row_of.as<int>(),
                           m->col.as<int>(), m->d.as<double>(), m->A.as<cd>(), m->Aeq.as<cd>());
    } else {
        hipLaunchKernelGGL(aphi_dev::k_copy_cd, dim3(blocks_for(m->nnz)), dim3(256), 0, s, m->nnz, m->A.as<cd>(), m->Aeq.as<cd>());
    }
    std::vector<double> h(static_cast<size_t>(n));
    HIPCHK(hipMemcpyAsync(h.data(), m->d.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    m->eq_min = n ? *std::min_element(h.begin(), h.end()) : 1.0;
    m->eq_max = n ? *std::max_element(h.begin(), h.end()) : 1.0;
    if (d_out) std::memcpy(d_out, h.data(), h.size() * sizeof(double));
    m->amax_eq = segment_absmax(m, m->Aeq.as<cd>());
    m->equilibrated = true;
    m->factored = false;
}

// Launch class of a trailing update for the tile table: its tallest region's
// rows and its widest region's columns, each by power-of-two bucket.
int height_class(int rows) {
    int c = 0;
    for (int v = 1024; c < aphi_mesh::kHeightClasses - 1 && rows > v; v *= 2) ++c;  // <=1024, <=2048, ..
    return c;
}
int width_class(int cols) {
    int c = 0;
    for (int v = 64; c < aphi_mesh::kWidthClasses - 1 && cols > v; v *= 2) ++c;  // <=64, <=128, ..
    return c;
}

// APHI_FACTOR_GRAPHS=0 launches the factorization one kernel at a time instead of
// replaying its captured graph
bool factor_graphs() {
    const char* e = std::getenv("APHI_FACTOR_GRAPHS");
    return !(e && e[0] == '0');
}

// Per-launch record of a profiled factorization (aphi_mesh_tune_update_tiles).
struct UpdateLaunch {
    int hclass, wclass;
    bool small;
    double flops;
    size_t ev;        // index of its start event; the stop event follows
    double ms = 0.0;  // measured duration
};

// The factorization's launches on the mesh's two streams, from the statistics
// reset to the inverse diagonal. Returns the number of trailing-update launches
// and their flops; records them in `log` when events are taken.
template <typename T>
int64_t factor_sequence(aphi_mesh* m, double pivot_floor, double& update_flops, std::vector<UpdateLaunch>* log,
                        size_t& ev_used) {
    hipStream_t s = m->st();
    const aphi_dev::FrontPlan fp = m->plan();
    aphi_dev::launch_stats_init(s, m->stats.as<aphi_dev::FactorStatsDev>());
    aphi_dev::launch_zero<T>(s, m->factor_entries + aphi_dev::kPanelTail, m->P.p);
    aphi_dev::launch_scatter_panel<T>(s, m->nnz, m->s2p.as<int64_t>(), m->Aeq.as<cd>(), m->P.p);
    const int* nodes = m->nodes.as<int>();
    auto* st = m->stats.as<aphi_dev::FactorStatsDev>();
    // complex float: the panels also write W = -L D into the level's W windows,
    // which the trailing updates read as their second operand
    void* wbuf = std::is_same<T, float>::value ? m->LW.p : nullptr;
    int64_t launches = 0;
    update_flops = 0.0;
    const auto next_event = [&]() {
        if (ev_used == m->prof_events.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            m->prof_events.push_back(e);
        }
        return m->prof_events[ev_used++];
    };
    for (const Level& L : m->levels) {
        if (L.small_count)
            aphi_dev::launch_front_small<T>(s, nodes + L.off, L.small_count, L.small_max_m, fp, m->P.p, m->CB.p, m->Xinv.p,
                                            pivot_floor, st);
        if (!L.large_count) continue;
        const int* lg = nodes + L.off + L.small_count;
        aphi_dev::launch_extend_add<T>(s, lg, m->ea_start.as<int>() + L.ea_off, L.large_count, L.ea_blocks,
                                       L.large_max_m, fp, m->P.p, m->CB.p);
        int maxk = 0;
        for (int i = 0; i < L.large_count; ++i) maxk = std::max(maxk, L.k[static_cast<size_t>(L.small_count + i)]);
        // one update launch on stream `on`: the largest tile count over the level's
        // fronts, and its flops (region semantics of launch_update); complex float
        // picks its tiles by the launch's class in the tile table, or by default
        // the small tiles when the tallest region is short or the widest is narrow
        // (the main stream's updates, on the block chain)
        const auto update = [&](hipStream_t on, int c0, int width, int tlo, int thi) {
            const auto region = [&](int i, int& t0, int& tend) {
                const int mm = L.m[static_cast<size_t>(L.small_count + i)], kk = L.k[static_cast<size_t>(L.small_count + i)];
                if (kk <= c0) return false;
                const int c1 = std::min(c0 + width, kk);
                t0 = std::max(c1, std::min(tlo, kk));
                tend = thi < 0 ? mm : std::min(thi, kk);
                return tend > t0;
            };
            int hmax = 0, wmax = 0;
            for (int i = 0; i < L.large_count; ++i) {
                int t0, tend;
                if (region(i, t0, tend)) {
                    hmax = std::max(hmax, L.m[static_cast<size_t>(L.small_count + i)] - t0);
                    wmax = std::max(wmax, tend - t0);
                }
            }
            if (hmax == 0) return;
            const int hc = height_class(hmax), wc = width_class(wmax);
            bool small = false;
            if (std::is_same<T, float>::value) {
                const int table = m->tile_table[hc][wc];
                const int mode = m->update_tiles != 0 ? m->update_tiles : table;
                small = mode == 0 ? (hmax <= aphi_dev::kSmallTileRows || wmax <= aphi_dev::kSmallTileCols) : mode == 1;
            }
            const int ts = aphi_dev::update_tile<T>(small);
            int tiles = 0;
            double fl = 0.0;
            for (int i = 0; i < L.large_count; ++i) {
                int t0, tend;
                if (!region(i, t0, tend)) continue;
                const int mm = L.m[static_cast<size_t>(L.small_count + i)], kk = L.k[static_cast<size_t>(L.small_count + i)];
                const int c1 = std::min(c0 + width, kk);
                const int nrt = (mm - t0 + ts - 1) / ts, nct = (tend - t0 + ts - 1) / ts;
                tiles = std::max(tiles, nct * nrt - nct * (nct - 1) / 2);
                const double kw = c1 - c0, w = tend - t0, h = mm - t0;
                fl += 8.0 * kw * (w * h - w * (w - 1.0) / 2.0);
            }
            if (log) {
                log->push_back(UpdateLaunch{hc, wc, small, fl, ev_used, 0.0});
                HIPCHK(hipEventRecord(next_event(), on));
            }
            aphi_dev::launch_update<T>(on, lg, L.large_count, c0, width, tlo, thi, tiles, 0, fp, m->P.p, m->CB.p, wbuf,
                                       small);
            if (log) HIPCHK(hipEventRecord(next_event(), on));
            ++launches;
            update_flops += fl;
        };
        // Outer blocks of kOuter columns. Inside a block, per kNB columns: the
        // diagonal block, the panel, and a narrow update of the block's remaining
        // columns. After a block, its rank-kOuter update is split: the next
        // block's columns on the main stream, the rest of the trailing front on
        // the side stream, so that the next block's factorization (main) runs
        // while the bulk of the update (side) does. The side stream's update starts
        // as soon as the block's panel is final; the next block's part waits for
        // the side stream's previous update, whose targets overlap it.
        const int nouter = (maxk + aphi_dev::kOuter - 1) / aphi_dev::kOuter;
        bool side_busy = false;
        for (int ob = 0; ob < nouter; ++ob) {
            const int bo = ob * aphi_dev::kOuter;
            const int be = std::min(bo + aphi_dev::kOuter, maxk);
            for (int b = bo; b < be; b += aphi_dev::kNB) {
                int rb = 0;
                for (int i = 0; i < L.large_count; ++i) {
                    const int mm = L.m[static_cast<size_t>(L.small_count + i)], kk = L.k[static_cast<size_t>(L.small_count + i)];
                    if (kk <= b) continue;
                    const int rem = mm - b - std::min(aphi_dev::kNB, kk - b);
                    if (rem > 0) rb = std::max(rb, (rem + 63) / 64);
                }
                aphi_dev::launch_block_step<T>(s, lg, L.large_count, b, rb, 0, fp, m->P.p, m->CB.p, m->mws.p, m->Xinv.p,
                                               pivot_floor, st, wbuf);
                update(s, b, aphi_dev::kNB, 0, bo + aphi_dev::kOuter);
            }
            // the previous block's rest update must be done before this block's
            // next-block part (they share the columns [bo + kOuter, bo + 2 kOuter))
            if (side_busy) HIPCHK(hipStreamWaitEvent(s, m->ctx->side_done, 0));
            if (ob + 1 < nouter && log && m->serial_updates) {
                // timed alone: the same two launches, one after the other
                update(s, bo, aphi_dev::kOuter, bo + 2 * aphi_dev::kOuter, -1);
                update(s, bo, aphi_dev::kOuter, 0, bo + 2 * aphi_dev::kOuter);
                side_busy = false;
            } else if (ob + 1 < nouter) {
                // the rest only needs this block's panel (final here): it starts now,
                // beside the next-block part, whose columns it does not touch
                HIPCHK(hipEventRecord(m->ctx->main_done, s));
                HIPCHK(hipStreamWaitEvent(m->ctx->side, m->ctx->main_done, 0));
                update(m->ctx->side, bo, aphi_dev::kOuter, bo + 2 * aphi_dev::kOuter, -1);
                HIPCHK(hipEventRecord(m->ctx->side_done, m->ctx->side));
                update(s, bo, aphi_dev::kOuter, 0, bo + 2 * aphi_dev::kOuter);
                side_busy = true;
            } else {
                update(s, bo, aphi_dev::kOuter, 0, -1);
                side_busy = false;
            }
        }
    }
    aphi_dev::launch_inv_blocks<T>(s, m->inv_pairs.as<int2>(), m->num_inv_pairs, fp, m->P.p, m->Xinv.p, m->Xsb.p,
                                   m->XTsb.p);
    aphi_dev::launch_dinv<T>(s, m->num_super, fp, m->P.p, m->dinv.p);
    return launches;
}

template <typename T>
void factorize_t(aphi_mesh* m, double pivot_floor, aphi_factor_stats* out, std::vector<UpdateLaunch>* log = nullptr) {
    hipStream_t s = m->st();
    const size_t cs = sizeof(typename aphi_dev::cplx<T>::type);
    m->P.reserve(static_cast<size_t>(m->factor_entries + aphi_dev::kPanelTail) * cs);
    m->CB.reserve(static_cast<size_t>(std::max<int64_t>(m->cb_entries, 1)) * cs);
    m->mws.reserve(static_cast<size_t>(std::max(m->max_level_count, 1)) * aphi_dev::kNB * aphi_dev::kNB * cs);
    m->Xinv.reserve(static_cast<size_t>(std::max<int64_t>(m->xinv_entries, 1)) * cs);
    m->dinv.reserve(static_cast<size_t>(m->n) * cs);
    m->Xsb.reserve(static_cast<size_t>(std::max<int64_t>(m->xsb_entries, 1)) * cs);
    m->XTsb.reserve(static_cast<size_t>(std::max<int64_t>(m->xsb_entries, 1)) * cs);
    m->stats.reserve(sizeof(aphi_dev::FactorStatsDev));
    if (std::is_same<T, float>::value) {
        m->LW.reserve(static_cast<size_t>(std::max<int64_t>(m->lw_entries, 1) + aphi_dev::kPanelTail) * cs);
        // the zero column (never written by the factorization; cleared outside any capture)
        HIPCHK(hipMemsetAsync(m->LW.p, 0, static_cast<size_t>(m->lw_zero) * cs, s));
    }
    int64_t launches = 0;
    double update_flops = 0.0;
    size_t ev_used = 0;
    const bool timed = m->profiling || log != nullptr;
    std::vector<UpdateLaunch> own;
    std::vector<UpdateLaunch>* rec = log ? log : (m->profiling ? &own : nullptr);
    HIPCHK(hipEventRecord(m->ctx->ev0, s));
    if (!timed && factor_graphs() && m->graphs_on) {
        const int prec = std::is_same<T, float>::value ? APHI_FP32 : APHI_FP64;
        const std::vector<const void*> key = {m->P.p,     m->CB.p,    m->mws.p, m->Xinv.p, m->dinv.p, m->Xsb.p,
                                              m->XTsb.p, m->stats.p, m->Aeq.p, m->s2p.p,  m->nodes.p, m->LW.p};
        auto& g = m->fgraph;
        if (!(g.exec && g.precision == prec && g.pivot_floor == pivot_floor && g.key == key &&
              g.tiles_version == m->tiles_version)) {
            if (g.exec) HIPCHK(hipGraphExecDestroy(g.exec));
            g.exec = nullptr;
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
            try {
                g.launches = factor_sequence<T>(m, pivot_floor, g.update_flops, nullptr, ev_used);
            } catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            HIPCHK(hipStreamEndCapture(s, &graph));
            const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCHK(ie);
            g.precision = prec;
            g.pivot_floor = pivot_floor;
            g.key = key;
            g.tiles_version = m->tiles_version;
        }
        HIPCHK(hipGraphLaunch(g.exec, s));
        launches = g.launches;
        update_flops = g.update_flops;
    } else {
        launches = factor_sequence<T>(m, pivot_floor, update_flops, rec, ev_used);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(m->ctx->ev1, s));
    std::vector<aphi_dev::StatSlot> slots(aphi_dev::kStatSlots);
    HIPCHK(hipMemcpyAsync(slots.data(), m->stats.p, sizeof(aphi_dev::FactorStatsDev), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    auto reduce = [&slots]() {
        aphi_dev::StatSlot h = slots[0];
        for (const aphi_dev::StatSlot& q : slots) {
            h.smallest_pivot = std::min(h.smallest_pivot, q.smallest_pivot);
            h.largest_pivot = std::max(h.largest_pivot, q.largest_pivot);
            h.largest_multiplier = std::max(h.largest_multiplier, q.largest_multiplier);
            h.fail_col = std::min(h.fail_col, q.fail_col);
        }
        return h;
    };
    aphi_dev::StatSlot h = reduce();
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, m->ctx->ev0, m->ctx->ev1));
    if (h.fail_col != INT_MAX) {
        // the extremes over what the reference's loop reaches before it stops
        const int fail = h.fail_col;
        aphi_dev::launch_failure_stats<T>(s, m->num_super, m->factor_entries, m->plan(), m->P.p, fail,
                                          m->stats.as<aphi_dev::FactorStatsDev>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(slots.data(), m->stats.p, sizeof(aphi_dev::FactorStatsDev), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        h = reduce();
        h.fail_col = fail;
    }
    aphi_factor_stats r{};
    r.ok = h.fail_col == INT_MAX ? 1 : 0;
    r.columns_done = r.ok ? m->n : h.fail_col + 1;
    r.smallest_pivot = m->n ? h.smallest_pivot : 0.0;
    r.largest_pivot = h.largest_pivot;
    r.largest_multiplier = h.largest_multiplier;
    r.nnz = m->predicted_nnz;
    r.milliseconds = ms;
    r.flops = m->flops;
    r.update_launches = launches;
    r.update_flops = update_flops;
    if (rec) {
        double total = 0.0;
        for (UpdateLaunch& u : *rec) {
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, m->prof_events[u.ev], m->prof_events[u.ev + 1]));
            u.ms = t;
            total += t;
        }
        r.update_ms = total;
    }
    *out = r;
    m->last_stats = r;
    m->factored = r.ok != 0;
}

// APHI_SOLVE_GRAPHS=0 launches the solve's level loops one by one instead of
// replaying their captured graph
// The library's streams do not wait for the legacy default stream on their own
// (they are non-blocking). An entry point that reads or writes caller device
// memory first orders its stream after the work the caller has queued there (a
// framework filling b, reading an earlier x), and synchronizes before it returns.
void after_caller(aphi_context* c) {
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventRecord(c->caller, nullptr));
    HIPCHK(hipStreamWaitEvent(c->stream, c->caller, 0));
}

bool solve_graphs() {
    const char* e = std::getenv("APHI_SOLVE_GRAPHS");
    return !(e && e[0] == '0');
}

void ensure_solve_buffers(aphi_mesh* m, int nrhs) {
    if (nrhs <= m->solve_nrhs_cap) return;
    const size_t cs = sizeof(cd);  // sized for the wider type; FP32 uses the front half
    const size_t q = static_cast<size_t>(aphi_dev::rhs_groups(nrhs)) * aphi_dev::kQ;  // padded right-hand sides
    m->bt.reserve(static_cast<size_t>(m->n) * q * cs);
    m->X.reserve(static_cast<size_t>(m->n) * q * cs);
    m->Yv.reserve(static_cast<size_t>(m->n) * q * cs);
    m->W.reserve(static_cast<size_t>(m->rows_total) * q * cs);
    m->xil.reserve(static_cast<size_t>(m->n) * q * sizeof(cd));
    m->rpart.reserve(static_cast<size_t>(std::max<int64_t>(m->max_rect_partials, 1)) * q * cs);
    m->trect.reserve(static_cast<size_t>(std::max<int64_t>(m->max_rect_sums, 1)) * q * cs);
    m->fpart.reserve(static_cast<size_t>(std::max<int64_t>(m->max_fwd_partials, 1)) * q * cs);
    m->colscale.reserve(nrhs * sizeof(double));
    m->res.reserve(static_cast<size_t>(m->n) * nrhs * sizeof(cd));
    const size_t cols = static_cast<size_t>(nrhs) * m->seg.count;  // (problem, right-hand side) pairs
    m->norms.reserve(static_cast<size_t>(64) * cols * 6 * sizeof(double));
    m->partials.reserve(std::max<size_t>(m->partials.bytes, static_cast<size_t>(64) * cols * sizeof(double)));
    m->solve_nrhs_cap = nrhs;
}

// x (+)= D * P^T * solve(P * D * b) with the current factor
template <typename T>
void solve_t(aphi_mesh* m, int nrhs, const cd* b, int64_t ldb, cd* x, int64_t ldx, int accumulate) {
    hipStream_t s = m->st();
    const aphi_dev::FrontPlan fp = m->plan();
    const int n = m->n;
    const int mode = std::is_same<T, float>::value ? 1 : 0;
    aphi_dev::launch_column_scale(s, n, nrhs, b, ldb, m->partials.as<double>(), m->colscale.as<double>(), mode);
    aphi_dev::launch_perm_in<T>(s, n, nrhs, m->perm.as<int>(), m->d.as<double>(), b, ldb, m->colscale.as<double>(), m->bt.p);
    const int* nodes = m->nodes.as<int>();
    constexpr int SB = aphi_dev::kSB;
    // triangle rows below block B of the level's supernodes, and rows below the
    // pivots (max over the level)
    const auto tri_rows = [](const Level& L, int B) {
        int rb = 0;
        for (int i = 0; i < L.count(); ++i) {
            const int kk = L.k[static_cast<size_t>(i)];
            if (kk <= B) continue;
            rb = std::max(rb, kk - B - std::min(aphi_dev::kSB, kk - B));
        }
        return rb;
    };
    const auto rect_rows = [](const Level& L) {
        int rb = 0;
        for (int i = 0; i < L.count(); ++i) rb = std::max(rb, L.m[static_cast<size_t>(i)] - L.k[static_cast<size_t>(i)]);
        return rb;
    };
    const auto levels = [&]() {
    for (const Level& L : m->levels) {
        const int cnt = L.count();
        int maxk = 0;
        for (int i = 0; i < cnt; ++i) maxk = std::max(maxk, L.k[static_cast<size_t>(i)]);
        const int maxm = std::max(L.small_max_m, L.large_max_m);
        if (aphi_dev::small_level<T>(maxk, maxm)) {
            // gather, diagonal step and update vector of every front in one launch
            aphi_dev::launch_fwd_small<T>(s, nodes + L.off, cnt, maxm, nrhs, fp, m->P.p, m->Xsb.p, m->bt.p, n,
                                          m->rows_total, m->W.p, m->Yv.p);
            continue;
        }
        aphi_dev::launch_fwd_gather<T>(s, nodes + L.off, cnt, maxm, nrhs, fp, m->bt.p, n, m->rows_total, m->W.p);
        for (int B = 0; B < maxk; B += SB)
            aphi_dev::launch_fwd_step<T>(s, nodes + L.off, cnt, B, std::min(SB, maxk - B), tri_rows(L, B), nrhs, fp,
                                         m->P.p, m->Xsb.p, m->W.p, m->Yv.p, n, m->rows_total);
        aphi_dev::launch_fwd_rect<T>(s, nodes + L.off, cnt, maxk, rect_rows(L), nrhs, fp, m->P.p, m->W.p, m->Yv.p, n,
                                     m->rows_total, m->fpart.p);
    }
    for (int li = m->num_levels - 1; li >= 0; --li) {
        const Level& L = m->levels[static_cast<size_t>(li)];
        const int cnt = L.count();
        int maxk = 0;
        for (int i = 0; i < cnt; ++i) maxk = std::max(maxk, L.k[static_cast<size_t>(i)]);
        if (maxk == 0) continue;
        const bool all = maxk > SB;  // the chain has updates
        const void* t = aphi_dev::launch_bwd_rect<T>(s, nodes + L.off, cnt, maxk, rect_rows(L), nrhs, all, fp, m->P.p,
                                                     m->X.p, n, m->rpart.p, m->trect.p);
        const int rk = aphi_dev::rect_cols(maxk);
        for (int B = ((maxk - 1) / SB) * SB; B >= 0; B -= SB)
            aphi_dev::launch_bwd_step<T>(s, nodes + L.off, cnt, B, std::min(SB, maxk - B), nrhs, fp, m->P.p, m->dinv.p,
                                         m->XTsb.p, m->Yv.p, m->X.p, n, const_cast<void*>(t), rk, all);
    }
    };
    if (solve_graphs() && m->solve_graphs_on) {
        const int prec = std::is_same<T, float>::value ? APHI_FP32 : APHI_FP64;
        const int groups = aphi_dev::rhs_groups(nrhs);
        const std::vector<const void*> key = {m->P.p,  m->Xsb.p, m->XTsb.p, m->dinv.p, m->bt.p, m->W.p,
                                              m->Yv.p, m->X.p,    m->rpart.p, m->trect.p, m->fpart.p, m->nodes.p};
        SolveGraph& sg = m->sgraph;
        if (!(sg.exec && sg.precision == prec && sg.groups == groups && sg.key == key)) {
            if (sg.exec) HIPCHK(hipGraphExecDestroy(sg.exec));
            sg.exec = nullptr;
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
            try {
                levels();
            } catch (...) {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(s, &g));
            const hipError_t ie = hipGraphInstantiate(&sg.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(ie);
            sg.precision = prec;
            sg.groups = groups;
            sg.key = key;
        }
        HIPCHK(hipGraphLaunch(sg.exec, s));
    } else {
        levels();
    }
    aphi_dev::launch_perm_out<T>(s, n, nrhs, m->perm.as<int>(), m->d.as<double>(), m->X.p, m->colscale.as<double>(), x, ldx,
                                 accumulate);
    HIPCHK(hipGetLastError());
}

void solve_any(aphi_mesh* m, int nrhs, const cd* b, int64_t ldb, cd* x, int64_t ldx, int accumulate) {
    require(m->factored, "solve: there is no usable factor (call aphi_factorize first)");
    ensure_solve_buffers(m, nrhs);
    if (m->factor_precision == APHI_FP64) solve_t<double>(m, nrhs, b, ldb, x, ldx, accumulate);
    else solve_t<float>(m, nrhs, b, ldb, x, ldx, accumulate);
}

// y = A x (original A, caller layouts), through the interleaved copy of x
void symv(aphi_mesh* m, int nrhs, const cd* x, int64_t ldx, cd* y, int64_t ldy) {
    hipStream_t s = m->st();
    const int n = m->n;
    const dim3 grid((n + 15) / 16, aphi_dev::rhs_groups(nrhs));
    hipLaunchKernelGGL(aphi_dev::k_interleave16, grid, dim3(256), 0, s, n, nrhs, x, ldx, m->xil.as<cd>());
    hipLaunchKernelGGL(aphi_dev::k_symv16, grid, dim3(256), 0, s, n, nrhs, m->row_ptr.as<int>(), m->col.as<int>(),
                       m->ttptr.as<int>(), m->tslot.as<int>(), m->trow.as<int>(), m->A.as<cd>(), m->xil.as<cd>(), y, ldy);
}

// backward error and relative residual per column c = r * count + s (problem s
// of the mesh, right-hand side r; c = r for a mesh of one problem), each on its
// own problem's rows and max |A_ij|, and (when bes is given) the backward error
// of the equilibrated system; leaves r = b - A x in m->res
void measure(aphi_mesh* m, int nrhs, const cd* x, int64_t ldx, const cd* b, int64_t ldb, double* be, double* rr,
             double* bes = nullptr, double* scaled_den = nullptr) {
    hipStream_t s = m->st();
    const int n = m->n, ns = m->seg.count, ncol = nrhs * ns;
    ensure_solve_buffers(m, nrhs);
    cd* res = m->res.as<cd>();
    symv(m, nrhs, x, ldx, res, static_cast<int64_t>(n));
    hipLaunchKernelGGL(aphi_dev::k_residual, dim3(blocks_for(n), nrhs), dim3(256), 0, s, n, nrhs, b, ldb, res,
                       static_cast<int64_t>(n));
    const int parts = ns == 1 ? 64 : 4;
    const int64_t* off = m->seg.d_rows.as<int64_t>();
    double* nrm = m->norms.as<double>();
    const size_t block = static_cast<size_t>(parts) * ncol;
    const dim3 grid(parts, ncol);
    hipLaunchKernelGGL(aphi_dev::k_seg_sumsq, grid, dim3(256), 0, s, ns, off, res, static_cast<int64_t>(n), nullptr, 0, nrm);
    hipLaunchKernelGGL(aphi_dev::k_seg_sumsq, grid, dim3(256), 0, s, ns, off, x, ldx, nullptr, 0, nrm + block);
    hipLaunchKernelGGL(aphi_dev::k_seg_sumsq, grid, dim3(256), 0, s, ns, off, b, ldb, nullptr, 0, nrm + 2 * block);
    const bool scaled = bes != nullptr && m->equilibrated;
    if (scaled) {
        const double* dd = m->d.as<double>();
        hipLaunchKernelGGL(aphi_dev::k_seg_sumsq, grid, dim3(256), 0, s, ns, off, res, static_cast<int64_t>(n), dd, 1,
                           nrm + 3 * block);
        hipLaunchKernelGGL(aphi_dev::k_seg_sumsq, grid, dim3(256), 0, s, ns, off, x, ldx, dd, -1, nrm + 4 * block);
        hipLaunchKernelGGL(aphi_dev::k_seg_sumsq, grid, dim3(256), 0, s, ns, off, b, ldb, dd, 1, nrm + 5 * block);
    }
    std::vector<double> h((scaled ? 6 : 3) * block);
    HIPCHK(hipMemcpyAsync(h.data(), nrm, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    const auto sum = [&](int which, int c) {
        double t = 0.0;
        for (int p = 0; p < parts; ++p) t += h[which * block + static_cast<size_t>(c) * parts + p];
        return t;
    };
    for (int c = 0; c < ncol; ++c) {
        const int sg = c % ns;
        const double sr = sum(0, c), sx = sum(1, c), sb = sum(2, c);
        const double num = std::sqrt(sr), den = m->amax[static_cast<size_t>(sg)] * std::sqrt(sx) + std::sqrt(sb);
        if (be) be[c] = den > 0.0 ? num / den : num;
        if (rr) rr[c] = sb > 0.0 ? num / std::sqrt(sb) : num;
        if (scaled) {
            const double tr = sum(3, c), tx = sum(4, c), tb = sum(5, c);
            const double dn = m->amax_eq[static_cast<size_t>(sg)] * std::sqrt(tx) + std::sqrt(tb);
            bes[c] = dn > 0.0 ? std::sqrt(tr) / dn : std::sqrt(tr);
            if (scaled_den) scaled_den[c] = dn;
        } else if (bes) {
            bes[c] = be ? be[c] : 0.0;
        }
    }
}

// Right-preconditioned restarted GMRES on the equilibrated system
// (D A D)(D^-1 x) = D b with the current factor as the preconditioner, in FP64:
// it minimises ||D (b - A x)||, the residual of the equilibrated backward error,
// over the Krylov space that classical refinement (one correction per solve)
// explores one direction at a time, so it never needs more solves and it keeps
// converging where the FP32 factor is too inaccurate for classical refinement.
// Classical Gram-Schmidt with one reorthogonalization; partial sums in a fixed
// order, so repeated runs give the same bits. Improves x in place, leaves the
// measured backward errors of the final x in be/rr/bs and returns the number of
// preconditioner solves.
constexpr int kGmresRestart = 48;
// Stall rule (see gmres_refine): the reduction per step is judged over the last
// kStallWindow steps, from kStallFrom steps into a cycle on.
constexpr int kStallWindow = 8;
constexpr int kStallFrom = 16;

// APHI_TRACE_REFINEMENT=1 prints the refinement's residual history to stderr.
bool trace_refinement() {
    static const bool on = [] {
        const char* e = std::getenv("APHI_TRACE_REFINEMENT");
        return e != nullptr && e[0] == '1';
    }();
    return on;
}
int gmres_refine(aphi_mesh* m, int nrhs, const cd* b, int64_t ldb, cd* x, int64_t ldx, double tol, int max_solves,
                 std::vector<double>& be, std::vector<double>& rr, std::vector<double>& bs) {
    using aphi_dev::cd;
    const int n = m->n;
    hipStream_t s = m->st();
    const int RS = kGmresRestart;
    // one Krylov process per column c = r * nseg + sg: right-hand side r of problem
    // sg (a batch's problems share the launches, never an inner product)
    const int nseg = m->seg.count, ncol = nrhs * nseg;
    const int64_t* off = m->seg.d_rows.as<int64_t>();
    const int* row_seg = nseg > 1 ? m->seg.row_seg.as<int>() : nullptr;
    const int64_t vs = static_cast<int64_t>(nrhs) * n;
    const int parts = nseg == 1 ? 64 : 4;
    m->gV.reserve(static_cast<size_t>(RS + 1) * vs * sizeof(cd));
    m->gZ.reserve(static_cast<size_t>(RS) * vs * sizeof(cd));
    m->gW.reserve(static_cast<size_t>(vs) * sizeof(cd));
    m->gU.reserve(static_cast<size_t>(vs) * sizeof(cd));
    m->gDots.reserve(static_cast<size_t>(RS + 1) * ncol * parts * sizeof(cd));
    m->gSmall.reserve(static_cast<size_t>(RS + 1) * ncol * sizeof(cd));
    cd* V = m->gV.as<cd>();
    cd* Z = m->gZ.as<cd>();
    cd* w = m->gW.as<cd>();
    cd* u = m->gU.as<cd>();
    cd* dots = m->gDots.as<cd>();
    const double* d = m->d.as<double>();
    const dim3 grid(blocks_for(n), nrhs);
    const size_t nc = static_cast<size_t>(ncol);
    std::vector<double> den(nc);
    std::vector<Complex> H(nc * (RS + 1) * RS), cs(nc * RS), g(nc * (RS + 1));
    std::vector<double> cr(nc * RS);
    std::vector<double> hn(nc), inv(nc);
    std::vector<cd> hup(static_cast<size_t>(RS + 1) * nc), hdot(static_cast<size_t>(RS + 1) * nc * parts);
    const auto worst = [&]() {
        double wv = 0.0;
        for (size_t c = 0; c < nc; ++c) {
            const double e = std::max(be[c], bs[c]);
            wv = std::isnan(e) ? std::numeric_limits<double>::infinity() : std::max(wv, e);
        }
        return wv;
    };
    // per-column 2-norms of v (host sums of fixed partials)
    const auto norms = [&](const cd* v, std::vector<double>& out) {
        double* nrm = m->norms.as<double>();
        hipLaunchKernelGGL(aphi_dev::k_seg_sumsq, dim3(parts, ncol), dim3(256), 0, s, nseg, off, v,
                           static_cast<int64_t>(n), nullptr, 0, nrm);
        std::vector<double> h(static_cast<size_t>(parts) * nc);
        HIPCHK(hipMemcpyAsync(h.data(), nrm, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (size_t c = 0; c < nc; ++c) {
            double t = 0.0;
            for (int p = 0; p < parts; ++p) t += h[c * parts + p];
            out[c] = std::sqrt(t);
        }
    };
    const auto scale_cols = [&](const cd* src, const std::vector<double>& f, cd* dst) {
        HIPCHK(hipMemcpyAsync(m->gSmall.p, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(aphi_dev::k_scale_cols, grid, dim3(256), 0, s, n, nrhs, nseg, row_seg, src,
                           m->gSmall.as<double>(), dst);
    };
    int solves = 0;
    for (;;) {
        measure(m, nrhs, x, ldx, b, ldb, be.data(), rr.data(), bs.data(), den.data());
        const double wv = worst();
        if (trace_refinement()) std::fprintf(stderr, "gmres restart after %d solves: worst backward error %.3e\n", solves, wv);
        if (wv <= tol || !std::isfinite(wv) || solves >= max_solves) return solves;
        // V_0 = D r / ||D r||
        hipLaunchKernelGGL(aphi_dev::k_scale_d, grid, dim3(256), 0, s, n, nrhs, d, 1, m->res.as<cd>(),
                           static_cast<int64_t>(n), V, static_cast<int64_t>(n));
        std::vector<double> beta(nc);
        norms(V, beta);
        double start = 0.0;  // estimated worst scaled backward error at this restart
        for (size_t c = 0; c < nc; ++c) start = std::max(start, den[c] > 0.0 ? beta[c] / den[c] : beta[c]);
        for (size_t c = 0; c < nc; ++c) inv[c] = beta[c] > 0.0 ? 1.0 / beta[c] : 0.0;
        scale_cols(V, inv, V);
        std::fill(H.begin(), H.end(), Complex(0.0, 0.0));
        std::fill(g.begin(), g.end(), Complex(0.0, 0.0));
        std::vector<double> hist(static_cast<size_t>(RS), 0.0);  // estimated worst error after each step
        for (size_t c = 0; c < nc; ++c) g[c * (RS + 1)] = Complex(beta[c], 0.0);
        int k = 0;
        for (int j = 0; j < RS && solves < max_solves; ++j) {
            cd* Vj = V + static_cast<int64_t>(j) * vs;
            cd* Zj = Z + static_cast<int64_t>(j) * vs;
            hipLaunchKernelGGL(aphi_dev::k_scale_d, grid, dim3(256), 0, s, n, nrhs, d, -1, Vj, static_cast<int64_t>(n), u,
                               static_cast<int64_t>(n));
            solve_any(m, nrhs, u, n, Zj, n, 0);
            ++solves;
            symv(m, nrhs, Zj, static_cast<int64_t>(n), w, static_cast<int64_t>(n));
            hipLaunchKernelGGL(aphi_dev::k_scale_d, grid, dim3(256), 0, s, n, nrhs, d, 1, w, static_cast<int64_t>(n), w,
                               static_cast<int64_t>(n));
            for (int pass = 0; pass < 2; ++pass) {
                hipLaunchKernelGGL(aphi_dev::k_dot_partial, dim3(parts, j + 1, ncol), dim3(256), 0, s, n, nrhs, nseg, off,
                                   V, vs, w, dots);
                HIPCHK(hipMemcpyAsync(hdot.data(), dots, static_cast<size_t>(j + 1) * nc * parts * sizeof(cd),
                                      hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                for (int v = 0; v <= j; ++v)
                    for (size_t c = 0; c < nc; ++c) {
                        double tr = 0.0, ti = 0.0;
                        for (int p = 0; p < parts; ++p) {
                            const cd& q = hdot[(static_cast<size_t>(v) * nc + c) * parts + p];
                            tr += q.re;
                            ti += q.im;
                        }
                        hup[static_cast<size_t>(v) * nc + c] = cd{tr, ti};
                        H[(c * (RS + 1) + v) * RS + j] += Complex(tr, ti);
                    }
                HIPCHK(hipMemcpyAsync(m->gSmall.p, hup.data(), static_cast<size_t>(j + 1) * nc * sizeof(cd),
                                      hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(aphi_dev::k_sub_basis, grid, dim3(256), 0, s, n, nrhs, nseg, row_seg, j + 1, V, vs,
                                   m->gSmall.as<cd>(), w);
            }
            norms(w, hn);
            for (size_t c = 0; c < nc; ++c) inv[c] = hn[c] > 0.0 ? 1.0 / hn[c] : 0.0;
            scale_cols(w, inv, V + static_cast<int64_t>(j + 1) * vs);
            // Givens rotations per column; |g_{j+1}| = ||D r|| of the updated x
            double ratio = 0.0;
            for (size_t c = 0; c < nc; ++c) {
                Complex* Hr = &H[c * (RS + 1) * RS];
                Complex* gr = &g[c * (RS + 1)];
                Hr[static_cast<size_t>(j + 1) * RS + j] = Complex(hn[c], 0.0);
                for (int i = 0; i < j; ++i) {
                    const double cc = cr[c * RS + i];
                    const Complex sn = cs[c * RS + i];
                    const Complex a = Hr[static_cast<size_t>(i) * RS + j], bb = Hr[static_cast<size_t>(i + 1) * RS + j];
                    Hr[static_cast<size_t>(i) * RS + j] = cc * a + sn * bb;
                    Hr[static_cast<size_t>(i + 1) * RS + j] = -std::conj(sn) * a + cc * bb;
                }
                const Complex a = Hr[static_cast<size_t>(j) * RS + j], bb = Hr[static_cast<size_t>(j + 1) * RS + j];
                const double rho = std::sqrt(std::norm(a) + std::norm(bb));
                double cc = 1.0;
                Complex sn(0.0, 0.0);
                if (rho > 0.0) {
                    if (std::abs(a) == 0.0) {
                        cc = 0.0;
                        sn = std::conj(bb) / rho;
                    } else {
                        cc = std::abs(a) / rho;
                        sn = (a / std::abs(a)) * std::conj(bb) / rho;
                    }
                }
                cr[c * RS + j] = cc;
                cs[c * RS + j] = sn;
                Hr[static_cast<size_t>(j) * RS + j] = cc * a + sn * bb;
                Hr[static_cast<size_t>(j + 1) * RS + j] = Complex(0.0, 0.0);
                const Complex gj = gr[j];
                gr[j] = cc * gj;
                gr[j + 1] = -std::conj(sn) * gj;
                ratio = std::max(ratio, den[c] > 0.0 ? std::abs(gr[j + 1]) / den[c] : std::abs(gr[j + 1]));
            }
            k = j + 1;
            hist[static_cast<size_t>(j)] = ratio;
            if (trace_refinement()) std::fprintf(stderr, "  gmres %2d: estimated scaled backward error %.3e\n", solves, ratio);
            if (!(ratio > 0.5 * tol)) break;
            // Stalled: stop here (the caller refactors in FP64) rather than spend the
            // rest of the budget. After three steps of this cycle without any
            // reduction; or, from kStallFrom steps on, when the reduction per step
            // over the last kStallWindow steps projects past twice the budget.
            // GMRES on these systems speeds up as the Krylov space grows (at 12.9 and
            // 17.3 MHz on the 32^3 cube the first ten steps gain less than a factor
            // of 3, and the attempts converge after 26 and 33 solves), so an early
            // rate would give up on attempts that succeed.
            if (start > 0.0 && ratio > 0.0) {
                double rate = 0.0, proj = 0.0;
                bool stalled = false;
                if (j >= 2) {
                    rate = std::pow(ratio / start, 1.0 / (j + 1));
                    stalled = !(rate < 1.0);
                }
                if (!stalled && j + 1 >= kStallFrom) {
                    rate = std::pow(ratio / hist[static_cast<size_t>(j - kStallWindow)], 1.0 / kStallWindow);
                    proj = solves + std::log(0.5 * tol / ratio) / std::log(rate);
                    stalled = !(rate < 1.0) || proj > 2.0 * max_solves;
                }
                if (stalled) {
                    if (trace_refinement()) std::fprintf(stderr, "  gmres stalled (rate %.3f per step)\n", rate);
                    return solves;
                }
            }
        }
        // y = R^-1 g per column, then x += Z y
        for (size_t c = 0; c < nc; ++c) {
            const Complex* Hr = &H[c * (RS + 1) * RS];
            const Complex* gr = &g[c * (RS + 1)];
            std::vector<Complex> y(static_cast<size_t>(k));
            for (int i = k - 1; i >= 0; --i) {
                Complex t = gr[i];
                for (int q = i + 1; q < k; ++q) t -= Hr[static_cast<size_t>(i) * RS + q] * y[static_cast<size_t>(q)];
                const Complex dg = Hr[static_cast<size_t>(i) * RS + i];
                y[static_cast<size_t>(i)] = std::abs(dg) > 0.0 ? t / dg : Complex(0.0, 0.0);
            }
            for (int i = 0; i < k; ++i) hup[static_cast<size_t>(i) * nc + c] = cd{y[static_cast<size_t>(i)].real(), y[static_cast<size_t>(i)].imag()};
        }
        HIPCHK(hipMemcpyAsync(m->gSmall.p, hup.data(), static_cast<size_t>(k) * nc * sizeof(cd), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(aphi_dev::k_add_basis, grid, dim3(256), 0, s, n, nrhs, nseg, row_seg, k, Z, vs,
                           m->gSmall.as<cd>(), x, ldx);
    }
}

void factorize(aphi_mesh* m, int precision, double pivot_floor, aphi_factor_stats* out) {
    require(m->assembled, "factorize: the system has not been assembled");
    require(precision == APHI_FP64 || precision == APHI_FP32, "factorize: unknown precision");
    if (!m->equilibrated) equilibrate(m, 0, nullptr);
    if (m->factor_precision != precision) {
        m->P.release();
        m->CB.release();
        m->mws.release();
    }
    m->factor_precision = precision;
    if (precision == APHI_FP64) factorize_t<double>(m, pivot_floor, out);
    else factorize_t<float>(m, pivot_floor, out);
}

}  // namespace

// ------------------------------------------------------------------ C-ABI

namespace {
// The device address of `scratch` on `device` (current), looked up once per
// device: the array's name in host code is only a stand-in for the symbol.
aphi_dev::cd* scratch_address(int device) {
    static std::mutex mu;
    static std::vector<aphi_dev::cd*> table;
    std::lock_guard<std::mutex> lock(mu);
    if (table.size() <= static_cast<size_t>(device)) table.resize(static_cast<size_t>(device) + 1, nullptr);
    if (!table[static_cast<size_t>(device)]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw DeviceError("hipGetSymbolAddress: no address for the scratch array", 1);
        table[static_cast<size_t>(device)] = static_cast<aphi_dev::cd*>(p);
    }
    return table[static_cast<size_t>(device)];
}
}  // namespace

extern "C" int aphi_context_create(int device, aphi_context** out) {
    return aphi_capi::guarded([&] {
        require(out != nullptr, "aphi_context_create: null output");
        *out = nullptr;
        int count = 0;
        HIPCHK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw std::invalid_argument("aphi_context_create: no HIP device " + std::to_string(device));
        HIPCHK(hipSetDevice(device));
        aphi_dev::cd* scr = scratch_address(device);
        auto* c = new aphi_context;
        c->device = device;
        c->scratch = scr;
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&c->ev0));
        HIPCHK(hipEventCreate(&c->ev1));
        HIPCHK(hipEventCreateWithFlags(&c->main_done, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->side_done, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->caller, hipEventDisableTiming));
        *out = c;
    });
}

extern "C" void aphi_context_destroy(aphi_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->side);
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipEventDestroy(c->main_done);
    (void)hipEventDestroy(c->side_done);
    (void)hipEventDestroy(c->caller);
    (void)hipStreamDestroy(c->side);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int aphi_context_synchronize(aphi_context* c) {
    return aphi_capi::guarded([&] {
        require(c != nullptr, "aphi_context_synchronize: null context");
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

namespace {
// The device plan of one mesh description (aphi_mesh_create), or of the
// block-diagonal union of a batch (aphi_batch_create); segments set by the caller.
std::unique_ptr<aphi_mesh> build_mesh(aphi_context* ctx, const aphi_mesh_desc* d) {
        require(d->n > 0 && d->nnz > 0 && d->num_super > 0, "aphi_mesh_create: empty mesh");
        // without an assembly plan (a problem made from a bare pattern) the system
        // arrives through aphi_system_upload instead of aphi_refill
        const bool plan = d->target_ptr != nullptr;
        require(!plan || (d->num_tets > 0 && d->entry && d->kind && d->coords && d->tet_nodes && d->tet_body),
                "aphi_mesh_create: an assembly plan needs the mesh's tets, coordinates and contributions");
        require(d->num_values >= 0 && d->num_values <= 8, "aphi_mesh_create: at most 8 prescribed values");
        HIPCHK(hipSetDevice(ctx->device));
        auto m = std::make_unique<aphi_mesh>();
        m->ctx = ctx;
        m->n = d->n;
        m->nnz = d->nnz;
        m->num_nodes = d->num_nodes;
        m->num_tets = d->num_tets;
        m->num_bodies = d->num_bodies;
        m->num_super = d->num_super;
        m->num_levels = d->num_levels;
        m->num_values = d->num_values;
        m->has_plan = plan;
        m->num_targets = plan ? d->nnz + d->n : 0;
        m->cb_entries = d->cb_pool_entries;
        m->flops = d->flops;
        m->predicted_nnz = d->predicted_nnz;
        m->sigma.assign(d->body_sigma, d->body_sigma + d->num_bodies);
        m->eps_r.assign(d->body_eps_r, d->body_eps_r + d->num_bodies);
        m->mu_r.assign(d->body_mu_r, d->body_mu_r + d->num_bodies);
        const int n = d->n, ns = d->num_super;
        if (plan) {
            const int64_t nc = d->target_ptr[m->num_targets];
            for (int t = 0; t < d->num_tets; ++t)
                if (d->tet_body[t] < 0 || d->tet_body[t] >= d->num_bodies)
                    throw std::invalid_argument("aphi_mesh_create: a tet's body index is out of range");
            upload(m->st(), m->coords, d->coords, static_cast<size_t>(d->num_nodes) * 3);
            upload(m->st(), m->tets, d->tet_nodes, static_cast<size_t>(d->num_tets) * 4);
            upload(m->st(), m->tet_body, d->tet_body, static_cast<size_t>(d->num_tets));
            upload(m->st(), m->tptr, d->target_ptr, static_cast<size_t>(m->num_targets) + 1);
            upload(m->st(), m->entry, d->entry, static_cast<size_t>(nc));
            upload(m->st(), m->kind, d->kind, static_cast<size_t>(nc));
            for (int64_t c = 0; c < nc; ++c) {
                const int kk = d->kind[c] & 0x0f;
                if (kk <= 5 && d->entry[c] >= static_cast<uint32_t>(d->num_tets) * 196u)
                    throw std::invalid_argument("aphi_mesh_create: an assembly contribution names a tet outside the mesh");
                if ((d->kind[c] >> 5) >= std::max(d->num_values, 1) && (kk == 4 || kk == 5 || kk == 7 || kk == 8))
                    throw std::invalid_argument("aphi_mesh_create: a contribution names an unknown prescribed value");
            }
        }
        upload(m->st(), m->row_ptr, d->row_ptr, static_cast<size_t>(n) + 1);
        upload(m->st(), m->col, d->col_index, static_cast<size_t>(d->nnz));
        {
            std::vector<int> row_of(static_cast<size_t>(d->nnz));
            std::vector<int> tp(static_cast<size_t>(n) + 1, 0);
            for (int r = 0; r < n; ++r)
                for (int k = d->row_ptr[r]; k < d->row_ptr[r + 1]; ++k) {
                    row_of[static_cast<size_t>(k)] = r;
                    const int c = d->col_index[k];
                    if (c < r || c >= n) throw std::invalid_argument("aphi_mesh_create: the pattern is not upper triangular");
                    if (c != r) ++tp[static_cast<size_t>(c) + 1];
                }
            for (int i = 0; i < n; ++i) tp[static_cast<size_t>(i) + 1] += tp[static_cast<size_t>(i)];
            std::vector<int> ts(static_cast<size_t>(tp[static_cast<size_t>(n)])), tr(ts.size());
            std::vector<int> cur(tp.begin(), tp.end() - 1);
            for (int r = 0; r < n; ++r)
                for (int k = d->row_ptr[r]; k < d->row_ptr[r + 1]; ++k) {
                    const int c = d->col_index[k];
                    if (c == r) continue;
                    ts[static_cast<size_t>(cur[static_cast<size_t>(c)])] = k;
                    tr[static_cast<size_t>(cur[static_cast<size_t>(c)]++)] = r;
                }
            upload(m->st(), m->row_of, row_of.data(), row_of.size());
            upload(m->st(), m->ttptr, tp.data(), tp.size());
            upload(m->st(), m->tslot, ts.data(), ts.size());
            upload(m->st(), m->trow, tr.data(), tr.size());
        }
        upload(m->st(), m->perm, d->factor_perm, static_cast<size_t>(n));
        upload(m->st(), m->col_ref, d->col_ref, static_cast<size_t>(n));
        upload(m->st(), m->first, d->super_first, static_cast<size_t>(ns) + 1);
        upload(m->st(), m->srow_ptr, d->super_row_ptr, static_cast<size_t>(ns) + 1);
        m->rows_total = d->super_row_ptr[ns];
        upload(m->st(), m->srows, d->super_rows, static_cast<size_t>(m->rows_total));
        upload(m->st(), m->rel, d->super_rel, static_cast<size_t>(m->rows_total));
        upload(m->st(), m->panel_off, d->panel_offset, static_cast<size_t>(ns) + 1);
        m->factor_entries = d->panel_offset[ns];
        upload(m->st(), m->cb_off, d->cb_offset, static_cast<size_t>(ns));
        upload(m->st(), m->s2p, d->slot_to_panel, static_cast<size_t>(d->nnz));
        for (int64_t k = 0; k < d->nnz; ++k)
            if (d->slot_to_panel[k] < 0 || d->slot_to_panel[k] >= m->factor_entries)
                throw std::invalid_argument("aphi_mesh_create: slot_to_panel points outside the factor");

        // children (ascending), levels (small fronts first, then large ones)
        std::vector<int> cptr(static_cast<size_t>(ns) + 1, 0);
        for (int s = 0; s < ns; ++s) {
            const int p = d->super_parent[s];
            if (p >= 0) {
                if (p <= s || p >= ns) throw std::invalid_argument("aphi_mesh_create: supernode parents must follow their children");
                ++cptr[static_cast<size_t>(p) + 1];
            }
        }
        for (int s = 0; s < ns; ++s) cptr[static_cast<size_t>(s) + 1] += cptr[static_cast<size_t>(s)];
        std::vector<int> ch(static_cast<size_t>(cptr[static_cast<size_t>(ns)]));
        {
            std::vector<int> cur(cptr.begin(), cptr.end() - 1);
            for (int s = 0; s < ns; ++s)
                if (d->super_parent[s] >= 0) ch[static_cast<size_t>(cur[static_cast<size_t>(d->super_parent[s])]++)] = s;
        }
        upload(m->st(), m->child_ptr, cptr.data(), cptr.size());
        upload(m->st(), m->child, ch.data(), ch.size());
        m->levels.assign(static_cast<size_t>(d->num_levels), Level{});
        std::vector<std::vector<int>> small(static_cast<size_t>(d->num_levels)), large(static_cast<size_t>(d->num_levels));
        const int small_max = small_front_max();
        for (int s = 0; s < ns; ++s) {
            const int L = d->super_level[s];
            if (L < 0 || L >= d->num_levels) throw std::invalid_argument("aphi_mesh_create: supernode level out of range");
            const int64_t mm = d->super_row_ptr[s + 1] - d->super_row_ptr[s];
            (mm <= small_max ? small : large)[static_cast<size_t>(L)].push_back(s);
        }
        std::vector<int> node_list, ea;
        std::vector<int64_t> wo(static_cast<size_t>(ns), -1);
        {
            int64_t maxm = 0;
            for (int s = 0; s < ns; ++s) maxm = std::max<int64_t>(maxm, d->super_row_ptr[s + 1] - d->super_row_ptr[s]);
            m->lw_zero = (maxm + aphi_dev::kPanelTail + 63) / 64 * 64;
        }
        for (int L = 0; L < d->num_levels; ++L) {
            Level& lv = m->levels[static_cast<size_t>(L)];
            lv.off = static_cast<int64_t>(node_list.size());
            lv.ea_off = static_cast<int64_t>(ea.size());
            lv.small_count = static_cast<int>(small[static_cast<size_t>(L)].size());
            lv.large_count = static_cast<int>(large[static_cast<size_t>(L)].size());
            int blk = 0, maxk = 0, maxrect = 0;
            int64_t wlevel = m->lw_zero;  // this level's W windows, after the zero column
            for (int pass = 0; pass < 2; ++pass)
                for (int s : (pass == 0 ? small : large)[static_cast<size_t>(L)]) {
                    node_list.push_back(s);
                    const int mm = static_cast<int>(d->super_row_ptr[s + 1] - d->super_row_ptr[s]);
                    const int kk = d->super_first[s + 1] - d->super_first[s];
                    lv.m.push_back(mm);
                    lv.k.push_back(kk);
                    if (pass == 0) lv.small_max_m = std::max(lv.small_max_m, mm);
                    else {
                        lv.large_max_m = std::max(lv.large_max_m, mm);
                        ea.push_back(blk);
                        blk += (mm + 63) / 64;
                        wo[static_cast<size_t>(s)] = wlevel;
                        wlevel += static_cast<int64_t>(std::min(kk, aphi_dev::kWCols)) * mm;
                    }
                    maxk = std::max(maxk, kk);
                    maxrect = std::max(maxrect, mm - kk);
                }
            lv.ea_blocks = blk;
            ea.push_back(blk);
            m->lw_entries = std::max(m->lw_entries, wlevel);
            m->max_level_count = std::max(m->max_level_count, lv.count());
            const int64_t rc = aphi_dev::rect_cols(maxk);
            m->max_rect_partials = std::max<int64_t>(m->max_rect_partials,
                                                     static_cast<int64_t>(lv.count()) * aphi_dev::rect_chunks(maxrect) * rc);
            m->max_rect_sums = std::max<int64_t>(m->max_rect_sums, static_cast<int64_t>(lv.count()) * rc);
            if (maxrect > 0)
                for (const bool f32 : {true, false}) {
                    const int ks = aphi_dev::fwd_rect_split(lv.count(), maxk, maxrect, f32, nullptr);
                    if (ks > 1)
                        m->max_fwd_partials = std::max<int64_t>(
                            m->max_fwd_partials, static_cast<int64_t>(lv.count()) * ks * maxrect);
                }
        }
        upload(m->st(), m->nodes, node_list.data(), node_list.size());
        upload(m->st(), m->woff, wo.data(), wo.size());
        {
            // diagonal-block inverses: full kNB x kNB blocks, then one nb x nb
            std::vector<int64_t> xo(static_cast<size_t>(ns));
            int64_t off = 0;
            for (int s = 0; s < ns; ++s) {
                const int64_t kk = d->super_first[s + 1] - d->super_first[s];
                xo[static_cast<size_t>(s)] = off;
                const int64_t rest = kk % aphi_dev::kNB;
                off += (kk / aphi_dev::kNB) * aphi_dev::kNB * aphi_dev::kNB + rest * rest;
            }
            m->xinv_entries = off;
            upload(m->st(), m->xinv_off, xo.data(), xo.size());
            // kSB-block inverses for the solves, and the (supernode, block) list that builds them
            std::vector<int64_t> xb(static_cast<size_t>(ns));
            std::vector<int2> pairs;
            off = 0;
            for (int s = 0; s < ns; ++s) {
                const int64_t kk = d->super_first[s + 1] - d->super_first[s];
                xb[static_cast<size_t>(s)] = off;
                const int64_t rest = kk % aphi_dev::kSB;
                off += (kk / aphi_dev::kSB) * aphi_dev::kSB * aphi_dev::kSB + rest * rest;
                for (int64_t B = 0; B < kk; B += aphi_dev::kSB) pairs.push_back(int2{s, static_cast<int>(B)});
            }
            m->xsb_entries = off;
            m->num_inv_pairs = static_cast<int>(pairs.size());
            upload(m->st(), m->xsb_off, xb.data(), xb.size());
            upload(m->st(), m->inv_pairs, pairs.data(), pairs.size());
        }
        upload(m->st(), m->ea_start, ea.data(), ea.size());
        {
            // extend-add bounds for the children of blocked fronts
            std::vector<int64_t> bo(static_cast<size_t>(ns), -1);
            std::vector<int> bv;
            for (int c = 0; c < ns; ++c) {
                const int p = d->super_parent[c];
                if (p < 0) continue;
                const int64_t pm = d->super_row_ptr[p + 1] - d->super_row_ptr[p];
                if (pm <= small_max) continue;
                const int ck = d->super_first[c + 1] - d->super_first[c];
                const int32_t* rl = d->super_rel + d->super_row_ptr[c] + ck;
                const int cu = static_cast<int>(d->super_row_ptr[c + 1] - d->super_row_ptr[c]) - ck;
                bo[static_cast<size_t>(c)] = static_cast<int64_t>(bv.size());
                const int nq = static_cast<int>((pm + 63) / 64);
                int pos = 0;
                for (int q = 0; q <= nq; ++q) {
                    while (pos < cu && rl[pos] < 64 * q) ++pos;
                    bv.push_back(pos);
                }
                if (pos != cu) throw std::invalid_argument("aphi_mesh_create: relative indices exceed the parent front");
            }
            upload(m->st(), m->ea_bound_off, bo.data(), bo.size());
            upload(m->st(), m->ea_bound, bv.data(), bv.size());
        }
        if (plan && d->tet_dofs && d->tet_edge_signs && d->tet_phi_coeff && d->tet_phi_fixed) {
            const size_t nt = static_cast<size_t>(d->num_tets);
            upload(m->st(), m->tdofs, d->tet_dofs, nt * 16);
            upload(m->st(), m->tsign, d->tet_edge_signs, nt * 6);
            upload(m->st(), m->tcoef, d->tet_phi_coeff, nt * 10);
            upload(m->st(), m->tfixed, d->tet_phi_fixed, nt * 20);
            upload(m->st(), m->bsigma, d->body_sigma, static_cast<size_t>(d->num_bodies));
        }
        m->flag.reserve(sizeof(int));
        return m;
}

// Uploads the problem offsets of the mesh's segments (and, for a batch, the
// problem of every unknown) and sizes the per-problem scale table.
void upload_segments(aphi_mesh* m) {
    auto& g = m->seg;
    upload(m->st(), g.d_rows, g.rows.data(), g.rows.size());
    upload(m->st(), g.d_slots, g.slots.data(), g.slots.size());
    upload(m->st(), g.d_tets, g.tets.data(), g.tets.size());
    g.scales.reserve(static_cast<size_t>(g.count) * sizeof(aphi_dev::FreqScales));
    if (g.count > 1) {
        std::vector<int> rs(static_cast<size_t>(m->n));
        for (int sg = 0; sg < g.count; ++sg)
            for (int64_t i = g.rows[static_cast<size_t>(sg)]; i < g.rows[static_cast<size_t>(sg) + 1]; ++i)
                rs[static_cast<size_t>(i)] = sg;
        upload(m->st(), g.row_seg, rs.data(), rs.size());
    }
    m->amax.assign(static_cast<size_t>(g.count), 0.0);
    m->amax_eq.assign(static_cast<size_t>(g.count), 1.0);
}
}  // namespace

extern "C" int aphi_mesh_create(aphi_context* ctx, const aphi_mesh_desc* d, aphi_mesh** out) {
    return aphi_capi::guarded([&] {
        require(ctx && d && out, "aphi_mesh_create: null argument");
        *out = nullptr;
        auto m = build_mesh(ctx, d);
        auto& g = m->seg;
        g.count = 1;
        g.rows = {0, d->n};
        g.slots = {0, d->nnz};
        g.tets = {0, std::max(d->num_tets, 0)};
        g.bodies = {0, std::max(d->num_bodies, 0)};
        g.values.assign(1, {});
        for (int i = 0; i < d->num_values; ++i) g.values[0].emplace_back(d->values[2 * i], d->values[2 * i + 1]);
        upload_segments(m.get());
        *out = m.release();
    });
}

extern "C" void aphi_mesh_destroy(aphi_mesh* m) {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    for (hipEvent_t e : m->prof_events) (void)hipEventDestroy(e);
    if (m->sgraph.exec) (void)hipGraphExecDestroy(m->sgraph.exec);
    if (m->fgraph.exec) (void)hipGraphExecDestroy(m->fgraph.exec);
    delete m;
}

extern "C" int aphi_mesh_release_workspace(aphi_mesh* m) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_mesh_release_workspace: null mesh");
        after_caller(m->ctx);
        HIPCHK(hipSetDevice(m->ctx->device));
        HIPCHK(hipStreamSynchronize(m->ctx->stream));
        HIPCHK(hipStreamSynchronize(m->ctx->side));
        // the captured graphs name the buffers released below
        if (m->sgraph.exec) HIPCHK(hipGraphExecDestroy(m->sgraph.exec));
        m->sgraph = SolveGraph{};
        if (m->fgraph.exec) HIPCHK(hipGraphExecDestroy(m->fgraph.exec));
        m->fgraph.exec = nullptr;
        m->fgraph.key.clear();
        m->fgraph.precision = -1;
        // (the assembly error flag stays: it is allocated with the plan)
        DevBuf* numeric[] = {&m->coef, &m->E,     &m->A,     &m->rhs,  &m->Aeq,      &m->work,  &m->rowmax,
                             &m->fac,  &m->d,     &m->partials, &m->stats, &m->P, &m->CB,       &m->mws,   &m->Xinv,
                             &m->Xsb, &m->XTsb, &m->dinv,  &m->bt,    &m->W,    &m->X,        &m->Yv,    &m->LW,
                             &m->rpart, &m->trect, &m->fpart, &m->colscale, &m->res, &m->norms, &m->xil, &m->gV,       &m->gZ,    &m->gW,
                             &m->gU,   &m->gDots, &m->gSmall, &m->tcontrib, &m->tcurrent, &m->fval, &m->fvol,
                             &m->fmu,  &m->fsig,  &m->fa_tet, &m->fb_tet, &m->fe_tet, &m->fj_tet, &m->fa_node,
                             &m->fb_node, &m->fh_node, &m->fe_node, &m->fj_node, &m->fweight, &m->fflag};
        for (DevBuf* b : numeric) b->release();
        m->assembled = m->equilibrated = m->factored = false;
        m->factor_precision = -1;
        m->solve_nrhs_cap = 0;
    });
}

extern "C" int aphi_mesh_set_solve_graphs(aphi_mesh* m, int on) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_mesh_set_solve_graphs: null mesh");
        m->solve_graphs_on = on != 0;
        m->graphs_on = on != 0;
    });
}

extern "C" int aphi_mesh_set_update_tiles(aphi_mesh* m, int mode) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_mesh_set_update_tiles: null mesh");
        require(mode >= 0 && mode <= 2, "aphi_mesh_set_update_tiles: mode must be 0, 1 or 2");
        if (m->update_tiles != mode) ++m->tiles_version;  // a captured factorization holds the old tiling
        m->update_tiles = mode;
    });
}

extern "C" int aphi_mesh_get_update_tile_table(const aphi_mesh* m, int8_t* table) {
    return aphi_capi::guarded([&] {
        require(m && table, "aphi_mesh_get_update_tile_table: null argument");
        std::memcpy(table, m->tile_table, sizeof(m->tile_table));
    });
}

extern "C" int aphi_mesh_set_update_tile_table(aphi_mesh* m, const int8_t* table) {
    return aphi_capi::guarded([&] {
        require(m && table, "aphi_mesh_set_update_tile_table: null argument");
        for (size_t i = 0; i < sizeof(m->tile_table); ++i)
            require(table[i] >= 0 && table[i] <= 2, "aphi_mesh_set_update_tile_table: entries must be 0, 1 or 2");
        std::memcpy(m->tile_table, table, sizeof(m->tile_table));
        ++m->tiles_version;
    });
}

extern "C" int aphi_mesh_tune_update_tiles(aphi_mesh* m, int repeats, int8_t* table_out, double* ms_default,
                                           double* ms_tuned) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_mesh_tune_update_tiles: null mesh");
        require(repeats >= 1 && repeats <= 16, "aphi_mesh_tune_update_tiles: repeats must be 1..16");
        require(m->assembled, "aphi_mesh_tune_update_tiles: the system has not been assembled");
        HIPCHK(hipSetDevice(m->ctx->device));
        if (!m->equilibrated) equilibrate(m, 0, nullptr);
        if (m->factor_precision != APHI_FP32) {
            m->P.release();
            m->CB.release();
            m->mws.release();
        }
        m->factor_precision = APHI_FP32;
        const int saved_mode = m->update_tiles;
        const bool saved_prof = m->profiling, saved_serial = m->serial_updates;
        m->profiling = false;
        m->serial_updates = false;  // tuned on the schedule the factorization runs
        constexpr int H = aphi_mesh::kHeightClasses, W = aphi_mesh::kWidthClasses;
        // 1. every launch timed with the small tiles, then with the large ones:
        //    the faster tiling per launch class (its summed launch time)
        double t[2][H][W] = {};
        bool seen[H][W] = {};
        for (int mode = 1; mode <= 2; ++mode) {
            m->update_tiles = mode;
            for (int rep = 0; rep < repeats; ++rep) {
                std::vector<UpdateLaunch> log;
                aphi_factor_stats st{};
                factorize_t<float>(m, 0.0, &st, &log);
                for (const UpdateLaunch& u : log) {
                    t[mode - 1][u.hclass][u.wclass] += u.ms;
                    seen[u.hclass][u.wclass] = true;
                }
            }
        }
        m->update_tiles = saved_mode;
        int8_t tuned[H][W] = {};
        for (int h = 0; h < H; ++h)
            for (int w = 0; w < W; ++w)
                if (seen[h][w]) tuned[h][w] = t[0][h][w] <= t[1][h][w] ? 1 : 2;
        // 2. whole factorizations (graph replay, as used) with the default rule and
        //    with the tuned table: the table is kept only if it is faster
        const auto best_of = [&](const int8_t (*tab)[W]) {
            std::memcpy(m->tile_table, tab, sizeof(m->tile_table));
            ++m->tiles_version;
            double best = std::numeric_limits<double>::infinity();
            for (int rep = 0; rep < repeats + 1; ++rep) {  // the first run captures the graph
                aphi_factor_stats st{};
                factorize_t<float>(m, 0.0, &st);
                if (rep > 0 || repeats == 0) best = std::min(best, st.milliseconds);
            }
            return best;
        };
        const int8_t zero[H][W] = {};
        const double t_default = best_of(zero);
        const double t_tuned = best_of(tuned);
        if (!(t_tuned < t_default)) {
            std::memcpy(m->tile_table, zero, sizeof(m->tile_table));
            ++m->tiles_version;
        }
        m->profiling = saved_prof;
        m->serial_updates = saved_serial;
        if (table_out) std::memcpy(table_out, m->tile_table, sizeof(m->tile_table));
        if (ms_default) *ms_default = t_default;
        if (ms_tuned) *ms_tuned = t_tuned;
    });
}

extern "C" int aphi_mesh_set_profiling(aphi_mesh* m, int on) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_mesh_set_profiling: null mesh");
        require(on >= 0 && on <= 2, "aphi_mesh_set_profiling: on must be 0, 1 or 2");
        m->profiling = on != 0;
        m->serial_updates = on == 2;
    });
}

extern "C" int64_t aphi_mesh_device_bytes(const aphi_mesh* m) {
    if (!m) return 0;
    const DevBuf* all[] = {&m->coords, &m->tets, &m->tet_body, &m->tptr, &m->entry, &m->kind, &m->row_ptr, &m->col,
                           &m->row_of, &m->ttptr, &m->tslot, &m->trow, &m->perm, &m->col_ref, &m->first, &m->srow_ptr,
                           &m->srows, &m->rel, &m->child_ptr, &m->child, &m->panel_off, &m->cb_off, &m->s2p, &m->nodes,
                           &m->ea_start, &m->xinv_off, &m->ea_bound_off, &m->ea_bound, &m->xsb_off, &m->woff, &m->LW, &m->inv_pairs, &m->Xinv, &m->Xsb, &m->XTsb, &m->rpart, &m->trect, &m->fpart, &m->tdofs, &m->tsign, &m->tcoef, &m->tfixed, &m->bsigma, &m->tcontrib, &m->tcurrent, &m->fptr, &m->fentry, &m->bmu, &m->fval, &m->fvol, &m->fmu, &m->fsig, &m->fa_tet, &m->fb_tet, &m->fe_tet, &m->fj_tet, &m->fa_node, &m->fb_node, &m->fh_node, &m->fe_node, &m->fj_node, &m->fweight, &m->fflag, &m->gV, &m->gZ, &m->gW, &m->gU, &m->gDots, &m->gSmall, &m->Yv, &m->coef, &m->E, &m->flag, &m->A, &m->rhs, &m->Aeq, &m->work, &m->rowmax,
                           &m->fac, &m->d, &m->partials, &m->stats, &m->P, &m->CB, &m->mws, &m->bt, &m->W, &m->X,
                           &m->colscale, &m->res, &m->norms, &m->xil, &m->dinv, &m->gf.coords,
                           &m->gf.tets, &m->gf.tet_body, &m->gf.tet_on, &m->gf.column, &m->gf.prescribed, &m->gf.tptr,
                           &m->gf.entry, &m->gf.kind, &m->gf.values, &m->gf.beta, &m->gf.terms};
    int64_t total = 0;
    for (const DevBuf* b : all) total += static_cast<int64_t>(b->bytes);
    return total;
}

namespace {
// formulation_scales + the symmetric guard of src/assembly.cpp:12-45, 484-491, and
// the per-frequency scalars of the scatter formed from them with std::complex
// std::complex<double> product and quotient as g++ evaluates them for finite
// operands on x86-64 (the reference's build): (ac - bd, ad + bc) with no fused
// multiply-add, and libgcc's __divdc3 (Smith's method, no scaling). This file's
// host code is compiled by clang, whose own complex division rounds differently
// (1 ulp apart at some frequencies), so the scatter's scalars are formed here.
Complex mul_like_gcc(Complex x, Complex y) {
#pragma clang fp contract(off)
    const double a = x.real(), b = x.imag(), c = y.real(), d = y.imag();
    return Complex(a * c - b * d, a * d + b * c);
}
Complex div_like_gcc(Complex x, Complex y) {
#pragma clang fp contract(off)
    const double a = x.real(), b = x.imag(), c = y.real(), d = y.imag();
    if (std::fabs(c) < std::fabs(d)) {
        const double ratio = c / d, denom = (c * ratio) + d;
        return Complex(((a * ratio) + b) / denom, ((b * ratio) - a) / denom);
    }
    const double ratio = d / c, denom = (d * ratio) + c;
    return Complex(((b * ratio) + a) / denom, (b - (a * ratio)) / denom);
}

aphi_dev::FreqScales freq_scales(double omega, int conditioning, const std::vector<Complex>& values) {
    const Complex jw(0.0, omega);
    Complex row(1.0, 0.0), column(1.0, 0.0);
    if (conditioning == APHI_NATURAL) {
        throw std::invalid_argument(
            "refill: conditioning = natural does not give a symmetric matrix, so only half of it cannot be "
            "stored. Use row_scaled or scaled_phi for symmetric storage.");
    } else if (conditioning == APHI_ROW_SCALED) {
        if (omega == 0.0)
            throw std::invalid_argument("formulation_scales: the row-scaled formulation divides the Phi rows by "
                                        "j*omega, which is undefined at DC.");
        row = div_like_gcc(Complex(1.0, 0.0), jw);
    } else if (conditioning == APHI_SCALED_PHI) {
        if (omega == 0.0)
            throw std::invalid_argument("formulation_scales: the scaled-potential formulation substitutes "
                                        "Phi = j*omega*Phi', which is undefined at DC.");
        column = jw;
    } else {
        throw std::invalid_argument("formulation_scales: unknown formulation");
    }
    aphi_dev::FreqScales sc{};
    const auto put = [](cd& dst, const Complex& v) { dst = cd{v.real(), v.imag()}; };
    put(sc.column, column);
    put(sc.row_jw, mul_like_gcc(row, jw));
    put(sc.row_column, mul_like_gcc(row, column));
    for (size_t i = 0; i < values.size(); ++i) {
        // prescribed(): a zero stays zero (src/assembly.cpp:188-190)
        put(sc.over_column[i], values[i] == Complex(0.0, 0.0) ? values[i] : div_like_gcc(values[i], column));
        put(sc.times_row[i], mul_like_gcc(row, values[i]));
    }
    return sc;
}

// refill of every problem of the mesh, problem s at angular frequency omegas[s]
void refill_all(aphi_mesh* m, const double* omegas, int conditioning) {
    if (!m->has_plan)
        throw std::logic_error("aphi_refill: the mesh was created from a bare pattern and has no assembly plan; "
                               "hand its values over with aphi_system_upload");
    HIPCHK(hipSetDevice(m->ctx->device));
    const auto& g = m->seg;
    std::vector<aphi_dev::FreqScales> sc(static_cast<size_t>(g.count));
    std::vector<aphi_dev::BodyCoef> coef(static_cast<size_t>(m->num_bodies));
    for (int sg = 0; sg < g.count; ++sg) {
        const double omega = omegas[sg];
        sc[static_cast<size_t>(sg)] = freq_scales(omega, conditioning, g.values[static_cast<size_t>(sg)]);
        // element_coefficients, src/element_matrix.cpp:10-25
        for (int64_t b = g.bodies[static_cast<size_t>(sg)]; b < g.bodies[static_cast<size_t>(sg) + 1]; ++b) {
            const double mu = m->mu_r[static_cast<size_t>(b)] * kMu0;
            const double eps = m->eps_r[static_cast<size_t>(b)] * kEps0;
            coef[static_cast<size_t>(b)].nu = 1.0 / mu;
            coef[static_cast<size_t>(b)].alpha = cd{-omega * omega * eps, omega * m->sigma[static_cast<size_t>(b)]};
            coef[static_cast<size_t>(b)].beta = cd{m->sigma[static_cast<size_t>(b)], omega * eps};
        }
    }
    hipStream_t s = m->st();
    m->coef.reserve(coef.size() * sizeof(aphi_dev::BodyCoef));
    HIPCHK(hipMemcpyAsync(m->coef.p, coef.data(), coef.size() * sizeof(aphi_dev::BodyCoef), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g.scales.p, sc.data(), sc.size() * sizeof(aphi_dev::FreqScales), hipMemcpyHostToDevice, s));
    m->E.reserve(static_cast<size_t>(m->num_tets) * 196 * sizeof(cd));
    m->A.reserve(static_cast<size_t>(m->nnz) * sizeof(cd));
    m->rhs.reserve(static_cast<size_t>(m->n) * sizeof(cd));
    m->flag.reserve(sizeof(int));
    hipLaunchKernelGGL(aphi_dev::k_fill_flag, dim3(1), dim3(64), 0, s, m->flag.as<int>());
    hipLaunchKernelGGL(aphi_dev::k_element, dim3(blocks_for(m->num_tets, 16)), dim3(256), 0, s, m->num_tets,
                       m->coords.as<double>(), m->tets.as<int>(), m->tet_body.as<int>(),
                       m->coef.as<aphi_dev::BodyCoef>(), m->E.as<cd>(), m->flag.as<int>());
    hipLaunchKernelGGL(aphi_dev::k_assemble_reduce, dim3(blocks_for(m->num_targets)), dim3(256), 0, s, m->num_targets,
                       m->nnz, m->tptr.as<int64_t>(), m->entry.as<uint32_t>(), m->kind.as<uint8_t>(), m->E.as<cd>(),
                       g.scales.as<aphi_dev::FreqScales>(), g.count, g.d_slots.as<int64_t>(), g.d_rows.as<int64_t>(),
                       m->A.as<cd>(), m->rhs.as<cd>());
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, m->flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    if (flag) throw std::runtime_error("compute_tet_geometry: degenerate (zero-volume) tet");
    m->assembled = true;
    m->equilibrated = false;
    m->factored = false;
    compute_amax(m);
}
}  // namespace

extern "C" int aphi_refill(aphi_mesh* m, double omega, int conditioning) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_refill: null mesh");
        // one frequency for every problem the mesh holds
        const std::vector<double> omegas(static_cast<size_t>(m->seg.count), omega);
        refill_all(m, omegas.data(), conditioning);
    });
}

extern "C" int aphi_system_download(aphi_mesh* m, double* values, double* rhs) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_system_download: null mesh");
        require(m->assembled, "aphi_system_download: nothing assembled");
        if (values) HIPCHK(hipMemcpyAsync(values, m->A.p, static_cast<size_t>(m->nnz) * sizeof(cd), hipMemcpyDeviceToHost, m->st()));
        if (rhs) HIPCHK(hipMemcpyAsync(rhs, m->rhs.p, static_cast<size_t>(m->n) * sizeof(cd), hipMemcpyDeviceToHost, m->st()));
        HIPCHK(hipStreamSynchronize(m->st()));
    });
}

extern "C" int aphi_system_upload(aphi_mesh* m, const double* values, const double* rhs) {
    return aphi_capi::guarded([&] {
        require(m && values && rhs, "aphi_system_upload: null argument");
        m->A.reserve(static_cast<size_t>(m->nnz) * sizeof(cd));
        m->rhs.reserve(static_cast<size_t>(m->n) * sizeof(cd));
        HIPCHK(hipMemcpyAsync(m->A.p, values, static_cast<size_t>(m->nnz) * sizeof(cd), hipMemcpyHostToDevice, m->st()));
        HIPCHK(hipMemcpyAsync(m->rhs.p, rhs, static_cast<size_t>(m->n) * sizeof(cd), hipMemcpyHostToDevice, m->st()));
        HIPCHK(hipStreamSynchronize(m->st()));
        m->assembled = true;
        m->equilibrated = false;
        m->factored = false;
        compute_amax(m);
    });
}

extern "C" const void* aphi_rhs_device(const aphi_mesh* m) { return m ? m->rhs.p : nullptr; }

extern "C" int aphi_rhs_copy(aphi_mesh* m, void* dst) {
    return aphi_capi::guarded([&] {
        require(m && dst, "aphi_rhs_copy: null argument");
        require(m->assembled, "aphi_rhs_copy: nothing assembled");
        after_caller(m->ctx);
        HIPCHK(hipMemcpyAsync(dst, m->rhs.p, static_cast<size_t>(m->n) * sizeof(cd), hipMemcpyDeviceToDevice, m->st()));
        HIPCHK(hipStreamSynchronize(m->st()));
    });
}


namespace {
// terminal_current + impedance of nrhs solutions of every problem of the mesh:
// problem s at omegas[s], length lengths[s], drive voltage voltages[s]; outputs
// per column c = r * count + s. The per-tet terms go to `terms` (num_tets * nrhs
// entries), the sums to `sums` (nrhs * count entries); `factors` (2 * count,
// device) carries a batch's per-problem -j w and Phi scale.
void terminal_current_into(aphi_mesh* m, int nrhs, const void* xv, int64_t ldx, const double* omegas, int conditioning,
                           const double* axis, const double* lengths, const double* voltages, double* current,
                           double* resistance, double* inductance, cd* terms, cd* sums, cd* factors,
                           double* d_lengths) {
    hipStream_t s = m->st();
    const int nt = m->num_tets, ns = m->seg.count;
    aphi_dev::TetFieldPlan tp{m->coords.as<double>(), m->tets.as<int>(), m->tet_body.as<int>(), m->bsigma.as<double>(),
                              m->tdofs.as<int>(), m->tsign.as<int8_t>(), m->tcoef.as<double>(), m->tfixed.as<cd>()};
    // -jw and the ScaledPhi substitution factor (Solution::phi_scale, src/postprocess.cpp:35-40)
    const auto minus_jw = [](double w) { return cd{-0.0, -w}; };
    const auto phi_scale = [&](double w) { return conditioning == APHI_SCALED_PHI ? cd{0.0, w} : cd{1.0, 0.0}; };
    if (ns > 1) {
        std::vector<cd> f(2 * static_cast<size_t>(ns));
        for (int sg = 0; sg < ns; ++sg) {
            f[2 * static_cast<size_t>(sg)] = minus_jw(omegas[sg]);
            f[2 * static_cast<size_t>(sg) + 1] = phi_scale(omegas[sg]);
        }
        HIPCHK(hipMemcpyAsync(factors, f.data(), f.size() * sizeof(cd), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_lengths, lengths, static_cast<size_t>(ns) * sizeof(double), hipMemcpyHostToDevice, s));
    }
    // every term the sums read is written here first: nothing to clear
    hipLaunchKernelGGL(aphi_dev::k_tet_current, dim3(blocks_for(nt), nrhs), dim3(256), 0, s, nt, tp, nrhs,
                       static_cast<const cd*>(xv), ldx, minus_jw(omegas[0]), phi_scale(omegas[0]), axis[0], axis[1],
                       axis[2], ns, m->seg.d_tets.as<int64_t>(), factors, terms);
    hipLaunchKernelGGL(aphi_dev::k_current_sum, dim3(nrhs * ns), dim3(64), 0, s, nt, ns, m->seg.d_tets.as<int64_t>(), tp,
                       terms, ns > 1 ? d_lengths : nullptr, lengths[0], sums);
    std::vector<cd> h(static_cast<size_t>(nrhs) * ns);
    HIPCHK(hipMemcpyAsync(h.data(), sums, h.size() * sizeof(cd), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    for (size_t c = 0; c < h.size(); ++c) {
        const int sg = static_cast<int>(c % static_cast<size_t>(ns));
        current[2 * c] = h[c].re;
        current[2 * c + 1] = h[c].im;
        const Complex z = div_like_gcc(Complex(voltages[sg], 0.0), Complex(h[c].re, h[c].im));
        if (resistance) resistance[c] = z.real();
        if (inductance) inductance[c] = z.imag() / omegas[sg];
    }
}

// the same for a mesh of one problem, sums in a fixed array (no allocation)
void terminal_current_one(aphi_mesh* m, int nrhs, const void* xv, int64_t ldx, double omega, int conditioning,
                          const double* axis, double length, double voltage, double* current, double* resistance,
                          double* inductance, cd* terms, cd* sums) {
    hipStream_t s = m->st();
    const int nt = m->num_tets;
    aphi_dev::TetFieldPlan tp{m->coords.as<double>(), m->tets.as<int>(), m->tet_body.as<int>(), m->bsigma.as<double>(),
                              m->tdofs.as<int>(), m->tsign.as<int8_t>(), m->tcoef.as<double>(), m->tfixed.as<cd>()};
    const cd minus_jw{-0.0, -omega};
    const cd scale = conditioning == APHI_SCALED_PHI ? cd{0.0, omega} : cd{1.0, 0.0};
    // every term the sum reads is written here first: nothing to clear
    hipLaunchKernelGGL(aphi_dev::k_tet_current, dim3(blocks_for(nt), nrhs), dim3(256), 0, s, nt, tp, nrhs,
                       static_cast<const cd*>(xv), ldx, minus_jw, scale, axis[0], axis[1], axis[2], 1,
                       m->seg.d_tets.as<int64_t>(), nullptr, terms);
    hipLaunchKernelGGL(aphi_dev::k_current_sum, dim3(nrhs), dim3(64), 0, s, nt, 1, m->seg.d_tets.as<int64_t>(), tp, terms,
                       nullptr, length, sums);
    cd h[64];
    HIPCHK(hipMemcpyAsync(h, sums, static_cast<size_t>(nrhs) * sizeof(cd), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    for (int r = 0; r < nrhs; ++r) {
        current[2 * r] = h[r].re;
        current[2 * r + 1] = h[r].im;
        const Complex z = div_like_gcc(Complex(voltage, 0.0), Complex(h[r].re, h[r].im));
        if (resistance) resistance[r] = z.real();
        if (inductance) inductance[r] = z.imag() / omega;
    }
}

void check_terminal_current(const aphi_mesh* m, const void* xv, int nrhs, int64_t ldx, int conditioning,
                            const double* axis, const double* current, const char* who) {
    const std::string w(who);
    require(m && xv && axis && current, (w + ": null argument").c_str());
    require(nrhs >= 1 && nrhs <= 64 && ldx >= m->n, (w + ": nrhs must be 1..64 and ldx >= n").c_str());
    require(conditioning >= APHI_NATURAL && conditioning <= APHI_SCALED_PHI, (w + ": unknown conditioning").c_str());
    require(m->tdofs.p != nullptr, (w + ": the mesh was created without per-tet unknowns").c_str());
}

void require_single(const aphi_mesh* m, const char* who) {
    if (m && m->seg.count != 1)
        throw std::invalid_argument(std::string(who) + ": the mesh is a batch of " + std::to_string(m->seg.count) +
                                    " problems; use the aphi_batch_ entry points");
}
}  // namespace

extern "C" int aphi_terminal_current(aphi_mesh* m, int nrhs, const void* xv, int64_t ldx, double omega, int conditioning,
                                     const double* axis, double length, double voltage, double* current,
                                     double* resistance, double* inductance) {
    return aphi_capi::guarded([&] {
        check_terminal_current(m, xv, nrhs, ldx, conditioning, axis, current, "aphi_terminal_current");
        require_single(m, "aphi_terminal_current");
        if (!(length > 0.0)) throw std::invalid_argument("terminal_current: length must be positive");
        after_caller(m->ctx);
        m->tcontrib.reserve(static_cast<size_t>(m->num_tets) * nrhs * sizeof(cd));
        m->tcurrent.reserve(static_cast<size_t>(nrhs) * sizeof(cd));
        terminal_current_one(m, nrhs, xv, ldx, omega, conditioning, axis, length, voltage, current, resistance,
                             inductance, m->tcontrib.as<cd>(), m->tcurrent.as<cd>());
    });
}

extern "C" int aphi_terminal_current_noalloc(aphi_mesh* m, int nrhs, const void* xv, int64_t ldx, double omega,
                                             int conditioning, const double* axis, double length, double voltage,
                                             double* current, double* resistance, double* inductance) {
    return aphi_capi::guarded([&] {
        check_terminal_current(m, xv, nrhs, ldx, conditioning, axis, current, "aphi_terminal_current_noalloc");
        require_single(m, "aphi_terminal_current_noalloc");
        if (!(length > 0.0)) throw std::invalid_argument("terminal_current: length must be positive");
        if (static_cast<int64_t>(m->num_tets) * nrhs > kNoAllocTerms)
            throw std::invalid_argument("aphi_terminal_current_noalloc: num_tets * nrhs = " +
                                        std::to_string(static_cast<int64_t>(m->num_tets) * nrhs) +
                                        " exceeds the scratch of " + std::to_string(kNoAllocTerms) +
                                        " terms; pass fewer right-hand sides per call");
        require(m->ctx->scratch != nullptr, "aphi_terminal_current_noalloc: the context has no scratch address");
        after_caller(m->ctx);
        terminal_current_one(m, nrhs, xv, ldx, omega, conditioning, axis, length, voltage, current, resistance,
                             inductance, m->ctx->scratch, m->ctx->scratch + kNoAllocTerms);
    });
}

extern "C" int aphi_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    return aphi_capi::guarded([&] {
        require(ctr && key && out, "aphi_philox4x32_10: null argument");
        const aphi_rng::Block b = aphi_rng::philox4x32_10(ctr, key);
        for (int i = 0; i < 4; ++i) out[i] = b.w[i];
    });
}

extern "C" int aphi_random_normal_host(uint64_t seed, uint64_t offset, int64_t count, double* out) {
    return aphi_capi::guarded([&] {
        require(count >= 0 && (count == 0 || out), "aphi_random_normal_host: bad arguments");
        for (int64_t i = 0; i < count; ++i) {
            const uint64_t k = offset + static_cast<uint64_t>(i);
            double z0, z1;
            aphi_rng::normal_pair(seed, k >> 1, z0, z1);
            out[i] = (k & 1) ? z1 : z0;
        }
    });
}

extern "C" int aphi_random_uniform_host(uint64_t seed, uint64_t offset, int64_t count, double* out) {
    return aphi_capi::guarded([&] {
        require(count >= 0 && (count == 0 || out), "aphi_random_uniform_host: bad arguments");
        for (int64_t i = 0; i < count; ++i) {
            const uint64_t k = offset + static_cast<uint64_t>(i);
            double u0, u1;
            aphi_rng::uniform_pair(seed, k >> 1, u0, u1);
            out[i] = (k & 1) ? u1 : u0;
        }
    });
}

extern "C" int aphi_random_normal(aphi_context* ctx, uint64_t seed, uint64_t offset, int64_t count, void* out) {
    return aphi_capi::guarded([&] {
        require(ctx && count >= 0 && (count == 0 || out), "aphi_random_normal: bad arguments");
        if (count == 0) return;
        HIPCHK(hipSetDevice(ctx->device));
        after_caller(ctx);
        hipLaunchKernelGGL(aphi_dev::k_random_normal, dim3(blocks_for(count)), dim3(256), 0, ctx->stream, seed, offset,
                           count, static_cast<double*>(out));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipGetLastError());
    });
}

extern "C" int aphi_tet_fields(aphi_mesh* m, const void* xv, double omega, int conditioning, void* e_tet) {
    return aphi_capi::guarded([&] {
        require(m && xv && e_tet, "aphi_tet_fields: null argument");
        require_single(m, "aphi_tet_fields");
        require(conditioning >= APHI_NATURAL && conditioning <= APHI_SCALED_PHI, "aphi_tet_fields: unknown conditioning");
        require(m->tdofs.p != nullptr, "aphi_tet_fields: the mesh was created without per-tet unknowns");
        after_caller(m->ctx);
        HIPCHK(hipSetDevice(m->ctx->device));
        const int nt = m->num_tets;
        aphi_dev::TetFieldPlan tp{m->coords.as<double>(), m->tets.as<int>(), m->tet_body.as<int>(), m->bsigma.as<double>(),
                                  m->tdofs.as<int>(), m->tsign.as<int8_t>(), m->tcoef.as<double>(), m->tfixed.as<cd>()};
        const cd minus_jw{-0.0, -omega};
        const cd scale = conditioning == APHI_SCALED_PHI ? cd{0.0, omega} : cd{1.0, 0.0};
        hipLaunchKernelGGL(aphi_dev::k_tet_efield, dim3(blocks_for(nt)), dim3(256), 0, m->st(), nt, tp,
                           static_cast<const cd*>(xv), minus_jw, scale, static_cast<cd*>(e_tet));
        HIPCHK(hipStreamSynchronize(m->st()));
        HIPCHK(hipGetLastError());
    });
}

extern "C" int aphi_fields_attach(aphi_mesh* m, const aphi_fields_desc* d) {
    return aphi_capi::guarded([&] {
        require(m && d, "aphi_fields_attach: null argument");
        require_single(m, "aphi_fields_attach");
        require(m->tdofs.p != nullptr, "aphi_fields_attach: the mesh was created without per-tet unknowns");
        require(d->num_tets == m->num_tets && d->num_p2 > 0, "aphi_fields_attach: the plan is not this mesh's");
        require(d->node_ptr && d->node_entry, "aphi_fields_attach: null array");
        const int np = d->num_p2, nt = d->num_tets;
        // the device indexes with these: check them here
        require(d->node_ptr[0] == 0 && d->node_ptr[np] == 10 * static_cast<int64_t>(nt),
                "aphi_fields_attach: node_ptr must run from 0 to 10 * num_tets");
        for (int u = 0; u < np; ++u)
            require(d->node_ptr[u] <= d->node_ptr[u + 1], "aphi_fields_attach: node_ptr must not decrease");
        for (int64_t e = 0; e < 10 * static_cast<int64_t>(nt); ++e)
            require(d->node_entry[e] >= 0 && d->node_entry[e] < 10 * nt, "aphi_fields_attach: an entry is out of range");
        HIPCHK(hipSetDevice(m->ctx->device));
        upload(m->st(), m->fptr, d->node_ptr, static_cast<size_t>(np) + 1);
        upload(m->st(), m->fentry, d->node_entry, 10 * static_cast<size_t>(nt));
        upload(m->st(), m->bmu, m->mu_r.data(), m->mu_r.size());
        m->fields_np = np;
    });
}

extern "C" int aphi_compute_fields(aphi_mesh* m, const void* xv, double omega, int conditioning, aphi_fields* out) {
    return aphi_capi::guarded([&] {
        require(m && xv && out, "aphi_compute_fields: null argument");
        require_single(m, "aphi_compute_fields");
        require(conditioning >= APHI_NATURAL && conditioning <= APHI_SCALED_PHI, "aphi_compute_fields: unknown conditioning");
        require(m->fields_np > 0, "aphi_compute_fields: attach the field plan first (aphi_fields_attach)");
        const double* host[] = {out->a_tet,  out->b_tet,  out->e_tet,  out->j_tet,  out->sigma_tet,    out->a_node,
                                out->b_node, out->h_node, out->e_node, out->j_node, out->vertex_weight};
        for (const double* h : host) require(h != nullptr, "aphi_compute_fields: a field array is null");
        require(out->on_material_interface != nullptr, "aphi_compute_fields: a field array is null");
        after_caller(m->ctx);
        HIPCHK(hipSetDevice(m->ctx->device));
        const int nt = m->num_tets, np = m->fields_np;
        const size_t ntc = 3 * static_cast<size_t>(nt) * sizeof(cd), npc = 3 * static_cast<size_t>(np) * sizeof(cd);
        m->fval.reserve(60 * static_cast<size_t>(nt) * sizeof(cd));
        for (DevBuf* b : {&m->fvol, &m->fmu, &m->fsig}) b->reserve(static_cast<size_t>(nt) * sizeof(double));
        for (DevBuf* b : {&m->fa_tet, &m->fb_tet, &m->fe_tet, &m->fj_tet}) b->reserve(ntc);
        for (DevBuf* b : {&m->fa_node, &m->fb_node, &m->fh_node, &m->fe_node, &m->fj_node}) b->reserve(npc);
        m->fweight.reserve(static_cast<size_t>(np) * sizeof(double));
        m->fflag.reserve(static_cast<size_t>(np));
        aphi_dev::TetFieldPlan tp{m->coords.as<double>(), m->tets.as<int>(), m->tet_body.as<int>(), m->bsigma.as<double>(),
                                  m->tdofs.as<int>(), m->tsign.as<int8_t>(), m->tcoef.as<double>(), m->tfixed.as<cd>()};
        const cd minus_jw{-0.0, -omega};
        const cd scale = conditioning == APHI_SCALED_PHI ? cd{0.0, omega} : cd{1.0, 0.0};
        hipStream_t s = m->st();
        hipLaunchKernelGGL(aphi_dev::k_fields_tet, dim3(blocks_for(nt)), dim3(256), 0, s, nt, tp, m->bmu.as<double>(),
                           static_cast<const cd*>(xv), minus_jw, scale, m->fa_tet.as<cd>(), m->fb_tet.as<cd>(),
                           m->fe_tet.as<cd>(), m->fj_tet.as<cd>(), m->fsig.as<double>(), m->fvol.as<double>(),
                           m->fmu.as<double>(), m->fval.as<cd>());
        hipLaunchKernelGGL(aphi_dev::k_fields_node, dim3(blocks_for(np)), dim3(256), 0, s, np, m->fptr.as<int64_t>(),
                           m->fentry.as<int>(), m->fsig.as<double>(), m->fmu.as<double>(), m->fvol.as<double>(),
                           m->fb_tet.as<cd>(), m->fval.as<cd>(), m->fa_node.as<cd>(), m->fb_node.as<cd>(),
                           m->fh_node.as<cd>(), m->fe_node.as<cd>(), m->fj_node.as<cd>(), m->fweight.as<double>(),
                           m->fflag.as<int8_t>());
        HIPCHK(hipGetLastError());
        const struct {
            double* dst;
            const DevBuf* src;
            size_t bytes;
        } copies[] = {{out->a_tet, &m->fa_tet, ntc},   {out->b_tet, &m->fb_tet, ntc},
                      {out->e_tet, &m->fe_tet, ntc},   {out->j_tet, &m->fj_tet, ntc},
                      {out->sigma_tet, &m->fsig, static_cast<size_t>(nt) * sizeof(double)},
                      {out->a_node, &m->fa_node, npc}, {out->b_node, &m->fb_node, npc},
                      {out->h_node, &m->fh_node, npc}, {out->e_node, &m->fe_node, npc},
                      {out->j_node, &m->fj_node, npc},
                      {out->vertex_weight, &m->fweight, static_cast<size_t>(np) * sizeof(double)}};
        for (const auto& c : copies) HIPCHK(hipMemcpyAsync(c.dst, c.src->p, c.bytes, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(out->on_material_interface, m->fflag.p, static_cast<size_t>(np), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        int32_t interface_nodes = 0, orphans = 0;
        for (int u = 0; u < np; ++u) {
            if (out->on_material_interface[u]) ++interface_nodes;
            if (!(out->vertex_weight[u] > 0.0)) ++orphans;
        }
        out->num_interface_nodes = interface_nodes;
        out->num_orphan_vertices = orphans;
    });
}

extern "C" int aphi_gauge_free_attach(aphi_mesh* m, const aphi_gauge_free_desc* d) {
    return aphi_capi::guarded([&] {
        require(m && d, "aphi_gauge_free_attach: null argument");
        require(!m->has_plan, "aphi_gauge_free_attach: the mesh must come from the gauge-free problem");
        require(d->unknowns == m->n && d->nnz == m->nnz, "aphi_gauge_free_attach: the plan is not this mesh's pattern");
        require(d->num_nodes > 0 && d->num_tets > 0 && d->num_bodies > 0 && d->num_ports >= 0 && d->num_values >= 0 &&
                    d->num_values <= 64,
                "aphi_gauge_free_attach: bad sizes");
        require(d->coords && d->tet_nodes && d->tet_body && d->body_sigma && d->body_eps_r && d->tet_on && d->column &&
                    d->prescribed && d->target_ptr && d->entry && d->kind && (d->num_values == 0 || d->values) &&
                    (d->num_ports == 0 || (d->port_column && d->port_value)),
                "aphi_gauge_free_attach: null array");
        HIPCHK(hipSetDevice(m->ctx->device));
        auto& g = m->gf;
        const int nv = d->num_nodes, nt = d->num_tets;
        const int64_t targets = d->nnz + d->unknowns;
        // the device indexes with these: check them here
        for (int64_t i = 0; i < 4 * static_cast<int64_t>(nt); ++i)
            require(d->tet_nodes[i] >= 0 && d->tet_nodes[i] < nv, "aphi_gauge_free_attach: a tet names a missing vertex");
        for (int t = 0; t < nt; ++t)
            require(d->tet_body[t] >= 0 && d->tet_body[t] < d->num_bodies, "aphi_gauge_free_attach: a tet's body is out of range");
        for (int v = 0; v < nv; ++v)
            require(d->column[v] >= -1 && d->column[v] < d->unknowns, "aphi_gauge_free_attach: a vertex column is out of range");
        for (int p = 0; p < d->num_ports; ++p)
            require(d->port_column[p] >= -1 && d->port_column[p] < d->unknowns, "aphi_gauge_free_attach: a port column is out of range");
        require(d->target_ptr[0] == 0, "aphi_gauge_free_attach: target_ptr must start at 0");
        for (int64_t t = 0; t < targets; ++t)
            require(d->target_ptr[t + 1] >= d->target_ptr[t], "aphi_gauge_free_attach: target_ptr must not decrease");
        const int64_t nc = d->target_ptr[targets];
        for (int64_t c = 0; c < nc; ++c) {
            require(d->entry[c] < static_cast<uint32_t>(nt) * 20u, "aphi_gauge_free_attach: a term names a tet outside the mesh");
            const int k = d->kind[c];
            require((k & 3) != 3 && ((k & 3) != 2 || (k >> 2) < d->num_values), "aphi_gauge_free_attach: bad term kind");
            const uint32_t local = d->entry[c] % 20u;
            require(c < d->target_ptr[d->nnz] ? (k == 0 && local < 16u) : ((k & 3) == 1 ? local >= 16u : ((k & 3) == 2 && local < 16u)),
                    "aphi_gauge_free_attach: a term does not fit its target");
            require(d->tet_on[d->entry[c] / 20u] != 0, "aphi_gauge_free_attach: a term names a tet that takes no part");
        }
        g.num_nodes = nv;
        g.num_tets = nt;
        g.num_bodies = d->num_bodies;
        g.num_targets = targets;
        g.sigma.assign(d->body_sigma, d->body_sigma + d->num_bodies);
        g.eps_r.assign(d->body_eps_r, d->body_eps_r + d->num_bodies);
        g.port_column.assign(d->port_column, d->port_column + d->num_ports);
        g.port_value.clear();
        for (int p = 0; p < d->num_ports; ++p) g.port_value.emplace_back(d->port_value[2 * p], d->port_value[2 * p + 1]);
        upload(m->st(), g.coords, d->coords, static_cast<size_t>(nv) * 3);
        upload(m->st(), g.tets, d->tet_nodes, static_cast<size_t>(nt) * 4);
        upload(m->st(), g.tet_body, d->tet_body, static_cast<size_t>(nt));
        upload(m->st(), g.tet_on, d->tet_on, static_cast<size_t>(nt));
        upload(m->st(), g.column, d->column, static_cast<size_t>(nv));
        upload(m->st(), g.prescribed, d->prescribed, static_cast<size_t>(nv) * 2);
        upload(m->st(), g.tptr, d->target_ptr, static_cast<size_t>(targets) + 1);
        upload(m->st(), g.entry, d->entry, static_cast<size_t>(nc));
        upload(m->st(), g.kind, d->kind, static_cast<size_t>(nc));
        upload(m->st(), g.values, d->values, static_cast<size_t>(d->num_values) * 2);
        g.terms.reserve(static_cast<size_t>(nt) * 20 * sizeof(cd));
        g.attached = true;
    });
}

extern "C" int aphi_gauge_free_refill(aphi_mesh* m, const void* e_tet, double omega) {
    return aphi_capi::guarded([&] {
        require(m && e_tet, "aphi_gauge_free_refill: null argument");
        require(m->gf.attached, "aphi_gauge_free_refill: no gauge-free plan attached");
        after_caller(m->ctx);
        if (omega == 0.0)
            throw std::invalid_argument("aphi_gauge_free_refill: omega must be nonzero (the symmetric A-Phi solve that "
                                        "produced E is undefined at DC)");
        HIPCHK(hipSetDevice(m->ctx->device));
        auto& g = m->gf;
        // beta = sigma + j w eps0 eps_r per body, as src/gauge_free_potential.cpp:117 forms it
        std::vector<cd> beta(static_cast<size_t>(g.num_bodies));
        for (int b = 0; b < g.num_bodies; ++b)
            beta[static_cast<size_t>(b)] = cd{g.sigma[static_cast<size_t>(b)], omega * kEps0 * g.eps_r[static_cast<size_t>(b)]};
        hipStream_t s = m->st();
        g.beta.reserve(beta.size() * sizeof(cd));
        HIPCHK(hipMemcpyAsync(g.beta.p, beta.data(), beta.size() * sizeof(cd), hipMemcpyHostToDevice, s));
        m->A.reserve(static_cast<size_t>(m->nnz) * sizeof(cd));
        m->rhs.reserve(static_cast<size_t>(m->n) * sizeof(cd));
        hipLaunchKernelGGL(aphi_dev::k_gf_tet, dim3(blocks_for(g.num_tets)), dim3(256), 0, s, g.num_tets,
                           g.coords.as<double>(), g.tets.as<int>(), g.tet_body.as<int>(), g.tet_on.as<int>(),
                           g.beta.as<cd>(), static_cast<const cd*>(e_tet), g.terms.as<cd>());
        hipLaunchKernelGGL(aphi_dev::k_gf_reduce, dim3(blocks_for(g.num_targets)), dim3(256), 0, s, g.num_targets, m->nnz,
                           g.tptr.as<int64_t>(), g.entry.as<uint32_t>(), g.kind.as<uint8_t>(), g.values.as<cd>(),
                           g.terms.as<cd>(), m->A.as<cd>(), m->rhs.as<cd>());
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipGetLastError());
        m->assembled = true;
        m->equilibrated = false;
        m->factored = false;
        compute_amax(m);
    });
}

extern "C" int aphi_gauge_free_potential(aphi_mesh* m, const void* xv, void* phi_vertex, double* port_potential) {
    return aphi_capi::guarded([&] {
        require(m && xv && phi_vertex, "aphi_gauge_free_potential: null argument");
        require(m->gf.attached, "aphi_gauge_free_potential: no gauge-free plan attached");
        after_caller(m->ctx);
        HIPCHK(hipSetDevice(m->ctx->device));
        auto& g = m->gf;
        hipStream_t s = m->st();
        hipLaunchKernelGGL(aphi_dev::k_gf_scatter, dim3(blocks_for(g.num_nodes)), dim3(256), 0, s, g.num_nodes,
                           g.column.as<int>(), g.prescribed.as<cd>(), static_cast<const cd*>(xv), static_cast<cd*>(phi_vertex));
        // port_potential (src/gauge_free_potential.cpp:172-176)
        std::vector<cd> px(g.port_column.size(), cd{0.0, 0.0});
        for (size_t p = 0; p < g.port_column.size(); ++p)
            if (g.port_column[p] >= 0)
                HIPCHK(hipMemcpyAsync(&px[p], static_cast<const cd*>(xv) + g.port_column[p], sizeof(cd),
                                      hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipGetLastError());
        if (port_potential)
            for (size_t p = 0; p < g.port_column.size(); ++p) {
                const bool fixed = g.port_column[p] < 0;
                port_potential[2 * p] = fixed ? g.port_value[p].real() : px[p].re;
                port_potential[2 * p + 1] = fixed ? g.port_value[p].imag() : px[p].im;
            }
    });
}

extern "C" int aphi_equilibrate(aphi_mesh* m, int iterations, double* d_out) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_equilibrate: null mesh");
        equilibrate(m, iterations, d_out);
    });
}

extern "C" int aphi_factorize(aphi_mesh* m, int precision, double pivot_floor, aphi_factor_stats* stats) {
    return aphi_capi::guarded([&] {
        require(m != nullptr, "aphi_factorize: null mesh");
        aphi_factor_stats st{};
        factorize(m, precision, pivot_floor, &st);
        if (stats) *stats = st;
        if (!st.ok) throw PivotFailure("factorize_ldlt: a pivot at or below the floor stopped the factorization");
    });
}

extern "C" int aphi_factor_diagonal(aphi_mesh* m, double* out) {
    return aphi_capi::guarded([&] {
        require(m && out, "aphi_factor_diagonal: null argument");
        require(m->factor_precision >= 0 && m->P.p, "aphi_factor_diagonal: nothing factored");
        // D sits on each panel's diagonal: column c of supernode s at panel_off[s] + c*m + c
        std::vector<int> first(static_cast<size_t>(m->num_super) + 1);
        std::vector<int64_t> rp(static_cast<size_t>(m->num_super) + 1), po(static_cast<size_t>(m->num_super) + 1);
        const size_t cs = m->factor_precision == APHI_FP64 ? sizeof(cd) : sizeof(aphi_dev::cf);
        std::vector<unsigned char> host(static_cast<size_t>(m->factor_entries) * cs);
        hipStream_t s = m->st();
        HIPCHK(hipMemcpyAsync(first.data(), m->first.p, first.size() * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(rp.data(), m->srow_ptr.p, rp.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(po.data(), m->panel_off.p, po.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(host.data(), m->P.p, host.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int s = 0; s < m->num_super; ++s) {
            const int64_t mm = rp[static_cast<size_t>(s) + 1] - rp[static_cast<size_t>(s)];
            for (int c = 0; c < first[static_cast<size_t>(s) + 1] - first[static_cast<size_t>(s)]; ++c) {
                const int64_t at = po[static_cast<size_t>(s)] + c * mm + c;
                const int j = first[static_cast<size_t>(s)] + c;
                if (m->factor_precision == APHI_FP64) {
                    const cd* v = reinterpret_cast<const cd*>(host.data()) + at;
                    out[2 * j] = v->re;
                    out[2 * j + 1] = v->im;
                } else {
                    const aphi_dev::cf* v = reinterpret_cast<const aphi_dev::cf*>(host.data()) + at;
                    out[2 * j] = v->re;
                    out[2 * j + 1] = v->im;
                }
            }
        }
    });
}

extern "C" int aphi_solve(aphi_mesh* m, int nrhs, const void* b, int64_t ldb, void* x, int64_t ldx) {
    return aphi_capi::guarded([&] {
        require(m && b && x, "aphi_solve: null argument");
        require(nrhs >= 1 && nrhs <= 64 && ldb >= m->n && ldx >= m->n, "aphi_solve: nrhs must be 1..64 and ld >= n");
        after_caller(m->ctx);
        solve_any(m, nrhs, static_cast<const cd*>(b), ldb, static_cast<cd*>(x), ldx, 0);
        HIPCHK(hipStreamSynchronize(m->st()));
    });
}

extern "C" int aphi_backward_error(aphi_mesh* m, int nrhs, const void* x, int64_t ldx, const void* b, int64_t ldb,
                                   double* be, double* rr) {
    return aphi_capi::guarded([&] {
        require(m && x && b, "aphi_backward_error: null argument");
        require_single(m, "aphi_backward_error");
        require(m->assembled, "aphi_backward_error: nothing assembled");
        require(nrhs >= 1 && nrhs <= 64 && ldb >= m->n && ldx >= m->n, "aphi_backward_error: nrhs must be 1..64 and ld >= n");
        after_caller(m->ctx);
        measure(m, nrhs, static_cast<const cd*>(x), ldx, static_cast<const cd*>(b), ldb, be, rr);
    });
}

namespace {
// solve_symmetric with refinement over every (problem, right-hand side) column;
// fills the report's shared fields (equilibration, factor, timings) and leaves
// the per-column measures in be / rr / bs. Throws PivotFailure after filling r.
void solve_symmetric_all(aphi_mesh* m, int nrhs, const cd* b, int64_t ldb, cd* x, int64_t ldx,
                         const aphi_solve_options* opt, aphi_solve_report& r, std::vector<double>& be,
                         std::vector<double>& rr, std::vector<double>& bs) {
    require(nrhs >= 1 && nrhs <= 64 && ldb >= m->n && ldx >= m->n, "aphi_solve_symmetric: nrhs must be 1..64 and ld >= n");
    require(opt->max_refinements >= 0, "aphi_solve_symmetric: max_refinements must be >= 0");
    after_caller(m->ctx);
    const auto t0 = std::chrono::steady_clock::now();
    r = aphi_solve_report{};
    equilibrate(m, opt->equilibration_iterations, nullptr);
    r.equilibration_min = m->eq_min;
    r.equilibration_max = m->eq_max;
    aphi_factor_stats fs{};
    factorize(m, opt->precision, opt->pivot_floor, &fs);
    r.smallest_pivot = fs.smallest_pivot;
    r.largest_pivot = fs.largest_pivot;
    r.largest_multiplier = fs.largest_multiplier;
    r.factor_nnz = fs.nnz;
    r.columns_done = fs.columns_done;
    r.factorize_ms = fs.milliseconds;
    if (!fs.ok) throw PivotFailure("solve_symmetric: the factorization hit a pivot at or below the floor");
    const auto s0 = std::chrono::steady_clock::now();
    solve_any(m, nrhs, b, ldb, x, ldx, 0);
    const size_t ncol = static_cast<size_t>(nrhs) * m->seg.count;
    be.assign(ncol, 0.0);
    rr.assign(ncol, 0.0);
    bs.assign(ncol, 0.0);
    int steps = 0, since_factor = 0;
    r.precision_used = opt->precision;
    if (opt->precision == APHI_FP32 && opt->max_refinements > 0) {
        // FP32 factor: GMRES refinement; if it does not converge within the
        // budget, the loop below refactors in FP64
        steps = gmres_refine(m, nrhs, b, ldb, x, ldx, opt->tolerance, opt->max_refinements, be, rr, bs);
        since_factor = opt->max_refinements;
    }
    // largest entry, +inf when any is NaN (so a non-finite solution never passes)
    const auto worst_of = [](const std::vector<double>& v) {
        double w = 0.0;
        for (double e : v) w = std::isnan(e) ? std::numeric_limits<double>::infinity() : std::max(w, e);
        return w;
    };
    for (;;) {
        measure(m, nrhs, x, ldx, b, ldb, be.data(), rr.data(), bs.data());
        const double worst = std::max(worst_of(be), worst_of(bs));
        const bool finite = std::isfinite(worst);
        if (opt->max_refinements == 0 && finite) break;
        if (worst <= opt->tolerance) break;
        if (since_factor >= opt->max_refinements || !finite) {
            if (m->factor_precision != APHI_FP32) break;
            // FP32 refinement did not converge (or overflowed): factor in FP64
            // and keep refining; a non-finite iterate restarts from b.
            aphi_factor_stats f64{};
            factorize(m, APHI_FP64, opt->pivot_floor, &f64);
            r.factorize_ms += f64.milliseconds;
            r.precision_used = APHI_FP64;
            since_factor = 0;
            if (!f64.ok) break;
            if (!finite) {
                solve_any(m, nrhs, b, ldb, x, ldx, 0);
                continue;
            }
            if (opt->max_refinements == 0) continue;
        }
        solve_any(m, nrhs, m->res.as<cd>(), m->n, x, ldx, 1);
        ++steps;
        ++since_factor;
    }
    const auto s1 = std::chrono::steady_clock::now();
    r.refinement_steps = steps;
    r.backward_error = worst_of(be);
    r.scaled_backward_error = worst_of(bs);
    r.residual = worst_of(rr);
    r.solve_ms = std::chrono::duration<double, std::milli>(s1 - s0).count();
    r.total_ms = std::chrono::duration<double, std::milli>(s1 - t0).count();
    r.ok = std::isfinite(r.backward_error) ? 1 : 0;
}
}  // namespace

extern "C" int aphi_solve_symmetric(aphi_mesh* m, int nrhs, const void* bv, int64_t ldb, void* xv, int64_t ldx,
                                    const aphi_solve_options* opt, aphi_solve_report* rep, double* backward_errors) {
    return aphi_capi::guarded([&] {
        require(m && bv && xv && opt && rep, "aphi_solve_symmetric: null argument");
        require_single(m, "aphi_solve_symmetric");
        aphi_solve_report r{};
        std::vector<double> be, rr, bs;
        try {
            solve_symmetric_all(m, nrhs, static_cast<const cd*>(bv), ldb, static_cast<cd*>(xv), ldx, opt, r, be, rr, bs);
        } catch (const PivotFailure&) {
            *rep = r;
            throw;
        }
        if (backward_errors) std::memcpy(backward_errors, be.data(), be.size() * sizeof(double));
        *rep = r;
    });
}

// ------------------------------------------------------------------ batches

namespace {
// The block-diagonal union of several meshes' descriptions: problem p's unknowns,
// stored entries, tets, bodies, supernodes, panels and contribution blocks follow
// those of problems 0 .. p-1; its integer references are shifted to match, its
// supernode levels are kept (so fronts of equal height of every problem share a
// level, children before parents), and its assembly targets keep their order
// (all problems' slots, then all problems' rhs rows).
struct UnionDesc {
    aphi_mesh_desc d{};
    std::vector<double> coords, sigma, eps_r, mu_r, values, phi_coeff, phi_fixed;
    std::vector<int32_t> tet_nodes, tet_body, row_ptr, col, fperm, col_ref, first, parent, rows, rel, level, dofs;
    std::vector<int64_t> tptr, srow_ptr, panel_off, cb_off, s2p;
    std::vector<uint32_t> entry;
    std::vector<uint8_t> kind;
    std::vector<int8_t> signs;
};

void build_union(const aphi_mesh_desc* ds, int count, UnionDesc& u, aphi_mesh::Segments& seg) {
    int64_t nodes = 0, tets = 0, bodies = 0, n = 0, nnz = 0, ns = 0, rows = 0, panel = 0, cb = 0, nc = 0;
    int levels = 0, nvals = 0;
    double flops = 0.0;
    int64_t pred = 0;
    seg.count = count;
    seg.rows.assign(1, 0);
    seg.slots.assign(1, 0);
    seg.tets.assign(1, 0);
    seg.bodies.assign(1, 0);
    seg.values.clear();
    for (int p = 0; p < count; ++p) {
        const aphi_mesh_desc& d = ds[p];
        require(d.n > 0 && d.nnz > 0 && d.num_super > 0 && d.num_tets > 0 && d.target_ptr && d.tet_dofs &&
                    d.tet_edge_signs && d.tet_phi_coeff && d.tet_phi_fixed,
                "aphi_batch_create: every problem needs a mesh with an assembly plan and per-tet unknowns");
        require(d.num_values >= 0 && d.num_values <= 8, "aphi_batch_create: at most 8 prescribed values per problem");
        nodes += d.num_nodes;
        tets += d.num_tets;
        bodies += d.num_bodies;
        n += d.n;
        nnz += d.nnz;
        ns += d.num_super;
        rows += d.super_row_ptr[d.num_super];
        panel += d.panel_offset[d.num_super];
        cb += d.cb_pool_entries;
        nc += d.target_ptr[d.nnz + d.n];
        levels = std::max(levels, d.num_levels);
        nvals = std::max(nvals, d.num_values);
        flops += d.flops;
        pred += d.predicted_nnz;
        seg.rows.push_back(n);
        seg.slots.push_back(nnz);
        seg.tets.push_back(tets);
        seg.bodies.push_back(bodies);
        std::vector<Complex> v;
        for (int i = 0; i < d.num_values; ++i) v.emplace_back(d.values[2 * i], d.values[2 * i + 1]);
        seg.values.push_back(v);
    }
    require(n < INT_MAX && nnz < INT_MAX && ns < INT_MAX && nodes < INT_MAX, "aphi_batch_create: the batch exceeds 32-bit indexing");
    require(tets * 196 < static_cast<int64_t>(UINT32_MAX), "aphi_batch_create: too many tets for the assembly plan");
    u.coords.reserve(static_cast<size_t>(3 * nodes));
    int64_t o_nodes = 0, o_tets = 0, o_bodies = 0, o_n = 0, o_nnz = 0, o_ns = 0, o_rows = 0, o_panel = 0, o_cb = 0;
    u.tptr.assign(1, 0);
    // slot targets of every problem first, then rhs targets of every problem
    for (int pass = 0; pass < 2; ++pass) {
        int64_t ot = 0;
        for (int p = 0; p < count; ++p) {
            const aphi_mesh_desc& d = ds[p];
            const int64_t t0 = pass == 0 ? 0 : d.nnz, t1 = pass == 0 ? d.nnz : d.nnz + d.n;
            for (int64_t t = t0; t < t1; ++t) {
                for (int64_t c = d.target_ptr[t]; c < d.target_ptr[t + 1]; ++c) {
                    const int kk = d.kind[c] & 0x0f;
                    u.entry.push_back(kk <= 5 ? d.entry[c] + static_cast<uint32_t>(ot) * 196u : d.entry[c]);
                    u.kind.push_back(d.kind[c]);
                }
                u.tptr.push_back(static_cast<int64_t>(u.entry.size()));
            }
            ot += d.num_tets;
        }
    }
    for (int p = 0; p < count; ++p) {
        const aphi_mesh_desc& d = ds[p];
        u.coords.insert(u.coords.end(), d.coords, d.coords + 3 * static_cast<int64_t>(d.num_nodes));
        for (int64_t i = 0; i < 4 * static_cast<int64_t>(d.num_tets); ++i)
            u.tet_nodes.push_back(d.tet_nodes[i] + static_cast<int32_t>(o_nodes));
        for (int t = 0; t < d.num_tets; ++t) u.tet_body.push_back(d.tet_body[t] + static_cast<int32_t>(o_bodies));
        u.sigma.insert(u.sigma.end(), d.body_sigma, d.body_sigma + d.num_bodies);
        u.eps_r.insert(u.eps_r.end(), d.body_eps_r, d.body_eps_r + d.num_bodies);
        u.mu_r.insert(u.mu_r.end(), d.body_mu_r, d.body_mu_r + d.num_bodies);
        for (int i = 0; i < d.n; ++i) u.row_ptr.push_back(d.row_ptr[i] + static_cast<int32_t>(o_nnz));
        for (int64_t k = 0; k < d.nnz; ++k) u.col.push_back(d.col_index[k] + static_cast<int32_t>(o_n));
        for (int i = 0; i < d.n; ++i) {
            u.fperm.push_back(d.factor_perm[i] + static_cast<int32_t>(o_n));
            u.col_ref.push_back(d.col_ref[i] + static_cast<int32_t>(o_n));
        }
        for (int sn = 0; sn < d.num_super; ++sn) {
            u.first.push_back(d.super_first[sn] + static_cast<int32_t>(o_n));
            u.parent.push_back(d.super_parent[sn] >= 0 ? d.super_parent[sn] + static_cast<int32_t>(o_ns) : -1);
            u.srow_ptr.push_back(d.super_row_ptr[sn] + o_rows);
            u.level.push_back(d.super_level[sn]);
            u.panel_off.push_back(d.panel_offset[sn] + o_panel);
            u.cb_off.push_back(d.cb_offset[sn] >= 0 ? d.cb_offset[sn] + o_cb : -1);
        }
        for (int64_t i = 0; i < d.super_row_ptr[d.num_super]; ++i) {
            u.rows.push_back(d.super_rows[i] + static_cast<int32_t>(o_n));
            u.rel.push_back(d.super_rel[i]);
        }
        for (int64_t k = 0; k < d.nnz; ++k) u.s2p.push_back(d.slot_to_panel[k] + o_panel);
        for (int64_t i = 0; i < 16 * static_cast<int64_t>(d.num_tets); ++i)
            u.dofs.push_back(d.tet_dofs[i] >= 0 ? d.tet_dofs[i] + static_cast<int32_t>(o_n) : -1);
        u.signs.insert(u.signs.end(), d.tet_edge_signs, d.tet_edge_signs + 6 * static_cast<int64_t>(d.num_tets));
        u.phi_coeff.insert(u.phi_coeff.end(), d.tet_phi_coeff, d.tet_phi_coeff + 10 * static_cast<int64_t>(d.num_tets));
        u.phi_fixed.insert(u.phi_fixed.end(), d.tet_phi_fixed, d.tet_phi_fixed + 20 * static_cast<int64_t>(d.num_tets));
        o_nodes += d.num_nodes;
        o_tets += d.num_tets;
        o_bodies += d.num_bodies;
        o_n += d.n;
        o_nnz += d.nnz;
        o_ns += d.num_super;
        o_rows += d.super_row_ptr[d.num_super];
        o_panel += d.panel_offset[d.num_super];
        o_cb += d.cb_pool_entries;
    }
    u.row_ptr.push_back(static_cast<int32_t>(nnz));
    u.first.push_back(static_cast<int32_t>(n));
    u.srow_ptr.push_back(rows);
    u.panel_off.push_back(panel);
    // the kind bytes keep each problem's own value ids; the refill takes the values
    // from the problem's scale table, so the union's table only sizes the check
    u.values.assign(2 * static_cast<size_t>(std::max(nvals, 1)), 0.0);
    aphi_mesh_desc& d = u.d;
    d.num_nodes = static_cast<int32_t>(nodes);
    d.num_tets = static_cast<int32_t>(tets);
    d.num_bodies = static_cast<int32_t>(bodies);
    d.n = static_cast<int32_t>(n);
    d.nnz = nnz;
    d.coords = u.coords.data();
    d.tet_nodes = u.tet_nodes.data();
    d.tet_body = u.tet_body.data();
    d.body_sigma = u.sigma.data();
    d.body_eps_r = u.eps_r.data();
    d.body_mu_r = u.mu_r.data();
    d.target_ptr = u.tptr.data();
    d.entry = u.entry.data();
    d.kind = u.kind.data();
    d.num_values = nvals;
    d.values = u.values.data();
    d.row_ptr = u.row_ptr.data();
    d.col_index = u.col.data();
    d.factor_perm = u.fperm.data();
    d.col_ref = u.col_ref.data();
    d.num_super = static_cast<int32_t>(ns);
    d.num_levels = levels;
    d.super_first = u.first.data();
    d.super_parent = u.parent.data();
    d.super_row_ptr = u.srow_ptr.data();
    d.super_rows = u.rows.data();
    d.super_rel = u.rel.data();
    d.super_level = u.level.data();
    d.panel_offset = u.panel_off.data();
    d.cb_offset = u.cb_off.data();
    d.cb_pool_entries = cb;
    d.slot_to_panel = u.s2p.data();
    d.flops = flops;
    d.predicted_nnz = pred;
    d.tet_dofs = u.dofs.data();
    d.tet_edge_signs = u.signs.data();
    d.tet_phi_coeff = u.phi_coeff.data();
    d.tet_phi_fixed = u.phi_fixed.data();
    (void)nc;
}

void require_batch_args(const aphi_mesh* m, const char* who) {
    require(m != nullptr, (std::string(who) + ": null mesh").c_str());
}
}  // namespace

extern "C" int aphi_batch_create(aphi_context* ctx, const aphi_mesh_desc* descs, int count, aphi_mesh** out) {
    return aphi_capi::guarded([&] {
        require(ctx && descs && out, "aphi_batch_create: null argument");
        require(count >= 1, "aphi_batch_create: count must be >= 1");
        *out = nullptr;
        UnionDesc u;
        aphi_mesh::Segments seg;
        build_union(descs, count, u, seg);
        auto m = build_mesh(ctx, &u.d);
        m->seg.count = seg.count;
        m->seg.rows = seg.rows;
        m->seg.slots = seg.slots;
        m->seg.tets = seg.tets;
        m->seg.bodies = seg.bodies;
        m->seg.values = seg.values;
        upload_segments(m.get());
        *out = m.release();
    });
}

extern "C" int aphi_batch_info(const aphi_mesh* m, int* count, int64_t* row_offsets, int64_t* tet_offsets) {
    return aphi_capi::guarded([&] {
        require_batch_args(m, "aphi_batch_info");
        if (count) *count = m->seg.count;
        if (row_offsets) std::memcpy(row_offsets, m->seg.rows.data(), m->seg.rows.size() * sizeof(int64_t));
        if (tet_offsets) std::memcpy(tet_offsets, m->seg.tets.data(), m->seg.tets.size() * sizeof(int64_t));
    });
}

extern "C" int aphi_batch_refill(aphi_mesh* m, const double* omegas, int conditioning) {
    return aphi_capi::guarded([&] {
        require_batch_args(m, "aphi_batch_refill");
        require(omegas != nullptr, "aphi_batch_refill: null frequencies");
        refill_all(m, omegas, conditioning);
    });
}

extern "C" int aphi_batch_solve_symmetric(aphi_mesh* m, int nrhs, const void* bv, int64_t ldb, void* xv, int64_t ldx,
                                          const aphi_solve_options* opt, aphi_solve_report* reports,
                                          double* backward_errors) {
    return aphi_capi::guarded([&] {
        require_batch_args(m, "aphi_batch_solve_symmetric");
        require(bv && xv && opt && reports, "aphi_batch_solve_symmetric: null argument");
        const int ns = m->seg.count;
        aphi_solve_report r{};
        std::vector<double> be, rr, bs;
        const auto fill = [&](bool have) {
            for (int sg = 0; sg < ns; ++sg) {
                aphi_solve_report q = r;
                if (have) {
                    double wb = 0.0, ws = 0.0, wr = 0.0;
                    const auto mx = [](double a, double e) { return std::isnan(e) ? std::numeric_limits<double>::infinity() : std::max(a, e); };
                    for (int c = sg; c < nrhs * ns; c += ns) {
                        wb = mx(wb, be[static_cast<size_t>(c)]);
                        ws = mx(ws, bs[static_cast<size_t>(c)]);
                        wr = mx(wr, rr[static_cast<size_t>(c)]);
                    }
                    q.backward_error = wb;
                    q.scaled_backward_error = ws;
                    q.residual = wr;
                    q.ok = std::isfinite(wb) ? 1 : 0;
                }
                reports[sg] = q;
            }
        };
        try {
            solve_symmetric_all(m, nrhs, static_cast<const cd*>(bv), ldb, static_cast<cd*>(xv), ldx, opt, r, be, rr, bs);
        } catch (const PivotFailure&) {
            fill(false);
            throw;
        }
        fill(true);
        if (backward_errors)
            for (int rhs = 0; rhs < nrhs; ++rhs)
                for (int sg = 0; sg < ns; ++sg)
                    backward_errors[static_cast<size_t>(sg) * nrhs + rhs] = be[static_cast<size_t>(rhs) * ns + sg];
    });
}

extern "C" int aphi_batch_backward_error(aphi_mesh* m, int nrhs, const void* x, int64_t ldx, const void* b, int64_t ldb,
                                         double* be, double* rr) {
    return aphi_capi::guarded([&] {
        require_batch_args(m, "aphi_batch_backward_error");
        require(x && b, "aphi_batch_backward_error: null argument");
        require(m->assembled, "aphi_batch_backward_error: nothing assembled");
        require(nrhs >= 1 && nrhs <= 64 && ldb >= m->n && ldx >= m->n,
                "aphi_batch_backward_error: nrhs must be 1..64 and ld >= n");
        after_caller(m->ctx);
        const int ns = m->seg.count;
        std::vector<double> e(static_cast<size_t>(nrhs) * ns), q(e.size());
        measure(m, nrhs, static_cast<const cd*>(x), ldx, static_cast<const cd*>(b), ldb, e.data(), q.data());
        for (int rhs = 0; rhs < nrhs; ++rhs)
            for (int sg = 0; sg < ns; ++sg) {
                const size_t c = static_cast<size_t>(rhs) * ns + sg, o = static_cast<size_t>(sg) * nrhs + rhs;
                if (be) be[o] = e[c];
                if (rr) rr[o] = q[c];
            }
    });
}

extern "C" int aphi_batch_terminal_current(aphi_mesh* m, int nrhs, const void* xv, int64_t ldx, const double* omegas,
                                           int conditioning, const double* axis, const double* lengths,
                                           const double* voltages, double* current, double* resistance,
                                           double* inductance) {
    return aphi_capi::guarded([&] {
        check_terminal_current(m, xv, nrhs, ldx, conditioning, axis, current, "aphi_batch_terminal_current");
        require(omegas && lengths && voltages, "aphi_batch_terminal_current: null argument");
        const int ns = m->seg.count;
        for (int sg = 0; sg < ns; ++sg)
            if (!(lengths[sg] > 0.0)) throw std::invalid_argument("terminal_current: length must be positive");
        after_caller(m->ctx);
        m->tcontrib.reserve(static_cast<size_t>(m->num_tets) * nrhs * sizeof(cd));
        m->tcurrent.reserve(static_cast<size_t>(nrhs) * ns * sizeof(cd) + static_cast<size_t>(2 * ns) * sizeof(cd) +
                            static_cast<size_t>(ns) * sizeof(double));
        cd* sums = m->tcurrent.as<cd>();
        cd* factors = sums + static_cast<size_t>(nrhs) * ns;
        double* dl = reinterpret_cast<double*>(factors + 2 * static_cast<size_t>(ns));
        std::vector<double> cur(2 * static_cast<size_t>(nrhs) * ns), res(static_cast<size_t>(nrhs) * ns),
            ind(static_cast<size_t>(nrhs) * ns);
        terminal_current_into(m, nrhs, xv, ldx, omegas, conditioning, axis, lengths, voltages, cur.data(), res.data(),
                              ind.data(), m->tcontrib.as<cd>(), sums, factors, dl);
        // per problem, then per right-hand side
        for (int rhs = 0; rhs < nrhs; ++rhs)
            for (int sg = 0; sg < ns; ++sg) {
                const size_t c = static_cast<size_t>(rhs) * ns + sg, o = static_cast<size_t>(sg) * nrhs + rhs;
                current[2 * o] = cur[2 * c];
                current[2 * o + 1] = cur[2 * c + 1];
                if (resistance) resistance[o] = res[c];
                if (inductance) inductance[o] = ind[c];
            }
    });
}

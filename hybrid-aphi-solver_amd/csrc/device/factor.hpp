// Host-callable launchers for the supernodal factorization and solve kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace aphi_dev {

constexpr int kNB = 64;       // block-column width of the blocked front factorization and solves
// rows per extend-add workgroup (64, 128, 256, 512 or 1024) and entries per lane
// and pass: a pass holds R / 64 rows of ea_entries / (R / 64) columns per lane
template <typename T> constexpr int ea_rows() { return 256; }
template <typename T> constexpr int ea_entries() { return sizeof(T) == 4 ? 16 : 8; }
constexpr int kEaChildren = 8;  // children an extend-add workgroup maps at once
constexpr int kOuter = 512;   // columns per outer block: the trailing update is rank kOuter
constexpr int kPanelTail = 64;  // zeroed entries after the factor storage (reads past a column's end)
constexpr int kWCols = 2 * kOuter;  // columns of a front's W window (complex float)
// complex float updates whose tallest region has at most kSmallTileRows rows, or
// whose widest region has at most kSmallTileCols columns (the block chain's own
// updates: four times the workgroups of the large tiles on a narrow region), use
// the 64 x 64 tiles (measured on the 32^3 cube: 591 ms with none, 586 ms by
// height at 8192, 577 ms adding width at 512, 623 ms with all)
constexpr int kSmallTileRows = 8192;
constexpr int kSmallTileCols = 512;
constexpr int kSB = 512;        // solve block: columns per forward/backward step (multiple of 128)
constexpr int kQ = 16;          // right-hand sides per interleaved solve group
constexpr int kGatherRows = 256;  // front rows per forward-gather workgroup (a multiple of 64)
constexpr int kRectRows = 1024;   // rows per backward partial sum below a front's pivots
constexpr int kFwdRectWgs = 4096; // workgroups a forward rectangle launch aims for (row tiles x column splits)
constexpr int kSmallK = 64;       // pivots per front of a level solved by the one-launch small-front kernels
constexpr int kSmallLds = 64 * 1024;  // LDS of a small-front forward workgroup: (max m + kSmallK) x kQ entries

void launch_stats_init(hipStream_t st, FactorStatsDev* stats);  // every slot to its identity
// Re-initialises the slots and gathers the pivot and multiplier extremes of the
// entries the reference's loop reaches before it stops at reference column
// `fail` (pivots of columns <= fail, multipliers of rows <= fail).
template <typename T>
void launch_failure_stats(hipStream_t st, int ns, int64_t entries, const FrontPlan& fp, const void* P, int fail,
                          FactorStatsDev* stats);
template <typename T> void launch_zero(hipStream_t st, int64_t n, void* p);
template <typename T> void launch_scatter_panel(hipStream_t st, int64_t nnz, const int64_t* s2p, const cd* a, void* P);
template <typename T> size_t small_lds_bytes(int max_m);
template <typename T>
void launch_front_small(hipStream_t st, const int* nodes, int count, int max_m, const FrontPlan& fp, void* P, void* CB,
                        void* xinv, double pivot_floor, FactorStatsDev* stats);
template <typename T>
void launch_extend_add(hipStream_t st, const int* nodes, const int* blk_start, int count, int total_blocks,
                       int max_m, const FrontPlan& fp, void* P, void* CB);
// Rank-(c1 - c0) update with the L columns [c0, c1 = min(c0 + width, k)) of each
// front, of the columns [max(c1, min(tlo, k)), thi < 0 ? m : min(thi, k)) and
// every row below them.
// max_wgs = 0: one workgroup per tile; > 0: a persistent grid of about that many
// workgroups (for an update that has the device to itself). small: the 64 x 64
// tiles of complex float (ignored for complex double); max_tiles counts tiles of
// update_tile<T>(small).
// Complex float reads the W = -L D operand from the level's W windows (W, see
// FrontPlan::woff); complex double forms it from the panel and D.
template <typename T>
void launch_update(hipStream_t st, const int* nodes, int count, int c0, int width, int tlo, int thi, int max_tiles,
                   int max_wgs, const FrontPlan& fp, void* P, void* CB, const void* W, bool small = false);
template <typename T> int update_tile(bool small = false);  // tile side of k_update
template <typename T> int update_resident_wgs();  // k_update workgroups the device holds at once
// Diagonal block and panel of kNB columns at b; complex float also writes the
// panel's W = -L D rows into the W windows (W non-null).
template <typename T>
void launch_block_step(hipStream_t st, const int* nodes, int count, int b, int max_row_blocks, int max_tiles,
                       const FrontPlan& fp, void* P, void* CB, void* mws, void* xinv, double pivot_floor,
                       FactorStatsDev* stats, void* W);

// solve.hip: vectors in groups of kQ right-hand sides, entry (i, q) of group g at
// [(g * rows + i) * kQ + q] (rows = n, or the total front row count for W)
int rhs_groups(int nrhs);
template <typename T>
void launch_perm_in(hipStream_t st, int n, int nrhs, const int* perm, const double* d, const cd* b, int64_t ldb,
                    const double* colscale, void* out);
template <typename T>
void launch_perm_out(hipStream_t st, int n, int nrhs, const int* perm, const double* d, const void* xt,
                     const double* colscale, cd* x, int64_t ldx, int accumulate);
template <typename T>
void launch_fwd_gather(hipStream_t st, const int* nodes, int count, int max_m, int nrhs, const FrontPlan& fp,
                       const void* b, int n, int64_t R, void* W);
// inverses of the kSB diagonal blocks (column-major X, row-major XT), from the
// kNB-block inverses; pairs = (supernode, first column of the block)
template <typename T>
void launch_inv_blocks(hipStream_t st, const int2* pairs, int npairs, const FrontPlan& fp, const void* P,
                       const void* xinv, void* X, void* XT);
// Step B of the block chain: y_B = X_B w_B, then the triangle rows [B + nb, k)
// take W -= L y_B; max_tri_rows: the most such rows among the level's supernodes
template <typename T>
void launch_fwd_step(hipStream_t st, const int* nodes, int count, int B, int max_nb, int max_tri_rows, int nrhs,
                     const FrontPlan& fp, const void* P, const void* Xsb, void* W, void* Y, int n, int64_t R);
// After a level's chain: the rows [k, m) below every front's pivots take
// W -= L[k:m, 0:k] y[0:k] (max_rect_rows: the most such rows, max_k the most
// pivots). The fronts' blocks are split fwd_rect_split ways when the row tiles
// alone would leave the device short of workgroups; the splits' partial sums go
// to part (count * split * max_rect_rows * kQ * rhs_groups(nrhs) entries) and are
// subtracted in split order.
int fwd_rect_split(int count, int max_k, int max_rect_rows, bool fp32, int* bps);
template <typename T>
void launch_fwd_rect(hipStream_t st, const int* nodes, int count, int max_k, int max_rect_rows, int nrhs,
                     const FrontPlan& fp, const void* P, void* W, const void* Y, int n, int64_t R, void* part);
// Before a level's backward chain: t = L[k:m, 0:k]^T x[rows k..m) for every front,
// in partial sums over kRectRows rows (rpart: count * rect_chunks(max_rect_rows) *
// rect_cols(max_k) * kQ * rhs_groups(nrhs) entries) reduced into t (count *
// rect_cols(max_k) * kQ * rhs_groups(nrhs), column stride rect_cols(max_k)).
// all: the level's chain has updates (a front with more than kSB pivots), so t is
// formed for every front, zero where there are no rows below the pivots. Returns
// the buffer holding t (rpart itself when one chunk suffices), null when there is
// nothing to add
int rect_chunks(int rect_rows);
int rect_cols(int max_k);
template <typename T>
const void* launch_bwd_rect(hipStream_t st, const int* nodes, int count, int max_k, int max_rect_rows, int nrhs,
                            bool all, const FrontPlan& fp, const void* P, const void* X, int n, void* rpart, void* t);
// Step B of the backward chain: x_B = X_B^T (D_B^-1 y_B - t[B:B+nb]), then
// t[0:B] += L[B:B+nb, 0:B]^T x_B (t from launch_bwd_rect, column stride rk)
template <typename T>
void launch_bwd_step(hipStream_t st, const int* nodes, int count, int B, int max_nb, int nrhs, const FrontPlan& fp,
                     const void* P, const void* dinv, const void* XTsb, const void* Y, void* X, int n, void* t, int rk,
                     bool all);
// Levels whose fronts have at most kSmallK pivots and whose front vectors fit in
// kSmallLds: the forward step in one launch, one workgroup per front (gather,
// y = X w, update vector)
template <typename T> bool small_level(int max_k, int max_m);
template <typename T>
void launch_fwd_small(hipStream_t st, const int* nodes, int count, int max_m, int nrhs, const FrontPlan& fp,
                      const void* P, const void* Xsb, const void* b, int n, int64_t R, void* W, void* Y);
// dinv[j] = 1 / D[j] for every column (n entries), after a factorization
template <typename T> void launch_dinv(hipStream_t st, int ns, const FrontPlan& fp, const void* P, void* dinv);
void launch_column_scale(hipStream_t st, int n, int nrhs, const cd* b, int64_t ldb, double* partials, double* colscale,
                         int mode);

}  // namespace aphi_dev

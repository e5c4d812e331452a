// Reference-side binding of the MI355X path (a translation unit the reference
// host adds, see aphi_gpu_backend.cpp). With it linked in place of the CPU
// definitions of aphi_solver::refill (symmetric overload, include/aphi_solver/
// assembly.hpp:197-199) and aphi_solver::solve_symmetric (include/aphi_solver/
// factorization.hpp:206-210), the reference's callers -- tools/solve_mesh.cpp,
// src/main.cpp, src/gauge_free_potential.cpp -- run on the GPU unchanged. This
// header adds only what the reference has no place for: the precision choice and
// a multi-right-hand-side solve.
#pragma once

#include <complex>
#include <vector>

#include "aphi_solver/factorization.hpp"

namespace aphi_solver {
namespace gpu {

struct Settings {
    int device = 0;
    // 0: complex128 factor (the reference's arithmetic); 1: complex64 factor on
    // MFMA + FP64 refinement to `tolerance` (falls back to complex128 when
    // refinement stalls)
    int precision = 0;
    int max_refinements = 0;  // 0 = the reference's behaviour: no refinement
    double tolerance = 1e-12;
};

// Settings for every later call (thread-safe).
void configure(const Settings& settings);
Settings settings();

// solve_symmetric for nrhs right-hand sides at once: B and X column-major, n x
// nrhs (column r of B at B[r * n]). The report's backward error and residual are
// the worst over the right-hand sides.
bool solve_symmetric_many(const SparseSymmetricZ& a, int nrhs, const std::vector<std::complex<double>>& B,
                          std::vector<std::complex<double>>& X, SolveReport& report, const SolveOptions& options = {});

// Releases every device mesh the binding holds (they are otherwise kept, one per
// pattern, for the next frequency).
void release();

}  // namespace gpu
}  // namespace aphi_solver

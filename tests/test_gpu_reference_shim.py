"""The reference-side binding as a drop-in: oracle/_ref/shim_driver is the
reference's own code (built from its sources by oracle/Makefile.ref) with its
refill and solve_symmetric taken from hybrid-aphi-solver_amd/shim (the GPU) and
the CPU definitions kept under other names for comparison. Through the unchanged
callers (the sweep step of tools/solve_mesh.cpp, recover_gauge_free_potential):
the refill is the CPU refill bit for bit (with the reference's ScatterMap and
without one), the solution within 1e-10 of the reference's complex128 one, every
SolveReport field filled as the reference fills it, the natural conditioning
refused with std::invalid_argument, the gauge-free potential's own solve on the
GPU, and four right-hand sides at once -- for a voltage-driven cube and one with
a current-driven port."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "oracle", "_ref", "shim_driver")


@pytest.mark.parametrize("name", ["config1", "cube4_current"])
def test_reference_callers_through_the_binding(name, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("oracle/_ref/shim_driver is not built (make -f oracle/Makefile.ref)")
    meta, arrays = golden(name)
    args = [DRIVER, str(tmp_path), str(meta["n"]), repr(meta["side"]), *map(str, meta["bar"]), repr(meta["sigma"]),
            repr(meta["freq"]), "1" if meta.get("mode") == 3 else "0"]
    print(" ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.load(open(tmp_path / "result.json"))
    print(json.dumps(res), flush=True)
    assert res["values_equal"] and res["ok_gpu"] and res["ok_cpu"] and res["refused_natural"]
    # the CPU side is the reference itself: its x is the golden one bit for bit
    xc = np.fromfile(tmp_path / "x_cpu.bin", dtype=np.complex128)
    assert hashlib.sha256(xc.tobytes()).hexdigest() == meta["sha256"]["x"]
    assert res["x_rel"] <= 1e-10
    g, c = res["report_gpu"], res["report_cpu"]
    assert g["ok"] and g["backward_error"] <= 1e-12
    assert g["equilibration_min"] == c["equilibration_min"] and g["equilibration_max"] == c["equilibration_max"]
    assert g["factor_nnz"] == c["factor_nnz"] and g["columns_done"] == c["columns_done"]
    for k in ("smallest_pivot", "largest_pivot", "largest_multiplier"):
        assert abs(g[k] / c[k] - 1) <= 1e-8, k
    assert g["factorize_ms"] > 0 and g["solve_ms"] > 0
    Ig, Ic = complex(*res["I_gpu"]), complex(*res["I_cpu"])
    assert abs(Ig - Ic) <= 1e-8 * abs(Ic)
    assert res["gf_converged"]
    if "gf_phi_vertex" in arrays:
        phi = np.fromfile(tmp_path / "gf_phi_vertex.bin", dtype=np.complex128)
        ref = arrays["gf_phi_vertex"]
        assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) <= 1e-10
    assert res["many_ok"] and res["many_max_rel"] <= 1e-10

"""Serving with continuous batching (aphi_amd.serving).

* Requests that overlap on the device give the same solutions, bit for bit, as
  the same requests served alone (same precision), also when a small HBM budget
  forces evictions.
* The serving workload's own sizes (meshes of 8, 16 and 24 hexes per edge and
  one of 32, meshes repeating with p = 0.7, 1 Hz-100 MHz, 1-8 right-hand sides),
  served with the precision chosen per mesh by frequency: every request reaches
  backward error <= 1e-12 on every right-hand side, recomputed on the CPU by the
  oracle (the reference's measure) and by a numpy restatement of the
  equilibrated measure; each mesh is uploaded once.
"""
import numpy as np
import pytest

import aphi_amd as A
from aphi_amd.serving import Request, Server, _plan_bytes, _workspace_bytes, serving_requests
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_stream():
    return serving_requests(11, 12, rate=200.0, sizes=(4, 6, 8), max_rhs=4)


def test_results_alone_equal_results_in_the_mix(gpu_context):
    reqs = small_stream()
    srv = Server(slots=3, hbm_budget=40e9, precision=A.APHI_FP32, keep_solutions=True)
    try:
        mix, summary = srv.run(reqs)
    finally:
        srv.close()
    assert summary["all_ok"], [r.get("error") for r in mix.values() if not r["ok"]]
    assert summary["max_backward_error"] <= 1e-12 and summary["max_scaled_backward_error"] <= 1e-12
    assert len({r["slot"] for r in mix.values()}) > 1          # the work was spread over slots
    assert summary["device_uploads"] == len({r.key() for r in reqs})  # one upload per distinct mesh
    alone_srv = Server(slots=1, hbm_budget=40e9, precision=A.APHI_FP32, keep_solutions=True)
    try:
        for r in reqs:
            solo, _ = alone_srv.run([Request(r.id, 0.0, r.config, r.frequency, r.nrhs, r.seed)])
            a, b = solo[r.id], mix[r.id]
            assert a["ok"] and b["ok"]
            assert np.array_equal(a["x"].view(np.uint64), b["x"].view(np.uint64)), r.id
    finally:
        alone_srv.close()


def test_budget_forces_evictions_without_changing_results(gpu_context):
    reqs = small_stream()
    big = Server(slots=2, hbm_budget=40e9, precision=A.APHI_FP32)
    try:
        ref, _ = big.run(reqs)
    finally:
        big.close()
    # room for the largest mesh and a little more: the meshes keep evicting each other
    need = max(_workspace_bytes(p, A.APHI_FP32) + _plan_bytes(p) for p in (A.Problem(r.config) for r in reqs))
    tight = Server(slots=2, hbm_budget=1.2 * need, precision=A.APHI_FP32)
    try:
        got, summary = tight.run(reqs)
    finally:
        tight.close()
    assert summary["all_ok"], [r.get("error") for r in got.values() if not r["ok"]]
    assert summary["evictions"] > 0
    for r in reqs:
        assert got[r.id]["digest"] == ref[r.id]["digest"], r.id


def test_release_workspace_keeps_the_plan(gpu_context):
    """After aphi_mesh_release_workspace the mesh holds only its plan, refuses what
    needs a factor, and its next refill and solve give the same bits as before."""
    import torch
    p = A.Problem(A.CubeConfig(n=6, bar=(2, 4, 2, 4)))
    m = A.Mesh(gpu_context, p)
    try:
        omega = 2 * np.pi * 1e4
        m.refill(omega)
        b = m.rhs_device()[None, :].contiguous()
        x1, _ = m.solve_symmetric(b, precision=A.APHI_FP32, max_refinements=10)
        full = m.device_bytes()
        m.release_workspace()
        plan = m.device_bytes()
        assert 0 < plan < full
        with pytest.raises(ValueError):
            m.solve(b, torch.empty_like(b))
        with pytest.raises(ValueError):
            m.backward_error(x1, b)
        m.refill(omega)
        x2, _ = m.solve_symmetric(b, precision=A.APHI_FP32, max_refinements=10)
        assert np.array_equal(x1.cpu().numpy().view(np.uint64), x2.cpu().numpy().view(np.uint64))
    finally:
        m.close()


def test_serving_workload_sizes_against_the_oracle(gpu_context):
    reqs = serving_requests(29, 14, rate=20.0, sizes=(8, 16, 24), max_rhs=8)
    last = reqs[-1]
    reqs.append(Request(len(reqs), last.arrival + 0.05, A.CubeConfig(n=32, side=0.1, bar=(6, 12, 20, 23), sigma=5.8e7),
                        3.3e4, 2, 777))
    srv = Server(slots=2, hbm_budget=200e9, keep_solutions=True)
    try:
        res, summary = srv.run(reqs)
        print({k: summary[k] for k in ("requests", "p50_ms", "p99_ms", "solves_per_s", "device_uploads",
                                       "fp64_requests", "fp32_attempts_refactored", "max_scaled_backward_error")},
              flush=True)
        assert summary["all_ok"], [r.get("error") for r in res.values() if not r["ok"]]
        assert summary["device_uploads"] == len({r.key() for r in reqs})
        for r in reqs:
            p = srv.problems[r.key()].result()
            omega = 2 * np.pi * r.frequency
            vals, _ = O.refill(p, omega)
            d = O.equilibrate(p, vals, 10)
            rp, ci = p.ints("row_ptr"), p.ints("col_index")
            rows = np.repeat(np.arange(p.n), np.diff(rp))
            amax_eq = np.max(np.abs(vals * d[rows] * d[ci]))
            B = A.random_normal_host(r.seed, 2 * r.nrhs * p.n).view(np.complex128).reshape(r.nrhs, p.n)
            X = res[r.id]["x"]
            for k in range(r.nrhs):
                be = O.backward_error(p, vals, X[k], B[k])[0]
                res_k = O.residual(p, vals, X[k], B[k])
                sbe = np.linalg.norm(d * res_k) / (amax_eq * np.linalg.norm(X[k] / d) + np.linalg.norm(d * B[k]))
                assert be <= 1e-12 and sbe <= 1e-12, (r.id, k, be, sbe)
            print(f"request {r.id}: n={r.config.n} f={r.frequency:.3g} Hz nrhs={r.nrhs} "
                  f"precision {res[r.id]['precision_requested']}->{res[r.id]['precision_used']} "
                  f"steps {res[r.id]['refinement_steps']} latency {res[r.id]['latency_ms']:.0f} ms", flush=True)
    finally:
        srv.close()

"""Serving with continuous batching (aphi_amd.serving): requests that overlap on
the device give the same solutions, bit for bit, as the same requests served
alone, also when a small HBM budget forces evictions; every request reaches
backward error 1e-12."""
import numpy as np
import pytest

import aphi_amd as A
from aphi_amd.serving import Request, Server, _estimate_bytes, serving_requests

pytestmark = pytest.mark.gpu


def small_stream():
    return serving_requests(11, 12, rate=200.0, sizes=(4, 6, 8), max_rhs=4)


def test_results_alone_equal_results_in_the_mix(gpu_context):
    reqs = small_stream()
    srv = Server(slots=3, hbm_budget=40e9, keep_solutions=True)
    try:
        mix, summary = srv.run(reqs)
    finally:
        srv.close()
    assert summary["all_ok"], [r.get("error") for r in mix.values() if not r["ok"]]
    assert summary["max_backward_error"] <= 1e-12
    assert len({r["slot"] for r in mix.values()}) > 1          # the work was spread over slots
    alone_srv = Server(slots=1, hbm_budget=40e9, keep_solutions=True)
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
    big = Server(slots=2, hbm_budget=40e9)
    try:
        ref, _ = big.run(reqs)
    finally:
        big.close()
    # room for the largest mesh and a little more: two slots keep evicting each other
    need = max(_estimate_bytes(A.Problem(r.config)) for r in reqs)
    tight = Server(slots=2, hbm_budget=1.2 * need)
    try:
        got, summary = tight.run(reqs)
    finally:
        tight.close()
    assert summary["all_ok"], [r.get("error") for r in got.values() if not r["ok"]]
    assert summary["evictions"] > 0
    for r in reqs:
        assert got[r.id]["digest"] == ref[r.id]["digest"], r.id

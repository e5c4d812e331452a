"""Gauge-free potential on the device: compute_fields' e_tet bit for bit with the
compiled reference, the P1 system's values and rhs bit for bit with the oracle,
the recovered vertex potentials within 1e-10 of the reference's (FP64; FP32 +
refinement to 1e-12 within the conditioning bound), and exact recovery of a P1
gradient field."""
import numpy as np
import pytest

from conftest import cube_config, golden
from test_gauge_free import p1_gradient_field, sha

import aphi_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NUMERIC = ["cube2", "cube3_offcentre", "config1", "cube8"]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


@pytest.fixture(scope="module", params=NUMERIC)
def gf_case(request, gpu_context):
    meta, arrays = golden(request.param)
    p = A.Problem(cube_config(meta))
    g = A.Problem.gauge_free(p)
    m = A.Mesh(gpu_context, p)
    mg = A.Mesh(gpu_context, g)
    mg.attach_gauge_free(g)
    yield dict(meta=meta, arrays=arrays, p=p, g=g, m=m, mg=mg)
    mg.close()
    m.close()


def test_tet_fields_bitwise_equal_reference(gf_case):
    c = gf_case
    e = c["m"].tet_fields(dev(c["arrays"]["x"]), c["meta"]["omega"])
    assert sha(e.cpu().numpy().reshape(-1)) == c["meta"]["sha256"]["e_tet"]


def test_gauge_free_refill_bitwise_equals_oracle(gf_case):
    c = gf_case
    omega = c["meta"]["omega"]
    e = O.tet_fields(c["p"], c["arrays"]["x"], omega)
    c["mg"].gauge_free_refill(dev(e), omega)
    vals, rhs = c["mg"].system()
    s = O.gauge_free_system(c["p"], e, omega)
    assert np.array_equal(vals.view(np.uint64), s["values"].view(np.uint64))
    assert np.array_equal(rhs.view(np.uint64), s["rhs"].view(np.uint64))


def test_recovered_potential_matches_reference(gf_case):
    c = gf_case
    r = A.recover_gauge_free_potential(c["m"], c["mg"], dev(c["arrays"]["x"]), c["meta"]["omega"])
    assert r["converged"] and r["unknowns"] == c["meta"]["gf_unknowns"]
    phi = r["phi_vertex"].cpu().numpy()
    ref = c["arrays"]["gf_phi_vertex"]
    assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) <= 1e-10
    assert np.array_equal(r["port_potential"], c["arrays"]["gf_port_potential"])


def test_recovered_potential_fp32_refined(gf_case):
    """FP32 factor + FP64 refinement to backward error 1e-12; distance to the
    reference within 4 kappa_2(D A D) (1e-12 + 1e-15), kappa from a dense SVD."""
    c = gf_case
    omega, tol = c["meta"]["omega"], 1e-12
    r = A.recover_gauge_free_potential(c["m"], c["mg"], dev(c["arrays"]["x"]), omega, precision=A.APHI_FP32,
                                       max_refinements=40, tolerance=tol)
    rep = r["report"]
    assert r["converged"] and rep["backward_error"] <= tol and rep["scaled_backward_error"] <= tol
    vals, _ = c["mg"].system()
    d = O.equilibrate(c["g"], vals, 10)
    rp, ci = c["g"].ints("row_ptr"), c["g"].ints("col_index")
    n = c["g"].n
    rows = np.repeat(np.arange(n), np.diff(rp))
    dense = np.zeros((n, n), dtype=np.complex128)
    dense[rows, ci] = vals * d[rows] * d[ci]
    dense[ci, rows] = vals * d[rows] * d[ci]
    sv = np.linalg.svd(dense, compute_uv=False)
    bound = 4.0 * (sv[0] / sv[-1]) * (tol + 1e-15)
    phi = r["phi_vertex"].cpu().numpy()
    ref = c["arrays"]["gf_phi_vertex"]
    assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) <= bound


def test_gradient_field_recovered_exactly_on_device(gf_case):
    c = gf_case
    omega = 2 * np.pi * 1e4
    nv = int(c["p"].counts["nodes"])
    col = c["g"].ints("gf_column")
    pres = c["g"].doubles("gf_prescribed").view(np.complex128)
    rng = np.random.default_rng(11)
    phi = rng.normal(size=nv) + 1j * rng.normal(size=nv)
    phi[col < 0] = pres[col < 0]
    e = p1_gradient_field(c["p"], phi)
    c["mg"].gauge_free_refill(dev(e), omega)
    _, rhs = c["mg"].system()
    x, rep = c["mg"].solve_symmetric(dev(rhs), precision=A.APHI_FP64)
    assert rep["ok"]
    got, _ = c["mg"].gauge_free_potential(x)
    assert np.abs(got.cpu().numpy() - phi).max() / np.abs(phi).max() < 1e-10


def test_gauge_free_refusals(gf_case):
    c = gf_case
    e = dev(np.zeros((int(c["p"].counts["tets"]), 3)))
    with pytest.raises(ValueError, match="omega must be nonzero"):
        c["mg"].gauge_free_refill(e, 0.0)
    with pytest.raises(ValueError, match="no gauge-free plan"):
        c["m"].gauge_free_refill(e, 1.0)

"""Gauge-free potential (src/gauge_free_potential.cpp:40-177), host side: the
oracle's restatement against the compiled reference's output (e_tet and the
recovered vertex potentials of the golden fixtures, bit for bit), against what
the mathematics fixes (a field that is the gradient of a P1 potential is
recovered exactly; a constant potential has no source), and the host plan's
vertex map and pattern against the oracle's."""
import numpy as np
import pytest

from conftest import cube_config, golden

import aphi_amd as A
from oracle import oracle as O

NUMERIC = ["cube2", "cube3_offcentre", "config1", "cube8"]


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module", params=NUMERIC)
def case(request):
    meta, arrays = golden(request.param)
    p = A.Problem(cube_config(meta))
    return dict(meta=meta, arrays=arrays, p=p)


def test_oracle_tet_fields_bitwise_equal_reference(case):
    e = O.tet_fields(case["p"], case["arrays"]["x"], case["meta"]["omega"])
    assert sha(e.reshape(-1)) == case["meta"]["sha256"]["e_tet"]


def test_oracle_recovers_reference_potential_bitwise(case):
    """The oracle's vertex map and system from the reference's own e_tet, solved by
    the oracle's restatement of solve_symmetric with the reference's AMD ordering
    of that pattern, give the reference's GaugeFreePotential bit for bit."""
    meta, p = case["meta"], case["p"]
    e = O.tet_fields(p, case["arrays"]["x"], meta["omega"])
    s = O.gauge_free_system(p, e, meta["omega"])
    assert s["unknowns"] == meta["gf_unknowns"]
    q = A.Problem.from_csr(s["unknowns"], s["row_ptr"], s["col"])
    x, _, st = O.csr_solve_symmetric_raw(s["unknowns"], s["row_ptr"], s["col"], s["values"], s["rhs"],
                                         q.ints("perm"), iterations=10)
    assert st["ok"]
    phi = np.where(s["column"] >= 0, x[np.maximum(s["column"], 0)], s["prescribed"])
    ref = case["arrays"]["gf_phi_vertex"]
    assert np.array_equal(phi.view(np.uint64), ref.view(np.uint64))
    ports = case["arrays"]["gf_port_potential"]
    pv = p.doubles("port_value").view(np.complex128)
    assert np.array_equal(ports, pv)  # voltage ports come back as prescribed


def p1_gradient_field(p, phi_v):
    """E = -grad of the P1 interpolant of the vertex values phi_v, per tet."""
    coords = p.doubles("coords").reshape(-1, 3)
    tets = p.ints("tet_nodes").reshape(-1, 4)
    M = np.ones((tets.shape[0], 4, 4))
    M[:, :, 1:] = coords[tets]
    G = np.linalg.inv(M)[:, 1:, :]          # grad L_i = column i of rows 1..3
    return -np.einsum("tci,ti->tc", G, phi_v[tets])


@pytest.mark.parametrize("name", ["config1", "cube8"])
def test_gradient_field_is_recovered_exactly(name):
    """Galerkin exactness: if E = -grad(phi) for phi in the P1 space that takes the
    prescribed values on the eliminated vertices, the system's solution is phi.
    A dropped term, a sign or a transposed index in the assembly breaks it."""
    meta, _ = golden(name)
    p = A.Problem(cube_config(meta))
    omega = 2 * np.pi * 1e4
    s0 = O.gauge_free_system(p, np.zeros((int(p.counts["tets"]), 3), dtype=np.complex128), omega)
    rng = np.random.default_rng(3)
    nv = int(p.counts["nodes"])
    phi = rng.normal(size=nv) + 1j * rng.normal(size=nv)
    fixed = s0["column"] < 0
    phi[fixed] = s0["prescribed"][fixed]
    s = O.gauge_free_system(p, p1_gradient_field(p, phi), omega)
    q = A.Problem.from_csr(s["unknowns"], s["row_ptr"], s["col"])
    x, _, st = O.csr_solve_symmetric_raw(s["unknowns"], s["row_ptr"], s["col"], s["values"], s["rhs"],
                                         q.ints("perm"), iterations=10)
    assert st["ok"]
    free = ~fixed
    assert np.abs(x[s["column"][free]] - phi[free]).max() / np.abs(phi).max() < 1e-10


def test_constant_potential_has_no_source():
    """Rows of the stiffness sum to zero over all four vertices of every tet
    (sum_j grad L_j = 0): with E = 0 and every eliminated vertex at the same value
    c, the potential is c everywhere."""
    meta, _ = golden("config1")
    cfg = cube_config(meta)
    cfg.v_top = cfg.v_bottom = 0.75 - 0.5j
    p = A.Problem(cfg)
    s = O.gauge_free_system(p, np.zeros((int(p.counts["tets"]), 3), dtype=np.complex128), 2 * np.pi * 50.0)
    q = A.Problem.from_csr(s["unknowns"], s["row_ptr"], s["col"])
    x, _, st = O.csr_solve_symmetric_raw(s["unknowns"], s["row_ptr"], s["col"], s["values"], s["rhs"],
                                         q.ints("perm"), iterations=10)
    assert st["ok"]
    assert np.abs(x - (0.75 - 0.5j)).max() < 1e-12


@pytest.mark.parametrize("name", NUMERIC)
def test_host_plan_matches_oracle_pattern(name):
    meta, _ = golden(name)
    p = A.Problem(cube_config(meta))
    g = A.Problem.gauge_free(p)
    s = O.gauge_free_system(p, np.zeros((int(p.counts["tets"]), 3), dtype=np.complex128), 1.0)
    assert g.n == s["unknowns"] == meta["gf_unknowns"]
    assert np.array_equal(g.ints("gf_column"), s["column"])
    assert np.array_equal(g.ints("row_ptr"), s["row_ptr"])
    assert np.array_equal(g.ints("col_index"), s["col"])
    assert np.array_equal(g.doubles("gf_prescribed").view(np.complex128), s["prescribed"])
    # every slot has at least one stiffness term, every rhs row its source terms
    tp = g.ints("gf_target_ptr")
    assert tp.size == g.counts["nnz"] + g.n + 1 and np.all(np.diff(tp)[: g.counts["nnz"]] > 0)
    kind = g.ints("gf_kind")
    assert np.all(kind[: tp[g.counts["nnz"]]] == 0)


def test_gauge_free_refuses_a_bare_pattern():
    n, rp, ci = 3, [0, 2, 3, 4], [0, 1, 1, 2]
    q = A.Problem.from_csr(n, rp, ci)
    with pytest.raises(ValueError, match="no mesh"):
        A.Problem.gauge_free(q)

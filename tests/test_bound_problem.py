"""aphi_problem_create_bound: a problem from the reference host's own bound
objects (Mesh nodes/tets/edge signs, BoundProblem bodies and ports, every tet's
DofMap::local_dofs, the upper SparsityPattern and optionally the ScatterMap),
flattened. The arrays a cube problem exports must rebuild exactly the cube's
device plan -- assembly contributions in the reference's order, pattern,
ordering, supernodes -- with or without the ScatterMap handed over, and the
oracle's refill from them must be the compiled reference's bit for bit
(including a current-driven port); malformed input is refused."""
import hashlib

import numpy as np
import pytest

from conftest import cube_config, golden

import aphi_amd as A
from oracle import oracle as O

CASES = ["cube2", "config1", "cube5_draw", "cube4_current"]

FIELDS = {  # aphi_mesh_desc pointer fields and their lengths
    "coords": lambda d: 3 * d.num_nodes, "tet_nodes": lambda d: 4 * d.num_tets, "tet_body": lambda d: d.num_tets,
    "body_sigma": lambda d: d.num_bodies, "body_eps_r": lambda d: d.num_bodies, "body_mu_r": lambda d: d.num_bodies,
    "target_ptr": lambda d: d.nnz + d.n + 1, "values": lambda d: 2 * d.num_values,
    "row_ptr": lambda d: d.n + 1, "col_index": lambda d: d.nnz, "factor_perm": lambda d: d.n,
    "col_ref": lambda d: d.n, "super_first": lambda d: d.num_super + 1, "super_parent": lambda d: d.num_super,
    "super_row_ptr": lambda d: d.num_super + 1, "super_level": lambda d: d.num_super,
    "panel_offset": lambda d: d.num_super + 1, "cb_offset": lambda d: d.num_super,
    "slot_to_panel": lambda d: d.nnz, "tet_dofs": lambda d: 16 * d.num_tets,
    "tet_edge_signs": lambda d: 6 * d.num_tets, "tet_phi_coeff": lambda d: 10 * d.num_tets,
    "tet_phi_fixed": lambda d: 20 * d.num_tets,
}


def desc_arrays(p):
    d = p.describe()
    out = {k: np.ctypeslib.as_array(getattr(d, k), shape=(int(f(d)),)).copy() for k, f in FIELDS.items()}
    nc = int(out["target_ptr"][-1])
    out["entry"] = np.ctypeslib.as_array(d.entry, shape=(nc,)).copy()
    out["kind"] = np.ctypeslib.as_array(d.kind, shape=(nc,)).copy()
    rows = int(out["super_row_ptr"][-1])
    out["super_rows"] = np.ctypeslib.as_array(d.super_rows, shape=(rows,)).copy()
    out["super_rel"] = np.ctypeslib.as_array(d.super_rel, shape=(rows,)).copy()
    for k in ("n", "nnz", "num_values", "num_super", "num_levels", "cb_pool_entries", "flops", "predicted_nnz"):
        out[k] = getattr(d, k)
    return out


@pytest.fixture(scope="module", params=CASES)
def case(request):
    meta, _ = golden(request.param)
    return meta, A.Problem(cube_config(meta))


@pytest.mark.parametrize("with_map", [True, False])
def test_bound_arrays_rebuild_the_cube_plan(case, with_map):
    meta, p = case
    arrays = p.bound_arrays()
    if not with_map:
        arrays = {k: v for k, v in arrays.items() if not k.startswith("scatter_")}
    q = A.Problem.from_bound(arrays)
    a, b = desc_arrays(p), desc_arrays(q)
    assert set(a) == set(b)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    for k in ("scatter_live", "scatter_pos", "scatter_offset", "scatter_slot", "perm", "parent", "col_count"):
        assert np.array_equal(p.ints(k), q.ints(k)), k


def test_oracle_refill_from_bound_arrays_is_the_reference(case):
    """The oracle reads the bound problem's exports: its refill is the compiled
    reference's (SHA-256 of the golden values and rhs)."""
    meta, p = case
    q = A.Problem.from_bound(p.bound_arrays())
    vals, rhs = O.refill(q, meta["omega"])
    assert hashlib.sha256(vals.tobytes()).hexdigest() == meta["sha256"]["values"]
    assert hashlib.sha256(rhs.tobytes()).hexdigest() == meta["sha256"]["rhs"]


def test_current_port_rhs_carries_the_current():
    """The current-driven port's row: rhs = row * I (src/assembly.cpp:396-403),
    no unit diagonal; the voltage port keeps its constraint row."""
    meta, _ = golden("cube4_current")
    p = A.Problem(cube_config(meta))
    fixed = p.ints("port_is_fixed")
    assert list(fixed) == [0, 1]
    rows = p.ints("port_index")
    vals, rhs = O.refill(p, meta["omega"])
    jw = 1j * meta["omega"]
    assert rhs[rows[0]] == (1.0 / jw) * 1.0
    ci = p.ints("col_index")
    # the port unknown couples to the terminal's tets: upper-stored, in the
    # columns of the rows before it (the A and Phi rows)
    assert np.count_nonzero(ci == rows[0]) > 1 and np.count_nonzero(ci == rows[1]) == 1


def mutate(arrays, **kw):
    out = {k: np.array(v, copy=True) for k, v in arrays.items()}
    for k, f in kw.items():
        f(out[k])
    return out


def test_bound_refusals():
    p = A.Problem(A.CubeConfig(n=2, bar=(0, 1, 0, 1)))
    base = p.bound_arrays()
    n = p.n

    def bad(**kw):
        with pytest.raises(ValueError):
            A.Problem.from_bound(mutate(base, **kw))

    bad(tet_nodes=lambda a: a.__setitem__(0, 10 ** 6))                      # a node that does not exist
    bad(edge_index=lambda a: a.reshape(-1).__setitem__(0, n))               # an unknown out of range
    bad(body_tets=lambda a: a.__setitem__(1, a[0]))                         # a tet in a body twice
    bad(col_index=lambda a: a.__setitem__(len(a) - 1, 0))                   # below the diagonal
    bad(scatter_slot=lambda a: a.__setitem__(0, 10 ** 7))                   # a slot outside the pattern
    bad(scatter_offset=lambda a: a.__setitem__(1, a[1] + 1))               # blocks that do not match
    bad(port_index=lambda a: a.__setitem__(0, -2))                          # a port row out of range
    part = {k: v for k, v in base.items() if k != "scatter_slot"}
    with pytest.raises(ValueError):
        A.Problem.from_bound(part)                                          # three of the four map arrays
    q = A.Problem.from_bound(base)
    with pytest.raises(ValueError):
        A.Problem.gauge_free(q)                                             # no mesh topology

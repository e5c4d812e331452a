"""Result files: potential_at_nodes + write_potential and write_vtk (without and
with compute_fields' data, no run information) written from the reference's own
solution give the compiled reference's files byte for byte (SHA-256 of the
golden fixtures). The field data here come from the oracle's compute_fields;
the device's are checked against the same bytes in test_gpu_parity.py."""
import hashlib

import numpy as np
import pytest

from conftest import cube_config, golden

import aphi_amd as A
from oracle import oracle as O

NUMERIC = ["cube2", "cube3_offcentre", "config1", "cube8", "cube5_draw", "cube4_current"]


@pytest.mark.parametrize("name", NUMERIC)
@pytest.mark.parametrize("vtk,fname", [(False, "potential.out"), (True, "solution.vtk")])
def test_files_equal_the_reference_bytes(tmp_path, name, vtk, fname):
    meta, arrays = golden(name)
    p = A.Problem(cube_config(meta))
    out = tmp_path / fname
    nbytes = p.write_potential(out, arrays["x"], meta["omega"], vtk=vtk)
    data = out.read_bytes()
    assert nbytes == len(data) == meta["bytes_" + fname]
    assert hashlib.sha256(data).hexdigest() == meta["sha256"][fname]


@pytest.mark.parametrize("name", NUMERIC)
def test_vtk_with_fields_equals_the_reference_bytes(tmp_path, name):
    meta, arrays = golden(name)
    p = A.Problem(cube_config(meta))
    f = O.compute_fields(p, arrays["x"], meta["omega"])
    out = tmp_path / "fields.vtk"
    nbytes = p.write_vtk_fields(out, arrays["x"], meta["omega"], f)
    data = out.read_bytes()
    assert nbytes == len(data) == meta["bytes_fields.vtk"]
    assert hashlib.sha256(data).hexdigest() == meta["sha256"]["fields.vtk"]
    # the oracle's e_tet is the one pinned bit for bit by the golden e_tet digest
    assert np.array_equal(f["e_tet"].view(np.uint64), O.tet_fields(p, arrays["x"], meta["omega"]).view(np.uint64))


def test_field_plan():
    """Each P2 node lists its (tet, local node) entries in tet order, every entry
    once; a tet's ten nodes are its vertices and its edges' midpoints."""
    p = A.Problem(A.CubeConfig(n=3, bar=(1, 2, 0, 2)))
    d = p.describe_fields()
    nt, npn = d.num_tets, d.num_p2
    assert npn == p.counts["nodes"] + p.counts["edges"]
    tp2 = np.ctypeslib.as_array(d.tet_p2, shape=(10 * nt,)).reshape(nt, 10)
    ptr = np.ctypeslib.as_array(d.node_ptr, shape=(npn + 1,))
    ent = np.ctypeslib.as_array(d.node_entry, shape=(10 * nt,))
    assert np.array_equal(tp2[:, :4], p.ints("tet_nodes").reshape(nt, 4))
    assert np.array_equal(tp2[:, 4:], p.counts["nodes"] + p.ints("tet_edges").reshape(nt, 6))
    assert ptr[0] == 0 and ptr[-1] == 10 * nt and np.all(np.diff(ptr) >= 0)
    assert np.array_equal(np.sort(ent), np.arange(10 * nt))
    for u in range(npn):
        e = ent[ptr[u]:ptr[u + 1]]
        assert np.all(tp2.reshape(-1)[e] == u) and np.all(np.diff(e // 10) > 0)


def test_writer_refusals(tmp_path):
    meta, arrays = golden("cube2")
    p = A.Problem(cube_config(meta))
    with pytest.raises(ValueError, match="n values"):
        p.write_potential(tmp_path / "x.out", arrays["x"][:-1], meta["omega"])
    q = A.Problem.from_csr(3, [0, 2, 3, 4], [0, 1, 1, 2])
    with pytest.raises(ValueError, match="no mesh"):
        q.write_potential(tmp_path / "y.out", np.zeros(3, dtype=complex), 1.0)
    with pytest.raises(RuntimeError, match="could not open"):
        p.write_potential(tmp_path / "missing" / "z.out", arrays["x"], meta["omega"])

"""Result files: potential_at_nodes + write_potential and write_vtk (no field
data, no run information) written from the reference's own solution give the
compiled reference's files byte for byte (SHA-256 of the golden fixtures)."""
import hashlib

import numpy as np
import pytest

from conftest import cube_config, golden

import aphi_amd as A

NUMERIC = ["cube2", "cube3_offcentre", "config1", "cube8"]


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

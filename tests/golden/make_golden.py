"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_driver (the reference's own sources, built by
oracle/Makefile.ref) on the cube configurations below and stores what it
produced: small arrays in full, large ones as SHA-256 digests of their int64 /
float64 little-endian bytes. Run in a container that has the reference:

    make -f oracle/Makefile.ref REF=/root/reference -j8
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

INT_ARRAYS = ["edges", "tet_edges", "tet_edge_signs", "tree_edge", "dirichlet_edge", "edge_index", "phi_index",
              "row_ptr", "col_index", "perm", "parent", "row_count", "col_count", "scatter_live", "scatter_pos",
              "scatter_offset", "scatter_slot"]
FULL_INTS = {"perm", "parent", "col_count", "tree_edge", "edge_index"}

# name, n, side, bar (i0, i1, j0, j1), sigma, frequency, mode (0 symbolic, 1 numeric, 2 one solve with I, R, L),
# keep everything in full
CASES = [
    ("cube2", 2, 0.1, (0, 1, 0, 1), 5.8e7, 1e3, 1, True),
    ("cube3_offcentre", 3, 0.05, (1, 2, 0, 2), 1e6, 1e6, 1, True),
    ("config1", 6, 0.1, (2, 4, 2, 4), 5.8e7, 50.0, 1, False),
    ("cube8", 8, 0.2, (3, 5, 3, 5), 5.8e7, 1e5, 1, False),
    # a draw of the batched workload: side, sigma and frequency with no short
    # binary expansion (the scatter's 1/(j w) and its products round differently
    # under other complex-arithmetic routines)
    ("cube5_draw", 5, 0.10592777681478216, (2, 4, 1, 5), 27506162.131823607, 540.9701531845078, 1, False),
    # the z = side port driven by a 1 A boundary current (mode 3)
    ("cube4_current", 4, 0.08, (1, 3, 1, 2), 3.0e7, 2.5e3, 3, True),
    ("config2", 24, 0.1, (10, 14, 10, 14), 5.8e7, 10.0, 0, False),
    ("config3", 32, 0.1, (6, 12, 20, 23), 5.8e7, 1.0, 0, False),
    # the 24^3 sweep's two end points (10 Hz and 10 MHz are exact in the log sweep)
    ("config2_10hz", 24, 0.1, (10, 14, 10, 14), 5.8e7, 10.0, 2, False),
    ("config2_10mhz", 24, 0.1, (10, 14, 10, 14), 5.8e7, 1e7, 2, False),
]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(case, outdir):
    name, n, side, bar, sigma, freq, mode, full = case
    cmd = [DRIVER, outdir, str(n), repr(side), *map(str, bar), repr(sigma), repr(freq), str(mode)]
    subprocess.run(cmd, check=True)
    meta = json.load(open(os.path.join(outdir, "meta.json")))
    meta.update(case=name, side=side, bar=list(bar), sigma=sigma, mode=mode)
    arrays, digests = {}, {}
    for a in INT_ARRAYS:
        v = np.fromfile(os.path.join(outdir, a + ".bin"), dtype=np.int64)
        digests[a] = digest(v)
        if full or (a in FULL_INTS and mode in (1, 3)):
            arrays[a] = v.astype(np.int32)
    if mode == 2:
        for a, dt in (("values", np.complex128), ("rhs", np.complex128), ("x", np.complex128)):
            v = np.fromfile(os.path.join(outdir, a + ".bin"), dtype=dt)
            digests[a] = digest(v)
            if a == "x":
                arrays[a] = v
    if mode in (1, 3):
        for a, dt in (("values", np.complex128), ("rhs", np.complex128), ("equil_d", np.float64),
                      ("diag", np.complex128), ("x", np.complex128), ("e_tet", np.complex128),
                      ("gf_phi_vertex", np.complex128), ("gf_port_potential", np.complex128)):
            v = np.fromfile(os.path.join(outdir, a + ".bin"), dtype=dt)
            digests[a] = digest(v)
            if full or a in ("x", "diag", "equil_d", "rhs", "gf_phi_vertex", "gf_port_potential"):
                arrays[a] = v
        for fname in ("potential.out", "solution.vtk", "fields.vtk"):
            data = open(os.path.join(outdir, fname), "rb").read()
            digests[fname] = hashlib.sha256(data).hexdigest()
            meta["bytes_" + fname] = len(data)
    meta["sha256"] = digests
    return meta, arrays


def main(names):
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -f oracle/Makefile.ref")
    for case in CASES:
        if names and case[0] not in names:
            continue
        with tempfile.TemporaryDirectory() as tmp:
            meta, arrays = run(case, tmp)
        with open(os.path.join(HERE, case[0] + ".json"), "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        if arrays:
            np.savez_compressed(os.path.join(HERE, case[0] + ".npz"), **arrays)
        print("wrote", case[0])


if __name__ == "__main__":
    main(sys.argv[1:])

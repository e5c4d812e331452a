"""Writes tests/golden/conditioning.json: the 2-norm condition numbers of the
golden numeric cases (dense SVD of the full symmetric matrix, and of the
equilibrated D A D with the reference's equilibration factors).

They bound the forward error a backward-stable solution may carry:
||x - x*|| / ||x*|| <= kappa * backward error (to first order), which is what
the solution-agreement tests allow. Needs only the committed fixtures and the
oracle:

    python tests/golden/make_conditioning.py [--all]

Cases already in the file are kept unless --all is given.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hybrid-aphi-solver_amd"))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

CASES = ["cube2", "cube3_offcentre", "config1", "cube8", "cube5_draw", "cube4_current"]


def main():
    import aphi_amd as A
    from conftest import cube_config, golden
    from oracle import oracle as O

    path = os.path.join(HERE, "conditioning.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for name in CASES:
        if name in out and "--all" not in sys.argv:
            continue  # kept from an earlier run (the dense SVDs take minutes)
        meta, g = golden(name)
        p = A.Problem(cube_config(meta))
        vals, _ = O.refill(p, meta["omega"])
        rp, ci = p.ints("row_ptr"), p.ints("col_index")
        n = p.n
        M = np.zeros((n, n), complex)
        for i in range(n):
            for k in range(rp[i], rp[i + 1]):
                M[i, ci[k]] = vals[k]
                M[ci[k], i] = vals[k]
        d = np.asarray(g["equil_d"])
        s = np.linalg.svd(M, compute_uv=False)
        se = np.linalg.svd(d[:, None] * M * d[None, :], compute_uv=False)
        out[name] = {"n": n, "cond2": float(s[0] / s[-1]), "cond2_equilibrated": float(se[0] / se[-1])}
        print(name, out[name], flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

"""The solve's level structure on cubes whose levels exercise every path: levels
of many small fronts (one-launch forward steps), block chains of 2 to 11 blocks of
512 pivots, rows below the pivots in one launch per level with the forward sums
split over column blocks, backward partial sums over several 1024-row chunks, and
more than one group of 16 right-hand sides.

* Complex-double factor: every right-hand side's backward error on the
  equilibrated system, ||D r|| / (max|D A D| ||D^-1 x|| + ||D b||) with r = b - A x
  from the oracle's residual and D the oracle's equilibration, is at the level of
  roundoff (<= 1e-14).
* Complex-float factor plus refinement: every right-hand side reaches 1e-12 on the
  same measure and on the reference's backward error, both recomputed on the CPU.
* The solution of a right-hand side does not depend on the group it sits in or on
  the other right-hand sides solved with it (bit for bit).
"""
import numpy as np
import pytest

import aphi_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[12, 16])
def cube(request, gpu_context):
    p = A.Problem(A.CubeConfig(n=request.param))
    m = A.Mesh(gpu_context, p)
    omega = 2 * np.pi * 2.0e5
    m.refill(omega)
    vals, rhs = m.system()
    yield p, m, vals, rhs
    m.close()


def scaled_backward_errors(p, vals, X, B):
    d = O.equilibrate(p, vals, 10)
    rp, ci = p.ints("row_ptr"), p.ints("col_index")
    rows = np.repeat(np.arange(p.n), np.diff(rp))
    amax_eq = np.max(np.abs(vals * d[rows] * d[ci]))
    out = []
    for x, b in zip(X, B):
        r = O.residual(p, vals, x, b)
        out.append(np.linalg.norm(d * r) / (amax_eq * np.linalg.norm(x / d) + np.linalg.norm(d * b)))
    return np.array(out)


def rhs_block(p, rhs, nrhs, seed):
    g = np.random.default_rng(seed)
    B = g.standard_normal((nrhs, p.n)) + 1j * g.standard_normal((nrhs, p.n))
    B[0] = rhs
    return B


def test_fp64_solve_backward_errors_two_groups(cube):
    import torch
    p, m, vals, rhs = cube
    B = rhs_block(p, rhs, 20, 3)
    x, rep = m.solve_symmetric(torch.from_numpy(B).cuda(), precision=A.APHI_FP64)
    assert rep["ok"]
    sbe = scaled_backward_errors(p, vals, x.cpu().numpy(), B)
    print(f"n={p.n}: FP64 scaled backward errors max {sbe.max():.2e}", flush=True)
    assert np.all(sbe <= 1e-14), sbe


def test_fp32_refined_backward_errors_two_groups(cube):
    import torch
    p, m, vals, rhs = cube
    B = rhs_block(p, rhs, 20, 5)
    x, rep = m.solve_symmetric(torch.from_numpy(B).cuda(), precision=A.APHI_FP32, max_refinements=40,
                               tolerance=1e-12)
    assert rep["precision_used"] == A.APHI_FP32
    X = x.cpu().numpy()
    sbe = scaled_backward_errors(p, vals, X, B)
    print(f"n={p.n}: FP32 + refinement scaled backward errors max {sbe.max():.2e}", flush=True)
    assert np.all(sbe <= 1e-12), sbe
    for r in range(B.shape[0]):
        assert O.backward_error(p, vals, X[r], B[r])[0] <= 1e-12, r


def test_solution_independent_of_group_and_companions(cube):
    import torch
    p, m, vals, rhs = cube
    m.equilibrate(10)
    assert m.factorize_ldlt(A.APHI_FP32)["ok"]
    B = rhs_block(p, rhs, 20, 9)
    X20 = m.solve(torch.from_numpy(B).cuda()).cpu().numpy()
    # the last right-hand side alone, and moved to the first group's first slot
    X1 = m.solve(torch.from_numpy(B[19:20].copy()).cuda()).cpu().numpy()
    assert np.array_equal(X20[19].view(np.uint64), X1[0].view(np.uint64))
    Bs = B[[19] + list(range(19))]
    Xs = m.solve(torch.from_numpy(Bs).cuda()).cpu().numpy()
    assert np.array_equal(Xs[0].view(np.uint64), X20[19].view(np.uint64))
    assert np.array_equal(Xs[1:].view(np.uint64), X20[:19].view(np.uint64))

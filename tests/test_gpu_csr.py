"""The device path on problems made from a bare CSR pattern (any complex-symmetric
matrix handed to solve_symmetric, e.g. by the gauge-free potential,
src/gauge_free_potential.cpp:177): the cube's own pattern and values give the
cube mesh's solution bit for bit, and matrices no mesh produced are solved to
the oracle's solution (FP64: relative error <= 1e-10) and to backward error
<= 1e-12 (FP32 factor + FP64 refinement)."""
import numpy as np
import pytest

from conftest import cube_config, golden
from test_csr_problem import grid_pattern

import aphi_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


def grid_matrix(k, seed):
    """Complex-symmetric 7-point matrix on a k^3 grid: diagonally dominant, then
    scaled symmetrically by factors spread over four decades (the equilibration's
    work)."""
    n, rp, ci = grid_pattern(k)
    return n, rp, ci, dominant_values(n, rp, ci, seed)


def star_pattern(hub, arms, arm, reach, seed):
    """Upper triangle of a clique of `hub` vertices with `arms` cliques of `arm`
    vertices around it, each arm vertex joined to about half of `reach` hub
    vertices drawn per arm: the hub's front has one child per arm."""
    rng = np.random.default_rng(seed)
    n = hub + arms * arm
    adj = [set(range(i + 1, hub)) if i < hub else set() for i in range(n)]
    for a in range(arms):
        base = hub + a * arm
        targets = rng.choice(hub, size=reach, replace=False)
        for i in range(arm):
            adj[base + i].update(range(base + i + 1, base + arm))
            for t in targets[rng.random(reach) < 0.5]:
                adj[min(base + i, int(t))].add(max(base + i, int(t)))
    rp, ci = [0], []
    for i in range(n):
        ci += [i] + sorted(adj[i])
        rp.append(len(ci))
    return n, np.array(rp, dtype=np.int64), np.array(ci, dtype=np.int64)


def dominant_values(n, rp, ci, seed):
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = rows != ci
    vals = np.zeros(ci.size, dtype=np.complex128)
    vals[off] = -(1.0 + 0.5j * rng.uniform(-1, 1, off.sum()))
    rowsum = np.zeros(n)
    np.add.at(rowsum, rows[off], np.abs(vals[off]))
    np.add.at(rowsum, ci[off], np.abs(vals[off]))
    diag = ~off
    vals[diag] = 1.05 * rowsum[rows[diag]] + 0.1 + 1j * rng.uniform(0.0, 2.0, diag.sum())
    s = 10.0 ** rng.uniform(-2, 2, n)
    vals *= s[rows] * s[ci]
    return vals


@pytest.fixture(scope="module")
def config1(gpu_context):
    meta, _ = golden("config1")
    p = A.Problem(cube_config(meta))
    q = A.Problem.from_csr(p.n, p.ints("row_ptr"), p.ints("col_index"))
    vals, rhs = O.refill(p, meta["omega"])
    mp, mq = A.Mesh(gpu_context, p), A.Mesh(gpu_context, q)
    yield dict(p=p, q=q, mp=mp, mq=mq, vals=vals, rhs=rhs)
    mq.close()
    mp.close()


@pytest.mark.parametrize("precision", [A.APHI_FP64, A.APHI_FP32])
def test_cube_pattern_solution_is_the_cube_meshs(config1, precision):
    c = config1
    c["mp"].upload_system(c["vals"], c["rhs"])
    c["mq"].upload_system(c["vals"], c["rhs"])
    kw = dict(precision=precision, max_refinements=40, tolerance=1e-12)
    xp, rp_ = c["mp"].solve_symmetric(dev(c["rhs"]), **kw)
    xq, rq = c["mq"].solve_symmetric(dev(c["rhs"]), **kw)
    assert rp_["ok"] and rq["ok"]
    assert np.array_equal(xp.cpu().numpy().view(np.uint64), xq.cpu().numpy().view(np.uint64))
    if precision == A.APHI_FP64:
        ox, _, _ = O.solve_symmetric(c["p"], c["vals"], c["rhs"])
        assert np.linalg.norm(xq.cpu().numpy() - ox) / np.linalg.norm(ox) <= 1e-10


def test_bare_pattern_refuses_mesh_work(config1):
    m = config1["mq"]
    with pytest.raises(A.LogicError, match="no assembly plan"):
        m.refill(2 * np.pi * 50.0)
    x = dev(np.zeros(m.n))
    with pytest.raises(ValueError, match="per-tet unknowns"):
        m.terminal_current(x, 2 * np.pi * 50.0, length=0.1)


@pytest.mark.parametrize("k,seed", [(6, 1), (10, 2), (14, 3), (22, 4)])
def test_grid_matrix_fp64_matches_oracle(gpu_context, k, seed):
    n, rp, ci, vals = grid_matrix(k, seed)
    q = A.Problem.from_csr(n, rp, ci)
    m = A.Mesh(gpu_context, q)
    try:
        rng = np.random.default_rng(seed + 100)
        B = rng.normal(size=(4, n)) + 1j * rng.normal(size=(4, n))
        m.upload_system(vals, B[0])
        X, rep = m.solve_symmetric(dev(B), precision=A.APHI_FP64)
        assert rep["ok"]
        ox, _, st = O.csr_solve_symmetric_raw(n, rp, ci, vals, B, q.ints("perm"), iterations=10)
        assert st["ok"]
        Xg = X.cpu().numpy()
        for r in range(B.shape[0]):
            assert np.linalg.norm(Xg[r] - ox[r]) / np.linalg.norm(ox[r]) <= 1e-10
        assert rep["backward_error"] < 1e-14
    finally:
        m.close()


def test_grid_matrix_fp32_refined(gpu_context):
    """FP32 factor + FP64 refinement to 1e-12, and the distance to the oracle's
    solution within the first-order bound 4 kappa_2(D A D) (tol + 1e-15), kappa
    from a dense SVD of the equilibrated matrix."""
    k, seed, tol = 10, 5, 1e-12
    n, rp, ci, vals = grid_matrix(k, seed)
    q = A.Problem.from_csr(n, rp, ci)
    m = A.Mesh(gpu_context, q)
    try:
        rng = np.random.default_rng(seed)
        B = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
        m.upload_system(vals, B[0])
        X, rep = m.solve_symmetric(dev(B), precision=A.APHI_FP32, max_refinements=40, tolerance=tol)
        assert rep["ok"] and rep["backward_error"] <= tol and rep["scaled_backward_error"] <= tol
        ox, _, _ = O.csr_solve_symmetric_raw(n, rp, ci, vals, B, q.ints("perm"), iterations=10)
        d = O.equilibrate(q, vals, 10)
        rows = np.repeat(np.arange(n), np.diff(rp))
        dense = np.zeros((n, n), dtype=np.complex128)
        dense[rows, ci] = vals * d[rows] * d[ci]
        dense[ci, rows] = vals * d[rows] * d[ci]
        sv = np.linalg.svd(dense, compute_uv=False)
        bound = 4.0 * (sv[0] / sv[-1]) * (tol + 1e-15)
        Xg = X.cpu().numpy()
        for r in range(B.shape[0]):
            assert np.linalg.norm(Xg[r] - ox[r]) / np.linalg.norm(ox[r]) <= bound
    finally:
        m.close()



@pytest.mark.parametrize("shape", [(120, 12, 30, 60, 1), (160, 20, 24, 80, 2)])
def test_hub_front_with_many_children(gpu_context, shape):
    """A front with more children than the extend-add maps at once (12 children
    of at most 96 rows; 20 children of 104 rows, themselves large fronts): FP64
    to the oracle's solution (relative error <= 1e-10), FP32 + refinement to
    backward error <= 1e-12."""
    n, rp, ci = star_pattern(*shape)
    vals = dominant_values(n, rp, ci, shape[-1])
    q = A.Problem.from_csr(n, rp, ci)
    parent, rows = q.ints("super_parent"), np.diff(q.ints("super_row_ptr"))
    children = np.bincount(parent[parent >= 0], minlength=rows.size)
    assert children[rows > 96].max() > 8
    m = A.Mesh(gpu_context, q)
    try:
        rng = np.random.default_rng(shape[-1] + 7)
        B = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
        m.upload_system(vals, B[0])
        X, rep = m.solve_symmetric(dev(B), precision=A.APHI_FP64)
        assert rep["ok"]
        ox, _, st = O.csr_solve_symmetric_raw(n, rp, ci, vals, B, q.ints("perm"), iterations=10)
        assert st["ok"]
        Xg = X.cpu().numpy()
        for r in range(B.shape[0]):
            assert np.linalg.norm(Xg[r] - ox[r]) / np.linalg.norm(ox[r]) <= 1e-10
        X, rep = m.solve_symmetric(dev(B), precision=A.APHI_FP32, max_refinements=40, tolerance=1e-12)
        assert rep["ok"] and rep["precision_used"] == A.APHI_FP32 and rep["backward_error"] <= 1e-12
    finally:
        m.close()

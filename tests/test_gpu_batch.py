"""Batched problems on the device (aphi_batch_*): the seeded draws of the batched
workload (cube side U[0.05, 0.2] m, 4-12 hexes per edge, a conductor block of
1-4 hexes per side spanning z, sigma log-uniform in [1e6, 6e7] S/m, frequency
log-uniform in [1 Hz, 1 MHz]) in one device plan. Each problem's result must be
the one it has alone: the assembled system bit for bit (and the oracle's), the
FP64 solution within rounding of the single-problem solve and within 1e-10 of
the oracle's complex128 solution, FP32 + refinement to backward error 1e-12 per
problem (recomputed by the oracle), and the port current of each problem bit for
bit what aphi_terminal_current gives for the same solution."""
import numpy as np
import pytest

import aphi_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED = 4242
COUNT = 10


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


@pytest.fixture(scope="module")
def batch(gpu_context):
    configs, freqs = A.batched_cube_configs(SEED, COUNT)
    problems = A.Problem.cubes(configs)
    print("batch sizes:", [c.n for c in configs], "unknowns:", [p.n for p in problems], flush=True)
    b = A.Batch(gpu_context, problems)
    omegas = 2 * np.pi * freqs
    yield dict(b=b, problems=problems, configs=configs, omegas=omegas)
    b.close()


def slot_offsets(problems):
    return np.concatenate([[0], np.cumsum([int(p.counts["nnz"]) for p in problems])])


def test_batch_layout(batch):
    b, problems = batch["b"], batch["problems"]
    assert b.count == COUNT
    assert list(np.diff(b.rows)) == [p.n for p in problems]
    assert list(np.diff(b.tets)) == [int(p.counts["tets"]) for p in problems]


def test_batch_refill_bitwise(gpu_context, batch):
    """One element launch and one scatter launch for every problem, each at its
    own frequency: every problem's values and rhs are its single-mesh refill's
    bit for bit, and the oracle's."""
    b, problems, omegas = batch["b"], batch["problems"], batch["omegas"]
    b.refill(omegas)
    vals, rhs = b.system()
    so = slot_offsets(problems)
    for p, prob in enumerate(problems):
        m = A.Mesh(gpu_context, prob)
        try:
            m.refill(omegas[p])
            v1, r1 = m.system()
        finally:
            m.close()
        vb, rb = vals[so[p]:so[p + 1]], b.part(rhs, p)
        assert np.array_equal(vb.view(np.uint64), v1.view(np.uint64)), p
        assert np.array_equal(rb.view(np.uint64), r1.view(np.uint64)), p
        if batch["configs"][p].n <= 8:
            ov, orhs = O.refill(prob, omegas[p])
            assert np.array_equal(vb.view(np.uint64), ov.view(np.uint64)), p
            assert np.array_equal(rb.view(np.uint64), orhs.view(np.uint64)), p


def test_batch_fp64_equals_single_and_oracle(gpu_context, batch):
    b, problems, omegas = batch["b"], batch["problems"], batch["omegas"]
    b.refill(omegas)
    vals, rhs = b.system()
    so = slot_offsets(problems)
    x, reps = b.solve_symmetric(dev(rhs[None, :]), precision=A.APHI_FP64)
    X = x.cpu().numpy()[0]
    for p, prob in enumerate(problems):
        assert reps[p]["ok"] and reps[p]["backward_error"] <= 1e-12
        m = A.Mesh(gpu_context, prob)
        try:
            m.refill(omegas[p])
            x1, r1 = m.solve_symmetric(m.rhs_device()[None, :], precision=A.APHI_FP64)
            xs = x1.cpu().numpy()[0]
        finally:
            m.close()
        xp = b.part(X, p)
        rel = np.linalg.norm(xp - xs) / np.linalg.norm(xs)
        print(f"problem {p} (n={batch['configs'][p].n}, {prob.n} unknowns): batch vs alone {rel:.1e}, "
              f"bitwise {np.array_equal(xp.view(np.uint64), xs.view(np.uint64))}", flush=True)
        assert rel <= 1e-13
        # the problem's own backward error, recomputed by the oracle
        ob = O.backward_error(prob, vals[so[p]:so[p + 1]], xp, b.part(rhs, p))[0]
        assert ob <= 1e-12 and abs(reps[p]["backward_errors"][0] / ob - 1) <= max(1e-12, 3 * prob.n * 2.0 ** -53)
        if batch["configs"][p].n <= 8:
            ox, _, _ = O.solve_symmetric(prob, vals[so[p]:so[p + 1]], b.part(rhs, p))
            assert np.linalg.norm(xp - ox) / np.linalg.norm(ox) <= 1e-10, p


def test_batch_fp32_refined(gpu_context, batch):
    """FP32-MFMA factor of every problem at once + per-problem GMRES refinement:
    every problem and right-hand side at backward error <= 1e-12 (both the
    reference's measure, recomputed by the oracle, and the equilibrated one), and
    each problem's solution within its conditioning bound of the FP64 one:
    ||D^-1 (x32 - x64)|| <= 4 ||(D A D)^-1|| (||D r32|| + ||D r64||), with the
    inverse norm estimated over the whole batch (so at least each problem's)."""
    import torch
    b, problems, omegas = batch["b"], batch["problems"], batch["omegas"]
    b.refill(omegas)
    vals, rhs = b.system()
    so = slot_offsets(problems)
    B = np.vstack([rhs[None, :], A.random_complex_rhs(gpu_context, SEED, 3, b.n).cpu().numpy()])
    x64, r64 = b.solve_symmetric(dev(B), precision=A.APHI_FP64)
    d = np.concatenate([O.equilibrate(prob, vals[so[p]:so[p + 1]], 10) for p, prob in enumerate(problems)])
    di = torch.from_numpy(1.0 / d).cuda()
    g = np.random.default_rng(1)
    v = dev(g.standard_normal((1, b.n)) + 1j * g.standard_normal((1, b.n)))
    v = v / torch.linalg.vector_norm(v)
    inv_norm = 0.0
    for _ in range(10):
        w = b.solve(v * di) * di
        inv_norm = float(torch.linalg.vector_norm(w))
        z = (b.solve(w.conj().resolve_conj() * di) * di).conj().resolve_conj()
        v = z / torch.linalg.vector_norm(z)
    x32, r32 = b.solve_symmetric(dev(B), precision=A.APHI_FP32, max_refinements=30, tolerance=1e-12)
    X32, X64 = x32.cpu().numpy(), x64.cpu().numpy()
    for p, prob in enumerate(problems):
        assert r32[p]["ok"] and r32[p]["backward_error"] <= 1e-12 and r32[p]["scaled_backward_error"] <= 1e-12, r32[p]
        vp, dp = vals[so[p]:so[p + 1]], b.part(d, p)
        worst = 0.0
        for r in range(B.shape[0]):
            xa, xb, bb = b.part(X32[r], p), b.part(X64[r], p), b.part(B[r], p)
            assert O.backward_error(prob, vp, xa, bb)[0] <= 1e-12
            s32 = np.linalg.norm(dp * O.residual(prob, vp, xa, bb))
            s64 = np.linalg.norm(dp * O.residual(prob, vp, xb, bb))
            worst = max(worst, np.linalg.norm((xa - xb) / dp) / (4.0 * inv_norm * (s32 + s64)))
        print(f"problem {p}: fp32+IR steps {r32[p]['refinement_steps']} precision {r32[p]['precision_used']} "
              f"be {r32[p]['backward_error']:.1e} scaled {r32[p]['scaled_backward_error']:.1e} "
              f"distance/bound {worst:.1e}", flush=True)
        assert worst <= 1.0


def test_batch_terminal_current_equals_single(gpu_context, batch):
    b, problems, omegas, configs = batch["b"], batch["problems"], batch["omegas"], batch["configs"]
    b.refill(omegas)
    _, rhs = b.system()
    x, _ = b.solve_symmetric(dev(rhs[None, :]), precision=A.APHI_FP64)
    sides = [c.side for c in configs]
    I, R, L = b.terminal_current(x, omegas, sides)
    for p, prob in enumerate(problems):
        m = A.Mesh(gpu_context, prob)
        try:
            I1, R1, L1 = m.terminal_current(b.part(x, p).contiguous(), omegas[p], sides[p])
        finally:
            m.close()
        assert I[p, 0] == I1[0] and R[p, 0] == R1[0] and L[p, 0] == L1[0], p
        # the low-frequency resistance of the conductor block, length / (sigma
        # area), where the skin depth sqrt(2 / (w mu0 sigma)) is much wider than it
        c = configs[p]
        h = c.side / c.n
        width = max(c.bar[1] - c.bar[0], c.bar[3] - c.bar[2]) * h
        area = (c.bar[1] - c.bar[0]) * (c.bar[3] - c.bar[2]) * h * h
        skin = np.sqrt(2.0 / (omegas[p] * 4e-7 * np.pi * c.sigma))
        if skin > 20 * width:
            print(f"problem {p}: R {R[p, 0]:.6e} vs dc {c.side / (c.sigma * area):.6e}", flush=True)
            assert abs(R[p, 0] / (c.side / (c.sigma * area)) - 1) < 0.01


def test_batch_refuses_single_problem_calls(gpu_context, batch):
    b = batch["b"]
    import torch
    x = torch.zeros(1, b.n, dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError):
        A.Mesh.backward_error(b, x, x)
    with pytest.raises(ValueError):
        A.Mesh.terminal_current(b, x, 1.0, 0.1)
    with pytest.raises(ValueError):
        A.Mesh.solve_symmetric(b, x)

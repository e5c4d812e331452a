"""Device path against the oracle, through the C-ABI: refill bit for bit,
equilibration, the factor's D (unique for the ordering), solutions of the plain
FP64 path (relative error <= 1e-10) and of the FP32 + FP64
refinement path (backward error <= 1e-12), pivot refusal, determinism."""
import numpy as np
import pytest

from conftest import conditioning, cube_config, golden

import aphi_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NUMERIC = ["cube2", "cube3_offcentre", "config1", "cube8", "cube5_draw", "cube4_current"]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


@pytest.fixture(scope="module", params=NUMERIC)
def case(request, gpu_context):
    meta, _ = golden(request.param)
    p = A.Problem(cube_config(meta))
    m = A.Mesh(gpu_context, p)
    vals, rhs = O.refill(p, meta["omega"])
    x, D, st = O.solve_symmetric(p, vals, rhs)
    yield dict(meta=meta, p=p, m=m, vals=vals, rhs=rhs, x=x, D=D, st=st)
    m.close()


def test_refill_bitwise_equals_oracle(case):
    m = case["m"]
    m.refill(case["meta"]["omega"])
    vals, rhs = m.system()
    assert np.array_equal(vals.view(np.uint64), case["vals"].view(np.uint64))
    assert np.array_equal(rhs.view(np.uint64), case["rhs"].view(np.uint64))


def test_refill_scaled_phi_bitwise(case):
    m, p = case["m"], case["p"]
    omega = 2 * np.pi * 3.3e5
    m.refill(omega, A.APHI_SCALED_PHI)
    vals, rhs = m.system()
    ov, orhs = O.refill(p, omega, conditioning=2)
    assert np.array_equal(vals.view(np.uint64), ov.view(np.uint64))
    assert np.array_equal(rhs.view(np.uint64), orhs.view(np.uint64))


@pytest.mark.parametrize("index", [5, 23, 47, 63])
def test_refill_bitwise_at_sweep_points(case, index):
    """At the log sweep's points (expand_frequency_sweep, 1 Hz-100 MHz), whose
    frequencies have no short binary expansion, the device refill is still the
    oracle's bit for bit (the oracle is the reference's, pinned at such a point by
    the cube5_draw fixture), in both symmetric conditionings."""
    m, p = case["m"], case["p"]
    omega = 2 * np.pi * A.frequency_sweep(1.0, 1e8, 64)[index]
    for cond in (A.APHI_ROW_SCALED, A.APHI_SCALED_PHI):
        m.refill(omega, cond)
        vals, rhs = m.system()
        ov, orhs = O.refill(p, omega, conditioning=cond)
        assert np.array_equal(vals.view(np.uint64), ov.view(np.uint64)), cond
        assert np.array_equal(rhs.view(np.uint64), orhs.view(np.uint64)), cond


def test_bound_problem_mesh_equals_cube_mesh(case, gpu_context):
    """A device mesh made from the cube's bound objects flattened
    (aphi_problem_create_bound, the drop-in boundary for any bound mesh): the same
    refill, solution and port current bit for bit as the cube's own mesh."""
    m, p, meta = case["m"], case["p"], case["meta"]
    q = A.Problem.from_bound(p.bound_arrays())
    mq = A.Mesh(gpu_context, q)
    try:
        m.refill(meta["omega"])
        mq.refill(meta["omega"])
        v1, r1 = m.system()
        v2, r2 = mq.system()
        assert np.array_equal(v1.view(np.uint64), v2.view(np.uint64))
        assert np.array_equal(r1.view(np.uint64), r2.view(np.uint64))
        b = dev(np.vstack([r1, r1 * (2 - 1j)]))
        x1, _ = m.solve_symmetric(b, precision=A.APHI_FP64)
        x2, _ = mq.solve_symmetric(b, precision=A.APHI_FP64)
        assert np.array_equal(x1.cpu().numpy().view(np.uint64), x2.cpu().numpy().view(np.uint64))
        I1, R1, L1 = m.terminal_current(x1, meta["omega"], meta["side"])
        I2, R2, L2 = mq.terminal_current(x2, meta["omega"], meta["side"])
        assert np.array_equal(I1, I2) and np.array_equal(R1, R2) and np.array_equal(L1, L2)
    finally:
        mq.close()


def test_refill_refusals(case):
    m = case["m"]
    with pytest.raises(ValueError):
        m.refill(100.0, A.APHI_NATURAL)
    with pytest.raises(ValueError):
        m.refill(0.0, A.APHI_ROW_SCALED)


def test_equilibration_matches_oracle(case):
    """Ruiz equilibration bit for bit (row maxima of |a_ij| with the host
    library's rounding of the modulus, 1/sqrt, products in the reference's
    order), so the equilibrated matrix is the reference's."""
    m, p = case["m"], case["p"]
    m.refill(case["meta"]["omega"])
    d = m.equilibrate(10)
    od = O.equilibrate(p, case["vals"], 10)
    assert np.array_equal(d.view(np.uint64), od.view(np.uint64))


@pytest.mark.parametrize("precision,tol", [(A.APHI_FP64, 1e-8), (A.APHI_FP32, 1e-2)])
def test_factor_diagonal_matches_oracle(case, precision, tol):
    """LDL^T without pivoting is unique for the ordering; the factorization order is
    a postorder of the reference's elimination tree, so D is the oracle's D permuted.
    FP64: normwise to 1e-10 of max|D| and every pivot to 1e-8 relative. FP32:
    normwise to 1e-2 of max|D| (small pivots lose digits to cancellation in single
    precision; the refinement recovers the solution, checked below)."""
    m, p = case["m"], case["p"]
    m.refill(case["meta"]["omega"])
    m.equilibrate(10)
    st = m.factorize_ldlt(precision)
    assert st["ok"]
    D = m.factor_diagonal()
    expect = case["D"][p.ints("col_ref")]
    o = case["st"]
    if precision == A.APHI_FP64:
        assert np.abs(D - expect).max() / np.abs(expect).max() < 1e-10
        assert np.max(np.abs(D - expect) / np.abs(expect)) < tol
        for k in ("smallest_pivot", "largest_pivot", "largest_multiplier"):
            assert abs(st[k] / o[k] - 1) < tol, k
    else:
        assert np.abs(D - expect).max() / np.abs(expect).max() < tol
        assert abs(st["largest_pivot"] / o["largest_pivot"] - 1) < tol
        assert abs(st["largest_multiplier"] / o["largest_multiplier"] - 1) < 1e-2
    assert st["nnz"] == p.counts["predicted_nnz"]


def test_fp64_solution_matches_oracle(case):
    m = case["m"]
    m.refill(case["meta"]["omega"])
    x, rep = m.solve_symmetric(dev(case["rhs"]), precision=A.APHI_FP64)
    xg = x.cpu().numpy()
    assert rep["ok"]
    assert np.linalg.norm(xg - case["x"]) / np.linalg.norm(case["x"]) <= 1e-10
    assert rep["backward_error"] < 1e-14


@pytest.mark.parametrize("tol", [1e-12, 1e-15])
def test_fp32_refined_solution(case, tol):
    """FP32 factor + FP64 refinement to backward error `tol` (normwise and on the
    equilibrated system). Its distance to the reference's complex128 solution is
    bounded by the conditioning: kappa_2(D A D) * (tol + the reference's own
    backward error, ~1e-15), with a factor 4 for the first-order bound and the
    change of norm under D. kappa comes from a dense SVD of the golden matrix
    (tests/golden/make_conditioning.py): 1.5e7 for cube8."""
    m = case["m"]
    m.refill(case["meta"]["omega"])
    x, rep = m.solve_symmetric(dev(case["rhs"]), precision=A.APHI_FP32, max_refinements=40, tolerance=tol)
    assert rep["ok"] and rep["backward_error"] <= tol and rep["scaled_backward_error"] <= tol
    kappa = conditioning(case["meta"]["case"])["cond2_equilibrated"]
    bound = 4.0 * kappa * (tol + 1e-15)
    assert np.linalg.norm(x.cpu().numpy() - case["x"]) / np.linalg.norm(case["x"]) <= bound


def test_multi_rhs_matches_oracle(case):
    m, p = case["m"], case["p"]
    m.refill(case["meta"]["omega"])
    rng = np.random.default_rng(17)
    B = rng.normal(size=(8, p.n)) + 1j * rng.normal(size=(8, p.n))
    X, rep = m.solve_symmetric(dev(B), precision=A.APHI_FP64)
    ox, _, _ = O.solve_symmetric(p, case["vals"], B)
    Xg = X.cpu().numpy()
    for r in range(8):
        assert np.linalg.norm(Xg[r] - ox[r]) / np.linalg.norm(ox[r]) <= 1e-10
    be, _ = m.backward_error(X, dev(B))
    assert np.all(be < 1e-14)


def test_pivot_floor_refusal(case):
    m, p = case["m"], case["p"]
    m.refill(case["meta"]["omega"])
    m.equilibrate(10)
    st = m.factorize_ldlt(A.APHI_FP64, pivot_floor=1e30)
    assert not st["ok"] and st["columns_done"] == 1
    with pytest.raises(ValueError):
        m.solve(dev(case["rhs"]))  # no usable factor


def test_pivot_failure_statistics(case):
    """A floor between the pivots: the factorization stops at the first reference
    column whose pivot is at or below it (src/factorization.cpp:191-210), and the
    report's extremes cover only what the reference's loop reached by then -- the
    oracle's figures for the same floor, not those of the whole factor."""
    m, p = case["m"], case["p"]
    mags = np.sort(np.abs(case["D"]))
    i = len(mags) // 2
    while i < len(mags) - 2 and mags[i + 1] < mags[i] * (1 + 1e-6):
        i += 1  # a clear gap, so both sides classify every pivot alike
    floor = 0.5 * (mags[i] + mags[i + 1])
    _, _, o = O.solve_symmetric(p, case["vals"], case["rhs"], pivot_floor=floor)
    assert not o["ok"] and o["columns_done"] < p.n
    m.refill(case["meta"]["omega"])
    m.equilibrate(10)
    st = m.factorize_ldlt(A.APHI_FP64, pivot_floor=floor)
    assert not st["ok"] and st["columns_done"] == o["columns_done"]
    for k in ("smallest_pivot", "largest_pivot", "largest_multiplier"):
        assert np.isfinite(st[k]) and abs(st[k] / o[k] - 1) < 1e-8, (k, st[k], o[k])
    # the same mesh factors in full again afterwards
    assert m.factorize_ldlt(A.APHI_FP64)["ok"]


def test_compute_fields_bitwise_and_vtk_bytes(case, tmp_path):
    """compute_fields on the device from the reference's own solution: every
    FieldOutput array bit for bit the oracle's (whose VTK is the reference's, byte
    for byte), and the VTK written from the device's fields is the reference's file."""
    import hashlib
    m, p, meta = case["m"], case["p"], case["meta"]
    _, arrays = golden(meta["case"])
    x = arrays["x"]
    f = m.compute_fields(dev(x), meta["omega"])
    o = O.compute_fields(p, x, meta["omega"])
    for k in ("a_tet", "b_tet", "e_tet", "j_tet", "sigma_tet", "a_node", "b_node", "h_node", "e_node", "j_node",
              "vertex_weight"):
        assert np.array_equal(np.asarray(f[k]).view(np.uint64), np.asarray(o[k]).view(np.uint64)), k
    assert np.array_equal(f["on_material_interface"], o["on_material_interface"])
    assert f["num_interface_nodes"] == o["num_interface_nodes"] > 0
    assert f["num_orphan_vertices"] == o["num_orphan_vertices"]
    out = tmp_path / "fields.vtk"
    p.write_vtk_fields(out, x, meta["omega"], f)
    assert hashlib.sha256(out.read_bytes()).hexdigest() == meta["sha256"]["fields.vtk"]


def test_deterministic_bits(case):
    m = case["m"]
    m.refill(case["meta"]["omega"])
    x1, _ = m.solve_symmetric(dev(case["rhs"]), precision=A.APHI_FP32, max_refinements=5)
    m.refill(case["meta"]["omega"])
    x2, _ = m.solve_symmetric(dev(case["rhs"]), precision=A.APHI_FP32, max_refinements=5)
    assert np.array_equal(x1.cpu().numpy().view(np.uint64), x2.cpu().numpy().view(np.uint64))


def test_precision_switch(case):
    """FP64 then FP32 then FP64 factorizations on one mesh reuse its device
    buffers; every solve must be finite and as accurate as a fresh one."""
    import torch
    m, p = case["m"], case["p"]
    m.refill(case["meta"]["omega"])
    m.equilibrate(10)
    B = dev(np.tile(case["rhs"], (3, 1)))
    errs = {}
    for prec in (A.APHI_FP64, A.APHI_FP32, A.APHI_FP64, A.APHI_FP32):
        X = torch.zeros_like(B)
        m.factorize_ldlt(prec)
        m.solve(B, X)
        x = X.cpu().numpy()
        assert np.isfinite(x).all()
        e = max(np.linalg.norm(x[r] - case["x"]) / np.linalg.norm(case["x"]) for r in range(3))
        errs.setdefault(prec, []).append(e)
    assert errs[A.APHI_FP64][0] == errs[A.APHI_FP64][1] and errs[A.APHI_FP32][0] == errs[A.APHI_FP32][1]
    assert errs[A.APHI_FP64][0] <= 1e-10


def test_terminal_current_bitwise(case):
    """Device port current / R / L on the reference's own solution: bit for bit
    the reference's (the per-tet terms and the tet-order sum are both exact)."""
    m = case["m"]
    _, arrays = golden(case["meta"]["case"])
    meta = case["meta"]
    I, R, L = m.terminal_current(dev(arrays["x"][None, :]), meta["omega"], meta["side"])
    assert I[0] == complex(meta["I_re"], meta["I_im"])
    assert R[0] == meta["R"] and L[0] == meta["L"]
    with pytest.raises(ValueError):
        m.terminal_current(dev(arrays["x"][None, :]), meta["omega"], -1.0)


def test_terminal_current_after_device_solve(case):
    """I, R and L from the device FP64 solution (several right-hand sides at
    once) against the reference's. The solution matches to 1e-10, and I is a
    linear functional of it, so I agrees to 1e-8 relative; Z = V/I then agrees
    to the same relative error in |Z|, which bounds R = Re Z and omega L = Im Z
    componentwise (R alone can be tiny next to |Z|: for cube8 Re I / |I| is
    2.5e-4, so R's own relative error is larger)."""
    m = case["m"]
    meta = case["meta"]
    m.refill(meta["omega"])
    B = np.tile(case["rhs"], (3, 1))
    x, rep = m.solve_symmetric(dev(B), precision=A.APHI_FP64)
    I, R, L = m.terminal_current(x, meta["omega"], meta["side"])
    I_ref = complex(meta["I_re"], meta["I_im"])
    Z_ref = 1.0 / I_ref
    w = meta["omega"]
    for r in range(3):
        assert abs(I[r] - I_ref) <= 1e-8 * abs(I_ref)
        assert abs(R[r] - meta["R"]) <= 1e-8 * abs(Z_ref)
        assert abs(w * L[r] - w * meta["L"]) <= 1e-8 * abs(Z_ref)


def test_upload_equals_refill(case):
    m = case["m"]
    m.upload_system(case["vals"], case["rhs"])
    x, rep = m.solve_symmetric(dev(case["rhs"]), precision=A.APHI_FP64)
    assert np.linalg.norm(x.cpu().numpy() - case["x"]) / np.linalg.norm(case["x"]) <= 1e-10


@pytest.mark.parametrize("freq", [10.0, 1e4, 1e7])
def test_config2_sweep_points(gpu_context, freq):
    """Config 2 (24^3 hexes, 196,978 unknowns) at sweep points: the plain FP64
    path and the FP32 + refinement path both reach backward error <= 1e-12 and
    agree with each other."""
    meta, _ = golden("config2")
    p = A.Problem(cube_config(meta))
    m = A.Mesh(gpu_context, p)
    m.refill(2 * np.pi * freq)
    vals, rhs = m.system()
    rng = np.random.default_rng(5)
    B = np.vstack([rhs[None, :], rng.normal(size=(7, p.n)) + 1j * rng.normal(size=(7, p.n))])
    x64, r64 = m.solve_symmetric(dev(B), precision=A.APHI_FP64)
    assert r64["backward_error"] <= 1e-12
    x32, r32 = m.solve_symmetric(dev(B), precision=A.APHI_FP32, max_refinements=30)
    assert r32["backward_error"] <= 1e-12 and r32["scaled_backward_error"] <= 1e-12
    a, b = x64.cpu().numpy(), x32.cpu().numpy()
    for r in range(B.shape[0]):
        assert np.linalg.norm(a[r] - b[r]) / np.linalg.norm(a[r]) <= 1e-5
    # Driven to backward error 1e-15, both precisions satisfy the equations to
    # that level, checked independently by the oracle's backward error (the
    # reference's formula, on the CPU). Their mutual distance is then bounded by
    # kappa * 2e-15 only; config2 at high frequency is too large for a dense
    # condition number, so the agreement bound here is the loose 1e-7.
    x64s, r64s = m.solve_symmetric(dev(B), precision=A.APHI_FP64, max_refinements=10, tolerance=1e-15)
    x32s, r32s = m.solve_symmetric(dev(B), precision=A.APHI_FP32, max_refinements=40, tolerance=1e-15)
    a, c = x64s.cpu().numpy(), x32s.cpu().numpy()
    print("tight", freq, r64s["backward_error"], r64s["refinement_steps"], r32s["backward_error"],
          r32s["refinement_steps"])
    for r in range(B.shape[0]):
        assert O.backward_error(p, vals, a[r], B[r])[0] <= 2e-15
        assert O.backward_error(p, vals, c[r], B[r])[0] <= 2e-15
        assert np.linalg.norm(a[r] - c[r]) / np.linalg.norm(a[r]) <= 1e-7
    m.close()


@pytest.mark.parametrize("precision", [A.APHI_FP32, A.APHI_FP64])
def test_solve_graph_replays_match_eager(case, precision, monkeypatch):
    """The solve's level loops run as a captured graph; every replay, with other
    solves (other right-hand-side counts) and synchronisations in between, equals
    the launch-by-launch solve bit for bit."""
    import torch
    m, n = case["m"], case["p"].n
    m.refill(case["meta"]["omega"])
    m.equilibrate(10)
    assert m.factorize_ldlt(precision)["ok"]
    gen = np.random.default_rng(7)
    b1 = dev(gen.standard_normal((16, n)) + 1j * gen.standard_normal((16, n)))
    b2 = dev(gen.standard_normal((3, n)) + 1j * gen.standard_normal((3, n)))
    b3 = dev(gen.standard_normal((20, n)) + 1j * gen.standard_normal((20, n)))
    monkeypatch.setenv("APHI_SOLVE_GRAPHS", "0")
    eager1 = m.solve(b1).cpu().numpy()
    eager3 = m.solve(b3).cpu().numpy()
    monkeypatch.delenv("APHI_SOLVE_GRAPHS")
    for other in (None, b2, b3, b2):
        x1 = m.solve(b1)
        torch.cuda.synchronize()
        assert np.array_equal(x1.cpu().numpy().view(np.uint64), eager1.view(np.uint64))
        if other is not None:
            xo = m.solve(other)
            torch.cuda.synchronize()
            if other is b3:
                assert np.array_equal(xo.cpu().numpy().view(np.uint64), eager3.view(np.uint64))


def test_potential_file_from_device_solution(case, tmp_path):
    """write_potential fed from the device solution (a tensor): the values it writes
    are within 1e-10 of those written from the reference's own solution."""
    m, p, meta = case["m"], case["p"], case["meta"]
    m.refill(meta["omega"])
    x, rep = m.solve_symmetric(dev(case["rhs"]), precision=A.APHI_FP64)
    assert rep["ok"]
    p.write_potential(tmp_path / "dev.out", x, meta["omega"])
    p.write_potential(tmp_path / "ref.out", case["x"], meta["omega"])

    def values(path):
        rows = [ln.split() for ln in open(path) if not ln.startswith("#")]
        return np.array([complex(float(r[4]), float(r[5])) for r in rows])

    a, b = values(tmp_path / "dev.out"), values(tmp_path / "ref.out")
    assert a.shape == b.shape and np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-10

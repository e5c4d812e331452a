"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it:
the reference's golden vectors bit for bit, closed forms of the element kernels
(tests/test_element_matrix.cpp:89-134, 555-575 of the reference), the hand 2x2
factor and the zero-pivot refusal (tests/test_factorization.cpp:111-136, 253-276),
no conjugation, and agreement with a dense LU solve."""
import hashlib

import numpy as np
import pytest

from conftest import cube_config, golden

import aphi_amd as A
from oracle import oracle as O

NUMERIC = ["cube2", "cube3_offcentre", "config1", "cube8", "cube5_draw", "cube4_current"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module", params=NUMERIC)
def case(request):
    meta, arrays = golden(request.param)
    p = A.Problem(cube_config(meta))
    vals, rhs = O.refill(p, meta["omega"])
    return meta, arrays, p, vals, rhs


def test_refill_bitwise_equals_reference(case):
    meta, arrays, _, vals, rhs = case
    assert sha(vals) == meta["sha256"]["values"]
    assert sha(rhs) == meta["sha256"]["rhs"]
    if "values" in arrays:
        assert np.array_equal(vals.view(np.uint64), arrays["values"].view(np.uint64))


def test_equilibration_bitwise_equals_reference(case):
    meta, arrays, p, vals, _ = case
    d = O.equilibrate(p, vals, 10)
    assert sha(d) == meta["sha256"]["equil_d"]
    assert d.min() == meta["equilibration_min"] and d.max() == meta["equilibration_max"]


def test_solve_bitwise_equals_reference(case):
    meta, arrays, p, vals, rhs = case
    x, D, st = O.solve_symmetric(p, vals, rhs)
    assert st["ok"]
    assert sha(x) == meta["sha256"]["x"]
    assert sha(D) == meta["sha256"]["diag"]
    assert st["backward_error"] == meta["backward_error"]
    assert st["smallest_pivot"] == meta["smallest_pivot"]
    assert st["largest_multiplier"] == meta["largest_multiplier"]


def unit_tet(scale=1.0):
    return np.array([0, 0, 0, scale, 0, 0, 0, scale, 0, 0, 0, scale], dtype=float)


def test_terminal_current_bitwise_equals_reference(case):
    """compute_fields + terminal_current + impedance on the reference's own x
    (src/postprocess.cpp:311-400, 526-547): I, R and L bit for bit."""
    meta, arrays, p, _, _ = case
    I, R, L = O.terminal_current(p, arrays["x"], meta["omega"], meta["side"])
    assert I == complex(meta["I_re"], meta["I_im"])
    assert R == meta["R"] and L == meta["L"]


def test_terminal_current_linear_and_refusals(case):
    """The current is linear in the free unknowns once the prescribed potentials
    are zero: I(x1 + x2) - I(x1) - I(x2) vanishes to rounding; a non-positive
    length is refused as in src/postprocess.cpp:528-530."""
    meta, arrays, p, _, _ = case
    x = arrays["x"]
    rng = np.random.default_rng(3)
    y = rng.normal(size=x.shape) + 1j * rng.normal(size=x.shape)
    _, _, _, t1 = O.terminal_current(p, x, meta["omega"], meta["side"], terms=True)
    _, _, _, t2 = O.terminal_current(p, y, meta["omega"], meta["side"], terms=True)
    _, _, _, t12 = O.terminal_current(p, x + y, meta["omega"], meta["side"], terms=True)
    _, _, _, t0 = O.terminal_current(p, np.zeros_like(x), meta["omega"], meta["side"], terms=True)
    # affine: terms(x) = M x + terms(0)
    lhs = t12 - t1 - t2 + t0
    assert np.max(np.abs(lhs)) <= 1e-9 * max(np.max(np.abs(t1)), np.max(np.abs(t2)))
    with pytest.raises(ValueError):
        O.terminal_current(p, x, meta["omega"], 0.0)


def test_element_curl_curl_hand_values():
    dots = np.array([[8, -4, -4, 4, 0, 4], [-4, 8, -4, -4, 4, 0], [-4, -4, 8, 0, -4, -4],
                     [4, -4, 0, 4, 0, 0], [0, 4, -4, 0, 4, 0], [4, 0, -4, 0, 0, 4]], dtype=float)
    el = O.element(unit_tet(), 1.0, 0j, 0j)
    np.testing.assert_allclose(el[:36].reshape(6, 6).real, dots / 6.0, atol=1e-14)
    assert np.all(el[36:] == 0)  # beta == 0 gives exact zeros (test_element_matrix.cpp:720-738)
    el7 = O.element(unit_tet(), 7.0, 0j, 0j)
    assert abs(el7[0] - 7.0 * 4.0 / 3.0) < 1e-13


def test_element_partition_of_unity_and_symmetry():
    rng = np.random.default_rng(3)
    p = unit_tet(1.7) + rng.normal(scale=0.1, size=12)
    el = O.element(p, 3.0, 2.0 - 5.0j, 1.5 + 0.25j)
    aa, pp = el[:36].reshape(6, 6), el[96:].reshape(10, 10)
    assert np.abs(aa - aa.T).max() < 1e-12 * np.abs(aa).max()
    assert np.abs(pp - pp.T).max() < 1e-12 * np.abs(pp).max()
    assert np.abs(pp.sum(axis=1)).max() < 1e-13 * np.abs(pp).max() * 10
    assert np.abs(pp.sum(axis=0)).max() < 1e-13 * np.abs(pp).max() * 10
    ap = el[36:96].reshape(6, 10)
    assert np.abs(ap.sum(axis=1)).max() < 1e-13 * np.abs(ap).max() * 10  # sum_b grad S_b = 0


def test_element_volume_scaling():
    """Curl-curl scales as 1/h, mass as h^3 (test_element_matrix.cpp:693-716)."""
    e1 = O.element(unit_tet(1.0), 1.0, 1j, 0j)
    e2 = O.element(unit_tet(2.0), 1.0, 0j, 0j)
    assert abs(e2[0] - e1[0].real / 2.0) < 1e-13
    e3 = O.element(unit_tet(2.0), 0.0, 1j, 0j)
    mass1 = e1[0] - e1[0].real
    assert abs(e3[0] - 2.0 * mass1) < 1e-13  # W ~ 1/h, V ~ h^3 -> h


def csr_upper(dense):
    n = dense.shape[0]
    rp, ci, v = [0], [], []
    for i in range(n):
        for j in range(i, n):
            if dense[i, j] != 0 or i == j:
                ci.append(j)
                v.append(dense[i, j])
        rp.append(len(ci))
    return np.array(rp), np.array(ci), np.array(v, dtype=complex)


def test_hand_two_by_two():
    rp, ci, v = csr_upper(np.array([[2.0, 3.0], [3.0, 5.0]], dtype=complex))
    x, D, st = O.csr_solve_symmetric_raw(2, rp, ci, v, np.array([[1.0, 0.0]], dtype=complex), [0, 1])
    assert st["ok"]
    assert abs(D[0] - 2.0) < 1e-15 and abs(D[1] - 0.5) < 1e-15
    np.testing.assert_allclose(x[0], np.linalg.solve([[2, 3], [3, 5]], [1, 0]), rtol=1e-14)


def test_zero_pivot_refused():
    rp, ci, v = csr_upper(np.array([[0.0, 1.0], [1.0, 0.0]], dtype=complex))
    _, _, st = O.csr_solve_symmetric_raw(2, rp, ci, v, np.array([[1.0, 1.0]], dtype=complex), [0, 1])
    assert not st["ok"] and st["columns_done"] == 1


def test_no_conjugation_and_dense_lu_agreement():
    rng = np.random.default_rng(424242)
    n = 40
    m = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    m = np.where(rng.random((n, n)) < 0.15, m, 0)
    a = m + m.T + np.diag(8.0 + 3j + rng.normal(size=n))
    rp, ci, v = csr_upper(a)
    b = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    perm = rng.permutation(n)
    x, _, st = O.csr_solve_symmetric_raw(n, rp, ci, v, b, perm, iterations=10)
    assert st["ok"]
    ref = np.linalg.solve(a, b.T).T
    assert np.abs(x - ref).max() / np.abs(ref).max() < 1e-12
    assert O.reconstruction_error(n, rp, ci, v, perm) < 1e-13

"""Problems made from a bare upper-triangle CSR pattern (aphi_problem_create_csr):
the pattern checks of SparseSymmetric::from_pattern
(include/aphi_solver/sparse_symmetric.hpp:38-83), analyze's factor budget
(src/symbolic.cpp:126-151), and the symbolic plan, which for the pattern of a
cube problem must be the cube problem's own (that plan is pinned by the golden
perm / parent / col_count fixtures). Host logic only."""
import numpy as np
import pytest

import aphi_amd as A

SYMBOLIC = ["perm", "iperm", "parent", "row_count", "col_count", "super_first", "super_parent", "super_row_ptr",
            "super_rows", "super_rel", "super_level", "factor_perm", "col_ref", "slot_to_panel", "row_ptr",
            "col_index"]


def grid_pattern(k):
    """Upper triangle of the 7-point stencil on a k^3 grid."""
    n = k ** 3
    rows = []
    for z in range(k):
        for y in range(k):
            for x in range(k):
                i = (z * k + y) * k + x
                cols = [i]
                if x + 1 < k:
                    cols.append(i + 1)
                if y + 1 < k:
                    cols.append(i + k)
                if z + 1 < k:
                    cols.append(i + k * k)
                rows.append(cols)
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return n, rp, np.concatenate([np.array(r) for r in rows])


@pytest.mark.parametrize("n_cube", [3, 4, 6])
def test_cube_pattern_gives_the_cube_plan(n_cube):
    p = A.Problem(A.CubeConfig(n=n_cube, bar=(1, 2, 1, 2) if n_cube < 6 else (2, 4, 2, 4)))
    q = A.Problem.from_csr(p.n, p.ints("row_ptr"), p.ints("col_index"))
    for name in SYMBOLIC:
        assert np.array_equal(p.ints(name), q.ints(name)), name
    for key in ("nnz", "predicted_nnz", "num_roots", "num_super", "num_levels", "factor_entries", "cb_pool_entries"):
        assert p.counts[key] == q.counts[key], key
    assert p.counts["flops"] == q.counts["flops"]
    assert q.counts["num_total"] == p.n and q.counts["tets"] == 0 and q.counts["contributions"] == 0


def test_grid_plan_invariants():
    """A matrix no mesh produced: the elimination tree is a forest whose parents
    follow their children, the supernodes cover every column once, and every
    stored slot lands in a distinct factor entry."""
    n, rp, ci = grid_pattern(7)
    q = A.Problem.from_csr(n, rp, ci)
    parent = q.ints("parent")
    assert np.all((parent == -1) | (parent > np.arange(n)))
    first = q.ints("super_first")
    assert first[0] == 0 and first[-1] == n and np.all(np.diff(first) > 0)
    assert np.array_equal(np.sort(q.ints("factor_perm")), np.arange(n))
    s2p = q.ints("slot_to_panel")
    assert s2p.size == ci.size and np.unique(s2p).size == s2p.size
    # nnz(L) with the diagonal is at least the stored upper triangle
    assert q.counts["predicted_nnz"] >= ci.size


def test_pattern_refusals_follow_from_pattern():
    n, rp, ci = grid_pattern(3)
    bad = ci.copy()
    bad[1] = 0  # row 0: (0, 0), (0, 0)
    with pytest.raises(ValueError, match="strictly ascending"):
        A.Problem.from_csr(n, rp, bad)
    below = ci.copy()
    below[rp[5]] = 2  # row 5 starts with column 2
    with pytest.raises(ValueError, match=r"entry \(5, 2\) is below the diagonal"):
        A.Problem.from_csr(n, rp, below)
    out = ci.copy()
    out[-1] = n
    with pytest.raises(ValueError, match="out of range"):
        A.Problem.from_csr(n, rp, out)
    dec = rp.copy()
    dec[1] = -1  # row 0 ends before it begins
    with pytest.raises(ValueError, match="non-decreasing"):
        A.Problem.from_csr(n, dec, ci)
    with pytest.raises(ValueError, match="rows \\+ 1"):
        A.Problem.from_csr(n + 1, rp, ci)
    with pytest.raises(ValueError, match="unknown ordering"):
        A.Problem.from_csr(n, rp, ci, ordering=7)


def test_factor_budget_refusal():
    n, rp, ci = grid_pattern(6)
    q = A.Problem.from_csr(n, rp, ci)
    need = int(q.counts["predicted_nnz"])
    A.Problem.from_csr(n, rp, ci, max_factor_nnz=need)
    with pytest.raises(ValueError, match="over the budget"):
        A.Problem.from_csr(n, rp, ci, max_factor_nnz=need - 1)


def test_missing_diagonal_is_accepted_by_the_pattern_checks():
    """from_pattern does not require a stored diagonal (the factorization then
    meets a zero pivot, as the reference's would)."""
    n, rp, ci = grid_pattern(3)
    keep = np.ones(ci.size, dtype=bool)
    keep[rp[4]] = False  # drop (4, 4)
    rp2 = rp.copy()
    rp2[5:] -= 1
    q = A.Problem.from_csr(n, rp2, ci[keep])
    assert q.counts["nnz"] == ci.size - 1

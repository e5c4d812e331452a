"""The integer half of the path against the reference's own outputs (fixtures of
tests/golden, produced by the reference's sources): mesh numbering, edge signs,
Dirichlet mask, tree-cotree gauge, DOF map, upper-triangle pattern, ScatterMap,
AMD permutation, elimination tree, row and column counts -- bit for bit."""
import hashlib

import numpy as np
import pytest

from conftest import cube_config, golden

import aphi_amd as A

CASES = ["cube2", "cube3_offcentre", "config1", "cube8", "cube5_draw", "cube4_current", "config2", "config3"]
ARRAYS = ["edges", "tet_edges", "tet_edge_signs", "tree_edge", "dirichlet_edge", "edge_index", "phi_index",
          "row_ptr", "col_index", "perm", "parent", "row_count", "col_count", "scatter_live", "scatter_pos",
          "scatter_offset", "scatter_slot"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int64).tobytes()).hexdigest()


@pytest.fixture(scope="module", params=CASES)
def case(request):
    meta, arrays = golden(request.param)
    return meta, arrays, A.Problem(cube_config(meta))


def test_counts_match_reference(case):
    meta, _, p = case
    c = p.counts
    for k in ("nodes", "edges", "tets", "num_a", "num_phi", "num_total", "nnz", "predicted_nnz", "num_roots"):
        key = "num_total" if k == "num_total" else k
        assert c[key] == meta[k], k


@pytest.mark.parametrize("name", ARRAYS)
def test_integer_arrays_bit_exact(case, name):
    meta, arrays, p = case
    mine = p.ints(name)
    if name in arrays:
        np.testing.assert_array_equal(mine, arrays[name].astype(np.int64))
    assert sha(mine) == meta["sha256"][name], f"{name} differs from the reference"


def test_tree_is_spanning(case):
    """tree_cotree.hpp:125-129: interior tree edges = nodes - boundary nodes for one piece."""
    _, _, p = case
    tree = p.ints("tree_edge")
    dirichlet = p.ints("dirichlet_edge")
    edges = p.ints("edges").reshape(-1, 2)
    boundary_nodes = np.unique(edges[dirichlet == 1].ravel())
    interior_tree = int(((tree == 1) & (dirichlet == 0)).sum())
    assert interior_tree == p.counts["nodes"] - len(boundary_nodes)
    assert p.counts["tree_edges"] == interior_tree


def test_permutation_and_supernodes_consistent(case):
    _, _, p = case
    n = p.n
    perm, iperm = p.ints("perm"), p.ints("iperm")
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.array_equal(iperm[perm], np.arange(n))
    fperm = p.ints("factor_perm")
    assert np.array_equal(np.sort(fperm), np.arange(n))
    # the factorization order is a postorder of the same elimination tree:
    # column counts are the reference's, relabelled
    col_ref = p.ints("col_ref")
    assert np.array_equal(np.sort(col_ref), np.arange(n))
    assert np.array_equal(iperm[fperm], col_ref)
    first = p.ints("super_first")
    assert first[0] == 0 and first[-1] == n and np.all(np.diff(first) > 0)
    sp = p.ints("super_parent")
    assert np.all((sp == -1) | (sp > np.arange(len(sp))))
    # supernodal storage covers the exact factor
    assert p.counts["factor_entries"] >= p.counts["predicted_nnz"]


def test_supernode_rows_cover_reference_column_counts(case):
    """Each supernode's row list is its top column's structure: its length is at
    least the reference col_count of every column it holds (minus the rows above)."""
    _, _, p = case
    first, srp, col_ref = p.ints("super_first"), p.ints("super_row_ptr"), p.ints("col_ref")
    cc = p.ints("col_count")
    for s in range(len(first) - 1):
        m = srp[s + 1] - srp[s]
        for c in range(first[s], first[s + 1]):
            assert cc[col_ref[c]] <= m - (c - first[s])


def test_frequency_sweep_formula():
    """src/input_file.cpp:19-50: endpoints exact, interior exp(log f0 + i*step)."""
    f = A.frequency_sweep(1.0, 1e8, 64)
    assert f[0] == 1.0 and f[-1] == 1e8
    step = (np.log(1e8) - np.log(1.0)) / 63
    assert f[10] == np.exp(np.log(1.0) + step * 10)
    lin = A.frequency_sweep(10.0, 20.0, 3, logarithmic=False)
    assert list(lin) == [10.0, 15.0, 20.0]
    with pytest.raises(ValueError):
        A.frequency_sweep(0.0, 1.0, 4)

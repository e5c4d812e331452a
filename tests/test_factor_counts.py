"""Row and column counts of L from the near-linear skeleton/least-common-ancestor
method against their definition -- row i of L is the union of the elimination-tree
paths from the columns j < i of A's row i up to i -- walked directly in Python, on
generated cubes beyond the golden fixtures (which pin the counts bit for bit to the
reference's own on five meshes)."""
import numpy as np
import pytest

import aphi_amd as A


def direct_counts(p):
    n = p.n
    rp, ci = p.ints("row_ptr"), p.ints("col_index")      # upper triangle of A, original order
    iperm = p.ints("iperm")
    parent = p.ints("parent")
    rows = [[] for _ in range(n)]                         # earlier neighbours in permuted order
    for r in range(n):
        for c in ci[rp[r]:rp[r + 1]]:
            a, b = iperm[r], iperm[c]
            if a != b:
                rows[max(a, b)].append(min(a, b))
    rc = np.ones(n, dtype=np.int64)
    cc = np.ones(n, dtype=np.int64)
    mark = np.full(n, -1)
    for k in range(n):
        mark[k] = k
        for j in rows[k]:
            while j >= 0 and mark[j] != k:
                mark[j] = k
                rc[k] += 1
                cc[j] += 1
                j = parent[j]
    return rc, cc


@pytest.mark.parametrize("n,bar", [(4, (1, 3, 1, 2)), (6, (1, 3, 2, 5))])
def test_counts_match_definition(n, bar):
    p = A.Problem(A.CubeConfig(n=n, side=0.1, bar=bar, sigma=5.8e7))
    rc, cc = direct_counts(p)
    assert np.array_equal(p.ints("row_count"), rc)
    assert np.array_equal(p.ints("col_count"), cc)
    assert p.counts["predicted_nnz"] == int(rc.sum()) == int(cc.sum())

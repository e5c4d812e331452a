"""Host symbolic phase of the batched / serving workloads: many cube problems
built on several host threads (aphi_problem_create_cubes), each identical to a
single build of its spec, with refusals reported by index; and the seeded
problem draws of the batched workload."""
import time

import numpy as np
import pytest

import aphi_amd as A

ARRAYS = ["row_ptr", "col_index", "perm", "parent", "col_count", "scatter_slot", "super_first", "super_rows",
          "super_rel", "factor_perm", "slot_to_panel", "tree_edge", "edge_index"]


def test_parallel_build_equals_single_builds():
    configs, _ = A.batched_cube_configs(7, 12)
    t0 = time.perf_counter()
    many = A.Problem.cubes(configs, threads=4)
    t1 = time.perf_counter()
    singles = [A.Problem(c) for c in configs]
    t2 = time.perf_counter()
    print(f"12 problems: 4 threads {t1 - t0:.2f} s, one by one {t2 - t1:.2f} s")
    for p, q in zip(many, singles):
        timing = ("analyze_ms", "build_ms")
        assert {k: v for k, v in p.counts.items() if k not in timing} == {k: v for k, v in q.counts.items() if k not in timing}
        for name in ARRAYS:
            assert np.array_equal(p.ints(name), q.ints(name)), name


def test_parallel_build_reports_the_first_refusal():
    configs, _ = A.batched_cube_configs(7, 6)
    configs[4].side = -1.0
    configs[2].ordering = 9
    with pytest.raises(ValueError, match="problem 2: .*unknown ordering"):
        A.Problem.cubes(configs, threads=3)


def test_batched_draws_follow_the_stated_ranges():
    configs, freqs = A.batched_cube_configs(20261203, 256)
    again, freqs2 = A.batched_cube_configs(20261203, 256)
    assert configs == again and np.array_equal(freqs, freqs2)
    ns = np.array([c.n for c in configs])
    assert ns.min() == 4 and ns.max() == 12
    for c in configs:
        i0, i1, j0, j1 = c.bar
        assert 0.05 <= c.side <= 0.2 and 1e6 <= c.sigma <= 6e7
        assert 1 <= i1 - i0 <= 4 and 1 <= j1 - j0 <= 4 and 0 <= i0 and i1 <= c.n and 0 <= j0 and j1 <= c.n
    assert freqs.min() >= 1.0 and freqs.max() <= 1e6
    # log-uniform: log10 f roughly uniform over [0, 6]
    assert abs(np.log10(freqs).mean() - 3.0) < 0.4

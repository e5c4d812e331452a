"""The factorization replayed from its captured HIP graph, and the update-tile
autotuner: neither may change a bit of the factor.

* Every replay of the captured factorization, with eager launches, other
  precisions and synchronisations between the replays, gives the factor (D) and
  the solutions of the launch-by-launch factorization bit for bit, in both
  precisions.
* The tile table (per launch class: default rule, 64 x 64 or 128 x 128 tiles)
  changes only speed: all-small, all-large and tuned tables give the same D; the
  tuner returns a table of valid entries and two positive timings; bad tables
  are refused. Switching the tiling mode between replays keeps the bits too.
"""
import numpy as np
import pytest

from conftest import cube_config, golden

import aphi_amd as A

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


@pytest.fixture(scope="module", params=["config1", "cube8"])
def mesh(request, gpu_context):
    meta, _ = golden(request.param)
    p = A.Problem(cube_config(meta))
    m = A.Mesh(gpu_context, p)
    m.refill(meta["omega"])
    m.equilibrate(10)
    yield p, m, meta
    m.close()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("precision", [A.APHI_FP32, A.APHI_FP64])
def test_factor_graph_replays_match_eager(mesh, precision, monkeypatch):
    import torch
    p, m, meta = mesh
    g = np.random.default_rng(11)
    B = dev(g.standard_normal((4, p.n)) + 1j * g.standard_normal((4, p.n)))
    monkeypatch.setenv("APHI_FACTOR_GRAPHS", "0")
    st = m.factorize_ldlt(precision)
    assert st["ok"]
    D0, X0 = m.factor_diagonal(), m.solve(B).cpu().numpy()
    monkeypatch.delenv("APHI_FACTOR_GRAPHS")
    other = A.APHI_FP64 if precision == A.APHI_FP32 else A.APHI_FP32
    for rep in range(4):
        st = m.factorize_ldlt(precision)
        torch.cuda.synchronize()
        assert st["ok"]
        assert np.array_equal(bits(m.factor_diagonal()), bits(D0)), rep
        assert np.array_equal(bits(m.solve(B).cpu().numpy()), bits(X0)), rep
        if rep == 1:
            m.factorize_ldlt(other)           # the other precision in between
        if rep == 2:
            monkeypatch.setenv("APHI_FACTOR_GRAPHS", "0")
            m.factorize_ldlt(precision)       # an eager factorization in between
            monkeypatch.delenv("APHI_FACTOR_GRAPHS")


def test_tile_tables_keep_the_bits(mesh):
    p, m, meta = mesh
    m.set_update_tile_table(np.zeros((8, 8), dtype=np.int8))
    assert m.factorize_ldlt(A.APHI_FP32)["ok"]
    D0 = m.factor_diagonal()
    for fill in (1, 2):
        m.set_update_tile_table(np.full((8, 8), fill, dtype=np.int8))
        assert np.array_equal(m.update_tile_table(), np.full((8, 8), fill, dtype=np.int8))
        assert m.factorize_ldlt(A.APHI_FP32)["ok"]
        assert np.array_equal(bits(m.factor_diagonal()), bits(D0)), fill
    table, t_default, t_tuned = m.tune_update_tiles(repeats=1)
    print(f"tuned table (rows: height class, columns: width class):\n{table}\n"
          f"default {t_default:.3f} ms, tuned {t_tuned:.3f} ms", flush=True)
    assert set(np.unique(table)) <= {0, 1, 2}
    assert t_default > 0 and t_tuned > 0
    assert np.array_equal(m.update_tile_table(), table)
    assert m.factorize_ldlt(A.APHI_FP32)["ok"]
    assert np.array_equal(bits(m.factor_diagonal()), bits(D0))
    with pytest.raises(ValueError):
        m.set_update_tile_table(np.full((8, 8), 3, dtype=np.int8))
    m.set_update_tile_table(np.zeros((8, 8), dtype=np.int8))


def test_tile_modes_between_graph_replays(mesh):
    # switching the tiling mode between replays of the captured factorization
    # recaptures it (the mode is part of what the graph holds) and keeps the bits
    p, m, meta = mesh
    m.set_update_tiles("auto")
    assert m.factorize_ldlt(A.APHI_FP32)["ok"]
    D0 = m.factor_diagonal()
    for mode in ("small", "large", "small", "auto"):
        m.set_update_tiles(mode)
        for _ in range(2):
            assert m.factorize_ldlt(A.APHI_FP32)["ok"]
            assert np.array_equal(bits(m.factor_diagonal()), bits(D0)), mode
    with pytest.raises(ValueError):
        m.set_update_tiles("medium")

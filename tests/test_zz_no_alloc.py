"""aphi_terminal_current_noalloc (the entry point that allocates no device
memory: its scratch is an array compiled into the library) through the C-ABI:
bitwise the results of aphi_terminal_current, on two calls in a row -- a larger
mesh with several right-hand sides, then a smaller mesh with fewer, so stale
scratch contents from the first call would show -- and a clear refusal when the
terms exceed the scratch."""
import ctypes as C

import numpy as np
import pytest

from conftest import cube_config, golden

import aphi_amd as A
from aphi_amd import _capi

pytestmark = pytest.mark.gpu


def _call(fn, mesh, x, omega, length):
    k = x.shape[0]
    cur, res, ind = np.zeros(2 * k), np.zeros(k), np.zeros(k)
    ax = np.array([0.0, 0.0, 1.0])
    st = fn(mesh._h, k, C.c_void_p(x.data_ptr()), mesh.n, float(omega), A.APHI_ROW_SCALED,
            ax.ctypes.data_as(_capi.DP), float(length), 1.0, cur.ctypes.data_as(_capi.DP),
            res.ctypes.data_as(_capi.DP), ind.ctypes.data_as(_capi.DP))
    return st, cur, res, ind


def _solutions(ctx, mesh, omega, k):
    import torch
    mesh.refill(omega)
    b = mesh.rhs_device()
    B = torch.stack([b * (1 + 0.25 * r) for r in range(k)])
    x, rep = mesh.solve_symmetric(B, precision=A.APHI_FP64)
    assert rep["ok"]
    return x


def test_noalloc_equals_allocating_entry_point(gpu_context):
    lib = A.lib()
    cases = [("cube8", 6), ("config1", 2)]  # the second call is smaller on both counts
    for name, k in cases:
        meta, _ = golden(name)
        p = A.Problem(cube_config(meta))
        m = A.Mesh(gpu_context, p)
        try:
            print(f"{name}: {p.counts['tets']} tets, {k} right-hand sides", flush=True)
            x = _solutions(gpu_context, m, meta["omega"], k)
            st0, c0, r0, l0 = _call(lib.aphi_terminal_current, m, x, meta["omega"], meta["side"])
            st1, c1, r1, l1 = _call(lib.aphi_terminal_current_noalloc, m, x, meta["omega"], meta["side"])
            assert st0 == 0 and st1 == 0, lib.aphi_last_error()
            for a, b in ((c0, c1), (r0, r1), (l0, l1)):
                assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
            # and the reference's value on its own solution (port right-hand side)
            assert complex(c1[0], c1[1]) != 0
        finally:
            m.close()


def test_noalloc_refuses_terms_beyond_the_scratch(gpu_context):
    """24^3 hexes = 82,944 tets: 26 right-hand sides need 2,156,544 terms, more
    than the 2,097,152 the scratch holds; 25 fit (2,073,600 terms) and give the
    allocating entry point's bits (x = 0 still carries current: the prescribed
    port potentials drive it)."""
    import torch
    lib = A.lib()
    meta, _ = golden("config2")
    p = A.Problem(cube_config(meta))
    m = A.Mesh(gpu_context, p)
    try:
        x = torch.zeros(26, p.n, dtype=torch.complex128, device="cuda")
        st, *_ = _call(lib.aphi_terminal_current_noalloc, m, x, 2 * np.pi * 10.0, 0.1)
        assert st == _capi.APHI_INVALID_ARGUMENT
        assert "exceeds the scratch" in lib.aphi_last_error().decode()
        x[:, ::7] = 1e-9 * torch.arange(26, dtype=torch.float64, device="cuda")[:, None]
        y = x[:25].contiguous()
        st0, c0, r0, l0 = _call(lib.aphi_terminal_current, m, y, 2 * np.pi * 10.0, 0.1)
        st1, c1, r1, l1 = _call(lib.aphi_terminal_current_noalloc, m, y, 2 * np.pi * 10.0, 0.1)
        assert st0 == 0 and st1 == 0, lib.aphi_last_error()
        for a, b in ((c0, c1), (r0, r1), (l0, l1)):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
        assert len(set(c1[0::2])) > 1
    finally:
        m.close()

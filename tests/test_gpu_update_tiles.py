"""The complex-float trailing update's two tilings (64 x 64 and 128 x 128 tiles)
accumulate every entry's products in the same order, so they give bitwise the
same factor: the solutions of one block of right-hand sides are compared bit for
bit across the forced tilings and the default choice, on a mesh whose fronts
span several tiles of either size."""
import numpy as np
import pytest

import aphi_amd as A

pytestmark = pytest.mark.gpu


def _solve(ctx, p, mode, b):
    import torch
    m = A.Mesh(ctx, p)
    try:
        m.set_update_tiles(mode)
        m.refill(2 * np.pi * 1e4)
        m.equilibrate(10)
        st = m.factorize_ldlt(A.APHI_FP32)
        assert st["ok"]
        x = torch.empty_like(b)
        m.solve(b, x)
        ctx.synchronize()
        return x.cpu().numpy()
    finally:
        m.close()


def test_update_tilings_bitwise_equal(gpu_context):
    import torch
    p = A.Problem(A.CubeConfig(n=12, side=0.1, bar=(2, 8, 8, 11), sigma=5.8e7))
    g = torch.Generator().manual_seed(7)
    b = torch.complex(torch.randn(4, p.n, dtype=torch.float64, generator=g),
                      torch.randn(4, p.n, dtype=torch.float64, generator=g)).cuda()
    xs = {mode: _solve(gpu_context, p, mode, b) for mode in ("small", "large", "auto")}
    assert np.all(np.isfinite(xs["small"]))
    assert np.array_equal(xs["small"].view(np.uint64), xs["large"].view(np.uint64))
    assert np.array_equal(xs["small"].view(np.uint64), xs["auto"].view(np.uint64))


def test_update_tiles_refuses_unknown_mode(gpu_context):
    p = A.Problem(A.CubeConfig(n=3, side=0.1, bar=(1, 2, 1, 2), sigma=5.8e7))
    m = A.Mesh(gpu_context, p)
    try:
        with pytest.raises(ValueError, match="update tiles"):
            m.set_update_tiles("medium")
    finally:
        m.close()

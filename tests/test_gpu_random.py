"""The counter-based generator on the device: the same bits as the host function,
and the complex right-hand-side layout the benchmark draws."""
import numpy as np
import pytest

import aphi_amd as A

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("offset,count", [(0, 1), (3, 1001), (2 ** 33 + 1, 70000)])
def test_device_normals_equal_host_bits(gpu_context, offset, count):
    """The generator's device kernel and host function give the same bits."""
    d = A.random_normal(gpu_context, 20261203, (count,), offset=offset).cpu().numpy()
    h = A.random_normal_host(20261203, count, offset=offset)
    assert np.array_equal(d.view(np.uint64), h.view(np.uint64))


def test_complex_rhs_layout(gpu_context):
    B = A.random_complex_rhs(gpu_context, 5, 3, 11).cpu().numpy()
    h = A.random_normal_host(5, 66)
    assert np.array_equal(B.real.reshape(-1), h[0::2]) and np.array_equal(B.imag.reshape(-1), h[1::2])

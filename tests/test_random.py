"""The counter-based generator behind every seeded input: Philox-4x32-10 against
the known-answer vectors published with the algorithm (Random123's kat_vectors),
the Box-Muller transform against numpy's log / sqrt / cos / sin on the same
uniforms, the distribution (moments, Kolmogorov-Smirnov), and counter
addressing (a draw at an offset is the slice of a longer draw)."""
import numpy as np
import pytest

import aphi_amd as A

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert A.philox4x32_10(ctr, key) == out


def test_box_muller_against_numpy():
    seed, nb = 987654321012345, 3000
    ref = []
    for b in range(nb):
        w = A.philox4x32_10([b & 0xffffffff, b >> 32, 0, 0], [seed & 0xffffffff, seed >> 32])
        a = ((w[1] << 32) | w[0]) >> 12
        c = ((w[3] << 32) | w[2]) >> 11
        u1, u2 = (a + 0.5) * 2.0 ** -52, c * 2.0 ** -53
        r = np.sqrt(-2.0 * np.log(u1))
        ref += [r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)]
    z = A.random_normal_host(seed, 2 * nb)
    assert np.abs(z - np.array(ref)).max() < 1e-14


def test_distribution():
    from scipy import stats
    n = 4_000_000
    z = A.random_normal_host(20261203, n)
    se = 1.0 / np.sqrt(n)
    assert abs(z.mean()) < 5 * se
    assert abs(z.var() - 1.0) < 5 * np.sqrt(2.0) * se
    assert abs(stats.skew(z)) < 5 * np.sqrt(6.0) * se
    assert abs(stats.kurtosis(z)) < 5 * np.sqrt(24.0) * se
    assert stats.kstest(z[:500_000], "norm").pvalue > 1e-3
    # different seeds are different streams
    w = A.random_normal_host(20261204, 1000)
    assert abs(np.corrcoef(z[:1000], w)[0, 1]) < 0.15


@pytest.mark.parametrize("offset", [0, 1, 7, 2 ** 33 + 3])
def test_offsets_address_the_same_stream(offset):
    long = A.random_normal_host(42, 64, offset=offset)
    for start in (0, 1, 5, 30):
        part = A.random_normal_host(42, 17, offset=offset + start)
        assert np.array_equal(part, long[start:start + 17])

"""The oracle's generator of layered Gaussian velocity models (a library
extension with no reference counterpart), pinned against an independent
integer restatement of its hash, against scipy.ndimage.gaussian_filter on
the same noise, and against the statistics the construction fixes."""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle

M64 = (1 << 64) - 1


def _mix64(x):
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def _uniform(seed, k):
    z = _mix64((seed + (k + 1) * 0x9E3779B97F4A7C15) & M64)
    return ((z >> 11) + 0.5) / 2.0**53


@pytest.mark.parametrize("seed", [0, 1, 12345, (1 << 64) - 1])
def test_noise_matches_an_integer_restatement_of_the_hash(seed):
    n = 64
    got = oracle.gaussian_noise(n, seed)
    want = [math.sqrt(-2.0 * math.log(_uniform(seed, 2 * i)))
            * math.cos(2.0 * math.pi * _uniform(seed, 2 * i + 1))
            for i in range(n)]
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)


def test_splitmix_known_answer():
    """splitmix64's published first output for state 0 (the generator's
    increment added once, then the finaliser)."""
    assert _mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF


def test_noise_statistics():
    x = oracle.gaussian_noise(400_000, 7)
    assert abs(x.mean()) < 0.01
    assert abs(x.var() - 1.0) < 0.01
    # fourth moment of a unit normal is 3
    assert abs((x**4).mean() - 3.0) < 0.05
    assert abs(np.corrcoef(x[:-1], x[1:])[0, 1]) < 0.01
    assert not np.array_equal(x[:100], oracle.gaussian_noise(100, 8))


@pytest.mark.parametrize("dims,d,l", [
    ((23, 17, 11), (10.0, 10.0, 10.0), 60.0),
    ((20, 9, 13), (5.0, 12.0, 7.0), 40.0),
    ((16, 8, 4), (2.0, 3.0, 1.0), 30.0),      # radius > n on every axis
    ((12, 10, 9), (10.0, 10.0, 10.0), 4.0),   # sigma < 1/8: no filter
])
def test_filter_matches_scipy_gaussian_filter(dims, d, l):
    nx, ny, nz = dims
    g = oracle.geometry3d(nx, ny, nz, *d)
    noise = oracle.gaussian_noise(nx * ny * nz, 3)
    got = oracle.gaussian_filter3d(g, l, noise).reshape(nz, ny, nx)
    sig = [l / (2 * d[2]), l / (2 * d[1]), l / (2 * d[0])]
    sig = [s if int(4 * s + 0.5) >= 1 else 0.0 for s in sig]
    want = ndimage.gaussian_filter(noise.reshape(nz, ny, nx), sig,
                                   mode="reflect", truncate=4.0)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def test_layers_and_normalisation():
    g = oracle.geometry3d(18, 14, 20, 10.0)
    lv = np.array([3000.0, 4000.0, 5500.0])
    flat = oracle.layered_gaussian3d(g, lv, 0.0, 50.0, 9)
    layer = (np.arange(20) * 3) // 20
    np.testing.assert_array_equal(flat.reshape(20, 14, 18),
                                  np.broadcast_to(lv[layer][:, None, None],
                                                  (20, 14, 18)))
    amp = 0.05
    v = oracle.layered_gaussian3d(g, lv, amp, 50.0, 9).reshape(20, 14, 18)
    G = (v / lv[layer][:, None, None] - 1.0) / amp
    assert abs(G.mean()) < 1e-12
    assert abs(G.std() - 1.0) < 1e-12
    # G is the filtered noise, standardised
    f = oracle.gaussian_filter3d(g, 50.0, oracle.gaussian_noise(g.nx * g.ny
                                                                * g.nz, 9))
    np.testing.assert_allclose(G.reshape(-1), (f - f.mean()) / f.std(),
                               rtol=0, atol=1e-11)


def test_correlation_length():
    """Filtered white noise with sigma = l / (2 d) nodes has correlation
    exp(-r^2 / l^2): about 1/e at lag l, about exp(-1/4) at lag l / 2."""
    g = oracle.geometry3d(160, 160, 40, 1.0)
    l = 8.0
    v = oracle.layered_gaussian3d(g, [1.0], 0.1, l, 5).reshape(40, 160, 160)
    G = (v - 1.0) / 0.1
    for lag, want in ((8, math.exp(-1.0)), (4, math.exp(-0.25))):
        r = np.mean(G[:, :, :-lag] * G[:, :, lag:])
        assert abs(r - want) < 0.05, (lag, r, want)
        r = np.mean(G[:, :-lag, :] * G[:, lag:, :])
        assert abs(r - want) < 0.05, (lag, r, want)

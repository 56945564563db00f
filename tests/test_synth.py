"""Synthetic inputs: the Gaussian smoothing restated from the reference and
the ellipsoid phantom (CPU only)."""
import numpy as np
import pytest

from claire_amd import synth


@pytest.mark.parametrize("k,sigma", [(3, 1.0), (7, 2.5), (16, 1.0)])
def test_smoothing_of_a_fourier_mode(k, sigma):
    """exp(−½·k²·(σh)²) on each mode (src/Preprocessing.cpp:2017-2025), the
    Nyquist mode (k = n/2) included."""
    n = 32
    x1 = synth.coords((n, n, n), np.float64)[0]
    f = np.broadcast_to(np.cos(k * x1) + 0.5 * np.sin(2 * x1), (n, n, n))
    h = 2 * np.pi / n
    want = (np.exp(-0.5 * (k * sigma * h) ** 2) * np.cos(k * x1)
            + 0.5 * np.exp(-0.5 * (2 * sigma * h) ** 2) * np.sin(2 * x1))
    assert np.abs(synth.gaussian_smoothing(f, sigma) - want).max() < 1e-12


def test_smoothing_of_a_point_is_a_gaussian():
    """A unit impulse smooths to the sampled Gaussian of σ voxels (unit sum),
    up to the periodic images and the sampling of a well-resolved σ."""
    n, sigma = 48, 3.0
    f = np.zeros((n, n, n))
    f[0, 0, 0] = 1.0
    g = synth.gaussian_smoothing(f, sigma)
    i = np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64)
    r2 = i[:, None, None] ** 2 + i[None, :, None] ** 2 + i[None, None, :] ** 2
    want = np.exp(-0.5 * r2 / sigma**2) / ((2 * np.pi) ** 1.5 * sigma**3)
    assert abs(g.sum() - 1.0) < 1e-12
    assert np.abs(g - want).max() < 1e-6 * want.max()


def test_phantom_range_determinism_and_volume():
    a = synth.ellipsoid_phantom((64, 64, 64), count=10, seed=5)
    b = synth.ellipsoid_phantom((64, 64, 64), count=10, seed=5)
    assert np.array_equal(a, b) and a.dtype == np.float32
    assert a.min() == 0.0 and a.max() == 1.0
    assert not np.array_equal(a, synth.ellipsoid_phantom((64, 64, 64), count=10, seed=6))
    # one unsmoothed ellipsoid: its voxel count is its volume 4/3·π·r1·r2·r3
    rng = np.random.default_rng(9)
    rng.uniform(0.0, 1.0, 3)
    r = rng.uniform(8.0, 14.0, 3)
    one = synth.ellipsoid_phantom((64, 64, 64), count=1, seed=9, radii=(8.0, 14.0), sigma=0.0)
    vol = 4.0 / 3.0 * np.pi * np.prod(r)
    assert set(np.unique(np.round(one, 6))) <= {0.0, 1.0}
    assert abs((one > 0.5).sum() - vol) < 0.03 * vol


def test_phantom_wider_than_the_grid_wraps():
    a = synth.ellipsoid_phantom((16, 16, 16), count=3, seed=2, radii=(6.0, 12.0))
    assert np.isfinite(a).all() and a.max() == 1.0

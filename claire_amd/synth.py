"""Synthetic inputs for the registration hot path (seeded, numpy only).

Grids are periodic on [0, 2π)^3 with h = 2π/n and nodal coordinates
x_i = h·i (src/RegOpt.cpp:935, apps/benchmark.cpp:413-416).
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi


def coords(nx, dtype=np.float32):
    """Nodal coordinates x1, x2, x3 (broadcastable), computed as hx·i in `dtype`."""
    out = []
    for a, n in enumerate(nx):
        h = dtype(np.pi * 2.0 / float(dtype(n)))
        x = (h * np.arange(n, dtype=dtype)).astype(dtype)
        shape = [1, 1, 1]
        shape[a] = n
        out.append(x.reshape(shape))
    return out


def benchmark_template(nx, dtype=np.float32):
    """m = (sin²x1 + sin²x2 + sin²x3)/3 (apps/benchmark.cpp:384-432, src/CLAIREBase.cpp:1123-1126)."""
    x1, x2, x3 = coords(nx, dtype)
    s1, s2, s3 = np.sin(x1), np.sin(x2), np.sin(x3)
    m = (s1 * s1 + s2 * s2 + s3 * s3) / dtype(3.0)
    return np.ascontiguousarray(np.broadcast_to(m, tuple(nx)), dtype=dtype)


def benchmark_velocity(nx, dtype=np.float32, v0=1.0):
    """v = (sin x3·cos x2·sin x2, sin x1·cos x3·sin x3, sin x2·cos x1·sin x1)
    (vcase 1 of apps/benchmark.cpp:484-489; with v0 this is synthetic problem 4,
    src/CLAIREBase.cpp:1094-1095,1132-1135)."""
    x1, x2, x3 = coords(nx, dtype)
    v = np.empty((3,) + tuple(nx), dtype=dtype)
    v[0] = dtype(v0) * np.sin(x3) * np.cos(x2) * np.sin(x2)
    v[1] = dtype(v0) * np.sin(x1) * np.cos(x3) * np.sin(x3)
    v[2] = dtype(v0) * np.sin(x2) * np.cos(x1) * np.sin(x1)
    return v


def cfl_number(v, nt):
    """max_x Σ_i |v_i|/h_i / nt (CLAIREBase::ComputeCFLCondition, src/CLAIREBase.cpp:1330-1351)."""
    nx = v.shape[1:]
    s = np.zeros(nx, dtype=np.float64)
    for a in range(3):
        s += np.abs(v[a].astype(np.float64)) / (TWO_PI / nx[a])
    return float(s.max()) / nt


def bandlimited_velocity(nx, kmax=8, cfl=2.0, nt=4, seed=0, dtype=np.float32):
    """Band-limited random velocity: N(0,1) Fourier coefficients for
    |w|_∞ <= kmax decayed by 1/(1+|w|²), Hermitian by construction (irfftn of
    a half spectrum), scaled to the given CFL number."""
    rng = np.random.default_rng(seed)
    n1, n2, n3 = nx
    w1 = np.fft.fftfreq(n1, 1.0 / n1).reshape(-1, 1, 1)
    w2 = np.fft.fftfreq(n2, 1.0 / n2).reshape(1, -1, 1)
    w3 = np.arange(n3 // 2 + 1).reshape(1, 1, -1)
    mask = (np.abs(w1) <= kmax) & (np.abs(w2) <= kmax) & (np.abs(w3) <= kmax)
    decay = 1.0 / (1.0 + w1 * w1 + w2 * w2 + w3 * w3)
    v = np.empty((3,) + tuple(nx), dtype=np.float64)
    for a in range(3):
        coef = (rng.standard_normal(mask.shape) + 1j * rng.standard_normal(mask.shape)) * decay * mask
        v[a] = np.fft.irfftn(coef, s=nx, axes=(0, 1, 2)) * (n1 * n2 * n3)
    c = cfl_number(v, nt)
    if c > 0:
        v *= cfl / c
    return v.astype(dtype)


def _periodic_gauss_1d(x, c, sigma):
    d = np.mod(x - c + np.pi, TWO_PI) - np.pi
    return np.exp(-0.5 * (d / sigma) ** 2)


def gaussian_blobs(nx, count=24, seed=1, dtype=np.float32):
    """Sum of `count` periodic Gaussian blobs: centres U[0,2π)^3, σ U[2,8]·h,
    amplitude U[0,1] (separable evaluation)."""
    rng = np.random.default_rng(seed)
    h = TWO_PI / np.asarray(nx, dtype=np.float64)
    xs = [np.arange(n) * h[a] for a, n in enumerate(nx)]
    f = np.zeros(tuple(nx), dtype=np.float64)
    for _ in range(count):
        c = rng.uniform(0.0, TWO_PI, 3)
        s = rng.uniform(2.0, 8.0) * h
        amp = rng.uniform(0.0, 1.0)
        g1 = _periodic_gauss_1d(xs[0], c[0], s[0])
        g2 = _periodic_gauss_1d(xs[1], c[1], s[1])
        g3 = _periodic_gauss_1d(xs[2], c[2], s[2])
        f += amp * (g1[:, None, None] * g2[None, :, None]) * g3[None, None, :]
    return f.astype(dtype)


def random_field(nx, seed=0, smooth=4, dtype=np.float32):
    """Smooth random scalar field (band-limited to |w|_∞ <= smooth)."""
    v = bandlimited_velocity(nx, kmax=smooth, cfl=1.0, nt=1, seed=seed, dtype=np.float64)[0]
    v /= np.abs(v).max()
    return v.astype(dtype)


def rescale(x, xminout=0.0, xmaxout=1.0):
    """Intensity rescaling of one image, restated from Rescale with nc = 1
    (src/CLAIREUtils.cpp:691-720): shift by −min only when min < −1e-6, then
    scale by xmaxout/(max − min) when max − min > 1e-6."""
    x = np.asarray(x)
    xmin, xmax = float(x.min()), float(x.max())
    y = x - xmin if -xmin > 1e-6 else x.copy()
    scale = xmaxout / (xmax - xmin) if abs(xmax - xmin) > 1e-6 else 1.0
    return (y * scale).astype(x.dtype)


def smooth_template_problem(nx, dtype=np.float32):
    """The reference's synthetic problem 4 (CLAIREBase::SetupSyntheticProb,
    src/CLAIREBase.cpp:1094-1095, 1123-1139, 1209): template
    mT = (sin²x1 + sin²x2 + sin²x3)/3 rescaled to [0, 1] and the velocity
    v = (sin x3·cos x2·sin x2, sin x1·cos x3·sin x3, sin x2·cos x1·sin x1).
    The reference image is mT transported by v (nt steps), rescaled."""
    x1, x2, x3 = coords(nx, np.float64)
    shape = tuple(nx)
    mT = np.broadcast_to((np.sin(x1) ** 2 + np.sin(x2) ** 2 + np.sin(x3) ** 2) / 3.0, shape)
    v = np.stack([np.broadcast_to(np.sin(x3) * np.cos(x2) * np.sin(x2), shape),
                  np.broadcast_to(np.sin(x1) * np.cos(x3) * np.sin(x3), shape),
                  np.broadcast_to(np.sin(x2) * np.cos(x1) * np.sin(x1), shape)])
    return rescale(mT.astype(dtype)), v.astype(dtype)


def gaussian_smoothing(x, sigma):
    """Preprocessing::GaussianSmoothing (src/Preprocessing.cpp:1956-2050):
    multiply the spectrum by exp(−½·Σ_i k_i²·(σ_i·h_i)²), σ in grid points,
    wavenumbers k > n/2 shifted to k − n (the Nyquist index keeps +n/2)."""
    x = np.asarray(x, dtype=np.float64)
    nx = x.shape
    sig = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (3,))
    c = [(sig[a] * TWO_PI / nx[a]) ** 2 for a in range(3)]
    k = []
    for a, n in enumerate(nx):
        kk = np.arange(n // 2 + 1 if a == 2 else n, dtype=np.float64)
        if a < 2:
            kk = np.where(kk > n // 2, kk - n, kk)
        shape = [1, 1, 1]
        shape[a] = kk.size
        k.append(kk.reshape(shape))
    mult = np.exp(-0.5 * (k[0] ** 2 * c[0] + k[1] ** 2 * c[1] + k[2] ** 2 * c[2]))
    return np.fft.irfftn(np.fft.rfftn(x) * mult, s=nx, axes=(0, 1, 2))


def _rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def ellipsoid_phantom(nx, count=30, seed=0, radii=(4.0, 20.0), sigma=1.0, dtype=np.float32):
    """Synthetic phantom: `count` solid ellipsoids painted in order (later ones
    overwrite earlier ones) on the periodic grid, each with centre uniform
    over the grid, semi-axes U[radii] voxels, a uniformly random rotation and
    intensity U[0,1]; then Gaussian smoothing with `sigma` voxels
    (gaussian_smoothing) and min-max normalization to [0, 1]."""
    rng = np.random.default_rng(seed)
    nx = tuple(int(n) for n in nx)
    f = np.zeros(nx, dtype=np.float64)
    for _ in range(count):
        c = rng.uniform(0.0, 1.0, 3) * np.asarray(nx)
        r = rng.uniform(radii[0], radii[1], 3)
        R = _rotation(rng)
        val = rng.uniform(0.0, 1.0)
        w = int(np.ceil(r.max())) + 1
        ci = np.floor(c).astype(int)
        off, disp = [], []
        for a in range(3):
            if 2 * w + 1 >= nx[a]:  # box wider than the grid: every index, nearest periodic image
                o = np.arange(nx[a])
                disp.append(np.mod(o - c[a] + 0.5 * nx[a], nx[a]) - 0.5 * nx[a])
            else:
                o = np.arange(ci[a] - w, ci[a] + w + 1)
                disp.append(o - c[a])
            off.append(o)
        d = np.stack(np.meshgrid(*disp, indexing="ij"), -1)
        q = d @ R  # coordinates in the ellipsoid's frame
        inside = ((q / r) ** 2).sum(-1) <= 1.0
        idx = np.ix_(*[np.mod(off[a], nx[a]) for a in range(3)])
        blk = f[idx]
        blk[inside] = val
        f[idx] = blk
    g = gaussian_smoothing(f, sigma)
    lo, hi = g.min(), g.max()
    return ((g - lo) / (hi - lo) if hi > lo else g - lo).astype(dtype)

"""Seeded Gaussian random fields on structured node lattices.

``g`` has zero mean and squared-exponential covariance
``C(r) = sigma^2 exp(-r^2 / (2 l^2))``.  Sampling uses circulant embedding
(Dietrich & Newsam, SIAM J. Sci. Comput. 18, 1997): the lattice is padded to
a periodic box at least ``pad_factor * l`` larger on every side, the
covariance is sampled with periodic distances, its DFT gives the circulant
eigenvalues (negatives from truncation are clipped), and
``g = Re(IFFT(sqrt(lambda) * FFT(w)))`` for white noise ``w`` drawn from
``numpy.random.default_rng(seed)``.  The covariance of ``g`` restricted to
the original lattice is then exactly ``C`` up to the clipped modes.
"""
from __future__ import annotations

import math

import numpy as np


def grf_lattice(shape, spacing: float, length: float, sigma: float = 1.0,
                seed: int = 0, pad_factor: float = 4.0) -> np.ndarray:
    """Sample ``g`` on a lattice of ``shape`` points with uniform ``spacing``.

    Returns a float64 array of ``shape`` (axis order as given).
    """
    shape = tuple(int(s) for s in shape)
    pad = int(math.ceil(pad_factor * length / spacing))
    big = tuple(s + pad for s in shape)
    axes = []
    for m in big:
        k = np.arange(m)
        axes.append(np.minimum(k, m - k) * spacing)
    r2 = np.zeros(big, dtype=np.float64)
    for ax, d in enumerate(axes):
        sh = [1] * len(big)
        sh[ax] = -1
        r2 = r2 + (d ** 2).reshape(sh)
    cov = sigma ** 2 * np.exp(-r2 / (2.0 * length ** 2))
    lam = np.fft.rfftn(cov).real
    np.maximum(lam, 0.0, out=lam)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(big)
    g = np.fft.irfftn(np.sqrt(lam) * np.fft.rfftn(w), s=big, axes=tuple(range(len(big))))
    sl = tuple(slice(0, s) for s in shape)
    return np.ascontiguousarray(g[sl])


def lognormal_kappa_cube(n: int, length: float = 0.1, sigma: float = 1.0, seed: int = 0) -> np.ndarray:
    """Nodal ``kappa = exp(g)`` on the ``(n+1)^3`` Kuhn lattice, x-fastest order."""
    g = grf_lattice((n + 1, n + 1, n + 1), 1.0 / n, length, sigma, seed)  # axes (z, y, x)
    return np.exp(g).ravel()


def lognormal_kappa_points(points: np.ndarray, length: float, sigma: float = 1.0,
                           seed: int = 0, grid: int = 128) -> np.ndarray:
    """``kappa = exp(g)`` at arbitrary 2D/3D points in the unit box.

    ``g`` is sampled on a ``(grid+1)^d`` lattice and multilinearly
    interpolated to the points (used for jittered meshes).
    """
    pts = np.asarray(points, dtype=np.float64)
    d = pts.shape[1]
    g = grf_lattice((grid + 1,) * d, 1.0 / grid, length, sigma, seed)
    # g axes are (.., y, x): reverse so that g[x, y, (z)]
    g = np.transpose(g, tuple(range(d - 1, -1, -1)))
    t = np.clip(pts * grid, 0.0, grid - 1e-9)
    i0 = np.floor(t).astype(np.int64)
    f = t - i0
    out = np.zeros(pts.shape[0])
    for corner in range(1 << d):
        idx = []
        wgt = np.ones(pts.shape[0])
        for ax in range(d):
            bit = (corner >> ax) & 1
            idx.append(i0[:, ax] + bit)
            wgt = wgt * (f[:, ax] if bit else 1.0 - f[:, ax])
        out += wgt * g[tuple(idx)]
    return np.exp(out)


def grf_lattice_device(shape, spacing: float, length: float, sigma: float, seed: int,
                       device, pad_factor: float = 4.0):
    """:func:`grf_lattice` on the GPU (torch FFT), white noise from a seeded
    torch generator on ``device``.  Returns a float64 tensor of ``shape``."""
    import torch
    shape = tuple(int(s) for s in shape)
    pad = int(math.ceil(pad_factor * length / spacing))
    big = tuple(s + pad for s in shape)
    r2 = torch.zeros(big, dtype=torch.float64, device=device)
    for ax, m in enumerate(big):
        k = torch.arange(m, device=device, dtype=torch.float64)
        d = torch.minimum(k, m - k) * spacing
        sh = [1] * len(big)
        sh[ax] = -1
        r2 = r2 + (d ** 2).reshape(sh)
    lam = torch.fft.rfftn(sigma ** 2 * torch.exp(-r2 / (2.0 * length ** 2))).real.clamp_(min=0.0)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    w = torch.randn(big, generator=gen, dtype=torch.float64, device=device)
    g = torch.fft.irfftn(torch.sqrt(lam) * torch.fft.rfftn(w), s=big)
    return g[tuple(slice(0, s) for s in shape)].contiguous()


def lognormal_kappa_cube_device(n: int, length: float, sigma: float, seed: int, device):
    """Nodal ``kappa = exp(g)`` (float32) on the ``(n+1)^3`` Kuhn lattice, x fastest."""
    import torch
    g = grf_lattice_device((n + 1,) * 3, 1.0 / n, length, sigma, seed, device)
    return torch.exp(g).to(torch.float32).reshape(-1)

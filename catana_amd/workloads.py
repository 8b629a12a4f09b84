"""Seeded synthetic catalogues for the SFB workloads (host side, numpy).

Each generator returns an (N, 3) float64 array of (r, theta, phi) rows, the
AoS layout of catana::PointContainer (Point.hpp:17-18).  The same arrays feed
the GPU path and the CPU oracle.
"""
import numpy as np


def to_spherical(xyz):
    """Cartesian -> (r, theta, phi) as catana::Point(x, y, z) does (Point.cpp:5-10):
    theta = atan2(sqrt(x^2+y^2), z), phi = atan2(y, x) wrapped to [0, 2 pi)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    r = np.sqrt((xyz * xyz).sum(axis=1))
    th = np.arctan2(np.hypot(xyz[:, 0], xyz[:, 1]), xyz[:, 2])
    ph = np.arctan2(xyz[:, 1], xyz[:, 0])
    ph = np.where(ph < 0, ph + 2 * np.pi, ph)
    return np.stack([r, th, ph], axis=1)


def uniform_ball(n, rmax, seed=0):
    """r = rmax u^(1/3), cos(theta) ~ U[-1, 1], phi ~ U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    r = rmax * rng.random(n) ** (1.0 / 3.0)
    th = np.arccos(rng.uniform(-1.0, 1.0, n))
    ph = rng.uniform(0.0, 2 * np.pi, n)
    return np.stack([r, th, ph], axis=1)


def clustered_ball(n_parents, rmax, mean_children=50, sigma=5.0, seed=0):
    """Parents uniform in the ball, Poisson(mean_children) children each at
    isotropic Gaussian offsets, children outside the ball dropped."""
    rng = np.random.default_rng(seed)
    par = to_cartesian(uniform_ball(n_parents, rmax, seed=seed + 1))
    counts = rng.poisson(mean_children, n_parents)
    centres = np.repeat(par, counts, axis=0)
    xyz = centres + rng.normal(0.0, sigma, centres.shape)
    keep = (xyz * xyz).sum(axis=1) < rmax * rmax
    return to_spherical(xyz[keep])


def to_cartesian(pts):
    r, th, ph = pts[:, 0], pts[:, 1], pts[:, 2]
    st = np.sin(th)
    return np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * np.cos(th)], axis=1)


def radial_selection(n, rmax, r0=400.0, seed=0):
    """p(r) ~ r^2 exp(-(r/r0)^2) on [0, rmax] (inverse CDF on a fine grid), isotropic angles."""
    rng = np.random.default_rng(seed)
    grid = np.linspace(0.0, rmax, 200001)
    pdf = grid ** 2 * np.exp(-(grid / r0) ** 2)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    r = np.interp(rng.random(n), cdf, grid)
    th = np.arccos(rng.uniform(-1.0, 1.0, n))
    ph = rng.uniform(0.0, 2 * np.pi, n)
    return np.stack([r, th, ph], axis=1)


def lognormal_ball(n_target, rmax=1000.0, ngrid=256, k0=0.1, seed=0):
    """Lognormal density on an ngrid^3 box of side 2 rmax: Gaussian field with
    P(k) ~ k exp(-(k/k0)^2) normalised to sigma_G = 1, delta = exp(G - 1/2) - 1,
    Poisson sampling of nbar (1 + delta) per cell at uniform in-cell positions,
    cut to r < rmax and trimmed to exactly n_target points by a seeded subset."""
    rng = np.random.default_rng(seed)
    box = 2.0 * rmax
    kf = 2 * np.pi / box
    kx = np.fft.fftfreq(ngrid, d=1.0 / ngrid) * kf
    kz = np.fft.rfftfreq(ngrid, d=1.0 / ngrid) * kf
    k = np.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2)
    pk = k * np.exp(-(k / k0) ** 2)
    noise = rng.normal(size=(ngrid, ngrid, ngrid)).astype(np.float32)
    g = np.fft.irfftn(np.fft.rfftn(noise) * np.sqrt(pk).astype(np.float32), s=noise.shape).astype(np.float32)
    del noise
    g /= g.std()
    dens = np.exp(g - 0.5)
    del g
    cell = box / ngrid
    ball_frac = (4.0 / 3.0) * np.pi / 8.0
    nbar = 1.08 * n_target / (ball_frac * ngrid ** 3)
    counts = rng.poisson(nbar * dens).astype(np.int64)
    del dens
    idx = np.repeat(np.arange(ngrid ** 3, dtype=np.int64), counts.ravel())
    ix, iy, iz = np.unravel_index(idx, (ngrid, ngrid, ngrid))
    del idx
    xyz = np.empty((ix.size, 3))
    xyz[:, 0] = (ix + rng.random(ix.size)) * cell - rmax
    xyz[:, 1] = (iy + rng.random(ix.size)) * cell - rmax
    xyz[:, 2] = (iz + rng.random(ix.size)) * cell - rmax
    keep = (xyz * xyz).sum(axis=1) < rmax * rmax
    xyz = xyz[keep]
    if xyz.shape[0] < n_target:
        raise RuntimeError(f"lognormal sample too small: {xyz.shape[0]} < {n_target}")
    sel = rng.choice(xyz.shape[0], n_target, replace=False)
    sel.sort()
    return to_spherical(xyz[sel])


def gaussian_fixture(path):
    """The reference's golden catalogue (test/test_data/gaussian_catalog.txt), read as float64
    (test/python/decomposition_test.py:16 reads it with CartesianTextSource = double records)."""
    return to_spherical(np.loadtxt(path))

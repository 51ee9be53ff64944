"""Seeded synthetic inputs: velocity models, wavelets and acquisition
geometry for the modelling and FWI workloads.

Vp is a linear gradient in depth times a product of Gaussian anomalies,
rho follows Gardner (0.31*Vp^0.25), Vs = Vp / ratio; the wavelet is a Ricker
(1 - 2(pi f t')^2) exp(-(pi f t')^2) with t' = t - 1/f.
"""
from __future__ import annotations

import numpy as np


def gradient_anomaly_vp(nx, nz, dx, dz, vtop=1500.0, vbot=4000.0, nanom=20, seed=0,
                        amp=0.1, sigma=(50.0, 200.0)):
    """Vp[nz, nx]: linear vtop->vbot in depth times prod(1 + a*exp(-r^2/2s^2))."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float64) * dz
    x = np.arange(nx, dtype=np.float64) * dx
    zz, xx = np.meshgrid(z, x, indexing="ij")
    zmax = max(z[-1], dz)
    vp = vtop + (vbot - vtop) * zz / zmax
    fac = np.ones_like(vp)
    for _ in range(nanom):
        a = rng.uniform(-amp, amp)
        s = rng.uniform(*sigma)
        cx = rng.uniform(0, x[-1])
        cz = rng.uniform(0, z[-1])
        fac *= 1.0 + a * np.exp(-((xx - cx) ** 2 + (zz - cz) ** 2) / (2 * s * s))
    return (vp * fac).astype(np.float32)


def layered_anomaly_vp3d(nx, ny, nz, dx, dy, dz, vels=(1500.0, 1900.0, 2300.0, 2800.0, 3300.0),
                         nanom=10, seed=0, amp=0.08, sigma=(50.0, 150.0)):
    """Vp[nz, ny, nx]: horizontal layers of equal thickness times anomalies."""
    rng = np.random.default_rng(seed)
    nl = len(vels)
    iz = np.minimum((np.arange(nz) * nl) // max(nz, 1), nl - 1)
    vp = np.asarray(vels, np.float64)[iz][:, None, None] * np.ones((1, ny, nx))
    z = np.arange(nz) * dz
    y = np.arange(ny) * dy
    x = np.arange(nx) * dx
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    for _ in range(nanom):
        a = rng.uniform(-amp, amp)
        s = rng.uniform(*sigma)
        c = (rng.uniform(0, x[-1]), rng.uniform(0, y[-1]), rng.uniform(0, z[-1]))
        vp *= 1.0 + a * np.exp(-((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) / (2 * s * s))
    return vp.astype(np.float32)


def gardner(vp):
    return (0.31 * np.asarray(vp, np.float64) ** 0.25).astype(np.float32)


def ricker(nt, dt, f0):
    t = np.arange(nt, dtype=np.float64) * dt - 1.0 / f0
    a = (np.pi * f0 * t) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


def gaussian_smooth(v, sigma_m, dx, dz):
    """Separable Gaussian smoothing (edge-replicated) of a 2D model."""
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(np.asarray(v, np.float64), (sigma_m / dz, sigma_m / dx),
                           mode="nearest").astype(np.float32)


def surface_line(n, d, z=0.0, x0=0.0):
    """n receivers along x at depth z: float32 (n, 2) coordinates (x, z)."""
    return np.stack([x0 + np.arange(n) * d, np.full(n, z)], 1).astype(np.float32)

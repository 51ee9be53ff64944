"""Synthetic acoustic 2D gradient cases for the exact-adjoint tests: a
layered medium with a Gaussian anomaly (+20% Vp, +10% rho), a Ricker source
near the surface and a line of receivers; observed data from the same grid
without the anomaly, so residuals are those of an inversion step.

FP32 conditioning: the residual mod - obs is a difference of two FP32 records
that agree to ~98% (the direct wave dominates both), so its relative error,
and with it the gradient's, is the records' own (~4e-7) times |d| / |r|
(~50 here; ~150 for a +8% anomaly, where the FP32 oracle's gradient error is
6e-5).  Gradients are checked against the FP64 oracle at 1e-4."""
import numpy as np


def ricker(nt, dt, f0):
    t = np.arange(nt) * dt - 1.0 / f0
    a = (np.pi * f0 * t) ** 2
    return (1 - 2 * a) * np.exp(-a)


def exact_case(seed=0, nx=48, nz=36, lpml=10, order=4, fs=False, nt=180, dx=10.0, dt=1e-3, nrec=14):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None] * dx
    x = np.arange(nx)[None, :] * dx
    vp0 = 1800.0 + 4.0 * z + 0.0 * x
    rho0 = 0.31e3 * (vp0 / 1000.0) ** 0.25
    cx, cz = rng.uniform(0.3, 0.7) * nx * dx, rng.uniform(0.4, 0.7) * nz * dx
    bump = np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * (6 * dx) ** 2))
    vp = vp0 * (1 + 0.2 * bump)
    rho = rho0 * (1 + 0.1 * bump)
    sc = np.array([[rng.uniform(0.3, 0.7) * (nx - 1) * dx, 4 * dx]])
    # unit-order pressure: the source term is dt * L * wavelet
    wav = (ricker(nt, dt, 18.0) / (dt * float(rho0.mean() * vp0.mean() ** 2)))[None, :]
    gc = np.stack([np.linspace(2 * dx, (nx - 3) * dx, nrec), np.full(nrec, 3 * dx)], 1)
    return dict(vp=vp, rho=rho, vp0=vp0, rho0=rho0, lpml=lpml, dx=dx, dt=dt, nt=nt, order=order, fs=fs,
                scoords=sc, wav=wav, gcoords=gc)


def observed(c):
    """Records of the background model (oracle, FP64)."""
    from oracle.adjoint import _forward
    _, _, d, _ = _forward(c["vp0"], c["rho0"], c["lpml"], c["dx"], c["dx"], c["dt"], c["nt"], c["order"], c["fs"],
                          c["scoords"], c["wav"], c["gcoords"], np.float64, False)
    return d


def oracle_args(c, dobs):
    return (c["vp"], c["rho"], c["lpml"], c["dx"], c["dx"], c["dt"], c["nt"], c["order"], c["fs"], c["scoords"],
            c["wav"], c["gcoords"], dobs)


def exact_case_e2d(seed=0, nx=44, nz=34, lpml=10, order=4, fs=False, nt=160, dx=10.0, dt=1e-3, nrec=12,
                   srcfield="P"):
    """Elastic version: the acoustic case's Vp and density, Vs = Vp / 1.8
    (+15% in the anomaly), an explosive (P) or force (VX / VZ) source."""
    c = exact_case(seed=seed, nx=nx, nz=nz, lpml=lpml, order=order, fs=fs, nt=nt, dx=dx, dt=dt, nrec=nrec)
    rng = np.random.default_rng(seed + 1000)
    z = np.arange(nz)[:, None] * dx
    x = np.arange(nx)[None, :] * dx
    cx, cz = rng.uniform(0.3, 0.7) * nx * dx, rng.uniform(0.4, 0.7) * nz * dx
    bump = np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * (5 * dx) ** 2))
    c["vs0"] = c["vp0"] / 1.8
    c["vs"] = c["vs0"] * (1 + 0.15 * bump)
    c["srcfield"] = srcfield
    # unit-order stress (explosive source: the term is dt * wavelet) or velocity
    amp = 1.0 / dt if srcfield == "P" else 1.0 / (dt * float(c["rho0"].mean()))
    c["wav"] = (ricker(nt, dt, 18.0) * amp)[None, :]
    return c


def observed_e2d(c):
    """P, Ux, Uz records of the background model (oracle, FP64)."""
    from oracle.adjoint import _forward_e2d
    _, _, d, _ = _forward_e2d(c["vp0"], c["vs0"], c["rho0"], c["lpml"], c["dx"], c["dx"], c["dt"], c["nt"],
                              c["order"], c["fs"], c["srcfield"], c["scoords"], c["wav"], c["gcoords"],
                              np.float64, False)
    return d


def oracle_args_e2d(c, dobs):
    return (c["vp"], c["vs"], c["rho"], c["lpml"], c["dx"], c["dx"], c["dt"], c["nt"], c["order"], c["fs"],
            c["srcfield"], c["scoords"], c["wav"], c["gcoords"], dobs)


def exact_case_a3d(seed=0, nx=20, ny=16, nz=18, lpml=6, order=4, fs=False, nt=110, dx=10.0, dt=1e-3):
    """Acoustic 3D version: a layered medium with a Gaussian anomaly (+20% Vp,
    +10% rho), a Ricker source near the surface and a 4 x 3 receiver grid."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None, None] * dx
    y = np.arange(ny)[None, :, None] * dx
    x = np.arange(nx)[None, None, :] * dx
    vp0 = 1800.0 + 6.0 * z + 0.0 * (x + y)
    rho0 = 0.31e3 * (vp0 / 1000.0) ** 0.25
    cx, cy, cz = (rng.uniform(0.35, 0.65) * n * dx for n in (nx, ny, nz))
    bump = np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * (3 * dx) ** 2))
    vp = vp0 * (1 + 0.2 * bump)
    rho = rho0 * (1 + 0.1 * bump)
    sc = np.array([[rng.uniform(0.4, 0.6) * (nx - 1) * dx, rng.uniform(0.4, 0.6) * (ny - 1) * dx, 3 * dx]])
    wav = (ricker(nt, dt, 20.0) / (dt * float(rho0.mean() * vp0.mean() ** 2)))[None, :]
    gx, gy = np.meshgrid(np.linspace(2 * dx, (nx - 3) * dx, 4), np.linspace(2 * dx, (ny - 3) * dx, 3))
    gc = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, 2 * dx)], 1)
    return dict(vp=vp, rho=rho, vp0=vp0, rho0=rho0, lpml=lpml, d3=(dx, dx, dx), dt=dt, nt=nt, order=order, fs=fs,
                scoords=sc, wav=wav, gcoords=gc)


def observed_a3d(c):
    from oracle.adjoint import _forward_a3d
    _, _, d, _ = _forward_a3d(c["vp0"], c["rho0"], c["lpml"], c["d3"], c["dt"], c["nt"], c["order"], c["fs"],
                              c["scoords"], c["wav"], c["gcoords"], np.float64, False)
    return d


def oracle_args_a3d(c, dobs):
    return (c["vp"], c["rho"], c["lpml"], c["d3"], c["dt"], c["nt"], c["order"], c["fs"], c["scoords"], c["wav"],
            c["gcoords"], dobs)

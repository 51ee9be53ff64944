"""Write the golden fixtures under tests/golden/ from the compiled reference.

Run in a container where /root/reference is mounted and built
(``make -C oracle ref``):  python tests/golden/make_golden.py

Every fixture is a compressed .npz holding the inputs and the reference's
outputs (float32 unless noted).  The reference itself never travels: only
these vectors do.  Sizes are kept small (a few MB in total).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from oracle import runs  # noqa: E402
from rockseis_amd import synth  # noqa: E402


def save(name, **arrs):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrs)
    print(f"{name}.npz  {os.path.getsize(path) / 1024:.1f} KiB")


def der_fixture(R):
    rng = np.random.default_rng(11)
    nx, ny, nz = 17, 18, 19
    f = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    out = {"f": f, "d": np.float32([10.0, 7.5, 5.0])}
    cases = [(o, w) for o in (1, 2, 4, 8) for w in range(6)]
    cases += [(o, w) for o in (3, 5, 6, 7) for w in (0, 5)]
    for order, which in cases:
        df = np.zeros_like(f)
        R.ref_der(f.ravel(), df.reshape(-1), nx, ny, nz, 10.0, 7.5, 5.0, order, which)
        out[f"o{order}_w{which}"] = df
    save("der", **out)


def pml_fixture(R):
    out = {}
    for lpml in (10, 20, 30, 40):
        for dt in (1e-3, 5e-4, 4.5e-4):
            for sign in (1, -1):
                t = np.zeros(12 * lpml, np.float32)
                R.ref_pml_abc(lpml, dt, 1200.0, sign, t)
                out[f"L{lpml}_dt{dt:g}_s{sign}"] = t
    save("pml", **out)


def stagger_fixture(R):
    rng = np.random.default_rng(12)
    nx, ny, nz, lpml = 23, 11, 17, 6
    vp = synth.gradient_anomaly_vp(nx, nz, 10.0, 10.0, nanom=4, seed=3)
    vs = (vp / 1.73).astype(np.float32)
    rho = synth.gardner(vp)
    vp3 = (2000 + 800 * rng.random((nz, ny, nx))).astype(np.float32)
    rho3 = synth.gardner(vp3)
    out = {"vp": vp, "vs": vs, "rho": rho, "vp3": vp3, "rho3": rho3, "lpml": np.int32(lpml)}
    shp = (nz + 2 * lpml, nx + 2 * lpml)
    shp3 = (nz + 2 * lpml, ny + 2 * lpml, nx + 2 * lpml)
    for fs in (0, 1):
        a = [np.zeros(shp, np.float32) for _ in range(3)]
        R.ref_stagger_a2d(vp.ravel(), rho.ravel(), nx, nz, lpml, fs, *(x.reshape(-1) for x in a))
        out.update({f"a2d_fs{fs}_{k}": v for k, v in zip(("L", "Rx", "Rz"), a)})
        a = [np.zeros(shp, np.float32) for _ in range(5)]
        R.ref_stagger_e2d(vp.ravel(), vs.ravel(), rho.ravel(), nx, nz, lpml, fs, *(x.reshape(-1) for x in a))
        out.update({f"e2d_fs{fs}_{k}": v for k, v in zip(("L", "L2M", "M", "Rx", "Rz"), a)})
        a = [np.zeros(shp3, np.float32) for _ in range(4)]
        R.ref_stagger_a3d(vp3.ravel(), rho3.ravel(), nx, ny, nz, lpml, fs, *(x.reshape(-1) for x in a))
        out.update({f"a3d_fs{fs}_{k}": v for k, v in zip(("L", "Rx", "Ry", "Rz"), a)})
    save("stagger", **out)


def makemap_fixture(R):
    rng = np.random.default_rng(13)
    n = 300
    coords = np.stack([rng.uniform(-30, 430, n), rng.uniform(-30, 330, n)], 1).astype(np.float32)
    # exact grid points, edges and values a hair either side of a node
    coords[:12] = np.float32([[0, 0], [10, 10], [390, 290], [395, 295], [400, 300], [5, 5],
                              [100, 0], [10.000001, 5], [9.99999, 5], [15, 15], [-0.0001, 0],
                              [389.9999, 289.9999]])
    out = {"coords": coords}
    for field in (0, 1, 3):
        m = np.zeros((n, 2), np.int32)
        s = np.zeros((n, 2), np.float32)
        R.ref_makemap2d(coords.ravel(), n, 40, 30, 10.0, 10.0, 0.0, 0.0, field, m.reshape(-1), s.reshape(-1))
        out[f"map_f{field}"] = m
        out[f"shift_f{field}"] = s
    save("makemap", **out)


def _model2d(nx, nz, dx, seed):
    vp = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3200.0, nanom=5, seed=seed,
                                   sigma=(30.0, 120.0))
    return vp, (vp / 1.73).astype(np.float32), synth.gardner(vp)


A2D_CASES = [  # order, fs, srcfield, recfield
    (4, 0, "P", "P"), (8, 1, "VX", "VZ"), (2, 0, "VZ", "VX"), (8, 0, "P", "P"), (6, 1, "P", "VX"),
]
E2D_CASES = [  # order, fs, srcfield, recfield, recip, nbw
    (8, 0, "P", "VX", 0, 0), (8, 1, "P", "P", 1, 25), (4, 0, "VX", "VZ", 0, 25), (4, 1, "VZ", "P", 0, 0),
]


def run_fixtures(R):
    nx, nz, lpml, dx, dt, nt = 60, 50, 10, 10.0, 1e-3, 300
    vp, vs, rho = _model2d(nx, nz, dx, 5)
    sc = np.float32([[300.0, 0.0], [297.5, 33.3]])
    wav = np.stack([synth.ricker(nt, dt, 15.0), 0.5 * synth.ricker(nt, dt, 20.0)])
    gc = np.stack([np.arange(0, 600, 10.0) + 2.5, np.full(60, 20.0)], 1).astype(np.float32)
    gc[::7, 1] = 0.0
    common = dict(vp=vp, vs=vs, rho=rho, scoords=sc, wav=wav, gcoords=gc,
                  grid=np.float32([nx, nz, lpml, dx, dt, nt]))
    for k, (order, fs, sf, rf) in enumerate(A2D_CASES):
        r = runs.ref_a2d_run(vp, rho, lpml, dx, dx, dt, nt, order, fs, sf, sc, wav, rf, gc,
                             snap_its=(10, 150))
        save(f"a2d_case{k}", rec=r["rec"], P=r["P"], Vx=r["Vx"], Vz=r["Vz"],
             snap10=r["snaps"][0], snap150=r["snaps"][1], **common)
    for k, (order, fs, sf, rf, recip, nbw) in enumerate(E2D_CASES):
        r = runs.ref_e2d_run(vp, vs, rho, lpml, dx, dx, dt, nt, order, fs, sf, sc, wav, rf, gc,
                             recip, nbw)
        save(f"e2d_case{k}", rec=r["rec"], **{f: r[f] for f in runs.E2DState.FIELDS}, **common)
    # acoustic 3D
    rng = np.random.default_rng(14)
    nx3, ny3, nz3, l3, nt3 = 22, 24, 20, 6, 80
    vp3 = (2000 + 500 * rng.random((nz3, ny3, nx3))).astype(np.float32)
    rho3 = synth.gardner(vp3)
    sc3 = np.float32([[100.0, 110.0, 0.0]])
    w3 = synth.ricker(nt3, dt, 15.0)[None]
    g = np.stack(np.meshgrid(np.arange(0, 220, 20.0), np.arange(0, 240, 30.0), indexing="ij"), -1).reshape(-1, 2)
    gc3 = np.concatenate([g, np.full((g.shape[0], 1), 15.0)], 1).astype(np.float32)
    for k, (order, sf, rf) in enumerate([(4, "P", "P"), (8, "VY", "VZ"), (2, "VX", "P")]):
        r = runs.ref_a3d_run(vp3, rho3, l3, (10.0, 10.0, 10.0), dt, nt3, order, 0, sf, sc3, w3, rf, gc3)
        save(f"a3d_case{k}", rec=r["rec"], P=r["P"], Vz=r["Vz"],
             vp=vp3, rho=rho3, scoords=sc3, wav=w3, gcoords=gc3,
             grid=np.float32([nx3, ny3, nz3, l3, 10.0, dt, nt3, order]))


FWI_CASES = [  # order, fs, ncheck
    (4, 0, 20), (8, 1, 7),
]


def fwi_fixtures(R):
    from oracle import fwi
    nx, nz, lpml, dx, dt, nt = 60, 50, 10, 10.0, 1e-3, 300
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=3)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    rho0 = synth.gardner(vp0)
    sc = np.float32([[300.0, 10.0]])
    wav = synth.ricker(nt, dt, 15.0)[None]
    gc = synth.surface_line(nx, dx, 20.0)
    for k, (order, fs, ncheck) in enumerate(FWI_CASES):
        obs = runs.ref_a2d_run(vtrue, rtrue, lpml, dx, dx, dt, nt, order, fs, "P", sc, wav, "P", gc)["rec"]
        r = fwi.ref_fwi_a2d(vp0, rho0, lpml, dx, dx, dt, nt, order, fs, ncheck, sc, wav, gc, obs)
        save(f"fwi_case{k}", vp=vp0, rho=rho0, scoords=sc, wav=wav, gcoords=gc, dobs=obs,
             grid=np.float32([nx, nz, lpml, dx, dt, nt]), misfit=np.float32(r["misfit"]),
             vpgrad=r["vpgrad"], rhograd=r["rhograd"], dmod=r["dmod"], dres=r["dres"],
             wavgrad=r["wavgrad"])


FWI_MISFIT_CASES = [  # order, fs, ncheck, misfit type (1 CORRELATION, 2/3 ADAPTIVE_GAUSS/LINEAR), data weights
    (4, 0, 20, 1, False), (8, 1, 7, 1, True), (4, 0, 20, 0, True), (4, 0, 20, 2, False), (8, 1, 7, 3, True),
]


def fwi_misfit_fixtures(R):
    """run_optimal with the CORRELATION and ADAPTIVE misfits / data weights (setDataweight)."""
    from oracle import fwi
    nx, nz, lpml, dx, dt, nt = 60, 50, 10, 10.0, 1e-3, 300
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=3)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    rho0 = synth.gardner(vp0)
    sc = np.float32([[300.0, 10.0]])
    wav = synth.ricker(nt, dt, 15.0)[None]
    gc = synth.surface_line(nx, dx, 20.0)
    rng = np.random.default_rng(1)
    for k, (order, fs, ncheck, mt, usew) in enumerate(FWI_MISFIT_CASES):
        obs = runs.ref_a2d_run(vtrue, rtrue, lpml, dx, dx, dt, nt, order, fs, "P", sc, wav, "P", gc)["rec"]
        w = (0.5 + rng.random((nx, nt))).astype(np.float32) if usew else np.ones((nx, nt), np.float32)
        r = fwi.ref_fwi_a2d_misfit(vp0, rho0, lpml, dx, dx, dt, nt, order, fs, ncheck, sc, wav, gc, obs, mt,
                                   w if usew else None)
        save(f"fwi_misfit_case{k}", vp=vp0, rho=rho0, scoords=sc, wav=wav, gcoords=gc, dobs=obs, weight=w,
             grid=np.float32([nx, nz, lpml, dx, dt, nt]), misfit=np.float32(r["misfit"]),
             vpgrad=r["vpgrad"], rhograd=r["rhograd"], dmod=r["dmod"], dres=r["dres"], wavgrad=r["wavgrad"])


FWI_FULL_CASES = [  # order, fs, snapinc (FwiAcoustic2D::run)
    (4, 0, 1), (8, 1, 3), (4, 0, 7),
]


def fwi_full_fixtures(R):
    from oracle import fwi
    nx, nz, lpml, dx, dt, nt = 60, 50, 10, 10.0, 1e-3, 300
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=3)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    rho0 = synth.gardner(vp0)
    sc = np.float32([[300.0, 10.0]])
    wav = synth.ricker(nt, dt, 15.0)[None]
    gc = synth.surface_line(nx, dx, 20.0)
    for k, (order, fs, snapinc) in enumerate(FWI_FULL_CASES):
        obs = runs.ref_a2d_run(vtrue, rtrue, lpml, dx, dx, dt, nt, order, fs, "P", sc, wav, "P", gc)["rec"]
        r = fwi.ref_fwi_a2d_full(vp0, rho0, lpml, dx, dx, dt, nt, order, fs, snapinc, sc, wav, gc, obs)
        save(f"fwi_full_case{k}", vp=vp0, rho=rho0, scoords=sc, wav=wav, gcoords=gc, dobs=obs,
             grid=np.float32([nx, nz, lpml, dx, dt, nt]), misfit=np.float32(r["misfit"]),
             vpgrad=r["vpgrad"], rhograd=r["rhograd"], dmod=r["dmod"], dres=r["dres"],
             wavgrad=r["wavgrad"])


FWI_E2D_CASES = [  # order, fs, ncheck, source field (FwiElastic2D::run_optimal, P/Ux/Uz data)
    (4, 0, 15, "P"), (8, 1, 6, "VZ"),
]


def fwi_e2d_fixtures(R):
    """The reference's scaled Vp/Vs/rho gradients, data, residuals, wavelet
    gradient and misfit; plus the oracle's lambda/mu/rho kernels before
    scaleGrad (oracle_*), whose scaleGrad reproduces the reference bit for
    bit (checked here before saving)."""
    from oracle import fwi
    nx, nz, lpml, dx, dt, nt = 60, 50, 10, 10.0, 1e-3, 250
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=3)
    vstrue = (vtrue / 1.8).astype(np.float32)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    vs0 = (vp0 / 1.8).astype(np.float32)
    rho0 = synth.gardner(vp0)
    sc = np.float32([[300.0, 20.0]])
    wav = synth.ricker(nt, dt, 15.0)[None]
    gc = synth.surface_line(nx, dx, 20.0)
    for k, (order, fs, ncheck, sf) in enumerate(FWI_E2D_CASES):
        obs = np.stack([runs.ref_e2d_run(vtrue, vstrue, rtrue, lpml, dx, dx, dt, nt, order, fs, sf, sc, wav, f,
                                         gc)["rec"] for f in ("P", "VX", "VZ")])
        r = fwi.ref_fwi_e2d(vp0, vs0, rho0, lpml, dx, dx, dt, nt, order, fs, ncheck, sf, sc, wav, gc, obs)
        o = fwi.fwi_e2d(vp0, vs0, rho0, lpml, dx, dx, dt, nt, order, fs, ncheck, sf, sc, wav, gc, obs)
        for key in ("vpgrad", "vsgrad", "rhograd", "dmod", "dres", "wavgrad"):
            assert np.array_equal(r[key], o[key]), key
        save(f"fwi_e2d_case{k}", vp=vp0, vs=vs0, rho=rho0, scoords=sc, wav=wav, gcoords=gc, dobs=obs,
             grid=np.float32([nx, nz, lpml, dx, dt, nt]), misfit=np.float32(r["misfit"]),
             vpgrad=r["vpgrad"], vsgrad=r["vsgrad"], rhograd=r["rhograd"], dmod=r["dmod"], dres=r["dres"],
             wavgrad=r["wavgrad"], oracle_lam=o["lam"], oracle_mu=o["mu"], oracle_rho=o["rhok"])


FWI_E2D_MISFIT_CASES = [  # order, fs, ncheck, source field, misfit type, data weights set (P, Ux, Uz)
    (4, 0, 15, "P", 1, ""), (8, 1, 6, "VZ", 1, "pxz"), (8, 1, 6, "VZ", 2, "xz"), (4, 0, 15, "P", 3, ""),
    (4, 0, 15, "VZ", 4, "xz"), (4, 0, 15, "P", 0, "pz"),
]


def fwi_e2d_misfit_fixtures(R):
    """FwiElastic2D::run_optimal with the CORRELATION / ADAPTIVE_GAUSS /
    ADAPTIVE_LINEAR / ADAPTIVE_SINGLE misfits and data weights
    (setDataweightp/x/z); the oracle's run is checked bit for bit first."""
    from oracle import fwi
    nx, nz, lpml, dx, dt, nt = 60, 50, 10, 10.0, 1e-3, 250
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=3)
    vstrue = (vtrue / 1.8).astype(np.float32)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    vs0 = (vp0 / 1.8).astype(np.float32)
    rho0 = synth.gardner(vp0)
    sc = np.float32([[300.0, 20.0]])
    wav = synth.ricker(nt, dt, 15.0)[None]
    gc = synth.surface_line(nx, dx, 20.0)
    rng = np.random.default_rng(4)
    for k, (order, fs, ncheck, sf, mt, wset) in enumerate(FWI_E2D_MISFIT_CASES):
        obs = np.stack([runs.ref_e2d_run(vtrue, vstrue, rtrue, lpml, dx, dx, dt, nt, order, fs, sf, sc, wav, f,
                                         gc)["rec"] for f in ("P", "VX", "VZ")])
        w = [(0.5 + rng.random((nx, nt))).astype(np.float32) if c in wset else None for c in "pxz"]
        r = fwi.ref_fwi_e2d(vp0, vs0, rho0, lpml, dx, dx, dt, nt, order, fs, ncheck, sf, sc, wav, gc, obs, mt, w)
        o = fwi.fwi_e2d(vp0, vs0, rho0, lpml, dx, dx, dt, nt, order, fs, ncheck, sf, sc, wav, gc, obs,
                        misfit_type=mt, weights=w)
        for key in ("vpgrad", "vsgrad", "rhograd", "dmod", "dres", "wavgrad"):
            assert np.array_equal(r[key], o[key]), key
        assert np.float32(r["misfit"]) == np.float32(o["misfit"])
        wts = {f"wei_{c}": w[i] for i, c in enumerate("pxz") if w[i] is not None}
        save(f"fwi_e2d_misfit_case{k}", vp=vp0, vs=vs0, rho=rho0, scoords=sc, wav=wav, gcoords=gc, dobs=obs,
             grid=np.float32([nx, nz, lpml, dx, dt, nt]), misfit=np.float32(r["misfit"]), vpgrad=r["vpgrad"],
             vsgrad=r["vsgrad"], rhograd=r["rhograd"], dmod=r["dmod"], dres=r["dres"], wavgrad=r["wavgrad"], **wts)


FWI_A3D_CASES = [  # order, fs, ncheck (FwiAcoustic3D::run_optimal)
    (4, 0, 10), (8, 1, 5),
]


def fwi_a3d_fixtures(R):
    from scipy.ndimage import gaussian_filter
    from oracle import fwi
    nx, ny, nz, lpml, dx, dt, nt = 26, 24, 22, 6, 10.0, 1e-3, 120
    vt = synth.layered_anomaly_vp3d(nx, ny, nz, dx, dx, dx, vels=(1800., 2200., 2600., 3000.), nanom=4, seed=2)
    rt = synth.gardner(vt)
    v0 = gaussian_filter(vt.astype(np.float64), 2.0, mode="nearest").astype(np.float32)
    r0 = synth.gardner(v0)
    sc = np.float32([[130., 120., 20.]])
    wav = synth.ricker(nt, dt, 25.0)[None]
    gx, gy = np.meshgrid(np.arange(2, nx - 2, 3) * dx, np.arange(2, ny - 2, 3) * dx)
    gc = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, 20.0)], 1).astype(np.float32)
    for k, (order, fs, ncheck) in enumerate(FWI_A3D_CASES):
        obs = runs.ref_a3d_run(vt, rt, lpml, (dx, dx, dx), dt, nt, order, fs, "P", sc, wav, "P", gc)["rec"]
        r = fwi.ref_fwi_a3d(v0, r0, lpml, (dx, dx, dx), dt, nt, order, fs, ncheck, sc, wav, gc, obs)
        save(f"fwi_a3d_case{k}", vp=v0, rho=r0, scoords=sc, wav=wav, gcoords=gc, dobs=obs,
             grid=np.float32([nx, ny, nz, lpml, dx, dt, nt]), misfit=np.float32(r["misfit"]),
             vpgrad=r["vpgrad"], rhograd=r["rhograd"], dmod=r["dmod"], dres=r["dres"], wavgrad=r["wavgrad"])


FWI_A3D_MISFIT_CASES = [  # order, fs, ncheck, misfit type (1 CORRELATION, 2/3 ADAPTIVE_GAUSS/LINEAR, 4 default), weights
    (4, 0, 10, 1, False), (4, 0, 10, 1, True), (4, 0, 10, 2, False), (4, 0, 10, 3, True), (4, 0, 10, 4, True),
]


def fwi_a3d_misfit_fixtures(R):
    """FwiAcoustic3D::run_optimal with the CORRELATION and ADAPTIVE misfits,
    the default branch (type 4) and data weights (setDataweight): the 3D class
    has no percentile clip of the residuals, unlike the 2D one."""
    from oracle import fwi
    g = np.load(os.path.join(HERE, "fwi_a3d_case0.npz"), allow_pickle=False)
    nx, ny, nz, lpml, dx, dt, nt = g["grid"]
    lpml, nt = int(lpml), int(nt)
    rng = np.random.default_rng(4)
    for k, (order, fs, ncheck, mt, usew) in enumerate(FWI_A3D_MISFIT_CASES):
        w = (0.5 + rng.random(g["dobs"].shape)).astype(np.float32) if usew else None
        r = fwi.ref_fwi_a3d(g["vp"], g["rho"], lpml, (dx, dx, dx), dt, nt, order, fs, ncheck, g["scoords"], g["wav"],
                            g["gcoords"], g["dobs"], mt, w)
        extra = {} if w is None else dict(weight=w)
        save(f"fwi_a3d_misfit_case{k}", misfit=np.float32(r["misfit"]), vpgrad=r["vpgrad"], rhograd=r["rhograd"],
             dmod=r["dmod"], dres=r["dres"], wavgrad=r["wavgrad"], **extra)


RTM_CASES = [  # order, fs, ncheck, nhx, nhz (RtmAcoustic2D::run_optimal)
    (4, 0, 20, 1, 1), (8, 1, 7, 5, 3), (4, 0, 20, 4, 2),
]


def rtm_fixtures(R):
    """RtmAcoustic2D::run_optimal on a smoothed model with the true model's
    records; even and odd subsurface-offset counts and a free surface."""
    from oracle import rtm
    nx, nz, lpml, dx, dt, nt = 60, 50, 10, 10.0, 1e-3, 300
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=3)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    rho0 = synth.gardner(vp0)
    sc = np.float32([[300.0, 10.0]])
    wav = synth.ricker(nt, dt, 15.0)[None]
    gc = synth.surface_line(nx, dx, 20.0)
    for k, (order, fs, ncheck, nhx, nhz) in enumerate(RTM_CASES):
        data = runs.ref_a2d_run(vtrue, rtrue, lpml, dx, dx, dt, nt, order, fs, "P", sc, wav, "P", gc)["rec"]
        img = rtm.ref_rtm_a2d(vp0, rho0, lpml, dx, dx, dt, nt, order, fs, ncheck, sc, wav, gc, data, nhx, nhz)
        save(f"rtm_case{k}", vp=vp0, rho=rho0, scoords=sc, wav=wav, gcoords=gc, data=data,
             grid=np.float32([nx, nz, lpml, dx, dt, nt]), image=img)


RTM_A3D_CASES = [  # order, fs, ncheck, nhx, nhy, nhz (RtmAcoustic3D::run_optimal)
    (4, 0, 10, 1, 1, 1), (4, 1, 5, 3, 2, 3),
]


def rtm_a3d_fixtures(R):
    """RtmAcoustic3D::run_optimal on the 3D FWI fixture's model with its
    observed records; subsurface offsets in all three directions."""
    from oracle import rtm
    g = np.load(os.path.join(HERE, "fwi_a3d_case0.npz"), allow_pickle=False)
    nx, ny, nz, lpml, dx, dt, nt = g["grid"]
    for k, (order, fs, ncheck, nhx, nhy, nhz) in enumerate(RTM_A3D_CASES):
        img = rtm.ref_rtm_a3d(g["vp"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), order, fs, ncheck,
                              g["scoords"], g["wav"], g["gcoords"], g["dobs"], nhx, nhy, nhz)
        save(f"rtm_a3d_case{k}", image=img)


INTERP_CASES = [  # ntr, n1, d1, o1, n2, d2, o2 (Interp::interp for every mode)
    (5, 300, 1e-3, 0.0, 700, 4e-4, 0.0), (3, 1000, 4.5e-4, 0.0, 150, 3e-3, 0.0),
    (4, 257, 2e-3, 0.01, 300, 1.7e-3, -0.003),
]


E3D_CASES = [  # order, fs, source field, record field (WavesElastic3D modelling)
    (4, 0, "P", "P"), (8, 1, "VZ", "VX"), (2, 1, "VY", "P"), (6, 0, "VX", "VY"),
]


def e3d_fixtures(R):
    """WavesElastic3D modelling runs of the compiled reference: records and
    four of the nine final fields (Vx, Vz, Szz, Sxy), plus the staggered model
    of ModelElastic3D::staggerModels with and without free surface."""
    rng = np.random.default_rng(21)
    nx, ny, nz, lpml, nt, dt, dx = 16, 14, 12, 5, 60, 1e-3, 10.0
    vp = (2500 + 400 * rng.random((nz, ny, nx))).astype(np.float32)
    vs = (vp / (1.7 + 0.2 * rng.random((nz, ny, nx)))).astype(np.float32)
    rho = synth.gardner(vp)
    sc = np.float32([[72.5, 61.0, 10.0]])
    wav = synth.ricker(nt, dt, 25.0)[None].astype(np.float32)
    g = np.stack(np.meshgrid(np.arange(0, 150, 20.0), np.arange(0, 130, 30.0), indexing="ij"), -1).reshape(-1, 2)
    gc = np.concatenate([g, np.full((g.shape[0], 1), 15.0)], 1).astype(np.float32)
    stag = {f"stag_fs{fs}": np.stack(runs.ref_stagger_e3d(vp, vs, rho, lpml, fs)) for fs in (0, 1)}
    save("e3d_stagger", vp=vp, vs=vs, rho=rho, lpml=np.int32(lpml), **stag)
    for k, (order, fs, sf, rf) in enumerate(E3D_CASES):
        r = runs.ref_e3d_run(vp, vs, rho, lpml, (dx, dx, dx), dt, nt, order, fs, sf, sc, wav, rf, gc)
        save(f"e3d_case{k}", rec=r["rec"], Vx=r["Vx"], Vz=r["Vz"], Szz=r["Szz"], Sxy=r["Sxy"], vp=vp, vs=vs,
             rho=rho, scoords=sc, wav=wav, gcoords=gc, grid=np.float32([nx, ny, nz, lpml, dx, dt, nt]))


def interp_fixture(R):
    rng = np.random.default_rng(17)
    out = {}
    for k, (ntr, n1, d1, o1, n2, d2, o2) in enumerate(INTERP_CASES):
        a = rng.standard_normal((ntr, n1)).astype(np.float32)
        out[f"in{k}"] = a
        for mode in (0, 1, 2):
            b = np.zeros((ntr, n2), np.float32)
            R.ref_interp(mode, a.ravel(), ntr, n1, d1, o1, b.reshape(-1), n2, d2, o2)
            out[f"out{k}_m{mode}"] = b
    save("interp", **out)


REVOLVE_CASES = [(10, 3), (100, 5), (7, 2), (50, 60), (1, 2), (257, 4), (3000, 50), (2500, 50), (4000, 50)]


def revolve_fixture(R):
    out = {}
    for nt, sn in REVOLVE_CASES:
        a = np.zeros(100000, np.int32)
        c = a.copy()
        k = a.copy()
        n = R.ref_revolve(nt, sn, a, c, k, a.size)
        out[f"n{nt}_s{sn}"] = np.stack([a[:n], c[:n], k[:n]], 1).astype(np.int32)
        out[f"n{nt}_s{sn}_numforw"] = np.int32(R.ref_revolve_numforw(nt, sn))
    for nt in (10, 100, 1000, 3000):
        out[f"adjust_{nt}"] = np.int32(R.ref_revolve_adjust(nt))
    save("revolve", **out)


GROUPS = dict(der=der_fixture, pml=pml_fixture, stagger=stagger_fixture, makemap=makemap_fixture,
              runs=run_fixtures, revolve=revolve_fixture, fwi=fwi_fixtures, fwi_full=fwi_full_fixtures,
              fwi_e2d=fwi_e2d_fixtures, fwi_a3d=fwi_a3d_fixtures,
              interp=interp_fixture, fwi_misfit=fwi_misfit_fixtures,
              fwi_e2d_misfit=fwi_e2d_misfit_fixtures, fwi_a3d_misfit=fwi_a3d_misfit_fixtures, rtm=rtm_fixtures,
              rtm_a3d=rtm_a3d_fixtures)

DOMDEC_CASES = [  # elastic, fs, map (0 SMAP, 1 GMAP), aperture, shot x, receiver x0, dx of the spread
    (0, 0, 0, 400.0, 230.0, 13.0, 20.0), (0, 1, 1, 150.0, 355.0, 101.0, 7.5), (0, 0, 0, 0.0, 230.0, 40.0, 25.0),
    (0, 1, 0, -120.0, 90.0, 3.0, 30.0), (1, 1, 0, 400.0, 230.0, 13.0, 20.0), (1, 0, 1, 250.0, 355.0, 101.0, 7.5),
]


def domdec_fixture(R):
    """Model::getLocalsize2d and Model*::getDomainmodel (one domain) of the
    reference on model files, for both map types, positive, zero and
    negative apertures, with and without free surface."""
    import ctypes as C
    import tempfile

    from rockseis_amd import rss
    nx, nz, lpml, dx = 47, 31, 8, 10.0
    vp, vs, rho = _model2d(nx, nz, dx, 9)
    out = {"vp": vp, "vs": vs, "rho": rho, "grid": np.float32([nx, nz, lpml, dx]),
           "cases": np.float32(DOMDEC_CASES)}
    fp = C.POINTER(C.c_float)
    with tempfile.TemporaryDirectory() as d:
        files = [os.path.join(d, f) for f in ("vp.rss", "vs.rss", "rho.rss")]
        for f, a in zip(files, (vp, vs, rho)):
            rss.write_model2d(f, a, dx, dx, 15.0, 5.0)
        for k, (el, fs, mp, ap, sx, g0, gd) in enumerate(DOMDEC_CASES):
            ntr = 12
            coords = np.zeros((ntr, 4), np.float32)
            coords[:, 0], coords[:, 1] = sx, 10.0
            coords[:, 2], coords[:, 3] = g0 + gd * np.arange(ntr), 20.0
            oi = (C.c_longlong * 4)()
            of = (C.c_float * 2)()
            vsf = files[1].encode() if el else b""
            args = (files[0].encode(), vsf, files[2].encode(), lpml, int(fs), coords.ctypes.data_as(fp), ntr,
                    C.c_float(ap), int(mp), 8, oi, of)
            R.ref_domain_model(*args, *([None] * 8))
            nxd, nzd = int(oi[2]), int(oi[3])
            arrs = {n: np.zeros((nzd, nxd), np.float32) for n in ("Vp", "Vs", "R", "L", "L2M", "M", "Rx", "Rz")}
            R.ref_domain_model(*args, *[arrs[n].ctypes.data_as(fp) for n in ("Vp", "Vs", "R", "L", "L2M", "M",
                                                                              "Rx", "Rz")])
            out[f"c{k}_coords"] = coords
            out[f"c{k}_int"] = np.int64(list(oi))
            out[f"c{k}_origin"] = np.float32(list(of))
            for n, a in arrs.items():
                if el or n in ("Vp", "R", "L", "Rx", "Rz"):
                    out[f"c{k}_{n}"] = a
    offs = []
    rng = np.random.default_rng(4)
    for _ in range(64):
        ol, og = np.float32(rng.uniform(-500, 500)), np.float32(rng.uniform(-500, 500))
        dl = np.float32(rng.choice([2.5, 5.0, 10.0, 12.5]))
        offs.append((ol, dl, og, dl, R.ref_image_offset(C.c_float(ol), C.c_float(dl), C.c_float(og),
                                                        C.c_float(dl))))
    out["image_offsets"] = np.float64(offs)
    save("domdec", **out)


GROUPS["domdec"] = domdec_fixture
GROUPS["e3d"] = e3d_fixtures

FWI_E3D_CASES = [  # order, fs, source field, misfit type, weights (x, y, z), ncheck
    (4, 0, "P", 0, "", 7), (8, 1, "VZ", 1, "", 5), (4, 1, "VX", 2, "", 9), (6, 0, "P", 3, "xz", 6),
    (2, 0, "VY", 0, "xyz", 60),
]


def fwi_e3d_fixtures(R):
    """FwiElastic3D::run_optimal of the compiled reference: Vp, Vs, rho and
    wavelet gradients, modelled records and residuals of the three
    components for five cases (orders 2-8, free surface, P / VX / VY / VZ
    sources, DIFFERENCE, CORRELATION, ADAPTIVE_GAUSS and the default branch,
    data weights, Revolve with few and with enough snapshots).  The observed
    records are the reference's own modelling in a 4% faster model."""
    from oracle import fwi
    rng = np.random.default_rng(23)
    nx, ny, nz, lpml, nt, dt, dx = 14, 12, 13, 4, 60, 1e-3, 10.0
    vp = (2300 + 300 * rng.random((nz, ny, nx))).astype(np.float32)
    vs = (vp / (1.7 + 0.2 * rng.random((nz, ny, nx)))).astype(np.float32)
    rho = synth.gardner(vp)
    sc = np.float32([[63.0, 55.0, 42.0]])
    wav = (synth.ricker(nt, dt, 40.0) * 1e8)[None].astype(np.float32)
    gc = np.float32([[20 + 13 * i, 30 + 7 * i, 30.0 + 3 * i] for i in range(6)])
    for k, (order, fs, sf, mt, wsel, ncheck) in enumerate(FWI_E3D_CASES):
        obs = np.stack([runs.ref_e3d_run(vp * np.float32(1.04), vs, rho, lpml, (dx, dx, dx), dt, nt, order, fs, sf,
                                         sc, wav, f, gc)["rec"] for f in ("VX", "VY", "VZ")])
        wts = [(0.5 + rng.random(obs[0].shape)).astype(np.float32) if a in wsel else None for a in "xyz"]
        r = fwi.ref_fwi_e3d(vp, vs, rho, lpml, (dx, dx, dx), dt, nt, order, fs, ncheck, sf, sc, wav, gc, obs,
                            misfit_type=mt, weights=wts)
        extra = {f"wei{a}": w for a, w in zip("xyz", wts) if w is not None}
        save(f"fwi_e3d_case{k}", vp=vp, vs=vs, rho=rho, scoords=sc, wav=wav, gcoords=gc, dobs=obs,
             grid=np.float32([nx, ny, nz, lpml, dx, dt, nt, order, fs, ncheck, mt]), srcfield=np.str_(sf),
             misfit=np.float32(r["misfit"]), vpgrad=r["vpgrad"], vsgrad=r["vsgrad"], rhograd=r["rhograd"],
             dmod=r["dmod"], dres=r["dres"], wavgrad=r["wavgrad"], **extra)


GROUPS["fwi_e3d"] = fwi_e3d_fixtures


if __name__ == "__main__":
    # python -m tests.golden.make_golden [group ...]   (default: every group)
    R = oracle.ref()
    for name in (sys.argv[1:] or list(GROUPS)):
        GROUPS[name](R)

"""The oracle (oracle/rsoracle.c, oracle/revolve.py) against the golden
fixtures written by the compiled reference (tests/golden/make_golden.py).
fp32 outputs must agree bit for bit; integer maps and schedules exactly."""
import numpy as np
import pytest

import oracle
from oracle import revolve as orv
from oracle import runs

from tests.golden.make_golden import A2D_CASES, E2D_CASES, E3D_CASES, FWI_E3D_CASES, REVOLVE_CASES


def test_der_all_orders_and_axes(golden):
    g = golden("der")
    f = g["f"]
    nz, ny, nx = f.shape
    d = g["d"]
    keys = [k for k in g.files if k.startswith("o")]
    assert len(keys) == 32
    for k in keys:
        order, which = int(k[1]), int(k.split("_w")[1])
        axis, fw = which % 3, int(which < 3)
        df = np.zeros_like(f)
        oracle.fn("deriv")(f.ravel(), df.reshape(-1), nx, ny, nz, axis, fw, order, d[axis])
        assert np.array_equal(df, g[k]), k


def test_pml_tables(golden):
    g = golden("pml")
    for k in g.files:
        lpml = int(k[1:].split("_")[0])
        dt = float(k.split("_dt")[1].split("_")[0])
        sign = int(k.split("_s")[1])
        t = runs.pml_tables(lpml, dt, sign=sign)
        assert np.array_equal(t, g[k]), k


def test_stagger(golden):
    g = golden("stagger")
    lpml = int(g["lpml"])
    for fs in (0, 1):
        for got, key in zip(runs.stagger_a2d(g["vp"], g["rho"], lpml, fs), ("L", "Rx", "Rz")):
            assert np.array_equal(got, g[f"a2d_fs{fs}_{key}"]), key
        for got, key in zip(runs.stagger_e2d(g["vp"], g["vs"], g["rho"], lpml, fs),
                            ("L", "L2M", "M", "Rx", "Rz")):
            assert np.array_equal(got, g[f"e2d_fs{fs}_{key}"]), key
        for got, key in zip(runs.stagger_a3d(g["vp3"], g["rho3"], lpml, fs), ("L", "Rx", "Ry", "Rz")):
            assert np.array_equal(got, g[f"a3d_fs{fs}_{key}"]), key


def test_makemap(golden):
    g = golden("makemap")
    for field, name in ((0, "P"), (1, "VX"), (3, "VZ")):
        m, s = runs.makemap2d(g["coords"], 40, 30, 10.0, 10.0, name)
        assert np.array_equal(m, g[f"map_f{field}"])
        ok = m >= 0
        assert np.array_equal(s[ok], g[f"shift_f{field}"][ok])


@pytest.mark.parametrize("case", range(len(A2D_CASES)))
def test_a2d_run(golden, case):
    g = golden(f"a2d_case{case}")
    order, fs, sf, rf = A2D_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    r = runs.a2d_run(g["vp"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, sf, g["scoords"],
                     g["wav"], rf, g["gcoords"], snap_its=(10, 150))
    for k in ("rec", "P", "Vx", "Vz"):
        assert np.array_equal(r[k], g[k]), k
    assert np.array_equal(r["snaps"][0], g["snap10"])
    assert np.array_equal(r["snaps"][1], g["snap150"])


@pytest.mark.parametrize("case", range(len(E2D_CASES)))
def test_e2d_run(golden, case):
    g = golden(f"e2d_case{case}")
    order, fs, sf, rf, recip, nbw = E2D_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    r = runs.e2d_run(g["vp"], g["vs"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, sf,
                     g["scoords"], g["wav"], rf, g["gcoords"], recip, nbw)
    for k in ("rec",) + runs.E2DState.FIELDS:
        assert np.array_equal(r[k], g[k]), k


@pytest.mark.parametrize("case", range(3))
def test_a3d_run(golden, case):
    g = golden(f"a3d_case{case}")
    nx, ny, nz, lpml, dx, dt, nt, order = g["grid"]
    sf, rf = [("P", "P"), ("VY", "VZ"), ("VX", "P")][case]
    r = runs.a3d_run(g["vp"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), int(order), 0, sf,
                     g["scoords"], g["wav"], rf, g["gcoords"])
    for k in ("rec", "P", "Vz"):
        assert np.array_equal(r[k], g[k]), k


@pytest.mark.parametrize("fs", [0, 1])
def test_e3d_stagger(golden, fs):
    g = golden("e3d_stagger")
    out = runs.stagger_e3d(g["vp"], g["vs"], g["rho"], int(g["lpml"]), fs)
    for i, (a, b) in enumerate(zip(out, g[f"stag_fs{fs}"])):
        assert np.array_equal(a, b), ("L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz")[i]


@pytest.mark.parametrize("case", range(len(E3D_CASES)))
def test_e3d_run(golden, case):
    """WavesElastic3D modelling (explosive and force sources, pressure and
    velocity gathers, free surface on and off, orders 2-8) bit for bit."""
    g = golden(f"e3d_case{case}")
    nx, ny, nz, lpml, dx, dt, nt = g["grid"]
    order, fs, sf, rf = E3D_CASES[case]
    r = runs.e3d_run(g["vp"], g["vs"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), order, fs, sf, g["scoords"],
                     g["wav"], rf, g["gcoords"])
    for k in ("rec", "Vx", "Vz", "Szz", "Sxy"):
        assert np.array_equal(r[k], g[k]), k


@pytest.mark.parametrize("case", range(len(FWI_E3D_CASES)))
def test_fwi_e3d_gradient(golden, case):
    """FwiElastic3D::run_optimal restated (oracle/fwi.py fwi_e3d: backward
    steps, crossCorr, scaleGrad, every misfit branch, data weights) vs the
    compiled reference, bit for bit: gradients, records, residuals, the
    wavelet gradient and the misfit."""
    from oracle.fwi import fwi_e3d
    g = golden(f"fwi_e3d_case{case}")
    nx, ny, nz, lpml, dx, dt, nt, order, fs, ncheck, mt = g["grid"]
    wts = [g[f"wei{a}"] if f"wei{a}" in g.files else None for a in "xyz"]
    r = fwi_e3d(g["vp"], g["vs"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), int(order), int(fs), int(ncheck),
                str(g["srcfield"]), g["scoords"], g["wav"], g["gcoords"], g["dobs"], misfit_type=int(mt), weights=wts)
    for k in ("vpgrad", "vsgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(r[k], g[k]), k
        assert np.abs(g[k]).max() > 0, k
    assert np.float32(r["misfit"]) == g["misfit"]


def test_revolve(golden):
    g = golden("revolve")
    for nt, sn in REVOLVE_CASES:
        ref = g[f"n{nt}_s{sn}"]
        got = np.int32(orv.schedule(nt, sn))
        assert np.array_equal(got, ref), (nt, sn)
        assert orv.numforw(nt, sn) == int(g[f"n{nt}_s{sn}_numforw"])
    for nt in (10, 100, 1000, 3000):
        assert orv.adjust(nt) == int(g[f"adjust_{nt}"])


@pytest.mark.parametrize("case", range(2))
def test_fwi_a2d_gradient(golden, case):
    """FwiAcoustic2D::run_optimal restated (oracle/fwi.py) vs the compiled
    reference's gradient, modelled data, residuals, wavelet gradient, misfit."""
    from oracle import fwi
    from tests.golden.make_golden import FWI_CASES
    g = golden(f"fwi_case{case}")
    order, fs, ncheck = FWI_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    r = fwi.fwi_a2d(g["vp"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, ncheck,
                    g["scoords"], g["wav"], g["gcoords"], g["dobs"])
    for k in ("vpgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(r[k], g[k]), k
    assert np.float32(r["misfit"]) == g["misfit"]


@pytest.mark.parametrize("case", range(3))
def test_fwi_a2d_full_gradient(golden, case):
    """FwiAcoustic2D::run with snapshots every snapinc steps, restated
    (oracle/fwi.py), vs the compiled reference: bit-exact, including the
    enddiff offset of the reverse-time snapshot reads (snapinc 3 and 7 do
    not divide nt-1 = 299)."""
    from oracle import fwi
    from tests.golden.make_golden import FWI_FULL_CASES
    g = golden(f"fwi_full_case{case}")
    order, fs, snapinc = FWI_FULL_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    r = fwi.fwi_a2d_full(g["vp"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, snapinc,
                         g["scoords"], g["wav"], g["gcoords"], g["dobs"])
    for k in ("vpgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(r[k], g[k]), k
    assert np.float32(r["misfit"]) == g["misfit"]


@pytest.mark.parametrize("case", range(2))
def test_fwi_e2d_gradient(golden, case):
    """FwiElastic2D::run_optimal restated (oracle/fwi.py: forward and
    adjoint-state elastic steps, 3-component misfit/residuals, crossCorr,
    scaleGrad) vs the compiled reference, bit for bit."""
    from oracle import fwi
    from tests.golden.make_golden import FWI_E2D_CASES
    g = golden(f"fwi_e2d_case{case}")
    order, fs, ncheck, sf = FWI_E2D_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    r = fwi.fwi_e2d(g["vp"], g["vs"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, ncheck, sf,
                    g["scoords"], g["wav"], g["gcoords"], g["dobs"])
    for k in ("vpgrad", "vsgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(r[k], g[k]), k
    assert np.float32(r["misfit"]) == g["misfit"]
    for k, o in (("lam", "oracle_lam"), ("mu", "oracle_mu"), ("rhok", "oracle_rho")):
        assert np.array_equal(r[k], g[o]), k


@pytest.mark.parametrize("case", range(2))
def test_fwi_a3d_gradient(golden, case):
    """FwiAcoustic3D::run_optimal restated (oracle/fwi.py, crossCorr with the
    reference's 3D loop bounds and wrz index) vs the compiled reference."""
    from oracle import fwi
    from tests.golden.make_golden import FWI_A3D_CASES
    g = golden(f"fwi_a3d_case{case}")
    order, fs, ncheck = FWI_A3D_CASES[case]
    nx, ny, nz, lpml, dx, dt, nt = g["grid"]
    r = fwi.fwi_a3d(g["vp"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), order, fs, ncheck, g["scoords"],
                    g["wav"], g["gcoords"], g["dobs"])
    for k in ("vpgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(r[k], g[k]), k
    assert np.float32(r["misfit"]) == g["misfit"]


@pytest.mark.parametrize("case", range(5))
def test_fwi_a3d_misfit_types(golden, case):
    """FwiAcoustic3D::run_optimal with CORRELATION, ADAPTIVE_GAUSS/LINEAR, the
    default branch (type 4: DIFFERENCE) and data weights, restated, vs the
    compiled reference, bit for bit: the 3D class leaves the residuals
    unclipped (the 2D class clips types 1-3 at the 97th percentile, so a
    restatement that clipped would fail here)."""
    from oracle import fwi
    from tests.golden.make_golden import FWI_A3D_MISFIT_CASES
    g = golden("fwi_a3d_case0")
    f = golden(f"fwi_a3d_misfit_case{case}")
    order, fs, ncheck, mt, usew = FWI_A3D_MISFIT_CASES[case]
    nx, ny, nz, lpml, dx, dt, nt = g["grid"]
    r = fwi.fwi_a3d(g["vp"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), order, fs, ncheck, g["scoords"],
                    g["wav"], g["gcoords"], g["dobs"], misfit_type=mt, weight=f["weight"] if usew else None)
    for k in ("vpgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(r[k], f[k]), k
    assert np.float32(r["misfit"]) == f["misfit"]


def test_interp_matches_reference(golden):
    """Interp::interp restated (LINEAR, BSPLINE, SINC) vs the reference."""
    from oracle import fn
    from tests.golden.make_golden import INTERP_CASES
    g = golden("interp")
    for k, (ntr, n1, d1, o1, n2, d2, o2) in enumerate(INTERP_CASES):
        for mode in (0, 1, 2):
            out = np.zeros((ntr, n2), np.float32)
            fn("interp")(mode, g[f"in{k}"].ravel(), ntr, n1, np.float32(d1), np.float32(o1), out.reshape(-1), n2,
                         np.float32(d2), np.float32(o2))
            assert np.array_equal(out, g[f"out{k}_m{mode}"]), (k, mode)


@pytest.mark.parametrize("case", range(5))
def test_fwi_a2d_misfit_types(golden, case):
    """run_optimal with the CORRELATION misfit (per-trace normalised
    correlation, 97th-percentile clip of the residuals), the ADAPTIVE_GAUSS and
    ADAPTIVE_LINEAR misfits (per-trace Wiener shaping filter by Levinson
    recursion, time-weighted) and data weights: the oracle vs the reference,
    bit for bit."""
    from oracle import fwi
    from tests.golden.make_golden import FWI_MISFIT_CASES
    g = golden(f"fwi_misfit_case{case}")
    order, fs, ncheck, mt, usew = FWI_MISFIT_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    r = fwi.fwi_a2d(g["vp"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, ncheck, g["scoords"], g["wav"],
                    g["gcoords"], g["dobs"], misfit_type=mt, weight=g["weight"] if usew else None)
    for k in ("vpgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(r[k], g[k]), k
    assert np.float32(r["misfit"]) == g["misfit"]


def _e2d_misfit_inputs(g, case):
    from tests.golden.make_golden import FWI_E2D_MISFIT_CASES
    order, fs, ncheck, sf, mt, wset = FWI_E2D_MISFIT_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    wts = [g[f"wei_{c}"] if f"wei_{c}" in g.files else None for c in "pxz"]
    return (order, fs, ncheck, sf, mt), (int(nx), int(nz), int(lpml), dx, dt, int(nt)), wts


@pytest.mark.parametrize("case", range(6))
def test_e2d_misfit_residuals_vs_reference(golden, case):
    """FwiElastic2D::computeMisfit / computeResiduals restated (orc_misfit_e,
    orc_residual_e) on the reference's own modelled P / Ux / Uz records: the
    reference's residuals bit for bit (CORRELATION with the combined Ux/Uz
    clip and the separate P clip, ADAPTIVE_GAUSS / LINEAR per component,
    ADAPTIVE_SINGLE with one joint shaper, data weights) and its misfit."""
    g = golden(f"fwi_e2d_misfit_case{case}")
    (order, fs, ncheck, sf, mt), (nx, nz, lpml, dx, dt, nt), wts = _e2d_misfit_inputs(g, case)
    gx, _ = runs.makemap2d(g["gcoords"], nx, nz, dx, dx, "VX")
    valid = ((gx[:, 0] >= 0) & (gx[:, 1] >= 0)).astype(np.int32)
    nrec = g["dmod"].shape[1]
    dmod, dobs = np.ascontiguousarray(g["dmod"]), np.ascontiguousarray(g["dobs"])
    md = [dmod[0], dobs[0], dmod[1], dobs[1], dmod[2], dobs[2]]
    md = [np.ascontiguousarray(a).reshape(-1) for a in md]
    wp = [None if w is None else np.ascontiguousarray(w).ctypes.data for w in wts]
    res = np.zeros_like(dmod)
    oracle.fn("residual_e")(mt, *md, *wp, res[0].reshape(-1), res[1].reshape(-1), res[2].reshape(-1), nrec, nt)
    assert np.array_equal(res, g["dres"])
    mis = oracle.fn("misfit_e")(mt, *md, *wp, valid, nrec, nt)
    assert np.float32(mis) == g["misfit"]


def test_fwi_e2d_adaptive_single_run(golden):
    """The whole elastic run_optimal with ADAPTIVE_SINGLE and Ux / Uz data
    weights (the oracle's fwi_e2d): gradients, records, residuals and wavelet
    gradient equal the reference's bit for bit."""
    from oracle import fwi
    g = golden("fwi_e2d_misfit_case4")
    (order, fs, ncheck, sf, mt), (nx, nz, lpml, dx, dt, nt), wts = _e2d_misfit_inputs(g, 4)
    r = fwi.fwi_e2d(g["vp"], g["vs"], g["rho"], lpml, dx, dx, dt, nt, order, fs, ncheck, sf, g["scoords"],
                    g["wav"], g["gcoords"], g["dobs"], misfit_type=mt, weights=wts)
    for k in ("vpgrad", "vsgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(r[k], g[k]), k
    assert np.float32(r["misfit"]) == g["misfit"]


@pytest.mark.parametrize("case", range(3))
def test_rtm_a2d_image(golden, case):
    """RtmAcoustic2D::run_optimal restated (oracle/rtm.py: the imaging
    condition with subsurface offsets, odd and even offset counts, free
    surface) vs the compiled reference, bit for bit."""
    from oracle import rtm
    from tests.golden.make_golden import RTM_CASES
    g = golden(f"rtm_case{case}")
    order, fs, ncheck, nhx, nhz = RTM_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    img = rtm.rtm_a2d(g["vp"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, ncheck, g["scoords"], g["wav"],
                      g["gcoords"], g["data"], nhx, nhz)
    assert img.shape == g["image"].shape
    assert np.array_equal(img, g["image"])


def test_rtm_image_condition_offsets():
    """The imaging condition at offset (hx, hz) is the product of the source
    field shifted by (-hx, -hz) and the receiver field shifted by (+hx, +hz),
    restricted to where both lie inside the grid: checked on delta fields,
    whose products land on exactly one cell per offset."""
    from oracle import rtm
    nx, nz, lpml, nhx, nhz = 9, 7, 2, 5, 3
    ws = np.zeros((nz + 2 * lpml, nx + 2 * lpml), np.float32)
    wr = np.zeros_like(ws)
    ws[lpml + 3, lpml + 2] = 2.0   # source field at (x 2, z 3)
    wr[lpml + 5, lpml + 6] = 3.0   # receiver field at (x 6, z 5)
    img = np.zeros((nhz, nhx, nz, nx), np.float32)
    rtm.image_a2d(img, ws, wr, lpml, nhx, nhz)
    # ix - hx = 2, ix + hx = 6 -> ix = 4, hx = 2 (ihx 4); iz - hz = 3, iz + hz = 5 -> iz = 4, hz = 1 (ihz 2)
    want = np.zeros_like(img)
    want[2, 4, 4, 4] = 6.0
    assert np.array_equal(img, want)


@pytest.mark.parametrize("case", range(2))
def test_rtm_a3d_image(golden, case):
    """RtmAcoustic3D::run_optimal restated (offsets along x, y and z, free
    surface) vs the compiled reference, bit for bit."""
    from oracle import rtm
    from tests.golden.make_golden import RTM_A3D_CASES
    g = golden("fwi_a3d_case0")
    f = golden(f"rtm_a3d_case{case}")
    order, fs, ncheck, nhx, nhy, nhz = RTM_A3D_CASES[case]
    nx, ny, nz, lpml, dx, dt, nt = g["grid"]
    img = rtm.rtm_a3d(g["vp"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), order, fs, ncheck, g["scoords"],
                      g["wav"], g["gcoords"], g["dobs"], nhx, nhy, nhz)
    assert np.array_equal(img, f["image"])


@pytest.mark.skipif(not __import__("oracle").ref_available(), reason="reference build absent")
def test_reference_window_timers_run():
    """The CPU-baseline timers (ref_a2d_time_window, ref_e2d_time_window:
    the compiled reference stepping from a given state) run on a small grid
    from a state a reference run produced and return a positive time; two
    threads at once also work (the shot-parallel figure)."""
    import concurrent.futures as cf
    from oracle import runs
    rng = np.random.default_rng(3)
    nx, nz, lpml, dx, dt, nt = 40, 30, 8, 10.0, 1e-3, 60
    vp = (2000.0 + 200.0 * rng.random((nz, nx))).astype(np.float32)
    rho = (2000.0 * np.ones((nz, nx))).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    sc = np.float32([[200.0, 20.0]])
    wav = rng.standard_normal((1, nt)).astype(np.float32)
    gc = np.stack([np.arange(nx) * dx, np.full(nx, 20.0)], 1).astype(np.float32)
    nxp, nzp = nx + 2 * lpml, nz + 2 * lpml
    e = runs.ref_e2d_run(vp, vs, rho, lpml, dx, dx, dt, 20, 4, False, "P", sc, wav[:, :20], "VX", gc)
    strips = np.zeros(8 * lpml * nzp + 8 * nxp * lpml, np.float32)
    st = np.concatenate([e[k].ravel() for k in ("Vx", "Vz", "Sxx", "Szz", "Sxz")] + [strips])
    with cf.ThreadPoolExecutor(2) as ex:
        ts = list(ex.map(lambda _: runs.ref_e2d_time_window(vp, vs, rho, lpml, dx, dx, dt, nt, 4, False, sc, wav,
                                                            gc, 20, 10, st), range(2)))
    assert all(t > 0 for t in ts)
    a = runs.ref_a2d_run(vp, rho, lpml, dx, dx, dt, 20, 4, False, "P", sc, wav[:, :20], "P", gc)
    st2 = np.concatenate([a[k].ravel() for k in ("P", "Vx", "Vz")] +
                         [np.zeros(4 * lpml * nzp + 4 * nxp * lpml, np.float32)])
    assert runs.ref_a2d_time_window(vp, rho, lpml, dx, dx, dt, nt, 4, sc, wav, gc, 20, 10, st2) > 0

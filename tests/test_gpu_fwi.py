"""FWI gradient on the HIP path (rsg_fwi_a2d_gradient through
rockseis_amd.fwi) against the compiled reference's FwiAcoustic2D::run_optimal
and FwiAcoustic2D::run (golden fixtures): gradients, modelled data,
residuals, wavelet gradient and misfit within relative L2 1e-5; Revolve and
store-every-step modes agree bit for bit; the torch.autograd.Function
returns the same gradient; FP16 snapshot / checkpoint storage stays within
relative L2 1e-4 to 2e-4 of the FP64 oracle's gradient."""
import numpy as np
import pytest

from tests.golden.make_golden import FWI_CASES, FWI_FULL_CASES
from tests.gpu_util import RTOL_L2, data2d, rel_l2

pytestmark = pytest.mark.gpu


def _run(g, case, mode, fp16=False, snapinc=None):
    from rockseis_amd import core, fwi
    if mode == "full":
        order, fs, si = FWI_FULL_CASES[case]
        snapinc, ncheck = snapinc or si, 0
    else:
        order, fs, ncheck = FWI_CASES[case]
        snapinc = snapinc or 1
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:], m.R[:] = g["vp"], g["rho"]
    src = data2d(g["scoords"], nt, dt, "P", g["wav"])
    src.makeMap(m, core.SMAP)
    obs = data2d(g["gcoords"], nt, dt, "P", g["dobs"])
    obs.makeMap(m, core.GMAP)
    mod = data2d(g["gcoords"], nt, dt, "P")
    res = data2d(g["gcoords"], nt, dt, "P")
    wgr = data2d(g["scoords"], nt, dt, "P")
    m.staggerModels()
    f = fwi.FwiAcoustic2D(m, src, obs, order, snapinc)
    f.setFp16Storage(fp16)
    vg, rg = fwi.Image2D(m), fwi.Image2D(m)
    f.setVpgrad(vg)
    f.setRhograd(rg)
    f.setDatamodP(mod)
    f.setDataresP(res)
    f.setWavgrad(wgr)
    f.setNcheck(ncheck)
    {"revolve": f.run_optimal, "store": f.run_stored, "full": f.run}[mode]()
    assert f.mode_used == mode
    return dict(misfit=f.getMisfit(), misfit64=f.misfit64, vpgrad=vg.imagedata, rhograd=rg.imagedata, dmod=mod.data,
                dres=res.data, wavgrad=wgr.data)


@pytest.mark.parametrize("case", range(len(FWI_CASES)))
def test_fwi_a2d_vs_reference(golden, case):
    g = golden(f"fwi_case{case}")
    print(f"fwi case {case}: {FWI_CASES[case]}", flush=True)
    r = _run(g, case, "revolve")
    for k in ("vpgrad", "rhograd", "dmod", "dres", "wavgrad"):
        e = rel_l2(r[k], g[k])
        print(f"  {k}: rel L2 {e:.2e}")
        assert e <= RTOL_L2, k
    # The reference sums 0.5*r^2 into a float (fwi.cpp:597), which drifts by
    # ~1e-5 relative over 18000 samples; the device reduces in double.  So the
    # device misfit is checked against the double sum of the reference's own
    # residuals, and against the reference's float total at 1e-4.
    exact = 0.5 * np.sum((g["dmod"].astype(np.float64) - g["dobs"].astype(np.float64)) ** 2)
    assert abs(r["misfit64"] - exact) <= RTOL_L2 * exact
    assert abs(r["misfit"] - float(g["misfit"])) <= 1e-4 * float(g["misfit"])
    s = _run(g, case, "store")
    for k in ("vpgrad", "rhograd", "dmod", "dres", "wavgrad"):
        assert np.array_equal(s[k], r[k]), k


@pytest.mark.parametrize("case", range(len(FWI_FULL_CASES)))
def test_fwi_a2d_full_vs_reference(golden, case):
    """FwiAcoustic2D::run: snapshots every snapinc steps held in HBM."""
    g = golden(f"fwi_full_case{case}")
    print(f"fwi full case {case}: {FWI_FULL_CASES[case]}", flush=True)
    r = _run(g, case, "full")
    for k in ("vpgrad", "rhograd", "dmod", "dres", "wavgrad"):
        e = rel_l2(r[k], g[k])
        print(f"  {k}: rel L2 {e:.2e}")
        assert e <= RTOL_L2, k
    exact = 0.5 * np.sum((g["dmod"].astype(np.float64) - g["dobs"].astype(np.float64)) ** 2)
    assert abs(r["misfit64"] - exact) <= RTOL_L2 * exact


def test_fp16_storage_error_vs_fp64_oracle(golden):
    """Gradient error against the FP64 oracle for FP32, FP16 and S16 storage of
    Revolve checkpoints, the store-every-step history and run()'s snapshots
    (order 4, 60x50, nt 300).  FP16 keeps 11 significant bits, so every
    stored value carries a relative rounding error up to 2^-11; the errors
    average out over the correlated time steps.  Measured on MI355X
    (vpgrad / rhograd): FP32 2.7e-6 / 2.6e-6 in every mode; FP16 Revolve
    7.4e-5 / 8.2e-5, FP16 store-every-step 3.9e-5 / 5.5e-5, FP16 run() with
    snapinc 3 8.1e-5 / 1.07e-4 (a third as many snapshots to average over).
    S16 (int16 mantissas, one power-of-two exponent per 256 values) keeps
    about four more bits of a smooth wavefield."""
    from oracle import fwi as ofwi
    g = golden("fwi_case0")
    order, fs, ncheck = FWI_CASES[0]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    args = (g["vp"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs)
    tail = (g["scoords"], g["wav"], g["gcoords"], g["dobs"])
    o_opt = ofwi.fwi_a2d(*args, ncheck, *tail, real=np.float64)
    o_full = ofwi.fwi_a2d_full(*args, 3, *tail, real=np.float64)
    # FP16 bounds: 1.5x the measured figures; S16 (16-bit block floating
    # point): the 1e-4 target
    bounds = {("revolve", 0): 1e-5, ("store", 0): 1e-5, ("full", 0): 1e-5,
              ("revolve", 1): 1.3e-4, ("store", 1): 1e-4, ("full", 1): 1.6e-4,
              ("revolve", 2): 1e-4, ("store", 2): 1e-4, ("full", 2): 1e-4}
    for (mode, fp16), bound in bounds.items():
        if mode == "full":
            r = _run_full_on_case0(g, fp16)
            ref = o_full
        else:
            r = _run(g, 0, mode, fp16=fp16)
            ref = o_opt
        ev = rel_l2(r["vpgrad"], ref["vpgrad"])
        er = rel_l2(r["rhograd"], ref["rhograd"])
        kind = ("fp32", "fp16", "s16")[fp16]
        print(f"  {mode:8s} {kind:4s} vpgrad {ev:.2e} rhograd {er:.2e} (bound {bound:.1e})", flush=True)
        assert ev <= bound and er <= bound, (mode, fp16)


def _run_full_on_case0(g, fp16):
    """run() with snapinc 3 on the fwi_case0 inputs (order 4, no free surface)."""
    from rockseis_amd import core, fwi
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, False)
    m.Vp[:], m.R[:] = g["vp"], g["rho"]
    src = data2d(g["scoords"], nt, dt, "P", g["wav"])
    src.makeMap(m, core.SMAP)
    obs = data2d(g["gcoords"], nt, dt, "P", g["dobs"])
    obs.makeMap(m, core.GMAP)
    m.staggerModels()
    f = fwi.FwiAcoustic2D(m, src, obs, FWI_CASES[0][0], 3)
    f.setFp16Storage(fp16)
    vg, rg = fwi.Image2D(m), fwi.Image2D(m)
    f.setVpgrad(vg)
    f.setRhograd(rg)
    f.run()
    return dict(vpgrad=vg.imagedata, rhograd=rg.imagedata)


@pytest.mark.parametrize("device_inputs", (True, False), ids=("cuda", "cpu"))
def test_autograd_function_matches_driver(golden, device_inputs):
    """AcousticFwiMisfit against the reference fixture, with CUDA inputs
    (model staggered on the device) and CPU inputs (gradients back on the
    CPU)."""
    import torch
    from rockseis_amd import fwi
    g = golden("fwi_case0")
    order, fs, ncheck = FWI_CASES[0]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    survey = fwi.Survey([fwi.Shot(g["scoords"], g["wav"], g["gcoords"], g["dobs"])], int(lpml), dx, dx, dt,
                        order=order, fs=bool(fs), ncheck=ncheck, mode="store")
    dev = "cuda" if device_inputs else "cpu"
    vp = torch.tensor(g["vp"], device=dev, requires_grad=True)
    rho = torch.tensor(g["rho"], device=dev, requires_grad=True)
    f = fwi.misfit(vp, rho, survey)
    f.backward()
    assert vp.grad.device == vp.device and rho.grad.device == rho.device
    exact = 0.5 * np.sum((g["dmod"].astype(np.float64) - g["dobs"].astype(np.float64)) ** 2)
    assert abs(float(f) - exact) <= RTOL_L2 * exact
    assert rel_l2(vp.grad.cpu().numpy(), g["vpgrad"]) <= RTOL_L2
    assert rel_l2(rho.grad.cpu().numpy(), g["rhograd"]) <= RTOL_L2


def test_batched_gradient_equals_sum_of_shots(golden):
    """Three shots in one batch == sum of the three single-shot gradients
    (both checkpoint modes), and the per-shot modelled data agree."""
    from oracle import runs
    from rockseis_amd import fwi, synth
    g = golden("fwi_case0")
    order, fs, ncheck = FWI_CASES[0]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=3)
    shots = []
    for xs in (150.0, 300.0, 450.0):
        sc = np.float32([[xs, 10.0]])
        obs = runs.a2d_run(vtrue, synth.gardner(vtrue), lpml, dx, dx, dt, nt, order, fs, "P", sc,
                           g["wav"], "P", g["gcoords"])["rec"]
        shots.append(fwi.Shot(sc, g["wav"], g["gcoords"], obs))
    for mode in ("store", "revolve"):
        sv = fwi.Survey(shots, lpml, dx, dx, dt, order=order, fs=bool(fs), ncheck=ncheck, mode=mode)
        mis, vg, rg, dmod = fwi.batch_gradient(g["vp"], g["rho"], sv)
        tot, gv, gr = 0.0, 0.0, 0.0
        for k, sh in enumerate(shots):
            m1, a, b = fwi.shot_gradient(g["vp"], g["rho"], sv, sh)
            tot, gv, gr = tot + m1, gv + a, gr + b
        assert abs(mis - tot) <= 1e-6 * tot
        assert rel_l2(vg.cpu().numpy(), gv) <= 1e-6
        assert rel_l2(rg.cpu().numpy(), gr) <= 1e-6


@pytest.mark.parametrize("case", range(5))
def test_misfit_types_vs_reference(golden, case):
    """CORRELATION and ADAPTIVE misfits (97th-percentile residual clip) and
    data weights against the reference's run_optimal."""
    from rockseis_amd import core, fwi
    from tests.golden.make_golden import FWI_MISFIT_CASES
    g = golden(f"fwi_misfit_case{case}")
    order, fs, ncheck, mt, usew = FWI_MISFIT_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:], m.R[:] = g["vp"], g["rho"]
    src = data2d(g["scoords"], nt, dt, "P", g["wav"])
    src.makeMap(m, core.SMAP)
    obs = data2d(g["gcoords"], nt, dt, "P", g["dobs"])
    obs.makeMap(m, core.GMAP)
    res = data2d(g["gcoords"], nt, dt, "P")
    m.staggerModels()
    f = fwi.FwiAcoustic2D(m, src, obs, order, 1)
    vg, rg = fwi.Image2D(m), fwi.Image2D(m)
    f.setVpgrad(vg)
    f.setRhograd(rg)
    f.setDataresP(res)
    f.setNcheck(ncheck)
    f.setMisfit_type(mt)
    if usew:
        f.setDataweight(data2d(g["gcoords"], nt, dt, "P", g["weight"]))
    f.run_optimal()
    for k, got in (("vpgrad", vg.imagedata), ("rhograd", rg.imagedata), ("dres", res.data)):
        e = rel_l2(got, g[k])
        print(f"  case {case} {k}: rel L2 {e:.2e}", flush=True)
        # CORRELATION residuals difference two nearly equal normalised terms:
        # the reference's own fp32 residuals are 1.2e-5 from FP64 in case 1.
        # ADAPTIVE residuals pass the records through two Levinson solves and
        # the 97th-percentile clip: a 1-ulp random perturbation of the
        # reference's own records moves its residuals by 5.2e-5 to 5.9e-5
        # (clip flips), so records agreeing to ~1e-6 bound them at 1e-4.
        bound = 3e-5 if (k == "dres" and mt == 1) else (1e-4 if mt >= 2 else RTOL_L2)
        assert e <= bound, k
    assert abs(f.misfit64 - float(g["misfit"])) <= 1e-4 * abs(float(g["misfit"]))


@pytest.mark.parametrize("case", range(5))
def test_misfit_residuals_bit_exact(golden, case):
    """The device misfit / residual kernels alone (rsg_misfit_residuals) on the
    reference's own modelled records: residuals equal the reference's bit for
    bit for DIFFERENCE, CORRELATION (per-trace sums in the reference's order,
    exact 97th-percentile clip) and ADAPTIVE (the reference's sequential
    convolutions and Levinson recursions, one lane per trace); the misfit
    (reduced in double, the reference sums in float) to 1e-4."""
    from rockseis_amd import core, fwi
    from tests.golden.make_golden import FWI_MISFIT_CASES
    g = golden(f"fwi_misfit_case{case}")
    order, fs, ncheck, mt, usew = FWI_MISFIT_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    m = core.ModelAcoustic2D(int(nx), int(nz), int(lpml), dx, dx, 0.0, 0.0, fs)
    obs = data2d(g["gcoords"], int(nt), dt, "P", g["dobs"])
    obs.makeMap(m, core.GMAP)
    mis, res = fwi.misfit_residuals(mt, g["dmod"], g["dobs"], obs.gmap, g["weight"] if usew else None)
    assert np.array_equal(res, g["dres"])
    assert abs(mis - float(g["misfit"])) <= 1e-4 * abs(float(g["misfit"]))


def test_offgrid_receivers_reproducible_and_vs_oracle(golden):
    """Receivers between grid nodes (x + 0.37 dx, z + 0.21 dz): the 2x2 cells
    of neighbouring receivers overlap with non-zero weights.  The residual
    injection adds each cell's contributions in trace order, so two runs of a
    three-shot batch give bit-identical gradients, and each shot's gradient
    matches the FP32 oracle (trace order, as the reference) within 1e-5."""
    from oracle import fwi as ofwi
    from oracle import runs
    from rockseis_amd import fwi, synth
    g = golden("fwi_case0")
    order, fs, ncheck = FWI_CASES[0][:3]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=9)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    rho0 = synth.gardner(vp0)
    gc = (synth.surface_line(nx - 1, dx, 20.0) + np.float32([[0.37 * dx, 0.21 * dx]])).astype(np.float32)
    shots, refs = [], []
    for sx in (150.0, 300.0, 450.0):
        sc = np.float32([[sx, 20.0]])
        dobs = runs.a2d_run(vtrue, rtrue, lpml, dx, dx, dt, nt, order, fs, "P", sc, g["wav"], "P", gc)["rec"]
        shots.append(fwi.Shot(sc, g["wav"], gc, dobs))
        refs.append(ofwi.fwi_a2d(vp0, rho0, lpml, dx, dx, dt, nt, order, fs, ncheck, sc, g["wav"], gc, dobs))
    sv = fwi.Survey(shots, lpml, dx, dx, dt, order=order, ncheck=ncheck, mode="revolve")
    m1, vg1, rg1, _ = fwi.batch_gradient(vp0, rho0, sv)
    m2, vg2, rg2, _ = fwi.batch_gradient(vp0, rho0, sv)
    assert np.array_equal(vg1.cpu().numpy(), vg2.cpu().numpy())
    assert np.array_equal(rg1.cpu().numpy(), rg2.cpu().numpy())
    want_v = sum(r["vpgrad"].astype(np.float64) for r in refs)
    want_r = sum(r["rhograd"].astype(np.float64) for r in refs)
    ev, er = rel_l2(vg1.cpu().numpy(), want_v), rel_l2(rg1.cpu().numpy(), want_r)
    print(f"  off-grid receivers: vpgrad {ev:.2e} rhograd {er:.2e} vs the oracle's sum of shots")
    assert ev <= RTOL_L2 and er <= RTOL_L2

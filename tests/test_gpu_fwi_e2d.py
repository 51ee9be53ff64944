"""Elastic 2D FWI gradient on the HIP path (rsg_fwi_e2d_gradient_batch
through rockseis_amd.fwi) against the compiled reference's
FwiElastic2D::run_optimal (golden fixtures): Vp / Vs / rho gradients,
modelled P / Ux / Uz records, residuals and wavelet gradient within
relative L2 1e-5; the lambda / mu / rho kernels before scaleGrad against
the oracle's; a ragged batch of windows equals the single-window runs
stacked; FP16 checkpoints against the FP64 oracle."""
import numpy as np
import pytest

from tests.golden.make_golden import FWI_E2D_CASES
from tests.gpu_util import RTOL_L2, data2d, rel_l2

pytestmark = pytest.mark.gpu


def _driver(g, case, fp16=False):
    from rockseis_amd import core, fwi
    order, fs, ncheck, sf = FWI_E2D_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    m = core.ModelElastic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:], m.Vs[:], m.R[:] = g["vp"], g["vs"], g["rho"]
    src = data2d(g["scoords"], nt, dt, sf, g["wav"])
    obs = [data2d(g["gcoords"], nt, dt, f, g["dobs"][c]) for c, f in enumerate(("P", "VX", "VZ"))]
    mod = [data2d(g["gcoords"], nt, dt, f) for f in ("P", "VX", "VZ")]
    res = [data2d(g["gcoords"], nt, dt, f) for f in ("P", "VX", "VZ")]
    wgr = data2d(g["scoords"], nt, dt, sf)
    f = fwi.FwiElastic2D(m, src, obs[0], obs[1], obs[2], order, 1)
    imgs = [fwi.Image2D(m) for _ in range(3)]
    f.setVpgrad(imgs[0])
    f.setVsgrad(imgs[1])
    f.setRhograd(imgs[2])
    f.setDatamodP(mod[0])
    f.setDatamodUx(mod[1])
    f.setDatamodUz(mod[2])
    f.setDataresP(res[0])
    f.setDataresUx(res[1])
    f.setDataresUz(res[2])
    f.setWavgrad(wgr)
    f.setNcheck(ncheck)
    f.setFp16Storage(fp16)
    f.run_optimal()
    return dict(misfit64=f.misfit64, vpgrad=imgs[0].imagedata, vsgrad=imgs[1].imagedata,
                rhograd=imgs[2].imagedata, dmod=np.stack([d.data for d in mod]),
                dres=np.stack([d.data for d in res]), wavgrad=wgr.data)


@pytest.mark.parametrize("case", range(len(FWI_E2D_CASES)))
def test_fwi_e2d_vs_reference(golden, case):
    g = golden(f"fwi_e2d_case{case}")
    print(f"fwi e2d case {case}: {FWI_E2D_CASES[case]}", flush=True)
    r = _driver(g, case)
    for k in ("vpgrad", "vsgrad", "rhograd", "dmod", "dres", "wavgrad"):
        e = rel_l2(r[k], g[k])
        print(f"  {k}: rel L2 {e:.2e}")
        # the wavelet gradient sums the adjoint stresses at the source; the
        # reference's own fp32 value is 4.4e-5 from the FP64 oracle in case 0,
        # so it is held to the gradient bound, 1e-4
        assert e <= (1e-4 if k == "wavgrad" else RTOL_L2), k
    # misfit: traces whose Ux map is valid (fwi.cpp:2911), the reference sums in float
    assert abs(r["misfit64"] - float(g["misfit"])) <= 1e-4 * float(g["misfit"])


def _survey_of(g, case, x0=0, nxw=None, fp16=False):
    from rockseis_amd import fwi
    order, fs, ncheck, sf = FWI_E2D_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    shot = fwi.ElasticShot(g["scoords"], g["wav"], g["gcoords"], g["dobs"], x0, nxw, sf)
    return fwi.ElasticSurvey([shot], int(lpml), dx, dx, dt, order, bool(fs), ncheck, fp16)


@pytest.mark.parametrize("case", range(len(FWI_E2D_CASES)))
def test_lambda_mu_rho_kernels_vs_oracle(golden, case):
    """scale=False: the kernels before scaleGrad (config-5 style output)."""
    from rockseis_amd import fwi
    g = golden(f"fwi_e2d_case{case}")
    mis, lam, mu, rho, _ = fwi.elastic_batch_gradient(g["vp"], g["vs"], g["rho"], _survey_of(g, case), scale=False)
    for name, got, want in (("lam", lam, g["oracle_lam"]), ("mu", mu, g["oracle_mu"]), ("rho", rho, g["oracle_rho"])):
        e = rel_l2(got.cpu().numpy(), want)
        print(f"  {name}: rel L2 {e:.2e}")
        assert e <= RTOL_L2, name


def test_ragged_windows_equal_single_windows(golden):
    """Three shots on windows of different widths of one model in one ragged
    batch == the three single-window runs, stacked at their offsets."""
    from oracle import runs
    from rockseis_amd import fwi, synth
    g = golden("fwi_e2d_case1")
    order, fs, ncheck, sf = FWI_E2D_CASES[1]
    _, nz, lpml, dx, dt, nt = g["grid"]
    nz, lpml, nt = int(nz), int(lpml), int(nt)
    nx = 150
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=8, seed=5)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    vs0, rho0 = (vp0 / 1.8).astype(np.float32), synth.gardner(vp0)
    wins = [(0, 60), (40, 90), (85, 65)]
    shots = []
    for x0, w in wins:
        sc = np.float32([[(x0 + w // 2) * dx, 20.0]])
        gc = synth.surface_line(w, dx, 20.0, x0 * dx)
        wv, ws, wr = (a[:, x0:x0 + w] for a in (vtrue, (vtrue / 1.8).astype(np.float32), synth.gardner(vtrue)))
        # observed data on the window of the true model (coordinates relative to the window)
        gcl = gc - np.float32([[x0 * dx, 0.0]])
        scl = sc - np.float32([[x0 * dx, 0.0]])
        obs = np.stack([runs.e2d_run(wv, ws, wr, lpml, dx, dx, dt, nt, order, fs, sf, scl, g["wav"], f, gcl)["rec"]
                        for f in ("P", "VX", "VZ")])
        shots.append(fwi.ElasticShot(sc, g["wav"], gc, obs, x0, w, sf))
    sv = fwi.ElasticSurvey(shots, lpml, dx, dx, dt, order, bool(fs), ncheck)
    mis, g0, g1, g2, outs = fwi.elastic_batch_gradient(vp0, vs0, rho0, sv, scale=True)
    tot, s0, s1, s2 = 0.0, 0.0, 0.0, 0.0
    for k, sh in enumerate(shots):
        one = fwi.ElasticSurvey([sh], lpml, dx, dx, dt, order, bool(fs), ncheck)
        m1, a, b, c, o1 = fwi.elastic_batch_gradient(vp0, vs0, rho0, one, scale=True)
        tot, s0, s1, s2 = tot + m1, s0 + a, s1 + b, s2 + c
        for comp in range(3):
            assert np.array_equal(o1[0]["dmod"][comp].cpu().numpy(), outs[k]["dmod"][comp].cpu().numpy())
    assert abs(mis - tot) <= 1e-12 * abs(tot)
    for got, want in ((g0, s0), (g1, s1), (g2, s2)):
        assert rel_l2(got.cpu().numpy(), want.cpu().numpy()) <= 1e-6


def test_fp16_checkpoints_error_vs_fp64_oracle(golden):
    """Vp / Vs / rho gradient error against the FP64 oracle with FP32, FP16
    and S16 Revolve checkpoints (order 8, free surface, Uz source)."""
    from oracle import fwi as ofwi
    g = golden("fwi_e2d_case1")
    order, fs, ncheck, sf = FWI_E2D_CASES[1]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    o64 = ofwi.fwi_e2d(g["vp"], g["vs"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, ncheck, sf,
                       g["scoords"], g["wav"], g["gcoords"], g["dobs"], real=np.float64)
    # FP16: 1.5x the measured 3.04e-4; S16 (16-bit block floating point): the 1e-4 target
    for fp16, bound in ((0, 1e-5), (1, 4.6e-4), (2, 1e-4)):
        r = _driver(g, 1, fp16=fp16)
        errs = [rel_l2(r[k], o64[k]) for k in ("vpgrad", "vsgrad", "rhograd")]
        print(f"  {('fp32', 'fp16', 's16')[fp16]}: vp {errs[0]:.2e} vs {errs[1]:.2e} rho {errs[2]:.2e} "
              f"(bound {bound:.1e})", flush=True)
        assert max(errs) <= bound, fp16


def _driver_misfit(g, case):
    """FwiElastic2D with setMisfit_type and setDataweightp/x/z."""
    from rockseis_amd import core, fwi
    from tests.golden.make_golden import FWI_E2D_MISFIT_CASES
    order, fs, ncheck, sf, mt, wset = FWI_E2D_MISFIT_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    m = core.ModelElastic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:], m.Vs[:], m.R[:] = g["vp"], g["vs"], g["rho"]
    src = data2d(g["scoords"], nt, dt, sf, g["wav"])
    comps = ("P", "VX", "VZ")
    obs = [data2d(g["gcoords"], nt, dt, f, g["dobs"][c]) for c, f in enumerate(comps)]
    mod = [data2d(g["gcoords"], nt, dt, f) for f in comps]
    res = [data2d(g["gcoords"], nt, dt, f) for f in comps]
    wgr = data2d(g["scoords"], nt, dt, sf)
    f = fwi.FwiElastic2D(m, src, obs[0], obs[1], obs[2], order, 1)
    imgs = [fwi.Image2D(m) for _ in range(3)]
    f.setVpgrad(imgs[0])
    f.setVsgrad(imgs[1])
    f.setRhograd(imgs[2])
    for c, (setm, setr) in enumerate(((f.setDatamodP, f.setDataresP), (f.setDatamodUx, f.setDataresUx),
                                      (f.setDatamodUz, f.setDataresUz))):
        setm(mod[c])
        setr(res[c])
    for c, setw in zip("pxz", (f.setDataweightp, f.setDataweightx, f.setDataweightz)):
        if f"wei_{c}" in g.files:
            setw(data2d(g["gcoords"], nt, dt, comps["pxz".index(c)], g[f"wei_{c}"]))
    f.setWavgrad(wgr)
    f.setNcheck(ncheck)
    f.setMisfit_type(mt)
    f.run_optimal()
    return dict(misfit64=f.misfit64, vpgrad=imgs[0].imagedata, vsgrad=imgs[1].imagedata,
                rhograd=imgs[2].imagedata, dmod=np.stack([d.data for d in mod]),
                dres=np.stack([d.data for d in res]), wavgrad=wgr.data)


@pytest.mark.parametrize("case", range(6))
def test_fwi_e2d_misfit_types_vs_reference(golden, case):
    """run_optimal with the CORRELATION, ADAPTIVE_GAUSS, ADAPTIVE_LINEAR and
    ADAPTIVE_SINGLE misfits and P / Ux / Uz data weights:

    * modelled records against the compiled reference within 1e-5;
    * the residuals the run computed equal, bit for bit, the restated
      computeResiduals (the oracle, itself bit-exact with the reference) on
      the run's own modelled records;
    * the gradients against the FP32 oracle's adjoint pass driven by those
      same residuals within 1e-5 (the wavelet gradient 1e-4, as for
      DIFFERENCE), and the misfit against the reference within 1e-4.

    Against the reference's gradients themselves the percentile clip (types
    1-3) and the Levinson solves make the residuals discontinuous / ill
    conditioned in the records: the FP64 oracle, the reference's algorithm,
    differs from the FP32 reference by up to 4.7e-3 on these cases, so the
    distance is printed and held to 1e-2 only."""
    import oracle
    from oracle import fwi as ofwi
    from oracle import runs
    from tests.golden.make_golden import FWI_E2D_MISFIT_CASES
    g = golden(f"fwi_e2d_misfit_case{case}")
    order, fs, ncheck, sf, mt, wset = FWI_E2D_MISFIT_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    print(f"fwi e2d misfit case {case}: {FWI_E2D_MISFIT_CASES[case]}", flush=True)
    r = _driver_misfit(g, case)
    # records: within 1e-5, or 3x the FP64 restatement's own distance from the
    # FP32 reference where that is larger (8e-6 in case 4, a VZ source)
    d64 = np.stack([runs.e2d_run(g["vp"], g["vs"], g["rho"], lpml, dx, dx, dt, nt, order, fs, sf, g["scoords"],
                                 g["wav"], f, g["gcoords"], real=np.float64)["rec"] for f in ("P", "VX", "VZ")])
    e, e64 = rel_l2(r["dmod"], g["dmod"]), rel_l2(d64, g["dmod"])
    print(f"  dmod vs reference: rel L2 {e:.2e} (FP64 restatement {e64:.2e})")
    assert e <= max(RTOL_L2, 3 * e64)
    # residuals of the run == computeResiduals restated, on the run's records
    wts = [g[f"wei_{c}"] if f"wei_{c}" in g.files else None for c in "pxz"]
    dmod = np.ascontiguousarray(r["dmod"], np.float32)
    dobs = np.ascontiguousarray(g["dobs"], np.float32)
    md = [np.ascontiguousarray(a).reshape(-1) for a in (dmod[0], dobs[0], dmod[1], dobs[1], dmod[2], dobs[2])]
    wp = [None if w is None else np.ascontiguousarray(w, np.float32).ctypes.data for w in wts]
    want = np.zeros_like(dmod)
    oracle.fn("residual_e")(mt, *md, *wp, want[0].reshape(-1), want[1].reshape(-1), want[2].reshape(-1),
                            dmod.shape[1], nt)
    assert np.array_equal(r["dres"], want)
    gx, _ = runs.makemap2d(g["gcoords"], nx, nz, dx, dx, "VX")
    valid = ((gx[:, 0] >= 0) & (gx[:, 1] >= 0)).astype(np.int32)
    om = oracle.fn("misfit_e")(mt, *md, *wp, valid, dmod.shape[1], nt)
    print(f"  misfit {r['misfit64']:.9e}, restated on the run's records {om:.9e}, reference {float(g['misfit']):.9e}")
    # the device reduces in double, the reference (and its restatement) sums in float
    assert abs(r["misfit64"] - om) <= 1e-4 * abs(om)
    assert abs(r["misfit64"] - float(g["misfit"])) <= 1e-4 * abs(float(g["misfit"]))
    # the adjoint pass on those residuals
    o = ofwi.fwi_e2d(g["vp"], g["vs"], g["rho"], lpml, dx, dx, dt, nt, order, fs, ncheck, sf, g["scoords"],
                     g["wav"], g["gcoords"], g["dobs"], misfit_type=mt, weights=wts, residuals=r["dres"])
    for k in ("vpgrad", "vsgrad", "rhograd", "wavgrad"):
        eo, er = rel_l2(r[k], o[k]), rel_l2(r[k], g[k])
        print(f"  {k}: rel L2 {eo:.2e} vs the oracle on the same residuals, {er:.2e} vs the reference")
        assert eo <= (1e-4 if k == "wavgrad" else RTOL_L2), k
        assert er <= 1e-2, k


@pytest.mark.parametrize("case", range(6))
def test_e2d_misfit_residuals_bit_exact(golden, case):
    """The device elastic misfit / residual kernels alone
    (rsg_misfit_residuals_e2d) on the reference's own modelled records:
    residuals equal the reference's bit for bit for every type (per-trace
    sums in the reference's order, exact per-shot percentile clips, the
    sequential Levinson recursions one lane per trace); the misfit (reduced
    in double, the reference sums its 15000 terms in float: 1.1e-5 apart in
    case 5) to 1e-4."""
    from rockseis_amd import core, fwi
    from tests.golden.make_golden import FWI_E2D_MISFIT_CASES
    g = golden(f"fwi_e2d_misfit_case{case}")
    order, fs, ncheck, sf, mt, wset = FWI_E2D_MISFIT_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    m = core.ModelElastic2D(int(nx), int(nz), int(lpml), dx, dx, 0.0, 0.0, fs)
    ux = data2d(g["gcoords"], int(nt), dt, "VX", g["dobs"][1])
    ux.makeMap(m, core.GMAP)
    wts = [g[f"wei_{c}"] if f"wei_{c}" in g.files else None for c in "pxz"]
    mis, res = fwi.elastic_misfit_residuals(mt, g["dmod"], g["dobs"], ux.gmap, wts)
    for c in range(3):
        assert np.array_equal(res[c], g["dres"][c]), c
    em = abs(mis - float(g["misfit"])) / abs(float(g["misfit"]))
    print(f"  misfit {mis:.9e} vs reference {float(g['misfit']):.9e} (rel {em:.1e})")
    assert em <= 1e-4


@pytest.mark.parametrize("fp16,misfit_type", [(True, 1), (False, 0)])
def test_plan_replays_equal_eager(golden, fp16, misfit_type):
    """The persistent plan (rsg_fwi_e2d_plan_*: the Revolve sweep captured as
    two hipGraphs) gives the eager batch's records, residuals and gradients
    bit for bit on three ragged windows; a second replay gives them again;
    and after set_model (the model restaggered in place) a replay equals the
    eager batch on the new model.  Bit for bit because every sum of the path
    has a fixed order (the residual injection adds each cell's contributions
    in trace order, although neighbouring staggered receivers share corner
    cells).  Misfits to 1e-12 (double atomics)."""
    from oracle import runs
    from rockseis_amd import fwi, synth
    g = golden("fwi_e2d_case1")
    order, fs, ncheck, sf = FWI_E2D_CASES[1]
    _, nz, lpml, dx, dt, nt = g["grid"]
    nz, lpml, nt = int(nz), int(lpml), 160
    nx = 150
    wav = np.ascontiguousarray(g["wav"][:, :nt])
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=8, seed=5)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    vs0, rho0 = (vp0 / 1.8).astype(np.float32), synth.gardner(vp0)
    shots = []
    for x0, w in ((0, 60), (40, 90), (85, 65)):
        sc = np.float32([[(x0 + w // 2) * dx, 20.0]])
        gc = synth.surface_line(w, dx, 20.0, x0 * dx)
        wv, ws, wr = (a[:, x0:x0 + w] for a in (vtrue, (vtrue / 1.8).astype(np.float32), synth.gardner(vtrue)))
        obs = np.stack([runs.e2d_run(wv, ws, wr, lpml, dx, dx, dt, nt, order, fs, sf,
                                     sc - np.float32([[x0 * dx, 0.0]]), wav, f,
                                     gc - np.float32([[x0 * dx, 0.0]]))["rec"] for f in ("P", "VX", "VZ")])
        shots.append(fwi.ElasticShot(sc, wav, gc, obs, x0, w, sf))
    sv = fwi.ElasticSurvey(shots, lpml, dx, dx, dt, order, bool(fs), ncheck, fp16, misfit_type=misfit_type)

    def same(a, b):
        ma, ga, oa = a[0], a[1:4], a[4]
        mb, gb, ob = b[0], b[1:4], b[4]
        assert abs(ma - mb) <= 1e-12 * abs(mb)
        for x, y in zip(ga, gb):
            assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())
        for p, q in zip(oa, ob):
            for key in ("dmod", "dres"):
                for c in range(3):
                    assert np.array_equal(p[key][c].cpu().numpy(), q[key][c].cpu().numpy()), key

    def snap(r):
        return (r[0], *[t.clone() for t in r[1:4]],
                [{k: [t.clone() for t in v] for k, v in o.items()} for o in r[4]])

    eager = fwi.elastic_batch_gradient(vp0, vs0, rho0, sv, scale=True)
    plan = fwi.ElasticFwiPlan(vp0, vs0, rho0, sv, scale=True)
    try:
        first = snap(plan.run())
        same(first, eager)
        same(snap(plan.run()), eager)
        vp1 = (vp0 * 1.02).astype(np.float32)
        vs1, rho1 = (vp1 / 1.8).astype(np.float32), synth.gardner(vp1)
        plan.set_model(vp1, vs1, rho1)
        same(snap(plan.run()), fwi.elastic_batch_gradient(vp1, vs1, rho1, sv, scale=True))
        print(f"  plan == eager (fp16 {fp16}, misfit type {misfit_type}), misfit {first[0]:.6e}")
    finally:
        plan.close()


@pytest.mark.parametrize("misfit_type", range(5))
def test_e2d_misfit_residuals_several_shots(golden, misfit_type):
    """rsg_misfit_residuals_e2d on three shots at once with 60, 37 and 51
    traces (the reference's records of two fixtures, cut and reordered):
    every shot's residuals equal, bit for bit, the restated computeResiduals
    of that shot alone (the percentile clips are per shot), and the misfit is
    the sum of the shots' misfits."""
    import oracle
    from rockseis_amd import fwi
    ga, gb = golden("fwi_e2d_misfit_case0"), golden("fwi_e2d_misfit_case1")
    parts = [(ga, slice(0, 60)), (gb, slice(5, 42)), (ga, slice(9, 60))]
    mods, obss, maps, wts = [], [], [], []
    for g, sl in parts:
        mods.append(np.ascontiguousarray(g["dmod"][:, sl]))
        obss.append(np.ascontiguousarray(g["dobs"][:, sl]))
        mp = np.zeros((mods[-1].shape[1], 2), np.int32)
        mp[::7, 0] = -1  # a few traces outside the model: they count in no misfit
        maps.append(mp)
        wts.append([g[f"wei_{c}"][sl] if f"wei_{c}" in g.files else None for c in "pxz"])
    mis, res = fwi.elastic_misfit_residuals(misfit_type, mods, obss, maps, wts)
    total = 0.0
    nt = mods[0].shape[2]
    for k in range(3):
        m, o = mods[k], obss[k]
        md = [np.ascontiguousarray(a).reshape(-1) for a in (m[0], o[0], m[1], o[1], m[2], o[2])]
        wk = [None if w is None else np.ascontiguousarray(w, np.float32) for w in wts[k]]
        wp = [None if w is None else w.ctypes.data for w in wk]
        want = np.zeros_like(m)
        oracle.fn("residual_e")(misfit_type, *md, *wp, want[0].reshape(-1), want[1].reshape(-1),
                                want[2].reshape(-1), m.shape[1], nt)
        assert np.array_equal(res[k], want), k
        valid = (maps[k][:, 0] >= 0).astype(np.int32)
        total += float(oracle.fn("misfit_e")(misfit_type, *md, *wp, valid, m.shape[1], nt))
    print(f"  misfit {mis:.9e} vs sum of the shots {total:.9e}")
    assert abs(mis - total) <= 1e-4 * abs(total)


@pytest.mark.parametrize("device_inputs", (True, False), ids=("cuda", "cpu"))
def test_elastic_autograd_vs_reference(golden, device_inputs):
    """ElasticFwiMisfit (Vp, Vs, rho) -> misfit on the reference fixture:
    the misfit and the backward pass's Vp / Vs / rho gradients against the
    compiled reference's FwiElastic2D::run_optimal, with CUDA inputs (the
    model staggered on the device) and with CPU inputs (gradients returned
    on the CPU, as autograd requires)."""
    import torch
    from rockseis_amd import fwi
    case = 0
    g = golden(f"fwi_e2d_case{case}")
    dev = "cuda" if device_inputs else "cpu"
    vp, vs, rho = (torch.tensor(np.asarray(g[k]), device=dev, requires_grad=True) for k in ("vp", "vs", "rho"))
    m = fwi.elastic_misfit(vp, vs, rho, _survey_of(g, case))
    m.backward()
    for t in (vp, vs, rho):
        assert t.grad.device == t.device
    for name, t, want in (("vp", vp, g["vpgrad"]), ("vs", vs, g["vsgrad"]), ("rho", rho, g["rhograd"])):
        e = rel_l2(t.grad.detach().cpu().numpy(), want)
        print(f"elastic autograd ({dev} inputs) {name}: rel L2 {e:.2e}", flush=True)
        assert e <= RTOL_L2, name
    assert abs(float(m) - float(g["misfit"])) <= 1e-4 * float(g["misfit"])


def test_device_staggering_matches_host():
    """rsg_stagger_a2d_dev / rsg_stagger_e2d_dev equal the host set-up bit
    for bit: whole models and a window of a wider model (row stride larger
    than the window), with and without free surface."""
    import torch
    from rockseis_amd import core, synth
    nxw, nz, lpml, x0 = 97, 61, 9, 13
    full = synth.gradient_anomaly_vp(150, nz, 5.0, 5.0, 1500, 4000, nanom=6, seed=3)
    vs_full = (full / 1.8).astype(np.float32)
    rho_full = synth.gardner(full)
    for fs in (False, True):
        for elastic in (False, True):
            cls = core.ModelElastic2D if elastic else core.ModelAcoustic2D
            host, devm = cls(nxw, nz, lpml, 5.0, 5.0, 0.0, 0.0, fs), cls(nxw, nz, lpml, 5.0, 5.0, 0.0, 0.0, fs)
            host.Vp[:], host.R[:] = full[:, x0:x0 + nxw], rho_full[:, x0:x0 + nxw]
            names = ("L", "Rx", "Rz")
            args = [torch.tensor(full, device="cuda")[:, x0:x0 + nxw]]
            if elastic:
                host.Vs[:] = vs_full[:, x0:x0 + nxw]
                names = ("L", "L2M", "M", "Rx", "Rz")
                args.append(torch.tensor(vs_full, device="cuda")[:, x0:x0 + nxw])
            args.append(torch.tensor(rho_full, device="cuda")[:, x0:x0 + nxw])
            host.staggerModels()
            devm.staggerModelsDevice(*args)
            torch.cuda.synchronize()
            for n in names:
                a = getattr(host, "_d" + n).cpu().numpy()
                b = getattr(devm, "_d" + n).cpu().numpy()
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (fs, elastic, n)
    print("device staggering: a2d and e2d, fs off/on, window of a wider model: bit-identical", flush=True)

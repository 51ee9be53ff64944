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
    """Vp / Vs / rho gradient error against the FP64 oracle with FP32 and
    FP16 Revolve checkpoints (order 8, free surface, Uz source)."""
    from oracle import fwi as ofwi
    g = golden("fwi_e2d_case1")
    order, fs, ncheck, sf = FWI_E2D_CASES[1]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    o64 = ofwi.fwi_e2d(g["vp"], g["vs"], g["rho"], int(lpml), dx, dx, dt, int(nt), order, fs, ncheck, sf,
                       g["scoords"], g["wav"], g["gcoords"], g["dobs"], real=np.float64)
    for fp16, bound in ((False, 1e-5), (True, 1e-3)):
        r = _driver(g, 1, fp16=fp16)
        errs = [rel_l2(r[k], o64[k]) for k in ("vpgrad", "vsgrad", "rhograd")]
        print(f"  fp16={fp16}: vp {errs[0]:.2e} vs {errs[1]:.2e} rho {errs[2]:.2e} (bound {bound:.0e})", flush=True)
        assert max(errs) <= bound, fp16

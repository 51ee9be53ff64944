"""FwiElastic3D::run_optimal on the HIP path (rsg_fwi_e3d_gradient through
fwi.FwiElastic3D) and the adjoint-state steps rsg_e3d_backward_velocity /
_stress:

* against the compiled reference's gradients (golden fixtures: orders 2-8,
  free surface, P / VX / VY / VZ sources, DIFFERENCE, CORRELATION,
  ADAPTIVE_GAUSS and the default branch, data weights): records within
  relative L2 1e-5; residuals, Vp / Vs / rho and wavelet gradients within
  the larger of 1e-5 and 3x the FP32 reference's own error against the FP64
  oracle (the oracle reproduces every fixture bit for bit);
* the backward steps against the oracle's restatement over a random state;
* Revolve with few snapshots bit-identical to one without recomputation;
* S16 checkpoints against the FP64 oracle within the larger of 1e-4 and
  1.5x the FP32-checkpoint error; the error paths."""
import numpy as np
import pytest

from tests.golden.make_golden import FWI_E3D_CASES
from tests.gpu_util import data3d, rel_l2

pytestmark = pytest.mark.gpu


def _gpu(g, ncheck=None, fp16=False, srcfield=None, cut_uy=False):
    from rockseis_amd import core, fwi
    nx, ny, nz, lpml, dx, dt, nt, order, fs, nchk, mt = g["grid"]
    nt = int(nt)
    sf = srcfield or str(g["srcfield"])
    m = core.ModelElastic3D(int(nx), int(ny), int(nz), int(lpml), dx, dx, dx, 0.0, 0.0, 0.0, bool(fs))
    m.Vp[:], m.Vs[:], m.R[:] = g["vp"], g["vs"], g["rho"]
    src = data3d(g["scoords"], nt, dt, sf, g["wav"])
    src.makeMap(m, core.SMAP)
    obs = []
    for c, f in enumerate(("VX", "VY", "VZ")):
        n = 3 if cut_uy and c == 1 else len(g["gcoords"])
        d = data3d(g["gcoords"][:n], nt, dt, f, g["dobs"][c][:n])
        d.makeMap(m, core.GMAP)
        obs.append(d)
    m.staggerModels()
    fw = fwi.FwiElastic3D(m, src, *obs, int(order), 1)
    imgs = [fwi.Image3D(m) for _ in range(3)]
    fw.setVpgrad(imgs[0])
    fw.setVsgrad(imgs[1])
    fw.setRhograd(imgs[2])
    mods = [data3d(g["gcoords"], nt, dt, f) for f in ("VX", "VY", "VZ")]
    ress = [data3d(g["gcoords"], nt, dt, f) for f in ("VX", "VY", "VZ")]
    fw.setDatamodUx(mods[0])
    fw.setDatamodUy(mods[1])
    fw.setDatamodUz(mods[2])
    fw.setDataresUx(ress[0])
    fw.setDataresUy(ress[1])
    fw.setDataresUz(ress[2])
    for a, setw in zip("xyz", (fw.setDataweightx, fw.setDataweighty, fw.setDataweightz)):
        if f"wei{a}" in g.files:
            setw(data3d(g["gcoords"], nt, dt, "V" + a.upper(), g[f"wei{a}"]))
    wg = data3d(g["scoords"], nt, dt, sf)
    fw.setWavgrad(wg)
    fw.setNcheck(int(nchk) if ncheck is None else ncheck)
    fw.setFp16Storage(fp16)
    fw.setMisfit_type(int(mt))
    fw.run_optimal()
    return dict(misfit=fw.misfit64, vpgrad=imgs[0].imagedata.reshape(g["vp"].shape).copy(),
                vsgrad=imgs[1].imagedata.reshape(g["vp"].shape).copy(),
                rhograd=imgs[2].imagedata.reshape(g["vp"].shape).copy(),
                dmod=np.stack([d.data for d in mods]), dres=np.stack([d.data for d in ress]), wavgrad=wg.data.copy())


@pytest.mark.parametrize("case", range(len(FWI_E3D_CASES)))
def test_fwi_e3d_vs_reference(golden, case):
    """Records within 1e-5 of the reference's; residuals and gradients within
    the larger of 1e-5 and 3x the reference's own FP32 error against the
    FP64 oracle (CORRELATION and ADAPTIVE_GAUSS residuals divide small
    differences of records: there the FP32 reference itself is 2e-5 - 4e-5
    from FP64)."""
    from oracle.fwi import fwi_e3d
    order, fs, sf, mt, wsel, ncheck = FWI_E3D_CASES[case]
    print(f"fwi_e3d case {case}: order {order} fs {fs} source {sf} misfit {mt} weights '{wsel}' ncheck {ncheck}",
          flush=True)
    g = golden(f"fwi_e3d_case{case}")
    r = _gpu(g)
    nx, ny, nz, lpml, dx, dt, nt, _, _, _, _ = g["grid"]
    wts = [g[f"wei{a}"] if f"wei{a}" in g.files else None for a in "xyz"]
    r64 = fwi_e3d(g["vp"], g["vs"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), order, fs, ncheck, sf,
                  g["scoords"], g["wav"], g["gcoords"], g["dobs"], real=np.float64, misfit_type=mt, weights=wts)
    for k in ("dmod", "dres", "vpgrad", "vsgrad", "rhograd", "wavgrad"):
        e, e32, e64 = rel_l2(r[k], g[k]), rel_l2(g[k], r64[k]), rel_l2(r[k], r64[k])
        bound = 1e-5 if k == "dmod" else max(1e-5, 3 * e32)
        print(f"  {k}: rel L2 {e:.2e} vs the reference (bound {bound:.1e}); vs FP64 {e64:.2e}, reference vs FP64 "
              f"{e32:.2e}")
        assert e <= bound, k
    m32, m64 = float(g["misfit"]), float(r64["misfit"])
    print(f"  misfit {r['misfit']:.9e} reference {m32:.9e} FP64 {m64:.9e}")
    assert abs(r["misfit"] - m32) <= max(1e-5, 3 * abs(m32 - m64) / abs(m64)) * abs(m32)


def test_e3d_backward_steps_vs_oracle():
    """Three backward velocity + stress steps from a random state (fields and
    strips) against the oracle's backwardstepVelocity / Stress, order 8, free
    surface on."""
    import torch
    from oracle import runs
    from rockseis_amd import core
    rng = np.random.default_rng(31)
    nx, ny, nz, lpml, dx, dt = 20, 18, 17, 5, 10.0, 1e-3
    vp = (2400 + 500 * rng.random((nz, ny, nx))).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = (1900 + 300 * rng.random((nz, ny, nx))).astype(np.float32)
    m = core.ModelElastic3D(nx, ny, nz, lpml, dx, dx, dx, 0.0, 0.0, 0.0, True)
    m.Vp[:], m.Vs[:], m.R[:] = vp, vs, rho
    m.staggerModels()
    w = core.WavesElastic3D(m, 10, dt)
    w.getPml().setSmax(core.SMAX)
    w.getPml().computeABC()
    o = runs.E3DState(runs.stagger_e3d(vp, vs, rho, lpml, True), lpml, 8, (dx, dx, dx), dt, True)
    for k in runs.E3D_FIELDS:
        a = (rng.standard_normal(getattr(o, k).shape) * 1e-3).astype(np.float32)
        getattr(o, k)[...] = a
        getattr(w, k).copy_(torch.from_numpy(a))
    for i, a in enumerate(o.strips):
        b = (rng.standard_normal(a.shape) * 1e-4).astype(np.float32)
        a[...] = b
        w.psi[i].copy_(torch.from_numpy(b))
    der = core.Der(w.nx_pml, w.ny_pml, w.nz_pml, dx, dx, dx, 8)
    for _ in range(3):
        w.backwardstepVelocity(m, der)
        w.backwardstepStress(m, der)
        o.velocity_bw()
        o.stress_bw()
    torch.cuda.synchronize()
    for k in runs.E3D_FIELDS:
        e = rel_l2(getattr(w, k).cpu().numpy(), getattr(o, k))
        print(f"  {k}: rel L2 {e:.2e}")
        assert e <= 1e-6, k
    worst = max(rel_l2(w.psi[i].cpu().numpy(), a) for i, a in enumerate(o.strips))
    print(f"  strips: worst rel L2 {worst:.2e}")
    assert worst <= 1e-6


def test_e3d_backward_velocity_entries_agree():
    """rsg_e3d_backward_velocity (weighted stresses formed at every tap) and
    rsg_e3d_backward_velocity_ws (formed once into a workspace) give the same
    step from a random state, order 6, free surface off."""
    import ctypes as C
    import torch
    from rockseis_amd import _lib as R
    from rockseis_amd import core
    rng = np.random.default_rng(41)
    nx, ny, nz, lpml, dx, dt = 19, 17, 21, 4, 10.0, 1e-3
    m = core.ModelElastic3D(nx, ny, nz, lpml, dx, dx, dx, 0.0, 0.0, 0.0, False)
    m.Vp[:] = 2400 + 500 * rng.random((nz, ny, nx))
    m.Vs[:] = m.Vp / 1.8
    m.R[:] = 1900 + 300 * rng.random((nz, ny, nx))
    m.staggerModels()
    ws = [core.WavesElastic3D(m, 10, dt) for _ in range(2)]
    for k in core.WavesElastic3D.FIELDS:
        a = torch.from_numpy((rng.standard_normal(getattr(ws[0], k).shape) * 1e-3).astype(np.float32))
        for w in ws:
            getattr(w, k).copy_(a)
    for i in range(R.E3D_NSTRIP):
        a = torch.from_numpy((rng.standard_normal(ws[0].psi[i].shape) * 1e-4).astype(np.float32))
        for w in ws:
            w.psi[i].copy_(a)
    der = core.Der(ws[0].nx_pml, ws[0].ny_pml, ws[0].nz_pml, dx, dx, dx, 6)
    R.check(R.lib().rsg_e3d_backward_velocity(C.byref(ws[0]._grid(6)), C.byref(ws[0]._state(m)), core._stream()),
            "rsg_e3d_backward_velocity")
    ws[1].backwardstepVelocity(m, der)
    torch.cuda.synchronize()
    for k in ("Vx", "Vy", "Vz"):
        e = rel_l2(getattr(ws[1], k).cpu().numpy(), getattr(ws[0], k).cpu().numpy())
        print(f"  {k}: rel L2 {e:.2e}")
        assert e <= 1e-6, k
    worst = max(rel_l2(ws[1].psi[i].cpu().numpy(), ws[0].psi[i].cpu().numpy()) for i in range(R.E3D_NSTRIP))
    assert worst <= 1e-6


def test_fwi_e3d_revolve_schedule_invariant(golden):
    g = golden("fwi_e3d_case0")
    a = _gpu(g, ncheck=3)
    b = _gpu(g, ncheck=int(g["grid"][6]))
    for k in ("vpgrad", "vsgrad", "rhograd", "wavgrad", "dres"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("store", ["s16", "fp16"])
def test_fwi_e3d_16bit_checkpoints(golden, store):
    """16-bit checkpoints vs the FP64 oracle, held to the larger of 1e-4 and
    1.5x the error of the same run with FP32 checkpoints (S16), and to 1e-3
    for FP16 (one scale per array: 11 significant bits)."""
    from oracle.fwi import fwi_e3d
    g = golden("fwi_e3d_case0")
    nx, ny, nz, lpml, dx, dt, nt, order, fs, ncheck, mt = g["grid"]
    ref = fwi_e3d(g["vp"], g["vs"], g["rho"], int(lpml), (dx, dx, dx), dt, int(nt), int(order), int(fs), int(ncheck),
                  str(g["srcfield"]), g["scoords"], g["wav"], g["gcoords"], g["dobs"], real=np.float64,
                  misfit_type=int(mt))
    r32 = _gpu(g, ncheck=4)
    r16 = _gpu(g, ncheck=4, fp16=store)
    for k in ("vpgrad", "vsgrad", "rhograd"):
        e, e32 = rel_l2(r16[k], ref[k]), rel_l2(r32[k], ref[k])
        bound = max(1e-4, 1.5 * e32) if store == "s16" else 1e-3
        print(f"  {store} {k}: rel L2 {e:.2e} (FP32 checkpoints {e32:.2e}, bound {bound:.1e})")
        assert e <= bound, k


def test_fwi_e3d_rejects_bad_records(golden):
    from rockseis_amd import _lib as R
    g = golden("fwi_e3d_case0")
    with pytest.raises(R.RsgError, match="equal count"):
        _gpu(g, cut_uy=True)  # Uy with fewer traces than Ux and Uz

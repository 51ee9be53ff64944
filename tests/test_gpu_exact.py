"""Exact-adjoint gradient (rsg_fwi_a2d_gradient with RSG_ADJOINT_EXACT,
rockseis_amd.fwi setAdjoint("exact")) on the HIP path:

* against the FP64 oracle's exact gradient (oracle/adjoint.py): Vp, rho and
  wavelet gradients within relative L2 1e-4, records within 1e-5 (where the
  case's residual is so small against the records that the FP32 restatement
  itself is further than 1e-4 from FP64, within 3x that FP32 error);
* the adjoint dot-product test on the device: <F w, r> = <w, F^T r> with
  F w the modelled records and F^T r the wavelet gradient, relative 1e-5;
* against central finite differences of the FP64 oracle misfit at 16 random
  model points: |g - fd| <= 1e-4 max|g|;
* a batch of shots (one launch per kernel) against the per-shot oracle sum,
  FP16 checkpoints within the stated bound, the torch.autograd.Function, and
  the error paths (modes and misfits the exact adjoint does not take)."""
import numpy as np
import pytest

from tests.exact_util import exact_case, observed, oracle_args
from tests.gpu_util import data2d, rel_l2

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-4  # gradient error vs the FP64 oracle


def _gpu(c, dobs, ncheck=6, fp16=False):
    from rockseis_amd import core, fwi
    nz, nx = c["vp"].shape
    nt, dt = c["nt"], c["dt"]
    m = core.ModelAcoustic2D(nx, nz, c["lpml"], c["dx"], c["dx"], 0.0, 0.0, c["fs"])
    m.Vp[:], m.R[:] = c["vp"], c["rho"]
    src = data2d(c["scoords"], nt, dt, "P", c["wav"])
    src.makeMap(m, core.SMAP)
    obs = data2d(c["gcoords"], nt, dt, "P", dobs)
    obs.makeMap(m, core.GMAP)
    mod = data2d(c["gcoords"], nt, dt, "P")
    res = data2d(c["gcoords"], nt, dt, "P")
    wgr = data2d(c["scoords"], nt, dt, "P")
    m.staggerModels()
    f = fwi.FwiAcoustic2D(m, src, obs, c["order"])
    f.setAdjoint("exact")
    f.setFp16Storage(fp16)
    vg, rg = fwi.Image2D(m), fwi.Image2D(m)
    f.setVpgrad(vg)
    f.setRhograd(rg)
    f.setDatamodP(mod)
    f.setDataresP(res)
    f.setWavgrad(wgr)
    f.setNcheck(ncheck)
    f.run_optimal()
    return dict(misfit=f.misfit64, vpgrad=vg.imagedata.copy(), rhograd=rg.imagedata.copy(), dmod=mod.data.copy(),
                dres=res.data.copy(), wavgrad=wgr.data.copy())


CASES = [(4, False, 6), (8, True, 5), (2, False, 150), (8, False, 3)]


@pytest.mark.parametrize("order,fs,ncheck", CASES)
def test_exact_vs_fp64_oracle(order, fs, ncheck):
    from oracle.adjoint import fwi_a2d_exact
    print(f"exact adjoint: order {order} fs {fs} ncheck {ncheck}", flush=True)
    c = exact_case(seed=order + 3 * fs, order=order, fs=fs)
    dobs = observed(c)
    ref = fwi_a2d_exact(*oracle_args(c, dobs))
    # the same restatement in FP32: its own error measures the case's FP32
    # conditioning (the residual is a difference of nearly equal records)
    ref32 = fwi_a2d_exact(*oracle_args(c, dobs), real=np.float32)
    r = _gpu(c, dobs, ncheck)
    for k, tol in (("dmod", 1e-5), ("dres", GRAD_RTOL), ("vpgrad", GRAD_RTOL), ("rhograd", GRAD_RTOL),
                   ("wavgrad", GRAD_RTOL)):
        e, e32 = rel_l2(r[k], ref[k]), rel_l2(ref32[k], ref[k])
        print(f"  {k}: rel L2 {e:.2e} (FP32 oracle {e32:.2e})")
        assert e <= max(tol, 3 * e32), k
    # J = |r|^2 / 2: a residual error dr moves J by r.dr + |dr|^2 / 2
    em, er = abs(r["misfit"] - ref["misfit"]) / ref["misfit"], rel_l2(r["dres"], ref["dres"])
    print(f"  misfit: rel {em:.2e} (bound from the residual error {2 * er + er * er:.2e})")
    assert em <= 2 * er + er * er + 1e-6


@pytest.mark.parametrize("order,fs", [(4, False), (8, True)])
def test_exact_dot_product(order, fs):
    c = exact_case(seed=20 + order, order=order, fs=fs)
    dobs = observed(c)
    r = _gpu(c, dobs)
    lhs = float(np.sum(r["dmod"].astype(np.float64) * r["dres"]))
    rhs = float(np.sum(np.asarray(c["wav"], np.float32).astype(np.float64) * r["wavgrad"]))
    print(f"  <Fw,r> {lhs:.9e}  <w,F^T r> {rhs:.9e}  rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_exact_vs_finite_differences():
    from oracle.adjoint import a2d_misfit
    c = exact_case(seed=7, order=8, fs=True)
    dobs = observed(c)
    r = _gpu(c, dobs)
    args = list(oracle_args(c, dobs))
    nz, nx = c["vp"].shape
    rng = np.random.default_rng(16)
    worst = 0.0
    for k in range(16):
        iz, ix = rng.integers(0, [nz, nx])
        idx, key = (0, "vpgrad") if k % 2 == 0 else (1, "rhograd")
        arr = args[idx]
        h = 1e-4 * arr[iz, ix]
        a, b = arr.copy(), arr.copy()
        a[iz, ix] += h
        b[iz, ix] -= h
        args[idx] = a
        jp = a2d_misfit(*args)
        args[idx] = b
        jm = a2d_misfit(*args)
        args[idx] = arr
        fd = (jp - jm) / (2 * h)
        g = r[key]
        err = abs(fd - g[iz, ix]) / np.abs(g).max()
        worst = max(worst, err)
        assert err <= 1e-4, (key, iz, ix, fd, g[iz, ix])
    print(f"  worst |g - fd| / max|g| over 16 points: {worst:.2e}")


def test_exact_batch_and_autograd():
    """Three shots in one batch (different source positions) against the sum
    of the per-shot FP64 oracle gradients; the autograd Function returns the
    same gradient."""
    import torch

    from oracle.adjoint import fwi_a2d_exact
    from rockseis_amd import fwi
    base = exact_case(seed=31, order=4)
    shots, gv, gr, mis = [], 0.0, 0.0, 0.0
    for k, sx in enumerate((120.0, 240.0, 330.0)):
        c = dict(base)
        c["scoords"] = np.array([[sx, 40.0]])
        dobs = observed(c)
        ref = fwi_a2d_exact(*oracle_args(c, dobs))
        gv, gr, mis = gv + ref["vpgrad"], gr + ref["rhograd"], mis + ref["misfit"]
        shots.append(fwi.Shot(c["scoords"], c["wav"], c["gcoords"], dobs))
    sv = fwi.Survey(shots, base["lpml"], base["dx"], base["dx"], base["dt"], order=4, fs=False, ncheck=7,
                    adjoint="exact")
    m, vg, rg, _ = fwi.batch_gradient(base["vp"], base["rho"], sv)
    print(f"  batch: vp {rel_l2(vg.cpu().numpy(), gv):.2e} rho {rel_l2(rg.cpu().numpy(), gr):.2e}")
    assert rel_l2(vg.cpu().numpy(), gv) <= GRAD_RTOL
    assert rel_l2(rg.cpu().numpy(), gr) <= GRAD_RTOL
    assert abs(m - mis) <= 1e-4 * mis
    vp = torch.tensor(base["vp"], dtype=torch.float32, device="cuda", requires_grad=True)
    rho = torch.tensor(base["rho"], dtype=torch.float32, device="cuda", requires_grad=True)
    fwi.misfit(vp, rho, sv).backward()
    assert rel_l2(vp.grad.cpu().numpy(), gv) <= GRAD_RTOL
    assert rel_l2(rho.grad.cpu().numpy(), gr) <= GRAD_RTOL


def test_exact_fp16_checkpoints():
    """16-bit Revolve checkpoints, gradient error vs the FP64 oracle.  This
    case's residual is a small difference of FP32 records, so FP32
    checkpoints already leave 1.5e-4 (measured on MI355X): FP16 (power-of-two
    scale per array, measured 3.4e-4 / 3.9e-4) is held to 6e-4, S16 (16-bit
    block floating point, measured 1.6e-4 / 1.7e-4) to the larger of 1e-4 and
    1.25x the FP32-checkpoint error."""
    from oracle.adjoint import fwi_a2d_exact
    c = exact_case(seed=41, order=8)
    dobs = observed(c)
    ref = fwi_a2d_exact(*oracle_args(c, dobs))
    keys = ("vpgrad", "rhograd", "wavgrad")
    r32 = _gpu(c, dobs, ncheck=5, fp16=0)
    e32 = {k: rel_l2(r32[k], ref[k]) for k in keys}
    for kind in (1, 2):
        r = _gpu(c, dobs, ncheck=5, fp16=kind)
        for k in keys:
            bound = 6e-4 if kind == 1 else max(1e-4, 1.25 * e32[k])
            e = rel_l2(r[k], ref[k])
            print(f"  {('fp32', 'fp16', 's16')[kind]} {k}: rel L2 {e:.2e} (FP32 checkpoints {e32[k]:.2e}, "
                  f"bound {bound:.1e})")
            assert e <= bound, (kind, k)


def test_exact_rejects_other_modes():
    from rockseis_amd import _lib as R
    c = exact_case(seed=1, order=4, nt=20)
    dobs = np.zeros((c["gcoords"].shape[0], c["nt"]))
    from rockseis_amd import core, fwi
    nz, nx = c["vp"].shape
    m = core.ModelAcoustic2D(nx, nz, c["lpml"], c["dx"], c["dx"], 0.0, 0.0, False)
    m.Vp[:], m.R[:] = c["vp"], c["rho"]
    src = data2d(c["scoords"], c["nt"], c["dt"], "P", c["wav"])
    src.makeMap(m, core.SMAP)
    obs = data2d(c["gcoords"], c["nt"], c["dt"], "P", dobs)
    obs.makeMap(m, core.GMAP)
    m.staggerModels()
    f = fwi.FwiAcoustic2D(m, src, obs, 4)
    f.setAdjoint("exact")
    f.setVpgrad(fwi.Image2D(m))
    f.setRhograd(fwi.Image2D(m))
    with pytest.raises(R.RsgError):
        f.run()
    f.setMisfit_type(fwi.CORRELATION)
    with pytest.raises(R.RsgError):
        f.run_optimal()
    with pytest.raises(R.RsgError):
        f.setAdjoint("transposed")

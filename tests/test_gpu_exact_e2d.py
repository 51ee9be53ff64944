"""Exact-adjoint elastic gradient (rsg_fwi_e2d_gradient_batch and the
elastic plan with RSG_ADJOINT_EXACT; ElasticSurvey(adjoint="exact")) on the
HIP path:

* against the FP64 oracle's exact gradient (oracle/adjoint.py
  fwi_e2d_exact, itself pinned by the dot-product test and finite
  differences in tests/test_oracle_adjoint_e2d.py): Vp / Vs / rho and
  lambda / mu / rho gradients and the wavelet gradient within relative L2
  1e-4 (or 3x the FP32 restatement's own error where the case's conditioning
  puts that higher), records within 1e-5;
* the adjoint dot-product test on the device: <F w, r> = <w, F^T r>,
  relative 1e-5;
* central finite differences of the FP64 oracle misfit at 16 model points;
* a ragged batch of two windows against the oracle's per-window gradients
  placed at their offsets; Revolve with few snapshots bit-identical to one
  without recomputation; the persistent plan against the batch call; S16
  checkpoints within the 1e-4 bound; the error paths."""
import numpy as np
import pytest

from tests.exact_util import exact_case_e2d, observed_e2d, oracle_args_e2d
from tests.gpu_util import rel_l2

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-4
COMP = ("P", "VX", "VZ")


def _shot(c, dobs, x0=0, nxw=None, dx_shift=0.0):
    from rockseis_amd import fwi
    sc = np.asarray(c["scoords"], np.float32) + np.float32([dx_shift, 0.0])
    gc = np.asarray(c["gcoords"], np.float32) + np.float32([dx_shift, 0.0])
    d = np.stack([np.asarray(dobs[f], np.float32) for f in COMP])
    return fwi.ElasticShot(sc, np.asarray(c["wav"], np.float32), gc, d, x0=x0, nxw=nxw, srcfield=c["srcfield"])


def _survey(c, shots, ncheck=6, fp16=False):
    from rockseis_amd import fwi
    return fwi.ElasticSurvey(shots, c["lpml"], c["dx"], c["dx"], c["dt"], order=c["order"], fs=c["fs"],
                             ncheck=ncheck, fp16=fp16, adjoint="exact")


def _gpu(c, dobs, ncheck=6, scale=True, fp16=False):
    import torch
    from rockseis_amd import fwi
    sv = _survey(c, [_shot(c, dobs)], ncheck, fp16)
    mis, g0, g1, g2, outs = fwi.elastic_batch_gradient(c["vp"], c["vs"], c["rho"], sv, scale=scale, wavgrad=True)
    torch.cuda.synchronize()
    o = outs[0]
    return dict(misfit=mis, g=[g.cpu().numpy().astype(np.float64) for g in (g0, g1, g2)],
                dmod={f: o["dmod"][i].cpu().numpy() for i, f in enumerate(COMP)},
                dres={f: o["dres"][i].cpu().numpy() for i, f in enumerate(COMP)},
                wavgrad=o["wavgrad"].cpu().numpy())


CASES = [(4, False, "P", 6), (8, True, "P", 5), (4, True, "VZ", 4), (2, False, "VX", 160)]


@pytest.mark.parametrize("order,fs,src,ncheck", CASES)
def test_e2d_exact_vs_fp64_oracle(order, fs, src, ncheck):
    from oracle.adjoint import fwi_e2d_exact
    print(f"elastic exact adjoint: order {order} fs {fs} source {src} ncheck {ncheck}", flush=True)
    c = exact_case_e2d(seed=40 + order + 3 * fs, order=order, fs=fs, srcfield=src)
    dobs = observed_e2d(c)
    ref = fwi_e2d_exact(*oracle_args_e2d(c, dobs))
    ref32 = fwi_e2d_exact(*oracle_args_e2d(c, dobs), real=np.float32)
    for scale, keys in ((True, ("vpgrad", "vsgrad", "rhograd")), (False, ("lamgrad", "mugrad", "rhograd_lame"))):
        r = _gpu(c, dobs, ncheck, scale=scale)
        for i, k in enumerate(keys):
            e, e32 = rel_l2(r["g"][i], ref[k]), rel_l2(ref32[k], ref[k])
            print(f"  {k}: rel L2 {e:.2e} (FP32 oracle {e32:.2e})")
            assert e <= max(GRAD_RTOL, 3 * e32), k
    for f in COMP:
        e = rel_l2(r["dmod"][f], ref["dmod"][f])
        print(f"  dmod {f}: rel L2 {e:.2e}")
        assert e <= 1e-5
    e, e32 = rel_l2(r["wavgrad"], ref["wavgrad"]), rel_l2(ref32["wavgrad"], ref["wavgrad"])
    print(f"  wavgrad: rel L2 {e:.2e} (FP32 oracle {e32:.2e})")
    assert e <= max(GRAD_RTOL, 3 * e32)
    assert abs(r["misfit"] - ref["misfit"]) <= 1e-4 * ref["misfit"]


@pytest.mark.parametrize("order,fs,src", [(4, False, "P"), (8, True, "VZ")])
def test_e2d_exact_dot_product(order, fs, src):
    c = exact_case_e2d(seed=60 + order, order=order, fs=fs, srcfield=src)
    dobs = observed_e2d(c)
    r = _gpu(c, dobs)
    lhs = sum(float(np.sum(r["dmod"][f].astype(np.float64) * r["dres"][f])) for f in COMP)
    rhs = float(np.sum(np.asarray(c["wav"], np.float32).astype(np.float64) * r["wavgrad"]))
    print(f"  <Fw,r> {lhs:.9e}  <w,F^T r> {rhs:.9e}  rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_e2d_exact_vs_finite_differences():
    from oracle.adjoint import e2d_misfit
    c = exact_case_e2d(seed=71, order=8, fs=True, srcfield="P")
    dobs = observed_e2d(c)
    r = _gpu(c, dobs)
    args = list(oracle_args_e2d(c, dobs))
    nz, nx = c["vp"].shape
    rng = np.random.default_rng(17)
    worst = 0.0
    for k in range(16):
        iz, ix = rng.integers(0, [nz, nx])
        idx = k % 3
        arr = args[idx]
        h = 1e-4 * arr[iz, ix]
        a, b = arr.copy(), arr.copy()
        a[iz, ix] += h
        b[iz, ix] -= h
        args[idx] = a
        jp = e2d_misfit(*args)
        args[idx] = b
        jm = e2d_misfit(*args)
        args[idx] = arr
        fd = (jp - jm) / (2 * h)
        g = r["g"][idx]
        err = abs(fd - g[iz, ix]) / np.abs(g).max()
        worst = max(worst, err)
        assert err <= 1e-4, (idx, iz, ix, fd, g[iz, ix])
    print(f"  worst |g - fd| / max|g| over 16 points: {worst:.2e}")


def test_e2d_exact_ragged_windows_vs_oracle():
    """Two shots on windows of different widths in one ragged batch: each
    window's exact gradient (oracle, FP64, on the window model) placed at its
    offset and summed."""
    import torch
    from oracle.adjoint import fwi_e2d_exact
    from rockseis_amd import fwi
    c = exact_case_e2d(seed=81, nx=60, order=4, fs=True, srcfield="P")
    nz, nx = c["vp"].shape
    wins = [(0, 44), (14, 46)]
    shots, refs = [], []
    for x0, nxw in wins:
        cw = dict(c)
        for k in ("vp", "vs", "rho", "vp0", "vs0", "rho0"):
            cw[k] = np.ascontiguousarray(c[k][:, x0:x0 + nxw])
        cw["scoords"] = np.array([[0.5 * (nxw - 1) * c["dx"], 4 * c["dx"]]])
        cw["gcoords"] = np.stack([np.linspace(2 * c["dx"], (nxw - 3) * c["dx"], 10), np.full(10, 3 * c["dx"])], 1)
        dobs = observed_e2d(cw)
        refs.append((x0, nxw, fwi_e2d_exact(*oracle_args_e2d(cw, dobs))))
        shots.append(_shot(cw, dobs, x0=x0, nxw=nxw, dx_shift=x0 * c["dx"]))
    mis, g0, g1, g2, _ = fwi.elastic_batch_gradient(c["vp"], c["vs"], c["rho"], _survey(c, shots), scale=True)
    torch.cuda.synchronize()
    want = [np.zeros((nz, nx)) for _ in range(3)]
    for x0, nxw, r in refs:
        for i, k in enumerate(("vpgrad", "vsgrad", "rhograd")):
            want[i][:, x0:x0 + nxw] += r[k]
    for i, g in enumerate((g0, g1, g2)):
        e = rel_l2(g.cpu().numpy(), want[i])
        print(f"  gradient {i}: rel L2 {e:.2e}")
        assert e <= GRAD_RTOL
    assert abs(mis - sum(r["misfit"] for _, _, r in refs)) <= 1e-4 * mis


def test_e2d_exact_revolve_schedule_invariant():
    """Few snapshots (recomputation) and one per step give the same bits."""
    c = exact_case_e2d(seed=90, order=8, fs=False, srcfield="P")
    dobs = observed_e2d(c)
    a = _gpu(c, dobs, ncheck=3)
    b = _gpu(c, dobs, ncheck=c["nt"])
    for x, y in zip(a["g"], b["g"]):
        assert np.array_equal(x, y)
    assert np.array_equal(a["wavgrad"], b["wavgrad"])


def test_e2d_exact_plan_matches_batch():
    import torch
    from rockseis_amd import fwi
    c = exact_case_e2d(seed=91, order=4, fs=True, srcfield="VX")
    dobs = observed_e2d(c)
    r = _gpu(c, dobs)
    plan = fwi.ElasticFwiPlan(c["vp"], c["vs"], c["rho"], _survey(c, [_shot(c, dobs)]), scale=True, tune=False)
    try:
        for rep in range(2):
            mis, g0, g1, g2, _ = plan.run()
            torch.cuda.synchronize()
            for x, y in zip((g0, g1, g2), r["g"]):
                assert np.array_equal(x.cpu().numpy().astype(np.float64), y), rep
            assert mis == r["misfit"]
    finally:
        plan.close()


def test_e2d_exact_s16_checkpoints():
    """16-bit block floating point checkpoints: the gradient stays within
    the 1e-4 bound of the FP64 oracle."""
    from oracle.adjoint import fwi_e2d_exact
    c = exact_case_e2d(seed=95, order=8, fs=True, srcfield="P")
    dobs = observed_e2d(c)
    ref = fwi_e2d_exact(*oracle_args_e2d(c, dobs))
    r = _gpu(c, dobs, ncheck=5, fp16="s16")
    for i, k in enumerate(("vpgrad", "vsgrad", "rhograd")):
        e = rel_l2(r["g"][i], ref[k])
        print(f"  S16 {k}: rel L2 {e:.2e}")
        assert e <= GRAD_RTOL


def test_e2d_exact_rejects_other_misfits():
    from rockseis_amd import _lib as R
    from rockseis_amd import fwi
    c = exact_case_e2d(seed=3, order=4, nt=20)
    dobs = {f: np.zeros((12, 20)) for f in COMP}
    sv = _survey(c, [_shot(c, dobs)])
    sv.misfit_type = fwi.CORRELATION
    with pytest.raises(R.RsgError, match="DIFFERENCE"):
        fwi.elastic_batch_gradient(c["vp"], c["vs"], c["rho"], sv)

"""Ragged batches (rsg_ragged2d_*): shots on grids of different sizes in one
launch per half step give, per shot, exactly the fields of the single-shot
calls, for every half step (elastic forward and adjoint-state, acoustic)
and every rows-per-thread setting."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(130, 70), (61, 90), (200, 45)]   # unpadded (nx, nz) of the three windows


def _setup(core, kind, seed):
    import torch
    from rockseis_amd import synth
    rng = np.random.default_rng(seed)
    waves, models = [], []
    for nx, nz in SIZES:
        vp = synth.gradient_anomaly_vp(nx, nz, 5.0, 5.0, 1500, 3000, nanom=3, seed=int(rng.integers(1000)))
        if kind == "e2d":
            m = core.ModelElastic2D(nx, nz, 12, 5.0, 5.0, 0.0, 0.0, nx == 61)
            m.Vs[:] = vp / 1.8
            w_cls = core.WavesElastic2D
        else:
            m = core.ModelAcoustic2D(nx, nz, 12, 5.0, 5.0, 0.0, 0.0, nx == 61)
            w_cls = core.WavesAcoustic2D
        m.Vp[:] = vp
        m.R[:] = synth.gardner(vp)
        m.staggerModels()
        w = w_cls(m, 10, 4e-4)
        fields = ("Sxx", "Szz", "Sxz", "Vx", "Vz") if kind == "e2d" else ("P", "Vx", "Vz")
        for f in fields:
            t = getattr(w, f)
            t.copy_(torch.from_numpy(rng.standard_normal(t.shape).astype(np.float32)))
        for t in w.psi.values():
            t.copy_(torch.from_numpy(0.1 * rng.standard_normal(t.shape).astype(np.float32)))
        waves.append(w)
        models.append(m)
    return waves, models, fields


@pytest.mark.parametrize("kind,order", [("e2d", 8), ("e2d", 3), ("a2d", 4)])
def test_ragged_equals_single_shot(kind, order):
    from rockseis_amd import _lib as R, core
    old = R.get_tuning("rz2d_ragged")
    try:
        for rz in (1, 2, 4, 8):
            R.set_tuning("rz2d_ragged", rz)
            wa, ma, fields = _setup(core, kind, 11)
            wb, mb, _ = _setup(core, kind, 11)
            rb = core.RaggedBatch2D(wb, mb, order)
            der = core.Der(1, 1, 1, 1, 1, 1, order)
            seq = (0, 1, 2, 3, 0, 1) if kind == "e2d" else (0, 1, 0, 1)
            for which in seq:
                rb.step(which)
                for w, m in zip(wa, ma):
                    if which == 0:
                        w.forwardstepVelocity(m, der)
                    elif which == 1:
                        w.forwardstepStress(m, der)
                    elif which == 2:
                        w.backwardstepVelocity(m, der)
                    else:
                        w.backwardstepStress(m, der)
            rb.close()
            for k, (x, y) in enumerate(zip(wa, wb)):
                for f in fields:
                    assert np.array_equal(getattr(x, f).cpu().numpy(), getattr(y, f).cpu().numpy()), (rz, k, f)
                for n in x.psi:
                    assert np.array_equal(x.psi[n].cpu().numpy(), y.psi[n].cpu().numpy()), (rz, k, n)
    finally:
        R.set_tuning("rz2d_ragged", old)

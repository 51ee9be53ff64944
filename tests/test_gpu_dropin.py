"""The reference's own driver programs with the rsgpu drop-in linked in
(oracle/_ref/bin/gpu), run end to end on the GPU on the inputs the
fixtures were made with, against the CPU programs' outputs
(tests/golden/make_dropin_golden.py).

Records and snapshots: relative L2 <= 1e-5.  The FWI gradient images (the
Vp and rho gradients stacked over the shots by the reference's own
Image2D::stackImage, the source illumination, the wavelet gradient) and
the misfit: relative L2 <= 1e-4 against the FP32 CPU program (the imaging
products sum thousands of steps; the acoustic gradient's own parity test
holds the C-ABI to the reference at <= 1e-5 on fixed residuals)."""
import os
import tempfile

import numpy as np
import pytest

import dropin_cases as dc
from tests.gpu_util import RTOL_L2, rel_l2

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(dc.binary("gpu", "Acousticfwi2d")),
                                 reason="drop-in driver programs not built (make -C oracle drivers)")]


def _compare(out, g, keys, tol, label, norm=None):
    """Relative L2 of each output against the fixture; with norm, the error
    of output k is divided by the norm of fixture norm[k] instead."""
    worst = 0.0
    for k in keys:
        if norm and k in norm:
            d = np.asarray(out[k], np.float64) - np.asarray(g[k], np.float64)
            e = float(np.linalg.norm(d) / np.linalg.norm(np.asarray(g[norm[k]], np.float64)))
            print(f"{label} {k}: L2 error / |{norm[k]}| {e:.2e} (rel L2 {rel_l2(out[k], g[k]):.2e})", flush=True)
        else:
            e = rel_l2(out[k], g[k])
            print(f"{label} {k}: rel L2 {e:.2e}", flush=True)
        worst = max(worst, e)
        assert e <= tol, (label, k, e)
    return worst


# A residual is the modelled record minus the observed one: its error is the
# record's error, so it is held to the records' bound relative to the
# modelled record's norm (its own norm can be a few percent of that).
_RES_NORM = {"pres": "pmod", "uxres": "uxmod", "uzres": "uzmod"}


@pytest.mark.parametrize("elastic", (False, True), ids=("Acousticmod2d", "Elasticmod2d"))
def test_modelling_driver_dropin(golden, elastic):
    name = "dropin_emod" if elastic else "dropin_amod"
    print(f"{name}: rs{'Elastic' if elastic else 'Acoustic'}mod2d with the drop-in", flush=True)
    g = golden(name)
    with tempfile.TemporaryDirectory() as d:
        out = dc.modelling("gpu", d, elastic)
    assert set(out) == set(g.files)
    _compare(out, g, sorted(out), RTOL_L2, name)


def test_fwi_driver_dropin(golden):
    print("dropin_fwi: mpiexec -n 2 rsAcousticfwi2d with the drop-in (domain model, srcilum)", flush=True)
    g = golden("dropin_fwi")
    with tempfile.TemporaryDirectory() as d:
        out = dc.fwi("gpu", d, g["obs"])
    _compare(out, g, ("pmod", "pres"), RTOL_L2, "dropin_fwi", _RES_NORM)
    _compare(out, g, ("vp_grad", "rho_grad", "src_ilum", "source_grad", "misfit"), 1e-4, "dropin_fwi")


def test_elastic_fwi_driver_dropin(golden):
    print("dropin_efwi: mpiexec -n 2 rsElasticfwi2d with the drop-in (domain model, free surface)", flush=True)
    g = golden("dropin_efwi")
    obs = {"p": g["obs_p"], "vx": g["obs_vx"], "vz": g["obs_vz"]}
    with tempfile.TemporaryDirectory() as d:
        out = dc.efwi("gpu", d, obs)
    _compare(out, g, ("pmod", "uxmod", "uzmod", "pres", "uxres", "uzres"), RTOL_L2, "dropin_efwi", _RES_NORM)
    _compare(out, g, ("vp_grad", "vs_grad", "rho_grad", "source_grad", "misfit"), 1e-4, "dropin_efwi")

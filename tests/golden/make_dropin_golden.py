"""Write the fixtures of the driver-level tests from the reference's own
programs (oracle/Makefile target `drivers`, CPU build under
oracle/_ref/bin/ref):

  dropin_amod.npz  rsAcousticmod2d: P / Vx / Vz records of two shots and
                   the P snapshots of each shot
  dropin_emod.npz  rsElasticmod2d (free surface): records, P (Sxx+Szz) and
                   Vx snapshots
  dropin_fwi.npz   mpiexec -n 2 rsAcousticfwi2d, one gradient evaluation
                   (L-BFGS with max_iterations 0) on a smoothed starting
                   model: the stacked Vp and rho gradients, the source
                   illumination, the wavelet gradient, the misfit of each
                   shot and the modelled and residual records; the observed
                   data (the acoustic case's P records) ride along as input
  dropin_efwi.npz  mpiexec -n 2 rsElasticfwi2d, one gradient evaluation on
                   the smoothed elastic model with the elastic case's P / Vx /
                   Vz records as observed data: Vp / Vs / rho gradients, the
                   wavelet gradient, misfits, modelled and residual records

The inputs are the seeded arrays of tests/dropin_cases.py written as RSS
files; run in a container where /root/reference is built:
    make -C oracle drivers && python tests/golden/make_dropin_golden.py
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import dropin_cases as dc  # noqa: E402


def save(name, **arrs):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrs)
    print(f"{name}.npz  {os.path.getsize(path) / 1024:.1f} KiB")


def main():
    with tempfile.TemporaryDirectory() as tmp:
        amod = dc.modelling("ref", dc.fresh_dir(tmp, "amod"), elastic=False)
        save("dropin_amod", **amod)
        emod = dc.modelling("ref", dc.fresh_dir(tmp, "emod"), elastic=True)
        save("dropin_emod", **emod)
        fwi = dc.fwi("ref", dc.fresh_dir(tmp, "fwi"), amod["p"])
        save("dropin_fwi", obs=amod["p"], **fwi)
        efwi = dc.efwi("ref", dc.fresh_dir(tmp, "efwi"), emod)
        save("dropin_efwi", obs_p=emod["p"], obs_vx=emod["vx"], obs_vz=emod["vz"], **efwi)


if __name__ == "__main__":
    main()

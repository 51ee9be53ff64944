"""The class-API drop-in at link level (CPU): the reference's own driver
programs link against librsdropin.a and librsgpu.so with no unresolved
rockseis symbols, the drop-in's definitions take the place of the
reference's template instantiations, the drop-in programs refuse to run
without a GPU (no CPU path), and the CPU programs reproduce the fixtures
of the driver-level GPU tests bit for bit."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import dropin_cases as dc

DRIVERS = ("Acousticmod2d", "Elasticmod2d", "Acousticfwi2d", "Acousticmod3d", "Elasticfwi2d", "Acousticfwi3d")
REPLACED = (
    "rockseis::WavesAcoustic2D<float>::forwardstepVelocity(",
    "rockseis::WavesAcoustic2D<float>::forwardstepStress(",
    "rockseis::WavesAcoustic2D<float>::insertSource(",
    "rockseis::WavesAcoustic2D<float>::recordData(",
    "rockseis::WavesElastic2D<float>::backwardstepVelocity(",
    "rockseis::WavesAcoustic3D<float>::forwardstepStress(",
    "rockseis::Der<float>::ddx_fw(",
    "rockseis::Pml<float>::computeABC(int)",
    "rockseis::Snapshot2D<float>::writeSnap(int)",
    "rockseis::Revolve<float>::revolve()",
    "rockseis::FwiAcoustic2D<float>::run_optimal()",
    "rockseis::FwiElastic2D<float>::run_optimal()",
)

have = all(os.path.exists(dc.binary(f, d)) for f in ("ref", "gpu") for d in DRIVERS)
pytestmark = pytest.mark.skipif(not have, reason="driver programs are built where the reference sources are mounted "
                                                 "(make -C oracle drivers)")


def _nm(path, *args):
    return subprocess.run(["nm", *args, path], capture_output=True, text=True, check=True).stdout


@pytest.mark.parametrize("name", DRIVERS)
def test_driver_links_with_no_unresolved_rockseis_symbol(name):
    und = [ln for ln in _nm(dc.binary("gpu", name), "-u", "-C").splitlines() if "rockseis::" in ln]
    assert und == [], und[:5]


def _binding(text, member):
    kinds = {ln.split()[1] for ln in text.splitlines() if member in ln and " [clone" not in ln}
    return kinds


@pytest.mark.parametrize("name", ("Acousticmod2d", "Acousticfwi2d", "Elasticfwi2d"))
def test_dropin_members_replace_the_reference(name):
    gpu, ref = _nm(dc.binary("gpu", name), "-C"), _nm(dc.binary("ref", name), "-C")
    for m in REPLACED:
        gk, rk = _binding(gpu, m), _binding(ref, m)
        if not rk:
            continue  # member not linked into this program
        assert gk == {"T"}, (m, gk)  # the drop-in's strong definition
        assert rk == {"W"}, (m, rk)  # the reference's weak template instantiation
    # the drop-in programs call the C-ABI
    assert "rsg_a2d_forward_velocity" in _nm(dc.binary("gpu", name), "-u")


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the behaviour without a GPU")
def test_dropin_program_fails_loudly_without_gpu():
    with tempfile.TemporaryDirectory() as d:
        vp, vs, rho = dc.model()
        dc.write_common(d, vp, rho)
        dc.mod_config(d, False)
        p = subprocess.run([dc.binary("gpu", "Acousticmod2d"), "mod.cfg"], cwd=d, capture_output=True, text=True,
                           timeout=120)
    assert p.returncode != 0
    assert "Fatal error: rsgpu drop-in" in p.stderr, p.stderr[-500:]


def test_reference_programs_reproduce_the_fixtures(golden):
    with tempfile.TemporaryDirectory() as tmp:
        amod = dc.modelling("ref", dc.fresh_dir(tmp, "amod"), elastic=False)
        emod = dc.modelling("ref", dc.fresh_dir(tmp, "emod"), elastic=True)
        fwi = dc.fwi("ref", dc.fresh_dir(tmp, "fwi"), golden("dropin_fwi")["obs"])
        ge = golden("dropin_efwi")
        efwi = dc.efwi("ref", dc.fresh_dir(tmp, "efwi"), {"p": ge["obs_p"], "vx": ge["obs_vx"], "vz": ge["obs_vz"]})
    for name, out in (("dropin_amod", amod), ("dropin_emod", emod), ("dropin_fwi", fwi), ("dropin_efwi", efwi)):
        g = golden(name)
        for k, v in out.items():
            np.testing.assert_array_equal(v, g[k], err_msg=f"{name}:{k}")

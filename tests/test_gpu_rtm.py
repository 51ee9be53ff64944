"""Acoustic 2D reverse-time migration on the HIP path (rsg_rtm_a2d_image
through rockseis_amd.rtm.RtmAcoustic2D) against the compiled reference's
RtmAcoustic2D::run_optimal (golden fixtures) within relative L2 1e-5."""
import numpy as np
import pytest

from tests.golden.make_golden import RTM_CASES
from tests.gpu_util import RTOL_L2, data2d, rel_l2

pytestmark = pytest.mark.gpu


def _run(g, case):
    from rockseis_amd import core, fwi, rtm
    order, fs, ncheck, nhx, nhz = RTM_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:], m.R[:] = g["vp"], g["rho"]
    src = data2d(g["scoords"], nt, dt, "P", g["wav"])
    src.makeMap(m, core.SMAP)
    dat = data2d(g["gcoords"], nt, dt, "P", g["data"])
    dat.makeMap(m, core.GMAP)
    m.staggerModels()
    img = fwi.Image2D(m, nhx, nhz)
    r = rtm.RtmAcoustic2D(m, img, src, dat, order, 1)
    r.setNcheck(ncheck)
    r.run_optimal()
    return img.imagedata.reshape(nhz, nhx, nz, nx)


@pytest.mark.parametrize("case", range(len(RTM_CASES)))
def test_rtm_a2d_vs_reference(golden, case):
    g = golden(f"rtm_case{case}")
    print(f"rtm a2d case {case}: {RTM_CASES[case]}", flush=True)
    img = _run(g, case)
    e = rel_l2(img, g["image"])
    print(f"  image: rel L2 {e:.2e}", flush=True)
    assert e <= RTOL_L2
    # each offset's panel on its own: the zero-offset panel dominates the norm
    for ihz in range(img.shape[0]):
        for ihx in range(img.shape[1]):
            want = g["image"][ihz, ihx]
            if np.any(want):
                assert rel_l2(img[ihz, ihx], want) <= RTOL_L2, (ihz, ihx)


def test_rtm_a2d_rerun_identical(golden):
    """Two runs give bit-identical images (the image is zeroed by each run)."""
    g = golden("rtm_case1")
    a, b = _run(g, 1), _run(g, 1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("case", range(2))
def test_rtm_a3d_vs_reference(golden, case):
    from rockseis_amd import core, fwi, rtm
    from tests.golden.make_golden import RTM_A3D_CASES
    from tests.gpu_util import data3d
    g = golden("fwi_a3d_case0")
    f = golden(f"rtm_a3d_case{case}")
    order, fs, ncheck, nhx, nhy, nhz = RTM_A3D_CASES[case]
    print(f"rtm a3d case {case}: {RTM_A3D_CASES[case]}", flush=True)
    nx, ny, nz, lpml, dx, dt, nt = g["grid"]
    nx, ny, nz, lpml, nt = int(nx), int(ny), int(nz), int(lpml), int(nt)
    m = core.ModelAcoustic3D(nx, ny, nz, lpml, dx, dx, dx, 0.0, 0.0, 0.0, fs)
    m.Vp[:], m.R[:] = g["vp"], g["rho"]
    src = data3d(g["scoords"], nt, dt, "P", g["wav"])
    src.makeMap(m, core.SMAP)
    dat = data3d(g["gcoords"], nt, dt, "P", g["dobs"])
    dat.makeMap(m, core.GMAP)
    m.staggerModels()
    img = fwi.Image3D(m, nhx, nhy, nhz)
    r = rtm.RtmAcoustic3D(m, img, src, dat, order, 1)
    r.setNcheck(ncheck)
    r.run_optimal()
    got = img.imagedata.reshape(f["image"].shape)
    e = rel_l2(got, f["image"])
    print(f"  image: rel L2 {e:.2e}", flush=True)
    assert e <= RTOL_L2

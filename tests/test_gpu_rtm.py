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


def test_rtm_a2d_batch_equals_stack_of_shots():
    """Three shots migrated as one batch (rsg_rtm_a2d_batch) against the sum
    of the FP32 oracle's single-shot images (relative L2 1e-5; the batch adds
    the shots' products cell by cell at every step, the stack adds whole
    images), with 3 x 1 subsurface offsets; and a batch of one shot equals
    the single-shot entry point bit for bit."""
    from oracle import rtm as ortm
    from oracle import runs
    from rockseis_amd import fwi, rtm, synth
    nx, nz, lpml, dx, dt, nt, order, ncheck = 60, 50, 10, 10.0, 1e-3, 300, 8, 7
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=3)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    rho0 = synth.gardner(vp0)
    wav = synth.ricker(nt, dt, 15.0)[None]
    gc = synth.surface_line(nx, dx, 20.0)
    shots, want = [], 0.0
    for sx in (150.0, 300.0, 450.0):
        sc = np.float32([[sx, 10.0]])
        dobs = runs.a2d_run(vtrue, rtrue, lpml, dx, dx, dt, nt, order, False, "P", sc, wav, "P", gc)["rec"]
        shots.append(fwi.Shot(sc, wav, gc, dobs))
        want = want + ortm.rtm_a2d(vp0, rho0, lpml, dx, dx, dt, nt, order, False, ncheck, sc, wav, gc, dobs, 3,
                                   1).astype(np.float64)
    sv = fwi.Survey(shots, lpml, dx, dx, dt, order=order, ncheck=ncheck, mode="revolve")
    img = rtm.batch_image(vp0, rho0, sv, nhx=3, nhz=1).cpu().numpy()
    e = rel_l2(img, want)
    print(f"  batch of 3 shots vs the oracle's stack: rel L2 {e:.2e}", flush=True)
    assert e <= RTOL_L2
    one = fwi.Survey(shots[:1], lpml, dx, dx, dt, order=order, ncheck=ncheck, mode="revolve")
    b1 = rtm.batch_image(vp0, rho0, one, nhx=3, nhz=1).cpu().numpy()
    g = dict(grid=np.float32([nx, nz, lpml, dx, dt, nt]), vp=vp0, rho=rho0, scoords=shots[0].scoords, wav=wav,
             gcoords=gc, data=shots[0].dobs)
    from rockseis_amd import core
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, False)
    m.Vp[:], m.R[:] = vp0, rho0
    src = data2d(g["scoords"], nt, dt, "P", wav)
    src.makeMap(m, core.SMAP)
    dat = data2d(gc, nt, dt, "P", g["data"])
    dat.makeMap(m, core.GMAP)
    m.staggerModels()
    im = fwi.Image2D(m, 3, 1)
    r = rtm.RtmAcoustic2D(m, im, src, dat, order, 1)
    r.setNcheck(ncheck)
    r.run_optimal()
    assert np.array_equal(b1, im.imagedata.reshape(b1.shape))


def test_rtm_offgrid_receivers_reproducible_and_vs_oracle():
    """Receivers between grid nodes (x + 0.37 dx, z + 0.21 dz): the 2x2 cells
    of neighbouring receivers overlap with non-zero weights.  The data are
    injected cell by cell in trace order, so two batched runs give
    bit-identical images, equal to the oracle's stack within 1e-5."""
    from oracle import rtm as ortm
    from oracle import runs
    from rockseis_amd import fwi, rtm, synth
    nx, nz, lpml, dx, dt, nt, order, ncheck = 60, 50, 10, 10.0, 1e-3, 300, 4, 20
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1800.0, 3000.0, nanom=5, seed=9)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 60.0, dx, dx)
    rho0 = synth.gardner(vp0)
    wav = synth.ricker(nt, dt, 15.0)[None]
    gc = (synth.surface_line(nx - 1, dx, 20.0) + np.float32([[0.37 * dx, 0.21 * dx]])).astype(np.float32)
    shots, want = [], 0.0
    for sx in (150.0, 450.0):
        sc = np.float32([[sx, 10.0]])
        dobs = runs.a2d_run(vtrue, rtrue, lpml, dx, dx, dt, nt, order, False, "P", sc, wav, "P", gc)["rec"]
        shots.append(fwi.Shot(sc, wav, gc, dobs))
        want = want + ortm.rtm_a2d(vp0, rho0, lpml, dx, dx, dt, nt, order, False, ncheck, sc, wav, gc, dobs, 1,
                                   1).astype(np.float64)
    sv = fwi.Survey(shots, lpml, dx, dx, dt, order=order, ncheck=ncheck, mode="revolve")
    a = rtm.batch_image(vp0, rho0, sv).cpu().numpy()
    b = rtm.batch_image(vp0, rho0, sv).cpu().numpy()
    assert np.array_equal(a, b)
    e = rel_l2(a[0, 0], want[0, 0])
    print(f"  off-grid receivers: rel L2 {e:.2e} vs the oracle's stack", flush=True)
    assert e <= RTOL_L2


def test_rtm_a3d_offgrid_receivers_reproducible_and_vs_oracle(golden):
    """A 12 x 10 receiver grid between grid nodes (spacing dx, offsets 0.37,
    0.43 and 0.21 cells) on the 3D fixture model: the data are injected cell
    by cell in trace order, so two runs give bit-identical images, equal to
    the FP32 oracle within 1e-5."""
    from oracle import rtm as ortm
    from oracle import runs
    from rockseis_amd import core, fwi, rtm
    from tests.gpu_util import data3d
    g = golden("fwi_a3d_case0")
    nx, ny, nz, lpml, dx, dt, nt = g["grid"]
    nx, ny, nz, lpml, nt = int(nx), int(ny), int(nz), int(lpml), int(nt)
    xs, ys = np.meshgrid(20.0 + dx * np.arange(12) + 0.37 * dx, 20.0 + dx * np.arange(10) + 0.43 * dx)
    gc = np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, 20.0 + 0.21 * dx)], axis=1).astype(np.float32)
    vtrue = g["vp"].copy()
    vtrue[3:19, 3:21, 3:23] *= 1.25
    data = runs.a3d_run(vtrue, g["rho"], lpml, (dx, dx, dx), dt, nt, 4, False, "P", g["scoords"], g["wav"], "P",
                        gc)["rec"]

    def run():
        m = core.ModelAcoustic3D(nx, ny, nz, lpml, dx, dx, dx, 0.0, 0.0, 0.0, False)
        m.Vp[:], m.R[:] = g["vp"], g["rho"]
        src = data3d(g["scoords"], nt, dt, "P", g["wav"])
        src.makeMap(m, core.SMAP)
        dat = data3d(gc, nt, dt, "P", data)
        dat.makeMap(m, core.GMAP)
        m.staggerModels()
        img = fwi.Image3D(m)
        r = rtm.RtmAcoustic3D(m, img, src, dat, 4, 1)
        r.setNcheck(10)
        r.run_optimal()
        return img.imagedata.copy()
    a, b = run(), run()
    assert np.array_equal(a, b)
    want = ortm.rtm_a3d(g["vp"], g["rho"], lpml, (dx, dx, dx), dt, nt, 4, False, 10, g["scoords"], g["wav"], gc,
                        data, 1, 1, 1)
    e = rel_l2(a, want[0, 0, 0])
    print(f"  3D off-grid receivers: rel L2 {e:.2e} vs the oracle", flush=True)
    assert e <= RTOL_L2

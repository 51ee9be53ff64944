"""Launch-shape knobs and the autotuner: every knob value gives bit-identical
fields (each cell is computed by the same arithmetic), and the tuner picks
one of its candidates, sets it and caches it."""
import numpy as np
import pytest

from tests.gpu_util import data2d, data3d

pytestmark = pytest.mark.gpu


def _a2d(core, order):
    from rockseis_amd import synth
    nx, nz, lpml, dx, dt, nt = 150, 90, 12, 10.0, 1e-3, 120
    vp = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1500, 3000, nanom=5, seed=3)
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, False)
    m.Vp[:] = vp
    m.R[:] = synth.gardner(vp)
    src = data2d(np.float32([[700.0, 20.0]]), nt, dt, "P", synth.ricker(nt, dt, 15.0)[None])
    src.makeMap(m, core.SMAP)
    rec = data2d(synth.surface_line(nx, dx, 10.0), nt, dt, "P")
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingAcoustic2D(m, src, order, 1)
    mod.setRecP(rec)
    mod.run()
    return rec.data.copy(), mod.waves.P.cpu().numpy()


def _a3d(core):
    from rockseis_amd import synth
    nx, ny, nz, lpml, dx, dt, nt = 40, 36, 44, 6, 10.0, 1e-3, 60
    vp = synth.layered_anomaly_vp3d(nx, ny, nz, dx, dx, dx, seed=4)
    m = core.ModelAcoustic3D(nx, ny, nz, lpml, dx, dx, dx, 0.0, 0.0, 0.0, False)
    m.Vp[:] = vp
    m.R[:] = synth.gardner(vp)
    src = data3d(np.float32([[200.0, 180.0, 30.0]]), nt, dt, "P", synth.ricker(nt, dt, 20.0)[None])
    src.makeMap(m, core.SMAP)
    gx, gy = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx)
    gc = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, 10.0)], 1).astype(np.float32)
    rec = data3d(gc, nt, dt, "P")
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingAcoustic3D(m, src, 4, 1)
    mod.setRecP(rec)
    mod.run()
    return rec.data.copy(), mod.waves.P.cpu().numpy()


def test_rz2d_values_give_identical_fields():
    from rockseis_amd import _lib as R, core
    old = R.get_tuning("rz2d")
    try:
        out = {}
        for v in (1, 2, 4, 8):
            R.set_tuning("rz2d", v)
            out[v] = _a2d(core, 8)
        for v in (4, 8):
            assert np.array_equal(out[v][0], out[2][0]), v
            assert np.array_equal(out[v][1], out[2][1]), v
        assert np.abs(out[2][0]).max() > 0
    finally:
        R.set_tuning("rz2d", old)


def test_pml_cost_values_give_identical_fields():
    """Every split of the 2D tiles into XCD ranges (pml_cost from 1 to 10000
    percent, the extremes leaving ranges of a few tiles or none) covers each
    tile exactly once: records and fields bit-identical, acoustic and elastic
    (free surface), at two tile heights."""
    from rockseis_amd import _lib as R, core
    old = R.get_tuning("pml_cost"), R.get_tuning("rz2d")
    try:
        for rz in (1, 2):
            R.set_tuning("rz2d", rz)
            out = {}
            for v in (100, 1, 200, 777, 10000):
                R.set_tuning("pml_cost", v)
                out[v] = list(_a2d(core, 8)) + _e2d_run(core, 0, fs=True)
            for v in (1, 200, 777, 10000):
                for a, b in zip(out[v], out[100]):
                    assert np.array_equal(a, b), (rz, v)
            assert np.abs(out[100][0]).max() > 0
    finally:
        R.set_tuning("pml_cost", old[0])
        R.set_tuning("rz2d", old[1])


def test_zc3d_values_give_identical_fields():
    """Every z chunk length and both 3D tile orders (grid order, XCD column
    groups) give bit-identical records and fields."""
    from rockseis_amd import _lib as R, core
    old = R.get_tuning("zc3d"), R.get_tuning("xcd3d")
    try:
        out = {}
        for xcd in (0, 1):
            R.set_tuning("xcd3d", xcd)
            for v in (0, 3, 16, 64):
                R.set_tuning("zc3d", v)
                out[(xcd, v)] = _a3d(core)
        for key in out:
            assert np.array_equal(out[key][0], out[(0, 0)][0]), key
            assert np.array_equal(out[key][1], out[(0, 0)][1]), key
        assert np.abs(out[(0, 0)][0]).max() > 0
    finally:
        R.set_tuning("zc3d", old[0])
        R.set_tuning("xcd3d", old[1])


@pytest.mark.parametrize("kind", ["a2d", "e2d", "a3d", "a2d_batch"])
def test_autotune_picks_and_sets_a_candidate(kind):
    from rockseis_amd import _lib as R, autotune, core
    nshot = 1
    if kind == "a3d":
        m = core.ModelAcoustic3D(60, 50, 40, 8, 10.0, 10.0, 10.0)
        w = core.WavesAcoustic3D(m, 10, 1e-3)
        knob = "zc3d"
    elif kind == "e2d":
        m = core.ModelElastic2D(300, 200, 20, 5.0, 5.0)
        w = core.WavesElastic2D(m, 10, 5e-4)
        knob = "rz2d"
    else:
        m = core.ModelAcoustic2D(300, 200, 20, 10.0, 10.0)
        w = core.WavesAcoustic2D(m, 10, 1e-3)
        knob = "rz2d"
        nshot = 4 if kind == "a2d_batch" else 1
    old = R.get_tuning(knob), R.get_tuning("pml_cost")
    try:
        res = autotune.tune_waves(w, 8 if kind != "a3d" else 4, nshot=nshot, reps=5, force=True)
        assert res[knob] in autotune.CANDIDATES[knob]
        assert R.get_tuning(knob) == res[knob]
        assert len(res["ms"]) == len(autotune.CANDIDATES[knob])
        assert all(t > 0 for t in res["ms"].values())
        if kind != "a3d":  # 2D: then the XCD split's PML weight at that tile height
            assert res["pml_cost"] in autotune.CANDIDATES["pml_cost"]
            assert R.get_tuning("pml_cost") == res["pml_cost"]
            assert len(res["ms_pml_cost"]) == len(autotune.CANDIDATES["pml_cost"])
        # cached: the second call returns the same choice without timing
        assert autotune.tune_waves(w, 8 if kind != "a3d" else 4, nshot=nshot) is res
    finally:
        R.set_tuning(knob, old[0])
        R.set_tuning("pml_cost", old[1])


def _e2d_run(core, graph, srcfield="P", order=8, fs=False):
    from rockseis_amd import synth
    nx, nz, lpml, dx, dt, nt = 200, 110, 15, 5.0, 5e-4, 150
    vp = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1500, 3500, nanom=6, seed=11)
    m = core.ModelElastic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:], m.Vs[:], m.R[:] = vp, vp / 1.73, synth.gardner(vp)
    # two sources (distinct cells: the separate kernel adds with atomics)
    src = data2d(np.float32([[500.0, 20.0], [702.5, 32.5]]), nt, dt, srcfield,
                 np.stack([synth.ricker(nt, dt, 20.0), 0.5 * synth.ricker(nt, dt, 25.0)]))
    src.makeMap(m, core.SMAP)
    recs = {}
    for f in ("VX", "VZ", "P"):
        r = data2d(synth.surface_line(nx, dx, 12.5), nt, dt, f)
        r.makeMap(m, core.GMAP)
        recs[f] = r
    m.staggerModels()
    mod = core.ModellingElastic2D(m, src, order, 1)
    mod.setRecVx(recs["VX"])
    mod.setRecVz(recs["VZ"])
    mod.setRecP(recs["P"])
    mod.setGraphChunk(graph)
    mod.run()
    w = mod.waves
    return [recs[f].data.copy() for f in ("VX", "VZ", "P")] + [getattr(w, k).cpu().numpy() for k in
                                                               ("Vx", "Vz", "Sxx", "Szz", "Sxz")]


@pytest.mark.parametrize("graph,fs", [(0, False), (25, True)])
def test_fused_io_matches_separate_kernels(graph, fs):
    """Source injection and velocity gathers fused into the stress kernel
    (tuning fuse_io = 1, the default) give the same records and fields, bit for
    bit, as the separate insertSource / recordData kernels (fuse_io = 0): an
    elastic run with two explosive sources, VX, VZ and P receivers; an
    acoustic run with a pressure source.  Tile heights 1, 2 and 4."""
    from rockseis_amd import _lib as R, core
    old = R.get_tuning("fuse_io"), R.get_tuning("rz2d")
    try:
        for rz in (1, 2, 4):
            R.set_tuning("rz2d", rz)
            out = {}
            for fuse in (0, 1):
                R.set_tuning("fuse_io", fuse)
                out[fuse] = _e2d_run(core, graph, fs=fs) + list(_a2d(core, 4))
            for a, b in zip(out[0], out[1]):
                assert np.array_equal(a, b)
    finally:
        R.set_tuning("fuse_io", old[0])
        R.set_tuning("rz2d", old[1])


def test_tune_ragged_picks_and_sets_a_candidate():
    """The ragged autotuner times the elastic half steps of three windows
    of different widths for every rows-per-thread candidate, sets the
    fastest as rz2d_ragged and caches it."""
    from rockseis_amd import _lib as R, autotune, core
    old = R.get_tuning("rz2d_ragged")
    pml = core.Pml(10, 5e-4)
    pml.setSmax(core.SMAX)
    pml.computeABC()
    grids = [R.Grid(nx=w + 20, ny=1, nz=80, lpml=10, order=8, fs=0, dx=5.0, dy=1.0, dz=5.0, dt=5e-4)
             for w in (60, 130, 95)]
    try:
        res = autotune.tune_ragged(R.KIND_E2D, grids, pml.tab.data_ptr(), reps=2, force=True)
        assert res["rz2d_ragged"] in autotune.CANDIDATES["rz2d"]
        assert R.get_tuning("rz2d_ragged") == res["rz2d_ragged"]
        assert len(res["ms"]) == len(autotune.CANDIDATES["rz2d"]) and all(t > 0 for t in res["ms"].values())
        assert autotune.tune_ragged(R.KIND_E2D, grids, pml.tab.data_ptr()) is res
    finally:
        R.set_tuning("rz2d_ragged", old)

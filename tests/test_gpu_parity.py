"""HIP path (librsgpu.so through rockseis_amd.core and the C-ABI) against
the reference fp32 results: the golden fixtures written by the compiled
reference and, at larger sizes, the oracle (bit-exact with the reference).
Tolerance: relative L2 <= 1e-5 for gathers and wavefields."""
import numpy as np
import pytest

from tests.golden.make_golden import A2D_CASES, E2D_CASES
from tests.gpu_util import RTOL_L2, data2d, data3d, rel_l2

pytestmark = pytest.mark.gpu


def _core():
    from rockseis_amd import core
    return core


def _a2d_modelling(core, vp, rho, lpml, dx, dt, nt, order, fs, sf, sc, wav, rf, gc, graph=0):
    nz, nx = vp.shape
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:] = vp
    m.R[:] = rho
    src = data2d(sc, nt, dt, sf, wav)
    src.makeMap(m, core.SMAP)
    rec = data2d(gc, nt, dt, rf)
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingAcoustic2D(m, src, order, 1)
    {"P": mod.setRecP, "VX": mod.setRecVx, "VZ": mod.setRecVz}[rf](rec)
    mod.setGraphChunk(graph)
    mod.run()
    return rec.data, mod.waves


@pytest.mark.parametrize("case", range(len(A2D_CASES)))
def test_a2d_modelling_vs_golden(golden, case):
    core = _core()
    g = golden(f"a2d_case{case}")
    order, fs, sf, rf = A2D_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    print(f"a2d case {case}: order {order} fs {fs} {sf}->{rf}", flush=True)
    rec, w = _a2d_modelling(core, g["vp"], g["rho"], int(lpml), dx, dt, int(nt), order, fs, sf,
                            g["scoords"], g["wav"], rf, g["gcoords"], graph=0)
    assert rel_l2(rec, g["rec"]) <= RTOL_L2
    for k in ("P", "Vx", "Vz"):
        assert rel_l2(getattr(w, k).cpu().numpy(), g[k]) <= RTOL_L2, k


def test_a2d_graph_replay_matches_eager(golden):
    core = _core()
    g = golden("a2d_case1")
    order, fs, sf, rf = A2D_CASES[1]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    args = (g["vp"], g["rho"], int(lpml), dx, dt, int(nt), order, fs, sf, g["scoords"], g["wav"], rf,
            g["gcoords"])
    r0, w0 = _a2d_modelling(core, *args, graph=0)
    r1, w1 = _a2d_modelling(core, *args, graph=32)   # graphs of 1..32 steps: 9 x 32 + 8 + 4
    assert np.array_equal(r0, r1)
    assert np.array_equal(w0.P.cpu().numpy(), w1.P.cpu().numpy())


@pytest.mark.parametrize("chunk", (1, 16, 128))
def test_modelling_plan_ranges_match_eager(chunk):
    """rsg_modelling_plan_run over consecutive step ranges of awkward lengths
    (1, 19, 37, the rest) from one plan, with a snapshot set, gives the
    records, fields and snapshots of the eager per-step loop bit for bit."""
    from rockseis_amd import synth
    core = _core()
    nx, nz, lpml, dx, dt, nt = 150, 90, 12, 5.0, 5e-4, 140
    vp = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1500, 3500, nanom=4, seed=5)
    out = {}
    for graph in (0, chunk):
        m = core.ModelElastic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, True)
        m.Vp[:], m.Vs[:], m.R[:] = vp, vp / 1.73, synth.gardner(vp)
        src = data2d(np.float32([[370.0, 20.0]]), nt, dt, "P", synth.ricker(nt, dt, 20.0))
        src.makeMap(m, core.SMAP)
        recs = {}
        for f in ("VX", "VZ", "P"):
            r = data2d(synth.surface_line(nx, dx, 12.5), nt, dt, f)
            r.makeMap(m, core.GMAP)
            recs[f] = r
        m.staggerModels()
        mod = core.ModellingElastic2D(m, src, 8, 3)
        mod.setRecVx(recs["VX"])
        mod.setRecVz(recs["VZ"])
        mod.setRecP(recs["P"])
        mod.setSnapSxz()
        mod.setGraphChunk(graph)
        mod.prepare()
        it = 0
        for n in (1, 19, 37, nt - 57):
            mod.launch(it, n)
            it += n
        mod.finish()
        w = mod.waves
        out[graph] = ([recs[f].data.copy() for f in ("VX", "VZ", "P")] + [mod.snaps["SXZ"]] +
                      [getattr(w, k).cpu().numpy() for k in ("Vx", "Vz", "Sxx", "Szz", "Sxz")])
    print(f"modelling plan, graphs up to {chunk} steps: {len(out[0])} arrays compared", flush=True)
    for a, b in zip(out[0], out[chunk]):
        assert np.array_equal(a, b)


def test_a2d_per_step_api_and_snapshots(golden):
    """forwardstepVelocity / forwardstepStress / insertSource / recordData
    one call at a time, as modelling.cpp:242-286 drives them."""
    core = _core()
    g = golden("a2d_case3")
    order, fs, sf, rf = A2D_CASES[3]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:] = g["vp"]
    m.R[:] = g["rho"]
    src = data2d(g["scoords"], nt, dt, sf, g["wav"])
    src.makeMap(m, core.SMAP)
    rec = data2d(g["gcoords"], nt, dt, rf)
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    w = core.WavesAcoustic2D(m, nt, dt, 0.0)
    der = core.Der(w.getNx_pml(), 1, w.getNz_pml(), dx, 1.0, dx, order)
    for it in range(nt):
        w.forwardstepVelocity(m, der)
        w.forwardstepStress(m, der)
        w.insertSource(m, src, core.SMAP, it)
        w.recordData(rec, core.GMAP, it)
        if it == 10:
            assert rel_l2(w.getP().cpu().numpy(), g["snap10"]) <= RTOL_L2
        if it == 150:
            assert rel_l2(w.getP().cpu().numpy(), g["snap150"]) <= RTOL_L2
    assert rel_l2(rec.data, g["rec"]) <= RTOL_L2


@pytest.mark.parametrize("case", range(len(E2D_CASES)))
def test_e2d_vs_golden(golden, case):
    core = _core()
    g = golden(f"e2d_case{case}")
    order, fs, sf, rf, recip, nbw = E2D_CASES[case]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    print(f"e2d case {case}: order {order} fs {fs} {sf}->{rf} recip {recip} nbw {nbw}", flush=True)
    m = core.ModelElastic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:] = g["vp"]
    m.Vs[:] = g["vs"]
    m.R[:] = g["rho"]
    src = data2d(g["scoords"], nt, dt, sf, g["wav"])
    src.makeMap(m, core.SMAP)
    rec = data2d(g["gcoords"], nt, dt, rf)
    rec.setReciprocity(recip)
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingElastic2D(m, src, order, 1)
    {"P": mod.setRecP, "VX": mod.setRecVx, "VZ": mod.setRecVz}[rf](rec)
    mod.run()
    w = mod.waves
    assert rel_l2(rec.data, g["rec"]) <= RTOL_L2
    for _ in range(nbw):
        w.backwardstepVelocity(m, mod.der)
        w.backwardstepStress(m, mod.der)
    for k in ("Vx", "Vz", "Sxx", "Szz", "Sxz"):
        assert rel_l2(getattr(w, k).cpu().numpy(), g[k]) <= RTOL_L2, k


@pytest.mark.parametrize("case", range(3))
def test_a3d_vs_golden(golden, case):
    core = _core()
    g = golden(f"a3d_case{case}")
    nx, ny, nz, lpml, dx, dt, nt, order = g["grid"]
    nx, ny, nz, lpml, nt, order = (int(v) for v in (nx, ny, nz, lpml, nt, order))
    sf, rf = [("P", "P"), ("VY", "VZ"), ("VX", "P")][case]
    print(f"a3d case {case}: order {order} {sf}->{rf}", flush=True)
    m = core.ModelAcoustic3D(nx, ny, nz, lpml, dx, dx, dx, 0.0, 0.0, 0.0, False)
    m.Vp[:] = g["vp"]
    m.R[:] = g["rho"]
    src = data3d(g["scoords"], nt, dt, sf, g["wav"])
    src.makeMap(m, core.SMAP)
    rec = data3d(g["gcoords"], nt, dt, rf)
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingAcoustic3D(m, src, order, 1)
    {"P": mod.setRecP, "VZ": mod.setRecVz}[rf](rec)
    mod.run()
    assert rel_l2(rec.data, g["rec"]) <= RTOL_L2
    for k in ("P", "Vz"):
        assert rel_l2(getattr(mod.waves, k).cpu().numpy(), g[k]) <= RTOL_L2, k


@pytest.mark.parametrize("order", [4, 8])
def test_e2d_larger_grid_vs_oracle(order):
    """300x200 interior, lpml 20, 800 steps, free surface, explosive source,
    Vx/Vz/P receivers -- against the oracle (bit-exact with the reference)."""
    from oracle import runs
    from rockseis_amd import synth
    core = _core()
    nx, nz, lpml, dx, dt, nt = 300, 200, 20, 5.0, 4.5e-4, 800
    vp = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1500, 4000, nanom=20, seed=7)
    vs = (vp / 1.73).astype(np.float32)
    rho = synth.gardner(vp)
    sc = np.float32([[750.0, 5.0]])
    wav = synth.ricker(nt, dt, 20.0)[None]
    gc = synth.surface_line(nx, dx, 10.0)
    ref = {}
    for rf in ("VX", "VZ"):
        ref[rf] = runs.e2d_run(vp, vs, rho, lpml, dx, dx, dt, nt, order, 1, "P", sc, wav, rf, gc)
    m = core.ModelElastic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, True)
    m.Vp[:], m.Vs[:], m.R[:] = vp, vs, rho
    src = data2d(sc, nt, dt, "P", wav)
    src.makeMap(m, core.SMAP)
    recs = {}
    for rf in ("VX", "VZ"):
        recs[rf] = data2d(gc, nt, dt, rf)
        recs[rf].makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingElastic2D(m, src, order, 1)
    mod.setRecVx(recs["VX"])
    mod.setRecVz(recs["VZ"])
    mod.setGraphChunk(50)
    mod.run()
    for rf in ("VX", "VZ"):
        e = rel_l2(recs[rf].data, ref[rf]["rec"])
        print(f"order {order} {rf}: rel L2 {e:.3e}")
        assert e <= RTOL_L2
    for k in ("Vx", "Vz", "Sxx", "Szz", "Sxz"):
        assert rel_l2(getattr(mod.waves, k).cpu().numpy(), ref["VZ"][k]) <= RTOL_L2, k

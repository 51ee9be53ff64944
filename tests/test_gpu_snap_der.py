"""Device snapshots of the modelling loop (rsg_modelling_run_snap, the
Modelling*::setSnap* path) and the standalone Der derivatives
(rsg_der_apply), against the reference's fixtures."""
import ctypes as C

import numpy as np
import pytest

from tests.golden.make_golden import A2D_CASES
from tests.gpu_util import RTOL_L2, data2d, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", (0, 16))
def test_modelling_snapshots_vs_golden(golden, graph):
    """Snapshots every 10 steps; the slots at steps 10 and 150 against the
    reference's pressure after those steps (interior cells)."""
    from rockseis_amd import core
    g = golden("a2d_case3")
    order, fs, sf, rf = A2D_CASES[3]
    nx, nz, lpml, dx, dt, nt = g["grid"]
    nx, nz, lpml, nt = int(nx), int(nz), int(lpml), int(nt)
    print(f"snapshots a2d_case3 graph {graph}", flush=True)
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, fs)
    m.Vp[:], m.R[:] = g["vp"], g["rho"]
    src = data2d(g["scoords"], nt, dt, sf, g["wav"])
    src.makeMap(m, core.SMAP)
    rec = data2d(g["gcoords"], nt, dt, rf)
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingAcoustic2D(m, src, order, 10)
    mod.setRecP(rec)
    mod.setSnapP()
    mod.setSnapVx()
    mod.setGraphChunk(graph)
    mod.run()
    snaps = mod.snaps["P"]
    assert snaps.shape == ((nt - 1) // 10 + 1, nz, nx)
    for it in (10, 150):
        ref = g[f"snap{it}"][lpml:-lpml, lpml:-lpml]
        e = rel_l2(snaps[it // 10], ref)
        print(f"  P snapshot at step {it}: rel L2 {e:.2e}", flush=True)
        assert e <= RTOL_L2
    # the last slot is the final state's interior
    assert np.array_equal(mod.snaps["VX"][-1], mod.waves.Vx.cpu().numpy()[lpml:-lpml, lpml:-lpml]) or \
        (nt - 1) % 10 != 0
    assert rel_l2(rec.data, g["rec"]) <= RTOL_L2


def test_der_apply_bit_exact(golden):
    """Der::dd{x,y,z}_{fw,bw} on a 17x18x19 array at every order, equal to
    the reference's floats."""
    import torch
    from rockseis_amd import _lib as R
    g = golden("der")
    f = torch.from_numpy(np.ascontiguousarray(g["f"])).cuda()
    nz, ny, nx = f.shape
    dxyz = [float(v) for v in g["d"]]
    df = torch.empty_like(f)
    n = 0
    for key in g.files:
        if not key.startswith("o"):
            continue
        order, which = int(key[1:key.index("_")]), int(key[key.index("w") + 1:])
        axis, bw = which % 3, which // 3
        if min(nx, ny, nz) < 2 * order + 1:
            continue
        R.check(R.lib().rsg_der_apply(axis, bw, f.data_ptr(), df.data_ptr(), nx, ny, nz, dxyz[axis], order, None),
                "rsg_der_apply")
        torch.cuda.synchronize()
        np.testing.assert_array_equal(df.cpu().numpy(), g[key], err_msg=key)
        n += 1
    print(f"rsg_der_apply: {n} order/direction cases bit-exact", flush=True)
    assert n == 32

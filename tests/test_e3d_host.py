"""ModelElastic3D::staggerModels on the host (rsg_stagger_e3d) bit for bit
against the compiled reference's staggered model (golden fixture), with and
without free surface; no GPU needed."""
import ctypes as C

import numpy as np
import pytest

from rockseis_amd import _lib as R


@pytest.mark.parametrize("fs", [0, 1])
def test_stagger_e3d_bit_exact(golden, fs):
    g = golden("e3d_stagger")
    vp, vs, rho = (np.ascontiguousarray(g[k], np.float32) for k in ("vp", "vs", "rho"))
    nz, ny, nx = vp.shape
    lpml = int(g["lpml"])
    out = [np.zeros((nz + 2 * lpml, ny + 2 * lpml, nx + 2 * lpml), np.float32) for _ in range(8)]
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    R.check(R.lib().rsg_stagger_e3d(p(vp), p(vs), p(rho), nx, ny, nz, lpml, fs, *(p(a) for a in out)),
            "rsg_stagger_e3d")
    for i, (a, b) in enumerate(zip(out, g[f"stag_fs{fs}"])):
        assert np.array_equal(a, b), ("L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz")[i]

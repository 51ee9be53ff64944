"""Record resampling on the device (rsg_interp through
rockseis_amd.core.Interp) against the reference's Interp::interp (golden
fixtures) for LINEAR, BSPLINE and SINC.  The kernel uses explicitly rounded
operations in the reference's order; the one difference is the device's
double sin, so the bound is relative L2 1e-6."""
import numpy as np
import pytest

from tests.golden.make_golden import INTERP_CASES
from tests.gpu_util import rel_l2

pytestmark = pytest.mark.gpu


def test_interp_vs_reference(golden):
    import torch
    from rockseis_amd import core
    g = golden("interp")
    for k, (ntr, n1, d1, o1, n2, d2, o2) in enumerate(INTERP_CASES):
        x = torch.from_numpy(g[f"in{k}"]).cuda()
        for mode in (core.LINEAR, core.BSPLINE, core.SINC):
            out = core.Interp(mode).interp_device(x, d1, o1, n2, d2, o2).cpu().numpy()
            want = g[f"out{k}_m{mode}"]
            assert np.array_equal(out == 0, want == 0), (k, mode)
            assert rel_l2(out, want) <= 1e-6, (k, mode)


def test_interp_data2d_roundtrip():
    """Interp.interp on Data2D objects copies the coordinates and resamples
    the traces as the oracle does (a 2x upsampling with SINC)."""
    from oracle import fn
    from rockseis_amd import core
    a = core.Data2D(3, 400, 1e-3)
    a.data[:] = np.sin(np.arange(400) * 0.05)[None] * np.float32([[1], [2], [3]])
    a.gcoords[:] = np.float32([[1, 2], [3, 4], [5, 6]])
    b = core.Data2D(3, 800, 5e-4)
    core.Interp(core.SINC).interp(a, b)
    assert np.array_equal(b.gcoords, a.gcoords)
    want = np.zeros((3, 800), np.float32)
    fn("interp")(2, np.ascontiguousarray(a.data).ravel(), 3, 400, np.float32(1e-3), np.float32(0), want.reshape(-1),
                 800, np.float32(5e-4), np.float32(0))
    assert rel_l2(b.data, want) <= 1e-6

"""rsg_misfit_residuals_noalloc (scratch in a fixed device array, no
allocation) against rsg_misfit_residuals through the C-ABI: residuals bit
for bit and the misfit to 1e-12 (both reduce in double with atomics, so the
last bits of the sum follow the scheduling), on two calls in a row, the
second with smaller inputs, for DIFFERENCE, CORRELATION with its
97th-percentile clip (2D) and without (3D), with and without data weights;
and the limits (ADAPTIVE types, more than 64 shots, inputs past the
scratch) are refused with a message."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(rng, ntr, nt, nshot, dim, weights):
    import torch
    mod = rng.standard_normal((nshot, ntr, nt)).astype(np.float32)
    rec = (mod + 0.3 * rng.standard_normal(mod.shape)).astype(np.float32)
    mp = rng.integers(0, 50, (nshot, ntr, dim)).astype(np.int32)
    mp[:, ::7, 0] = -1                          # traces off the grid count in no misfit
    wei = rng.uniform(0.5, 1.5, mod.shape).astype(np.float32) if weights else None
    dev = [torch.from_numpy(a).cuda() for a in (mod, rec, mp)]
    dev.append(None if wei is None else torch.from_numpy(wei).cuda())
    return dev


def _call(fn, typ, dev, dim, scaling):
    import torch
    from rockseis_amd import _lib as R
    mod, rec, mp, wei = dev
    nshot, ntr, nt = mod.shape
    res = torch.zeros_like(mod)
    mis = C.c_double(0.0)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    R.check(fn(typ, mod.data_ptr(), rec.data_ptr(), None if wei is None else wei.data_ptr(), res.data_ptr(),
               mp.data_ptr(), dim, ntr, nt, nshot, scaling, C.byref(mis), st), fn.__name__)
    torch.cuda.synchronize()
    return mis.value, res.cpu().numpy()


@pytest.mark.parametrize("typ,dim,weights,scaling", [(0, 2, False, 1), (0, 2, True, -1), (1, 2, False, 1),
                                                     (1, 2, True, 1), (1, 3, False, -1), (4, 3, True, 1)])
def test_noalloc_equals_allocating_entry_point(typ, dim, weights, scaling):
    from rockseis_amd import _lib as R
    L = R.lib()
    rng = np.random.default_rng(100 + 10 * typ + dim)
    for ntr, nt, nshot in ((90, 400, 5), (33, 170, 2)):   # the second call smaller
        dev = _inputs(rng, ntr, nt, nshot, dim, weights)
        m_ref, r_ref = _call(L.rsg_misfit_residuals, typ, dev, dim, scaling)
        m_na, r_na = _call(L.rsg_misfit_residuals_noalloc, typ, dev, dim, scaling)
        print(f"noalloc type {typ} dim {dim} weights {weights} {nshot}x{ntr}x{nt}: misfit {m_na:.9e} "
              f"(allocating {m_ref:.9e})", flush=True)
        assert np.array_equal(r_na.view(np.uint32), r_ref.view(np.uint32))
        assert abs(m_na - m_ref) <= 1e-12 * abs(m_ref)
        if typ == 1 and dim == 2:
            assert np.count_nonzero(r_na == 0) > nshot * ntr * nt * 0.03 * 0.5  # the clip zeroed the top 3%


def test_noalloc_refuses_past_its_limits():
    import torch
    from rockseis_amd import _lib as R
    L = R.lib()
    rng = np.random.default_rng(5)
    small = _inputs(rng, 8, 16, 2, 2, False)
    for typ in (2, 3):
        with pytest.raises(R.RsgError, match="ADAPTIVE"):
            _call(L.rsg_misfit_residuals_noalloc, typ, small, 2, 1)
    many = _inputs(rng, 2, 4, 65, 2, False)
    with pytest.raises(R.RsgError, match="64 shots"):
        _call(L.rsg_misfit_residuals_noalloc, 0, many, 2, 1)
    # CORRELATION norms of 64 x 6000 traces need 4.6 MB, past the 4 MiB scratch
    big = _inputs(rng, 6000, 1, 64, 2, False)
    with pytest.raises(R.RsgError, match="scratch"):
        _call(L.rsg_misfit_residuals_noalloc, 1, big, 2, 1)
    # still usable afterwards
    m0, r0 = _call(L.rsg_misfit_residuals, 1, small, 2, 1)
    m1, r1 = _call(L.rsg_misfit_residuals_noalloc, 1, small, 2, 1)
    assert np.array_equal(r0, r1)
    torch.cuda.synchronize()

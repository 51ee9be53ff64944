"""RSS files into and out of device memory (rsg_rss_*, rockseis_amd.rss) and
the in-memory image stacking (rsg_stack_image_2d) against the compiled
reference (oracle/_ref): files the reference's Data2D / Data3D::write
produced are read here bit for bit, files written here are read by the
reference's Data2D / Data3D::read bit for bit and are byte-identical to the
reference's own, and stacking local images in device memory gives the
global image of Image2D::createEmpty + stackImage_parallel + read bit for
bit (offsets from truncated origin ratios, pads, images partly outside the
global one, two subsurface offsets)."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) not present")]


def _ref():
    R = oracle.ref()
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    R.ref_rss_write_data.argtypes = [C.c_char_p, C.c_int, f32, f32, C.c_int, C.c_int, C.c_float, C.c_float]
    R.ref_rss_read_data.argtypes = [C.c_char_p, C.c_int, f32, f32, C.c_int, C.c_int]
    R.ref_stack_images_2d.argtypes = ([C.c_char_p] + [C.c_int] * 4 + [C.c_float] * 4 + [C.c_int, i32, i32, f32, f32,
                                                                                         f32] + [C.c_int] * 4 + [f32])
    return R


def _traces(dim, ntrace, nt, seed):
    rng = np.random.default_rng(seed)
    coords = (rng.random((ntrace, 2 * dim)) * 1000).astype(np.float32)
    data = rng.standard_normal((ntrace, nt)).astype(np.float32)
    return coords, data


@pytest.mark.parametrize("dim", [2, 3])
def test_read_reference_written_traces(tmp_path, dim):
    import torch
    from rockseis_amd import rss
    R = _ref()
    coords, data = _traces(dim, 37, 251, dim)
    path = str(tmp_path / f"ref{dim}.rss")
    assert R.ref_rss_write_data(path.encode(), dim == 3, coords, data, 37, 251, 0.002, -0.1) == 0
    c, d, h = rss.read_traces_device(path, dim)
    torch.cuda.synchronize()
    assert np.array_equal(c, coords)
    assert np.array_equal(d.cpu().numpy(), data)
    assert h["n"][:2] == [251, 37] and h["nheader"] == 2 * dim and h["start"] == 249
    # the numpy reader sees the same traces
    r = rss.read_data(path)
    assert np.array_equal(np.asarray(r["data"]), data)


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_reads_traces_written_here(tmp_path, dim):
    import torch
    from rockseis_amd import rss
    R = _ref()
    coords, data = _traces(dim, 29, 180, 10 + dim)
    path, rpath = str(tmp_path / f"ours{dim}.rss"), str(tmp_path / f"ref{dim}.rss")
    dev = torch.from_numpy(data).cuda()
    rss.write_traces_device(path, coords, dev, 0.004, 0.25, dim)
    c, d = np.zeros_like(coords), np.zeros_like(data)
    assert R.ref_rss_read_data(path.encode(), dim == 3, c, d, 29, 180) == 0
    assert np.array_equal(c, coords) and np.array_equal(d, data)
    # byte-identical to the reference's own file of the same traces
    assert R.ref_rss_write_data(rpath.encode(), dim == 3, coords, data, 29, 180, 0.004, 0.25) == 0
    assert open(path, "rb").read() == open(rpath, "rb").read()


def test_data2d_file_methods(tmp_path):
    import torch
    from rockseis_amd import core
    R = _ref()
    coords, data = _traces(2, 11, 90, 5)
    path = str(tmp_path / "d.rss")
    R.ref_rss_write_data(path.encode(), 0, coords, data, 11, 90, 0.001, 0.0)
    d = core.Data2D.fromFile(path, "VX")
    assert (d.ntrace, d.nt) == (11, 90) and d.dt == pytest.approx(0.001)
    assert np.array_equal(d.scoords, coords[:, :2]) and np.array_equal(d.gcoords, coords[:, 2:])
    assert np.array_equal(d.data, data)
    assert d.device_data.is_cuda
    d.data *= 2
    out = str(tmp_path / "e.rss")
    d.setFile(out)
    d.write()
    torch.cuda.synchronize()
    c2, d2 = np.zeros_like(coords), np.zeros_like(data)
    assert R.ref_rss_read_data(out.encode(), 0, c2, d2, 11, 90) == 0
    assert np.array_equal(d2, 2 * data) and np.array_equal(c2, coords)


def test_regular_round_trip(tmp_path):
    import torch
    from rockseis_amd import rss
    x = torch.randn(7, 13, 21, device="cuda")
    path = str(tmp_path / "m.rss")
    rss.write_regular_device(path, x, (21, 13, 7), (5.0, 5.0, 5.0), (0.0, 10.0, -5.0))
    h = rss.read_header(path)  # the numpy reader agrees on the header
    assert tuple(h.n[:3]) == (21, 13, 7) and h.o[1] == 10.0
    assert np.array_equal(np.asarray(rss.read(path)[1]).reshape(7, 13, 21), x.cpu().numpy())
    y = rss.read_regular_device(path)
    torch.cuda.synchronize()
    assert torch.equal(x, y)


def test_wrong_kind_is_rejected(tmp_path):
    import torch
    from rockseis_amd import _lib as RL
    from rockseis_amd import rss
    path = str(tmp_path / "m.rss")
    rss.write_regular_device(path, torch.zeros(4, 5, device="cuda"), (5, 4), (1.0, 1.0), (0.0, 0.0))
    with pytest.raises(RL.RsgError):
        rss.read_traces_device(path, 2)
    with pytest.raises(RL.RsgError):
        rss.read_traces_device(str(tmp_path / "missing.rss"), 2)


@pytest.mark.parametrize("pads", [(0, 0, 0, 0), (3, 2, 1, 4)])
def test_stack_images_vs_reference(tmp_path, pads):
    import torch
    from rockseis_amd import rss
    R = _ref()
    nxg, nzg, nhx, nhz, dx, dz = 60, 25, 2, 1, 7.5, 5.0
    oxg, ozg = 13.0, -2.5
    # local windows: inside, overlapping the left edge, past the right edge,
    # and an origin whose ratio truncates toward zero
    loc = [(20, 25, 13.0 + 7.5 * 5, -2.5), (18, 20, 13.0 - 7.5 * 6, 2.5), (25, 25, 13.0 + 7.5 * 45, -2.5),
           (16, 22, 13.0 + 7.5 * 10.6, -2.5 + 5.0 * 1.7)]
    rng = np.random.default_rng(7)
    imgs = [rng.standard_normal((nhz, nhx, nz, nx)).astype(np.float32) for nx, nz, _, _ in loc]
    want = np.zeros(nxg * nzg * nhx * nhz, np.float32)
    assert R.ref_stack_images_2d(str(tmp_path / "g.rss").encode(), nxg, nzg, nhx, nhz, dx, dz, oxg, ozg, len(loc),
                                 np.int32([l[0] for l in loc]), np.int32([l[1] for l in loc]),
                                 np.float32([l[2] for l in loc]), np.float32([l[3] for l in loc]),
                                 np.concatenate([i.ravel() for i in imgs]), *pads, want) == 0
    g = torch.zeros((nhz, nhx, nzg, nxg), device="cuda")
    gd = rss.image_desc(nxg, nzg, dx, dz, oxg, ozg, nhx, nhz)
    for (nx, nz, ox, oz), im in zip(loc, imgs):
        rss.stack_image_2d(g, gd, torch.from_numpy(im).cuda(), rss.image_desc(nx, nz, dx, dz, ox, oz, nhx, nhz), *pads)
    torch.cuda.synchronize()
    assert np.array_equal(g.cpu().numpy().ravel(), want)
    assert np.count_nonzero(want) > 0

"""RockSeis RSS files: read and write with numpy, memory-mapped.

The format is the reference's File class (file.cpp:1-330, file.h:15-25):
a fixed header of

    magic "R0CKS" (5 bytes), data_format int32, header_format int32,
    type int32, Nheader uint64, Ndims uint64,
    n[9] uint64, d[9] float64, o[9] float64

(249 bytes, startofdata), then the samples.  Files of type DATA2D / DATA3D
hold one record per trace: Nheader header values (source x, z, receiver
x, z in 2D, data.cpp:600-630; source x, y, z, receiver x, y, z in 3D) then
n[0] samples.  Other types are a flat array, n[0] fastest.

Models and images are REGULAR with the axes the reference reads
(model.cpp:509-514): n1 = nx, n2 = ny (1 in 2D), n3 = nz; images add the
subsurface offsets n4 = nhx, n5 = nhy, n6 = nhz (image.cpp:545-550).
Snapshots are SNAPSHOT with n4 = the number of snapshots (snap.cpp:26-42).

Reading returns numpy memmaps of the sample block (no copy), so a large
record or snapshot file is paged in only where it is used.

The device path (read_traces_device, write_traces_device,
read_regular_device, write_regular_device) goes through the C-ABI
(rsg_rss_*): one read or write of the whole file through pinned host memory
and one strided copy, the samples landing on or leaving the GPU.
stack_image_2d replaces the inversion driver's file stacking
(Image2D::stackImage_parallel, image.cpp:533-600) with images in device
memory (rsg_stack_image_2d).
"""
from __future__ import annotations

import dataclasses
import os
import struct

import numpy as np

MAGIC = b"R0CKS"
MAXDIMS = 9
HEADER_BYTES = 5 + 3 * 4 + 2 * 8 + MAXDIMS * (8 + 8 + 8)

# rs_datatype (file.h:24)
GENERIC, REGULAR, DATA2D, DATA3D, SNAPSHOT, EDGESNAP, CHECKPOINT, KEYMAP, SORTMAP, TTABLE = range(10)

_FMT = {4: np.float32, 8: np.float64}


@dataclasses.dataclass
class Header:
    data_format: int = 4
    header_format: int = 0
    type: int = GENERIC
    nheader: int = 0
    ndims: int = 1
    n: tuple = (0,) * MAXDIMS
    d: tuple = (0.0,) * MAXDIMS
    o: tuple = (0.0,) * MAXDIMS

    def pack(self) -> bytes:
        n = list(self.n) + [0] * (MAXDIMS - len(self.n))
        d = list(self.d) + [0.0] * (MAXDIMS - len(self.d))
        o = list(self.o) + [0.0] * (MAXDIMS - len(self.o))
        return (MAGIC + struct.pack("<iiiQQ", self.data_format, self.header_format, self.type, self.nheader,
                                    self.ndims)
                + struct.pack("<%dQ" % MAXDIMS, *[int(v) for v in n])
                + struct.pack("<%dd" % MAXDIMS, *[float(v) for v in d])
                + struct.pack("<%dd" % MAXDIMS, *[float(v) for v in o]))

    @property
    def ntot(self) -> int:
        """Geometry::getNtot (geometry.cpp:61-72): product of the non-zero sizes."""
        t = 1
        for v in self.n:
            if v > 0:
                t *= int(v)
        return t


def read_header(path: str) -> Header:
    with open(path, "rb") as fh:
        raw = fh.read(HEADER_BYTES)
    if len(raw) < HEADER_BYTES or raw[:5] != MAGIC:
        raise ValueError(f"{path}: not an RSS file")
    df, hf, ty, nh, nd = struct.unpack_from("<iiiQQ", raw, 5)
    off = 5 + 3 * 4 + 2 * 8
    n = struct.unpack_from("<%dQ" % MAXDIMS, raw, off)
    d = struct.unpack_from("<%dd" % MAXDIMS, raw, off + 8 * MAXDIMS)
    o = struct.unpack_from("<%dd" % MAXDIMS, raw, off + 16 * MAXDIMS)
    return Header(df, hf, ty, nh, nd, tuple(n), tuple(d), tuple(o))


def _dtype(fmt: int):
    if fmt not in _FMT:
        raise ValueError(f"unsupported sample size {fmt}")
    return np.dtype(_FMT[fmt]).newbyteorder("<")


def read(path: str, mode: str = "r"):
    """(header, samples).  Trace files give samples [ntrace][nt] and the
    trace headers [ntrace][nheader] as a second memmap: (header, samples,
    trace_headers).  Other files give the flat sample array."""
    h = read_header(path)
    dt = _dtype(h.data_format)
    size = os.path.getsize(path) - HEADER_BYTES
    if h.nheader:
        n1 = int(h.n[0])
        ntr = h.ntot // max(n1, 1)
        hdt = _dtype(h.header_format)
        if hdt != dt:
            raise ValueError(f"{path}: mixed header and sample formats are not supported")
        rec = h.nheader + n1
        if size < rec * ntr * dt.itemsize:
            raise ValueError(f"{path}: file shorter than its header says")
        m = np.memmap(path, dtype=dt, mode=mode, offset=HEADER_BYTES, shape=(ntr, rec))
        return h, m[:, h.nheader:], m[:, :h.nheader]
    if size < h.ntot * dt.itemsize:
        raise ValueError(f"{path}: file shorter than its header says")
    return h, np.memmap(path, dtype=dt, mode=mode, offset=HEADER_BYTES, shape=(h.ntot,))


def write(path: str, h: Header, samples: np.ndarray, trace_headers: np.ndarray | None = None) -> None:
    dt = _dtype(h.data_format)
    with open(path, "wb") as fh:
        fh.write(h.pack())
        if h.nheader:
            s = np.ascontiguousarray(samples, dtype=dt).reshape(-1, int(h.n[0]))
            th = np.ascontiguousarray(trace_headers, dtype=dt).reshape(s.shape[0], h.nheader)
            fh.write(np.concatenate([th, s], axis=1).tobytes())
        else:
            fh.write(np.ascontiguousarray(samples, dtype=dt).tobytes())


# ---- typed helpers ---------------------------------------------------------

def write_model2d(path: str, a: np.ndarray, dx: float, dz: float, ox: float = 0.0, oz: float = 0.0) -> None:
    """A [nz][nx] model (ModelAcoustic2D reads n1 = nx, n3 = nz)."""
    nz, nx = a.shape
    h = Header(4, 0, REGULAR, 0, 1, (nx, 1, nz), (dx, 1.0, dz), (ox, 0.0, oz))
    write(path, h, a)


def write_model3d(path: str, a: np.ndarray, dx, dy, dz, ox=0.0, oy=0.0, oz=0.0) -> None:
    nz, ny, nx = a.shape
    h = Header(4, 0, REGULAR, 0, 1, (nx, ny, nz), (dx, dy, dz), (ox, oy, oz))
    write(path, h, a)


def read_model(path: str) -> tuple[np.ndarray, Header]:
    """[nz][nx] (2D) or [nz][ny][nx] (3D) copy of a REGULAR file."""
    h, s = read(path)
    nx, ny, nz = (max(int(v), 1) for v in h.n[:3])
    a = np.array(s[: nx * ny * nz])
    return (a.reshape(nz, nx) if ny == 1 else a.reshape(nz, ny, nx)), h


def write_data2d(path: str, data: np.ndarray, scoords: np.ndarray, gcoords: np.ndarray, dt: float,
                 ot: float = 0.0) -> None:
    """Data2D::write (data.cpp:590-630): data [ntrace][nt], coordinates
    [ntrace][2] (x, z); each trace carries sx, sz, gx, gz."""
    data = np.asarray(data, np.float32)
    ntr, nt = data.shape
    h = Header(4, 4, DATA2D, 4, 1, (nt, ntr), (dt, 1.0), (ot, 0.0))
    th = np.concatenate([np.asarray(scoords, np.float32).reshape(ntr, 2),
                         np.asarray(gcoords, np.float32).reshape(ntr, 2)], axis=1)
    write(path, h, data, th)


def write_data3d(path: str, data: np.ndarray, scoords: np.ndarray, gcoords: np.ndarray, dt: float,
                 ot: float = 0.0) -> None:
    data = np.asarray(data, np.float32)
    ntr, nt = data.shape
    h = Header(4, 4, DATA3D, 6, 1, (nt, ntr), (dt, 1.0), (ot, 0.0))
    th = np.concatenate([np.asarray(scoords, np.float32).reshape(ntr, 3),
                         np.asarray(gcoords, np.float32).reshape(ntr, 3)], axis=1)
    write(path, h, data, th)


def read_data(path: str) -> dict:
    """Trace file -> dict(data [ntr][nt], scoords, gcoords, dt, ot) (copies)."""
    h, s, th = read(path)
    nh = h.nheader
    th = np.array(th)
    half = nh // 2
    return {"data": np.array(s), "scoords": th[:, :half], "gcoords": th[:, half:], "dt": h.d[0],
            "ot": h.o[0], "header": h}


def read_snapshots(path: str) -> tuple[np.ndarray, Header]:
    """SNAPSHOT file -> [snapnt][nz][nx] (2D) or [snapnt][nz][ny][nx] view."""
    h, s = read(path)
    nx, ny, nz, ns = (max(int(v), 1) for v in h.n[:4])
    a = s[: nx * ny * nz * ns]
    return (a.reshape(ns, nz, nx) if ny == 1 else a.reshape(ns, nz, ny, nx)), h


def read_image2d(path: str) -> tuple[np.ndarray, Header]:
    """Image2D file -> [nhz][nhx][nz][nx] copy (image.cpp ki2D order)."""
    h, s = read(path)
    nx, _, nz, nhx, _, nhz = (max(int(v), 1) for v in h.n[:6])
    return np.array(s[: nx * nz * nhx * nhz]).reshape(nhz, nhx, nz, nx), h


# --------------------------------------------------------------------------
# Device path (C-ABI)
# --------------------------------------------------------------------------

# (the torch and librsgpu imports stay inside these functions, so that the
# numpy reader above needs neither)
_KIND = {2: (DATA2D, 4, "rsg_rss_read_data2d", "rsg_rss_write_data2d"),
         3: (DATA3D, 6, "rsg_rss_read_data3d", "rsg_rss_write_data3d")}


def _rt():
    import ctypes as C

    import torch

    from . import _lib as R
    return C, torch, R


def _stream():
    C, torch, _ = _rt()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev():
    _, torch, _ = _rt()
    return torch.device("cuda", torch.cuda.current_device())


def _header_c(path: str) -> dict:
    """The file's header: data_format, header_format, type, nheader, ndims,
    n, d, o (nine each) and the byte offset of the data."""
    C, torch, R = _rt()
    h = R.RssHeader()
    R.check(R.lib().rsg_rss_read_header(str(path).encode(), C.byref(h)), "rsg_rss_read_header")
    return dict(data_format=h.data_format, header_format=h.header_format, type=h.type, nheader=h.nheader,
                ndims=h.ndims, n=list(h.n), d=list(h.d), o=list(h.o), start=h.start)


def read_traces_device(path: str, dim: int = 2):
    """A DATA2D (dim 2) or DATA3D (dim 3) file through rsg_rss_read_data2d/3d:
    (coords [ntrace, 4 or 6] on the host: sx, (sy,) sz, gx, (gy,) gz;
    samples [ntrace, nt] as a device tensor; header dict)."""
    C, torch, R = _rt()
    kind, nhead, fread, _ = _KIND[dim]
    h = _header_c(path)
    if h["type"] != kind:
        raise R.RsgError(f"{path}: not a DATA{dim}D file")
    nt, ntrace = int(h["n"][0]), int(h["n"][1])
    coords = np.zeros((ntrace, nhead), np.float32)
    data = torch.empty((ntrace, nt), dtype=torch.float32, device=_dev())
    R.check(getattr(R.lib(), fread)(str(path).encode(), coords.ctypes.data, data.data_ptr(), ntrace, nt, _stream()),
            fread)
    return coords, data, h


def write_traces_device(path: str, coords, data, dt: float, ot: float = 0.0, dim: int = 2):
    """Data2D<float>::write / Data3D<float>::write through
    rsg_rss_write_data2d/3d from device samples (the 3D file records o1 = 0,
    as the reference's)."""
    C, torch, R = _rt()
    _, nhead, _, fwrite = _KIND[dim]
    coords = np.ascontiguousarray(coords, np.float32).reshape(-1, nhead)
    if not (isinstance(data, torch.Tensor) and data.is_cuda):
        data = torch.from_numpy(np.ascontiguousarray(data, np.float32)).to(_dev())
    data = data.to(torch.float32).contiguous()
    ntrace, nt = data.shape
    if coords.shape[0] != ntrace:
        raise R.RsgError("write: one coordinate row per trace")
    args = [str(path).encode(), coords.ctypes.data, data.data_ptr(), ntrace, nt, float(dt)]
    if dim == 2:
        args.append(float(ot))
    R.check(getattr(R.lib(), fwrite)(*args, _stream()), fwrite)


def read_regular_device(path: str):
    """A REGULAR float file (model, image) as a device tensor, dimensions
    slowest first (n9 .. n1) with the size-0 dimensions dropped."""
    C, torch, R = _rt()
    h = _header_c(path)
    shape = [int(n) for n in reversed(h["n"]) if n > 0]
    t = torch.empty(shape or [1], dtype=torch.float32, device=_dev())
    R.check(R.lib().rsg_rss_read_regular(str(path).encode(), t.data_ptr(), C.c_size_t(t.numel()), _stream()),
            "rsg_rss_read_regular")
    return t


def write_regular_device(path: str, data, n, d, o, type_=REGULAR):
    """A REGULAR float file with sizes n, steps d and origins o (up to nine
    each, n1 first) from a device tensor."""
    C, torch, R = _rt()
    h = R.RssHeader()
    h.data_format = h.header_format = 4
    h.type, h.nheader, h.ndims = int(type_), 0, 1
    for i, (a, b, c) in enumerate(zip(n, d, o)):
        h.n[i], h.d[i], h.o[i] = int(a), float(b), float(c)
    if not (isinstance(data, torch.Tensor) and data.is_cuda):
        data = torch.from_numpy(np.ascontiguousarray(data, np.float32)).to(_dev())
    data = data.to(torch.float32).contiguous()
    R.check(R.lib().rsg_rss_write_regular(str(path).encode(), C.byref(h), data.data_ptr(), _stream()),
            "rsg_rss_write_regular")


def image_desc(nx, nz, dx, dz, ox, oz, nhx=1, nhz=1):
    """The rsg_image2d description of an image (sizes, steps, origins)."""
    C, torch, R = _rt()
    return R.Image2DDesc(nx=int(nx), nz=int(nz), nhx=int(nhx), nhz=int(nhz), dx=float(dx), dz=float(dz),
                         ox=float(ox), oz=float(oz))


def stack_image_2d(gimg, g, limg, l, padlx=0, padhx=0, padlz=0, padhz=0):
    """Image2D::stackImage_parallel in device memory: limg added into gimg
    at the offset its origin gives (both [nhz][nhx][nz][nx] device tensors)."""
    C, torch, R = _rt()
    for t, desc in ((gimg, g), (limg, l)):
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise R.RsgError("stack_image_2d: contiguous float32 device tensors")
        if t.numel() != desc.nx * desc.nz * desc.nhx * desc.nhz:
            raise R.RsgError("stack_image_2d: tensor size differs from its description")
    R.check(R.lib().rsg_stack_image_2d(gimg.data_ptr(), C.byref(g), limg.data_ptr(), C.byref(l), int(padlx),
                                       int(padhx), int(padlz), int(padhz), _stream()), "rsg_stack_image_2d")

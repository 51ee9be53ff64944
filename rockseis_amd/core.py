"""Python mirror of RockSeis's modelling interface on the MI355X path.

Class and method names follow the reference (include/der.h, pml.h, model.h,
data.h, waves.h, modelling.h) so that a caller written against the
reference reads the same: Model*::staggerModels(), Data2D::makeMap(),
Waves*::forwardstepVelocity(model, der) / forwardstepStress / insertSource /
recordData, Modelling*::run().  Arrays are numpy on the host side
(models, wavelets, records, as the reference's host classes hold them) and
torch tensors on cuda:0 for the propagator state; all arithmetic runs in
librsgpu.so.  Domain decomposition (Domain::getStatus()) is not offered:
one GPU holds a whole shot window.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _lib as R

SMAP, GMAP = 0, 1          # utils.h:12-13
SMAX = 1200.0              # pml.h:16
FIELDS = {"P": R.PRESSURE, "PRESSURE": R.PRESSURE, "VX": R.VX, "VY": R.VY, "VZ": R.VZ}


def _dev():
    if not torch.cuda.is_available():
        raise R.RsgError("rockseis_amd needs a GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _np_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _field(f):
    if isinstance(f, str):
        return FIELDS[f.upper()]
    return int(f)


# --------------------------------------------------------------------------
# Der / Pml
# --------------------------------------------------------------------------

class Der:
    """Der(nx, ny, nz, dx, dy, dz, order): the stencil description; the
    derivative itself is evaluated inside the fused step kernels."""

    def __init__(self, nx, ny, nz, dx, dy, dz, order):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.order = min(max(int(order), 1), 8)   # der.cpp:94-95

    def getOrder(self):
        return self.order

    @staticmethod
    def coefficients():
        out = np.zeros(36, np.float32)
        R.check(R.lib().rsg_der_coeffs(_np_ptr(out)), "rsg_der_coeffs")
        return out


def _dev_window(t, nz, nx):
    """A float32 device tensor [nz, >= nx] with unit column stride (a window
    of a wider model may keep the wider row stride)."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise R.RsgError("staggerModelsDevice: model arrays must be CUDA tensors")
    t = t.detach()
    if t.dtype != torch.float32:
        t = t.float()
    if t.dim() != 2 or t.shape[0] != nz or t.shape[1] < nx:
        raise R.RsgError(f"staggerModelsDevice: expected [{nz}, >= {nx}], got {tuple(t.shape)}")
    if t.stride(1) != 1 or t.stride(0) < t.shape[1] and t.shape[0] > 1:
        t = t.contiguous()
    return t


class Pml:
    """CFS-PML constants (Pml<T>, pml.h:31-81); tables computed on the host
    exactly as computeABC() and kept on the device."""

    def __init__(self, lpml, dt):
        self.lpml, self.dt = int(lpml), float(dt)
        self.Smax, self.Amax, self.Kmax = SMAX, 0.0, 2.0
        self.tab_host = np.zeros(12 * self.lpml, np.float32)
        self.tab = None
        self.computeABC()

    def setSmax(self, smax):
        self.Smax = float(smax)

    def computeABC(self, sign=1):
        R.check(R.lib().rsg_pml_tables(self.lpml, self.dt, self.Smax, sign, _np_ptr(self.tab_host)),
                "rsg_pml_tables")
        if torch.cuda.is_available():
            self.tab = torch.from_numpy(self.tab_host).to(_dev())

    def table(self, name):
        names = ("A_ltf", "B_ltf", "C_ltf", "A_ltf_stag", "B_ltf_stag", "C_ltf_stag",
                 "A_rbb", "B_rbb", "C_rbb", "A_rbb_stag", "B_rbb_stag", "C_rbb_stag")
        j = names.index(name)
        return self.tab_host[j * self.lpml:(j + 1) * self.lpml]


# --------------------------------------------------------------------------
# Models
# --------------------------------------------------------------------------

class _Model:
    dim = 2

    def getLpml(self):
        return self.lpml

    def getFs(self):
        return self.fs

    def _device_arrays(self, names):
        """The padded device arrays _d<name>, allocated unless they exist
        with this model's shape (captured launches keep their pointers)."""
        shp = (self.getNz_pml(), self.getNx_pml())
        out = []
        for n in names:
            d = getattr(self, "_d" + n, None)
            if d is None or tuple(d.shape) != shp or d.dtype != torch.float32:
                d = torch.empty(shp, dtype=torch.float32, device=_dev())
                setattr(self, "_d" + n, d)
            out.append(d)
        return out

    def _upload(self, names):
        # in place when the device array exists with the same shape, so that
        # captured launches (FWI plans) keep pointing at the model
        dev = _dev()
        for n in names:
            h = torch.from_numpy(np.ascontiguousarray(getattr(self, n)))
            d = getattr(self, "_d" + n, None)
            if d is not None and d.shape == h.shape and d.dtype == h.dtype:
                d.copy_(h)
            else:
                setattr(self, "_d" + n, h.to(dev))


class ModelAcoustic2D(_Model):
    """ModelAcoustic2D(nx, nz, lpml, dx, dz, ox, oz, fs); Vp/R are [nz, nx]."""

    def __init__(self, nx, nz, lpml, dx, dz, ox=0.0, oz=0.0, fs=False):
        self.nx, self.nz, self.lpml = int(nx), int(nz), int(lpml)
        self.dx, self.dz, self.ox, self.oz, self.fs = float(dx), float(dz), float(ox), float(oz), bool(fs)
        self.Vp = np.zeros((nz, nx), np.float32)
        self.R = np.zeros((nz, nx), np.float32)

    def getNx_pml(self):
        return self.nx + 2 * self.lpml

    def getNz_pml(self):
        return self.nz + 2 * self.lpml

    def staggerModels(self):
        shp = (self.getNz_pml(), self.getNx_pml())
        self.L, self.Rx, self.Rz = (np.zeros(shp, np.float32) for _ in range(3))
        vp, rho = (np.ascontiguousarray(a, np.float32) for a in (self.Vp, self.R))
        R.check(R.lib().rsg_stagger_a2d(_np_ptr(vp), _np_ptr(rho), self.nx, self.nz, self.lpml,
                                        int(self.fs), _np_ptr(self.L), _np_ptr(self.Rx),
                                        _np_ptr(self.Rz)), "staggerModels")
        self._upload(("L", "Rx", "Rz"))

    def staggerModelsDevice(self, vp, rho):
        """staggerModels() from device tensors: vp, rho [nz, ld >= nx] float32
        (the first nx columns are this model's), staggered on the device into
        the model's device arrays (in place when they exist), bit-identical
        to the host set-up; the host arrays are left as they are."""
        vp, rho = _dev_window(vp, self.nz, self.nx), _dev_window(rho, self.nz, self.nx)
        d = self._device_arrays(("L", "Rx", "Rz"))
        R.check(R.lib().rsg_stagger_a2d_dev(vp.data_ptr(), rho.data_ptr(), vp.stride(0), self.nx, self.nz,
                                            self.lpml, int(self.fs), *(t.data_ptr() for t in d), _stream()),
                "staggerModels")


class ModelElastic2D(_Model):
    def __init__(self, nx, nz, lpml, dx, dz, ox=0.0, oz=0.0, fs=False):
        self.nx, self.nz, self.lpml = int(nx), int(nz), int(lpml)
        self.dx, self.dz, self.ox, self.oz, self.fs = float(dx), float(dz), float(ox), float(oz), bool(fs)
        self.Vp = np.zeros((nz, nx), np.float32)
        self.Vs = np.zeros((nz, nx), np.float32)
        self.R = np.zeros((nz, nx), np.float32)

    def getNx_pml(self):
        return self.nx + 2 * self.lpml

    def getNz_pml(self):
        return self.nz + 2 * self.lpml

    def staggerModels(self):
        shp = (self.getNz_pml(), self.getNx_pml())
        self.L, self.L2M, self.M, self.Rx, self.Rz = (np.zeros(shp, np.float32) for _ in range(5))
        vp, vs, rho = (np.ascontiguousarray(a, np.float32) for a in (self.Vp, self.Vs, self.R))
        R.check(R.lib().rsg_stagger_e2d(_np_ptr(vp), _np_ptr(vs), _np_ptr(rho), self.nx, self.nz,
                                        self.lpml, int(self.fs), _np_ptr(self.L), _np_ptr(self.L2M),
                                        _np_ptr(self.M), _np_ptr(self.Rx), _np_ptr(self.Rz)),
                "staggerModels")
        self._upload(("L", "L2M", "M", "Rx", "Rz"))

    def staggerModelsDevice(self, vp, vs, rho):
        """staggerModels() from device tensors (ModelAcoustic2D.staggerModelsDevice)."""
        a = [_dev_window(t, self.nz, self.nx) for t in (vp, vs, rho)]
        if len({t.stride(0) for t in a}) != 1:
            raise R.RsgError("staggerModelsDevice: vp, vs and rho must share their row stride")
        d = self._device_arrays(("L", "L2M", "M", "Rx", "Rz"))
        R.check(R.lib().rsg_stagger_e2d_dev(*(t.data_ptr() for t in a), a[0].stride(0), self.nx, self.nz,
                                            self.lpml, int(self.fs), *(t.data_ptr() for t in d), _stream()),
                "staggerModels")


class ModelAcoustic3D(_Model):
    dim = 3

    def __init__(self, nx, ny, nz, lpml, dx, dy, dz, ox=0.0, oy=0.0, oz=0.0, fs=False):
        self.nx, self.ny, self.nz, self.lpml = int(nx), int(ny), int(nz), int(lpml)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.ox, self.oy, self.oz, self.fs = float(ox), float(oy), float(oz), bool(fs)
        self.Vp = np.zeros((nz, ny, nx), np.float32)
        self.R = np.zeros((nz, ny, nx), np.float32)

    def getNx_pml(self):
        return self.nx + 2 * self.lpml

    def getNy_pml(self):
        return self.ny + 2 * self.lpml

    def getNz_pml(self):
        return self.nz + 2 * self.lpml

    def staggerModels(self):
        shp = (self.getNz_pml(), self.getNy_pml(), self.getNx_pml())
        self.L, self.Rx, self.Ry, self.Rz = (np.zeros(shp, np.float32) for _ in range(4))
        vp, rho = (np.ascontiguousarray(a, np.float32) for a in (self.Vp, self.R))
        R.check(R.lib().rsg_stagger_a3d(_np_ptr(vp), _np_ptr(rho), self.nx, self.ny, self.nz,
                                        self.lpml, int(self.fs), _np_ptr(self.L), _np_ptr(self.Rx),
                                        _np_ptr(self.Ry), _np_ptr(self.Rz)), "staggerModels")
        self._upload(("L", "Rx", "Ry", "Rz"))


class ModelElastic3D(ModelAcoustic3D):
    """ModelElastic3D (model.h:441-470): Vp, Vs, R [nz][ny][nx];
    staggerModels gives L, L2M, M_xz, M_yz, M_xy, Rx, Ry, Rz
    (model.cpp:3371-3473, bit-exact, rsg_stagger_e3d)."""

    def __init__(self, nx, ny, nz, lpml, dx, dy, dz, ox=0.0, oy=0.0, oz=0.0, fs=False):
        super().__init__(nx, ny, nz, lpml, dx, dy, dz, ox, oy, oz, fs)
        self.Vs = np.zeros((nz, ny, nx), np.float32)

    _E3D_MODEL = ("L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz")

    def staggerModels(self):
        shp = (self.getNz_pml(), self.getNy_pml(), self.getNx_pml())
        for n in self._E3D_MODEL:
            setattr(self, n, np.zeros(shp, np.float32))
        vp, vs, rho = (np.ascontiguousarray(a, np.float32) for a in (self.Vp, self.Vs, self.R))
        R.check(R.lib().rsg_stagger_e3d(_np_ptr(vp), _np_ptr(vs), _np_ptr(rho), self.nx, self.ny, self.nz,
                                        self.lpml, int(self.fs), *(_np_ptr(getattr(self, n))
                                                                   for n in self._E3D_MODEL)), "staggerModels")
        self._upload(self._E3D_MODEL + ("Vp", "Vs", "R"))


# --------------------------------------------------------------------------
# Data (traces + maps)
# --------------------------------------------------------------------------

class Data2D:
    """Data2D(ntrace, nt, dt, ot): trace-major data [ntrace, nt] with source
    and receiver coordinates (x, z) and the maps makeMap builds."""

    dim = 2

    def __init__(self, ntrace, nt, dt, ot=0.0, field="P"):
        self.ntrace, self.nt, self.dt, self.ot = int(ntrace), int(nt), float(dt), float(ot)
        self.data = np.zeros((self.ntrace, self.nt), np.float32)
        self.scoords = np.zeros((self.ntrace, self.dim), np.float32)
        self.gcoords = np.zeros((self.ntrace, self.dim), np.float32)
        self.field = _field(field)
        self.reciprocity = False
        self.smap = self.gmap = None

    def setField(self, f):
        self.field = _field(f)

    def getField(self):
        return self.field

    def setReciprocity(self, v):
        self.reciprocity = bool(v)

    def getData(self):
        return self.data

    def makeMap(self, model, maptype=SMAP):
        coords = np.ascontiguousarray(self.scoords if maptype == SMAP else self.gcoords, np.float32)
        m = np.full((self.ntrace, self.dim), -1, np.int32)
        s = np.zeros((self.ntrace, self.dim), np.float32)
        if self.dim == 2:
            R.check(R.lib().rsg_makemap2d(_np_ptr(coords), self.ntrace, model.nx, model.nz, model.dx,
                                          model.dz, model.ox, model.oz, self.field, _np_ptr(m),
                                          _np_ptr(s)), "makeMap")
        else:
            R.check(R.lib().rsg_makemap3d(_np_ptr(coords), self.ntrace, model.nx, model.ny, model.nz,
                                          model.dx, model.dy, model.dz, model.ox, model.oy, model.oz,
                                          self.field, _np_ptr(m), _np_ptr(s)), "makeMap")
        if maptype == SMAP:
            self.smap, self.sshift = m, s
        else:
            self.gmap, self.gshift = m, s

    def _device(self, maptype, factor=None, zero=False):
        """rsg_traces view on the device (map, shift, data, factor tensors kept alive)."""
        dev = _dev()
        m, s = (self.smap, self.sshift) if maptype == SMAP else (self.gmap, self.gshift)
        if m is None:
            raise R.RsgError("makeMap() must be called before using the traces")
        t = R.Traces()
        keep = [torch.from_numpy(m).to(dev), torch.from_numpy(s).to(dev)]
        d = torch.zeros_like(torch.from_numpy(self.data), device=dev) if zero else \
            torch.from_numpy(np.ascontiguousarray(self.data)).to(dev)
        keep.append(d)
        t.field, t.ntrace, t.nt = self.field, self.ntrace, self.nt
        t.map, t.shift, t.data = keep[0].data_ptr(), keep[1].data_ptr(), d.data_ptr()
        if factor is not None:
            f = torch.from_numpy(np.ascontiguousarray(factor, np.float32)).to(dev)
            keep.append(f)
            t.factor = f.data_ptr()
        return t, keep


    # File input / output (Data2D::read / write, data.cpp:463-634; the RSS
    # layout of file.cpp) through the device: rockseis_amd.rss.
    def setFile(self, path):
        self.file = str(path)

    def getFile(self):
        return getattr(self, "file", "")

    @classmethod
    def fromFile(cls, path, field="P"):
        """Data2D(datafile): sizes, sampling and traces from the file."""
        from . import rss
        coords, data, h = rss.read_traces_device(path, cls.dim)
        d = cls(data.shape[0], data.shape[1], h["d"][0], h["o"][0], field)
        d._take(coords, data)
        d.setFile(path)
        return d

    def _take(self, coords, data):
        k = self.dim
        self.scoords[:] = coords[:, :k]
        self.gcoords[:] = coords[:, k:]
        self.device_data = data
        self.data[:] = data.cpu().numpy()

    def read(self):
        """Traces and coordinates from the file set by setFile (sizes must
        match, as in the reference); the samples also stay on the device as
        device_data."""
        from . import rss
        coords, data, h = rss.read_traces_device(self.getFile(), self.dim)
        if data.shape != (self.ntrace, self.nt):
            raise R.RsgError(f"Data{self.dim}D::read: trace count or nt differs from the file's")
        self._take(coords, data)
        return True

    def write(self):
        from . import rss
        coords = np.concatenate([self.scoords, self.gcoords], axis=1)
        rss.write_traces_device(self.getFile(), coords, self.data, self.dt, self.ot, self.dim)
        return True


class Data3D(Data2D):
    dim = 3


# --------------------------------------------------------------------------
# Waves
# --------------------------------------------------------------------------

class _Waves:
    kind = 0
    strips = ()

    def _alloc(self, shape, fields, lpml, dt, strip_shapes):
        dev = _dev()
        self.shape = shape
        for f in fields:
            setattr(self, f, torch.zeros(shape, dtype=torch.float32, device=dev))
        self.Pml = Pml(lpml, dt)
        self.psi = {n: torch.zeros(strip_shapes[i], dtype=torch.float32, device=dev)
                    for i, n in enumerate(self.strips)}

    def getPml(self):
        return self.Pml

    def getLpml(self):
        return self.lpml

    def getDt(self):
        return self.dt

    def _state(self, model):
        s = R.State()
        for f in ("P", "Vx", "Vy", "Vz", "Sxx", "Szz", "Sxz"):
            t = getattr(self, f, None)
            if t is not None:
                setattr(s, f, t.data_ptr())
        for f in ("L", "L2M", "M", "Rx", "Ry", "Rz"):
            t = getattr(model, "_d" + f, None)
            if t is not None:
                setattr(s, f, t.data_ptr())
        if self.Pml.tab is None:
            self.Pml.computeABC()
        s.tab = self.Pml.tab.data_ptr()
        for i, n in enumerate(self.strips):
            s.psi[i] = self.psi[n].data_ptr()
        return s

    def _grid(self, der):
        g = R.Grid()
        g.nx, g.ny, g.nz = self.nx_pml, self.ny_pml, self.nz_pml
        g.lpml, g.order, g.fs = self.lpml, der.order, int(self.fs)
        g.dx, g.dy, g.dz, g.dt = self.dx, self.dy, self.dz, self.dt
        return g

    def _step(self, fname, model, der):
        g = self._grid(der)
        s = self._state(model)
        R.check(getattr(R.lib(), fname)(C.byref(g), C.byref(s), _stream()), fname)

    def insertSource(self, model, source, maptype, it):
        g = self._grid(Der(1, 1, 1, 1, 1, 1, 4))
        t, keep = source._device(maptype)
        R.check(R.lib().rsg_insert_source(self.kind, C.byref(g), C.byref(self._state(model)),
                                          C.byref(t), int(it), _stream()), "insertSource")
        torch.cuda.current_stream().synchronize()

    def recordData(self, *args):
        """recordData([model,] data, maptype, it): accumulates into data.data
        (host) through a device round trip of that sample."""
        if len(args) == 4:
            model, data, maptype, it = args
        else:
            model, (data, maptype, it) = None, args
        factor = _record_factor(model, data, maptype) if model is not None else None
        g = self._grid(Der(1, 1, 1, 1, 1, 1, 4))
        t, keep = data._device(maptype, factor)
        R.check(R.lib().rsg_record_data(self.kind, C.byref(g), C.byref(self._state(model or self._model)),
                                        C.byref(t), int(it), _stream()), "recordData")
        data.data[:, it] = keep[2][:, it].cpu().numpy()


def _record_factor(model, data, maptype):
    """Elastic pressure divisor rho*(Vp^2 - Vs^2) at the receiver cell under
    reciprocity, else 1 (waves.cpp:2219-2228)."""
    if not isinstance(model, ModelElastic2D) or data.field != R.PRESSURE:
        return None
    m = data.smap if maptype == SMAP else data.gmap
    f = np.ones(data.ntrace, np.float32)
    if data.reciprocity:
        ok = (m[:, 0] >= 0) & (m[:, 1] >= 0)
        x, z = m[ok, 0], m[ok, 1]
        r, vp, vs = model.R[z, x], model.Vp[z, x], model.Vs[z, x]
        f[ok] = r * vp * vp - r * vs * vs
    return f


class WavesAcoustic2D(_Waves):
    kind = R.KIND_A2D
    strips = R.A2D_STRIPS

    def __init__(self, model: ModelAcoustic2D, nt, dt, ot=0.0):
        self.nt, self.dt, self.ot = int(nt), float(dt), float(ot)
        self.lpml, self.fs = model.lpml, model.fs
        self.nx_pml, self.ny_pml, self.nz_pml = model.getNx_pml(), 1, model.getNz_pml()
        self.dx, self.dy, self.dz = model.dx, 1.0, model.dz
        self._model = model
        nxp, nzp, L = self.nx_pml, self.nz_pml, self.lpml
        self._alloc((nzp, nxp), ("P", "Vx", "Vz"), L, dt,
                    [(nzp, L)] * 2 + [(L, nxp)] * 2 + [(nzp, L)] * 2 + [(L, nxp)] * 2)

    def forwardstepVelocity(self, model, der):
        self._step("rsg_a2d_forward_velocity", model, der)

    def forwardstepStress(self, model, der):
        self._step("rsg_a2d_forward_stress", model, der)

    def getP(self):
        return self.P

    def getVx(self):
        return self.Vx

    def getVz(self):
        return self.Vz

    def getNx_pml(self):
        return self.nx_pml

    def getNz_pml(self):
        return self.nz_pml


class WavesElastic2D(_Waves):
    kind = R.KIND_E2D
    strips = R.E2D_STRIPS

    def __init__(self, model: ModelElastic2D, nt, dt, ot=0.0):
        self.nt, self.dt, self.ot = int(nt), float(dt), float(ot)
        self.lpml, self.fs = model.lpml, model.fs
        self.nx_pml, self.ny_pml, self.nz_pml = model.getNx_pml(), 1, model.getNz_pml()
        self.dx, self.dy, self.dz = model.dx, 1.0, model.dz
        self._model = model
        nxp, nzp, L = self.nx_pml, self.nz_pml, self.lpml
        x, z = (nzp, L), (L, nxp)
        self._alloc((nzp, nxp), ("Sxx", "Szz", "Sxz", "Vx", "Vz"), L, dt,
                    [x] * 4 + [z] * 4 + [x] * 4 + [z] * 4)

    def forwardstepVelocity(self, model, der):
        self._step("rsg_e2d_forward_velocity", model, der)

    def forwardstepStress(self, model, der):
        self._step("rsg_e2d_forward_stress", model, der)

    def backwardstepVelocity(self, model, der):
        self._step("rsg_e2d_backward_velocity", model, der)

    def backwardstepStress(self, model, der):
        self._step("rsg_e2d_backward_stress", model, der)

    def getSxx(self):
        return self.Sxx

    def getSzz(self):
        return self.Szz

    def getSxz(self):
        return self.Sxz

    def getVx(self):
        return self.Vx

    def getVz(self):
        return self.Vz

    def getNx_pml(self):
        return self.nx_pml

    def getNz_pml(self):
        return self.nz_pml


class WavesAcoustic3D(_Waves):
    kind = R.KIND_A3D
    strips = R.A3D_STRIPS

    def __init__(self, model: ModelAcoustic3D, nt, dt, ot=0.0):
        self.nt, self.dt, self.ot = int(nt), float(dt), float(ot)
        self.lpml, self.fs = model.lpml, model.fs
        self.nx_pml, self.ny_pml, self.nz_pml = model.getNx_pml(), model.getNy_pml(), model.getNz_pml()
        self.dx, self.dy, self.dz = model.dx, model.dy, model.dz
        self._model = model
        nxp, nyp, nzp, L = self.nx_pml, self.ny_pml, self.nz_pml, self.lpml
        lr, fb, tb = (nzp, nyp, L), (nzp, L, nxp), (L, nyp, nxp)
        self._alloc((nzp, nyp, nxp), ("P", "Vx", "Vy", "Vz"), L, dt,
                    [lr, lr, fb, fb, tb, tb, lr, lr, fb, fb, tb, tb])

    def forwardstepVelocity(self, model, der):
        self._step("rsg_a3d_forward_velocity", model, der)

    def forwardstepStress(self, model, der):
        self._step("rsg_a3d_forward_stress", model, der)

    def getP(self):
        return self.P

    def getVx(self):
        return self.Vx

    def getVy(self):
        return self.Vy

    def getVz(self):
        return self.Vz

    def getNx_pml(self):
        return self.nx_pml

    def getNy_pml(self):
        return self.ny_pml

    def getNz_pml(self):
        return self.nz_pml


class WavesElastic3D:
    """WavesElastic3D (waves.h:215-250): the nine device fields Vx, Vy, Vz,
    Sxx, Syy, Szz, Syz, Sxz, Sxy and the 36 PML strips of PmlElastic3D on
    the padded grid, stepped by rsg_e3d_forward_velocity / _stress."""

    FIELDS = ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")

    def __init__(self, model: ModelElastic3D, nt, dt, ot=0.0):
        self.nt, self.dt, self.ot = int(nt), float(dt), float(ot)
        self.lpml, self.fs = model.lpml, model.fs
        self.nx_pml, self.ny_pml, self.nz_pml = model.getNx_pml(), model.getNy_pml(), model.getNz_pml()
        self.dx, self.dy, self.dz = model.dx, model.dy, model.dz
        self._model = model
        dev = _dev()
        nxp, nyp, nzp, L = self.nx_pml, self.ny_pml, self.nz_pml, self.lpml
        for f in self.FIELDS:
            setattr(self, f, torch.zeros((nzp, nyp, nxp), dtype=torch.float32, device=dev))
        face = {0: (nzp, nyp, L), 1: (nzp, L, nxp), 2: (L, nyp, nxp)}
        self.psi = [torch.zeros(face[i // 12], dtype=torch.float32, device=dev) for i in range(R.E3D_NSTRIP)]
        self.Pml = Pml(L, dt)

    def getPml(self):
        return self.Pml

    def getNx_pml(self):
        return self.nx_pml

    def getNy_pml(self):
        return self.ny_pml

    def getNz_pml(self):
        return self.nz_pml

    def _grid(self, order):
        g = R.Grid()
        g.nx, g.ny, g.nz = self.nx_pml, self.ny_pml, self.nz_pml
        g.lpml, g.order, g.fs = self.lpml, int(order), int(self.fs)
        g.dx, g.dy, g.dz, g.dt = self.dx, self.dy, self.dz, self.dt
        return g

    def _state(self, model):
        s = R.StateE3D()
        for f in self.FIELDS:
            setattr(s, f, getattr(self, f).data_ptr())
        for n in ModelElastic3D._E3D_MODEL:
            setattr(s, n, getattr(model, "_d" + n).data_ptr())
        if self.Pml.tab is None:
            self.Pml.computeABC()
        s.tab = self.Pml.tab.data_ptr()
        for i, t in enumerate(self.psi):
            s.psi[i] = t.data_ptr()
        return s

    def forwardstepVelocity(self, model, der):
        R.check(R.lib().rsg_e3d_forward_velocity(C.byref(self._grid(der.order)), C.byref(self._state(model)),
                                                 _stream()), "rsg_e3d_forward_velocity")

    def forwardstepStress(self, model, der):
        R.check(R.lib().rsg_e3d_forward_stress(C.byref(self._grid(der.order)), C.byref(self._state(model)),
                                               _stream()), "rsg_e3d_forward_stress")

    def backwardstepVelocity(self, model, der):
        """The adjoint-state velocity step of FwiElastic3D (waves.cpp:3260-3713),
        with a device workspace of six padded grids kept by this object (the
        reference's wrk array)."""
        ws = getattr(self, "_wrk", None)
        if ws is None:
            ws = self._wrk = torch.empty((6, self.nz_pml, self.ny_pml, self.nx_pml), dtype=torch.float32,
                                         device=self.Vx.device)
        R.check(R.lib().rsg_e3d_backward_velocity_ws(C.byref(self._grid(der.order)), C.byref(self._state(model)),
                                                     C.c_void_p(ws.data_ptr()), _stream()),
                "rsg_e3d_backward_velocity_ws")

    def backwardstepStress(self, model, der):
        """The adjoint-state stress step of FwiElastic3D (waves.cpp:3715-4103)."""
        R.check(R.lib().rsg_e3d_backward_stress(C.byref(self._grid(der.order)), C.byref(self._state(model)),
                                                _stream()), "rsg_e3d_backward_stress")

    def getVx(self):
        return self.Vx

    def getVy(self):
        return self.Vy

    def getVz(self):
        return self.Vz

    def getSxx(self):
        return self.Sxx

    def getSyy(self):
        return self.Syy

    def getSzz(self):
        return self.Szz

    def getSyz(self):
        return self.Syz

    def getSxz(self):
        return self.Sxz

    def getSxy(self):
        return self.Sxy


class ModellingElastic3D:
    """ModellingElastic3D::run (modelling.cpp:727-): the elastic 3D time loop,
    velocity step, stress step, insertSource(it), recordData(it) of every
    record set, on the device; records are copied to the host Data3D
    objects at the end."""

    def __init__(self, model: ModelElastic3D, source: Data3D, order=4, snapinc=1):
        self.model, self.source = model, source
        self.order = int(order) if 1 < int(order) < 9 else 4
        self.snapinc = max(int(snapinc), 1)
        self.recs = {}
        self.waves = None

    def getOrder(self):
        return self.order

    def setRecP(self, d):
        self.recs["P"] = d

    def setRecVx(self, d):
        self.recs["VX"] = d

    def setRecVy(self, d):
        self.recs["VY"] = d

    def setRecVz(self, d):
        self.recs["VZ"] = d

    def checkStability(self):
        # modelling.cpp:710-724
        m = self.model
        dx, dy, dz = np.float32(m.dx), np.float32(m.dy), np.float32(m.dz)
        dt = np.float32(self.source.dt)
        vpmax = np.float32(np.max(m.Vp))
        dt_stab = 2.0 / (3.1415 * np.sqrt((1.0 / (dx * dx)) + (1 / (dy * dy)) + (1 / (dz * dz))) * vpmax)
        return bool(dt < np.float32(dt_stab))

    def run(self):
        if not self.checkStability():
            raise R.RsgError("Wavelet sampling interval (dt) does not match the stability criteria.")
        m, src = self.model, self.source
        if not hasattr(m, "_dL2M"):
            m.staggerModels()
        w = WavesElastic3D(m, src.nt, src.dt, src.ot)
        self.waves = w
        w.getPml().setSmax(SMAX)
        w.getPml().computeABC()
        g = w._grid(self.order)
        s = w._state(m)
        t_src, keep = src._device(SMAP)
        recs = []
        for f, d in self.recs.items():
            if d.field != _field(f):
                d.setField(f)
            t, k = d._device(GMAP, zero=True)
            recs.append((d, t, k))
        L = R.lib()
        st = _stream()
        vp, vs, rho = (getattr(m, "_d" + n).data_ptr() for n in ("Vp", "Vs", "R"))
        for it in range(src.nt):
            R.check(L.rsg_e3d_forward_velocity(C.byref(g), C.byref(s), st), "rsg_e3d_forward_velocity")
            R.check(L.rsg_e3d_forward_stress(C.byref(g), C.byref(s), st), "rsg_e3d_forward_stress")
            R.check(L.rsg_e3d_insert_source(C.byref(g), C.byref(s), C.byref(t_src), it, st), "insertSource")
            for d, t, k in recs:
                R.check(L.rsg_e3d_record_data(C.byref(g), C.byref(s), C.byref(t), vp, vs, rho, it, st),
                        "recordData")
        for d, t, k in recs:
            d.data[...] = k[2].cpu().numpy()
        return 1


LINEAR, BSPLINE, SINC = 0, 1, 2  # rs_interpmode (utils.h:25)


class Interp:
    """Interp(mode).interp(from, to): resample the traces of a Data2D/Data3D
    onto the time axis of another (interp.cpp:110-186) with rsg_interp on
    the device; coordinates are copied as the reference does."""

    def __init__(self, mode=LINEAR):
        self.mode = int(mode)

    def interp(self, frm, to):
        if frm.ntrace != to.ntrace:
            raise R.RsgError("Interp::interp: Number of traces in the two datasets mismatch.")
        to.scoords[:] = frm.scoords
        to.gcoords[:] = frm.gcoords
        out = self.interp_device(torch.from_numpy(np.ascontiguousarray(frm.data)).to(_dev()), frm.dt, frm.ot,
                                 to.nt, to.dt, to.ot)
        to.data[...] = out.cpu().numpy()

    def interp_device(self, data, d1, o1, n2, d2, o2):
        """Resample device traces data[ntr, n1] (float32); returns [ntr, n2]."""
        data = data.contiguous()
        ntr, n1 = data.shape
        out = torch.empty((ntr, int(n2)), dtype=torch.float32, device=data.device)
        R.check(R.lib().rsg_interp(self.mode, data.data_ptr(), ntr, n1, float(d1), float(o1), out.data_ptr(),
                                   int(n2), float(d2), float(o2), _stream()), "Interp::interp")
        return out


class RaggedBatch2D:
    """Shots on 2D grids of different sizes (aperture windows) advanced by
    one launch per half step (rsg_ragged2d_*).  `waves` and `models` are the
    per-shot WavesAcoustic2D / WavesElastic2D objects and their models, all
    of one kind and one stencil order; step(which) with which = 0 forward
    velocity, 1 forward stress, 2 / 3 the elastic adjoint-state steps."""

    def __init__(self, waves, models, order):
        kinds = {w.kind for w in waves}
        if len(kinds) != 1:
            raise R.RsgError("RaggedBatch2D: every shot must be of one propagator kind")
        n = len(waves)
        der = Der(1, 1, 1, 1, 1, 1, order)
        self._grids = (R.Grid * n)(*[w._grid(der) for w in waves])
        self._states = (R.State * n)(*[w._state(m) for w, m in zip(waves, models)])
        self._keep = (list(waves), list(models))
        self._plan = C.c_void_p()
        R.check(R.lib().rsg_ragged2d_create(kinds.pop(), n, self._grids, self._states, C.byref(self._plan)),
                "rsg_ragged2d_create")

    def step(self, which):
        R.check(R.lib().rsg_ragged2d_step(self._plan, int(which), _stream()), "rsg_ragged2d_step")

    def close(self):
        if self._plan:
            R.lib().rsg_ragged2d_destroy(self._plan)
            self._plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# Modelling drivers
# --------------------------------------------------------------------------

class _Modelling:
    waves_cls = None

    def __init__(self, model, source, order=4, snapinc=1):
        self.model, self.source = model, source
        self.order = int(order) if 1 < int(order) < 9 else 4      # modelling.cpp:17-23
        self.snapinc = max(int(snapinc), 1)
        self.recs = {}
        self.snaps = {}
        self.snapfiles = {}
        self.graph_chunk = 0
        self.waves = None
        self._plan = None

    def getOrder(self):
        return self.order

    def getSnapinc(self):
        return self.snapinc

    # Snapshots (Modelling*::setSnapP/Vx/Vz/Sxx/Szz/Sxz, modelling.h:79-81,
    # 155-160): every snapinc-th step the interior of the field (P of the
    # elastic propagator is Sxx + Szz, as modelling.cpp snaps it) is captured
    # on the device inside the time loop; after run() it is in
    # self.snaps[name] as a [snapnt][nz][nx] (3D [snapnt][nz][ny][nx]) numpy
    # array, and written as an RSS SNAPSHOT file when a file name is given.
    _SNAP_FIELDS = {"P": ("P",), "VX": ("Vx",), "VY": ("Vy",), "VZ": ("Vz",)}

    def _set_snap(self, name, filename):
        self.snapfiles[name] = filename
        self.snaps[name] = None

    def setSnapP(self, filename=None):
        self._set_snap("P", filename)

    def setSnapVx(self, filename=None):
        self._set_snap("VX", filename)

    def setSnapVy(self, filename=None):
        self._set_snap("VY", filename)

    def setSnapVz(self, filename=None):
        self._set_snap("VZ", filename)

    def setRecP(self, d):
        self.recs["P"] = d

    def setRecVx(self, d):
        self.recs["VX"] = d

    def setRecVy(self, d):
        self.recs["VY"] = d

    def setRecVz(self, d):
        self.recs["VZ"] = d

    def setGraphChunk(self, n):
        """Run the time loop from hipGraphs of up to n steps (0 = plain
        per-step launches).  The graphs (1, 2, 4, ... n steps) are captured
        once, at the first launch after prepare(), with the launch shapes in
        force then, and replayed for every later launch."""
        self.graph_chunk = int(n)
        self._drop_plan()

    def _drop_plan(self):
        plan = getattr(self, "_plan", None)
        self._plan = None
        if plan:
            R.check(R.lib().rsg_modelling_plan_destroy(plan), "rsg_modelling_plan_destroy")

    def __del__(self):
        try:
            self._drop_plan()
        except Exception:
            pass

    def _vpmax(self):
        return float(np.max(self.model.Vp))

    def checkStability(self):
        raise NotImplementedError

    def prepare(self):
        """Allocate the propagator and upload traces; returns the launch
        arguments of rsg_modelling_run (kept so run() can be timed alone)."""
        if not self.checkStability():
            raise R.RsgError("Wavelet sampling interval (dt) does not match the stability criteria.")
        self._drop_plan()
        m, src = self.model, self.source
        self.waves = self.waves_cls(m, src.nt, src.dt, src.ot)
        w = self.waves
        self.der = Der(w.nx_pml, w.ny_pml, w.nz_pml, w.dx, w.dy, w.dz, self.order)
        w.getPml().setSmax(SMAX)
        w.getPml().computeABC()
        g = w._grid(self.der)
        s = w._state(m)
        t_src, keep = src._device(SMAP)
        recs = (R.Traces * max(len(self.recs), 1))()
        self._keep = [keep]
        self._rec_dev = {}
        for i, (k, d) in enumerate(self.recs.items()):
            t, kk = d._device(GMAP, _record_factor(m, d, GMAP), zero=True)
            recs[i] = t
            self._keep.append(kk)
            self._rec_dev[k] = kk[2]
        self._args = (g, s, t_src, recs, len(self.recs))
        self._snapsets = self._prepare_snaps(w)
        return self._args

    def _prepare_snaps(self, w):
        if not self.snaps:
            return None
        pad = 0 if getattr(w, "domdec", False) else w.lpml
        shape = (w.nz_pml - 2 * pad,) + ((w.ny_pml - 2 * pad,) if w.ny_pml > 1 else ()) + (w.nx_pml - 2 * pad,)
        snapnt = (self.source.nt - 1) // self.snapinc + 1
        sets = (R.SnapSet * len(self.snaps))()
        self._snap_dev = {}
        for i, name in enumerate(self.snaps):
            fields = [getattr(w, f) for f in self._SNAP_FIELDS[name]]
            out = torch.zeros((snapnt,) + shape, dtype=torch.float32, device=fields[0].device)
            for k, f in enumerate(fields):
                sets[i].f[k] = f.data_ptr()
            sets[i].out = out.data_ptr()
            sets[i].snapinc = self.snapinc
            self._snap_dev[name] = out
        return sets

    def capture(self):
        """Capture the time loop's hipGraphs (setGraphChunk) now rather than
        at the first launch; a no-op without graphs or when captured."""
        if self.graph_chunk <= 0 or self._plan is not None:
            return
        g, s, t_src, recs, nrec = self._args
        plan = C.c_void_p()
        snaps = self._snapsets
        R.check(R.lib().rsg_modelling_plan_create(
            self.waves.kind, C.byref(g), C.byref(s), C.byref(t_src), nrec, recs,
            0 if snaps is None else len(snaps), snaps, self.graph_chunk, _stream(), C.byref(plan)),
            "rsg_modelling_plan_create")
        self._plan = plan

    def launch(self, it0=0, nsteps=None):
        g, s, t_src, recs, nrec = self._args
        nsteps = self.source.nt - it0 if nsteps is None else nsteps
        if self.graph_chunk > 0:
            self.capture()
            R.check(R.lib().rsg_modelling_plan_run(self._plan, it0, nsteps, _stream()), "rsg_modelling_plan_run")
            return
        if self._snapsets is not None:
            R.check(R.lib().rsg_modelling_run_snap(self.waves.kind, C.byref(g), C.byref(s), C.byref(t_src), nrec,
                                                   recs, len(self._snapsets), self._snapsets, it0, nsteps,
                                                   self.graph_chunk, _stream()),
                    "rsg_modelling_run_snap")
            return
        R.check(R.lib().rsg_modelling_run(self.waves.kind, C.byref(g), C.byref(s), C.byref(t_src), nrec,
                                          recs, it0, nsteps, self.graph_chunk, _stream()),
                "rsg_modelling_run")

    def finish(self):
        torch.cuda.current_stream().synchronize()
        for k, d in self.recs.items():
            d.data[...] = self._rec_dev[k].cpu().numpy()
        if self._snapsets is not None:
            from . import rss
            m, w = self.model, self.waves
            for name, out in self._snap_dev.items():
                a = out.cpu().numpy()
                self.snaps[name] = a
                fn = self.snapfiles.get(name)
                if fn:
                    nx, nz = a.shape[-1], a.shape[1]
                    ny = a.shape[2] if a.ndim == 4 else 1
                    h = rss.Header(4, 0, rss.SNAPSHOT, 0, 1, (nx, ny, nz, a.shape[0]),
                                   (m.dx, getattr(m, "dy", 1.0), m.dz, w.dt * self.snapinc),
                                   (m.ox, getattr(m, "oy", 0.0), m.oz, w.ot))
                    rss.write(fn, h, a)

    def run(self):
        self.prepare()
        self.launch()
        self.finish()
        return 1


class ModellingAcoustic2D(_Modelling):
    waves_cls = WavesAcoustic2D

    def checkStability(self):
        # modelling.cpp:181-194
        dx, dz, dt = np.float32(self.model.dx), np.float32(self.model.dz), np.float32(self.source.dt)
        vpmax = np.float32(self._vpmax())
        dt_stab = 2.0 / (3.1415 * np.sqrt((1.0 / (dx * dx)) + (1 / (dz * dz))) * vpmax)
        return bool(dt < np.float32(dt_stab))


class ModellingElastic2D(ModellingAcoustic2D):
    waves_cls = WavesElastic2D
    _SNAP_FIELDS = {"P": ("Sxx", "Szz"), "VX": ("Vx",), "VZ": ("Vz",), "SXX": ("Sxx",), "SZZ": ("Szz",),
                    "SXZ": ("Sxz",)}

    def setSnapSxx(self, filename=None):
        self._set_snap("SXX", filename)

    def setSnapSzz(self, filename=None):
        self._set_snap("SZZ", filename)

    def setSnapSxz(self, filename=None):
        self._set_snap("SXZ", filename)


class ModellingAcoustic3D(_Modelling):
    waves_cls = WavesAcoustic3D

    def checkStability(self):
        # modelling.cpp:346-360
        m = self.model
        dx, dy, dz = np.float32(m.dx), np.float32(m.dy), np.float32(m.dz)
        dt = np.float32(self.source.dt)
        vpmax = np.float32(self._vpmax())
        dt_stab = 2.0 / (3.1415 * np.sqrt((1.0 / (dx * dx)) + (1 / (dy * dy)) + (1 / (dz * dz))) * vpmax)
        return bool(dt < np.float32(dt_stab))

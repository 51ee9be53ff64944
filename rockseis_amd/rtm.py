"""RtmAcoustic2D / RtmAcoustic3D (rtm.h, rtm.cpp) on the device:
reverse-time migration of one shot with optimal (Revolve) checkpointing and
the imaging condition with subsurface offsets, through rsg_rtm_a2d_image and
rsg_rtm_a3d_image."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as R
from .core import GMAP, SMAP, SMAX, Data2D, Data3D, ModelAcoustic2D, ModelAcoustic3D, Pml, _dev, _stream
from .fwi import Image2D, Image3D


class RtmAcoustic2D:
    """Mirror of RtmAcoustic2D<float>: model, pressure image (Image2D with
    nhx x nhz subsurface offsets), PRESSURE source and recorded PRESSURE
    data; run_optimal() fills pimage.imagedata."""

    def __init__(self, model: ModelAcoustic2D, pimage: Image2D, source: Data2D, dataP: Data2D, order=4,
                 snapinc=1):
        self.model, self.pimage, self.source, self.dataP = model, pimage, source, dataP
        # Rtm::Rtm (rtm.cpp:19-39): orders outside 2..8 fall back to 4
        self.order = int(order) if 1 < int(order) < 9 else 4
        self.snapinc = int(snapinc) if int(snapinc) > 0 else 1
        self.ncheck = 0
        self.incore = True

    def setNcheck(self, n):
        self.ncheck = int(n)

    def setIncore(self, v):
        self.incore = bool(v)

    def setOrder(self, order):
        self.order = int(order) if 1 < int(order) < 9 else 4

    def getOrder(self):
        return self.order

    def run_optimal(self):
        dev = _dev()
        m, src, dat, img = self.model, self.source, self.dataP, self.pimage
        if dat.field != R.PRESSURE or src.field != R.PRESSURE:
            raise R.RsgError("RtmAcoustic2D on this path takes PRESSURE sources and data")
        if src.ntrace != 1:
            raise R.RsgError("RtmAcoustic2D: one source trace per shot")
        if img.nx != m.nx or img.nz != m.nz:
            raise R.RsgError("RtmAcoustic2D: image and model sizes differ")
        if self.ncheck < 1:
            raise R.RsgError("RtmAcoustic2D::run_optimal: set the number of checkpoints (setNcheck)")
        if not hasattr(m, "_dL"):
            m.staggerModels()
        pml = Pml(m.lpml, src.dt)
        pml.setSmax(SMAX)
        pml.computeABC()
        g = R.Grid(nx=m.getNx_pml(), ny=1, nz=m.getNz_pml(), lpml=m.lpml, order=self.order,
                   fs=int(m.fs), dx=m.dx, dy=1.0, dz=m.dz, dt=src.dt)
        t_src, k1 = src._device(SMAP)
        t_dat, k2 = dat._device(GMAP)
        image = torch.zeros(img.imagedata.size, dtype=torch.float32, device=dev)
        r = R.RtmA2D()
        r.L, r.Rx, r.Rz, r.tab = m._dL.data_ptr(), m._dRx.data_ptr(), m._dRz.data_ptr(), pml.tab.data_ptr()
        r.nx, r.nz = m.nx, m.nz
        r.src, r.data = t_src, t_dat
        r.image = image.data_ptr()
        r.nhx, r.nhz, r.ncheck = img.nhx, img.nhz, self.ncheck
        R.check(R.lib().rsg_rtm_a2d_image(C.byref(g), C.byref(r), _stream()), "rsg_rtm_a2d_image")
        img.imagedata[...] = image.cpu().numpy().reshape(img.imagedata.shape)
        return 1


class RtmAcoustic3D(RtmAcoustic2D):
    """Mirror of RtmAcoustic3D<float>: model, pressure image (Image3D with
    nhx x nhy x nhz subsurface offsets), PRESSURE source and recorded
    PRESSURE data; run_optimal() fills pimage.imagedata."""

    def __init__(self, model: ModelAcoustic3D, pimage: Image3D, source: Data3D, dataP: Data3D, order=4,
                 snapinc=1):
        super().__init__(model, pimage, source, dataP, order, snapinc)

    def run_optimal(self):
        dev = _dev()
        m, src, dat, img = self.model, self.source, self.dataP, self.pimage
        if dat.field != R.PRESSURE or src.field != R.PRESSURE:
            raise R.RsgError("RtmAcoustic3D on this path takes PRESSURE sources and data")
        if src.ntrace != 1:
            raise R.RsgError("RtmAcoustic3D: one source trace per shot")
        if (img.nx, img.ny, img.nz) != (m.nx, m.ny, m.nz):
            raise R.RsgError("RtmAcoustic3D: image and model sizes differ")
        if self.ncheck < 1:
            raise R.RsgError("RtmAcoustic3D::run_optimal: set the number of checkpoints (setNcheck)")
        if not hasattr(m, "_dL"):
            m.staggerModels()
        pml = Pml(m.lpml, src.dt)
        pml.setSmax(SMAX)
        pml.computeABC()
        g = R.Grid(nx=m.getNx_pml(), ny=m.getNy_pml(), nz=m.getNz_pml(), lpml=m.lpml, order=self.order,
                   fs=int(m.fs), dx=m.dx, dy=m.dy, dz=m.dz, dt=src.dt)
        t_src, k1 = src._device(SMAP)
        t_dat, k2 = dat._device(GMAP)
        image = torch.zeros(img.imagedata.size, dtype=torch.float32, device=dev)
        r = R.RtmA3D()
        r.L, r.Rx, r.Ry, r.Rz = m._dL.data_ptr(), m._dRx.data_ptr(), m._dRy.data_ptr(), m._dRz.data_ptr()
        r.tab = pml.tab.data_ptr()
        r.nx, r.ny, r.nz = m.nx, m.ny, m.nz
        r.src, r.data = t_src, t_dat
        r.image = image.data_ptr()
        r.nhx, r.nhy, r.nhz, r.ncheck = img.nhx, img.nhy, img.nhz, self.ncheck
        R.check(R.lib().rsg_rtm_a3d_image(C.byref(g), C.byref(r), _stream()), "rsg_rtm_a3d_image")
        img.imagedata[...] = image.cpu().numpy().reshape(img.imagedata.shape)
        return 1

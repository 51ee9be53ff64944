"""RtmAcoustic2D / RtmAcoustic3D (rtm.h, rtm.cpp) on the device:
reverse-time migration of one shot with optimal (Revolve) checkpointing and
the imaging condition with subsurface offsets, through rsg_rtm_a2d_image and
rsg_rtm_a3d_image."""
from __future__ import annotations

import ctypes as C

import numpy as np
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


def batch_image(vp, rho, survey, nhx=1, nhz=1):
    """Stacked RTM image of every shot of a survey (fwi.Survey; each shot's
    dobs is its recorded data) as one batch through rsg_rtm_a2d_batch: every
    kernel of a time step runs all shots in one launch, on the survey's
    Revolve schedule (ncheck checkpoints).  Shots must share nt and their
    trace counts.  Returns the image [nhz, nhx, nz, nx] (a device tensor)."""
    dev = _dev()
    nz, nx = vp.shape
    shots = survey.shots
    ns = len(shots)
    nsrc, nrec, nt = shots[0].scoords.shape[0], shots[0].gcoords.shape[0], shots[0].wavelet.shape[1]
    for sh in shots:
        if sh.scoords.shape[0] != nsrc or sh.gcoords.shape[0] != nrec or sh.wavelet.shape[1] != nt:
            raise R.RsgError("batch_image: shots must share nt and trace counts")
    if survey.ncheck < 1:
        raise R.RsgError("batch_image: the survey needs ncheck >= 1 Revolve checkpoints")
    m = ModelAcoustic2D(nx, nz, survey.lpml, survey.dx, survey.dz, 0.0, 0.0, survey.fs)
    m.Vp[:] = vp
    m.R[:] = rho
    smaps, sshifts, gmaps, gshifts = [], [], [], []
    for sh in shots:
        d = Data2D(nsrc, nt, survey.dt, 0.0, "P")
        d.scoords[:] = sh.scoords
        d.makeMap(m, SMAP)
        smaps.append(d.smap)
        sshifts.append(d.sshift)
        d = Data2D(nrec, nt, survey.dt, 0.0, "P")
        d.gcoords[:] = sh.gcoords
        d.makeMap(m, GMAP)
        gmaps.append(d.gmap)
        gshifts.append(d.gshift)
    m.staggerModels()
    pml = Pml(m.lpml, survey.dt)
    pml.setSmax(SMAX)
    pml.computeABC()

    def dev_t(a, dtype=torch.float32):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)

    keep = dict(smap=dev_t(np.stack(smaps), torch.int32), sshift=dev_t(np.stack(sshifts)),
                gmap=dev_t(np.stack(gmaps), torch.int32), gshift=dev_t(np.stack(gshifts)),
                wav=dev_t(np.stack([sh.wavelet for sh in shots])), data=dev_t(np.stack([sh.dobs for sh in shots])))
    image = torch.zeros((int(nhz), int(nhx), nz, nx), dtype=torch.float32, device=dev)
    g = R.Grid(nx=m.getNx_pml(), ny=1, nz=m.getNz_pml(), lpml=m.lpml, order=survey.order,
               fs=int(m.fs), dx=m.dx, dy=1.0, dz=m.dz, dt=survey.dt)
    r = R.RtmA2D()
    r.L, r.Rx, r.Rz, r.tab = m._dL.data_ptr(), m._dRx.data_ptr(), m._dRz.data_ptr(), pml.tab.data_ptr()
    r.nx, r.nz = nx, nz
    r.src = R.Traces(field=R.PRESSURE, ntrace=nsrc, nt=nt, map=keep["smap"].data_ptr(),
                     shift=keep["sshift"].data_ptr(), data=keep["wav"].data_ptr(), factor=None)
    r.data = R.Traces(field=R.PRESSURE, ntrace=nrec, nt=nt, map=keep["gmap"].data_ptr(),
                      shift=keep["gshift"].data_ptr(), data=keep["data"].data_ptr(), factor=None)
    r.image = image.data_ptr()
    r.nhx, r.nhz, r.ncheck = int(nhx), int(nhz), survey.ncheck
    R.check(R.lib().rsg_rtm_a2d_batch(C.byref(g), C.byref(r), ns, _stream()), "rsg_rtm_a2d_batch")
    return image

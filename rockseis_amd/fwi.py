"""Acoustic 2D FWI gradient on the MI355X path, mirroring FwiAcoustic2D
(include/fwi.h:61-140, lib/fwi.cpp:449-1337).

FwiAcoustic2D(model, source, dataP, order, snapinc) with setDatamodP /
setDataresP / setVpgrad / setRhograd / setWavgrad / setNcheck / setIncore,
then run_optimal() (Revolve checkpoints kept in HBM), run() (the reference's
full-snapshot gradient: the pressure every snapinc-th step kept in HBM in
place of the snapshot file) or run_stored() (run_optimal's gradient with
every forward pressure kept in HBM: one forward sweep), getMisfit().
setFp16Storage(True) stores those snapshots / checkpoints as FP16 with a
power-of-two scale per stored array.  The per-shot work is one call of
rsg_fwi_a2d_gradient.

AcousticFwiMisfit is the torch.autograd.Function (Vp, rho) -> misfit summed
over shots, whose backward returns the reference's Vp and rho gradients.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as R
from .core import GMAP, SMAP, Data2D, ModelAcoustic2D, Pml, SMAX, _dev, _stream

DIFFERENCE = 0  # rs_fwimisfit (utils.h:27)
_MODES = {"revolve": R.FWI_REVOLVE, "store": R.FWI_STORE, "full": R.FWI_FULL}


class Image2D:
    """Gradient image on the unpadded model grid ([nz, nx] float32)."""

    def __init__(self, model):
        self.nx, self.nz = model.nx, model.nz
        self.imagedata = np.zeros((self.nz, self.nx), np.float32)

    def getImagedata(self):
        return self.imagedata


class FwiAcoustic2D:
    def __init__(self, model: ModelAcoustic2D, source: Data2D, dataP: Data2D, order=4, snapinc=1):
        self.model, self.source, self.dataP = model, source, dataP
        self.order = int(order) if 1 < int(order) < 9 else 4      # fwi.cpp:28-33
        self.snapinc = max(int(snapinc), 1)
        self.ncheck, self.incore = 0, True
        self.datamodP = self.dataresP = self.vpgrad = self.rhograd = self.wavgrad = None
        self.misfit = 0.0
        self.misfit_type = DIFFERENCE
        self.fp16 = False

    def setDatamodP(self, d):
        self.datamodP = d

    def setDataresP(self, d):
        self.dataresP = d

    def setVpgrad(self, img):
        self.vpgrad = img

    def setRhograd(self, img):
        self.rhograd = img

    def setWavgrad(self, d):
        self.wavgrad = d

    def setNcheck(self, n):
        self.ncheck = int(n)

    def setIncore(self, v):
        self.incore = bool(v)

    def setMisfit_type(self, t):
        if int(t) != DIFFERENCE:
            raise R.RsgError("only the DIFFERENCE misfit is implemented on this path")
        self.misfit_type = DIFFERENCE

    def setFp16Storage(self, v):
        self.fp16 = bool(v)

    def getMisfit(self):
        return self.misfit

    def run_optimal(self):
        return self._run(R.FWI_REVOLVE)

    def run(self):
        return self._run(R.FWI_FULL)

    def run_stored(self):
        return self._run(R.FWI_STORE)

    def _run(self, mode):
        dev = _dev()
        m, src, obs = self.model, self.source, self.dataP
        if obs.field != R.PRESSURE or src.field != R.PRESSURE:
            raise R.RsgError("FwiAcoustic2D on this path takes PRESSURE sources and data")
        if self.vpgrad is None or self.rhograd is None:
            raise R.RsgError("FwiAcoustic2D:crossCorr: No gradient set in fwi class")
        if not hasattr(m, "_dL"):
            m.staggerModels()
        pml = Pml(m.lpml, src.dt)
        pml.setSmax(SMAX)
        pml.computeABC()
        g = R.Grid(nx=m.getNx_pml(), ny=1, nz=m.getNz_pml(), lpml=m.lpml, order=self.order,
                   fs=int(m.fs), dx=m.dx, dy=1.0, dz=m.dz, dt=src.dt)
        vp = torch.from_numpy(np.ascontiguousarray(m.Vp, np.float32)).to(dev)
        rho = torch.from_numpy(np.ascontiguousarray(m.R, np.float32)).to(dev)
        t_src, k1 = src._device(SMAP)
        t_obs, k2 = obs._device(GMAP)
        ntr, nt = obs.ntrace, obs.nt
        dmod = torch.zeros((ntr, nt), dtype=torch.float32, device=dev)
        dres = torch.zeros_like(dmod)
        wg = torch.zeros((src.ntrace, src.nt), dtype=torch.float32, device=dev)
        vg = torch.from_numpy(self.vpgrad.imagedata).to(dev)
        rg = torch.from_numpy(self.rhograd.imagedata).to(dev)
        misfit = C.c_double(0.0)
        f = R.Fwi()
        f.L, f.Rx, f.Rz, f.tab = m._dL.data_ptr(), m._dRx.data_ptr(), m._dRz.data_ptr(), pml.tab.data_ptr()
        f.Vp, f.Rho = vp.data_ptr(), rho.data_ptr()
        f.nx, f.nz = m.nx, m.nz
        f.src, f.obs = t_src, t_obs
        f.dmod, f.dres, f.wavgrad = dmod.data_ptr(), dres.data_ptr(), wg.data_ptr()
        f.vpgrad, f.rhograd = vg.data_ptr(), rg.data_ptr()
        f.misfit = C.pointer(misfit)
        f.ncheck, f.mode = self.ncheck, mode
        f.snapinc, f.fp16 = self.snapinc, int(self.fp16)
        used = C.c_int(-1)
        f.mode_used = C.pointer(used)
        R.check(R.lib().rsg_fwi_a2d_gradient(C.byref(g), C.byref(f), _stream()), "rsg_fwi_a2d_gradient")
        self.misfit = float(np.float32(misfit.value))
        self.vpgrad.imagedata[...] = vg.cpu().numpy()
        self.rhograd.imagedata[...] = rg.cpu().numpy()
        if self.datamodP is not None:
            self.datamodP.data[...] = dmod.cpu().numpy()
        if self.dataresP is not None:
            self.dataresP.data[...] = dres.cpu().numpy()
        if self.wavgrad is not None:
            self.wavgrad.data[...] = wg.cpu().numpy()
        self.misfit64 = misfit.value
        self.mode_used = {R.FWI_REVOLVE: "revolve", R.FWI_STORE: "store", R.FWI_FULL: "full"}[used.value]
        return 1


class Shot:
    """One shot of a survey: source coordinates (nsrc, 2), wavelet (nsrc, nt),
    receiver coordinates (nrec, 2) and observed pressure (nrec, nt)."""

    def __init__(self, scoords, wavelet, gcoords, dobs):
        self.scoords = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
        self.wavelet = np.ascontiguousarray(wavelet, np.float32).reshape(self.scoords.shape[0], -1)
        self.gcoords = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
        self.dobs = np.ascontiguousarray(dobs, np.float32).reshape(self.gcoords.shape[0], -1)


class Survey:
    """Shots on one model grid plus the gradient settings: mode "revolve"
    (run_optimal), "store" (run_optimal's gradient, every forward pressure in
    HBM) or "full" (run with snapshots every snapinc steps); fp16 stores the
    snapshots / checkpoints as scaled FP16."""

    def __init__(self, shots, lpml, dx, dz, dt, order=4, fs=False, ncheck=50, mode="revolve",
                 batched=True, snapinc=1, fp16=False):
        self.shots = list(shots)
        self.snapinc, self.fp16 = int(snapinc), bool(fp16)
        self.batched = bool(batched)
        self.lpml, self.dx, self.dz, self.dt = int(lpml), float(dx), float(dz), float(dt)
        self.order, self.fs, self.ncheck = int(order), bool(fs), int(ncheck)
        self.mode = mode


def shot_gradient(vp, rho, survey: Survey, shot: Shot):
    """Misfit and (Vp, rho) gradients of one shot, as numpy arrays."""
    nz, nx = vp.shape
    m = ModelAcoustic2D(nx, nz, survey.lpml, survey.dx, survey.dz, 0.0, 0.0, survey.fs)
    m.Vp[:] = vp
    m.R[:] = rho
    nt = shot.wavelet.shape[1]
    src = Data2D(shot.scoords.shape[0], nt, survey.dt, 0.0, "P")
    src.scoords[:] = shot.scoords
    src.data[:] = shot.wavelet
    src.makeMap(m, SMAP)
    obs = Data2D(shot.gcoords.shape[0], nt, survey.dt, 0.0, "P")
    obs.gcoords[:] = shot.gcoords
    obs.data[:] = shot.dobs
    obs.makeMap(m, GMAP)
    m.staggerModels()
    fwi = FwiAcoustic2D(m, src, obs, survey.order, survey.snapinc)
    fwi.setFp16Storage(survey.fp16)
    vg, rg = Image2D(m), Image2D(m)
    fwi.setVpgrad(vg)
    fwi.setRhograd(rg)
    fwi.setNcheck(survey.ncheck)
    if survey.mode == "store":
        fwi.run_stored()
    elif survey.mode == "full":
        fwi.run()
    else:
        fwi.run_optimal()
    return fwi.misfit64, vg.imagedata, rg.imagedata


def batch_gradient(vp, rho, survey: Survey, mode=None):
    """Misfit and (Vp, rho) gradients summed over all shots of the survey,
    computed as one batch (rsg_fwi_a2d_gradient_batch): every kernel of a time
    step runs all shots in one launch.  Shots must share nt and their source
    and receiver counts.  Returns (misfit, vpgrad, rhograd, dmod)."""
    dev = _dev()
    nz, nx = vp.shape
    shots = survey.shots
    ns = len(shots)
    nsrc, nrec, nt = shots[0].scoords.shape[0], shots[0].gcoords.shape[0], shots[0].wavelet.shape[1]
    for sh in shots:
        if sh.scoords.shape[0] != nsrc or sh.gcoords.shape[0] != nrec or sh.wavelet.shape[1] != nt:
            raise R.RsgError("batch_gradient: shots must share nt and trace counts")
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
                wav=dev_t(np.stack([sh.wavelet for sh in shots])),
                dobs=dev_t(np.stack([sh.dobs for sh in shots])),
                vp=dev_t(m.Vp), rho=dev_t(m.R))
    dmod = torch.zeros((ns, nrec, nt), dtype=torch.float32, device=dev)
    dres = torch.zeros_like(dmod)
    vg = torch.zeros((nz, nx), dtype=torch.float32, device=dev)
    rg = torch.zeros_like(vg)
    g = R.Grid(nx=m.getNx_pml(), ny=1, nz=m.getNz_pml(), lpml=m.lpml, order=survey.order,
               fs=int(m.fs), dx=m.dx, dy=1.0, dz=m.dz, dt=survey.dt)
    f = R.Fwi()
    f.L, f.Rx, f.Rz, f.tab = m._dL.data_ptr(), m._dRx.data_ptr(), m._dRz.data_ptr(), pml.tab.data_ptr()
    f.Vp, f.Rho = keep["vp"].data_ptr(), keep["rho"].data_ptr()
    f.nx, f.nz = nx, nz
    f.src = R.Traces(field=R.PRESSURE, ntrace=nsrc, nt=nt, map=keep["smap"].data_ptr(),
                     shift=keep["sshift"].data_ptr(), data=keep["wav"].data_ptr(), factor=None)
    f.obs = R.Traces(field=R.PRESSURE, ntrace=nrec, nt=nt, map=keep["gmap"].data_ptr(),
                     shift=keep["gshift"].data_ptr(), data=keep["dobs"].data_ptr(), factor=None)
    f.dmod, f.dres, f.wavgrad = dmod.data_ptr(), dres.data_ptr(), None
    f.vpgrad, f.rhograd = vg.data_ptr(), rg.data_ptr()
    mis = C.c_double(0.0)
    used = C.c_int(-1)
    f.misfit = C.pointer(mis)
    f.mode_used = C.pointer(used)
    mode = mode or survey.mode
    f.ncheck, f.mode = survey.ncheck, _MODES[mode]
    f.snapinc, f.fp16 = survey.snapinc, int(survey.fp16)
    R.check(R.lib().rsg_fwi_a2d_gradient_batch(C.byref(g), C.byref(f), ns, _stream()),
            "rsg_fwi_a2d_gradient_batch")
    survey.mode_used = {v: k for k, v in _MODES.items()}[used.value]
    return mis.value, vg, rg, dmod


class AcousticFwiMisfit(torch.autograd.Function):
    """(Vp, rho) -> sum over shots of the DIFFERENCE misfit.  The backward
    pass returns the reference's FwiAcoustic2D gradients (crossCorr of the
    forward pressure with the residual-driven wavefield)."""

    @staticmethod
    def forward(ctx, vp: torch.Tensor, rho: torch.Tensor, survey: Survey):
        v = vp.detach().to("cpu", torch.float32).numpy()
        r = rho.detach().to("cpu", torch.float32).numpy()
        if getattr(survey, "batched", True):
            total, gv, gr, _ = batch_gradient(v, r, survey)
            gv, gr = gv.to(vp.dtype), gr.to(rho.dtype)
        else:
            total = 0.0
            gvn = np.zeros_like(v)
            grn = np.zeros_like(r)
            for shot in survey.shots:
                mis, a, b = shot_gradient(v, r, survey, shot)
                total += mis
                gvn += a
                grn += b
            gv = torch.from_numpy(gvn).to(vp.device, vp.dtype)
            gr = torch.from_numpy(grn).to(rho.device, rho.dtype)
        ctx.save_for_backward(gv, gr)
        return torch.tensor(total, dtype=torch.float64, device=vp.device)

    @staticmethod
    def backward(ctx, gout):
        gv, gr = ctx.saved_tensors
        s = gout.to(gv.dtype)
        return s * gv, s * gr, None


def misfit(vp, rho, survey):
    return AcousticFwiMisfit.apply(vp, rho, survey)

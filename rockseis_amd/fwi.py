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

setAdjoint("exact") replaces the reference's adjoint (the forward operator
plus crossCorr) by the exact discrete adjoint of the step and of the model
staggering: the gradients are then those of the DIFFERENCE misfit of the
discrete forward map (Revolve only), and the wavelet gradient dJ/d(wavelet).

AcousticFwiMisfit is the torch.autograd.Function (Vp, rho) -> misfit summed
over shots, whose backward returns the survey's gradients (the reference's by
default, the exact ones with Survey(adjoint="exact")).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as R
from .core import (GMAP, SMAP, Data2D, Data3D, ModelAcoustic2D, ModelAcoustic3D, ModelElastic2D, ModelElastic3D, Pml,
                   SMAX, _dev, _stream)

DIFFERENCE, CORRELATION, ADAPTIVE_GAUSS, ADAPTIVE_LINEAR, ADAPTIVE_SINGLE = 0, 1, 2, 3, 4  # rs_fwimisfit (utils.h:26)
# the acoustic classes have no ADAPTIVE_SINGLE branch: it takes their default one, DIFFERENCE
_ACOUSTIC_MISFITS = (DIFFERENCE, CORRELATION, ADAPTIVE_GAUSS, ADAPTIVE_LINEAR, ADAPTIVE_SINGLE)
_ELASTIC_MISFITS = (DIFFERENCE, CORRELATION, ADAPTIVE_GAUSS, ADAPTIVE_LINEAR, ADAPTIVE_SINGLE)
_MODES = {"revolve": R.FWI_REVOLVE, "store": R.FWI_STORE, "full": R.FWI_FULL}
_ADJOINTS = {"reference": R.ADJOINT_REFERENCE, "exact": R.ADJOINT_EXACT}


def _adjoint_kind(kind):
    if kind not in _ADJOINTS:
        raise R.RsgError('adjoint must be "reference" or "exact"')
    return _ADJOINTS[kind]



# Storage of snapshots and Revolve checkpoints (RSG_STORE_* of rsgpu.h):
# False / 0 / "fp32", True / 1 / "fp16" (one power-of-two scale per array),
# 2 / "s16" (16-bit block floating point, 256 values per exponent).
STORE_FP32, STORE_FP16, STORE_S16 = 0, 1, 2


def storage_kind(v):
    if isinstance(v, str):
        kinds = {"fp32": STORE_FP32, "fp16": STORE_FP16, "s16": STORE_S16}
        if v.lower() not in kinds:
            raise R.RsgError(f"storage kind {v!r}: use 'fp32', 'fp16' or 's16'")
        return kinds[v.lower()]
    k = int(v)
    if k not in (STORE_FP32, STORE_FP16, STORE_S16):
        raise R.RsgError(f"storage kind {v!r}: use 0 (FP32), 1 (FP16) or 2 (S16)")
    return k

class Image2D:
    """Image on the unpadded model grid: [nz, nx] float32, or
    [nhz, nhx, nz, nx] with subsurface offsets (Image2D's ki2D order)."""

    def __init__(self, model, nhx=1, nhz=1):
        self.nx, self.nz = model.nx, model.nz
        self.nhx, self.nhz = int(nhx), int(nhz)
        if self.nhx < 1 or self.nhz < 1:
            raise R.RsgError("Image2D: nhx and nhz must be >= 1")
        shape = (self.nz, self.nx) if self.nhx * self.nhz == 1 else (self.nhz, self.nhx, self.nz, self.nx)
        self.imagedata = np.zeros(shape, np.float32)

    def getNhx(self):
        return self.nhx

    def getNhz(self):
        return self.nhz

    def getImagedata(self):
        return self.imagedata


class Image3D:
    """Image on the unpadded 3D model grid: [nz, ny, nx] float32, or
    [nhz, nhy, nhx, nz, ny, nx] with subsurface offsets (Image3D's ki3D
    order)."""

    def __init__(self, model, nhx=1, nhy=1, nhz=1):
        self.nx, self.ny, self.nz = model.nx, model.ny, model.nz
        self.nhx, self.nhy, self.nhz = int(nhx), int(nhy), int(nhz)
        if min(self.nhx, self.nhy, self.nhz) < 1:
            raise R.RsgError("Image3D: nhx, nhy and nhz must be >= 1")
        shape = (self.nz, self.ny, self.nx)
        if self.nhx * self.nhy * self.nhz > 1:
            shape = (self.nhz, self.nhy, self.nhx) + shape
        self.imagedata = np.zeros(shape, np.float32)

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
        self.dataweight = None
        self.fp16 = False
        self.adjoint = "reference"

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
        if int(t) not in _ACOUSTIC_MISFITS:
            raise R.RsgError("the acoustic misfit type is one of rs_fwimisfit's 0..4")
        self.misfit_type = int(t)

    def setDataweight(self, d):
        self.dataweight = d

    def setFp16Storage(self, v):
        self.fp16 = storage_kind(v)

    def setAdjoint(self, kind):
        """"reference" (the reference's gradient) or "exact" (the exact
        discrete adjoint; run_optimal and DIFFERENCE only)."""
        _adjoint_kind(kind)
        self.adjoint = kind

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
        f.misfit_type = self.misfit_type
        f.adjoint = _adjoint_kind(self.adjoint)
        if self.dataweight is not None:
            wt = torch.from_numpy(np.ascontiguousarray(self.dataweight.data, np.float32)).to(dev)
            f.weight = wt.data_ptr()
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


def misfit_residuals(misfit_type, dmod, dobs, gmap, weight=None, scaling=1):
    """FwiAcoustic2D/3D::computeMisfit and computeResiduals on the device for
    modelled records dmod and observed dobs (ntrace, nt) or (nshot, ntrace,
    nt), receiver maps gmap (..., ntrace, 2 or 3; traces with a negative entry
    count in no misfit), optional data weights.  Maps with 2 indices follow
    FwiAcoustic2D (types 1-3 clipped at the 97th percentile), with 3
    FwiAcoustic3D (no clip).  Returns (misfit, residuals)."""
    dev = _dev()
    mod = torch.as_tensor(np.ascontiguousarray(dmod, np.float32)).to(dev)
    rec = torch.as_tensor(np.ascontiguousarray(dobs, np.float32)).to(dev)
    mp = torch.as_tensor(np.ascontiguousarray(gmap, np.int32)).to(dev)
    if int(misfit_type) not in _ACOUSTIC_MISFITS:
        raise R.RsgError("the acoustic misfit type is one of rs_fwimisfit's 0..4")
    nt, ntr = mod.shape[-1], mod.shape[-2]
    nshot = mod.numel() // (nt * ntr) if nt * ntr else 1
    dim = mp.shape[-1]
    wt = None if weight is None else torch.as_tensor(np.ascontiguousarray(weight, np.float32)).to(dev)
    res = torch.zeros_like(mod)
    mis = C.c_double(0.0)
    R.check(R.lib().rsg_misfit_residuals(int(misfit_type), mod.data_ptr(), rec.data_ptr(),
                                         None if wt is None else wt.data_ptr(), res.data_ptr(), mp.data_ptr(), dim,
                                         ntr, nt, nshot, int(scaling), C.byref(mis), _stream()),
            "rsg_misfit_residuals")
    return mis.value, res.cpu().numpy()


def elastic_misfit_residuals(misfit_type, dmod, dobs, gmap, weights=None):
    """FwiElastic2D::computeMisfit and computeResiduals on the device, shot
    by shot as the reference runs them (percentile clips per shot): modelled
    and observed records dmod / dobs [3, nrec, nt] (P, Ux, Uz), the Ux
    receiver map gmap [nrec, 2] (traces with a negative entry count in no
    misfit), weights None or three optional [nrec, nt].  Lists of these (one
    entry per shot, nrec may differ) run as one batch.  Returns (misfit,
    residuals [3, nrec, nt]) or, for lists, (summed misfit, list of
    residuals)."""
    if int(misfit_type) not in _ELASTIC_MISFITS:
        raise R.RsgError(f"elastic misfit type {misfit_type} is not one of rs_fwimisfit's 0..4")
    single = not isinstance(dmod, (list, tuple))
    if single:
        dmod, dobs, gmap, weights = [dmod], [dobs], [gmap], [weights]
    elif weights is None:
        weights = [None] * len(dmod)
    dev = _dev()
    ns = len(dmod)
    recs = (R.E2DRecords * ns)()
    keep, res_all = [], []
    nt = None
    for k in range(ns):
        mod = torch.as_tensor(np.ascontiguousarray(dmod[k], np.float32)).to(dev)
        rec = torch.as_tensor(np.ascontiguousarray(dobs[k], np.float32)).to(dev)
        if mod.dim() != 3 or mod.shape[0] != 3 or rec.shape != mod.shape:
            raise R.RsgError("elastic_misfit_residuals: records must be [3, nrec, nt]")
        if nt is None:
            nt = mod.shape[2]
        elif mod.shape[2] != nt:
            raise R.RsgError("elastic_misfit_residuals: every shot needs the same nt")
        mp = torch.as_tensor(np.ascontiguousarray(gmap[k], np.int32).reshape(-1, 2)).to(dev)
        res = torch.zeros_like(mod)
        w = weights[k]
        wts = [None if w is None or w[c] is None else
               torch.as_tensor(np.ascontiguousarray(w[c], np.float32)).to(dev) for c in range(3)]
        keep += [mod, rec, mp, res] + [t for t in wts if t is not None]
        r = recs[k]
        r.nrec, r.map = mod.shape[1], mp.data_ptr()
        for c in range(3):
            r.mod[c], r.obs[c], r.res[c] = mod[c].data_ptr(), rec[c].data_ptr(), res[c].data_ptr()
            r.weight[c] = None if wts[c] is None else wts[c].data_ptr()
        res_all.append(res)
    mis = C.c_double(0.0)
    R.check(R.lib().rsg_misfit_residuals_e2d(int(misfit_type), recs, ns, nt, C.byref(mis), _stream()),
            "rsg_misfit_residuals_e2d")
    out = [r.cpu().numpy() for r in res_all]
    return mis.value, (out[0] if single else out)


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
    snapshots / checkpoints as scaled FP16; adjoint "reference" or "exact"
    (exact discrete adjoint, mode "revolve")."""

    def __init__(self, shots, lpml, dx, dz, dt, order=4, fs=False, ncheck=50, mode="revolve",
                 batched=True, snapinc=1, fp16=False, adjoint="reference"):
        self.shots = list(shots)
        _adjoint_kind(adjoint)
        self.adjoint = adjoint
        self.snapinc, self.fp16 = int(snapinc), storage_kind(fp16)
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
    fwi.setAdjoint(getattr(survey, "adjoint", "reference"))
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
    on_dev = isinstance(vp, torch.Tensor) and vp.is_cuda
    nz, nx = vp.shape
    shots = survey.shots
    ns = len(shots)
    nsrc, nrec, nt = shots[0].scoords.shape[0], shots[0].gcoords.shape[0], shots[0].wavelet.shape[1]
    for sh in shots:
        if sh.scoords.shape[0] != nsrc or sh.gcoords.shape[0] != nrec or sh.wavelet.shape[1] != nt:
            raise R.RsgError("batch_gradient: shots must share nt and trace counts")
    m = ModelAcoustic2D(nx, nz, survey.lpml, survey.dx, survey.dz, 0.0, 0.0, survey.fs)
    if not on_dev:
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
    if on_dev:
        # the model stays in HBM: staggered on the device, bit-identical to
        # the host set-up
        vp_d = vp.detach().to(dev, torch.float32).contiguous()
        rho_d = rho.detach().to(dev, torch.float32).contiguous()
        m.staggerModelsDevice(vp_d, rho_d)
    else:
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
                vp=vp_d if on_dev else dev_t(m.Vp), rho=rho_d if on_dev else dev_t(m.R))
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
    f.adjoint = _adjoint_kind(getattr(survey, "adjoint", "reference"))
    R.check(R.lib().rsg_fwi_a2d_gradient_batch(C.byref(g), C.byref(f), ns, _stream()),
            "rsg_fwi_a2d_gradient_batch")
    survey.mode_used = {v: k for k, v in _MODES.items()}[used.value]
    return mis.value, vg, rg, dmod


class AcousticFwiMisfit(torch.autograd.Function):
    """(Vp, rho) -> sum over shots of the DIFFERENCE misfit.  The backward
    pass returns the reference's FwiAcoustic2D gradients (crossCorr of the
    forward pressure with the residual-driven wavefield) or, for a survey with
    adjoint="exact", the exact gradient of the misfit."""

    @staticmethod
    def forward(ctx, vp: torch.Tensor, rho: torch.Tensor, survey: Survey):
        if getattr(survey, "batched", True):
            # CUDA inputs stay on the device (staggered there); CPU inputs
            # go through the host set-up
            a = (vp, rho) if vp.is_cuda and rho.is_cuda else \
                tuple(t.detach().to("cpu", torch.float32).numpy() for t in (vp, rho))
            total, gv, gr, _ = batch_gradient(*a, survey)
            gv, gr = gv.to(vp.device, vp.dtype), gr.to(rho.device, rho.dtype)
        else:
            v = vp.detach().to("cpu", torch.float32).numpy()
            r = rho.detach().to("cpu", torch.float32).numpy()
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


# --------------------------------------------------------------------------
# Elastic 2D (FwiElastic2D, include/fwi.h:243-300, lib/fwi.cpp:2203-4143)
# --------------------------------------------------------------------------

E2D_COMPONENTS = ("P", "VX", "VZ")


class ElasticShot:
    """One shot of an elastic survey: source coordinates (nsrc, 2) and
    wavelet (nsrc, nt) of field `srcfield` (P explosive, VX or VZ), receiver
    coordinates (nrec, 2) and observed records dobs[3, nrec, nt] (P, Ux, Uz),
    all in the coordinates of the full model; the shot runs on the window of
    columns [x0, x0 + nxw) (the whole model when nxw is None).  weights: None
    or three data weights (P, Ux, Uz), each None or [nrec, nt]
    (setDataweightp/x/z)."""

    def __init__(self, scoords, wavelet, gcoords, dobs, x0=0, nxw=None, srcfield="P", weights=None):
        self.scoords = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
        self.wavelet = np.ascontiguousarray(wavelet, np.float32).reshape(self.scoords.shape[0], -1)
        self.gcoords = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
        self.dobs = np.ascontiguousarray(dobs, np.float32).reshape(3, self.gcoords.shape[0], -1)
        self.x0, self.nxw, self.srcfield = int(x0), nxw, srcfield
        if weights is None:
            weights = (None, None, None)
        if len(weights) != 3:
            raise R.RsgError("ElasticShot: weights must be three arrays (P, Ux, Uz) or None each")
        self.weights = [None if w is None else np.ascontiguousarray(w, np.float32).reshape(self.dobs.shape[1:])
                        for w in weights]


class ElasticSurvey:
    """Shots of one elastic model plus the gradient settings (order, PML,
    Revolve snapshots, FP16 checkpoints, misfit type 0..4 as rs_fwimisfit).
    adjoint "reference" (the reference's adjoint-state steps and crossCorr)
    or "exact" (the exact discrete adjoint of the forward step and of the
    model staggering: the gradient of the DIFFERENCE misfit, with scale=False
    the gradients of lambda / mu / rho as independent parameters)."""

    def __init__(self, shots, lpml, dx, dz, dt, order=8, fs=False, ncheck=50, fp16=False, ox=0.0, oz=0.0,
                 misfit_type=DIFFERENCE, adjoint="reference"):
        if int(misfit_type) not in _ELASTIC_MISFITS:
            raise R.RsgError(f"ElasticSurvey: misfit type {misfit_type} is not one of rs_fwimisfit's 0..4")
        _adjoint_kind(adjoint)
        self.adjoint = adjoint
        self.misfit_type = int(misfit_type)
        self.shots = list(shots)
        self.lpml, self.dx, self.dz, self.dt = int(lpml), float(dx), float(dz), float(dt)
        self.order, self.fs, self.ncheck, self.fp16 = int(order), bool(fs), int(ncheck), storage_kind(fp16)
        self.ox, self.oz = float(ox), float(oz)  # origin of the full model


class _E2DBatch:
    """Device set-up of an ElasticSurvey on the model (vp, vs, rho [nz, nx]):
    per shot a staggered window model, device records, residual and
    optional wavelet-gradient arrays, and the C structures of
    rsg_fwi_e2d_gradient_batch / rsg_fwi_e2d_plan_create."""

    def __init__(self, vp, vs, rho, survey: ElasticSurvey, scale=True, wavgrad=False):
        dev = _dev()
        self.on_dev = isinstance(vp, torch.Tensor) and vp.is_cuda
        if self.on_dev:
            vp, vs, rho = (t.detach().to(dev, torch.float32).contiguous() for t in (vp, vs, rho))
        else:
            vp, vs, rho = (np.ascontiguousarray(a, np.float32) for a in (vp, vs, rho))
        nz, nx = vp.shape
        self.survey, self.nx, self.nz = survey, nx, nz
        ns = len(survey.shots)
        self.keep, self.models, self.windows, self.outs = [], [], [], []
        self.shots = (R.FwiE2DShot * ns)()

        def dev_t(a, dtype=torch.float32):
            t = torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)
            self.keep.append(t)
            return t

        for k, sh in enumerate(survey.shots):
            nxw = nx - sh.x0 if sh.nxw is None else int(sh.nxw)
            if sh.x0 < 0 or sh.x0 + nxw > nx:
                raise R.RsgError(f"elastic_batch_gradient: window of shot {k} outside the model")
            m = ModelElastic2D(nxw, nz, survey.lpml, survey.dx, survey.dz, survey.ox + sh.x0 * survey.dx,
                               survey.oz, survey.fs)
            win = [a[:, sh.x0:sh.x0 + nxw] for a in (vp, vs, rho)]
            if self.on_dev:
                m.staggerModelsDevice(*win)
            else:
                m.Vp[:], m.Vs[:], m.R[:] = win
                m.staggerModels()
            self.models.append(m)
            self.windows.append((sh.x0, nxw))
            nt = sh.wavelet.shape[1]
            src = Data2D(sh.scoords.shape[0], nt, survey.dt, 0.0, sh.srcfield)
            src.scoords[:] = sh.scoords
            src.data[:] = sh.wavelet
            src.makeMap(m, SMAP)
            t_src, k1 = src._device(SMAP)
            self.keep.append(k1)
            s = self.shots[k]
            s.g = R.Grid(nx=m.getNx_pml(), ny=1, nz=m.getNz_pml(), lpml=m.lpml, order=survey.order, fs=int(m.fs),
                         dx=m.dx, dy=1.0, dz=m.dz, dt=survey.dt)
            s.L, s.L2M, s.M, s.Rx, s.Rz = (getattr(m, "_d" + n).data_ptr() for n in ("L", "L2M", "M", "Rx", "Rz"))
            if self.on_dev:
                tv = [w.contiguous() for w in win]
                self.keep.extend(tv)
            else:
                tv = [dev_t(a) for a in (m.Vp, m.Vs, m.R)]
            m._dwin = tv  # the unpadded window model the gradient scaling reads
            s.Vp, s.Vs, s.Rho = (t.data_ptr() for t in tv)
            s.ox = sh.x0
            s.src = t_src
            o = {"dmod": [], "dres": []}
            for c, f in enumerate(E2D_COMPONENTS):
                d = Data2D(sh.gcoords.shape[0], nt, survey.dt, 0.0, f)
                d.gcoords[:] = sh.gcoords
                d.data[:] = sh.dobs[c]
                d.makeMap(m, GMAP)
                t, kk = d._device(GMAP)
                self.keep.append(kk)
                s.obs[c] = t
                dm = torch.zeros((d.ntrace, nt), dtype=torch.float32, device=dev)
                dr = torch.zeros_like(dm)
                o["dmod"].append(dm)
                o["dres"].append(dr)
                s.dmod[c], s.dres[c] = dm.data_ptr(), dr.data_ptr()
                s.weight[c] = None if sh.weights[c] is None else dev_t(sh.weights[c]).data_ptr()
            if wavgrad:
                wg = torch.zeros((src.ntrace, nt), dtype=torch.float32, device=dev)
                o["wavgrad"] = wg
                s.wavgrad = wg.data_ptr()
            self.outs.append(o)
        self.pml = Pml(survey.lpml, survey.dt)
        self.pml.setSmax(SMAX)
        self.pml.computeABC()
        self.grads = [torch.zeros((nz, nx), dtype=torch.float32, device=dev) for _ in range(3)]
        f = R.FwiE2D()
        f.tab, f.ncheck, f.fp16, f.scale, f.nx, f.nz = self.pml.tab.data_ptr(), survey.ncheck, int(survey.fp16), \
            int(bool(scale)), nx, nz
        for c in range(3):
            f.grad[c] = self.grads[c].data_ptr()
        f.misfit_type = survey.misfit_type
        f.adjoint = _adjoint_kind(getattr(survey, "adjoint", "reference"))
        self.f = f

    def tune(self, reps=5):
        """Pick the ragged launches' rows per thread for these windows
        (autotune.tune_ragged); call before a plan is built."""
        from . import autotune
        return autotune.tune_ragged(R.KIND_E2D, [s.g for s in self.shots], self.f.tab, reps)

    def set_model(self, vp, vs, rho):
        """A new model into the same device arrays (restaggered in place;
        on the device when the model is given as CUDA tensors)."""
        if isinstance(vp, torch.Tensor) and vp.is_cuda:
            vp, vs, rho = (t.detach().to(vp.device, torch.float32).contiguous() for t in (vp, vs, rho))
            if tuple(vp.shape) != (self.nz, self.nx):
                raise R.RsgError("set_model: model shape differs from the plan's")
            for m, (x0, nxw) in zip(self.models, self.windows):
                win = [a[:, x0:x0 + nxw] for a in (vp, vs, rho)]
                m.staggerModelsDevice(*win)
                for t, w in zip(m._dwin, win):
                    t.copy_(w)
            return
        vp, vs, rho = (np.ascontiguousarray(a, np.float32) for a in (vp, vs, rho))
        if vp.shape != (self.nz, self.nx):
            raise R.RsgError("set_model: model shape differs from the plan's")
        for m, (x0, nxw) in zip(self.models, self.windows):
            m.Vp[:], m.Vs[:], m.R[:] = vp[:, x0:x0 + nxw], vs[:, x0:x0 + nxw], rho[:, x0:x0 + nxw]
            m.staggerModels()
            for t, a in zip(m._dwin, (m.Vp, m.Vs, m.R)):
                t.copy_(torch.from_numpy(np.ascontiguousarray(a)))


def elastic_batch_gradient(vp, vs, rho, survey: ElasticSurvey, scale=True, wavgrad=False):
    """Misfit and gradients of all shots of `survey` on the model (vp, vs,
    rho [nz, nx]), every shot on its own window, as one ragged batch
    (rsg_fwi_e2d_gradient_batch).  scale=True returns the Vp / Vs / rho
    gradients of FwiElastic2D::run_optimal, scale=False the lambda / mu / rho
    kernels before scaleGrad.  Returns (misfit, g0, g1, g2, per-shot dict of
    dmod / dres / wavgrad device tensors)."""
    bt = _E2DBatch(vp, vs, rho, survey, scale, wavgrad)
    mis = C.c_double(0.0)
    bt.f.misfit = C.pointer(mis)
    R.check(R.lib().rsg_fwi_e2d_gradient_batch(C.byref(bt.f), bt.shots, len(survey.shots), _stream()),
            "rsg_fwi_e2d_gradient_batch")
    return mis.value, bt.grads[0], bt.grads[1], bt.grads[2], bt.outs


class ElasticFwiPlan:
    """The batched elastic gradient as a persistent plan
    (rsg_fwi_e2d_plan_*): built once for a survey and a model shape, the
    Revolve sweep captured as two hipGraphs; run() replays them and returns
    (misfit, g0, g1, g2, per-shot outputs) like elastic_batch_gradient;
    set_model() changes the model between runs (an inversion's iterations).
    The gradient tensors are zeroed by run() and owned by the plan.  With
    tune (default) the ragged launches' rows per thread are first picked by
    the autotuner for these windows (autotune.tune_ragged)."""

    def __init__(self, vp, vs, rho, survey: ElasticSurvey, scale=True, wavgrad=False, tune=True):
        self._b = _E2DBatch(vp, vs, rho, survey, scale, wavgrad)
        # the launch shape is fixed in the captured graphs: choose it first
        self.tuned = self._b.tune() if tune else None
        torch.cuda.current_stream().synchronize()  # inputs in place before the capture
        h = C.c_void_p()
        R.check(R.lib().rsg_fwi_e2d_plan_create(C.byref(self._b.f), self._b.shots, len(survey.shots), C.byref(h)),
                "rsg_fwi_e2d_plan_create")
        self._h = h

    def set_model(self, vp, vs, rho):
        self._b.set_model(vp, vs, rho)

    def run(self):
        if self._h is None:
            raise R.RsgError("ElasticFwiPlan: closed")
        for g in self._b.grads:
            g.zero_()
        mis = C.c_double(0.0)
        R.check(R.lib().rsg_fwi_e2d_plan_run(self._h, C.byref(mis), _stream()), "rsg_fwi_e2d_plan_run")
        g = self._b.grads
        return mis.value, g[0], g[1], g[2], self._b.outs

    def close(self):
        if self._h is not None:
            R.check(R.lib().rsg_fwi_e2d_plan_destroy(self._h), "rsg_fwi_e2d_plan_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FwiElastic2D:
    """FwiElastic2D(model, source, dataP, dataUx, dataUz, order, snapinc)
    with the reference's setters and run_optimal() (Revolve checkpoints in
    HBM); misfit types DIFFERENCE, CORRELATION, ADAPTIVE_GAUSS,
    ADAPTIVE_LINEAR and ADAPTIVE_SINGLE over the P, Ux and Uz records, with
    optional data weights."""

    def __init__(self, model: ModelElastic2D, source: Data2D, dataP: Data2D, dataUx: Data2D, dataUz: Data2D,
                 order=4, snapinc=1):
        self.model, self.source = model, source
        self.data = [dataP, dataUx, dataUz]
        self.order = int(order) if 1 < int(order) < 9 else 4      # fwi.cpp:28-33
        self.snapinc = max(int(snapinc), 1)
        self.ncheck, self.incore, self.fp16 = 0, True, False
        self.datamod = [None, None, None]
        self.datares = [None, None, None]
        self.vpgrad = self.vsgrad = self.rhograd = self.wavgrad = None
        self.misfit = 0.0
        self.misfit_type = DIFFERENCE
        self.dataweight = [None, None, None]

    def setDataweightp(self, d):
        self.dataweight[0] = d

    def setDataweightx(self, d):
        self.dataweight[1] = d

    def setDataweightz(self, d):
        self.dataweight[2] = d

    def setDatamodP(self, d):
        self.datamod[0] = d

    def setDatamodUx(self, d):
        self.datamod[1] = d

    def setDatamodUz(self, d):
        self.datamod[2] = d

    def setDataresP(self, d):
        self.datares[0] = d

    def setDataresUx(self, d):
        self.datares[1] = d

    def setDataresUz(self, d):
        self.datares[2] = d

    def setVpgrad(self, img):
        self.vpgrad = img

    def setVsgrad(self, img):
        self.vsgrad = img

    def setRhograd(self, img):
        self.rhograd = img

    def setWavgrad(self, d):
        self.wavgrad = d

    def setNcheck(self, n):
        self.ncheck = int(n)

    def setIncore(self, v):
        self.incore = bool(v)

    def setFp16Storage(self, v):
        self.fp16 = storage_kind(v)

    def setMisfit_type(self, t):
        if int(t) not in _ELASTIC_MISFITS:
            raise R.RsgError(f"FwiElastic2D: misfit type {t} is not one of rs_fwimisfit's 0..4")
        self.misfit_type = int(t)

    def getMisfit_type(self):
        return self.misfit_type

    def getMisfit(self):
        return self.misfit

    def run_optimal(self):
        m = self.model
        if any(d is None for d in self.data):
            raise R.RsgError("FwiElastic2D: P, Ux and Uz data must be set")
        wts = [None if w is None else w.data for w in self.dataweight]
        shot = ElasticShot(self.source.scoords, self.source.data, self.data[0].gcoords,
                           np.stack([d.data for d in self.data]), 0, m.nx, _field_name(self.source.field), wts)
        sv = ElasticSurvey([shot], m.lpml, m.dx, m.dz, self.source.dt, self.order, m.fs, self.ncheck, self.fp16,
                           m.ox, m.oz, self.misfit_type)
        mis, g0, g1, g2, outs = elastic_batch_gradient(m.Vp, m.Vs, m.R, sv, scale=True,
                                                        wavgrad=self.wavgrad is not None)
        self.misfit64 = mis
        self.misfit = float(np.float32(mis))
        for img, g in ((self.vpgrad, g0), (self.vsgrad, g1), (self.rhograd, g2)):
            if img is not None:
                img.imagedata[...] = g.cpu().numpy()
        for c in range(3):
            if self.datamod[c] is not None:
                self.datamod[c].data[...] = outs[0]["dmod"][c].cpu().numpy()
            if self.datares[c] is not None:
                self.datares[c].data[...] = outs[0]["dres"][c].cpu().numpy()
        if self.wavgrad is not None:
            self.wavgrad.data[...] = outs[0]["wavgrad"].cpu().numpy()
        return 1


def _field_name(f):
    return {R.PRESSURE: "P", R.VX: "VX", R.VZ: "VZ"}[f]


class ElasticFwiMisfit(torch.autograd.Function):
    """(Vp, Vs, rho) -> elastic misfit (the survey's misfit type) summed over the shots of an
    ElasticSurvey (each on its window, one ragged batch); the backward pass
    returns the reference's FwiElastic2D Vp / Vs / rho gradients."""

    @staticmethod
    def forward(ctx, vp: torch.Tensor, vs: torch.Tensor, rho: torch.Tensor, survey: ElasticSurvey):
        t3 = (vp, vs, rho)
        a = t3 if all(t.is_cuda for t in t3) else [t.detach().to("cpu", torch.float32).numpy() for t in t3]
        total, g0, g1, g2, _ = elastic_batch_gradient(*a, survey, scale=True)
        ctx.save_for_backward(g0.to(vp.device, vp.dtype), g1.to(vs.device, vs.dtype), g2.to(rho.device, rho.dtype))
        return torch.tensor(total, dtype=torch.float64, device=vp.device)

    @staticmethod
    def backward(ctx, gout):
        g0, g1, g2 = ctx.saved_tensors
        s = gout.to(g0.dtype)
        return s * g0, s * g1, s * g2, None


def elastic_misfit(vp, vs, rho, survey):
    return ElasticFwiMisfit.apply(vp, vs, rho, survey)


# --------------------------------------------------------------------------
# Acoustic 3D (FwiAcoustic3D, include/fwi.h:186-240, lib/fwi.cpp:1355-2171)
# --------------------------------------------------------------------------

class FwiAcoustic3D:
    """FwiAcoustic3D(model, source, dataP, order, snapinc) with the
    reference's setters and run_optimal() (Revolve checkpoints in HBM, FP32
    or, with setFp16Storage(True), FP16 with one scale per array).
    setAdjoint("exact") replaces the reference's adjoint by the exact
    discrete adjoint of the step and of the model staggering (DIFFERENCE
    misfit only)."""

    def __init__(self, model: ModelAcoustic3D, source: Data3D, dataP: Data3D, order=4, snapinc=1):
        self.model, self.source, self.dataP = model, source, dataP
        self.order = int(order) if 1 < int(order) < 9 else 4      # fwi.cpp:28-33
        self.snapinc = max(int(snapinc), 1)
        self.ncheck, self.incore, self.fp16 = 0, True, False
        self.datamodP = self.dataresP = self.vpgrad = self.rhograd = self.wavgrad = None
        self.misfit = 0.0
        self.misfit_type, self.dataweight = DIFFERENCE, None
        self.adjoint = "reference"

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

    def setFp16Storage(self, v):
        self.fp16 = storage_kind(v)

    def setMisfit_type(self, t):
        if int(t) not in _ACOUSTIC_MISFITS:
            raise R.RsgError("the acoustic misfit type is one of rs_fwimisfit's 0..4")
        self.misfit_type = int(t)

    def setDataweight(self, d):
        self.dataweight = d

    def setAdjoint(self, kind):
        """"reference" (the reference's gradient) or "exact" (the exact
        discrete adjoint; DIFFERENCE only)."""
        _adjoint_kind(kind)
        self.adjoint = kind

    def getMisfit(self):
        return self.misfit

    def run_optimal(self):
        dev = _dev()
        m, src, obs = self.model, self.source, self.dataP
        if obs.field != R.PRESSURE or src.field != R.PRESSURE:
            raise R.RsgError("FwiAcoustic3D on this path takes PRESSURE sources and data")
        if self.vpgrad is None or self.rhograd is None:
            raise R.RsgError("FwiAcoustic3D:crossCorr: No gradient set in fwi class")
        if not hasattr(m, "_dL"):
            m.staggerModels()
        pml = Pml(m.lpml, src.dt)
        pml.setSmax(SMAX)
        pml.computeABC()
        g = R.Grid(nx=m.getNx_pml(), ny=m.getNy_pml(), nz=m.getNz_pml(), lpml=m.lpml, order=self.order,
                   fs=int(m.fs), dx=m.dx, dy=m.dy, dz=m.dz, dt=src.dt)
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
        f = R.FwiA3D()
        f.L, f.Rx, f.Ry, f.Rz = (getattr(m, "_d" + n).data_ptr() for n in ("L", "Rx", "Ry", "Rz"))
        f.tab, f.Vp, f.Rho = pml.tab.data_ptr(), vp.data_ptr(), rho.data_ptr()
        f.nx, f.ny, f.nz = m.nx, m.ny, m.nz
        f.src, f.obs = t_src, t_obs
        f.dmod, f.dres, f.wavgrad = dmod.data_ptr(), dres.data_ptr(), wg.data_ptr()
        f.vpgrad, f.rhograd = vg.data_ptr(), rg.data_ptr()
        f.misfit = C.pointer(misfit)
        f.ncheck, f.fp16 = self.ncheck, int(self.fp16)
        f.misfit_type = self.misfit_type
        f.adjoint = _adjoint_kind(self.adjoint)
        if self.dataweight is not None:
            wt = torch.from_numpy(np.ascontiguousarray(self.dataweight.data, np.float32)).to(dev)
            f.weight = wt.data_ptr()
        R.check(R.lib().rsg_fwi_a3d_gradient(C.byref(g), C.byref(f), _stream()), "rsg_fwi_a3d_gradient")
        self.misfit64 = misfit.value
        self.misfit = float(np.float32(misfit.value))
        self.vpgrad.imagedata[...] = vg.cpu().numpy()
        self.rhograd.imagedata[...] = rg.cpu().numpy()
        if self.datamodP is not None:
            self.datamodP.data[...] = dmod.cpu().numpy()
        if self.dataresP is not None:
            self.dataresP.data[...] = dres.cpu().numpy()
        if self.wavgrad is not None:
            self.wavgrad.data[...] = wg.cpu().numpy()
        return 1


class FwiElastic3D:
    """FwiElastic3D(model, source, dataUx, dataUy, dataUz, order, snapinc)
    with the reference's setters and run_optimal() (fwi.cpp:6770-6991):
    Revolve checkpoints of the nine fields and 36 strips in HBM (FP32, or
    16-bit with setFp16Storage), misfit types DIFFERENCE, CORRELATION,
    ADAPTIVE_GAUSS and the reference's default branch, Vp / Vs / rho and
    wavelet gradients.  The observed records are VX, VY, VZ traces with
    their makeMap(GMAP) maps."""

    def __init__(self, model: ModelElastic3D, source: Data3D, dataUx: Data3D, dataUy: Data3D, dataUz: Data3D,
                 order=4, snapinc=1):
        self.model, self.source = model, source
        self.dataUx, self.dataUy, self.dataUz = dataUx, dataUy, dataUz
        self.order = int(order) if 1 < int(order) < 9 else 4      # fwi.cpp:28-33
        self.snapinc = max(int(snapinc), 1)
        self.ncheck, self.incore, self.fp16 = 0, True, False
        self.datamodUx = self.datamodUy = self.datamodUz = None
        self.dataresUx = self.dataresUy = self.dataresUz = None
        self.vpgrad = self.vsgrad = self.rhograd = self.wavgrad = None
        self.dataweightx = self.dataweighty = self.dataweightz = None
        self.misfit = 0.0
        self.misfit_type = DIFFERENCE

    def setDatamodUx(self, d):
        self.datamodUx = d

    def setDatamodUy(self, d):
        self.datamodUy = d

    def setDatamodUz(self, d):
        self.datamodUz = d

    def setDataresUx(self, d):
        self.dataresUx = d

    def setDataresUy(self, d):
        self.dataresUy = d

    def setDataresUz(self, d):
        self.dataresUz = d

    def setDataweightx(self, d):
        self.dataweightx = d

    def setDataweighty(self, d):
        self.dataweighty = d

    def setDataweightz(self, d):
        self.dataweightz = d

    def setVpgrad(self, img):
        self.vpgrad = img

    def setVsgrad(self, img):
        self.vsgrad = img

    def setRhograd(self, img):
        self.rhograd = img

    def setWavgrad(self, d):
        self.wavgrad = d

    def setNcheck(self, n):
        self.ncheck = int(n)

    def setIncore(self, v):
        self.incore = bool(v)

    def setFp16Storage(self, v):
        self.fp16 = storage_kind(v)

    def setMisfit_type(self, t):
        if int(t) not in range(5):
            raise R.RsgError("the elastic misfit type is one of rs_fwimisfit's 0..4")
        self.misfit_type = int(t)

    def getMisfit(self):
        return self.misfit

    def run_optimal(self):
        dev = _dev()
        m, src = self.model, self.source
        obs = (self.dataUx, self.dataUy, self.dataUz)
        if tuple(d.field for d in obs) != (R.VX, R.VY, R.VZ):
            raise R.RsgError("FwiElastic3D takes Ux, Uy, Uz data as VX, VY, VZ traces")
        if self.vpgrad is None and self.vsgrad is None and self.rhograd is None and self.wavgrad is None:
            raise R.RsgError("FwiElastic3D::crossCorr: No gradient set for computation.")
        if getattr(m, "_dL2M", None) is None or tuple(m._dL2M.shape) != (m.getNz_pml(), m.getNy_pml(), m.getNx_pml()):
            m.staggerModels()
        pml = Pml(m.lpml, src.dt)
        pml.setSmax(SMAX)
        pml.computeABC()
        g = R.Grid(nx=m.getNx_pml(), ny=m.getNy_pml(), nz=m.getNz_pml(), lpml=m.lpml, order=self.order,
                   fs=int(m.fs), dx=m.dx, dy=m.dy, dz=m.dz, dt=src.dt)
        t_src, keep = src._device(SMAP)
        f = R.FwiE3D()
        for n in ("L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz", "Vp", "Vs"):
            setattr(f, n, getattr(m, "_d" + n).data_ptr())
        f.Rho, f.tab = m._dR.data_ptr(), pml.tab.data_ptr()
        f.nx, f.ny, f.nz = m.nx, m.ny, m.nz
        f.src = t_src
        ntr, nt = obs[0].ntrace, obs[0].nt
        dmod = torch.zeros((3, ntr, nt), dtype=torch.float32, device=dev)
        dres = torch.zeros_like(dmod)
        for c, d in enumerate(obs):
            t, k = d._device(GMAP)
            keep.append(k)
            f.obs[c] = t
            f.dmod[c], f.dres[c] = dmod[c].data_ptr(), dres[c].data_ptr()
        for c, w in enumerate((self.dataweightx, self.dataweighty, self.dataweightz)):
            if w is not None:
                wt = torch.from_numpy(np.ascontiguousarray(w.data, np.float32)).to(dev)
                keep.append(wt)
                f.weight[c] = wt.data_ptr()
        grads = [torch.zeros((m.nz, m.ny, m.nx), dtype=torch.float32, device=dev) for _ in range(3)]
        f.vpgrad, f.vsgrad, f.rhograd = (t.data_ptr() for t in grads)
        wg = torch.zeros((src.ntrace, src.nt), dtype=torch.float32, device=dev)
        f.wavgrad = wg.data_ptr()
        misfit = C.c_double(0.0)
        f.misfit = C.pointer(misfit)
        f.ncheck, f.fp16, f.misfit_type = self.ncheck, int(self.fp16), self.misfit_type
        R.check(R.lib().rsg_fwi_e3d_gradient(C.byref(g), C.byref(f), _stream()), "rsg_fwi_e3d_gradient")
        self.misfit64 = misfit.value
        self.misfit = float(np.float32(misfit.value))
        for img, t in zip((self.vpgrad, self.vsgrad, self.rhograd), grads):
            if img is not None:
                img.imagedata[...] = t.cpu().numpy()
        for c, (dm, dr) in enumerate(((self.datamodUx, self.dataresUx), (self.datamodUy, self.dataresUy),
                                      (self.datamodUz, self.dataresUz))):
            if dm is not None:
                dm.data[...] = dmod[c].cpu().numpy()
            if dr is not None:
                dr.data[...] = dres[c].cpu().numpy()
        if self.wavgrad is not None:
            self.wavgrad.data[...] = wg.cpu().numpy()
        return 1

"""Oracle restatement of FwiAcoustic2D::run_optimal (fwi.cpp:1159-1332): the
reference's per-shot gradient with Revolve checkpointing, DIFFERENCE misfit
(fwi.cpp:574-608), residuals (fwi.cpp:772-800) and the zero-lag
cross-correlation imaging condition crossCorr (fwi.cpp:511-571).  The
adjoint wavefield is advanced with the forward operator, as in the
reference.  Also the same run on the compiled reference.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import numpy as np

from . import fn, ref
from . import revolve as orv
from .runs import A2DState, inject2d, makemap2d, record2d, stagger_a2d


def _snapshot(s):
    return [s.P.copy(), s.Vx.copy(), s.Vz.copy()] + [a.copy() for a in s.strips.values()]


def _restore(s, snap):
    s.P[...] = snap[0]
    s.Vx[...] = snap[1]
    s.Vz[...] = snap[2]
    for a, b in zip(s.strips.values(), snap[3:]):
        a[...] = b


def fwi_a2d(vp, rho, lpml, dx, dz, dt, nt, order, fs, ncheck, scoords, wav, gcoords, dobs,
            real=np.float32):
    """Returns dict(misfit, vpgrad, rhograd, dmod, dres, wavgrad)."""
    nz, nx = vp.shape
    smap, sshift = makemap2d(scoords, nx, nz, dx, dz, "P")
    gmap, gshift = makemap2d(gcoords, nx, nz, dx, dz, "P")
    vp = np.ascontiguousarray(vp, real)
    rho = np.ascontiguousarray(rho, real)
    L, Rx, Rz = stagger_a2d(vp, rho, lpml, fs, real)
    fw = A2DState(L, Rx, Rz, lpml, order, dx, dz, dt, real)
    bw = A2DState(L, Rx, Rz, lpml, order, dx, dz, dt, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    dobs = np.ascontiguousarray(dobs, real).reshape(gmap.shape[0], nt)
    dmod = np.zeros_like(dobs)
    dres = np.zeros_like(dobs)
    wgrad = np.zeros_like(wav)
    vpgrad = np.zeros((nz, nx), real)
    rhograd = np.zeros((nz, nx), real)
    valid = ((gmap[:, 0] >= 0) & (gmap[:, 1] >= 0)).astype(np.int32)
    misfit = None
    checks = {}
    rev = orv.Revolve(nt, ncheck)
    reverse = False
    xc = fn("a2d_crosscorr", real)

    def corr():
        xc(fw.P.reshape(-1), lpml, bw.Vx.reshape(-1), bw.Vz.reshape(-1), lpml, vp.reshape(-1),
           rho.reshape(-1), fw.Rx.reshape(-1), fw.Rz.reshape(-1), nx, nz, real(dx), real(dz),
           vpgrad.reshape(-1), rhograd.reshape(-1))

    def fstep(it, rec):
        fw.velocity()
        fw.stress()
        inject2d(fw.P, fw.L, lpml, dt, smap, sshift, wav, it, real)
        if rec:
            record2d(fw.P, lpml, gmap, gshift, dmod, it, real)

    while True:
        oldcapo = rev.capo
        a = rev.step()
        capo = rev.capo
        if a == orv.ADVANCE:
            for it in range(oldcapo, capo):
                fstep(it, not reverse)
        elif a == orv.FIRSTURN:
            fstep(capo, True)
            misfit = fn("misfit_diff", real)(dmod.reshape(-1), dobs.reshape(-1), valid, dobs.shape[0], nt)
            fn("residual_diff", real)(dmod.reshape(-1), dobs.reshape(-1), dres.reshape(-1), dobs.shape[0], nt, 1)
            inject2d(bw.P, bw.L, lpml, dt, gmap, gshift, dres, capo, real)
            corr()
            record2d(bw.P, lpml, smap, sshift, wgrad, capo, real)
            reverse = True
        elif a == orv.YOUTURN:
            bw.velocity()
            bw.stress()
            inject2d(bw.P, bw.L, lpml, dt, gmap, gshift, dres, capo, real)
            corr()
            record2d(bw.P, lpml, smap, sshift, wgrad, capo, real)
        elif a == orv.TAKESHOT:
            checks[rev.check] = _snapshot(fw)
        elif a == orv.RESTORE:
            _restore(fw, checks[rev.check])
        if a in (orv.TERMINATE, orv.ERROR):
            break
    return dict(misfit=misfit, vpgrad=vpgrad, rhograd=rhograd, dmod=dmod, dres=dres, wavgrad=wgrad)


def ref_fwi_a2d(vp, rho, lpml, dx, dz, dt, nt, order, fs, ncheck, scoords, wav, gcoords, dobs):
    nz, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    nsrc, nrec = sc.shape[0], gc.shape[0]
    vg = np.zeros((nz, nx), np.float32)
    rg = np.zeros((nz, nx), np.float32)
    dmod = np.zeros((nrec, nt), np.float32)
    dres = np.zeros((nrec, nt), np.float32)
    wg = np.zeros((nsrc, nt), np.float32)
    m = ref().ref_fwi_a2d(np.ascontiguousarray(vp, np.float32).ravel(),
                          np.ascontiguousarray(rho, np.float32).ravel(), nx, nz, lpml, dx, dz, dt, nt,
                          order, int(fs), ncheck, nsrc, sc.ravel(),
                          np.ascontiguousarray(wav, np.float32).ravel(), nrec, gc.ravel(),
                          np.ascontiguousarray(dobs, np.float32).ravel(), vg.reshape(-1),
                          rg.reshape(-1), dmod.reshape(-1), dres.reshape(-1), wg.reshape(-1))
    return dict(misfit=np.float32(m), vpgrad=vg, rhograd=rg, dmod=dmod, dres=dres, wavgrad=wg)

"""Oracle restatement of FwiAcoustic2D::run_optimal (fwi.cpp:1159-1332) and
FwiAcoustic2D::run (fwi.cpp:1030-1156): the reference's per-shot gradient
with Revolve checkpointing or with snapshots every snapinc steps (Snapshot2D
cadence, snap.cpp:330-335, 638-689), DIFFERENCE misfit (fwi.cpp:574-608),
residuals (fwi.cpp:772-800) and the zero-lag cross-correlation imaging
condition crossCorr (fwi.cpp:511-571).  The adjoint wavefield is advanced
with the forward operator, as in the reference.  Also the same runs on the
compiled reference.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import numpy as np

from . import fn, ref
from . import revolve as orv
from .runs import (E3D_FIELDS, A2DState, A3DState, E2DState, E3DState, e2d_record_factor, inject2d,
                   makemap2d, makemap3d, record2d, stagger_a2d, stagger_a3d, stagger_e2d, stagger_e3d)


def _fields(s):
    if isinstance(s, E3DState):
        return [getattr(s, k) for k in E3D_FIELDS] + s.strips
    if isinstance(s, A3DState):
        return [s.P, s.Vx, s.Vy, s.Vz] + list(s.strips.values())
    if isinstance(s, E2DState):
        return [s.Vx, s.Vz, s.Sxx, s.Szz, s.Sxz] + list(s.strips.values())
    return [s.P, s.Vx, s.Vz] + list(s.strips.values())


def _snapshot(s):
    """Revolve payload: fields then PML strips (revolve.cpp:648-1271)."""
    return [a.copy() for a in _fields(s)]


def _restore(s, snap):
    for a, b in zip(_fields(s), snap):
        a[...] = b


def snap_cadence(nt, dt, snapinc):
    """(snapnt, enddiff) of Snapshot2D (snap.cpp:330-335), in the
    reference's float32 arithmetic."""
    f = np.float32
    snapnt = (nt - 1) // snapinc + 1
    dt = f(dt)
    snapdt = f(dt * f(snapinc))
    a = f(f(nt - 1) * dt)
    b = f(f(snapnt - 1) * snapdt)
    return snapnt, int(np.rint(f(f(a - b) / dt)))


def fwi_a2d(vp, rho, lpml, dx, dz, dt, nt, order, fs, ncheck, scoords, wav, gcoords, dobs,
            real=np.float32, misfit_type=0, weight=None):
    """Returns dict(misfit, vpgrad, rhograd, dmod, dres, wavgrad).
    misfit_type 0 DIFFERENCE, 1 CORRELATION; weight: optional data weights
    (nrec, nt) as setDataweight."""
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
    wei = None if weight is None else np.ascontiguousarray(weight, real).reshape(gmap.shape[0], nt)
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
            wp = None if wei is None else wei.ctypes.data
            misfit = fn("misfit_a", real)(misfit_type, dmod.reshape(-1), dobs.reshape(-1), wp, valid,
                                          dobs.shape[0], nt)
            fn("residual_a", real)(misfit_type, dmod.reshape(-1), dobs.reshape(-1), wp, dres.reshape(-1),
                                   dobs.shape[0], nt, 1, 1)
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


def ref_fwi_a2d_misfit(vp, rho, lpml, dx, dz, dt, nt, order, fs, ncheck, scoords, wav, gcoords, dobs,
                       misfit_type=0, weight=None):
    """The reference's run_optimal with a misfit type and data weights."""
    nz, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    nsrc, nrec = sc.shape[0], gc.shape[0]
    vg = np.zeros((nz, nx), np.float32)
    rg = np.zeros((nz, nx), np.float32)
    dmod = np.zeros((nrec, nt), np.float32)
    dres = np.zeros((nrec, nt), np.float32)
    wg = np.zeros((nsrc, nt), np.float32)
    w = None if weight is None else np.ascontiguousarray(weight, np.float32)
    m = ref().ref_fwi_a2d_misfit(np.ascontiguousarray(vp, np.float32).ravel(),
                                 np.ascontiguousarray(rho, np.float32).ravel(), nx, nz, lpml, dx, dz, dt, nt,
                                 order, int(fs), ncheck, nsrc, sc.ravel(),
                                 np.ascontiguousarray(wav, np.float32).ravel(), nrec, gc.ravel(),
                                 np.ascontiguousarray(dobs, np.float32).ravel(), int(misfit_type),
                                 None if w is None else w.ctypes.data, vg.reshape(-1), rg.reshape(-1),
                                 dmod.reshape(-1), dres.reshape(-1), wg.reshape(-1))
    return dict(misfit=np.float32(m), vpgrad=vg, rhograd=rg, dmod=dmod, dres=dres, wavgrad=wg)


def fwi_a2d_full(vp, rho, lpml, dx, dz, dt, nt, order, fs, snapinc, scoords, wav, gcoords, dobs,
                 real=np.float32):
    """FwiAcoustic2D::run (fwi.cpp:1030-1156): forward loop writing the
    unpadded pressure when it % snapinc == 0 (snap.cpp:645-665), misfit and
    residuals, then a fresh wavefield stepped nt times; at reverse time t:
    step, inject the residual at t, crossCorr with snapshot
    (t - enddiff) / snapinc when (t - enddiff) % snapinc == 0 (the read order
    of snap.cpp:672-689), record the wavelet gradient at t."""
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
    snapnt, enddiff = snap_cadence(nt, dt, snapinc)
    snaps = []
    for it in range(nt):
        fw.velocity()
        fw.stress()
        inject2d(fw.P, fw.L, lpml, dt, smap, sshift, wav, it, real)
        record2d(fw.P, lpml, gmap, gshift, dmod, it, real)
        if it % snapinc == 0:
            snaps.append(np.ascontiguousarray(fw.P[lpml:lpml + nz, lpml:lpml + nx]))
    assert len(snaps) == snapnt
    misfit = fn("misfit_diff", real)(dmod.reshape(-1), dobs.reshape(-1), valid, dobs.shape[0], nt)
    fn("residual_diff", real)(dmod.reshape(-1), dobs.reshape(-1), dres.reshape(-1), dobs.shape[0], nt, 1)
    xc = fn("a2d_crosscorr", real)
    for t in range(nt - 1, -1, -1):
        bw.velocity()
        bw.stress()
        inject2d(bw.P, bw.L, lpml, dt, gmap, gshift, dres, t, real)
        if (t - enddiff) % snapinc == 0:
            sn = snaps[(t - enddiff) // snapinc]
            xc(sn.reshape(-1), 0, bw.Vx.reshape(-1), bw.Vz.reshape(-1), lpml, vp.reshape(-1),
               rho.reshape(-1), fw.Rx.reshape(-1), fw.Rz.reshape(-1), nx, nz, real(dx), real(dz),
               vpgrad.reshape(-1), rhograd.reshape(-1))
        record2d(bw.P, lpml, smap, sshift, wgrad, t, real)
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


def ref_fwi_a2d_full(vp, rho, lpml, dx, dz, dt, nt, order, fs, snapinc, scoords, wav, gcoords, dobs):
    """The reference's FwiAcoustic2D::run; its snapshot file goes to a
    temporary directory and is removed by run()."""
    import os
    import tempfile
    nz, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    nsrc, nrec = sc.shape[0], gc.shape[0]
    vg = np.zeros((nz, nx), np.float32)
    rg = np.zeros((nz, nx), np.float32)
    dmod = np.zeros((nrec, nt), np.float32)
    dres = np.zeros((nrec, nt), np.float32)
    wg = np.zeros((nsrc, nt), np.float32)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "psnap.rss").encode()
        m = ref().ref_fwi_a2d_full(np.ascontiguousarray(vp, np.float32).ravel(),
                                   np.ascontiguousarray(rho, np.float32).ravel(), nx, nz, lpml, dx, dz, dt,
                                   nt, order, int(fs), snapinc, path, nsrc, sc.ravel(),
                                   np.ascontiguousarray(wav, np.float32).ravel(), nrec, gc.ravel(),
                                   np.ascontiguousarray(dobs, np.float32).ravel(), vg.reshape(-1),
                                   rg.reshape(-1), dmod.reshape(-1), dres.reshape(-1), wg.reshape(-1))
    return dict(misfit=np.float32(m), vpgrad=vg, rhograd=rg, dmod=dmod, dres=dres, wavgrad=wg)


E2D_FIELDS = ("P", "VX", "VZ")


def _e2d_inject(s, field, lpml, dt, m, sh, data, it, real):
    """WavesElastic2D::insertSource (waves.cpp:2085-2171): PRESSURE is an
    explosive source into Sxx and Szz (no modulus), VX / VZ scale by Rx / Rz."""
    if field == "P":
        inject2d(s.Sxx, None, lpml, dt, m, sh, data, it, real)
        inject2d(s.Szz, None, lpml, dt, m, sh, data, it, real)
    elif field == "VX":
        inject2d(s.Vx, s.Rx, lpml, dt, m, sh, data, it, real)
    else:
        inject2d(s.Vz, s.Rz, lpml, dt, m, sh, data, it, real)


def _e2d_record(s, field, lpml, m, sh, factor, data, it, real):
    """WavesElastic2D::recordData (waves.cpp:2174-2288)."""
    if field == "P":
        fn("record2d_pressure_el", real)(s.Sxx.reshape(-1), s.Szz.reshape(-1), s.nx, lpml, m.reshape(-1),
                                         np.ascontiguousarray(sh, real).reshape(-1), factor, data.reshape(-1),
                                         m.shape[0], data.shape[-1], it)
    else:
        record2d(s.Vx if field == "VX" else s.Vz, lpml, m, sh, data, it, real)


def _weights3(weights, nrec, nt, real):
    """Three optional [nrec, nt] weight arrays (P, Ux, Uz) or None."""
    if weights is None:
        return [None] * 3
    return [None if w is None else np.ascontiguousarray(w, real).reshape(nrec, nt) for w in weights]


def fwi_e2d(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, ncheck, srcfield, scoords, wav, gcoords, dobs,
            real=np.float32, misfit_type=0, weights=None, residuals=None):
    """FwiElastic2D::run_optimal (fwi.cpp:3937-4143) with P, Ux and Uz data
    (dobs[3, nrec, nt]): forward elastic steps, the misfit and residuals of
    the three records (fwi.cpp:2852-3773; misfit_type 0 DIFFERENCE,
    1 CORRELATION, 2 ADAPTIVE_GAUSS, 3 ADAPTIVE_LINEAR, 4 ADAPTIVE_SINGLE;
    weights: None or three optional [nrec, nt] data weights), adjoint with
    backwardstepVelocity/Stress (waves.cpp:1738-2080) driven by the three
    residuals, crossCorr (fwi.cpp:2591-2775) at every reverse step, then
    scaleGrad (fwi.cpp:2778-2849).  Returns the lambda/mu/rho kernels before
    scaleGrad as well as the Vp/Vs/rho gradients after it.  residuals: None,
    or [3, nrec, nt] residuals that drive the adjoint in place of the
    computed ones (to check an adjoint pass on given residuals)."""
    nz, nx = vp.shape
    smap, sshift = makemap2d(scoords, nx, nz, dx, dz, srcfield)
    maps = {f: makemap2d(gcoords, nx, nz, dx, dz, f) for f in E2D_FIELDS}
    vp, vs, rho = (np.ascontiguousarray(a, real) for a in (vp, vs, rho))
    L, L2M, M, Rx, Rz = stagger_e2d(vp, vs, rho, lpml, fs, real)
    fw = E2DState(L, L2M, M, Rx, Rz, lpml, order, dx, dz, dt, fs, real)
    bw = E2DState(L, L2M, M, Rx, Rz, lpml, order, dx, dz, dt, fs, real)
    nsrc, nrec = smap.shape[0], maps["P"][0].shape[0]
    wav = np.ascontiguousarray(wav, real).reshape(nsrc, nt)
    dobs = np.ascontiguousarray(dobs, real).reshape(3, nrec, nt)
    dmod = np.zeros_like(dobs)
    dres = np.zeros_like(dobs)
    wgrad = np.zeros_like(wav)
    lam = np.zeros((nz, nx), real)
    mu = np.zeros((nz, nx), real)
    rhok = np.zeros((nz, nx), real)
    factor = e2d_record_factor(vp, vs, rho, maps["P"][0], False, real)
    gx = maps["VX"][0]
    valid = ((gx[:, 0] >= 0) & (gx[:, 1] >= 0)).astype(np.int32)
    wf = {"P": 0, "VX": 1, "VZ": 3}[srcfield]
    xc = fn("e2d_crosscorr", real)

    def corr(capo):
        xc(fw.Vx.reshape(-1), fw.Vz.reshape(-1), fw.Sxx.reshape(-1), fw.Szz.reshape(-1), fw.Sxz.reshape(-1),
           bw.Vx.reshape(-1), bw.Vz.reshape(-1), bw.Sxx.reshape(-1), bw.Szz.reshape(-1), bw.Sxz.reshape(-1),
           fw.Rx.reshape(-1), fw.Rz.reshape(-1), lpml, nx, nz, real(dx), real(dz), lam.reshape(-1),
           mu.reshape(-1), rhok.reshape(-1), wgrad.ctypes.data, wf, smap.ctypes.data,
           np.ascontiguousarray(sshift, real).ctypes.data, nsrc, nt, capo)

    def fstep(it, rec):
        fw.velocity()
        fw.stress()
        _e2d_inject(fw, srcfield, lpml, dt, smap, sshift, wav, it, real)
        if rec:
            for k, f in enumerate(E2D_FIELDS):
                _e2d_record(fw, f, lpml, maps[f][0], maps[f][1], factor, dmod[k], it, real)

    def inject_res(capo):
        for k, f in enumerate(E2D_FIELDS):
            _e2d_inject(bw, f, lpml, dt, maps[f][0], maps[f][1], dres[k], capo, real)

    misfit = None
    checks = {}
    rev = orv.Revolve(nt, ncheck)
    reverse = False
    while True:
        oldcapo = rev.capo
        a = rev.step()
        capo = rev.capo
        if a == orv.ADVANCE:
            for it in range(oldcapo, capo):
                fstep(it, not reverse)
        elif a == orv.FIRSTURN:
            fstep(capo, True)
            wts = _weights3(weights, nrec, nt, real)
            wp = [None if w is None else w.ctypes.data for w in wts]
            md = [dmod[k].reshape(-1) for k in range(3)] + [dobs[k].reshape(-1) for k in range(3)]
            md = [md[0], md[3], md[1], md[4], md[2], md[5]]
            misfit = fn("misfit_e", real)(misfit_type, *md, *wp, valid, nrec, nt)
            fn("residual_e", real)(misfit_type, *md, *wp, dres[0].reshape(-1), dres[1].reshape(-1),
                                   dres[2].reshape(-1), nrec, nt)
            if residuals is not None:
                dres[...] = np.asarray(residuals, real).reshape(dres.shape)
            inject_res(capo)
            corr(capo)
            reverse = True
        elif a == orv.YOUTURN:
            bw.velocity_bw()
            bw.stress_bw()
            inject_res(capo)
            corr(capo)
        elif a == orv.TAKESHOT:
            checks[rev.check] = _snapshot(fw)
        elif a == orv.RESTORE:
            _restore(fw, checks[rev.check])
        if a in (orv.TERMINATE, orv.ERROR):
            break
    vpg, vsg, rhg = lam.copy(), mu.copy(), rhok.copy()
    fn("e2d_scalegrad", real)(vp.reshape(-1), vs.reshape(-1), rho.reshape(-1), nx, nz, vpg.reshape(-1),
                              vsg.reshape(-1), rhg.reshape(-1))
    return dict(misfit=misfit, lam=lam, mu=mu, rhok=rhok, vpgrad=vpg, vsgrad=vsg, rhograd=rhg, dmod=dmod,
                dres=dres, wavgrad=wgrad)


def ref_fwi_e2d(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, ncheck, srcfield, scoords, wav, gcoords, dobs,
                misfit_type=0, weights=None):
    nz, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    nsrc, nrec = sc.shape[0], gc.shape[0]
    out = {k: np.zeros((nz, nx), np.float32) for k in ("vpgrad", "vsgrad", "rhograd")}
    dmod = np.zeros((3, nrec, nt), np.float32)
    dres = np.zeros((3, nrec, nt), np.float32)
    wg = np.zeros((nsrc, nt), np.float32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32).ravel()  # noqa: E731
    wts = _weights3(weights, nrec, nt, np.float32)  # kept alive across the call
    m = ref().ref_fwi_e2d(f32(vp), f32(vs), f32(rho), nx, nz, lpml, dx, dz, dt, nt, order, int(fs), ncheck,
                          {"P": 0, "VX": 1, "VZ": 3}[srcfield], nsrc, sc.ravel(), f32(wav), nrec, gc.ravel(),
                          f32(dobs), out["vpgrad"].reshape(-1), out["vsgrad"].reshape(-1),
                          out["rhograd"].reshape(-1), dmod.reshape(-1), dres.reshape(-1), wg.reshape(-1),
                          int(misfit_type), *[None if w is None else w.ctypes.data for w in wts])
    return dict(misfit=np.float32(m), dmod=dmod, dres=dres, wavgrad=wg, **out)


def fwi_a3d(vp, rho, lpml, d3, dt, nt, order, fs, ncheck, scoords, wav, gcoords, dobs, real=np.float32,
            misfit_type=0, weight=None):
    """FwiAcoustic3D::run_optimal (fwi.cpp:1991-2171): PRESSURE source and
    data, misfit over traces with all three map indices valid
    (computeMisfit, fwi.cpp:1468-1649) and residuals (computeResiduals,
    fwi.cpp:1651-1857) of misfit_type 0 DIFFERENCE, 1 CORRELATION,
    2 ADAPTIVE_GAUSS, 3 ADAPTIVE_LINEAR with optional data weights (the
    same arithmetic as the 2D driver's, so the same C functions), adjoint
    with the forward operator, crossCorr (fwi.cpp:1403-1466, quirks kept)
    and the wavelet gradient recorded at the source from the adjoint
    pressure."""
    nz, ny, nx = vp.shape
    smap, sshift = makemap3d(scoords, (nx, ny, nz), d3, "P")
    gmap, gshift = makemap3d(gcoords, (nx, ny, nz), d3, "P")
    vp = np.ascontiguousarray(vp, real)
    rho = np.ascontiguousarray(rho, real)
    L, Rx, Ry, Rz = stagger_a3d(vp, rho, lpml, fs, real)
    fw = A3DState(L, Rx, Ry, Rz, lpml, order, *d3, dt, real)
    bw = A3DState(L, Rx, Ry, Rz, lpml, order, *d3, dt, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    dobs = np.ascontiguousarray(dobs, real).reshape(gmap.shape[0], nt)
    dmod = np.zeros_like(dobs)
    dres = np.zeros_like(dobs)
    wgrad = np.zeros_like(wav)
    vpgrad = np.zeros((nz, ny, nx), real)
    rhograd = np.zeros((nz, ny, nx), real)
    valid = ((gmap[:, 0] >= 0) & (gmap[:, 1] >= 0) & (gmap[:, 2] >= 0)).astype(np.int32)
    wei = None if weight is None else np.ascontiguousarray(weight, real).reshape(gmap.shape[0], nt)
    sh32 = np.ascontiguousarray(sshift, real).reshape(-1)
    gh32 = np.ascontiguousarray(gshift, real).reshape(-1)

    def inject(s, m, sh, data, it):
        fn("inject3d", real)(s.P.reshape(-1), s.L.reshape(-1), s.nx, s.ny, lpml, real(dt), m.reshape(-1), sh,
                             data.reshape(-1), m.shape[0], nt, it)

    def record(s, m, sh, data, it):
        fn("record3d", real)(s.P.reshape(-1), s.nx, s.ny, lpml, m.reshape(-1), sh, data.reshape(-1), m.shape[0],
                             nt, it)

    xc = fn("a3d_crosscorr", real)

    def corr():
        xc(fw.P.reshape(-1), lpml, bw.Vx.reshape(-1), bw.Vy.reshape(-1), bw.Vz.reshape(-1), lpml,
           vp.reshape(-1), rho.reshape(-1), fw.Rx.reshape(-1), fw.Ry.reshape(-1), fw.Rz.reshape(-1), nx, ny, nz,
           real(d3[0]), real(d3[1]), real(d3[2]), vpgrad.reshape(-1), rhograd.reshape(-1))

    def fstep(it, rec):
        fw.velocity()
        fw.stress()
        inject(fw, smap, sh32, wav, it)
        if rec:
            record(fw, gmap, gh32, dmod, it)

    misfit = None
    checks = {}
    rev = orv.Revolve(nt, ncheck)
    reverse = False
    while True:
        oldcapo = rev.capo
        a = rev.step()
        capo = rev.capo
        if a == orv.ADVANCE:
            for it in range(oldcapo, capo):
                fstep(it, not reverse)
        elif a == orv.FIRSTURN:
            fstep(capo, True)
            wp = None if wei is None else wei.ctypes.data
            misfit = fn("misfit_a", real)(misfit_type, dmod.reshape(-1), dobs.reshape(-1), wp, valid,
                                          dobs.shape[0], nt)
            fn("residual_a", real)(misfit_type, dmod.reshape(-1), dobs.reshape(-1), wp, dres.reshape(-1),
                                   dobs.shape[0], nt, 1, 0)  # the 3D class has no percentile clip
            inject(bw, gmap, gh32, dres, capo)
            corr()
            record(bw, smap, sh32, wgrad, capo)
            reverse = True
        elif a == orv.YOUTURN:
            bw.velocity()
            bw.stress()
            inject(bw, gmap, gh32, dres, capo)
            corr()
            record(bw, smap, sh32, wgrad, capo)
        elif a == orv.TAKESHOT:
            checks[rev.check] = _snapshot(fw)
        elif a == orv.RESTORE:
            _restore(fw, checks[rev.check])
        if a in (orv.TERMINATE, orv.ERROR):
            break
    return dict(misfit=misfit, vpgrad=vpgrad, rhograd=rhograd, dmod=dmod, dres=dres, wavgrad=wgrad)


def ref_fwi_a3d(vp, rho, lpml, d3, dt, nt, order, fs, ncheck, scoords, wav, gcoords, dobs, misfit_type=0,
                weight=None):
    nz, ny, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 3)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 3)
    nrec = gc.shape[0]
    vg = np.zeros((nz, ny, nx), np.float32)
    rg = np.zeros((nz, ny, nx), np.float32)
    dmod = np.zeros((nrec, nt), np.float32)
    dres = np.zeros((nrec, nt), np.float32)
    wg = np.zeros((1, nt), np.float32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32).ravel()  # noqa: E731
    wts = None if weight is None else f32(weight)  # bound here: the pointer must outlive the call
    m = ref().ref_fwi_a3d(f32(vp), f32(rho), nx, ny, nz, lpml, *d3, dt, nt, order, int(fs), ncheck, sc.ravel(),
                          f32(wav), nrec, gc.ravel(), f32(dobs), vg.reshape(-1), rg.reshape(-1), dmod.reshape(-1),
                          dres.reshape(-1), wg.reshape(-1), int(misfit_type),
                          None if wts is None else wts.ctypes.data)
    return dict(misfit=np.float32(m), vpgrad=vg, rhograd=rg, dmod=dmod, dres=dres, wavgrad=wg)


_E3_REC = ("VX", "VY", "VZ")


def fwi_e3d(vp, vs, rho, lpml, d3, dt, nt, order, fs, ncheck, srcfield, scoords, wav, gcoords, dobs,
            real=np.float32, misfit_type=0, weights=None):
    """FwiElastic3D::run_optimal (fwi.cpp:6770-6991): forward steps with the
    source (PRESSURE explosive or VX / VY / VZ force) and the Ux, Uy, Uz
    records (each with the makeMap of its field); at the first turn the
    misfit (computeMisfit, fwi.cpp:5632-5899) and residuals
    (computeResiduals, fwi.cpp:5901-6196); the adjoint state of
    backwardstepVelocity / Stress with the residuals injected as VX, VY, VZ
    sources (insertSource in trace order); crossCorr (fwi.cpp:5046-5291) and
    the wavelet gradient of the source's field at every reverse step;
    scaleGrad (fwi.cpp:5553-5630) at the end.  dobs: [3][nrec][nt];
    weights: None or three arrays / Nones."""
    nz, ny, nx = vp.shape
    smap, sshift = makemap3d(scoords, (nx, ny, nz), d3, srcfield)
    gmaps = [makemap3d(gcoords, (nx, ny, nz), d3, f) for f in _E3_REC]
    fw = E3DState(stagger_e3d(vp, vs, rho, lpml, fs, real), lpml, order, d3, dt, fs, real)
    bw = E3DState(stagger_e3d(vp, vs, rho, lpml, fs, real), lpml, order, d3, dt, fs, real)
    nsrc, nrec = smap.shape[0], gmaps[0][0].shape[0]
    wav = np.ascontiguousarray(wav, real).reshape(nsrc, nt)
    dobs = np.ascontiguousarray(dobs, real).reshape(3, nrec, nt)
    dmod = np.zeros_like(dobs)
    dres = np.zeros_like(dobs)
    wgrad = np.zeros_like(wav)
    lam, mu, rg = (np.zeros((nz, ny, nx), real) for _ in range(3))
    wts = [None] * 3 if weights is None else [None if w is None else np.ascontiguousarray(w, real).reshape(nrec, nt)
                                               for w in weights]
    sh = np.ascontiguousarray(sshift, real).reshape(-1)
    gh = [np.ascontiguousarray(g[1], real).reshape(-1) for g in gmaps]
    model = [np.ascontiguousarray(a, real).ravel() for a in (vp, vs, rho)]

    def src_inject(s, it):
        if srcfield == "P":
            fn("inject3d_explosive", real)(s.Sxx.reshape(-1), s.Syy.reshape(-1), s.Szz.reshape(-1), s.nx, s.ny,
                                           lpml, real(dt), smap.reshape(-1), sh, wav.reshape(-1), nsrc, nt, it)
        else:
            F, M = {"VX": (s.Vx, s.Rx), "VY": (s.Vy, s.Ry), "VZ": (s.Vz, s.Rz)}[srcfield]
            fn("inject3d", real)(F.reshape(-1), M.reshape(-1), s.nx, s.ny, lpml, real(dt), smap.reshape(-1), sh,
                                 wav.reshape(-1), nsrc, nt, it)

    def record(it):
        for c, F in enumerate((fw.Vx, fw.Vy, fw.Vz)):
            fn("record3d", real)(F.reshape(-1), fw.nx, fw.ny, lpml, gmaps[c][0].reshape(-1), gh[c], dmod[c].reshape(-1),
                                 nrec, nt, it)

    def res_inject(it):
        for c, (F, M) in enumerate(((bw.Vx, bw.Rx), (bw.Vy, bw.Ry), (bw.Vz, bw.Rz))):
            fn("inject3d", real)(F.reshape(-1), M.reshape(-1), bw.nx, bw.ny, lpml, real(dt),
                                 gmaps[c][0].reshape(-1), gh[c], dres[c].reshape(-1), nrec, nt, it)

    xc = fn("e3d_crosscorr", real)
    nxp, nxyp = fw.nx, fw.nx * fw.ny

    def corr(it):
        xc(*(getattr(fw, k).reshape(-1) for k in ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")),
           *(getattr(bw, k).reshape(-1) for k in ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")),
           fw.Rx.reshape(-1), fw.Ry.reshape(-1), fw.Rz.reshape(-1), nx, ny, nz, lpml, real(d3[0]), real(d3[1]),
           real(d3[2]), lam.reshape(-1), mu.reshape(-1), rg.reshape(-1))
        # wavelet gradient (fwi.cpp:5143-5185): traces mapped to interior cells
        for i in range(nsrc):
            m = smap[i]
            if not (1 <= m[0] < nx - 1 and 1 <= m[1] < ny - 1 and 1 <= m[2] < nz - 1):
                continue
            c = ((m[2] + lpml) * fw.ny + m[1] + lpml) * nxp + m[0] + lpml
            if srcfield == "P":
                v = bw.Sxx.reshape(-1)[c] + bw.Syy.reshape(-1)[c] + bw.Szz.reshape(-1)[c]
            else:
                F, st = {"VX": (bw.Vx, 1), "VY": (bw.Vy, nxp), "VZ": (bw.Vz, nxyp)}[srcfield]
                v = real(0.5) * (F.reshape(-1)[c] + F.reshape(-1)[c - st])
            wgrad[i, it] = v

    def fstep(it, rec):
        fw.velocity()
        fw.stress()
        src_inject(fw, it)
        if rec:
            record(it)

    misfit = None
    checks = {}
    rev = orv.Revolve(nt, ncheck)
    reverse = False
    wp = [None if w is None else w.ctypes.data for w in wts]
    while True:
        oldcapo = rev.capo
        a = rev.step()
        capo = rev.capo
        if a == orv.ADVANCE:
            for it in range(oldcapo, capo):
                fstep(it, not reverse)
        elif a == orv.FIRSTURN:
            fstep(capo, True)
            args = [misfit_type] + [x for c in range(3) for x in (dmod[c].reshape(-1), dobs[c].reshape(-1))] + wp
            misfit = fn("misfit_e3", real)(*args, nrec, nt)
            fn("residual_e3", real)(*args, *(dres[c].reshape(-1) for c in range(3)), nrec, nt)
            res_inject(capo)
            corr(capo)
            reverse = True
        elif a == orv.YOUTURN:
            bw.velocity_bw()
            bw.stress_bw()
            res_inject(capo)
            corr(capo)
        elif a == orv.TAKESHOT:
            checks[rev.check] = _snapshot(fw)
        elif a == orv.RESTORE:
            _restore(fw, checks[rev.check])
        if a in (orv.TERMINATE, orv.ERROR):
            break
    fn("e3d_scalegrad", real)(*model, nx * ny * nz, lam.reshape(-1), mu.reshape(-1), rg.reshape(-1))
    return dict(misfit=misfit, vpgrad=lam, vsgrad=mu, rhograd=rg, dmod=dmod, dres=dres, wavgrad=wgrad)


def ref_fwi_e3d(vp, vs, rho, lpml, d3, dt, nt, order, fs, ncheck, srcfield, scoords, wav, gcoords, dobs,
                misfit_type=0, weights=None):
    """FwiElastic3D::run_optimal of the compiled reference, the same outputs."""
    nz, ny, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 3)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 3)
    nsrc, nrec = sc.shape[0], gc.shape[0]
    g = [np.zeros((nz, ny, nx), np.float32) for _ in range(3)]
    dmod = np.zeros((3, nrec, nt), np.float32)
    dres = np.zeros((3, nrec, nt), np.float32)
    wg = np.zeros((nsrc, nt), np.float32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32).ravel()  # noqa: E731
    wts = [None] * 3 if weights is None else [None if w is None else f32(w) for w in weights]
    from .runs import _FIELD_CODE
    m = ref().ref_fwi_e3d(f32(vp), f32(vs), f32(rho), nx, ny, nz, lpml, *(np.float32(v) for v in d3),
                          np.float32(dt), nt, order, int(fs), ncheck, _FIELD_CODE[srcfield], nsrc, sc.ravel(),
                          f32(wav), nrec, gc.ravel(), f32(dobs), *(a.reshape(-1) for a in g), dmod.reshape(-1),
                          dres.reshape(-1), wg.reshape(-1), int(misfit_type),
                          *(None if w is None else w.ctypes.data for w in wts))
    return dict(misfit=np.float32(m), vpgrad=g[0], vsgrad=g[1], rhograd=g[2], dmod=dmod, dres=dres, wavgrad=wg)

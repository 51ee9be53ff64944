"""Oracle of the exact-gradient mode of the acoustic 2D FWI gradient: the
exact discrete adjoint of the modelling time loop (modelling.cpp:242-286:
velocity step, stress step, insertSource(it), recordData(it)) and of the
model staggering (model.cpp:1350-1411), so that the gradient is that of the
DIFFERENCE misfit (fwi.cpp:574-608) of the discrete forward map itself.

The reference has no such mode: FwiAcoustic2D::run_optimal advances its
adjoint wavefield with the forward operator and images with crossCorr
(fwi.cpp:511-571), which is not the transpose of the discrete step.  These
functions are pinned by what the mathematics fixes instead: the adjoint
dot-product test <F w, r> = <w, F^T r> (source wavelet to records) and
central finite differences of the FP64 misfit.

Every forward state is kept in memory (small grids only).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import STRUCTS, fn
from .runs import A2DState, _p, inject2d, makemap2d, record2d, stagger_a2d


class _Adj:
    """Adjoint state: the fields and strips of an A2DState (no model)."""

    def __init__(self, like: A2DState):
        real = like.real
        self.real = real
        nzp, nxp = like.P.shape
        self.P, self.Vx, self.Vz, self.df = (np.zeros((nzp, nxp), real) for _ in range(4))
        self.strips = {k: np.zeros_like(v) for k, v in like.strips.items()}
        ct = C.c_float if real == np.float32 else C.c_double
        st = STRUCTS[real][0]()
        st.nx, st.nz, st.lpml, st.order = like.st.nx, like.st.nz, like.st.lpml, like.st.order
        st.dx, st.dz, st.dt = like.st.dx, like.st.dz, like.st.dt
        for k in ("P", "Vx", "Vz", "df"):
            setattr(st, k, _p(getattr(self, k), ct))
        for k, a in self.strips.items():
            setattr(st, k, _p(a, ct))
        self.st = st


def _bilinear(mp, sh, nrm):
    """(cells, weights) of each trace's 2x2 corners, None for maps < 0."""
    out = []
    for i in range(mp.shape[0]):
        mx, mz = int(mp[i, 0]), int(mp[i, 1])
        if mx < 0 or mz < 0:
            out.append(None)
            continue
        sx, sz = float(sh[i, 0]), float(sh[i, 1])
        xt, zt = (1 - sx, sx), (1 - sz, sz)
        out.append([(nrm + mz + i1, nrm + mx + i2, xt[i2] * zt[i1]) for i1 in range(2) for i2 in range(2)])
    return out


def a2d_model_grad(gL, gRx, gRz, vp, rho, lpml, fs):
    """Chain rule of ModelAcoustic2D::staggerModels (L = rho*vp^2, Rx/Rz =
    1/mean of the edge-padded density along x/z, last plane copied; free
    surface: L = 0 and Rx doubled on row lpml), in float64: gradients of
    (L, Rx, Rz) on the padded grid -> gradients of (vp, rho)."""
    nz, nx = vp.shape
    vpp = np.pad(np.asarray(vp, np.float64), lpml, mode="edge")
    rp = np.pad(np.asarray(rho, np.float64), lpml, mode="edge")
    gL, gRx, gRz = (np.asarray(a, np.float64).copy() for a in (gL, gRx, gRz))
    sfac = np.ones_like(vpp)
    if fs:
        gL[lpml, :] = 0.0  # L is forced to 0 there
        sfac[lpml, :] = 2.0
    gvp = 2.0 * gL * rp * vpp
    grho = gL * vpp ** 2
    for g, ax in ((gRx, 1), (gRz, 0)):
        a = 0.5 * (rp + np.roll(rp, -1, axis=ax))
        if ax == 1:
            a[:, -1] = a[:, -2]
        else:
            a[-1, :] = a[-2, :]
        ga = -g * (sfac if ax == 1 else 1.0) / a ** 2  # the free surface doubles Rx only
        if ax == 1:
            ga[:, -2] += ga[:, -1]
            ga[:, -1] = 0.0
            grho[:, :-1] += 0.5 * ga[:, :-1]
            grho[:, 1:] += 0.5 * ga[:, :-1]
        else:
            ga[-2, :] += ga[-1, :]
            ga[-1, :] = 0.0
            grho[:-1, :] += 0.5 * ga[:-1, :]
            grho[1:, :] += 0.5 * ga[:-1, :]

    def unpad(g):
        out = np.zeros((nz, nx))
        iz = np.clip(np.arange(g.shape[0]) - lpml, 0, nz - 1)
        ix = np.clip(np.arange(g.shape[1]) - lpml, 0, nx - 1)
        np.add.at(out, (iz[:, None], ix[None, :]), g)
        return out

    return unpad(gvp), unpad(grho)


def _forward(vp, rho, lpml, dx, dz, dt, nt, order, fs, scoords, wav, gcoords, real, keep):
    nz, nx = vp.shape
    smap, sshift = makemap2d(scoords, nx, nz, dx, dz, "P")
    gmap, gshift = makemap2d(gcoords, nx, nz, dx, dz, "P")
    L, Rx, Rz = stagger_a2d(np.asarray(vp, real), np.asarray(rho, real), lpml, fs, real)
    s = A2DState(L, Rx, Rz, lpml, order, dx, dz, dt, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    dmod = np.zeros((gmap.shape[0], nt), real)
    states = []
    for it in range(nt):
        if keep:
            states.append([a.copy() for a in [s.P, s.Vx, s.Vz] + list(s.strips.values())])
        s.velocity()
        s.stress()
        inject2d(s.P, s.L, lpml, dt, smap, sshift, wav, it, real)
        record2d(s.P, lpml, gmap, gshift, dmod, it, real)
    return s, states, dmod, (smap, sshift, gmap, gshift, wav)


def _misfit_res(dmod, dobs, gmap, weight, real):
    nt = dmod.shape[1]
    dobs = np.ascontiguousarray(dobs, real).reshape(dmod.shape)
    valid = ((gmap[:, 0] >= 0) & (gmap[:, 1] >= 0)).astype(np.int32)
    wei = None if weight is None else np.ascontiguousarray(weight, real).reshape(dmod.shape)
    wp = None if wei is None else wei.ctypes.data
    dres = np.zeros_like(dmod)
    mis = fn("misfit_a", real)(0, dmod.reshape(-1), dobs.reshape(-1), wp, valid, dmod.shape[0], nt)
    fn("residual_a", real)(0, dmod.reshape(-1), dobs.reshape(-1), wp, dres.reshape(-1), dmod.shape[0], nt, 1, 1)
    return float(mis), dres


def a2d_misfit(vp, rho, lpml, dx, dz, dt, nt, order, fs, scoords, wav, gcoords, dobs, real=np.float64,
               weight=None):
    """DIFFERENCE misfit of one shot after the forward run alone."""
    _, _, dmod, (_, _, gmap, _, _) = _forward(vp, rho, lpml, dx, dz, dt, nt, order, fs, scoords, wav, gcoords,
                                              real, False)
    return _misfit_res(dmod, dobs, gmap, weight, real)[0]


def fwi_a2d_exact(vp, rho, lpml, dx, dz, dt, nt, order, fs, scoords, wav, gcoords, dobs, real=np.float64,
                  weight=None):
    """Exact gradient of the DIFFERENCE misfit of one shot.  Returns
    dict(misfit, vpgrad, rhograd, dmod, dres, wavgrad, gL, gRx, gRz);
    wavgrad = dJ/d(wavelet), gL/gRx/gRz = dJ/d(L, Rx, Rz) on the padded grid."""
    s, states, dmod, (smap, sshift, gmap, gshift, wav) = _forward(vp, rho, lpml, dx, dz, dt, nt, order, fs,
                                                                  scoords, wav, gcoords, real, True)
    misfit, dres = _misfit_res(dmod, dobs, gmap, weight, real)
    a = _Adj(s)
    gL, gRx, gRz = (np.zeros_like(s.P) for _ in range(3))
    wavgrad = np.zeros_like(wav)
    rec = _bilinear(gmap, gshift, lpml)
    src = _bilinear(smap, sshift, lpml)
    dtr = real(dt)
    for it in range(nt - 1, -1, -1):
        for dst, v in zip([s.P, s.Vx, s.Vz] + list(s.strips.values()), states[it]):
            dst[...] = v
        s.velocity()  # V of step it, P strips after the velocity half step
        for i, cells in enumerate(rec):  # adjoint of recordData
            if cells:
                for (z, x, w) in cells:
                    a.P[z, x] += w * dres[i, it]
        for i, cells in enumerate(src):  # insertSource: P += dt*L*wav*w
            if cells:
                for (z, x, w) in cells:
                    wavgrad[i, it] += dtr * s.L[z, x] * w * a.P[z, x]
                    gL[z, x] += dtr * wav[i, it] * w * a.P[z, x]
        fn("a2d_adj_stress", real)(C.byref(s.st), C.byref(a.st), gL.reshape(-1))
        fn("a2d_adj_velocity", real)(C.byref(s.st), C.byref(a.st), gRx.reshape(-1), gRz.reshape(-1))
    gvp, grho = a2d_model_grad(gL, gRx, gRz, vp, rho, lpml, fs)
    return dict(misfit=misfit, vpgrad=gvp, rhograd=grho, dmod=dmod, dres=dres, wavgrad=wavgrad, gL=gL, gRx=gRx,
                gRz=gRz)


# --------------------------------------------------------------------------
# Elastic 2D: exact adjoint of the modelling time loop (modelling.cpp:
# velocity step, stress step, insertSource, recordData; WavesElastic2D
# forwardstepVelocity / forwardstepStress, waves.cpp:1416-1735) and of
# ModelElastic2D::staggerModels (model.cpp:2720-2801); the gradient of the
# DIFFERENCE misfit of FwiElastic2D::computeMisfit (fwi.cpp:2852-2904) over
# the P, Ux and Uz records.  The reference's FwiElastic2D back-propagates
# with the adjoint-state steps (waves.cpp:1738-2080) and images with
# crossCorr; this restatement is pinned by the dot-product test and by
# finite differences of the FP64 misfit instead.
# --------------------------------------------------------------------------

def _stagger_t(g, axis):
    """Transpose of the arithmetic-mean staggering along `axis` (out[i] =
    (in[i] + in[i+1]) / 2, last plane copying its neighbour)."""
    g = np.moveaxis(np.asarray(g, np.float64).copy(), axis, 0)
    g[-2] += g[-1]
    g[-1] = 0.0
    out = np.zeros_like(g)
    out[:-1] += 0.5 * g[:-1]
    out[1:] += 0.5 * g[:-1]
    return np.moveaxis(out, 0, axis)


def _unpad_t(g, nz, nx, lpml):
    """Transpose of the edge padding: each model cell sums its replicas."""
    out = np.zeros((nz, nx))
    iz = np.clip(np.arange(g.shape[0]) - lpml, 0, nz - 1)
    ix = np.clip(np.arange(g.shape[1]) - lpml, 0, nx - 1)
    np.add.at(out, (iz[:, None], ix[None, :]), g)
    return out


def e2d_model_grad(gL, gL2M, gM, gRx, gRz, vp, vs, rho, lpml, fs, lame=False):
    """Chain rule of ModelElastic2D::staggerModels in float64: L2M = rho vp^2,
    M0 = rho vs^2, L = L2M - 2 M0 on the edge-padded grid (free surface: L = 0
    and L2M = M0 on row lpml), M = M0 staggered along z then x, Rx / Rz =
    1 / mean of the density along x / z (Rx doubled on the free-surface row).
    Gradients of (L, L2M, M, Rx, Rz) on the padded grid -> (vp, vs, rho), or
    with lame -> (lambda, mu, rho) as independent parameters (lambda = L,
    mu = M0, L2M = lambda + 2 mu)."""
    nz, nx = vp.shape
    v = np.pad(np.asarray(vp, np.float64), lpml, mode="edge")
    w = np.pad(np.asarray(vs, np.float64), lpml, mode="edge")
    r = np.pad(np.asarray(rho, np.float64), lpml, mode="edge")
    gL, gL2M = (np.asarray(a, np.float64).copy() for a in (gL, gL2M))
    gM0 = _stagger_t(_stagger_t(gM, 1), 0)  # M = stagger_x(stagger_z(M0))
    if fs:
        gM0[lpml, :] += gL2M[lpml, :]  # L2M = M0 there
        gL[lpml, :] = 0.0
        gL2M[lpml, :] = 0.0
    if lame:
        gv = gL + gL2M
        gw = 2.0 * gL2M + gM0
        gr = np.zeros_like(r)
    else:
        gv = 2.0 * r * v * (gL2M + gL)
        gw = 2.0 * r * w * (gM0 - 2.0 * gL)
        gr = v * v * gL2M + (v * v - 2.0 * w * w) * gL + w * w * gM0
    sfac = np.ones_like(r)
    if fs:
        sfac[lpml, :] = 2.0
    for g, ax in ((gRx, 1), (gRz, 0)):
        a = 0.5 * (r + np.roll(r, -1, axis=ax))
        if ax == 1:
            a[:, -1] = a[:, -2]
        else:
            a[-1, :] = a[-2, :]
        ga = -np.asarray(g, np.float64) * (sfac if ax == 1 else 1.0) / a ** 2
        gr += _stagger_t(ga, ax)
    return _unpad_t(gv, nz, nx, lpml), _unpad_t(gw, nz, nx, lpml), _unpad_t(gr, nz, nx, lpml)


class _AdjE:
    """Adjoint state of an E2DState: its fields and strips (no model)."""

    def __init__(self, like):
        from .runs import E2DState
        real = like.real
        self.real = real
        nzp, nxp = like.Sxx.shape
        for k in E2DState.FIELDS + ("df", "wrk"):
            setattr(self, k, np.zeros((nzp, nxp), real))
        self.strips = {k: np.zeros_like(v) for k, v in like.strips.items()}
        ct = C.c_float if real == np.float32 else C.c_double
        st = STRUCTS[real][1]()
        for k in ("nx", "nz", "lpml", "order", "fs", "dx", "dz", "dt"):
            setattr(st, k, getattr(like.st, k))
        for k in E2DState.FIELDS + ("df", "wrk"):
            setattr(st, k, _p(getattr(self, k), ct))
        for k, a in self.strips.items():
            setattr(st, k, _p(a, ct))
        self.st = st


_E2D_FIELDS = ("P", "VX", "VZ")


def _forward_e2d(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, srcfield, scoords, wav, gcoords, real, keep):
    from .runs import E2DState, e2d_record_factor, stagger_e2d
    nz, nx = vp.shape
    smap, sshift = makemap2d(scoords, nx, nz, dx, dz, srcfield)
    maps = {f: makemap2d(gcoords, nx, nz, dx, dz, f) for f in _E2D_FIELDS}
    L, L2M, M, Rx, Rz = stagger_e2d(np.asarray(vp, real), np.asarray(vs, real), np.asarray(rho, real), lpml,
                                    fs, real)
    s = E2DState(L, L2M, M, Rx, Rz, lpml, order, dx, dz, dt, fs, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    ntr = maps["VX"][0].shape[0]
    dmod = {f: np.zeros((ntr, nt), real) for f in _E2D_FIELDS}
    factor = e2d_record_factor(vp, vs, rho, maps["P"][0], False, real)
    states = []
    for it in range(nt):
        if keep:
            states.append([getattr(s, k).copy() for k in E2DState.FIELDS] + [a.copy() for a in s.strips.values()])
        s.velocity()
        s.stress()
        if srcfield == "P":
            inject2d(s.Sxx, None, lpml, dt, smap, sshift, wav, it, real)
            inject2d(s.Szz, None, lpml, dt, smap, sshift, wav, it, real)
        elif srcfield == "VX":
            inject2d(s.Vx, s.Rx, lpml, dt, smap, sshift, wav, it, real)
        else:
            inject2d(s.Vz, s.Rz, lpml, dt, smap, sshift, wav, it, real)
        gm, gs = maps["P"]
        fn("record2d_pressure_el", real)(s.Sxx.reshape(-1), s.Szz.reshape(-1), s.nx, lpml, gm.reshape(-1),
                                         np.ascontiguousarray(gs, real).reshape(-1), factor,
                                         dmod["P"].reshape(-1), ntr, nt, it)
        record2d(s.Vx, lpml, *maps["VX"], dmod["VX"], it, real)
        record2d(s.Vz, lpml, *maps["VZ"], dmod["VZ"], it, real)
    return s, states, dmod, (smap, sshift, maps, wav, factor)


def _misfit_res_e(dmod, dobs, maps, weight, real):
    nt = dmod["VX"].shape[1]
    valid = ((maps["VX"][0][:, 0] >= 0) & (maps["VX"][0][:, 1] >= 0)).astype(np.int32)
    obs = {f: np.ascontiguousarray(dobs[f], real).reshape(dmod[f].shape) for f in _E2D_FIELDS}
    wei = {f: None if weight is None or weight.get(f) is None else
           np.ascontiguousarray(weight[f], real).reshape(dmod[f].shape) for f in _E2D_FIELDS}
    wp = [None if wei[f] is None else wei[f].ctypes.data for f in _E2D_FIELDS]
    res = {f: np.zeros_like(dmod[f]) for f in _E2D_FIELDS}
    args = []
    for f in _E2D_FIELDS:
        args += [dmod[f].reshape(-1), obs[f].reshape(-1)]
    mis = fn("misfit_e", real)(0, *args, *wp, valid, valid.size, nt)
    fn("residual_e", real)(0, *args, *wp, *(res[f].reshape(-1) for f in _E2D_FIELDS), valid.size, nt)
    return float(mis), res


def e2d_misfit(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, srcfield, scoords, wav, gcoords, dobs,
               real=np.float64, weight=None):
    """DIFFERENCE misfit of one elastic shot (P, Ux, Uz records) after the
    forward run alone; dobs / weight: dicts over "P", "VX", "VZ"."""
    _, _, dmod, (_, _, maps, _, _) = _forward_e2d(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, srcfield,
                                                  scoords, wav, gcoords, real, False)
    return _misfit_res_e(dmod, dobs, maps, weight, real)[0]


def fwi_e2d_exact(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, srcfield, scoords, wav, gcoords, dobs,
                  real=np.float64, weight=None):
    """Exact gradient of the DIFFERENCE misfit of one elastic shot.  Returns
    dict(misfit, vpgrad, vsgrad, rhograd, lamgrad, mugrad, rhograd_lame, dmod,
    dres, wavgrad, gL, gL2M, gM, gRx, gRz) with the moduli / buoyancy
    gradients on the padded grid; lamgrad / mugrad / rhograd_lame are the
    gradients with lambda, mu and rho as the independent parameters."""
    from .runs import E2DState
    s, states, dmod, (smap, sshift, maps, wav, factor) = _forward_e2d(
        vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, srcfield, scoords, wav, gcoords, real, True)
    misfit, dres = _misfit_res_e(dmod, dobs, maps, weight, real)
    a = _AdjE(s)
    gL, gL2M, gM, gRx, gRz = (np.zeros_like(s.Sxx) for _ in range(5))
    wavgrad = np.zeros_like(wav)
    rec = {f: _bilinear(maps[f][0], maps[f][1], lpml) for f in _E2D_FIELDS}
    src = _bilinear(smap, sshift, lpml)
    dtr = real(dt)
    names = list(E2DState.FIELDS) + list(s.strips.keys())
    for it in range(nt - 1, -1, -1):
        for k, v in zip(names, states[it]):
            (s.strips[k] if k in s.strips else getattr(s, k))[...] = v
        s.velocity()
        for i, cells in enumerate(rec["P"]):  # adjoint of the pressure gather (Sxx + Szz) / 2 / factor
            if cells:
                for (z, x, w) in cells:
                    c = 0.5 * w * dres["P"][i, it] / factor[i]
                    a.Sxx[z, x] += c
                    a.Szz[z, x] += c
        for f, fld in (("VX", a.Vx), ("VZ", a.Vz)):
            for i, cells in enumerate(rec[f]):
                if cells:
                    for (z, x, w) in cells:
                        fld[z, x] += w * dres[f][i, it]
        for i, cells in enumerate(src):  # insertSource
            if not cells:
                continue
            for (z, x, w) in cells:
                if srcfield == "P":  # Sxx, Szz -= dt wav w
                    wavgrad[i, it] -= dtr * w * (a.Sxx[z, x] + a.Szz[z, x])
                elif srcfield == "VX":  # Vx += dt Rx wav w
                    wavgrad[i, it] += dtr * s.Rx[z, x] * w * a.Vx[z, x]
                    gRx[z, x] += dtr * wav[i, it] * w * a.Vx[z, x]
                else:
                    wavgrad[i, it] += dtr * s.Rz[z, x] * w * a.Vz[z, x]
                    gRz[z, x] += dtr * wav[i, it] * w * a.Vz[z, x]
        fn("e2d_adj_stress", real)(C.byref(s.st), C.byref(a.st), gL.reshape(-1), gL2M.reshape(-1),
                                   gM.reshape(-1))
        fn("e2d_adj_velocity", real)(C.byref(s.st), C.byref(a.st), gRx.reshape(-1), gRz.reshape(-1))
    gvp, gvs, grho = e2d_model_grad(gL, gL2M, gM, gRx, gRz, vp, vs, rho, lpml, fs)
    glam, gmu, grl = e2d_model_grad(gL, gL2M, gM, gRx, gRz, vp, vs, rho, lpml, fs, lame=True)
    return dict(misfit=misfit, vpgrad=gvp, vsgrad=gvs, rhograd=grho, lamgrad=glam, mugrad=gmu, rhograd_lame=grl,
                dmod=dmod, dres=dres, wavgrad=wavgrad, gL=gL, gL2M=gL2M, gM=gM, gRx=gRx, gRz=gRz)


# --------------------------------------------------------------------------
# Acoustic 3D: exact adjoint of the modelling time loop (WavesAcoustic3D
# forwardstepVelocity / forwardstepStress, waves.cpp:710-996; insertSource /
# recordData, waves.cpp:998-1275) and of ModelAcoustic3D::staggerModels
# (model.cpp:1965-2040); the gradient of the DIFFERENCE misfit of
# FwiAcoustic3D::computeMisfit (fwi.cpp:1468-1649).  The reference
# back-propagates with the forward operator and images with crossCorr
# (fwi.cpp:1403-1466); this restatement is pinned by the dot-product test and
# finite differences of the FP64 misfit.
# --------------------------------------------------------------------------

def a3d_model_grad(gL, gRx, gRy, gRz, vp, rho, lpml, fs):
    """Chain rule of ModelAcoustic3D::staggerModels in float64: L = rho vp^2
    on the edge-padded grid (0 on the free-surface plane z = lpml), Rx / Ry /
    Rz = 1 / mean of the density along x / y / z (Rx and Ry doubled on the
    free-surface plane).  Gradients on the padded [z][y][x] grid -> (vp,
    rho) of the [nz][ny][nx] model."""
    nz, ny, nx = vp.shape
    v = np.pad(np.asarray(vp, np.float64), lpml, mode="edge")
    r = np.pad(np.asarray(rho, np.float64), lpml, mode="edge")
    gL = np.asarray(gL, np.float64).copy()
    sfac = np.ones_like(r)
    if fs:
        gL[lpml] = 0.0
        sfac[lpml] = 2.0
    gv = 2.0 * gL * r * v
    gr = gL * v * v
    for g, ax, s in ((gRx, 2, sfac), (gRy, 1, sfac), (gRz, 0, 1.0)):
        a = 0.5 * (r + np.roll(r, -1, axis=ax))
        a = np.moveaxis(a, ax, 0)
        a[-1] = a[-2]
        a = np.moveaxis(a, 0, ax)
        gr += _stagger_t(-np.asarray(g, np.float64) * s / a ** 2, ax)

    def unpad(g):
        out = np.zeros((nz, ny, nx))
        iz = np.clip(np.arange(g.shape[0]) - lpml, 0, nz - 1)
        iy = np.clip(np.arange(g.shape[1]) - lpml, 0, ny - 1)
        ix = np.clip(np.arange(g.shape[2]) - lpml, 0, nx - 1)
        np.add.at(out, (iz[:, None, None], iy[None, :, None], ix[None, None, :]), g)
        return out

    return unpad(gv), unpad(gr)


class _AdjA3:
    """Adjoint state of an A3DState: fields and strips (no model)."""

    def __init__(self, like):
        real = like.real
        self.real = real
        for k in ("P", "Vx", "Vy", "Vz", "df"):
            setattr(self, k, np.zeros_like(like.P))
        self.strips = {k: np.zeros_like(v) for k, v in like.strips.items()}
        ct = C.c_float if real == np.float32 else C.c_double
        st = STRUCTS[real][2]()
        for k in ("nx", "ny", "nz", "lpml", "order", "dx", "dy", "dz", "dt"):
            setattr(st, k, getattr(like.st, k))
        for k in ("P", "Vx", "Vy", "Vz", "df"):
            setattr(st, k, _p(getattr(self, k), ct))
        for k, a in self.strips.items():
            setattr(st, k, _p(a, ct))
        self.st = st


def _trilinear(mp, sh, lpml):
    out = []
    for i in range(mp.shape[0]):
        if (mp[i] < 0).any():
            out.append(None)
            continue
        xt = (1 - float(sh[i, 0]), float(sh[i, 0]))
        yt = (1 - float(sh[i, 1]), float(sh[i, 1]))
        zt = (1 - float(sh[i, 2]), float(sh[i, 2]))
        out.append([(lpml + int(mp[i, 2]) + i1, lpml + int(mp[i, 1]) + i2, lpml + int(mp[i, 0]) + i3,
                     xt[i3] * yt[i2] * zt[i1]) for i1 in range(2) for i2 in range(2) for i3 in range(2)])
    return out


def _forward_a3d(vp, rho, lpml, d3, dt, nt, order, fs, scoords, wav, gcoords, real, keep):
    from .runs import A3DState, makemap3d, stagger_a3d
    nz, ny, nx = vp.shape
    smap, sshift = makemap3d(scoords, (nx, ny, nz), d3, "P")
    gmap, gshift = makemap3d(gcoords, (nx, ny, nz), d3, "P")
    L, Rx, Ry, Rz = stagger_a3d(np.asarray(vp, real), np.asarray(rho, real), lpml, fs, real)
    s = A3DState(L, Rx, Ry, Rz, lpml, order, *d3, dt, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    dmod = np.zeros((gmap.shape[0], nt), real)
    sh32 = np.ascontiguousarray(sshift, real).reshape(-1)
    gh32 = np.ascontiguousarray(gshift, real).reshape(-1)
    states = []
    for it in range(nt):
        if keep:
            states.append([a.copy() for a in [s.P, s.Vx, s.Vy, s.Vz] + list(s.strips.values())])
        s.velocity()
        s.stress()
        fn("inject3d", real)(s.P.reshape(-1), s.L.reshape(-1), s.nx, s.ny, lpml, real(dt), smap.reshape(-1), sh32,
                             wav.reshape(-1), smap.shape[0], nt, it)
        fn("record3d", real)(s.P.reshape(-1), s.nx, s.ny, lpml, gmap.reshape(-1), gh32, dmod.reshape(-1),
                             gmap.shape[0], nt, it)
    return s, states, dmod, (smap, sshift, gmap, gshift, wav)


def _misfit_res3(dmod, dobs, gmap, weight, real):
    nt = dmod.shape[1]
    dobs = np.ascontiguousarray(dobs, real).reshape(dmod.shape)
    valid = (gmap >= 0).all(axis=1).astype(np.int32)
    wei = None if weight is None else np.ascontiguousarray(weight, real).reshape(dmod.shape)
    wp = None if wei is None else wei.ctypes.data
    dres = np.zeros_like(dmod)
    mis = fn("misfit_a", real)(0, dmod.reshape(-1), dobs.reshape(-1), wp, valid, dmod.shape[0], nt)
    fn("residual_a", real)(0, dmod.reshape(-1), dobs.reshape(-1), wp, dres.reshape(-1), dmod.shape[0], nt, 1, 0)
    return float(mis), dres


def a3d_misfit(vp, rho, lpml, d3, dt, nt, order, fs, scoords, wav, gcoords, dobs, real=np.float64, weight=None):
    """DIFFERENCE misfit of one 3D shot after the forward run alone."""
    _, _, dmod, (_, _, gmap, _, _) = _forward_a3d(vp, rho, lpml, d3, dt, nt, order, fs, scoords, wav, gcoords,
                                                  real, False)
    return _misfit_res3(dmod, dobs, gmap, weight, real)[0]


def fwi_a3d_exact(vp, rho, lpml, d3, dt, nt, order, fs, scoords, wav, gcoords, dobs, real=np.float64,
                  weight=None):
    """Exact gradient of the DIFFERENCE misfit of one acoustic 3D shot.
    Returns dict(misfit, vpgrad, rhograd, dmod, dres, wavgrad, gL, gRx, gRy,
    gRz)."""
    s, states, dmod, (smap, sshift, gmap, gshift, wav) = _forward_a3d(vp, rho, lpml, d3, dt, nt, order, fs,
                                                                      scoords, wav, gcoords, real, True)
    misfit, dres = _misfit_res3(dmod, dobs, gmap, weight, real)
    a = _AdjA3(s)
    gL, gRx, gRy, gRz = (np.zeros_like(s.P) for _ in range(4))
    wavgrad = np.zeros_like(wav)
    rec = _trilinear(gmap, gshift, s.lpml)
    src = _trilinear(smap, sshift, s.lpml)
    dtr = real(dt)
    for it in range(nt - 1, -1, -1):
        for dst, v in zip([s.P, s.Vx, s.Vy, s.Vz] + list(s.strips.values()), states[it]):
            dst[...] = v
        s.velocity()
        for i, cells in enumerate(rec):
            if cells:
                for (z, y, x, w) in cells:
                    a.P[z, y, x] += w * dres[i, it]
        for i, cells in enumerate(src):  # insertSource: P += dt L wav w
            if cells:
                for (z, y, x, w) in cells:
                    wavgrad[i, it] += dtr * s.L[z, y, x] * w * a.P[z, y, x]
                    gL[z, y, x] += dtr * wav[i, it] * w * a.P[z, y, x]
        fn("a3d_adj_stress", real)(C.byref(s.st), C.byref(a.st), gL.reshape(-1))
        fn("a3d_adj_velocity", real)(C.byref(s.st), C.byref(a.st), gRx.reshape(-1), gRy.reshape(-1),
                                     gRz.reshape(-1))
    gvp, grho = a3d_model_grad(gL, gRx, gRy, gRz, vp, rho, s.lpml, fs)
    return dict(misfit=misfit, vpgrad=gvp, rhograd=grho, dmod=dmod, dres=dres, wavgrad=wavgrad, gL=gL, gRx=gRx,
                gRy=gRy, gRz=gRz)

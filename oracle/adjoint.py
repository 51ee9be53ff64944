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

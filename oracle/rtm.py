"""Oracle restatement of RtmAcoustic2D::run_optimal (rtm.cpp:423-544) and
RtmAcoustic3D::run_optimal (rtm.cpp:735-): the
source wavefield on the Revolve schedule, the receiver wavefield advanced
with the forward operator from the recorded data (insertSource of dataP at
GMAP), and the imaging condition with subsurface offsets
RtmAcoustic2D::crossCorr (rtm.cpp:191-219) at every reverse step.  Also the
same run on the compiled reference.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import numpy as np

from . import fn, ref
from . import revolve as orv
from .fwi import _restore, _snapshot
from .runs import A2DState, A3DState, inject2d, makemap2d, makemap3d, stagger_a2d, stagger_a3d


def image_a2d(image, ws, wr, lpml, nhx, nhz):
    """crossCorr (rtm.cpp:191-219): image[ihz, ihx, iz, ix] += ws(ix-hx,
    iz-hz) * wr(ix+hx, iz+hz), hx = -(nhx-1)//2 + ihx (C division of a
    non-negative number), both points inside the unpadded grid; ws and wr
    padded by lpml.  Each cell's product and sum are rounded in the image's
    type, one term per call as in the reference's loop."""
    nz, nx = image.shape[2:]
    for ihx in range(nhx):
        hx = -((nhx - 1) // 2) + ihx
        for ihz in range(nhz):
            hz = -((nhz - 1) // 2) + ihz
            x0, x1 = max(0, hx, -hx), min(nx, nx - hx, nx + hx)
            z0, z1 = max(0, hz, -hz), min(nz, nz - hz, nz + hz)
            if x0 >= x1 or z0 >= z1:
                continue
            s = ws[z0 - hz + lpml:z1 - hz + lpml, x0 - hx + lpml:x1 - hx + lpml]
            r = wr[z0 + hz + lpml:z1 + hz + lpml, x0 + hx + lpml:x1 + hx + lpml]
            image[ihz, ihx, z0:z1, x0:x1] += s * r


def rtm_a2d(vp, rho, lpml, dx, dz, dt, nt, order, fs, ncheck, scoords, wav, gcoords, data, nhx, nhz,
            real=np.float32):
    """Returns the image [nhz][nhx][nz][nx] of one shot."""
    nz, nx = vp.shape
    smap, sshift = makemap2d(scoords, nx, nz, dx, dz, "P")
    gmap, gshift = makemap2d(gcoords, nx, nz, dx, dz, "P")
    vp = np.ascontiguousarray(vp, real)
    rho = np.ascontiguousarray(rho, real)
    L, Rx, Rz = stagger_a2d(vp, rho, lpml, fs, real)
    fw = A2DState(L, Rx, Rz, lpml, order, dx, dz, dt, real)
    bw = A2DState(L, Rx, Rz, lpml, order, dx, dz, dt, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    data = np.ascontiguousarray(data, real).reshape(gmap.shape[0], nt)
    image = np.zeros((nhz, nhx, nz, nx), real)
    checks = {}
    rev = orv.Revolve(nt, ncheck)

    def fstep(it):
        fw.velocity()
        fw.stress()
        inject2d(fw.P, fw.L, lpml, dt, smap, sshift, wav, it, real)

    def image_step(capo):
        inject2d(bw.P, bw.L, lpml, dt, gmap, gshift, data, capo, real)
        image_a2d(image, fw.P, bw.P, lpml, nhx, nhz)

    while True:
        oldcapo = rev.capo
        a = rev.step()
        capo = rev.capo
        if a == orv.ADVANCE:
            for it in range(oldcapo, capo):
                fstep(it)
        elif a == orv.FIRSTURN:
            fstep(capo)
            image_step(capo)
        elif a == orv.YOUTURN:
            bw.velocity()
            bw.stress()
            image_step(capo)
        elif a == orv.TAKESHOT:
            checks[rev.check] = _snapshot(fw)
        elif a == orv.RESTORE:
            _restore(fw, checks[rev.check])
        if a in (orv.TERMINATE, orv.ERROR):
            break
    return image


def ref_rtm_a2d(vp, rho, lpml, dx, dz, dt, nt, order, fs, ncheck, scoords, wav, gcoords, data, nhx, nhz):
    nz, nx = vp.shape
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    img = np.zeros((nhz, nhx, nz, nx), np.float32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32).ravel()  # noqa: E731
    ref().ref_rtm_a2d(f32(vp), f32(rho), nx, nz, lpml, dx, dz, dt, nt, order, int(fs), ncheck, f32(scoords),
                      f32(wav), gc.shape[0], gc.ravel(), f32(data), nhx, nhz, img.reshape(-1))
    return img


def image_a3d(image, ws, wr, lpml, nhx, nhy, nhz):
    """RtmAcoustic3D::crossCorr (rtm.cpp:601-640): image[ihz, ihy, ihx, iz,
    iy, ix] += ws(ix-hx, iy-hy, iz-hz) * wr(ix+hx, iy+hy, iz+hz), all six
    points' coordinates inside the unpadded grid."""
    nz, ny, nx = image.shape[3:]

    def rng(h, n):
        return max(0, h, -h), min(n, n - h, n + h)
    for ihx in range(nhx):
        hx = -((nhx - 1) // 2) + ihx
        for ihy in range(nhy):
            hy = -((nhy - 1) // 2) + ihy
            for ihz in range(nhz):
                hz = -((nhz - 1) // 2) + ihz
                (x0, x1), (y0, y1), (z0, z1) = rng(hx, nx), rng(hy, ny), rng(hz, nz)
                if x0 >= x1 or y0 >= y1 or z0 >= z1:
                    continue
                s = ws[z0 - hz + lpml:z1 - hz + lpml, y0 - hy + lpml:y1 - hy + lpml, x0 - hx + lpml:x1 - hx + lpml]
                r = wr[z0 + hz + lpml:z1 + hz + lpml, y0 + hy + lpml:y1 + hy + lpml, x0 + hx + lpml:x1 + hx + lpml]
                image[ihz, ihy, ihx, z0:z1, y0:y1, x0:x1] += s * r


def rtm_a3d(vp, rho, lpml, d3, dt, nt, order, fs, ncheck, scoords, wav, gcoords, data, nhx, nhy, nhz,
            real=np.float32):
    """Returns the image [nhz][nhy][nhx][nz][ny][nx] of one shot."""
    nz, ny, nx = vp.shape
    smap, sshift = makemap3d(scoords, (nx, ny, nz), d3, "P")
    gmap, gshift = makemap3d(gcoords, (nx, ny, nz), d3, "P")
    vp = np.ascontiguousarray(vp, real)
    rho = np.ascontiguousarray(rho, real)
    L, Rx, Ry, Rz = stagger_a3d(vp, rho, lpml, fs, real)
    fw = A3DState(L, Rx, Ry, Rz, lpml, order, *d3, dt, real)
    bw = A3DState(L, Rx, Ry, Rz, lpml, order, *d3, dt, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    data = np.ascontiguousarray(data, real).reshape(gmap.shape[0], nt)
    sh_s = np.ascontiguousarray(sshift, real).reshape(-1)
    sh_g = np.ascontiguousarray(gshift, real).reshape(-1)
    image = np.zeros((nhz, nhy, nhx, nz, ny, nx), real)
    checks = {}
    rev = orv.Revolve(nt, ncheck)

    def inject(s, m, sh, d, it):
        fn("inject3d", real)(s.P.reshape(-1), s.L.reshape(-1), s.nx, s.ny, lpml, real(dt), m.reshape(-1), sh,
                             d.reshape(-1), m.shape[0], nt, it)

    def fstep(it):
        fw.velocity()
        fw.stress()
        inject(fw, smap, sh_s, wav, it)

    def image_step(capo):
        inject(bw, gmap, sh_g, data, capo)
        image_a3d(image, fw.P, bw.P, lpml, nhx, nhy, nhz)

    while True:
        oldcapo = rev.capo
        a = rev.step()
        capo = rev.capo
        if a == orv.ADVANCE:
            for it in range(oldcapo, capo):
                fstep(it)
        elif a == orv.FIRSTURN:
            fstep(capo)
            image_step(capo)
        elif a == orv.YOUTURN:
            bw.velocity()
            bw.stress()
            image_step(capo)
        elif a == orv.TAKESHOT:
            checks[rev.check] = _snapshot(fw)
        elif a == orv.RESTORE:
            _restore(fw, checks[rev.check])
        if a in (orv.TERMINATE, orv.ERROR):
            break
    return image


def ref_rtm_a3d(vp, rho, lpml, d3, dt, nt, order, fs, ncheck, scoords, wav, gcoords, data, nhx, nhy, nhz):
    nz, ny, nx = vp.shape
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 3)
    img = np.zeros((nhz, nhy, nhx, nz, ny, nx), np.float32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32).ravel()  # noqa: E731
    ref().ref_rtm_a3d(f32(vp), f32(rho), nx, ny, nz, lpml, *d3, dt, nt, order, int(fs), ncheck, f32(scoords),
                      f32(wav), gc.shape[0], gc.ravel(), f32(data), nhx, nhy, nhz, img.reshape(-1))
    return img

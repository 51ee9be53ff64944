"""Oracle restatement of RtmAcoustic2D::run_optimal (rtm.cpp:423-544): the
source wavefield on the Revolve schedule, the receiver wavefield advanced
with the forward operator from the recorded data (insertSource of dataP at
GMAP), and the imaging condition with subsurface offsets
RtmAcoustic2D::crossCorr (rtm.cpp:191-219) at every reverse step.  Also the
same run on the compiled reference.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import numpy as np

from . import ref
from . import revolve as orv
from .fwi import _restore, _snapshot
from .runs import A2DState, inject2d, makemap2d, stagger_a2d


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

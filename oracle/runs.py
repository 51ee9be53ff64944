"""Whole-run drivers over the oracle (and the same runs on the compiled
reference), following the modelling time loop of modelling.cpp:242-286:
velocity step, stress step, insertSource(it), recordData(it).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import STRUCTS, fn, lib, ref

FIELD = {"P": 0, "VX": 1, "VY": 2, "VZ": 3}
SMAX = 1200.0  # pml.h:16, forced by modelling.cpp:216-218 and fwi.cpp:1048


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def pml_tables(lpml, dt, real=np.float32, smax=SMAX, sign=1):
    tab = np.zeros(12 * lpml, real)
    fn("pml_abc", real)(lpml, real(dt), real(0.0), real(2.0), real(smax), sign, tab)
    return tab


def makemap2d(coords, nx, nz, dx, dz, field="P", ox=0.0, oz=0.0):
    """Data2D::makeMap against an unpadded model geometry (data.cpp:440)."""
    coords = np.ascontiguousarray(coords, np.float32).reshape(-1, 2)
    n = coords.shape[0]
    m = np.full((n, 2), -1, np.int32)
    s = np.zeros((n, 2), np.float32)
    shx = -0.5 if field == "VX" else 0.0
    shz = -0.5 if field == "VZ" else 0.0
    lib().orc_makemap2d(coords.ravel(), n, nx, nz, dx, dz, ox, oz, 0, 0, 0, 0, shx, shz,
                        m.reshape(-1), s.reshape(-1))
    return m, s


def makemap3d(coords, n3, d3, field="P", o3=(0.0, 0.0, 0.0)):
    coords = np.ascontiguousarray(coords, np.float32).reshape(-1, 3)
    n = coords.shape[0]
    m = np.full((n, 3), -1, np.int32)
    s = np.zeros((n, 3), np.float32)
    sh = np.zeros(3, np.float32)
    if field in ("VX", "VY", "VZ"):
        sh["XYZ".index(field[1])] = -0.5
    lib().orc_makemap3d(coords.ravel(), n, np.asarray(n3, np.int32), np.asarray(d3, np.float32),
                        np.asarray(o3, np.float32), np.zeros(3, np.int32), np.zeros(3, np.int32),
                        sh, m.reshape(-1), s.reshape(-1))
    return m, s


def stagger_a2d(vp, rho, lpml, fs, real=np.float32):
    nz, nx = vp.shape
    shp = (nz + 2 * lpml, nx + 2 * lpml)
    L, Rx, Rz = (np.zeros(shp, real) for _ in range(3))
    rc = fn("stagger_a2d", real)(np.ascontiguousarray(vp, real).ravel(),
                                 np.ascontiguousarray(rho, real).ravel(), nx, nz, lpml, int(fs),
                                 L.reshape(-1), Rx.reshape(-1), Rz.reshape(-1))
    if rc:
        raise ValueError("zero density")
    return L, Rx, Rz


def stagger_e2d(vp, vs, rho, lpml, fs, real=np.float32):
    nz, nx = vp.shape
    shp = (nz + 2 * lpml, nx + 2 * lpml)
    L, L2M, M, Rx, Rz = (np.zeros(shp, real) for _ in range(5))
    rc = fn("stagger_e2d", real)(*(np.ascontiguousarray(a, real).ravel() for a in (vp, vs, rho)),
                                 nx, nz, lpml, int(fs), *(a.reshape(-1) for a in (L, L2M, M, Rx, Rz)))
    if rc:
        raise ValueError("zero density")
    return L, L2M, M, Rx, Rz


def stagger_a3d(vp, rho, lpml, fs, real=np.float32):
    nz, ny, nx = vp.shape
    shp = (nz + 2 * lpml, ny + 2 * lpml, nx + 2 * lpml)
    L, Rx, Ry, Rz = (np.zeros(shp, real) for _ in range(4))
    rc = fn("stagger_a3d", real)(np.ascontiguousarray(vp, real).ravel(),
                                 np.ascontiguousarray(rho, real).ravel(), nx, ny, nz, lpml, int(fs),
                                 *(a.reshape(-1) for a in (L, Rx, Ry, Rz)))
    if rc:
        raise ValueError("zero density")
    return L, Rx, Ry, Rz


# --------------------------------------------------------------------------
# Acoustic 2D
# --------------------------------------------------------------------------

class A2DState:
    """Padded-grid state of WavesAcoustic2D plus its PML strips."""

    STRIPS_X = ("P_left", "P_right", "Vxx_left", "Vxx_right")
    STRIPS_Z = ("P_top", "P_bottom", "Vzz_top", "Vzz_bottom")

    def __init__(self, L, Rx, Rz, lpml, order, dx, dz, dt, real=np.float32):
        self.real = real
        nzp, nxp = L.shape
        self.nx, self.nz, self.lpml = nxp, nzp, lpml
        self.L, self.Rx, self.Rz = (np.ascontiguousarray(a, real) for a in (L, Rx, Rz))
        self.P, self.Vx, self.Vz, self.df = (np.zeros((nzp, nxp), real) for _ in range(4))
        self.strips = {k: np.zeros((nzp, lpml), real) for k in self.STRIPS_X}
        self.strips.update({k: np.zeros((lpml, nxp), real) for k in self.STRIPS_Z})
        self.tab = pml_tables(lpml, dt, real)
        ct = C.c_float if real == np.float32 else C.c_double
        st = STRUCTS[real][0]()
        st.nx, st.nz, st.lpml, st.order = nxp, nzp, lpml, order
        st.dx, st.dz, st.dt = dx, dz, dt
        for k in ("P", "Vx", "Vz", "L", "Rx", "Rz", "df", "tab"):
            setattr(st, k, _p(getattr(self, k), ct))
        for k, a in self.strips.items():
            setattr(st, k, _p(a, ct))
        self.st = st
        self.dt = real(dt)

    def velocity(self):
        fn("a2d_velocity", self.real)(C.byref(self.st))

    def stress(self):
        fn("a2d_stress", self.real)(C.byref(self.st))

    def field(self, name):
        return {"P": self.P, "VX": self.Vx, "VZ": self.Vz}[name]

    def modulus(self, name):
        return {"P": self.L, "VX": self.Rx, "VZ": self.Rz}[name]


def inject2d(F, Mod, lpml, dt, smap, sshift, wav, it, real=np.float32):
    nt = wav.shape[-1]
    fn("inject2d", real)(F.reshape(-1), None if Mod is None else Mod.ctypes.data, F.shape[-1], lpml,
                         real(dt), smap.reshape(-1), np.ascontiguousarray(sshift, real).reshape(-1),
                         np.ascontiguousarray(wav, real).reshape(-1), smap.shape[0], nt, it)


def record2d(F, lpml, gmap, gshift, data, it, real=np.float32):
    fn("record2d", real)(F.reshape(-1), F.shape[-1], lpml, gmap.reshape(-1),
                         np.ascontiguousarray(gshift, real).reshape(-1), data.reshape(-1),
                         gmap.shape[0], data.shape[-1], it)


def a2d_run(vp, rho, lpml, dx, dz, dt, nt, order, fs, srcfield, scoords, wav, recfield, gcoords,
            real=np.float32, snap_its=()):
    """Acoustic 2D modelling with the oracle; returns dict(rec, P, Vx, Vz, snaps)."""
    nz, nx = vp.shape
    smap, sshift = makemap2d(scoords, nx, nz, dx, dz, srcfield)
    gmap, gshift = makemap2d(gcoords, nx, nz, dx, dz, recfield)
    L, Rx, Rz = stagger_a2d(vp, rho, lpml, fs, real)
    s = A2DState(L, Rx, Rz, lpml, order, dx, dz, dt, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    rec = np.zeros((gmap.shape[0], nt), real)
    snaps = []
    snap_its = list(snap_its)
    for it in range(nt):
        s.velocity()
        s.stress()
        inject2d(s.field(srcfield), s.modulus(srcfield), lpml, dt, smap, sshift, wav, it, real)
        record2d(s.field(recfield), lpml, gmap, gshift, rec, it, real)
        if snap_its and it == snap_its[0]:
            snaps.append(s.P.copy())
            snap_its.pop(0)
    return dict(rec=rec, P=s.P, Vx=s.Vx, Vz=s.Vz, snaps=snaps, state=s)


def ref_a2d_run(vp, rho, lpml, dx, dz, dt, nt, order, fs, srcfield, scoords, wav, recfield, gcoords,
                snap_its=()):
    """The same run on the compiled reference (WavesAcoustic2D itself)."""
    nz, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    nsrc, nrec = sc.shape[0], gc.shape[0]
    npad = (nz + 2 * lpml, nx + 2 * lpml)
    rec = np.zeros((nrec, nt), np.float32)
    si = np.asarray(list(snap_its) or [-1], np.int32)
    snaps = np.zeros((max(len(snap_its), 1),) + npad, np.float32)
    P, Vx, Vz = (np.zeros(npad, np.float32) for _ in range(3))
    ref().ref_a2d_run(np.ascontiguousarray(vp, np.float32).ravel(),
                      np.ascontiguousarray(rho, np.float32).ravel(), nx, nz, lpml, dx, dz, dt, nt,
                      order, int(fs), FIELD[srcfield], nsrc, sc.ravel(),
                      np.ascontiguousarray(wav, np.float32).ravel(), FIELD[recfield], nrec,
                      gc.ravel(), rec.reshape(-1), len(snap_its), si, snaps.reshape(-1),
                      P.reshape(-1), Vx.reshape(-1), Vz.reshape(-1))
    return dict(rec=rec, P=P, Vx=Vx, Vz=Vz, snaps=list(snaps[:len(snap_its)]))


# --------------------------------------------------------------------------
# Elastic 2D
# --------------------------------------------------------------------------

class E2DState:
    STRIPS_X = ("Sxx_left", "Sxx_right", "Sxzx_left", "Sxzx_right",
                "Vxx_left", "Vxx_right", "Vzx_left", "Vzx_right")
    STRIPS_Z = ("Szz_top", "Szz_bottom", "Sxzz_top", "Sxzz_bottom",
                "Vzz_top", "Vzz_bottom", "Vxz_top", "Vxz_bottom")
    FIELDS = ("Vx", "Vz", "Sxx", "Szz", "Sxz")

    def __init__(self, L, L2M, M, Rx, Rz, lpml, order, dx, dz, dt, fs, real=np.float32):
        self.real = real
        nzp, nxp = L.shape
        self.nx, self.nz, self.lpml = nxp, nzp, lpml
        self.L, self.L2M, self.M, self.Rx, self.Rz = (np.ascontiguousarray(a, real)
                                                      for a in (L, L2M, M, Rx, Rz))
        for k in self.FIELDS + ("df", "wrk"):
            setattr(self, k, np.zeros((nzp, nxp), real))
        self.strips = {k: np.zeros((nzp, lpml), real) for k in self.STRIPS_X}
        self.strips.update({k: np.zeros((lpml, nxp), real) for k in self.STRIPS_Z})
        self.tab = pml_tables(lpml, dt, real)
        ct = C.c_float if real == np.float32 else C.c_double
        st = STRUCTS[real][1]()
        st.nx, st.nz, st.lpml, st.order, st.fs = nxp, nzp, lpml, order, int(fs)
        st.dx, st.dz, st.dt = dx, dz, dt
        for k in self.FIELDS + ("L", "L2M", "M", "Rx", "Rz", "df", "wrk", "tab"):
            setattr(st, k, _p(getattr(self, k), ct))
        for k, a in self.strips.items():
            setattr(st, k, _p(a, ct))
        self.st = st

    def velocity(self):
        fn("e2d_velocity", self.real)(C.byref(self.st))

    def stress(self):
        fn("e2d_stress", self.real)(C.byref(self.st))

    def velocity_bw(self):
        fn("e2d_velocity_bw", self.real)(C.byref(self.st))

    def stress_bw(self):
        fn("e2d_stress_bw", self.real)(C.byref(self.st))


def e2d_record_factor(vp, vs, rho, gmap, recip, real=np.float32):
    """Per-trace pressure scaling: rho*Vp^2 - rho*Vs^2 at the receiver cell of
    the unpadded model when reciprocity is set, else 1 (waves.cpp:2219-2228)."""
    f = np.ones(gmap.shape[0], real)
    if recip:
        for i, (x, z) in enumerate(gmap):
            if x >= 0 and z >= 0:
                r, p, s = real(rho[z, x]), real(vp[z, x]), real(vs[z, x])
                f[i] = r * p * p - r * s * s
    return f


def e2d_run(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, srcfield, scoords, wav, recfield,
            gcoords, recip=False, nbw=0, real=np.float32):
    nz, nx = vp.shape
    smap, sshift = makemap2d(scoords, nx, nz, dx, dz, srcfield)
    gmap, gshift = makemap2d(gcoords, nx, nz, dx, dz, recfield)
    L, L2M, M, Rx, Rz = stagger_e2d(vp, vs, rho, lpml, fs, real)
    s = E2DState(L, L2M, M, Rx, Rz, lpml, order, dx, dz, dt, fs, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    rec = np.zeros((gmap.shape[0], nt), real)
    factor = e2d_record_factor(vp, vs, rho, gmap, recip, real)
    for it in range(nt):
        s.velocity()
        s.stress()
        if srcfield == "P":
            inject2d(s.Sxx, None, lpml, dt, smap, sshift, wav, it, real)
            inject2d(s.Szz, None, lpml, dt, smap, sshift, wav, it, real)
        elif srcfield == "VX":
            inject2d(s.Vx, s.Rx, lpml, dt, smap, sshift, wav, it, real)
        else:
            inject2d(s.Vz, s.Rz, lpml, dt, smap, sshift, wav, it, real)
        if recfield == "P":
            fn("record2d_pressure_el", real)(s.Sxx.reshape(-1), s.Szz.reshape(-1), s.nx, lpml,
                                             gmap.reshape(-1), np.ascontiguousarray(gshift, real).reshape(-1),
                                             factor, rec.reshape(-1), gmap.shape[0], nt, it)
        else:
            record2d(s.Vx if recfield == "VX" else s.Vz, lpml, gmap, gshift, rec, it, real)
    for _ in range(nbw):
        s.velocity_bw()
        s.stress_bw()
    return dict(rec=rec, state=s, **{k: getattr(s, k) for k in E2DState.FIELDS})


def ref_e2d_run(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, srcfield, scoords, wav, recfield,
                gcoords, recip=False, nbw=0):
    nz, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    rec = np.zeros((gc.shape[0], nt), np.float32)
    fields = np.zeros((5, nz + 2 * lpml, nx + 2 * lpml), np.float32)
    ref().ref_e2d_run(*(np.ascontiguousarray(a, np.float32).ravel() for a in (vp, vs, rho)),
                      nx, nz, lpml, dx, dz, dt, nt, order, int(fs), FIELD[srcfield], sc.shape[0],
                      sc.ravel(), np.ascontiguousarray(wav, np.float32).ravel(), FIELD[recfield],
                      int(recip), gc.shape[0], gc.ravel(), rec.reshape(-1), nbw, fields.reshape(-1))
    return dict(rec=rec, **dict(zip(E2DState.FIELDS, fields)))


def ref_a2d_time_window(vp, rho, lpml, dx, dz, dt, nt, order, scoords, wav, gcoords, it0, n, state):
    """Seconds the compiled reference takes for time steps it0 .. it0+n-1 of
    an acoustic 2D run (PRESSURE source and receivers, no free surface)
    started from `state`: P, Vx, Vz then the 8 PML strips, flat float32
    (ref_harness.cpp ref_a2d_time_window)."""
    nz, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    return ref().ref_a2d_time_window(*(np.ascontiguousarray(a, np.float32).ravel() for a in (vp, rho)),
                                     nx, nz, lpml, dx, dz, dt, nt, order, sc.shape[0], sc.ravel(),
                                     np.ascontiguousarray(wav, np.float32).ravel(), gc.shape[0], gc.ravel(),
                                     int(it0), int(n), np.ascontiguousarray(state, np.float32).ravel())


def ref_e2d_time_window(vp, vs, rho, lpml, dx, dz, dt, nt, order, fs, scoords, wav, gcoords, it0, n, state):
    """Seconds the compiled reference takes for time steps it0 .. it0+n-1 of
    an elastic 2D run (explosive source, VX and VZ receivers) started from
    `state`: Vx, Vz, Sxx, Szz, Sxz then the 16 PML strips, flat float32
    (ref_harness.cpp ref_e2d_time_window).  Releases the GIL: calls on
    separate threads run in parallel."""
    nz, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 2)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 2)
    return ref().ref_e2d_time_window(*(np.ascontiguousarray(a, np.float32).ravel() for a in (vp, vs, rho)),
                                     nx, nz, lpml, dx, dz, dt, nt, order, int(fs), sc.shape[0], sc.ravel(),
                                     np.ascontiguousarray(wav, np.float32).ravel(), gc.shape[0], gc.ravel(),
                                     int(it0), int(n), np.ascontiguousarray(state, np.float32).ravel())


# --------------------------------------------------------------------------
# Acoustic 3D
# --------------------------------------------------------------------------

class A3DState:
    STRIPS = {"P_left": 0, "P_right": 0, "Vxx_left": 0, "Vxx_right": 0,
              "P_front": 1, "P_back": 1, "Vyy_front": 1, "Vyy_back": 1,
              "P_top": 2, "P_bottom": 2, "Vzz_top": 2, "Vzz_bottom": 2}

    def __init__(self, L, Rx, Ry, Rz, lpml, order, dx, dy, dz, dt, real=np.float32):
        self.real = real
        nzp, nyp, nxp = L.shape
        self.nx, self.ny, self.nz, self.lpml = nxp, nyp, nzp, lpml
        self.L, self.Rx, self.Ry, self.Rz = (np.ascontiguousarray(a, real) for a in (L, Rx, Ry, Rz))
        for k in ("P", "Vx", "Vy", "Vz", "df"):
            setattr(self, k, np.zeros((nzp, nyp, nxp), real))
        shp = {0: (nzp, nyp, lpml), 1: (nzp, lpml, nxp), 2: (lpml, nyp, nxp)}
        self.strips = {k: np.zeros(shp[a], real) for k, a in self.STRIPS.items()}
        self.tab = pml_tables(lpml, dt, real)
        ct = C.c_float if real == np.float32 else C.c_double
        st = STRUCTS[real][2]()
        st.nx, st.ny, st.nz, st.lpml, st.order = nxp, nyp, nzp, lpml, order
        st.dx, st.dy, st.dz, st.dt = dx, dy, dz, dt
        for k in ("P", "Vx", "Vy", "Vz", "L", "Rx", "Ry", "Rz", "df", "tab"):
            setattr(st, k, _p(getattr(self, k), ct))
        for k, a in self.strips.items():
            setattr(st, k, _p(a, ct))
        self.st = st

    def velocity(self):
        fn("a3d_velocity", self.real)(C.byref(self.st))

    def stress(self):
        fn("a3d_stress", self.real)(C.byref(self.st))


def a3d_run(vp, rho, lpml, d3, dt, nt, order, fs, srcfield, scoords, wav, recfield, gcoords,
            real=np.float32):
    nz, ny, nx = vp.shape
    smap, sshift = makemap3d(scoords, (nx, ny, nz), d3, srcfield)
    gmap, gshift = makemap3d(gcoords, (nx, ny, nz), d3, recfield)
    L, Rx, Ry, Rz = stagger_a3d(vp, rho, lpml, fs, real)
    s = A3DState(L, Rx, Ry, Rz, lpml, order, *d3, dt, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    rec = np.zeros((gmap.shape[0], nt), real)
    fld = {"P": (s.P, s.L), "VX": (s.Vx, s.Rx), "VY": (s.Vy, s.Ry), "VZ": (s.Vz, s.Rz)}
    for it in range(nt):
        s.velocity()
        s.stress()
        F, Mod = fld[srcfield]
        fn("inject3d", real)(F.reshape(-1), Mod.reshape(-1), s.nx, s.ny, lpml, real(dt),
                             smap.reshape(-1), np.ascontiguousarray(sshift, real).reshape(-1),
                             wav.reshape(-1), smap.shape[0], nt, it)
        F = fld[recfield][0]
        fn("record3d", real)(F.reshape(-1), s.nx, s.ny, lpml, gmap.reshape(-1),
                             np.ascontiguousarray(gshift, real).reshape(-1), rec.reshape(-1),
                             gmap.shape[0], nt, it)
    return dict(rec=rec, P=s.P, Vx=s.Vx, Vy=s.Vy, Vz=s.Vz, state=s)


def ref_a3d_run(vp, rho, lpml, d3, dt, nt, order, fs, srcfield, scoords, wav, recfield, gcoords):
    nz, ny, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 3)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 3)
    rec = np.zeros((gc.shape[0], nt), np.float32)
    fields = np.zeros((4, nz + 2 * lpml, ny + 2 * lpml, nx + 2 * lpml), np.float32)
    ref().ref_a3d_run(np.ascontiguousarray(vp, np.float32).ravel(),
                      np.ascontiguousarray(rho, np.float32).ravel(), nx, ny, nz, lpml, *d3, dt, nt,
                      order, int(fs), FIELD[srcfield], sc.shape[0], sc.ravel(),
                      np.ascontiguousarray(wav, np.float32).ravel(), FIELD[recfield], gc.shape[0],
                      gc.ravel(), rec.reshape(-1), fields.reshape(-1))
    return dict(rec=rec, P=fields[0], Vx=fields[1], Vy=fields[2], Vz=fields[3])


# --------------------------------------------------------------------------
# Elastic 3D
# --------------------------------------------------------------------------

E3D_FIELDS = ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")
# strip slots of orc_e3d_t: (name, face axis), left/front/top first
E3D_STRIPS = [(n, 0) for n in ("Sxx", "Sxzx", "Sxyx", "Vxx", "Vzx", "Vyx")] + \
             [(n, 1) for n in ("Syy", "Sxyy", "Syzy", "Vyy", "Vxy", "Vzy")] + \
             [(n, 2) for n in ("Szz", "Sxzz", "Syzz", "Vzz", "Vxz", "Vyz")]


def stagger_e3d(vp, vs, rho, lpml, fs, real=np.float32):
    nz, ny, nx = vp.shape
    shp = (nz + 2 * lpml, ny + 2 * lpml, nx + 2 * lpml)
    out = [np.zeros(shp, real) for _ in range(8)]
    rc = fn("stagger_e3d", real)(*(np.ascontiguousarray(a, real).ravel() for a in (vp, vs, rho)), nx, ny, nz, lpml,
                                 int(fs), *(a.reshape(-1) for a in out))
    if rc:
        raise ValueError("zero density")
    return out  # L, L2M, Mxz, Myz, Mxy, Rx, Ry, Rz


class E3DState:
    """Padded-grid state of WavesElastic3D plus its 36 PML strips."""

    def __init__(self, model, lpml, order, d3, dt, fs, real=np.float32):
        self.real = real
        L = model[0]
        nzp, nyp, nxp = L.shape
        self.nx, self.ny, self.nz, self.lpml = nxp, nyp, nzp, lpml
        (self.L, self.L2M, self.Mxz, self.Myz, self.Mxy, self.Rx, self.Ry,
         self.Rz) = (np.ascontiguousarray(a, real) for a in model)
        for k in E3D_FIELDS + ("df", "wrk"):
            setattr(self, k, np.zeros((nzp, nyp, nxp), real))
        shp = {0: (nzp, nyp, lpml), 1: (nzp, lpml, nxp), 2: (lpml, nyp, nxp)}
        self.strips = []
        for name, ax in E3D_STRIPS:
            self.strips += [np.zeros(shp[ax], real), np.zeros(shp[ax], real)]
        self.tab = pml_tables(lpml, dt, real)
        ct = C.c_float if real == np.float32 else C.c_double
        st = STRUCTS[real][3]()
        st.nx, st.ny, st.nz, st.lpml, st.order, st.fs = nxp, nyp, nzp, lpml, order, int(fs)
        st.dx, st.dy, st.dz, st.dt = d3[0], d3[1], d3[2], dt
        for k in E3D_FIELDS + ("L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz", "df", "wrk", "tab"):
            setattr(st, k, _p(getattr(self, k), ct))
        for i, a in enumerate(self.strips):
            st.psi[i] = _p(a, ct)
        self.st = st

    def velocity(self):
        fn("e3d_velocity", self.real)(C.byref(self.st))

    def stress(self):
        fn("e3d_stress", self.real)(C.byref(self.st))

    def velocity_bw(self):
        fn("e3d_velocity_bw", self.real)(C.byref(self.st))

    def stress_bw(self):
        fn("e3d_stress_bw", self.real)(C.byref(self.st))


def e3d_run(vp, vs, rho, lpml, d3, dt, nt, order, fs, srcfield, scoords, wav, recfield, gcoords, real=np.float32):
    """Elastic 3D modelling with the oracle; returns dict(rec, state, fields)."""
    nz, ny, nx = vp.shape
    smap, sshift = makemap3d(scoords, (nx, ny, nz), d3, srcfield)
    gmap, gshift = makemap3d(gcoords, (nx, ny, nz), d3, recfield)
    s = E3DState(stagger_e3d(vp, vs, rho, lpml, fs, real), lpml, order, d3, dt, fs, real)
    wav = np.ascontiguousarray(wav, real).reshape(smap.shape[0], nt)
    rec = np.zeros((gmap.shape[0], nt), real)
    sh32 = np.ascontiguousarray(sshift, real).reshape(-1)
    gh32 = np.ascontiguousarray(gshift, real).reshape(-1)
    mod = [np.ascontiguousarray(a, real).ravel() for a in (vp, vs, rho)]
    for it in range(nt):
        s.velocity()
        s.stress()
        if srcfield == "P":
            fn("inject3d_explosive", real)(s.Sxx.reshape(-1), s.Syy.reshape(-1), s.Szz.reshape(-1), s.nx, s.ny,
                                           lpml, real(dt), smap.reshape(-1), sh32, wav.reshape(-1), smap.shape[0],
                                           nt, it)
        else:
            F, M = {"VX": (s.Vx, s.Rx), "VY": (s.Vy, s.Ry), "VZ": (s.Vz, s.Rz)}[srcfield]
            fn("inject3d", real)(F.reshape(-1), M.reshape(-1), s.nx, s.ny, lpml, real(dt), smap.reshape(-1), sh32,
                                 wav.reshape(-1), smap.shape[0], nt, it)
        if recfield == "P":
            fn("record3d_pressure_el", real)(s.Sxx.reshape(-1), s.Syy.reshape(-1), s.Szz.reshape(-1), s.nx, s.ny,
                                             lpml, *mod, nx, ny, gmap.reshape(-1), gh32, rec.reshape(-1),
                                             gmap.shape[0], nt, it)
        else:
            F = {"VX": s.Vx, "VY": s.Vy, "VZ": s.Vz}[recfield]
            fn("record3d", real)(F.reshape(-1), s.nx, s.ny, lpml, gmap.reshape(-1), gh32, rec.reshape(-1),
                                 gmap.shape[0], nt, it)
    return dict(rec=rec, state=s, **{k: getattr(s, k) for k in E3D_FIELDS})


_FIELD_CODE = {"P": 0, "VX": 1, "VY": 2, "VZ": 3}


def ref_e3d_run(vp, vs, rho, lpml, d3, dt, nt, order, fs, srcfield, scoords, wav, recfield, gcoords):
    """The compiled reference's WavesElastic3D modelling run."""
    nz, ny, nx = vp.shape
    sc = np.ascontiguousarray(scoords, np.float32).reshape(-1, 3)
    gc = np.ascontiguousarray(gcoords, np.float32).reshape(-1, 3)
    rec = np.zeros((gc.shape[0], nt), np.float32)
    fields = np.zeros((9, nz + 2 * lpml, ny + 2 * lpml, nx + 2 * lpml), np.float32)
    ref().ref_e3d_run(*(np.ascontiguousarray(a, np.float32).ravel() for a in (vp, vs, rho)), nx, ny, nz, lpml,
                      *(np.float32(v) for v in d3), np.float32(dt), nt, order, int(fs), _FIELD_CODE[srcfield],
                      sc.shape[0], sc.ravel(), np.ascontiguousarray(wav, np.float32).ravel(), _FIELD_CODE[recfield],
                      gc.shape[0], gc.ravel(), rec.reshape(-1), fields.reshape(-1))
    return dict(rec=rec, **{k: fields[i] for i, k in enumerate(E3D_FIELDS)})


def ref_stagger_e3d(vp, vs, rho, lpml, fs):
    nz, ny, nx = vp.shape
    out = np.zeros((8, nz + 2 * lpml, ny + 2 * lpml, nx + 2 * lpml), np.float32)
    ref().ref_stagger_e3d(*(np.ascontiguousarray(a, np.float32).ravel() for a in (vp, vs, rho)), nx, ny, nz, lpml,
                          int(fs), out.reshape(-1))
    return list(out)

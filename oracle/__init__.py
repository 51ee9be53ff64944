"""CPU oracle for the RockSeis staggered-grid FD step and FWI gradient.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg may import this package, and only as the
checker (or the CPU timer).  The product (``rockseis_amd``) never imports it.

Two back ends are wrapped here with ctypes:

* ``liboracle.so`` -- ``rsoracle.c``, a from-scratch C restatement of the
  reference's der.cpp / pml.cpp / waves.cpp / model.cpp / geometry.cpp /
  fwi.cpp arithmetic in fp32 (``*_f32``) and fp64 (``*_f64``).
* ``_ref/librsref.so`` -- the reference itself, compiled from
  /root/reference by ``make -C oracle ref`` (present only where it was built),
  driven through ``ref_harness.cpp``.

``revolve.py`` holds the oracle's restatement of the Revolve schedule.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def lib():
    """Load liboracle.so (build with ``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        _lib = C.CDLL(path)
        _declare(_lib)
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "librsref.so"))


def ref():
    """Load the compiled reference harness (``make -C oracle ref``)."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "librsref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path}")
        _ref = C.CDLL(path)
        _declare_ref(_ref)
    return _ref


# --------------------------------------------------------------------------
# ctypes structures mirroring rsoracle_impl.h
# --------------------------------------------------------------------------

def _ptr_fields(names, ctype):
    return [(n, C.POINTER(ctype)) for n in names]


def _make_structs(real):
    ct = C.c_float if real == np.float32 else C.c_double

    class A2D(C.Structure):
        _fields_ = ([("nx", C.c_int), ("nz", C.c_int), ("lpml", C.c_int), ("order", C.c_int),
                     ("dx", ct), ("dz", ct), ("dt", ct)]
                    + _ptr_fields(["tab", "P", "Vx", "Vz", "P_left", "P_right", "P_top", "P_bottom",
                                   "Vxx_left", "Vxx_right", "Vzz_top", "Vzz_bottom",
                                   "L", "Rx", "Rz", "df"], ct))

    class E2D(C.Structure):
        _fields_ = ([("nx", C.c_int), ("nz", C.c_int), ("lpml", C.c_int), ("order", C.c_int),
                     ("fs", C.c_int), ("dx", ct), ("dz", ct), ("dt", ct)]
                    + _ptr_fields(["tab", "Sxx", "Szz", "Sxz", "Vx", "Vz",
                                   "Sxx_left", "Sxx_right", "Sxzx_left", "Sxzx_right",
                                   "Szz_top", "Szz_bottom", "Sxzz_top", "Sxzz_bottom",
                                   "Vxx_left", "Vxx_right", "Vzx_left", "Vzx_right",
                                   "Vzz_top", "Vzz_bottom", "Vxz_top", "Vxz_bottom",
                                   "L", "L2M", "M", "Rx", "Rz", "df", "wrk"], ct))

    class A3D(C.Structure):
        _fields_ = ([("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("lpml", C.c_int),
                     ("order", C.c_int), ("dx", ct), ("dy", ct), ("dz", ct), ("dt", ct)]
                    + _ptr_fields(["tab", "P", "Vx", "Vy", "Vz",
                                   "P_left", "P_right", "P_front", "P_back", "P_top", "P_bottom",
                                   "Vxx_left", "Vxx_right", "Vyy_front", "Vyy_back",
                                   "Vzz_top", "Vzz_bottom", "L", "Rx", "Ry", "Rz", "df"], ct))

    class E3D(C.Structure):
        _fields_ = ([("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("lpml", C.c_int), ("order", C.c_int),
                     ("fs", C.c_int), ("dx", ct), ("dy", ct), ("dz", ct), ("dt", ct)]
                    + _ptr_fields(["tab", "Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy"], ct)
                    + [("psi", C.POINTER(ct) * 36)]
                    + _ptr_fields(["L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz", "df", "wrk"], ct))

    return A2D, E2D, A3D, E3D


STRUCTS = {np.float32: _make_structs(np.float32), np.float64: _make_structs(np.float64)}
_SUF = {np.float32: "_f32", np.float64: "_f64"}


def _declare(L):
    for real, suf in _SUF.items():
        rp = _f32p if real == np.float32 else _f64p
        ct = C.c_float if real == np.float32 else C.c_double
        A2D, E2D, A3D, E3D = STRUCTS[real]
        f = getattr(L, "orc_der_coeffs" + suf)
        f.argtypes = [rp]
        f = getattr(L, "orc_deriv" + suf)
        f.argtypes = [rp, rp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ct]
        f = getattr(L, "orc_pml_abc" + suf)
        f.argtypes = [C.c_int, ct, ct, ct, ct, C.c_int, rp]
        for nm, st in (("a2d_velocity", A2D), ("a2d_stress", A2D), ("e2d_velocity", E2D),
                       ("e2d_stress", E2D), ("e2d_velocity_bw", E2D), ("e2d_stress_bw", E2D),
                       ("a3d_velocity", A3D), ("a3d_stress", A3D)):
            f = getattr(L, "orc_" + nm + suf)
            f.argtypes = [C.POINTER(st)]
        f = getattr(L, "orc_inject2d" + suf)
        f.argtypes = [rp, C.c_void_p, C.c_int, C.c_int, ct, _i32p, rp, rp, C.c_int, C.c_int, C.c_int]
        f = getattr(L, "orc_record2d" + suf)
        f.argtypes = [rp, C.c_int, C.c_int, _i32p, rp, rp, C.c_int, C.c_int, C.c_int]
        f = getattr(L, "orc_record2d_pressure_el" + suf)
        f.argtypes = [rp, rp, C.c_int, C.c_int, _i32p, rp, rp, rp, C.c_int, C.c_int, C.c_int]
        f = getattr(L, "orc_inject3d" + suf)
        f.argtypes = [rp, rp, C.c_int, C.c_int, C.c_int, ct, _i32p, rp, rp, C.c_int, C.c_int, C.c_int]
        f = getattr(L, "orc_record3d" + suf)
        f.argtypes = [rp, C.c_int, C.c_int, C.c_int, _i32p, rp, rp, C.c_int, C.c_int, C.c_int]
        f = getattr(L, "orc_stagger_a2d" + suf)
        f.argtypes = [rp, rp, C.c_int, C.c_int, C.c_int, C.c_int, rp, rp, rp]
        f.restype = C.c_int
        f = getattr(L, "orc_stagger_e2d" + suf)
        f.argtypes = [rp, rp, rp, C.c_int, C.c_int, C.c_int, C.c_int, rp, rp, rp, rp, rp]
        f.restype = C.c_int
        f = getattr(L, "orc_stagger_a3d" + suf)
        f.argtypes = [rp, rp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, rp, rp, rp, rp]
        f.restype = C.c_int
        f = getattr(L, "orc_conv" + suf)
        f.argtypes = [C.c_int, C.c_int, rp, C.c_int, C.c_int, rp, C.c_int, C.c_int, rp]
        f = getattr(L, "orc_stoep" + suf)
        f.argtypes = [C.c_int, rp, rp, rp, rp]
        f = getattr(L, "orc_deriv_t" + suf)
        f.argtypes = [rp, rp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ct]
        f = getattr(L, "orc_a2d_adj_stress" + suf)
        f.argtypes = [C.POINTER(A2D), C.POINTER(A2D), rp]
        f = getattr(L, "orc_a2d_adj_velocity" + suf)
        f.argtypes = [C.POINTER(A2D), C.POINTER(A2D), rp, rp]
        f = getattr(L, "orc_e2d_adj_stress" + suf)
        f.argtypes = [C.POINTER(E2D), C.POINTER(E2D), rp, rp, rp]
        f = getattr(L, "orc_e2d_adj_velocity" + suf)
        f.argtypes = [C.POINTER(E2D), C.POINTER(E2D), rp, rp]
        for nm in ("e3d_velocity", "e3d_stress", "e3d_velocity_bw", "e3d_stress_bw"):
            f = getattr(L, "orc_" + nm + suf)
            f.argtypes = [C.POINTER(E3D)]
        f = getattr(L, "orc_e3d_crosscorr" + suf)
        f.argtypes = [rp] * 21 + [C.c_int] * 4 + [ct] * 3 + [rp] * 3
        f = getattr(L, "orc_e3d_scalegrad" + suf)
        f.argtypes = [rp, rp, rp, C.c_long, rp, rp, rp]
        f = getattr(L, "orc_misfit_e3" + suf)
        f.argtypes = [C.c_int] + [rp] * 6 + [C.c_void_p] * 3 + [C.c_int, C.c_int]
        f.restype = ct
        f = getattr(L, "orc_residual_e3" + suf)
        f.argtypes = [C.c_int] + [rp] * 6 + [C.c_void_p] * 3 + [rp] * 3 + [C.c_int, C.c_int]
        f = getattr(L, "orc_inject3d_explosive" + suf)
        f.argtypes = [rp, rp, rp, C.c_int, C.c_int, C.c_int, ct, _i32p, rp, rp, C.c_int, C.c_int, C.c_int]
        f = getattr(L, "orc_record3d_pressure_el" + suf)
        f.argtypes = [rp, rp, rp, C.c_int, C.c_int, C.c_int, rp, rp, rp, C.c_int, C.c_int, _i32p, rp, rp, C.c_int,
                      C.c_int, C.c_int]
        f = getattr(L, "orc_stagger_e3d" + suf)
        f.argtypes = [rp, rp, rp] + [C.c_int] * 5 + [rp] * 8
        f.restype = C.c_int
        f = getattr(L, "orc_a3d_adj_stress" + suf)
        f.argtypes = [C.POINTER(A3D), C.POINTER(A3D), rp]
        f = getattr(L, "orc_a3d_adj_velocity" + suf)
        f.argtypes = [C.POINTER(A3D), C.POINTER(A3D), rp, rp, rp]
        f = getattr(L, "orc_a2d_crosscorr" + suf)
        f.argtypes = [rp, C.c_int, rp, rp, C.c_int, rp, rp, rp, rp, C.c_int, C.c_int, ct, ct, rp, rp]
        f = getattr(L, "orc_misfit_a" + suf)
        f.argtypes = [C.c_int, rp, rp, C.c_void_p, _i32p, C.c_int, C.c_int]
        f.restype = ct
        f = getattr(L, "orc_residual_a" + suf)
        f.argtypes = [C.c_int, rp, rp, C.c_void_p, rp, C.c_int, C.c_int, C.c_int, C.c_int]
        f = getattr(L, "orc_misfit_e" + suf)
        f.argtypes = [C.c_int] + [rp] * 6 + [C.c_void_p] * 3 + [_i32p, C.c_int, C.c_int]
        f.restype = ct
        f = getattr(L, "orc_residual_e" + suf)
        f.argtypes = [C.c_int] + [rp] * 6 + [C.c_void_p] * 3 + [rp] * 3 + [C.c_int, C.c_int]
        f = getattr(L, "orc_interp" + suf)
        f.argtypes = [C.c_int, rp, C.c_int, C.c_int, ct, ct, rp, C.c_int, ct, ct]
        f = getattr(L, "orc_a3d_crosscorr" + suf)
        f.argtypes = [rp, C.c_int, rp, rp, rp, C.c_int] + [rp] * 5 + [C.c_int] * 3 + [ct] * 3 + [rp, rp]
        f = getattr(L, "orc_e2d_crosscorr" + suf)
        f.argtypes = [rp] * 12 + [C.c_int] * 3 + [ct, ct, rp, rp, rp, C.c_void_p, C.c_int, C.c_void_p,
                                                    C.c_void_p, C.c_int, C.c_int, C.c_int]
        f = getattr(L, "orc_e2d_scalegrad" + suf)
        f.argtypes = [rp, rp, rp, C.c_int, C.c_int, rp, rp, rp]
        f = getattr(L, "orc_misfit_diff3" + suf)
        f.argtypes = [rp] * 6 + [_i32p, C.c_int, C.c_int]
        f.restype = ct
        f = getattr(L, "orc_misfit_diff" + suf)
        f.argtypes = [rp, rp, _i32p, C.c_int, C.c_int]
        f.restype = ct
        f = getattr(L, "orc_residual_diff" + suf)
        f.argtypes = [rp, rp, rp, C.c_int, C.c_int, C.c_int]
    L.orc_makemap2d.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                _i32p, _f32p]
    L.orc_makemap3d.argtypes = [_f32p, C.c_int, _i32p, _f32p, _f32p, _i32p, _i32p, _f32p, _i32p, _f32p]
    L.orc_kfw_export.restype = C.c_int
    L.orc_kbw_export.restype = C.c_int


def _declare_ref(R):
    R.ref_e3d_run.argtypes = ([_f32p, _f32p, _f32p] + [C.c_int] * 4 + [C.c_float] * 4 + [C.c_int] * 5
                              + [_f32p, _f32p, C.c_int, C.c_int, _f32p, _f32p, _f32p])
    R.ref_stagger_e3d.argtypes = [_f32p, _f32p, _f32p] + [C.c_int] * 5 + [_f32p]
    R.ref_fwi_e3d.argtypes = ([_f32p] * 3 + [C.c_int] * 4 + [C.c_float] * 4 + [C.c_int] * 6 + [_f32p, _f32p, C.c_int,
                              _f32p, _f32p] + [_f32p] * 6 + [C.c_int] + [C.c_void_p] * 3)
    R.ref_fwi_e3d.restype = C.c_float
    R.ref_der.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                          C.c_int, C.c_int]
    R.ref_pml_abc.argtypes = [C.c_int, C.c_float, C.c_float, C.c_int, _f32p]
    R.ref_stagger_a2d.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p]
    R.ref_stagger_e2d.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int,
                                  _f32p, _f32p, _f32p, _f32p, _f32p]
    R.ref_stagger_a3d.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  _f32p, _f32p, _f32p, _f32p]
    R.ref_makemap2d.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                C.c_float, C.c_int, _i32p, _f32p]
    R.ref_a2d_run.argtypes = ([_f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3 + [C.c_int] * 5
                              + [_f32p, _f32p, C.c_int, C.c_int, _f32p, _f32p, C.c_int, _i32p,
                                 _f32p, _f32p, _f32p, _f32p])
    R.ref_e2d_run.argtypes = ([_f32p, _f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3
                              + [C.c_int] * 5 + [_f32p, _f32p, C.c_int, C.c_int, C.c_int, _f32p,
                                                 _f32p, C.c_int, _f32p])
    R.ref_a2d_time_window.argtypes = ([_f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3 + [C.c_int] * 3
                                      + [_f32p, _f32p, C.c_int, _f32p, C.c_int, C.c_int, _f32p])
    R.ref_a2d_time_window.restype = C.c_double
    R.ref_e2d_time_window.argtypes = ([_f32p, _f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3
                                      + [C.c_int] * 4 + [_f32p, _f32p, C.c_int, _f32p, C.c_int, C.c_int,
                                                         _f32p])
    R.ref_e2d_time_window.restype = C.c_double
    R.ref_a3d_run.argtypes = ([_f32p, _f32p] + [C.c_int] * 4 + [C.c_float] * 4 + [C.c_int] * 5
                              + [_f32p, _f32p, C.c_int, C.c_int, _f32p, _f32p, _f32p])
    R.ref_fwi_a2d.argtypes = ([_f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3 + [C.c_int] * 5
                              + [_f32p, _f32p, C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p])
    R.ref_fwi_a2d.restype = C.c_float
    R.ref_fwi_a2d_misfit.argtypes = ([_f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3 + [C.c_int] * 5
                                     + [_f32p, _f32p, C.c_int, _f32p, _f32p, C.c_int, C.c_void_p]
                                     + [_f32p] * 5)
    R.ref_fwi_a2d_misfit.restype = C.c_float
    R.ref_fwi_a2d_full.argtypes = ([_f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3 + [C.c_int] * 4
                                   + [C.c_char_p, C.c_int, _f32p, _f32p, C.c_int, _f32p, _f32p, _f32p,
                                      _f32p, _f32p, _f32p, _f32p])
    R.ref_fwi_a2d_full.restype = C.c_float
    R.ref_fwi_e2d.argtypes = ([_f32p, _f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3 + [C.c_int] * 6
                              + [_f32p, _f32p, C.c_int] + [_f32p] * 8 + [C.c_int] + [C.c_void_p] * 3)
    R.ref_fwi_e2d.restype = C.c_float
    R.ref_fwi_a3d.argtypes = ([_f32p, _f32p] + [C.c_int] * 4 + [C.c_float] * 4 + [C.c_int] * 4
                              + [_f32p, _f32p, C.c_int, _f32p, _f32p] + [_f32p] * 5 + [C.c_int, C.c_void_p])
    R.ref_fwi_a3d.restype = C.c_float
    R.ref_rtm_a2d.argtypes = ([_f32p, _f32p] + [C.c_int] * 3 + [C.c_float] * 3 + [C.c_int] * 4
                              + [_f32p, _f32p, C.c_int, _f32p, _f32p, C.c_int, C.c_int, _f32p])
    R.ref_rtm_a2d.restype = None
    R.ref_rtm_a3d.argtypes = ([_f32p, _f32p] + [C.c_int] * 4 + [C.c_float] * 4 + [C.c_int] * 4
                              + [_f32p, _f32p, C.c_int, _f32p, _f32p] + [C.c_int] * 3 + [_f32p])
    R.ref_rtm_a3d.restype = None
    R.ref_interp.argtypes = [C.c_int, _f32p, C.c_int, C.c_int, C.c_float, C.c_float, _f32p, C.c_int, C.c_float,
                             C.c_float]
    R.ref_revolve.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, C.c_int]
    R.ref_revolve.restype = C.c_int
    R.ref_revolve_numforw.argtypes = [C.c_int, C.c_int]
    R.ref_revolve_numforw.restype = C.c_int
    R.ref_revolve_adjust.argtypes = [C.c_int]
    R.ref_revolve_adjust.restype = C.c_int


def fn(name, real=np.float32):
    """Typed oracle function ``orc_<name>_<f32|f64>``."""
    return getattr(lib(), "orc_" + name + _SUF[real])

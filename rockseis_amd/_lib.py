"""ctypes binding of librsgpu.so (include/rsgpu.h).

The library is loaded on first use from the package directory; a missing
library raises immediately -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "librsgpu.so")

PRESSURE, VX, VY, VZ = 0, 1, 2, 3
KIND_A2D, KIND_E2D, KIND_A3D = 1, 2, 3
MAX_STRIPS = 16

A2D_STRIPS = ("P_left", "P_right", "P_top", "P_bottom", "Vxx_left", "Vxx_right", "Vzz_top",
              "Vzz_bottom")
E2D_STRIPS = ("Sxx_left", "Sxx_right", "Sxzx_left", "Sxzx_right", "Szz_top", "Szz_bottom",
              "Sxzz_top", "Sxzz_bottom", "Vxx_left", "Vxx_right", "Vzx_left", "Vzx_right",
              "Vzz_top", "Vzz_bottom", "Vxz_top", "Vxz_bottom")
A3D_STRIPS = ("P_left", "P_right", "P_front", "P_back", "P_top", "P_bottom", "Vxx_left",
              "Vxx_right", "Vyy_front", "Vyy_back", "Vzz_top", "Vzz_bottom")


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("lpml", C.c_int),
                ("order", C.c_int), ("fs", C.c_int), ("dx", C.c_float), ("dy", C.c_float),
                ("dz", C.c_float), ("dt", C.c_float), ("domdec", C.c_int)]


class State(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in ("P", "Vx", "Vy", "Vz", "Sxx", "Szz", "Sxz",
                                           "L", "L2M", "M", "Rx", "Ry", "Rz", "tab")]
                + [("psi", C.c_void_p * MAX_STRIPS), ("wrk", C.c_void_p)])


class Traces(C.Structure):
    _fields_ = [("field", C.c_int), ("ntrace", C.c_int), ("nt", C.c_int), ("map", C.c_void_p),
                ("shift", C.c_void_p), ("data", C.c_void_p), ("factor", C.c_void_p)]


class Batch(C.Structure):
    _fields_ = [("nshot", C.c_int), ("field_stride", C.c_size_t), ("xstrip_stride", C.c_size_t),
                ("zstrip_stride", C.c_size_t)]


class Fwi(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in ("L", "Rx", "Rz", "tab", "Vp", "Rho")]
                + [("nx", C.c_int), ("nz", C.c_int), ("src", Traces), ("obs", Traces)]
                + [(n, C.c_void_p) for n in ("dmod", "dres", "wavgrad", "vpgrad", "rhograd")]
                + [("misfit", C.POINTER(C.c_double)), ("ncheck", C.c_int), ("mode", C.c_int),
                   ("mode_used", C.POINTER(C.c_int)), ("snapinc", C.c_int), ("fp16", C.c_int),
                   ("misfit_type", C.c_int), ("weight", C.c_void_p), ("adjoint", C.c_int),
                   ("srcilum", C.c_void_p)])


class FwiA3D(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in ("L", "Rx", "Ry", "Rz", "tab", "Vp", "Rho")]
                + [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("src", Traces), ("obs", Traces)]
                + [(n, C.c_void_p) for n in ("dmod", "dres", "wavgrad", "vpgrad", "rhograd")]
                + [("misfit", C.POINTER(C.c_double)), ("ncheck", C.c_int), ("fp16", C.c_int),
                   ("misfit_type", C.c_int), ("weight", C.c_void_p), ("adjoint", C.c_int)])


class FwiE3D(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in ("L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz", "tab", "Vp", "Vs",
                                          "Rho")]
                + [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("src", Traces), ("obs", Traces * 3),
                   ("dmod", C.c_void_p * 3), ("dres", C.c_void_p * 3), ("weight", C.c_void_p * 3)]
                + [(n, C.c_void_p) for n in ("wavgrad", "vpgrad", "vsgrad", "rhograd")]
                + [("misfit", C.POINTER(C.c_double)), ("ncheck", C.c_int), ("fp16", C.c_int),
                   ("misfit_type", C.c_int)])


class RtmA2D(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in ("L", "Rx", "Rz", "tab")]
                + [("nx", C.c_int), ("nz", C.c_int), ("src", Traces), ("data", Traces), ("image", C.c_void_p),
                   ("nhx", C.c_int), ("nhz", C.c_int), ("ncheck", C.c_int)])


class RtmA3D(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in ("L", "Rx", "Ry", "Rz", "tab")]
                + [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("src", Traces), ("data", Traces),
                   ("image", C.c_void_p), ("nhx", C.c_int), ("nhy", C.c_int), ("nhz", C.c_int), ("ncheck", C.c_int)])


class FwiE2DShot(C.Structure):
    _fields_ = ([("g", Grid)] + [(n, C.c_void_p) for n in ("L", "L2M", "M", "Rx", "Rz", "Vp", "Vs", "Rho")]
                + [("ox", C.c_int), ("src", Traces), ("obs", Traces * 3), ("dmod", C.c_void_p * 3),
                   ("dres", C.c_void_p * 3), ("wavgrad", C.c_void_p), ("weight", C.c_void_p * 3)])


class SnapSet(C.Structure):
    _fields_ = [("f", C.c_void_p * 3), ("out", C.c_void_p), ("snapinc", C.c_int)]


class E2DRecords(C.Structure):
    _fields_ = [("nrec", C.c_int), ("map", C.c_void_p), ("mod", C.c_void_p * 3), ("obs", C.c_void_p * 3),
                ("weight", C.c_void_p * 3), ("res", C.c_void_p * 3)]


class FwiE2D(C.Structure):
    _fields_ = [("tab", C.c_void_p), ("ncheck", C.c_int), ("fp16", C.c_int), ("scale", C.c_int),
                ("nx", C.c_int), ("nz", C.c_int), ("grad", C.c_void_p * 3), ("misfit", C.POINTER(C.c_double)),
                ("misfit_type", C.c_int), ("adjoint", C.c_int)]


E3D_NSTRIP = 36
E3D_STRIPS = tuple(f"{n}_{s}" for n, sides in
                   ((n, ("left", "right")) for n in ("Sxx", "Sxzx", "Sxyx", "Vxx", "Vzx", "Vyx")) for s in sides) + \
    tuple(f"{n}_{s}" for n in ("Syy", "Sxyy", "Syzy", "Vyy", "Vxy", "Vzy") for s in ("front", "back")) + \
    tuple(f"{n}_{s}" for n in ("Szz", "Sxzz", "Syzz", "Vzz", "Vxz", "Vyz") for s in ("top", "bottom"))


class StateE3D(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy",
                                          "L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz", "tab")]
                + [("psi", C.c_void_p * E3D_NSTRIP)])


class RssHeader(C.Structure):
    _fields_ = [("data_format", C.c_int), ("header_format", C.c_int), ("type", C.c_int), ("nheader", C.c_size_t),
                ("ndims", C.c_size_t), ("n", C.c_size_t * 9), ("d", C.c_double * 9), ("o", C.c_double * 9),
                ("start", C.c_size_t)]


class Image2DDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("nhx", C.c_int), ("nhz", C.c_int), ("dx", C.c_float),
                ("dz", C.c_float), ("ox", C.c_float), ("oz", C.c_float)]


RSS_GENERIC, RSS_REGULAR, RSS_DATA2D, RSS_DATA3D = 0, 1, 2, 3
FWI_REVOLVE, FWI_STORE, FWI_FULL = 0, 1, 2
ADJOINT_REFERENCE, ADJOINT_EXACT = 0, 1


class RsgError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RsgError(f"librsgpu.so not built ({LIBPATH}); run __graft_entry__.build()")
        # torch ships its own HIP runtime under the soname the library needs
        # (libamdhip64.so.7).  With torch loaded first the library binds to
        # that one runtime; loaded the other way round, the process holds two
        # HIP/HSA runtimes and the library's finds no device.
        import torch  # noqa: F401
        L = C.CDLL(LIBPATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, ip, fp = C.c_void_p, C.c_int, C.c_float
    GP, SP, TP = C.POINTER(Grid), C.POINTER(State), C.POINTER(Traces)
    sig = {
        "rsg_last_error": ([], C.c_char_p),
        "rsg_device_count": ([C.POINTER(C.c_int)], ip),
        "rsg_set_device": ([ip], ip),
        "rsg_stream_sync": ([vp], ip),
        "rsg_malloc": ([C.POINTER(C.c_void_p), C.c_size_t], ip),
        "rsg_free": ([vp], ip),
        "rsg_memcpy_h2d": ([vp, vp, C.c_size_t, vp], ip),
        "rsg_memcpy_d2h": ([vp, vp, C.c_size_t, vp], ip),
        "rsg_memcpy_d2d": ([vp, vp, C.c_size_t, vp], ip),
        "rsg_zero": ([vp, C.c_size_t, vp], ip),
        "rsg_pml_tables": ([ip, fp, fp, ip, vp], ip),
        "rsg_der_coeffs": ([vp], ip),
        "rsg_stagger_a2d": ([vp, vp, ip, ip, ip, ip, vp, vp, vp], ip),
        "rsg_stagger_e2d": ([vp, vp, vp, ip, ip, ip, ip, vp, vp, vp, vp, vp], ip),
        "rsg_stagger_a3d": ([vp, vp, ip, ip, ip, ip, ip, vp, vp, vp, vp], ip),
        "rsg_makemap2d": ([vp, ip, ip, ip, fp, fp, fp, fp, ip, vp, vp], ip),
        "rsg_makemap3d": ([vp, ip, ip, ip, ip, fp, fp, fp, fp, fp, fp, ip, vp, vp], ip),
        "rsg_insert_source": ([ip, GP, SP, TP, ip, vp], ip),
        "rsg_record_data": ([ip, GP, SP, TP, ip, vp], ip),
        "rsg_modelling_run": ([ip, GP, SP, TP, ip, TP, ip, ip, ip, vp], ip),
        "rsg_repeat_step": ([ip, ip, GP, SP, ip, vp], ip),
        "rsg_set_tuning": ([C.c_char_p, ip], ip),
        "rsg_fwi_a2d_gradient": ([GP, C.POINTER(Fwi), vp], ip),
        "rsg_fwi_a2d_gradient_batch": ([GP, C.POINTER(Fwi), ip, vp], ip),
        "rsg_a2d_forward_velocity_batch": ([GP, SP, C.POINTER(Batch), vp], ip),
        "rsg_a2d_forward_stress_batch": ([GP, SP, C.POINTER(Batch), vp], ip),
        "rsg_e2d_velocity_batch": ([GP, SP, C.POINTER(Batch), ip, vp], ip),
        "rsg_e2d_stress_batch": ([GP, SP, C.POINTER(Batch), ip, vp], ip),
        "rsg_revolve_schedule": ([ip, ip, vp, vp, vp, ip, C.POINTER(C.c_int)], ip),
        "rsg_revolve_numforw": ([ip, ip], ip),
        "rsg_snap_cadence": ([ip, C.c_float, ip, C.POINTER(C.c_int), C.POINTER(C.c_int)], ip),
        "rsg_get_tuning": ([C.c_char_p, C.POINTER(C.c_int)], ip),
        "rsg_ragged2d_create": ([ip, ip, GP, SP, C.POINTER(C.c_void_p)], ip),
        "rsg_ragged2d_step": ([vp, ip, vp], ip),
        "rsg_ragged2d_destroy": ([vp], ip),
        "rsg_fwi_e2d_gradient_batch": ([C.POINTER(FwiE2D), C.POINTER(FwiE2DShot), ip, vp], ip),
        "rsg_fwi_e2d_plan_create": ([C.POINTER(FwiE2D), C.POINTER(FwiE2DShot), ip, C.POINTER(C.c_void_p)], ip),
        "rsg_fwi_e2d_plan_run": ([vp, C.POINTER(C.c_double), vp], ip),
        "rsg_fwi_e2d_plan_destroy": ([vp], ip),
        "rsg_fwi_a3d_gradient": ([GP, C.POINTER(FwiA3D), vp], ip),
        "rsg_rtm_a2d_image": ([GP, C.POINTER(RtmA2D), vp], ip),
        "rsg_rtm_a2d_batch": ([GP, C.POINTER(RtmA2D), ip, vp], ip),
        "rsg_rtm_a3d_image": ([GP, C.POINTER(RtmA3D), vp], ip),
        "rsg_interp": ([ip, vp, ip, ip, fp, fp, vp, ip, fp, fp, vp], ip),
        "rsg_misfit_residuals": ([ip, vp, vp, vp, vp, vp, ip, ip, ip, ip, ip, C.POINTER(C.c_double), vp], ip),
        "rsg_misfit_residuals_e2d": ([ip, C.POINTER(E2DRecords), ip, ip, C.POINTER(C.c_double), vp], ip),
        "rsg_modelling_run_snap": ([ip, GP, SP, TP, ip, TP, ip, C.POINTER(SnapSet), ip, ip, ip, vp], ip),
        "rsg_snap_interior": ([vp, vp, vp, GP, ip, vp, vp], ip),
        "rsg_crosscorr_a2d": ([vp, ip, vp, vp, ip, vp, vp, vp, vp, ip, ip, fp, fp, ip, vp, vp, vp, vp], ip),
        "rsg_local_window_2d": ([vp, ip, fp, fp, fp, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)], ip),
        "rsg_domain_model_a2d": ([vp, vp, ip, ip, ip, ip, fp, fp, fp, fp, C.c_longlong, C.c_longlong,
                                  vp, vp, vp, vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float)], ip),
        "rsg_domain_model_e2d": ([vp, vp, vp, ip, ip, ip, ip, fp, fp, fp, fp, C.c_longlong, C.c_longlong,
                                  vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float)], ip),
        "rsg_image_offset": ([fp, fp, fp, fp], ip),
        "rsg_der_apply": ([ip, ip, vp, vp, ip, ip, ip, fp, ip, vp], ip),
        "rsg_revolve_adjust": ([ip], ip),
        "rsg_revolve_init": ([vp, ip, ip, C.POINTER(C.c_int)], ip),
        "rsg_revolve_next": ([vp], ip),
        "rsg_modelling_plan_create": ([ip, GP, SP, TP, ip, TP, ip, C.POINTER(SnapSet), ip, vp,
                                       C.POINTER(C.c_void_p)], ip),
        "rsg_modelling_plan_run": ([vp, ip, ip, vp], ip),
        "rsg_modelling_plan_destroy": ([vp], ip),
        "rsg_stagger_a2d_dev": ([vp, vp, ip, ip, ip, ip, ip, vp, vp, vp, vp], ip),
        "rsg_stagger_e2d_dev": ([vp, vp, vp, ip, ip, ip, ip, ip, vp, vp, vp, vp, vp, vp], ip),
        "rsg_misfit_residuals_noalloc": ([ip, vp, vp, vp, vp, vp, ip, ip, ip, ip, ip, C.POINTER(C.c_double), vp],
                                         ip),
        "rsg_stagger_e3d": ([vp, vp, vp, ip, ip, ip, ip, ip] + [vp] * 8, ip),
        "rsg_e3d_forward_velocity": ([GP, C.POINTER(StateE3D), vp], ip),
        "rsg_e3d_forward_stress": ([GP, C.POINTER(StateE3D), vp], ip),
        "rsg_e3d_backward_velocity": ([GP, C.POINTER(StateE3D), vp], ip),
        "rsg_e3d_backward_stress": ([GP, C.POINTER(StateE3D), vp], ip),
        "rsg_e3d_backward_velocity_ws": ([GP, C.POINTER(StateE3D), vp, vp], ip),
        "rsg_fwi_e3d_gradient": ([GP, C.POINTER(FwiE3D), vp], ip),
        "rsg_e3d_insert_source": ([GP, C.POINTER(StateE3D), TP, ip, vp], ip),
        "rsg_e3d_record_data": ([GP, C.POINTER(StateE3D), TP, vp, vp, vp, ip, vp], ip),
        "rsg_rss_read_header": ([C.c_char_p, C.POINTER(RssHeader)], ip),
        "rsg_rss_read_data2d": ([C.c_char_p, vp, vp, ip, ip, vp], ip),
        "rsg_rss_read_data3d": ([C.c_char_p, vp, vp, ip, ip, vp], ip),
        "rsg_rss_write_data2d": ([C.c_char_p, vp, vp, ip, ip, fp, fp, vp], ip),
        "rsg_rss_write_data3d": ([C.c_char_p, vp, vp, ip, ip, fp, vp], ip),
        "rsg_rss_read_regular": ([C.c_char_p, vp, C.c_size_t, vp], ip),
        "rsg_rss_write_regular": ([C.c_char_p, C.POINTER(RssHeader), vp, vp], ip),
        "rsg_stack_image_2d": ([vp, C.POINTER(Image2DDesc), vp, C.POINTER(Image2DDesc), ip, ip, ip, ip, vp], ip),
    }
    for name in ("a2d_forward_velocity", "a2d_forward_stress", "e2d_forward_velocity",
                 "e2d_forward_stress", "e2d_backward_velocity", "e2d_backward_stress",
                 "a3d_forward_velocity", "a3d_forward_stress"):
        sig["rsg_" + name] = ([GP, SP, vp], ip)
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


# Every symbol include/rsgpu.h declares (checked by tests/test_abi.py).
EXPORTS = (
    "rsg_last_error", "rsg_device_count", "rsg_set_device", "rsg_stream_sync", "rsg_malloc",
    "rsg_free", "rsg_memcpy_h2d", "rsg_memcpy_d2h", "rsg_memcpy_d2d", "rsg_zero",
    "rsg_pml_tables", "rsg_der_coeffs", "rsg_stagger_a2d", "rsg_stagger_e2d", "rsg_stagger_a3d",
    "rsg_makemap2d", "rsg_makemap3d", "rsg_a2d_forward_velocity", "rsg_a2d_forward_stress",
    "rsg_e2d_forward_velocity", "rsg_e2d_forward_stress", "rsg_e2d_backward_velocity",
    "rsg_e2d_backward_stress", "rsg_a3d_forward_velocity", "rsg_a3d_forward_stress",
    "rsg_insert_source", "rsg_record_data", "rsg_modelling_run", "rsg_repeat_step",
    "rsg_set_tuning", "rsg_get_tuning", "rsg_fwi_a2d_gradient", "rsg_revolve_schedule",
    "rsg_revolve_numforw", "rsg_snap_cadence", "rsg_fwi_a2d_gradient_batch", "rsg_a2d_forward_velocity_batch",
    "rsg_a2d_forward_stress_batch", "rsg_e2d_velocity_batch", "rsg_e2d_stress_batch",
    "rsg_ragged2d_create", "rsg_ragged2d_step", "rsg_ragged2d_destroy", "rsg_fwi_e2d_gradient_batch",
    "rsg_fwi_a3d_gradient", "rsg_rtm_a2d_image", "rsg_rtm_a2d_batch", "rsg_rtm_a3d_image", "rsg_interp", "rsg_misfit_residuals", "rsg_misfit_residuals_e2d",
    "rsg_fwi_e2d_plan_create", "rsg_fwi_e2d_plan_run", "rsg_fwi_e2d_plan_destroy",
    "rsg_modelling_run_snap", "rsg_snap_interior", "rsg_crosscorr_a2d", "rsg_local_window_2d",
    "rsg_domain_model_a2d", "rsg_domain_model_e2d", "rsg_image_offset", "rsg_der_apply",
    "rsg_revolve_adjust", "rsg_revolve_init", "rsg_revolve_next",
    "rsg_modelling_plan_create", "rsg_modelling_plan_run", "rsg_modelling_plan_destroy",
    "rsg_stagger_a2d_dev", "rsg_stagger_e2d_dev", "rsg_misfit_residuals_noalloc",
    "rsg_rss_read_header", "rsg_rss_read_data2d", "rsg_rss_read_data3d", "rsg_rss_write_data2d",
    "rsg_rss_write_data3d", "rsg_rss_read_regular", "rsg_rss_write_regular", "rsg_stack_image_2d",
    "rsg_stagger_e3d", "rsg_e3d_forward_velocity", "rsg_e3d_forward_stress", "rsg_e3d_backward_velocity",
    "rsg_e3d_backward_stress", "rsg_e3d_backward_velocity_ws", "rsg_fwi_e3d_gradient", "rsg_e3d_insert_source",
    "rsg_e3d_record_data",
)


def set_tuning(key: str, value: int):
    check(lib().rsg_set_tuning(key.encode(), int(value)), f"rsg_set_tuning({key})")


def get_tuning(key: str) -> int:
    v = C.c_int()
    check(lib().rsg_get_tuning(key.encode(), C.byref(v)), f"rsg_get_tuning({key})")
    return v.value


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().rsg_last_error()
        raise RsgError(f"{what}: rsgpu error {rc}: {msg.decode() if msg else ''}")

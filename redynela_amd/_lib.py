"""ctypes binding of libdnlgpu.so (the C ABI declared in include/dnlg.h).

The library is loaded from this package directory only; there is no CPU
fallback: if it is missing, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdnlgpu.so")

DNLG_OK = 0
ERRORS = {1: "DNLG_E_ARG", 2: "DNLG_E_POLAR_NOCONV", 3: "DNLG_E_RR_NOCONV", 4: "DNLG_E_NEG_DETJ",
          5: "DNLG_E_HIP", 6: "DNLG_E_STATE", 7: "DNLG_E_PULSE_NOCONV"}

F_COORDS, F_V, F_A, F_DU, F_D, F_FINT, F_MASS, F_S, F_EVP, F_P, F_EPS, F_EPSPLAS, F_FE, F_TEMP, F_DEPS = range(15)
OPT_KINETIC_ENERGY, OPT_STORE_FINT, OPT_TRACK_DISP, OPT_TRACK_STRAIN, OPT_TIME_STEP_METHOD = 1, 2, 3, 4, 5
OPT_PLOT_HISTORY = 6
TS_COURANT, TS_PULSE = 0, 1
# Node::getNodalValue field names -> DNLG_HF_* codes; True for vector fields
HF_FIELDS = {"density": (0, False), "densityInc": (1, False), "energy": (2, False), "energyInc": (3, False),
             "speed": (4, True), "speedInc": (5, True), "force": (6, True), "temperature": (7, False),
             "displacementInc": (8, True), "displacement": (9, True), "coords": (10, True), "mass": (11, False),
             "temperatureInit": (12, False),
             # element fields: components 0 (equivalent) or 11 .. 33
             "stress": (13, "tensor"), "strain": (14, "tensor"), "strainInc": (15, "tensor"),
             "plasticStrain": (16, "tensor")}

# every symbol include/dnlg.h declares
EXPORTS = [
    "dnlg_abi_version", "dnlg_create", "dnlg_destroy", "dnlg_set_option", "dnlg_compute_internal_forces",
    "dnlg_begin", "dnlg_step", "dnlg_solve", "dnlg_step_async", "dnlg_synchronize", "dnlg_get_field",
    "dnlg_set_field", "dnlg_field_size", "dnlg_get_time", "dnlg_set_time", "dnlg_pulse_info", "dnlg_plot_rows", "dnlg_history_file_add", "dnlg_history_file_rows", "dnlg_get_stream", "dnlg_enable_kernel_timing",
    "dnlg_kernel_stats", "dnlg_last_error", "dnlg_mesh_axi_bar", "dnlg_mesh_ogrid_sizes", "dnlg_mesh_ogrid_bar",
    "dnlg_batch_create", "dnlg_batch_destroy", "dnlg_batch_solve", "dnlg_batch_get_time", "dnlg_batch_get_field",
    "dnlg_batch_last_error", "dnlg_batch_taylor_params", "dnlg_batch_set_engine", "dnlg_batch_get_tuning",
    "dnlg_adjoint_forward", "dnlg_adjoint_backward", "dnlg_compute_internal_forces_from", "dnlg_courant_time_step",
    "dnlg_courant_time_step_noalloc", "dnlg_noalloc_last_error",
]
BATCH_AUTO, BATCH_PERSISTENT, BATCH_STEPPED, BATCH_STEPPED_PLAIN, BATCH_STEPPED_WIDE = 0, 1, 2, 3, 4


class Mesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_nodes", C.c_longlong), ("n_elems", C.c_longlong),
                ("coords", C.c_void_p), ("conn", C.c_void_p)]


class Material(C.Structure):
    _fields_ = [("young", C.c_double), ("poisson", C.c_double), ("density", C.c_double),
                ("A", C.c_double), ("B", C.c_double), ("n", C.c_double), ("hardening", C.c_int),
                ("C", C.c_double), ("epsd0", C.c_double), ("m", C.c_double), ("T0", C.c_double),
                ("Tm", C.c_double), ("cp", C.c_double), ("taylor_quinney", C.c_double)]


class BC(C.Structure):
    _fields_ = [("restrain", C.c_void_p), ("v0", C.c_void_p), ("speed_mask", C.c_void_p), ("speed", C.c_void_p)]


class SolverParams(C.Structure):
    _fields_ = [("rho_b", C.c_double), ("safety", C.c_double), ("ts_freq", C.c_int), ("end_time", C.c_double),
                ("form", C.c_int)]


FORM_FP64, FORM_MIXED = 0, 1


class AdjointStats(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("backward_ms", C.c_double), ("recompute_ms", C.c_double),
                ("device_bytes", C.c_longlong), ("checkpoint_bytes", C.c_longlong), ("n_checkpoints", C.c_int),
                ("steps", C.c_longlong)]


class DnlgError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DnlgError("libdnlgpu.so is not built: run `python -m redynela_amd.build` "
                            "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, i, ll, d = C.c_void_p, C.c_int, C.c_longlong, C.c_double
        sig = {
            "dnlg_abi_version": (i, []),
            "dnlg_create": (i, [C.POINTER(Mesh), C.POINTER(Material), C.POINTER(BC), C.POINTER(SolverParams),
                                C.POINTER(vp)]),
            "dnlg_destroy": (i, [vp]),
            "dnlg_set_option": (i, [vp, i, i]),
            "dnlg_compute_internal_forces": (i, [vp, vp]),
            "dnlg_begin": (i, [vp, d, C.POINTER(i)]),
            "dnlg_step": (i, [vp, i, C.POINTER(i)]),
            "dnlg_solve": (i, [vp, d, C.POINTER(ll)]),
            "dnlg_step_async": (i, [vp, i]),
            "dnlg_synchronize": (i, [vp]),
            "dnlg_get_field": (i, [vp, i, vp, ll]),
            "dnlg_set_field": (i, [vp, i, vp, ll]),
            "dnlg_field_size": (ll, [vp, i]),
            "dnlg_get_time": (i, [vp, vp]),
            "dnlg_set_time": (i, [vp, d]),
            "dnlg_pulse_info": (i, [vp, vp]),
            "dnlg_plot_rows": (ll, [vp, vp, ll]),
            "dnlg_history_file_add": (i, [vp, d, d, d, i, vp, vp, vp, d, C.POINTER(i)]),
            "dnlg_history_file_rows": (ll, [vp, i, vp, ll]),
            "dnlg_get_stream": (vp, [vp]),
            "dnlg_enable_kernel_timing": (i, [vp, i]),
            "dnlg_kernel_stats": (i, [vp, i, C.POINTER(ll), C.POINTER(d)]),
            "dnlg_last_error": (C.c_char_p, [vp]),
            "dnlg_mesh_axi_bar": (i, [i, i, d, d, d, vp, vp, vp, vp]),
            "dnlg_mesh_ogrid_sizes": (i, [i, i, i, C.POINTER(ll), C.POINTER(ll)]),
            "dnlg_mesh_ogrid_bar": (i, [i, i, i, d, d, d, d, vp, vp, vp, vp]),
            "dnlg_batch_create": (i, [i, vp, vp, vp, C.POINTER(SolverParams), C.POINTER(vp)]),
            "dnlg_batch_destroy": (i, [vp]),
            "dnlg_batch_solve": (i, [vp, d, ll, vp, C.POINTER(d)]),
            "dnlg_batch_get_time": (i, [vp, i, vp]),
            "dnlg_batch_get_field": (i, [vp, i, i, vp, ll]),
            "dnlg_batch_last_error": (C.c_char_p, [vp]),
            "dnlg_batch_taylor_params": (i, [i, C.c_ulonglong, vp, vp, vp, vp]),
            "dnlg_batch_set_engine": (i, [vp, i]),
            "dnlg_batch_get_tuning": (i, [vp, C.POINTER(i), vp]),
            "dnlg_adjoint_forward": (i, [vp, ll, i]),
            "dnlg_adjoint_backward": (i, [vp, vp, vp, C.POINTER(AdjointStats)]),
            "dnlg_compute_internal_forces_from": (i, [vp, vp, vp, vp]),
            "dnlg_courant_time_step": (i, [vp, C.POINTER(d)]),
            "dnlg_courant_time_step_noalloc": (i, [i, ll, vp, vp, C.POINTER(Material), vp, vp]),
            "dnlg_noalloc_last_error": (C.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, h=None):
    if rc != DNLG_OK:
        msg = lib().dnlg_last_error(h)
        raise DnlgError("%s: %s" % (ERRORS.get(rc, rc), msg.decode() if msg else ""))


def ptr(a):
    return a.ctypes.data if a is not None else None

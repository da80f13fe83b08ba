"""ctypes mirror of include/sw2_hip.h and the loader of libsw2.so.

The structures below follow the C declarations field by field; the library
is loaded from this package's directory only, and a missing library is an
error (there is no fallback implementation).
"""
import ctypes
import os
from ctypes import POINTER, c_char, c_double, c_int, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsw2.so")

P = POINTER(c_double)


MAXP = 1920  # per-process table slots of the sppde map (sppde/sppde.h:44)


class Map(ctypes.Structure):
    """sppde map (sppde/sppde.h:46-67), one process."""

    _fields_ = [
        ("sh", c_int), ("nd", c_int),
        ("l0", c_int * 4), ("l1", c_int * 4), ("gl0", c_int * 4), ("gl1", c_int * 4),
        ("al0", (c_int * 4) * MAXP), ("al1", (c_int * 4) * MAXP), ("pco", (c_int * 4) * MAXP),
        ("s", c_int * 4), ("S", c_int), ("gpc", c_int * 4), ("pc", c_int * 4),
        ("nb_e", c_int), ("nb_w", c_int), ("nb_n", c_int), ("nb_s", c_int), ("nb_t", c_int), ("nb_b", c_int),
        ("bc0", c_int * 4), ("bc1", c_int * 4), ("per", c_int * 4),
    ]


_FIELD_PTRS = ["windx", "windy", "u", "v", "eta", "hB", "tracer", "perturbs",
               "xe", "xc", "xv", "yn", "yc", "yv", "Dxe", "Dxc", "Dxv", "Dyn", "Dyc", "Dyv",
               "ge", "gn", "gc", "gv", "fe", "fn", "fc", "fv"]


class SW(ctypes.Structure):
    """The world struct (sw.h:22-95)."""

    _fields_ = [
        ("sch", c_int), ("polar", c_int), ("N", c_int), ("loadFrom", c_int), ("saveEvery", c_int),
        ("IteInfo", c_int), ("prn", c_int),
        ("t0", c_double), ("t1", c_double), ("Dt", c_double), ("Courant_max", c_double),
        ("nx", c_int), ("ny", c_int), ("nz", c_int),
        ("lon0", c_double), ("lat0", c_double), ("lon1", c_double), ("lat1", c_double),
        ("m", Map), ("M", POINTER(Map)),
        ("Omega", c_double), ("rE", c_double), ("rP", c_double), ("epsilon2", c_double), ("gE", c_double),
        ("depth", c_double), ("sigma_dis", c_double), ("tracer_dis", c_double),
        ("beta_sponge", c_double), ("sigma_sponge", c_double), ("tau_sponge", c_double),
    ] + [(n, P) for n in _FIELD_PTRS] + [
        ("nu2", c_double), ("nu4", c_double), ("nu6", c_double),
        ("reftype", c_int), ("refname", c_char * 1000), ("reftable", P), ("reftable_nr", c_int), ("reftable_nc", c_int),
        ("geoeq", c_int), ("num_Gaussians", c_int), ("Gaussians", P), ("num_vortices", c_int), ("vortices", P),
    ]


class Params(ctypes.Structure):
    """Case parameters as written in a case file (include/sw2_hip.h)."""

    _fields_ = [
        ("sch", c_int), ("polar", c_int),
        ("lon0_deg", c_double), ("lon1_deg", c_double), ("lat0_deg", c_double), ("lat1_deg", c_double),
        ("nx", c_int), ("ny", c_int),
        ("t0", c_double), ("t1", c_double), ("Dt", c_double), ("CFL", c_double), ("Tau", c_double),
        ("hypernu", c_double * 3),
        ("IteInfo", c_int), ("SaveEvery", c_int), ("LoadFrom", c_int), ("reftype", c_int),
        ("Omega", c_double), ("rE", c_double), ("rP", c_double), ("gE", c_double), ("depth", c_double),
        ("geoeq", c_int),
        ("num_vortices", c_int), ("vortices", P),
        ("num_Gaussians", c_int), ("Gaussians", P),
        ("TracerDis", c_double),
        ("wind_rows", c_int), ("wind", P),
    ]


# symbols declared in include/sw2_hip.h (checked by tests/test_abi.py)
EXPORTS = [
    "initialize", "loop", "destroy",
    "checkr", "crash", "init_pprintf", "end_pprintf", "pprintf", "quants", "quisoc", "dmem",
    "cr_reset", "cr_start", "cr_end", "cr_time", "cr_info",
    "sw2_init_params", "sw2_read_case", "sw2_free_params", "sw2_eval_expr",
    "halo_update", "do_channel", "calc_energy_balance", "sw2_mass",
    "solve_tracer", "solve_cons_h", "dissipation_effect", "do_eta_geofactor", "solve_adv_u", "solve_adv_v",
    "solve_pres_u", "solve_pres_v", "solve_Coriolis_u", "solve_Coriolis_v", "do_AdamsBashforth", "compute_Dt",
    "laplacian_vels", "laplacian_eta",
    "save_binary_sw", "load_binary_sw",
    "sw2_session_create", "sw2_session_begin", "sw2_session_step", "sw2_session_time", "sw2_session_kernel_times",
    "sw2_session_energy", "sw2_session_energy_noalloc", "sw2_session_download", "sw2_session_form",
    "sw2_session_failed_step", "sw2_session_last_cfl", "sw2_session_fused", "sw2_session_state", "sw2_session_destroy",
    "sw2_device_cus", "sw2_batch_footprint", "sw2_batch_create", "sw2_batch_add", "sw2_batch_set_launch", "sw2_batch_get_launch",
    "sw2_batch_set_tiles", "sw2_batch_get_tiles",
    "sw2_batch_step", "sw2_batch_sync", "sw2_batch_pause", "sw2_batch_member", "sw2_batch_finished", "sw2_batch_download",
    "sw2_batch_energy", "sw2_batch_remove", "sw2_batch_count", "sw2_batch_destroy",
    "sw2_last_error", "sw2_build_info",
]

_lib = None


def lib():
    """Load libsw2.so from this package; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C sw2_amd` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    L.sw2_init_params.argtypes = [POINTER(Params), POINTER(SW)]
    L.sw2_read_case.argtypes = [ctypes.c_char_p, POINTER(Params)]
    L.sw2_free_params.argtypes = [POINTER(Params)]
    L.sw2_eval_expr.argtypes = [ctypes.c_char_p, POINTER(c_int)]
    L.sw2_eval_expr.restype = c_double
    L.halo_update.argtypes = [P, POINTER(Map)]
    L.do_channel.argtypes = [POINTER(SW)]
    L.calc_energy_balance.argtypes = [P, P, POINTER(SW)]
    L.sw2_mass.argtypes = [POINTER(SW)]
    L.sw2_mass.restype = c_double
    for name in ("solve_tracer", "solve_cons_h", "do_eta_geofactor", "solve_adv_u", "solve_adv_v", "solve_pres_u",
                 "solve_pres_v", "solve_Coriolis_u", "solve_Coriolis_v"):
        getattr(L, name).argtypes = [P, c_double, POINTER(SW)]
        getattr(L, name).restype = None
    L.dissipation_effect.argtypes = [P, P, POINTER(SW)]
    L.dissipation_effect.restype = None
    L.do_AdamsBashforth.argtypes = [P, P, P, P, c_int, POINTER(Map)]
    L.do_AdamsBashforth.restype = None
    L.compute_Dt.argtypes = [c_double, POINTER(SW)]
    L.compute_Dt.restype = c_double
    L.laplacian_vels.argtypes = [P, P, P, P, POINTER(SW)]
    L.laplacian_vels.restype = None
    L.laplacian_eta.argtypes = [P, P, POINTER(SW)]
    L.laplacian_eta.restype = None
    L.destroy.argtypes = [POINTER(SW)]
    L.save_binary_sw.argtypes = [c_int, c_int, c_double, ctypes.c_char_p, ctypes.c_char_p, POINTER(SW)]
    L.load_binary_sw.argtypes = [c_int, POINTER(c_int), POINTER(c_double), ctypes.c_char_p, ctypes.c_char_p, POINTER(SW)]
    L.sw2_session_create.argtypes = [POINTER(SW), c_int, POINTER(c_void_p)]
    L.sw2_session_begin.argtypes = [c_void_p, c_int, c_double]
    L.sw2_session_step.argtypes = [c_void_p, c_int]
    L.sw2_session_time.argtypes = [c_void_p, c_int, P]
    L.sw2_session_kernel_times.argtypes = [c_void_p, c_int, P, c_int]
    L.sw2_session_download.argtypes = [c_void_p, POINTER(SW)]
    L.sw2_session_energy.argtypes = [c_void_p, P, P, P]
    L.sw2_session_energy_noalloc.argtypes = [c_void_p, c_void_p, c_void_p]
    L.sw2_session_form.argtypes = [c_void_p]
    L.sw2_session_failed_step.argtypes = [c_void_p, ctypes.c_char_p, c_int]
    L.sw2_session_last_cfl.argtypes = [c_void_p, P]
    L.sw2_session_fused.argtypes = [c_void_p]
    L.sw2_session_state.argtypes = [c_void_p, POINTER(c_int), P]
    L.sw2_session_destroy.argtypes = [c_void_p]
    PI = POINTER(c_int)
    L.sw2_device_cus.argtypes = []
    L.sw2_device_cus.restype = c_int
    L.sw2_batch_footprint.argtypes = [POINTER(SW), c_int]
    L.sw2_batch_footprint.restype = ctypes.c_size_t
    L.sw2_batch_create.argtypes = [c_int, c_int, c_double, POINTER(c_void_p)]
    L.sw2_batch_add.argtypes = [c_void_p, POINTER(SW), c_int, PI]
    L.sw2_batch_set_launch.argtypes = [c_void_p, c_int, c_int, c_int]
    L.sw2_batch_get_launch.argtypes = [c_void_p, PI, PI, PI]
    L.sw2_batch_set_tiles.argtypes = [c_void_p, c_int, c_int]
    L.sw2_batch_get_tiles.argtypes = [c_void_p, PI, PI]
    L.sw2_batch_step.argtypes = [c_void_p, c_int]
    L.sw2_batch_sync.argtypes = [c_void_p]
    L.sw2_batch_pause.argtypes = [c_void_p, c_int, c_int]
    L.sw2_batch_member.argtypes = [c_void_p, c_int, PI, PI, PI]
    L.sw2_batch_finished.argtypes = [c_void_p, PI, c_int]
    L.sw2_batch_download.argtypes = [c_void_p, c_int, POINTER(SW)]
    L.sw2_batch_energy.argtypes = [c_void_p, c_int, P, P, P]
    L.sw2_batch_remove.argtypes = [c_void_p, c_int]
    L.sw2_batch_count.argtypes = [c_void_p, POINTER(ctypes.c_size_t)]
    L.sw2_batch_destroy.argtypes = [c_void_p]
    L.sw2_batch_destroy.restype = None
    L.sw2_last_error.restype = ctypes.c_char_p
    L.sw2_build_info.restype = ctypes.c_char_p
    _lib = L
    return L


def last_error():
    return lib().sw2_last_error().decode(errors="replace")


def as_array(ptr, n):
    """numpy view of n doubles behind a ctypes pointer (no copy)."""
    return np.ctypeslib.as_array(ptr, shape=(n,))

"""Python binding of the oracle (oracle/sw2_oracle.c).

TEST INFRASTRUCTURE ONLY: the CPU restatement of SW2's FP64 step that the
GPU library is checked against.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module; the product never does.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsw2oracle.so")
FIELD_ID = {"eta": 0, "u": 1, "v": 2, "tracer": 3, "perturbs": 4, "windx": 5, "windy": 6, "hB": 7}
ROW_ID = {n: k for k, n in enumerate(["Dxe", "Dxc", "Dxv", "Dyn", "Dyc", "Dyv", "ge", "gn", "gc", "gv", "fe", "fn", "fc", "fv"])}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.argtypes = [c_void_p]
        L.orc_create.restype = c_void_p
        L.orc_free.argtypes = [c_void_p]
        L.orc_field.argtypes = [c_void_p, c_int]
        L.orc_field.restype = POINTER(c_double)
        L.orc_rows.argtypes = [c_void_p, c_int]
        L.orc_rows.restype = POINTER(c_double)
        L.orc_size.argtypes = [c_void_p]
        L.orc_set_threads.argtypes = [c_void_p, c_int]
        L.orc_set_threads.restype = None
        for name in ("orc_solve_tracer", "orc_solve_cons_h", "orc_do_eta_geofactor", "orc_solve_adv_u", "orc_solve_adv_v",
                     "orc_solve_pres_u", "orc_solve_pres_v", "orc_solve_coriolis_u", "orc_solve_coriolis_v"):
            getattr(L, name).argtypes = [c_void_p, POINTER(c_double), c_double]
            getattr(L, name).restype = None
        L.orc_dissipation_effect.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double)]
        L.orc_dissipation_effect.restype = None
        L.orc_adams_bashforth.argtypes = [c_void_p] + [POINTER(c_double)] * 4 + [c_int]
        L.orc_adams_bashforth.restype = None
        L.orc_compute_dt.argtypes = [c_void_p, c_double]
        L.orc_compute_dt.restype = c_double
        L.orc_laplacian_vels.argtypes = [c_void_p] + [POINTER(c_double)] * 4
        L.orc_laplacian_vels.restype = None
        L.orc_laplacian_eta.argtypes = [c_void_p] + [POINTER(c_double)] * 2
        L.orc_laplacian_eta.restype = None
        L.orc_nsteps.argtypes = [c_void_p]
        L.orc_time.argtypes = [c_void_p]
        L.orc_time.restype = c_double
        L.orc_step_index.argtypes = [c_void_p]
        L.orc_failmsg.argtypes = [c_void_p]
        L.orc_failmsg.restype = ctypes.c_char_p
        L.orc_last_dtcfl.argtypes = [c_void_p]
        L.orc_last_dtcfl.restype = c_double
        L.orc_preloop.argtypes = [c_void_p]
        L.orc_after_load.argtypes = [c_void_p, c_int, c_double]
        L.orc_begin.argtypes = [c_void_p, c_int, c_double]
        L.orc_energy.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double)]
        L.orc_mass.argtypes = [c_void_p]
        L.orc_mass.restype = c_double
        L.orc_step.argtypes = [c_void_p, c_int, POINTER(c_double), POINTER(c_int), c_int]
        L.orc_stats.argtypes = [c_void_p, c_int, POINTER(c_double)]
        L.orc_psi.argtypes = [c_int, c_double]
        L.orc_psi.restype = c_double
        _lib = L
    return _lib


class Oracle:
    """One single-rank SW2 run on the CPU, from sw2_amd Params."""

    def __init__(self, params):
        self.params = params
        self.h = lib().orc_create(ctypes.addressof(params))
        if not self.h:
            raise RuntimeError("oracle: unsupported case (polar or wind table out of range)")
        self.nx, self.ny = params.nx, params.ny
        self.S = lib().orc_size(self.h)

    def field(self, name):
        p = lib().orc_field(self.h, FIELD_ID[name])
        return np.ctypeslib.as_array(p, shape=(self.S,)).reshape(self.ny + 4, self.nx + 4)

    def interior(self, name):
        return self.field(name)[2:-2, 2:-2]

    def rows(self, name):
        p = lib().orc_rows(self.h, ROW_ID[name])
        return np.ctypeslib.as_array(p, shape=(self.ny + 4,)).copy()

    def set_threads(self, n):
        """Sweep rows on n threads (same results bit for bit; the timed CPU
        baseline uses this)."""
        lib().orc_set_threads(self.h, int(n))

    def call(self, name, *args):
        """One of SW2's per-step functions on the oracle's fields
        (orc_<name>); numpy arrays are passed as pointers."""
        conv = [a.ctypes.data_as(POINTER(c_double)) if isinstance(a, np.ndarray) else a for a in args]
        return getattr(lib(), "orc_" + name)(self.h, *conv)

    def preloop(self):
        if lib().orc_preloop(self.h):
            raise RuntimeError(self.failmsg())

    def load(self, eta, u, v, tracer=None, perturbs=None, nstep=0, t=None):
        """Pre-loop, then the state a frame load leaves (LoadFrom)."""
        self.preloop()
        self.interior("eta")[:] = eta
        self.interior("u")[:] = u
        self.interior("v")[:] = v
        self.interior("tracer")[:] = 0.0 if tracer is None else tracer
        self.interior("perturbs")[:] = 0.0 if perturbs is None else perturbs
        lib().orc_after_load(self.h, nstep, self.params.t0 if t is None else t)

    def begin(self, nstart=0, t=None):
        lib().orc_begin(self.h, nstart, self.params.t0 if t is None else t)

    def step(self, n, energy=False, maxrec=100000):
        """Run n steps; returns (failed_step, records) with records an array of
        (t, KE, APE) at every step SW2 prints ENERGY."""
        if energy:
            rec = (c_double * (3 * maxrec))()
            nrec = c_int(0)
            f = lib().orc_step(self.h, n, rec, byref(nrec), maxrec)
            return f, np.array(rec[: 3 * nrec.value]).reshape(-1, 3)
        return lib().orc_step(self.h, n, None, None, 0), None

    def energy(self):
        ke, ape = c_double(), c_double()
        lib().orc_energy(self.h, byref(ke), byref(ape))
        return ke.value, ape.value

    def mass(self):
        return lib().orc_mass(self.h)

    def stats(self, name):
        out = (c_double * 3)()
        bad = lib().orc_stats(self.h, FIELD_ID[name], out)
        return bad, tuple(out)

    def failmsg(self):
        return lib().orc_failmsg(self.h).decode(errors="replace")

    @property
    def time(self):
        return lib().orc_time(self.h)

    @property
    def nsteps(self):
        return lib().orc_nsteps(self.h)

    def last_dtcfl(self):
        return lib().orc_last_dtcfl(self.h)

    def close(self):
        if self.h:
            lib().orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def psi(sch, r):
    return lib().orc_psi(sch, r)

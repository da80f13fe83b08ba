"""Python binding of SW2 itself, built from the reference's sources into
oracle/_ref (make -C oracle ref; see ref_harness.c).

TEST INFRASTRUCTURE ONLY: tests/ use it to check the oracle against the real
SW2, and bench.py's cpu_baseline leg times its run driver (sw2ref) on the
host cores.  The product never imports it.

SW2 is not re-entrant (global output handle, timer table, the comparator's
static file) and aborts the process on any error (CRASH -> MPI_Abort), so a
caller runs one world at a time and only cases SW2 runs to the end.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
LIB_PATH = os.path.join(REF_DIR, "libsw2ref.so")
RUN_PATH = os.path.join(REF_DIR, "sw2ref")
FIELD_ID = {"eta": 0, "u": 1, "v": 2, "tracer": 3, "perturbs": 4, "windx": 5, "windy": 6, "hB": 7}
ROW_ID = {n: 10 + k for k, n in enumerate(["Dxe", "Dxc", "Dxv", "Dyn", "Dyc", "Dyv", "ge", "gn", "gc", "gv", "fe", "fn",
                                             "fc", "fv"])}
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib(outdir):
    """The library, started once per process with SW2's output in outdir."""
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle ref` where the reference sources are")
        L = ctypes.CDLL(LIB_PATH)
        L.ref_start.argtypes = [c_char_p]
        L.ref_world.argtypes = [c_int, c_int, c_char_p]
        L.ref_world.restype = c_void_p
        L.ref_loop.argtypes = [c_void_p, c_char_p, c_char_p]
        L.ref_loop.restype = None
        L.ref_free.argtypes = [c_void_p]
        L.ref_free.restype = None
        L.ref_field.argtypes = [c_void_p, c_int]
        L.ref_field.restype = POINTER(c_double)
        L.ref_size.argtypes = [c_void_p]
        L.ref_nsteps.argtypes = [c_void_p]
        L.ref_linterp.argtypes = [c_double, POINTER(c_double), POINTER(c_double), c_int]
        L.ref_linterp.restype = c_double
        P = POINTER(c_double)
        for name in ("solve_tracer", "solve_cons_h", "do_eta_geofactor", "solve_adv_u", "solve_adv_v", "solve_pres_u",
                     "solve_pres_v", "solve_Coriolis_u", "solve_Coriolis_v"):
            getattr(L, name).argtypes = [P, c_double, c_void_p]
            getattr(L, name).restype = None
        L.dissipation_effect.argtypes = [P, P, c_void_p]
        L.dissipation_effect.restype = None
        L.laplacian_vels.argtypes = [P, P, P, P, c_void_p]
        L.laplacian_vels.restype = None
        L.laplacian_eta.argtypes = [P, P, c_void_p]
        L.laplacian_eta.restype = None
        L.compute_Dt.argtypes = [c_double, c_void_p]
        L.compute_Dt.restype = c_double
        L.add_vortices.argtypes = [c_double, c_double, c_void_p]
        L.add_vortices.restype = None
        L.add_Gaussians.argtypes = [c_double, c_double, c_void_p]
        L.add_Gaussians.restype = None
        L.calc_energy_balance.argtypes = [P, P, c_void_p]
        L.calc_energy_balance.restype = None
        L.halo_update.argtypes = [P, c_void_p]
        L.halo_update.restype = None
        L.ref_start(os.fsencode(outdir))
        _lib = L
    return _lib


def write_case(directory, bname, params, state):
    """<directory>/<bname>.sw without m4 constructs and the initial frame
    <directory>/<bname>.0001.bin (LoadFrom 1 reads it from the output
    folder, which is `directory` here)."""
    from sw2_amd import cases
    os.makedirs(directory, exist_ok=True)
    txt, _ = cases.case_text(params, bname, flat=True)
    path = os.path.join(directory, bname + ".sw")
    with open(path, "w") as f:
        f.write(txt)
    eta, u, v = state
    z = np.zeros_like(eta)
    cases.write_frame(os.path.join(directory, bname + ".0001.bin"), params.nx, params.ny, [eta, u, v, z, z])
    return path


class Ref:
    """One single-rank SW2 world built by SW2's own code from Params and an
    initial state (eta, u, v interiors), run with SW2's loop()."""

    def __init__(self, params, state, directory, bname="refcase"):
        self.params, self.dir, self.bname = params, str(directory), bname
        self.case = write_case(self.dir, bname, params, state)
        self.L = lib(self.dir)
        self.h = self.L.ref_world(1, 1, os.fsencode(self.case))
        self.nx, self.ny = params.nx, params.ny
        # createM sizes fields for a third axis too (sppde_core.c:542, with
        # gl1[3] = nz = 1): the 2-D field is the first (nx+4)(ny+4) values
        self.S = (self.nx + 4) * (self.ny + 4)
        assert self.L.ref_size(self.h) >= self.S

    def field(self, name):
        p = self.L.ref_field(self.h, FIELD_ID[name] if name in FIELD_ID else ROW_ID[name])
        return np.ctypeslib.as_array(p, shape=(self.S,)).reshape(self.ny + 4, self.nx + 4)

    def interior(self, name):
        return self.field(name)[2:-2, 2:-2]

    def ptr(self, a):
        return a.ctypes.data_as(POINTER(c_double))

    def loop(self):
        """SW2's loop(): pre-loop, frame 1 loaded, steps 1..N."""
        self.L.ref_loop(self.h, self.bname.encode(), os.fsencode(self.dir))

    def energy(self):
        ke, ape = c_double(), c_double()
        self.L.calc_energy_balance(ctypes.byref(ke), ctypes.byref(ape), self.h)
        return ke.value, ape.value

    def close(self):
        if self.h:
            self.L.ref_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

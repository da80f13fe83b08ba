"""sw2_amd -- MI355X (gfx950) time step of SW2 (Shallow Worlds 2).

Python mirror of the library's C-ABI (include/sw2_hip.h).  The product is
libsw2.so: host C for set-up, case files and frames; HIP kernels for every
per-step computation of SW2's loop() (core.c:134-338).  This module only
marshals arguments; it never computes a step itself.

    World    -- a world built by the library (sw struct, host arrays with halos)
    Session  -- the device copy of a world and its step loop
    read_case, build_params, load_state -- case files and initial states
"""
import ctypes
from ctypes import byref, c_double, c_int, c_void_p

import numpy as np

from ._ffi import EXPORTS, LIB_PATH, Params, SW, as_array, last_error, lib  # noqa: F401

FIELDS = ("eta", "u", "v", "tracer", "perturbs", "windx", "windy", "hB")
FORM_FP64, FORM_MIXED, PHASED = 0, 1, 0x100


class SW2Error(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise SW2Error(f"{what}: {last_error()}")


def build_params(**kw):
    """Params from keyword arguments; tables are numpy arrays kept alive on
    the returned object (attribute _keep)."""
    p = Params()
    keep = []
    defaults = dict(sch=2, polar=0, lon0_deg=0.0, lon1_deg=360.0, lat0_deg=-80.0, lat1_deg=80.0, nx=36, ny=18,
                    t0=0.0, t1=600.0, Dt=60.0, CFL=0.25, Tau=1e11, hypernu=(0.0, 0.0, 0.0), IteInfo=100,
                    SaveEvery=100000, LoadFrom=0, reftype=0, Omega=2 * np.pi / 35730.0, rE=7.1492e7, rP=7.1492e7,
                    gE=24.79, depth=1e5 / 24.79, geoeq=0, TracerDis=0.0)
    tables = {"vortices": (7, "num_vortices"), "Gaussians": (12, "num_Gaussians"), "wind": (2, "wind_rows")}
    d = dict(defaults)
    d.update({k: v for k, v in kw.items() if k not in tables})
    for k, v in d.items():
        if k == "hypernu":
            p.hypernu = (c_double * 3)(*v)
        else:
            setattr(p, k, v)
    wind = kw.get("wind", np.array([[-90.0, 0.0], [90.0, 0.0]]))
    for name, (ncol, cnt) in tables.items():
        arr = kw.get(name, wind if name == "wind" else None)
        if arr is None or len(arr) == 0:
            setattr(p, cnt, 0)
            continue
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1, ncol))
        keep.append(a)
        setattr(p, cnt, a.shape[0])
        setattr(p, name, a.ctypes.data_as(ctypes.POINTER(c_double)))
    p._keep = keep
    p._dict = dict(d, **{k: kw[k] for k in tables if k in kw})
    return p


def read_case(path):
    """Parse an SW2 case file with the library's reader; returns Params."""
    p = Params()
    _check(lib().sw2_read_case(str(path).encode(), byref(p)), f"read_case({path})")
    return p


def eval_expr(text):
    err = c_int(0)
    v = lib().sw2_eval_expr(text.encode(), byref(err))
    if err.value:
        raise SW2Error(f"cannot evaluate <{text}>")
    return v


class World:
    """A world built by the library from Params (sw2_init_params)."""

    def __init__(self, params):
        self.params = params
        self.sw = SW()
        _check(lib().sw2_init_params(byref(params), byref(self.sw)), "sw2_init_params")
        self.nx, self.ny = self.sw.nx, self.sw.ny
        self.S = self.sw.m.S
        self._closed = False

    def field(self, name):
        """(ny+4, nx+4) view of a host field; [j+1, i+1] is SW2's (i, j)."""
        return as_array(getattr(self.sw, name), self.S).reshape(self.ny + 4, self.nx + 4)

    def interior(self, name):
        return self.field(name)[2:-2, 2:-2]

    def rows(self, name):
        """per-row value of a j-only table, index j+1 for j=-1..ny+2"""
        return self.field(name)[:, 2].copy()

    def halo_update(self, name):
        lib().halo_update(getattr(self.sw, name), self.sw.M)

    def do_channel(self):
        lib().do_channel(byref(self.sw))

    def energy(self):
        ke, ape = c_double(), c_double()
        lib().calc_energy_balance(byref(ke), byref(ape), byref(self.sw))
        return ke.value, ape.value

    def mass(self):
        return lib().sw2_mass(byref(self.sw))

    def close(self):
        if not self._closed:
            lib().destroy(byref(self.sw))
            self._closed = True

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_state(world, eta, u, v, tracer=None, perturbs=None):
    """Put a world in the state loop() reaches when it starts from a frame
    (LoadFrom, core.c:87-118): pre-loop halo updates and do_channel, then the
    frame's interiors followed by a halo update of each loaded field
    (sppde_binary_files.c:147-180)."""
    for f in ("tracer", "u", "v", "eta", "hB"):
        world.halo_update(f)
    world.field("perturbs")[:] = 0.0
    world.do_channel()
    world.interior("eta")[:] = eta
    world.interior("u")[:] = u
    world.interior("v")[:] = v
    world.interior("tracer")[:] = 0.0 if tracer is None else tracer
    world.interior("perturbs")[:] = 0.0 if perturbs is None else perturbs
    for f in ("eta", "u", "v", "tracer", "perturbs"):
        world.halo_update(f)


class Session:
    """Device copy of a world; steps run on the GPU (sw2_session_*).
    form: FORM_FP64 (SW2's results) or FORM_MIXED (FP32 anomalies)."""

    def __init__(self, world, form=0):
        self.world = world
        self.h = c_void_p()
        _check(lib().sw2_session_create(byref(world.sw), form, byref(self.h)), "sw2_session_create")

    def begin(self, nstart=0, t=None):
        t = self.world.sw.t0 if t is None else t
        _check(lib().sw2_session_begin(self.h, nstart, t), "sw2_session_begin")

    def step(self, n):
        _check(lib().sw2_session_step(self.h, n), "sw2_session_step")

    def timed_steps(self, n):
        """Advance n steps; device milliseconds from HIP events on the
        session's stream."""
        ms = c_double()
        _check(lib().sw2_session_time(self.h, n, byref(ms)), "sw2_session_time")
        return ms.value

    def kernel_times(self, n):
        """Advance n steps on the fused path; summed device ms of the height
        kernel and of the momentum kernel (HIP events around each launch)."""
        ms = (c_double * 2)()
        _check(lib().sw2_session_kernel_times(self.h, n, ms, 2), "sw2_session_kernel_times")
        return ms[0], ms[1]

    def download(self):
        _check(lib().sw2_session_download(self.h, byref(self.world.sw)), "sw2_session_download")

    def failed_step(self):
        buf = ctypes.create_string_buffer(512)
        n = lib().sw2_session_failed_step(self.h, buf, 512)
        return n, buf.value.decode(errors="replace")

    def last_cfl(self):
        d = c_double()
        _check(lib().sw2_session_last_cfl(self.h, byref(d)), "sw2_session_last_cfl")
        return d.value

    def energy(self):
        """(KE, APE, mass) of the current state, reduced on the device in FP64."""
        ke, ape, m = c_double(), c_double(), c_double()
        _check(lib().sw2_session_energy(self.h, byref(ke), byref(ape), byref(m)), "sw2_session_energy")
        return ke.value, ape.value, m.value

    @property
    def form(self):
        return lib().sw2_session_form(self.h)

    @property
    def fused(self):
        return bool(lib().sw2_session_fused(self.h))

    def state(self):
        n, t = c_int(), c_double()
        lib().sw2_session_state(self.h, byref(n), byref(t))
        return n.value, t.value

    def close(self):
        if self.h:
            lib().sw2_session_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many worlds stepped together in the mixed-precision form
    (sw2_batch_*): one persistent launch per phase over all members' tiles,
    the step captured in a hipGraph, members admitted by HBM footprint and
    retired between steps."""

    def __init__(self, capacity, hbm_budget=0.0, form=FORM_MIXED):
        self.h = c_void_p()
        _check(lib().sw2_batch_create(form, capacity, float(hbm_budget), byref(self.h)), "sw2_batch_create")
        self.worlds = {}

    @staticmethod
    def footprint(world, form=FORM_MIXED):
        return lib().sw2_batch_footprint(byref(world.sw), form)

    def add(self, world, nsteps):
        """Admit a world for nsteps steps; the member id, or None when there
        is no room (slots or HBM budget)."""
        mid = c_int(-1)
        rc = lib().sw2_batch_add(self.h, byref(world.sw), nsteps, byref(mid))
        if rc == 1:
            return None
        _check(rc, "sw2_batch_add")
        self.worlds[mid.value] = world
        return mid.value

    def set_launch(self, blocks_eta, blocks_uv, steps_per_graph):
        _check(lib().sw2_batch_set_launch(self.h, blocks_eta, blocks_uv, steps_per_graph), "sw2_batch_set_launch")

    def launch(self):
        a, b, c = c_int(), c_int(), c_int()
        lib().sw2_batch_get_launch(self.h, byref(a), byref(b), byref(c))
        return a.value, b.value, c.value

    def set_tiles(self, tile_x, tile_y):
        """Tile shape of the batch kernels: (32, 16), the default, (64, 8) or (64, 16)."""
        _check(lib().sw2_batch_set_tiles(self.h, tile_x, tile_y), "sw2_batch_set_tiles")

    def tiles(self):
        a, b = c_int(), c_int()
        lib().sw2_batch_get_tiles(self.h, byref(a), byref(b))
        return a.value, b.value

    def step(self, n):
        """Queue n steps of every active member (returns without waiting)."""
        _check(lib().sw2_batch_step(self.h, n), "sw2_batch_step")

    def sync(self):
        _check(lib().sw2_batch_sync(self.h), "sw2_batch_sync")

    def pause(self, mid, paused=True):
        """Hold a member between steps (its tiles skipped, its step count
        kept) or release it."""
        _check(lib().sw2_batch_pause(self.h, mid, 1 if paused else 0), "sw2_batch_pause")

    def member(self, mid):
        a, b, c = c_int(), c_int(), c_int()
        _check(lib().sw2_batch_member(self.h, mid, byref(a), byref(b), byref(c)), "sw2_batch_member")
        return {"steps": a.value, "done": bool(b.value), "failed_step": c.value}

    def finished(self):
        ids = (c_int * 4096)()
        n = lib().sw2_batch_finished(self.h, ids, 4096)
        if n < 0:
            raise SW2Error(last_error())
        return list(ids[:n])

    def download(self, mid):
        w = self.worlds[mid]
        _check(lib().sw2_batch_download(self.h, mid, byref(w.sw)), "sw2_batch_download")
        return w

    def energy(self, mid):
        ke, ape, m = c_double(), c_double(), c_double()
        _check(lib().sw2_batch_energy(self.h, mid, byref(ke), byref(ape), byref(m)), "sw2_batch_energy")
        return ke.value, ape.value, m.value

    def remove(self, mid):
        _check(lib().sw2_batch_remove(self.h, mid), "sw2_batch_remove")
        self.worlds.pop(mid, None)

    def count(self):
        used = ctypes.c_size_t()
        n = lib().sw2_batch_count(self.h, byref(used))
        return n, used.value

    def close(self):
        if self.h:
            lib().sw2_batch_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Synthetic SW2 cases: the jet workloads and their storm bumps.

The zonal jet u(lat) = 60 cos(6 lat) m/s is SW2's background wind: a wind
table (lat_deg, u) read by read_wind (io.c:173-205) and linearly
interpolated at the cell-centre latitudes; SW2 applies no Coriolis force to
it (calc.c:563), so the jet is balanced by the deep layer it stands for.
Prognostic fields start as eta = storm bump(s), u = v = 0.

(The other reading -- the jet in the prognostic u with eta in discrete
geostrophic balance -- needs a height swing of about 1e4 m against the
4034 m mean depth: H turns negative on 20% of the 360x180 grid and SW2's
own CFL guard stops that run at step 127.)

A case here is what SW2 runs: a case file with LoadFrom 1 plus the initial
frame <bname>.0001.bin (MAGICSW2, io.c:27-64).  A bump of amplitude A and
e-fold radius R (degrees) is A * exp(-(d/R)^2), d the great-circle angle in
degrees from its centre.
"""
import os
import struct

import numpy as np

from . import World, build_params, load_state

PLANET = dict(Omega=2 * np.pi / 35730.0, rE=7.1492e7, rP=7.1492e7, gE=24.79)
DEPTH = 1e5 / 24.79

_LAT = np.round(np.arange(-900, 901) * 0.1, 1)
JET_WIND = np.stack([_LAT, 60.0 * np.cos(6.0 * np.deg2rad(_LAT))], axis=1)  # 0.1 deg table
ZERO_WIND = np.array([[-90.0, 0.0], [90.0, 0.0]])

# grid, latitude span, Dt, steps and del-4 coefficient (as entered) of the
# single-run workloads
CONFIGS = {
    "c0": dict(nx=360, ny=180, lat=80.0, Dt=60.0, steps=1000, nu4=1e15),
    "c1": dict(nx=4096, ny=2048, lat=85.0, Dt=10.0, steps=2000, nu4=1e15 * (360.0 / 4096) ** 4),
    "c2": dict(nx=4096, ny=2048, lat=85.0, Dt=10.0, steps=5000, nu4=1e15 * (360.0 / 4096) ** 4),
}


def jet_params(nx, ny, lat, Dt, steps, nu4, sch=2, IteInfo=None, SaveEvery=None, CFL=0.25, **extra):
    """Params of a jet case on an nx x ny grid over latitudes -lat..lat."""
    kw = dict(sch=sch, lon0_deg=0.0, lon1_deg=360.0, lat0_deg=-lat, lat1_deg=lat, nx=nx, ny=ny,
              t0=0.0, t1=steps * Dt, Dt=Dt, CFL=CFL, Tau=1e11, hypernu=(0.0, nu4, 0.0),
              IteInfo=IteInfo or max(steps, 1), SaveEvery=SaveEvery or 10 * steps + 1, LoadFrom=1,
              depth=DEPTH, wind=JET_WIND, **PLANET)
    kw.update(extra)
    return build_params(**kw)


def config_params(name, steps=None, **kw):
    c = CONFIGS[name]
    return jet_params(c["nx"], c["ny"], c["lat"], c["Dt"], c["steps"] if steps is None else steps, c["nu4"], **kw)


def scaled_params(nx, ny, steps, lat=80.0, Dt=None, **kw):
    """Jet case on any grid: del-4 coefficient scaled by (dx ratio)^4 from
    the 360x180 case, Dt scaled with the grid spacing from 60 s at 360."""
    Dt = Dt if Dt is not None else 60.0 * 360.0 / nx
    return jet_params(nx, ny, lat, Dt, steps, 1e15 * (360.0 / nx) ** 4, **kw)


def add_bumps(world, eta, bumps):
    """Add Gaussian bumps (lon_deg, lat_deg, amplitude_m, efold_radius_deg)."""
    w = world
    i = np.arange(1, w.nx + 1)
    j = np.arange(1, w.ny + 1)
    lone = lambda q: w.sw.lon0 + (q * (w.sw.lon1 - w.sw.lon0)) / w.nx
    latn = lambda q: w.sw.lat0 + (q * (w.sw.lat1 - w.sw.lat0)) / w.ny
    lonc = (lone(i) + lone(i - 1)) / 2.0
    latc = (latn(j) + latn(j - 1)) / 2.0
    LON, LAT = np.meshgrid(lonc, latc)
    for lon_d, lat_d, amp, rad in bumps:
        lo, la = np.deg2rad(lon_d), np.deg2rad(lat_d)
        c = np.sin(la) * np.sin(LAT) + np.cos(la) * np.cos(LAT) * np.cos(LON - lo)
        d = np.rad2deg(np.arccos(np.clip(c, -1.0, 1.0)))
        eta = eta + amp * np.exp(-(d / rad) ** 2)
    return eta


def single_bump(seed):
    """The storm of the single-run workloads: 5% of the mean depth, e-fold
    radius 2 deg, at lat 20, lon 0, centre jittered by +-1 deg (seeded)."""
    rng = np.random.default_rng(seed)
    dl, dp = rng.uniform(-1.0, 1.0, size=2)
    return [(0.0 + dl, 20.0 + dp, 0.05 * DEPTH, 2.0)]


def ensemble_bumps(rng):
    """1-5 bumps, amplitude 1-10% of the mean depth, radius 1-4 deg,
    latitude -60..60, any longitude."""
    k = int(rng.integers(1, 6))
    return [(float(rng.uniform(0, 360)), float(rng.uniform(-60, 60)), float(rng.uniform(0.01, 0.10)) * DEPTH,
             float(rng.uniform(1.0, 4.0))) for _ in range(k)]


def initial_state(world, bumps):
    eta = add_bumps(world, np.zeros((world.ny, world.nx)), bumps)
    z = np.zeros_like(eta)
    return eta, z, z.copy()


# ------------------------------------------------------------ ensembles
ENSEMBLE_SIZES = [(256, 128), (512, 256), (1024, 512), (2048, 1024)]
SERVED_SIZES = [(512, 256), (1024, 512), (2048, 1024), (4096, 2048)]
SERVED_WEIGHTS = [0.4, 0.3, 0.2, 0.1]


def storm_bumps(rng):
    """Seeded storm injections of the ensemble and served workloads: 1-5
    bumps, amplitude 1-10% of the mean depth, e-fold radius 1-4 deg, centre
    latitude uniform in -60..60 deg, longitude uniform."""
    k = int(rng.integers(1, 6))
    return [(float(rng.uniform(0.0, 360.0)), float(rng.uniform(-60.0, 60.0)), float(rng.uniform(0.01, 0.10)) * DEPTH,
             float(rng.uniform(1.0, 4.0))) for _ in range(k)]


def member_params(nx, ny, steps, **kw):
    """The 360x180 workload's physics on an nx x ny grid (lat -80..80): Dt
    and the del-4 coefficient scaled with the grid spacing; keyword
    arguments as jet_params takes them (an in-loop storm table, say)."""
    return scaled_params(nx, ny, steps=steps, lat=80.0, **kw)


def ensemble(n=64, steps=3000, seed=2026):
    """Members of the batched ensemble: grid sizes uniform over
    ENSEMBLE_SIZES, seeded storm bumps; a list of (params, bumps)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx, ny = ENSEMBLE_SIZES[int(rng.integers(0, len(ENSEMBLE_SIZES)))]
        out.append((member_params(nx, ny, steps), storm_bumps(rng)))
    return out


def served_requests(n, rate=4.0, seed=2027, min_steps=500, max_steps=20000):
    """Requests of the served stream: Poisson arrivals at `rate` per second,
    grid sizes from SERVED_SIZES with SERVED_WEIGHTS, step counts uniform in
    min_steps..max_steps, seeded storm bumps; a list of (t_arrival, nx, ny,
    steps, bumps)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate))
        nx, ny = SERVED_SIZES[int(rng.choice(len(SERVED_SIZES), p=SERVED_WEIGHTS))]
        steps = int(rng.integers(min_steps, max_steps + 1))
        out.append((t, nx, ny, steps, storm_bumps(rng)))
    return out


def make_world(params, bumps):
    """World in the state loop() starts from: initial frame loaded."""
    w = World(params)
    eta, u, v = initial_state(w, bumps)
    load_state(w, eta, u, v)
    return w, (eta, u, v)


# ------------------------------------------------------------ files
def params_dict(p):
    """The values of Params p as case_text writes them: the keyword
    dictionary of build_params, or for Params read from a case file
    (read_case) the struct's fields and copies of its tables."""
    if hasattr(p, "_dict"):
        return p._dict
    d = {n: getattr(p, n) for n, _ in p._fields_ if n not in ("hypernu", "vortices", "Gaussians", "wind")}
    d["hypernu"] = tuple(p.hypernu)
    for name, cnt, ncol in (("vortices", "num_vortices", 7), ("Gaussians", "num_Gaussians", 12), ("wind", "wind_rows", 2)):
        n = getattr(p, cnt)
        d[name] = np.ctypeslib.as_array(getattr(p, name), shape=(n * ncol,)).reshape(n, ncol).copy() if n else None
    return d


def _table(rows, ncol, label):
    rows = np.zeros((0, ncol)) if rows is None else np.asarray(rows, dtype=float).reshape(-1, ncol)
    return [f"{len(rows)} # {label} rows", f"{ncol}"] + [" ".join(repr(float(x)) for x in r) for r in rows]


def case_text(p, bname, flat=False):
    """SW2 case file (tag order of init.c:48-229) for Params p, and the
    planet file it includes (flat: the planet's tags written in place and
    no m4 construct at all, so the file reads without pre-processing)."""
    d = params_dict(p)
    hn = d["hypernu"]
    wind = np.asarray(ZERO_WIND if d.get("wind") is None else d["wind"])
    planet_lines = [f"Omega {d['Omega']!r}", f"rE {d['rE']!r}", f"rP {d['rP']!r}", f"gE {d['gE']!r}"]
    lines = [
        f"# {bname}: jet as background wind, storm bump(s) in frame 1",
        f"sch {d['sch']}", "polar 0",
        f"lon0 {d['lon0_deg']!r}", f"lon1 {d['lon1_deg']!r}", f"lat0 {d['lat0_deg']!r}", f"lat1 {d['lat1_deg']!r}",
        f"nx {d['nx']}", f"ny {d['ny']}",
        f"t0 {d['t0']!r}", f"t1 {d['t1']!r}", f"Dt {d['Dt']!r}", f"CFL {d['CFL']!r}", f"Tau {d['Tau']!r}",
        f"Hypernu {hn[0]!r} {hn[1]!r} {hn[2]!r}",
        f"IteInfo {d['IteInfo']}", f"SaveEvery {d['SaveEvery']}", f"LoadFrom {d['LoadFrom']}", "reftype 0",
    ] + (planet_lines if flat else [f"include({bname}.planet)"]) + [
        f"depth {d['depth']!r}", f"geoeq {d.get('geoeq', 0)}",
    ] + _table(d.get("vortices"), 7, "vortex") + _table(d.get("Gaussians"), 12, "Gaussian") + [
        f"TracerDis {float(d.get('TracerDis', 0.0))!r}",
        f"{len(wind)} # wind rows", "2",
    ] + [f"{float(a)!r} {float(b)!r}" for a, b in wind]
    planet = "\n".join(planet_lines) + "\n"
    return "\n".join(lines) + "\n", planet


def write_frame(path, nx, ny, fields, nstep=0, t=0.0):
    """MAGICSW2 frame: header then eta, u, v, tracer, perturbs (io.c:27-64);
    each field is int 15 and (nx+4)(ny+4) doubles, halos included."""
    with open(path, "wb") as f:
        f.write(struct.pack("=iidii", 28, nstep, t, nx, ny))
        for a in fields:
            full = np.zeros((ny + 4, nx + 4))
            full[2:-2, 2:-2] = a
            f.write(struct.pack("=i", 15))
            f.write(full.astype("<f8").tobytes())


def write_case(directory, bname, params, state, outdir):
    """Write <bname>.sw, <bname>.planet and <outdir>/<bname>.0001.bin."""
    os.makedirs(directory, exist_ok=True)
    os.makedirs(outdir, exist_ok=True)
    txt, planet = case_text(params, bname)
    with open(os.path.join(directory, bname + ".sw"), "w") as f:
        f.write(txt)
    with open(os.path.join(directory, bname + ".planet"), "w") as f:
        f.write(planet)
    eta, u, v = state
    z = np.zeros_like(eta)
    write_frame(os.path.join(outdir, bname + ".0001.bin"), params.nx, params.ny, [eta, u, v, z, z])

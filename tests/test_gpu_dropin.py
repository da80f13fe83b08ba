"""The drop-in boundary end to end: SW2's driver program (sw2_amd/SW, the
library's build of main/Main.c) runs the reference's own JupiterEnergyTest
and SaturnSmoothInjection case files through initialize / loop / destroy on
the GPU.  Its printed
ENERGY lines and its last frame must be those of the oracle's run of the
same case (core.c:134-338, energy.c:25-55, io.c:27-76)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import sw2_amd
from sw2_amd import World
from oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEPS = 600
FIELDS = ("eta", "u", "v", "tracer", "perturbs")


def shortened_case(golden, tmp_path):
    src = os.path.join(golden, "jupiter")
    for f in ("JupiterEGM.planet", "nowind.wind"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    txt = open(os.path.join(src, "jupiters.sw")).read()
    txt = re.sub(r"(?m)^t1 .*$", f"t1  <DT*{STEPS}>", txt)
    txt = re.sub(r"(?m)^SaveEvery .*$", "SaveEvery 200", txt)
    (tmp_path / "jupiters.sw").write_text(txt)


def test_sw_driver_runs_reference_case(golden, tmp_path):
    shortened_case(golden, tmp_path)
    print(f"SW 1 1 jupiters ({STEPS} steps)", flush=True)
    r = subprocess.run([os.path.join(ROOT, "sw2_amd", "SW"), "1", "1", "jupiters", "out"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "PROGRAM ENDED" in r.stdout
    printed = [line.split()[1:] for line in r.stdout.splitlines() if line.startswith("ENERGY ")]

    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        p = sw2_amd.read_case("jupiters.sw")
    finally:
        os.chdir(cwd)
    o = Oracle(p)
    o.preloop()
    ke0, ape0 = o.energy()
    fo, rec = o.step(STEPS, energy=True)
    assert fo == 0
    expect = [(p.t0, ke0, ape0)] + [tuple(x) for x in rec]
    assert len(printed) == len(expect) == 1 + 1 + STEPS // 100
    for got, (t, ke, ape) in zip(printed, expect):
        assert got == ["%e" % t, "%e" % ke, "%e" % ape, "%e" % (ke + ape)]

    w = World(p)
    n, t = ctypes_load(w, 3, str(tmp_path / "out"))
    assert n == STEPS and t == o.time
    for f in FIELDS:
        a, b = w.interior(f), o.interior(f)
        assert np.array_equal(a, b), (f, np.argwhere(a != b)[:3])


def ctypes_load(world, frame, folder, name=b"jupiters"):
    from ctypes import byref, c_double, c_int
    n, t = c_int(), c_double()
    rc = sw2_amd.lib().load_binary_sw(frame, byref(n), byref(t), name, folder.encode(), byref(world.sw))
    assert rc == 0
    return n.value, t.value


def test_sw_driver_runs_saturn_case(golden, tmp_path):
    """The reference's SaturnSmoothInjection case (600 columns: the fused
    step, with its storm injected after the momentum kernel) through the
    driver: an ENERGY line every step (IteInfo 1), as the oracle's run, and
    the last frame's fields.  Frames are saved every 50 steps instead of
    every step."""
    src = os.path.join(golden, "saturn")
    for f in ("SaturnEGM2.planet", "Saturn_Cassini.wind", "nowind.wind"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    txt = open(os.path.join(src, "minicanal.sw")).read()
    txt = re.sub(r"(?m)^SaveEvery .*$", "SaveEvery 50", txt)
    (tmp_path / "minicanal.sw").write_text(txt)
    print("SW 1 1 minicanal (100 steps)", flush=True)
    r = subprocess.run([os.path.join(ROOT, "sw2_amd", "SW"), "1", "1", "minicanal", "out"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "PROGRAM ENDED" in r.stdout
    printed = [line.split()[1:] for line in r.stdout.splitlines() if line.startswith("ENERGY ")]
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        p = sw2_amd.read_case("minicanal.sw")
    finally:
        os.chdir(cwd)
    o = Oracle(p)
    o.preloop()
    ke0, ape0 = o.energy()
    fo, rec = o.step(100, energy=True)
    assert fo == 0
    expect = [(p.t0, ke0, ape0)] + [tuple(x) for x in rec]
    assert len(printed) == len(expect), (len(printed), len(expect))
    for got, (t, ke, ape) in zip(printed, expect):
        assert got == ["%e" % t, "%e" % ke, "%e" % ape, "%e" % (ke + ape)]
    w = sw2_amd.World(p)
    n, t = ctypes_load(w, 2, str(tmp_path / "out"), b"minicanal")
    assert n == 100 and t == o.time
    for f in FIELDS:
        a, b = w.interior(f), o.interior(f)
        assert np.array_equal(a, b), (f, np.argwhere(a != b)[:3])


HDR_DRIVER = os.path.join(ROOT, "tests", "_bin", "sw2_headers_driver")


@pytest.mark.skipif(not os.path.exists(HDR_DRIVER), reason="the SW2-headers caller is built only where SW2's headers are")
def test_caller_built_against_sw2_headers_runs_loop(golden, tmp_path):
    """An object compiled against SW2's own sw.h / sppde.h (tests/
    sw2_headers_driver.c, built by __graft_entry__.build()) runs initialize /
    loop / destroy through libsw2.so: after 300 steps of SW2's Jupiter case
    its eta, u and v, read through SW2's struct definitions, are the
    oracle's."""
    n = 300
    for f in os.listdir(os.path.join(golden, "jupiter")):
        shutil.copy(os.path.join(golden, "jupiter", f), tmp_path / f)
    raw = tmp_path / "state.raw"
    print(f"sw2_headers_driver jupiters {n} steps", flush=True)
    r = subprocess.run([HDR_DRIVER, "jupiters", "out", str(n), str(raw)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        p = sw2_amd.read_case("jupiters.sw")
    finally:
        os.chdir(cwd)
    o = Oracle(p)
    o.preloop()
    assert o.step(n)[0] == 0
    S = (p.nx + 4) * (p.ny + 4)
    got = np.fromfile(raw, dtype=np.float64).reshape(3, p.ny + 4, p.nx + 4)
    assert got.size == 3 * S
    for k, f in enumerate(("eta", "u", "v")):
        assert np.array_equal(got[k], o.field(f)), f
    ke, ape = o.energy()
    e1 = [float(x) for l in r.stdout.splitlines() if l.startswith("ENERGY1") for x in l.split()[1:]]
    assert e1 == [ke, ape]

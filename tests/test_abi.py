"""The C-ABI library: it loads, exports every function include/sw2_hip.h
declares, and its host-side pieces (case reader, expression evaluator, frame
IO, world set-up) behave as SW2's.  CPU only: no kernel is launched."""
import os
import re
import subprocess

import numpy as np
import pytest

import sw2_amd
from sw2_amd import World, cases
from sw2_amd._ffi import EXPORTS, LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "sw2_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"#define[^\n]*", "", txt)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{}]*\)\s*;", txt)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_library_exports_every_declared_function():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    decl = declared_functions()
    assert len(decl) >= 35
    missing = [n for n in decl if n not in exported]
    assert not missing, missing
    assert set(EXPORTS) <= set(decl)


def test_library_is_built_for_gfx950():
    out = subprocess.run(["strings", LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out
    assert b"gfx950" in sw2_amd.lib().sw2_build_info()


def test_sin_and_cos_stay_separate_libm_calls():
    """gcc merges sin(x) and cos(x) of one argument into sincos(x) when it
    sees both (after inlining), and glibc's sincos differs from sin in the
    last bit for some x: that moved sin(latc) of three rows by an ulp and
    the 1000-step run of the CPU-runnable case off the oracle.  Neither the
    library nor the oracle may import sincos, and the table of sin(latc)
    rows equals libm's sin (Python's math.sin calls it) bit for bit."""
    import ctypes
    import math
    from oracle import LIB_PATH as ORACLE_LIB
    from sw2_amd import _ffi
    for path in (LIB_PATH, ORACLE_LIB):
        out = subprocess.run(["nm", "-D", "--undefined-only", path], capture_output=True, text=True, check=True).stdout
        assert not any(line.split()[-1].startswith("sincos") for line in out.splitlines() if line.split()), path
    w, _ = cases.make_world(cases.config_params("c0"), [])
    L = sw2_amd.lib()
    nt, R = 64, w.sw.ny + 4
    tab = np.zeros(nt * R)
    fn = L.sw2_row_tables
    fn.argtypes = [ctypes.POINTER(_ffi.SW), ctypes.POINTER(ctypes.c_double)]
    assert fn(ctypes.byref(w.sw), tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 1
    SINC = 11  # SW2_T_SINC in sw2_internal.h
    lat0, lat1, ny = w.sw.lat0, w.sw.lat1, w.sw.ny
    latn = lambda j: lat0 + j * (lat1 - lat0) / ny
    for j in range(-1, ny + 3):
        latc = (latn(j) + latn(j - 1)) / 2.0
        assert tab[SINC * R + j + 1] == math.sin(latc), j


@pytest.mark.parametrize("expr,val", [
    ("2+3^2", 11.0), ("-2^2", 4.0), ("2^3^2", 64.0), ("log(100)", 2.0), ("ln(e)", 1.0),
    ("10 % 4", 2.0), ("pow(2,10)", 1024.0), ("sin 0", 0.0), ("(1,2,3)", 3.0), ("86400*20", 1728000.0),
    ("12*3600/50", 864.0), ("fac 5", 120.0), ("ncr(5,2)", 10.0), ("atan2(1,1)*4", np.pi), ("--3", 3.0),
])
def test_expression_evaluator_tinyexpr_semantics(expr, val):
    """'^' left-associative with unary minus binding tighter, log = log10
    (the compile flags of sppde_tinyexpr.c:25-35)."""
    assert sw2_amd.eval_expr(expr) == pytest.approx(val, rel=1e-15)


def test_expression_errors():
    with pytest.raises(sw2_amd.SW2Error):
        sw2_amd.eval_expr("2+")
    with pytest.raises(sw2_amd.SW2Error):
        sw2_amd.eval_expr("foo(3)")


def test_reads_reference_case_file(golden):
    cwd = os.getcwd()
    os.chdir(os.path.join(golden, "jupiter"))
    try:
        p = sw2_amd.read_case("jupiters.sw")
    finally:
        os.chdir(cwd)
    assert (p.sch, p.polar, p.nx, p.ny) == (1, 0, 50, 50)
    assert (p.lon0_deg, p.lon1_deg, p.lat0_deg, p.lat1_deg) == (0.0, 90.0, -35.0, -10.0)
    assert (p.t0, p.t1, p.Dt, p.CFL, p.Tau) == (0.0, 1728000.0, 50.0, 0.25, 1e7)
    assert list(p.hypernu) == [0.0, 0.0, 0.0]
    assert (p.IteInfo, p.SaveEvery, p.LoadFrom, p.reftype) == (100, 864, 0, 0)
    assert (p.Omega, p.rE, p.rP, p.gE, p.depth, p.geoeq) == (0.000176, 7.1492e7, 6.6852e7, 22.58, 1000.0, 0)
    assert p.num_vortices == 0 and p.num_Gaussians == 1 and p.wind_rows == 2
    g = np.ctypeslib.as_array(p.Gaussians, shape=(12,))
    assert list(g) == [0.0, 604800.0, 0.0, 45.0, -22.5, 0.0, -3.6, 4.52e11, 0.0, 3.5e6, 3.5e6, 0.0]


def test_reads_saturn_smooth_injection_case(golden):
    """The reference's SaturnSmoothInjection case: macros from define(),
    <expr> times, the planet include and the 861-row Cassini wind table
    (the case file's own nowind include is commented out)."""
    cwd = os.getcwd()
    os.chdir(os.path.join(golden, "saturn"))
    try:
        p = sw2_amd.read_case("minicanal.sw")
    finally:
        os.chdir(cwd)
    assert (p.sch, p.polar, p.nx, p.ny) == (1, 0, 600, 100)
    assert (p.lon0_deg, p.lon1_deg, p.lat0_deg, p.lat1_deg) == (0.0, 60.0, 65.0, 75.0)
    assert (p.t0, p.t1, p.Dt, p.CFL, p.Tau) == (0.0, 6000.0, 60.0, 0.25, 1e9)
    assert list(p.hypernu) == [0.0, 0.0, 0.0] and p.depth == 500.0
    assert p.num_vortices == 0 and p.num_Gaussians == 1 and p.wind_rows == 861
    g = np.ctypeslib.as_array(p.Gaussians, shape=(12,))
    # <20*DT> = 1200 s, P1 = 68.8, P3 = 1.5e9 (define()), radius and sigma in metres, no tracer
    assert list(g) == [0.0, 1200.0, 200.0, 30.0, 68.8, 0.0, 14.2, 1.5e9, 0.0, 114e3, 114e3, 0.0]


def test_case_file_macro_subset(tmp_path):
    (tmp_path / "inc.txt").write_text("rE 7e7 # radius\nrP 7e7\n")
    txt = cases.case_text(cases.scaled_params(8, 6, steps=2), "x")[0]
    txt = "define(NX,8)\ndefine(TWICE,<2*NX>)\n# NX in a comment stays NX\n" + txt.replace("nx 8", "nx NX")
    txt = txt.replace("include(x.planet)", "Omega 1e-4\ninclude(inc.txt)\ngE 20 dnl deleted with its newline\n")
    (tmp_path / "m.sw").write_text(txt)
    p = sw2_amd.read_case(str(tmp_path / "m.sw"))
    assert p.nx == 8 and p.rE == 7e7 and p.gE == 20.0 and p.Omega == 1e-4


def test_case_file_tag_order_is_enforced(tmp_path):
    txt, planet = cases.case_text(cases.scaled_params(8, 6, steps=2), "x")
    (tmp_path / "x.planet").write_text(planet)
    (tmp_path / "bad.sw").write_text(txt.replace("ny 6", "nz 6"))
    with pytest.raises(sw2_amd.SW2Error, match="expecting 'ny'"):
        sw2_amd.read_case(str(tmp_path / "bad.sw"))


def test_written_case_round_trips(tmp_path):
    p = cases.scaled_params(20, 12, steps=7, lat=65.0)
    w, state = cases.make_world(p, cases.single_bump(3))
    cases.write_case(str(tmp_path), "rt", p, state, str(tmp_path / "out"))
    q = sw2_amd.read_case(str(tmp_path / "rt.sw"))
    for f in ("sch", "nx", "ny", "lon0_deg", "lon1_deg", "lat0_deg", "lat1_deg", "t0", "t1", "Dt", "CFL", "Tau",
              "IteInfo", "SaveEvery", "LoadFrom", "Omega", "rE", "rP", "gE", "depth", "wind_rows"):
        assert getattr(q, f) == getattr(p, f), f
    assert list(q.hypernu) == list(p.hypernu)
    assert np.array_equal(World(q).field("windx"), World(p).field("windx"))


def test_frame_io_round_trip(tmp_path):
    import ctypes
    p = cases.scaled_params(20, 12, steps=7, lat=60.0)
    w, (eta, u, v) = cases.make_world(p, cases.single_bump(4))
    L = sw2_amd.lib()
    L.save_binary_sw(3, 42, ctypes.c_double(126.0), b"fr", str(tmp_path).encode(), ctypes.byref(w.sw))
    w2 = World(p)
    n, t = ctypes.c_int(), ctypes.c_double()
    assert L.load_binary_sw(3, ctypes.byref(n), ctypes.byref(t), b"fr", str(tmp_path).encode(), ctypes.byref(w2.sw)) == 0
    assert (n.value, t.value) == (42, 126.0)
    for f in ("eta", "u", "v", "tracer", "perturbs"):
        assert np.array_equal(w2.field(f), w.field(f)), f   # interior + halo_update reproduces the halos
    assert L.load_binary_sw(9, ctypes.byref(n), ctypes.byref(t), b"fr", str(tmp_path).encode(), ctypes.byref(w2.sw)) == 1


def test_world_rejects_bad_parameters():
    with pytest.raises(sw2_amd.SW2Error, match="too small"):
        World(cases.scaled_params(2, 2, steps=1))
    with pytest.raises(sw2_amd.SW2Error, match="a la parra"):
        World(cases.scaled_params(8, 8, steps=1, sch=11))
    with pytest.raises(sw2_amd.SW2Error, match="out of bounds"):
        World(cases.scaled_params(8, 8, steps=1, wind=np.array([[-10.0, 0.0], [10.0, 0.0]])))


def test_sw_driver_usage_and_rank_check(tmp_path, golden):
    """The SW driver (main/Main.c without MPI) keeps SW2's usage message and
    CRASH exit status, and refuses a split domain before any GPU work."""
    exe = os.path.join(ROOT, "sw2_amd", "SW")
    r = subprocess.run([exe], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 255
    assert "SW npx npy input_file output_folder" in r.stderr
    case = os.path.join(golden, "jupiter", "jupiters")
    r = subprocess.run([exe, "2", "1", case, str(tmp_path / "out")], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 255
    assert "npx=2 npy=1" in r.stdout
    assert "use npx=npy=1" in r.stderr
    assert (tmp_path / "out" / "jupiters.sw.prem4").exists()
    assert (tmp_path / "out" / "jupiters.sw").read_bytes() == open(case + ".sw", "rb").read()


REFERENCE = "/root/reference"
HDR_DRIVER_SRC = os.path.join(ROOT, "tests", "sw2_headers_driver.c")


def test_struct_layout_is_sw2s():
    """sizeof and every offsetof of map and sw, from include/sw2_hip.h and
    from the ctypes mirror, against SW2's own headers (tests/golden/
    sw_layout.json, written by tests/golden/make_layout.py)."""
    import json
    from sw2_amd._ffi import SW, Map
    from tests.golden.make_layout import MAP_FIELDS, SW_FIELDS, measure
    with open(os.path.join(ROOT, "tests", "golden", "sw_layout.json")) as f:
        ref = json.load(f)
    ours = measure(['#include "sw2_hip.h"'], [os.path.join(ROOT, "include")])
    assert ours["map"] == ref["map"] and ours["sw"] == ref["sw"]
    import ctypes
    assert ctypes.sizeof(Map) == ref["map"]["sizeof"] and ctypes.sizeof(SW) == ref["sw"]["sizeof"]
    for f in MAP_FIELDS:
        assert getattr(Map, f).offset == ref["map"][f], f
    for f in SW_FIELDS:
        assert getattr(SW, f).offset == ref["sw"][f], f


def build_headers_driver(out):
    """The caller of tests/sw2_headers_driver.c, compiled against SW2's own
    sw.h and sppde.h and linked against libsw2.so."""
    subprocess.run(["gcc", "-std=gnu99", "-O2", f"-I{REFERENCE}", f"-I{REFERENCE}/sppde", "-o", out, HDR_DRIVER_SRC,
                    f"-L{os.path.dirname(LIB_PATH)}", "-lsw2", f"-Wl,-rpath,{os.path.dirname(LIB_PATH)}", "-lm"],
                   check=True)


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="SW2's headers are not here")
@pytest.mark.parametrize("case", [("jupiter", "jupiters"), ("saturn", "minicanal")])
def test_caller_built_against_sw2_headers(tmp_path, golden, case):
    """initialize, do_channel, calc_energy_balance and destroy called from an
    object compiled against SW2's headers: the values it reads through SW2's
    struct definitions are the library's."""
    exe = str(tmp_path / "drv")
    build_headers_driver(exe)
    sub, bname = case
    r = subprocess.run([exe, bname, str(tmp_path), "0", str(tmp_path / "x.raw")], cwd=os.path.join(golden, sub),
                       capture_output=True, text=True, check=True)
    tag = {l.split()[0]: l.split()[1:] for l in r.stdout.splitlines() if l.split() and l.split()[0].isupper()}
    cwd = os.getcwd()
    os.chdir(os.path.join(golden, sub))
    try:
        w = World(sw2_amd.read_case(bname + ".sw"))
    finally:
        os.chdir(cwd)
    s, M = w.sw, w.sw.m
    kv = lambda key: dict(zip(tag[key][0::2], tag[key][1::2]))
    lay = tag["LAYOUT"]
    assert lay[:8] == ["nx", str(s.nx), "ny", str(s.ny), "N", str(s.N), "S", str(M.S)]
    assert " ".join(lay) == (f"nx {s.nx} ny {s.ny} N {s.N} S {M.S} sh 2 l0 1 1 l1 {s.nx} {s.ny} s {s.nx + 4} {s.ny + 4} "
                             f"per 1 0 bc 0 0 1 1 nb 0 0 -1 -1 -1 -1")
    v = kv("VALUES")
    for name in ("depth", "Omega", "gE", "nu4", "Dt"):
        assert float(v[name]) == getattr(s, name), name
    assert float(v["sigma"]) == s.sigma_dis
    assert int(v["gauss"]) == s.num_Gaussians and int(v["vort"]) == s.num_vortices
    rows = tag["ROWS"]
    assert float(rows[1]) == w.field("Dxc")[2, 2] and float(rows[2]) == w.field("Dxc")[s.ny + 1, 2]
    assert float(rows[4]) == w.field("gc")[2, 2] and float(rows[6]) == w.field("fc")[2, 2]
    w.do_channel()
    ke, ape = w.energy()
    assert [float(x) for x in tag["ENERGY0"]] == [ke, ape]
    w.close()


@pytest.mark.parametrize("expr,val", [("fac 20", 2432902008176640000.0), ("fac 3.7", 6.0), ("fac 0", 1.0),
                                      ("ncr(60,30)", 118264581564861424.0), ("ncr(10,0)", 1.0), ("ncr(7,7)", 1.0),
                                      ("npr(6,3)", 120.0)])
def test_factorial_and_binomial_exact(expr, val):
    """fac and ncr are exact integers wherever the value fits in 64 bits."""
    assert sw2_amd.eval_expr(expr) == val


def test_factorial_and_binomial_limits():
    assert np.isinf(sw2_amd.eval_expr("fac 21")) and np.isnan(sw2_amd.eval_expr("fac(-1)"))
    assert np.isnan(sw2_amd.eval_expr("ncr(2,3)")) and np.isinf(sw2_amd.eval_expr("ncr(1e10,2)"))
    assert np.isinf(sw2_amd.eval_expr("ncr(68,34)"))  # 2.8e19 > 2^64


def _case_with_line(tmp_path, old, new):
    p = cases.scaled_params(160, 64, steps=5, lat=70.0)
    txt, _ = cases.case_text(p, "esc", flat=True)
    assert old in txt
    path = tmp_path / "esc.sw"
    path.write_text(txt.replace(old, new, 1))
    return str(path)


def test_expression_lines_and_escapes(tmp_path):
    """<expr> values replaced by their %.15e value before a line is read;
    an escaped bracket makes the rest of the line literal; a '<' without
    '>' or a '>' without '<' is an error."""
    q = sw2_amd.read_case(_case_with_line(tmp_path, "nx 160", "nx <2*80> \\< not an <expr> \\> here"))
    assert q.nx == 160
    q = sw2_amd.read_case(_case_with_line(tmp_path, "CFL 0.25", "CFL <1/3>"))
    assert q.CFL == float("%.15e" % (1 / 3))
    with pytest.raises(sw2_amd.SW2Error, match="missing '>'"):
        sw2_amd.read_case(_case_with_line(tmp_path, "ny 64", "ny <64"))
    with pytest.raises(sw2_amd.SW2Error, match="without its '<'"):
        sw2_amd.read_case(_case_with_line(tmp_path, "ny 64", "ny 64 >"))
    with pytest.raises(sw2_amd.SW2Error, match="can not be evaluated"):
        sw2_amd.read_case(_case_with_line(tmp_path, "ny 64", "ny <64+>"))
    with pytest.raises(sw2_amd.SW2Error, match="expecting an integer"):
        sw2_amd.read_case(_case_with_line(tmp_path, "ny 64", "ny 64.5"))

"""The oracle against SW2 itself, CPU only.

oracle/_ref holds SW2 compiled from the reference's own sources (make -C
oracle ref, where they are present; see oracle/ref_harness.c).  Each test
builds the same case in both, runs SW2's loop() and the oracle's steps, and
requires every field -- eta, u, v, tracer, perturbs, windx and windy, whole
arrays with their halos -- to be equal value for value (a zero's sign is the
only freedom, as np.array_equal allows; NaNs must sit in the same cells).

Between them the cases reach every function of the oracle: the ten
limiters, BODYH / BODYU / BODYV and the tracer body (calc.c:70-237) with its
non-conservative term and dissipation, the three hyperviscosity orders and
the Rayleigh drag, the Coriolis update, the halo and channel rules, storm
injection with and without tracer (perturbs.c:22-395), vortex insertion
(perturbs.c:397-460), the wind-table interpolation (tools.c:21-74 through
io.c:173-205), the energy diagnostics, and SW2's own two example cases.
The reference build is skipped where it is absent (on a box without the
reference sources)."""
import json
import os

import numpy as np
import pytest

import sw2_amd
from oracle import Oracle
from oracle import ref as refmod
from sw2_amd import cases

pytestmark = pytest.mark.skipif(not refmod.available(), reason="oracle/_ref not built (no reference sources here)")

FIELDS = ("eta", "u", "v", "tracer", "perturbs", "windx", "windy")


def run_both(tmp_path, p, state, n, bname="refcase"):
    r = refmod.Ref(p, state, tmp_path, bname)
    assert r.L.ref_nsteps(r.h) == n  # SW2's N = (int)((t1 - t0) / Dt), init.c:99
    r.loop()
    o = Oracle(p)
    if p.LoadFrom:
        o.load(*state)
    else:
        o.preloop()
    failed, _ = o.step(n)
    return r, o, failed


def assert_same(r, o, label=""):
    for f in FIELDS:
        a, b = r.field(f), o.field(f)
        assert np.array_equal(a, b, equal_nan=True), (label, f, int(np.sum(~((a == b) | (np.isnan(a) & np.isnan(b))))))


def jet(nx, ny, steps, **kw):
    p = cases.scaled_params(nx, ny, steps=steps, lat=75.0, **kw)
    w, state = cases.make_world(p, cases.single_bump(nx + ny))
    w.close()
    return p, state


@pytest.mark.parametrize("sch", list(range(10)))
def test_limiters_against_sw2(tmp_path, sch):
    # van Leer turns NaN in the first step of this case and SW2's checks
    # abort it at step 3: two steps, NaN cells compared
    n = 2 if sch == 3 else 25
    p, state = jet(128, 48, n, sch=sch)
    r, o, failed = run_both(tmp_path, p, state, n)
    assert failed == 0
    assert_same(r, o, f"sch {sch}")
    if sch == 3:
        assert np.isnan(o.interior("v")).any()  # NaN cells compared above
    r.close()


def test_tracer_storm_with_dissipation_against_sw2(tmp_path):
    """A moving storm that injects tracer (the tracer then advects with its
    non-conservative term, calc.c:147-161) and tracer dissipation."""
    G = np.array([[0.0, 1e9, 120.0, 200.0, 15.0, 0, 25.0, 4e11, 1, 3.0, 6.0, 1.0],
                  [600.0, 1e9, 90.0, 355.0, -10.0, 1, 0.5, 2e11, 1, 2.5, 5.0, 1.0]])
    p, state = jet(160, 64, 60, Gaussians=G, TracerDis=2e-6)
    r, o, failed = run_both(tmp_path, p, state, 60)
    assert failed == 0
    assert r.interior("tracer").max() > 0 and np.count_nonzero(r.interior("tracer")) > 100
    assert_same(r, o, "tracer storm")
    assert r.energy() == o.energy()
    r.close()


def test_vortices_against_sw2(tmp_path):
    """Two vortices inserted during the run (perturbs.c:397-460), one across
    the periodic seam."""
    Vt = np.array([[100.0, 120.0, 20.0, 8.0, 5.0, -60.0, 2.0],
                   [250.0, 356.0, -25.0, 6.0, 4.0, 40.0, 1.5]])
    p, state = jet(160, 64, 40, vortices=Vt)
    r, o, failed = run_both(tmp_path, p, state, 40)
    assert failed == 0
    assert_same(r, o, "vortices")
    r.close()


def test_hyperviscosity_orders_and_drag_against_sw2(tmp_path):
    p, state = jet(144, 56, 30, hypernu=(2e5, 3e14, 1e22), Tau=2e5, sch=1)
    r, o, failed = run_both(tmp_path, p, state, 30)
    assert failed == 0
    assert_same(r, o, "del-2, del-4, del-6, drag")
    r.close()


def _golden_case(golden, sub, fname):
    cwd = os.getcwd()
    os.chdir(os.path.join(golden, sub))
    try:
        return sw2_amd.read_case(fname)
    finally:
        os.chdir(cwd)


def test_jupiter_example_case_against_sw2(tmp_path, golden):
    """SW2's JupiterEnergyTest case (superbee, drag, a moving storm), the
    first 400 of its 34,560 steps."""
    p = _golden_case(golden, "jupiter", "jupiters.sw")
    p.t1 = p.t0 + 400 * p.Dt
    z = np.zeros((p.ny, p.nx))
    r, o, failed = run_both(tmp_path, p, (z, z, z), 400)
    assert failed == 0
    assert_same(r, o, "jupiter")
    assert r.energy() == o.energy()
    r.close()


def test_saturn_example_case_against_sw2(tmp_path, golden):
    """SW2's SaturnSmoothInjection case: the 861-row Cassini wind table
    interpolated at every row (linterp, tools.c:21-74), 12 steps."""
    p = _golden_case(golden, "saturn", "minicanal.sw")
    p.t1 = p.t0 + 12 * p.Dt
    z = np.zeros((p.ny, p.nx))
    # the product's world as initialize() leaves it (wind table included) is SW2's
    r = refmod.Ref(p, (z, z, z), tmp_path, "saturn0")
    w = sw2_amd.World(p)
    for f in FIELDS:
        assert np.array_equal(w.field(f), r.field(f)), f
    w.close()
    r.close()
    r, o, failed = run_both(tmp_path, p, (z, z, z), 12)
    assert failed == 0
    assert_same(r, o, "saturn")
    r.close()


def test_linterp_closed_forms():
    """SW2's interpolation: exact at the nodes, the mean at midpoints of a
    uniform table, linear in between."""
    L = refmod.lib(os.getcwd())
    xs = np.linspace(-90.0, 90.0, 37)
    ys = np.cos(np.deg2rad(xs)) * 7.0 + xs * 0.25
    P = lambda a: a.ctypes.data_as(refmod.POINTER(refmod.c_double))
    for k, x in enumerate(xs):
        assert L.ref_linterp(float(x), P(xs), P(ys), len(xs)) == ys[k]
    mid = (xs[:-1] + xs[1:]) / 2
    got = np.array([L.ref_linterp(float(x), P(xs), P(ys), len(xs)) for x in mid])
    np.testing.assert_allclose(got, (ys[:-1] + ys[1:]) / 2, rtol=0, atol=1e-13)


@pytest.mark.slow
def test_config0_1000_steps_digest_equals_sw2(golden):
    """The 360x180 workload over its 1000 steps: the oracle's final state has
    the digest SW2's own run of the case has (tests/golden/digests.json,
    written by tests/golden/make_digests.py with the reference build)."""
    from tests.golden.make_digests import summary
    with open(os.path.join(golden, "digests.json")) as f:
        dg = json.load(f)["c0_1000_ref"]
    p = cases.config_params("c0")
    w, state = cases.make_world(p, cases.single_bump(2024))
    w.close()
    o = Oracle(p)
    o.load(*state)
    o.set_threads(8)
    failed, _ = o.step(1000)
    assert failed == dg["failed_step"] == 0
    for f in ("eta", "u", "v", "tracer", "perturbs"):
        assert summary(o.field(f))["sha256"] == dg[f]["sha256"], f

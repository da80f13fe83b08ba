"""SW2's per-step functions (sw.h:229-297) as the library exports them, on
the GPU, against the oracle's code for the same functions: every output bit
for bit (signs of zero aside), on states where every term is active (drag,
del-2/4/6 hyperviscosity, a tracer, a moving flow), for all ten limiters,
at the case's Dt and at another Dt, and -- for the in-place Coriolis
updates -- in SW2's order u then v.  Cells the reference function does not
write must keep what the caller had there."""
import ctypes

import numpy as np
import pytest

from oracle import Oracle
from sw2_amd import cases
from sw2_amd._ffi import lib

pytestmark = pytest.mark.gpu

P = ctypes.POINTER(ctypes.c_double)
FIELDS = ("eta", "u", "v", "tracer", "windx", "windy")


def ptr(a):
    return a.ctypes.data_as(P)


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def params(sch=2, nx=160, ny=64):
    return cases.scaled_params(nx, ny, steps=20, lat=75.0, sch=sch, Tau=2e6, TracerDis=1e-6,
                               hypernu=(3e4, 1e15 * (360.0 / nx) ** 4, 1e24))


def stepped_pair(sch=2, steps=7):
    """World and oracle in the same state: the oracle's after `steps` steps
    of the jet + storm case, plus a tracer field."""
    p = params(sch)
    w, (eta, u, v) = cases.make_world(p, cases.single_bump(3))
    o = Oracle(p)
    o.load(eta, u, v)
    assert o.step(steps)[0] == 0
    rng = np.random.default_rng(5)
    o.field("tracer")[:] = rng.uniform(0.0, 1.0, size=o.field("tracer").shape)
    for f in FIELDS:
        w.field(f)[:] = o.field(f)
    return p, w, o


def random_pair(sch, seed):
    """World and oracle holding the same random fields (every limiter ratio
    generic, so van Leer's 0/0 does not occur)."""
    p = params(sch, nx=130, ny=40)
    w, _ = cases.make_world(p, [])
    o = Oracle(p)
    rng = np.random.default_rng(seed)
    for f, scale in (("eta", 30.0), ("u", 8.0), ("v", 8.0), ("tracer", 1.0)):
        a = rng.standard_normal(o.field(f).shape) * scale
        o.field(f)[:] = a
        w.field(f)[:] = a
    for f in ("windx", "windy"):
        w.field(f)[:] = o.field(f)
    return p, w, o


def run_both(w, o, name, oname, nargs_out, *scalars, inputs=()):
    """Call the library's `name` and the oracle's `oname` with identical
    output arrays (prefilled with the same junk) and inputs; returns the
    outputs of both."""
    rng = np.random.default_rng(11)
    outs_p = [rng.standard_normal(w.S) for _ in range(nargs_out)]
    outs_o = [a.copy() for a in outs_p]
    ins_p = [np.ascontiguousarray(a) for a in inputs]
    ins_o = [a.copy() for a in ins_p]
    r1 = getattr(lib(), name)(*[ptr(a) for a in ins_p + outs_p], *scalars, ctypes.byref(w.sw))
    r2 = o.call(oname, *(ins_o + outs_o), *scalars)
    return outs_p, outs_o, r1, r2


@pytest.mark.parametrize("name,oname", [("solve_cons_h", "solve_cons_h"), ("solve_tracer", "solve_tracer"),
                                        ("solve_adv_u", "solve_adv_u"), ("solve_adv_v", "solve_adv_v"),
                                        ("solve_pres_u", "solve_pres_u"), ("solve_pres_v", "solve_pres_v")])
def test_tendencies_equal_oracle(name, oname):
    p, w, o = stepped_pair()
    for Dt in (p.Dt, 0.7 * p.Dt):
        a, b, _, _ = run_both(w, o, name, oname, 1, Dt)
        assert same(a[0], b[0]), (name, Dt, int(np.sum(a[0] != b[0])))
        assert np.any(a[0] != 0)


@pytest.mark.parametrize("sch", list(range(10)))
def test_flux_bodies_all_limiters(sch):
    p, w, o = random_pair(sch, 100 + sch)
    for name in ("solve_cons_h", "solve_tracer", "solve_adv_u", "solve_adv_v"):
        a, b, _, _ = run_both(w, o, name, name, 1, p.Dt)
        assert same(a[0], b[0]), (sch, name, int(np.sum(a[0] != b[0])))


def test_geofactor_dissipation_and_ab_equal_oracle():
    p, w, o = stepped_pair()
    rng = np.random.default_rng(2)
    for Dt in (p.Dt, 1.3 * p.Dt):
        d = rng.standard_normal(w.S)
        d_o = d.copy()
        lib().do_eta_geofactor(ptr(d), Dt, ctypes.byref(w.sw))
        o.call("do_eta_geofactor", d_o, Dt)
        assert same(d, d_o)
    du, dv = rng.standard_normal(w.S), rng.standard_normal(w.S)
    du_o, dv_o = du.copy(), dv.copy()
    lib().dissipation_effect(ptr(du), ptr(dv), ctypes.byref(w.sw))
    o.call("dissipation_effect", du_o, dv_o)
    assert same(du, du_o) and same(dv, dv_o)
    for k in (0, 1, 2, 7, -1):
        nab = rng.standard_normal(w.S)
        d, m1, m2 = (rng.standard_normal(w.S) for _ in range(3))
        nab_o = nab.copy()
        lib().do_AdamsBashforth(ptr(nab), ptr(d), ptr(m1), ptr(m2), k, w.sw.M)
        o.call("adams_bashforth", nab_o, d, m1, m2, k)
        assert same(nab, nab_o), k


def test_coriolis_u_then_v_equal_oracle():
    """In place: u from Du_nAB, then v from Dv_nAB and the new u (whose halo
    still holds the previous values, calc.c:591)."""
    p, w, o = stepped_pair()
    rng = np.random.default_rng(3)
    duab, dvab = rng.standard_normal(w.S) * 1e-2, rng.standard_normal(w.S) * 1e-2
    lib().solve_Coriolis_u(ptr(duab.copy()), p.Dt, ctypes.byref(w.sw))
    o.call("solve_coriolis_u", duab.copy(), p.Dt)
    assert same(w.field("u"), o.field("u"))
    lib().solve_Coriolis_v(ptr(dvab.copy()), p.Dt, ctypes.byref(w.sw))
    o.call("solve_coriolis_v", dvab.copy(), p.Dt)
    assert same(w.field("v"), o.field("v"))
    assert same(w.field("u"), o.field("u"))


def test_laplacians_equal_oracle():
    p, w, o = stepped_pair()
    rng = np.random.default_rng(4)
    uin, vin = w.field("u").ravel().copy(), w.field("v").ravel().copy()
    a, b, _, _ = run_both(w, o, "laplacian_vels", "laplacian_vels", 2, inputs=(uin, vin))
    assert same(a[0], b[0]) and same(a[1], b[1])
    ein = rng.standard_normal(w.S)
    a, b, _, _ = run_both(w, o, "laplacian_eta", "laplacian_eta", 1, inputs=(ein,))
    assert same(a[0], b[0])


def test_compute_dt_equal_oracle():
    p, w, o = stepped_pair()
    for courant in (0.25, 0.9):
        assert lib().compute_Dt(courant, ctypes.byref(w.sw)) == o.call("compute_dt", courant)
    # flow at rest: the gravity-wave speeds decide (calc.c:721-729)
    for f in ("u", "v", "windx", "windy"):
        w.field(f)[:] = 0.0
        o.field(f)[:] = 0.0
    assert lib().compute_Dt(0.25, ctypes.byref(w.sw)) == o.call("compute_dt", 0.25)

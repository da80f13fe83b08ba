"""The mixed-precision form (anomalies about an FP64 mean state, FP32 fluxes)
against the FP64 oracle: relative L2 and max errors of eta, u and v, printed for
every case and held to the stated tolerance; the device FP64 reductions of
energy and mass against the host's exact sums."""
import numpy as np
import pytest

from oracle import Oracle
from sw2_amd import FORM_FP64, FORM_MIXED, Session, SW2Error, World, cases

pytestmark = pytest.mark.gpu

# the mixed form's tolerance on eta, u and v against the FP64 result, both
# relative L2 (||a - b||_2 / ||b||_2) and relative max (max|a - b| / max|b|)
# over the interior
TOL = 1e-5


def errors(a, b):
    d = a - b
    nb = np.linalg.norm(b)
    l2 = np.linalg.norm(d) / nb if nb > 0 else np.linalg.norm(d)
    mb = np.max(np.abs(b))
    mx = np.max(np.abs(d)) / mb if mb > 0 else np.max(np.abs(d))
    return l2, mx


def run_pair(p, bumps, n):
    w, (eta, u, v) = cases.make_world(p, bumps)
    o = Oracle(p)
    o.load(eta, u, v)
    s = Session(w, FORM_MIXED)
    assert s.form == FORM_MIXED
    s.begin(0, p.t0)
    s.step(n)
    fs, msg = s.failed_step()
    s.download()
    fo, _ = o.step(n)
    return w, o, s, fs, fo


def report(label, w, ref, fields=("eta", "u", "v")):
    out = {}
    for f in fields:
        l2, mx = errors(w.interior(f), ref(f))
        out[f] = (l2, mx)
        print(f"{label}: {f} relative L2 {l2:.3e} max {mx:.3e}", flush=True)
    return out


def test_mixed_config0_1000_steps():
    """360x180 jet + storm, 1000 steps against the FP64 oracle."""
    p = cases.config_params("c0")
    w, o, s, fs, fo = run_pair(p, cases.single_bump(2024), 1000)
    assert fs == fo == 0
    e = report("mixed c0 1000 steps", w, o.interior)
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (f, l2, mx)


@pytest.mark.parametrize("sch", list(range(10)))
def test_mixed_all_limiters(sch):
    p = cases.scaled_params(128, 24, steps=40, lat=70.0, sch=sch)
    n = 40
    w, o, s, fs, fo = run_pair(p, cases.single_bump(sch), n)
    if sch == 3:
        # Van Leer's (r+|r|)/(1+r) is 0/0 at r = -1 (calc.c:57): on this jet
        # case SW2 turns NaN in the first step and its checks stop it at
        # step 3 (the FP64 parity tests reproduce that NaN bit for bit).  The
        # mixed form takes the limiter's value 0 at r = -1 and runs on: there
        # is no SW2 result to compare it with.
        assert fo > 0 and fs == 0, (fo, fs)
        for f in ("eta", "u", "v"):
            assert np.isfinite(w.interior(f)).all()
        return
    assert fs == fo == 0
    e = report(f"mixed sch={sch} 128x24 {n} steps", w, o.interior)
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (sch, f, l2, mx)


def test_mixed_against_fp64_gpu_at_4096x2048():
    """The mixed-precision workload at its size and step count (4096x2048,
    5000 steps): mixed against the FP64 GPU form, which the full-size tests
    hold bit-equal to the oracle."""
    n = 5000
    p = cases.config_params("c2", steps=n)
    bumps = cases.single_bump(2024)
    w64, _ = cases.make_world(p, bumps)
    s64 = Session(w64, FORM_FP64)
    s64.begin(0, p.t0)
    s64.step(n)
    s64.download()
    w32, _ = cases.make_world(p, bumps)
    s32 = Session(w32, FORM_MIXED)
    s32.begin(0, p.t0)
    s32.step(n)
    s32.download()
    assert s64.failed_step()[0] == s32.failed_step()[0] == 0
    e = report(f"mixed c2 {n} steps", w32, w64.interior)
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (f, l2, mx)
    ke64, ape64, m64 = s64.energy()
    ke32, ape32, m32 = s32.energy()
    print(f"energy fp64 {ke64:.6e} {ape64:.6e} mass {m64:.12e}; mixed {ke32:.6e} {ape32:.6e} mass {m32:.12e}", flush=True)
    assert abs(m32 - m64) <= 1e-12 * abs(m64)
    assert abs(ke32 - ke64) <= 1e-8 * abs(ke64)


def test_device_reductions_match_host_sums():
    p = cases.scaled_params(200, 60, steps=30, lat=70.0)
    w, _ = cases.make_world(p, cases.single_bump(5))
    s = Session(w, FORM_FP64)
    s.begin(0, p.t0)
    s.step(30)
    s.download()
    ke, ape, m = s.energy()
    hke, hape = w.energy()
    hm = w.mass()
    assert abs(ke - hke) <= 1e-12 * abs(hke)
    assert abs(ape - hape) <= 1e-12 * abs(hape)
    assert abs(m - hm) <= 1e-13 * abs(hm)


def test_mixed_refuses_cases_the_fused_step_does_not_take():
    p = cases.scaled_params(64, 32, steps=5, lat=60.0)  # narrower than 128 columns
    w, _ = cases.make_world(p, cases.single_bump(1))
    with pytest.raises(SW2Error, match="mixed-precision form"):
        Session(w, FORM_MIXED)
    # vortices and del-6 run in a mixed session, not in a batch
    from sw2_amd import Batch
    V = np.array([[100.0, 200.0, -20.0, 8.0, 5.0, -60.0, 2.0]])
    for kw in (dict(vortices=V), dict(hypernu=(0.0, 3e14, 1e22))):
        p = cases.scaled_params(128, 32, steps=5, lat=60.0, **kw)
        w, _ = cases.make_world(p, cases.single_bump(1))
        assert Session(w, FORM_MIXED).form == FORM_MIXED
        b = Batch(capacity=2)
        with pytest.raises(SW2Error, match="without vortices or del-6"):
            b.add(w, 5)
        b.close()


def test_mixed_del6_against_fp64():
    """Del-2, del-4 and del-6 together: the mixed step takes the three
    passes of both Laplacians (calc.c:324-351) from the per-phase FP64
    kernels, run before the step on an FP64 copy of u and v, and adds SW2's
    three terms in its order.  Within the tolerance of the FP64 run, whose
    fused step takes them the same way."""
    n = 100
    p = cases.scaled_params(144, 56, steps=n, lat=75.0, hypernu=(2e5, 3e14, 1e22))
    bumps = cases.single_bump(9)
    runs = {}
    for form in (FORM_FP64, FORM_MIXED):
        w, _ = cases.make_world(p, bumps)
        s = Session(w, form)
        assert s.form == form and s.fused
        s.begin(0, p.t0)
        s.step(n)
        s.download()
        runs[form] = (w, s.failed_step()[0])
    (w64, f64), (w32, f32) = runs[FORM_FP64], runs[FORM_MIXED]
    assert f64 == f32 == 0
    e = report(f"mixed del-6 {n} steps", w32, w64.interior)
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (f, l2, mx)


def test_mixed_del2_alone_against_fp64():
    """Del-2 hyperviscosity without del-4: the FP64 step takes its Laplacian
    from the per-phase kernels (SW2's sum carries no del-4 term at all); the
    mixed kernels form it in LDS with a zero del-4 coefficient.  Within the
    tolerance of the FP64 run."""
    n = 200
    p = cases.scaled_params(160, 64, steps=n, lat=75.0, hypernu=(2e5, 0.0, 0.0))
    bumps = cases.single_bump(5)
    runs = {}
    for form in (FORM_FP64, FORM_MIXED):
        w, _ = cases.make_world(p, bumps)
        s = Session(w, form)
        assert s.form == form and s.fused
        s.begin(0, p.t0)
        s.step(n)
        s.download()
        runs[form] = (w, s.failed_step()[0])
    (w64, f64), (w32, f32) = runs[FORM_FP64], runs[FORM_MIXED]
    assert f64 == f32 == 0
    e = report(f"mixed del-2 alone {n} steps", w32, w64.interior)
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (f, l2, mx)


def test_mixed_vortices_against_fp64():
    """Vortices inserted during the run (perturbs.c:397-460) in the mixed
    form: the increments added to eta', u', v' (value and residual) after
    the momentum kernel, whose wall and u, v checks wait for them, then the
    halo images, do_channel and the checks, and the next step's CFL maxima
    reduced again, as the fused FP64 step orders them.  One vortex crosses
    the periodic seam; one sits near the channel's southern half.  Against
    the FP64 GPU run (equal to the oracle) within the tolerance, and the
    vortices moved the flow."""
    Vt = np.array([[100.0, 120.0, 20.0, 8.0, 5.0, -60.0, 2.0],
                   [250.0, 356.0, -25.0, 6.0, 4.0, 40.0, 1.5]])
    n = 200
    p = cases.scaled_params(160, 64, steps=n, lat=75.0, vortices=Vt)
    bumps = cases.single_bump(160 + 64)
    runs = {}
    for form in (FORM_FP64, FORM_MIXED):
        w, _ = cases.make_world(p, bumps)
        s = Session(w, form)
        assert s.form == form and s.fused
        s.begin(0, p.t0)
        s.step(n)
        s.download()
        runs[form] = (w, s.failed_step()[0])
    (w64, f64), (w32, f32) = runs[FORM_FP64], runs[FORM_MIXED]
    assert f64 == f32 == 0
    e = report(f"mixed vortices {n} steps", w32, w64.interior)
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (f, l2, mx)
    w0, _ = cases.make_world(cases.scaled_params(160, 64, steps=n, lat=75.0), bumps)
    s0 = Session(w0, FORM_FP64)
    s0.begin(0, p.t0)
    s0.step(n)
    s0.download()
    assert errors(w64.interior("v"), w0.interior("v"))[0] > 100 * TOL


def test_mixed_tracer_storms_against_fp64():
    """Storms that inject tracer (perturbs.c:348-395) in the mixed form: the
    tracer advected with the FP32 flux-limited scheme of the height plus its
    non-conservative term and dissipation (calc.c:117-163), updated in FP64
    as value + residual, and written by the injections.  One storm moves
    across the periodic seam, one stops a third of the way in; tracer
    dissipation on.
    Against the FP64 GPU run (equal to the oracle): the same storm cells
    marked, eta, u, v within the tolerance, and the tracer within it too.
    (The second storm is a sink that also writes tracer.)"""
    G = np.array([[0.0, 1e5, 200.0, 358.0, 20.0, 1, 0.4, 3e12, 1, 3.0, 8.0, 1.0],
                  [300.0, 4000.0, 100.0, 150.0, -30.0, 0, -40.0, -1e12, 1, 2.5, 6.0, 1.0]])
    # limiter sch 0: with SW2's sch 2 (MIN(r,1), unbounded below) the tracer's
    # near-zero tails overflow the ratio and SW2 itself turns the tracer NaN
    # at step 82 of this case (the oracle, equal to SW2, does too)
    p = cases.scaled_params(512, 256, steps=300, lat=75.0, Gaussians=G, TracerDis=1e-6, sch=0)
    bumps = cases.single_bump(33)
    runs = {}
    for form in (FORM_FP64, FORM_MIXED):
        w, _ = cases.make_world(p, bumps)
        s = Session(w, form)
        assert s.form == form and s.fused
        s.begin(0, p.t0)
        s.step(300)
        s.download()
        runs[form] = (w, s.failed_step()[0])
    (w64, f64), (w32, f32) = runs[FORM_FP64], runs[FORM_MIXED]
    assert f64 == f32 == 0
    assert np.array_equal(w32.field("perturbs"), w64.field("perturbs"))
    assert w64.interior("tracer").max() > 0.5
    e = report("mixed tracer storms 300 steps", w32, w64.interior, fields=("eta", "u", "v", "tracer"))
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (f, l2, mx)


def test_mixed_eta_only_storms_against_fp64():
    """Storms that leave the tracer alone (inject 0, as in the reference's
    JupiterEnergyTest and SaturnSmoothInjection cases) in the mixed form:
    the same storm cells marked as the FP64 run's, and eta, u, v within the
    mixed form's tolerance of the FP64 GPU run, which the parity tests hold
    bit-equal to the oracle.  One storm crosses the periodic seam, one is a
    sink."""
    # the first storm is still injecting at the last step; the second stops
    # a third of the way in
    G = np.array([[0.0, 1e5, 200.0, 358.0, 20.0, 1, 0.4, 3e12, 1, 3.0, 8.0, 0.0],
                  [300.0, 4000.0, 100.0, 150.0, -30.0, 0, -40.0, -1e12, 1, 2.5, 6.0, 0.0]])
    p = cases.scaled_params(512, 256, steps=300, lat=75.0, Gaussians=G)
    bumps = cases.single_bump(31)
    runs = {}
    for form in (FORM_FP64, FORM_MIXED):
        w, _ = cases.make_world(p, bumps)
        s = Session(w, form)
        assert s.form == form and s.fused
        s.begin(0, p.t0)
        s.step(300)
        s.download()
        runs[form] = (w, s.failed_step()[0])
    (w64, f64), (w32, f32) = runs[FORM_FP64], runs[FORM_MIXED]
    assert f64 == f32 == 0
    assert np.array_equal(w32.field("perturbs"), w64.field("perturbs"))
    assert w64.field("perturbs").max() == 1.0
    e = report("mixed eta-only storms 300 steps", w32, w64.interior)
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (f, l2, mx)
    # the storms moved the height: the comparison is not of two quiet runs
    w0, _ = cases.make_world(cases.scaled_params(512, 256, steps=300, lat=75.0), bumps)
    s0 = Session(w0, FORM_FP64)
    s0.begin(0, p.t0)
    s0.step(300)
    s0.download()
    assert errors(w64.interior("eta"), w0.interior("eta"))[0] > 100 * TOL


def test_mixed_saturn_smooth_injection_case(golden):
    """The reference's SaturnSmoothInjection case (its storm injects height
    only) in the mixed form against the FP64 GPU run of the same case."""
    import os
    import sw2_amd
    cwd = os.getcwd()
    os.chdir(os.path.join(golden, "saturn"))
    try:
        p = sw2_amd.read_case("minicanal.sw")
    finally:
        os.chdir(cwd)
    o = Oracle(p)
    o.preloop()
    runs = {}
    for form in (FORM_FP64, FORM_MIXED):
        w = World(p)
        for f in ("eta", "u", "v", "tracer", "perturbs", "windx", "windy"):
            w.field(f)[:] = o.field(f)
        s = Session(w, form)
        assert s.form == form and s.fused
        s.begin(0, p.t0)
        s.step(100)
        s.download()
        runs[form] = (w, s.failed_step()[0])
    (w64, f64), (w32, f32) = runs[FORM_FP64], runs[FORM_MIXED]
    assert f64 == f32 == 0
    e = report("mixed minicanal.sw 100 steps", w32, w64.interior)
    for f, (l2, mx) in e.items():
        assert l2 <= TOL and mx <= TOL, (f, l2, mx)

"""The workloads at their own sizes and step counts.

  * FP64 step against the oracle, bit for bit: the 4096x2048 workload over
    its 2000 steps and one member of each ensemble size over 3000 steps,
    through digests of the oracle's final state (tests/golden/digests.json,
    tests/golden/make_digests.py): the oracle takes minutes at these sizes
    and its final fields are too large to commit.
  * The batched ensemble (one member of each size 256x128 .. 2048x1024,
    3000 steps, mixed form) and the served stream (requests of the served
    sizes 512x256 .. 4096x2048 arriving while others run) against the FP64
    runs of the same members, which the digests hold equal to the oracle:
    relative L2 and max error of eta, u and v, printed for every result.

Tolerance.  The relative L2 error is held to 1e-5 for every result.  The
mixed form stores its anomalies h', u', v' in FP32 every step and forms the
tendencies in FP32, so rounding accumulates step by step; the max error, a
pointwise measure set by a few cells at sharp features, grows with the step
count and with the time step of the coarser grids.  Measured (MI355X): L2 at
most 5.5e-6 for every result here; max below 1e-5 on 2048x1024 and
4096x2048, up to 2.9e-5 on the coarser members after 3000 steps and 5.4e-5
on a 512x256 request (Dt 42 s) after 1192 steps.  The max is therefore held
to 1e-4 here; the mixed-form tests at the single-run workloads
(tests/test_gpu_mixed.py: 360x180 over 1000 steps, 4096x2048 over 5000,
every limiter, storms) hold it to 1e-5.
"""
import json
import os

import numpy as np
import pytest

from sw2_amd import Batch, FORM_FP64, FORM_MIXED, Session, cases
from sw2_amd.serve import Server
from tests.golden.make_digests import summary

pytestmark = pytest.mark.gpu

TOL = 1e-5      # relative L2, ||a-b|| / ||b|| over the interior
TOL_MAX = 1e-4  # relative max, max|a-b| / max|b|, over these long runs (see above)
DIGESTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "digests.json")
FIELDS = ("eta", "u", "v")


def digests():
    with open(DIGESTS) as f:
        return json.load(f)


def rel(a, b):
    d = a - b
    return float(np.linalg.norm(d) / np.linalg.norm(b)), float(np.max(np.abs(d)) / np.max(np.abs(b)))


def fp64_run(p, bumps, n):
    w, _ = cases.make_world(p, bumps)
    s = Session(w, FORM_FP64)
    s.begin(0, p.t0)
    s.step(n)
    failed = s.failed_step()[0]
    s.download()
    s.close()
    out = {f: w.field(f).copy() for f in ("eta", "u", "v", "tracer", "perturbs")}
    w.close()
    return failed, out


def check_digest(name, failed, out):
    dg = digests()[name]
    assert failed == dg["failed_step"] == 0
    fields = [f for f in ("eta", "u", "v", "tracer", "perturbs") if f in dg]
    assert fields[:3] == ["eta", "u", "v"]
    for f in fields:
        got = summary(out[f])
        print(f"{name} {f}: l2 {got['l2']:.15e} (oracle {dg[f]['l2']:.15e})", flush=True)
        assert got["sha256"] == dg[f]["sha256"], (name, f)


def test_fp64_4096x2048_2000_steps_equals_oracle():
    p = cases.config_params("c1", steps=2000)
    failed, out = fp64_run(p, cases.single_bump(2024), 2000)
    check_digest("c1_2000", failed, out)


def ensemble_members():
    """The first member of each size of the 64-member ensemble."""
    seen, out = set(), []
    for p, bumps in cases.ensemble(64, 3000):
        if (p.nx, p.ny) not in seen:
            seen.add((p.nx, p.ny))
            out.append((p, bumps))
    assert len(out) == 4
    return out


@pytest.fixture(scope="module")
def fp64_members():
    res = {}
    for p, bumps in ensemble_members():
        res[(p.nx, p.ny)] = fp64_run(p, bumps, 3000)
    return res


def test_fp64_ensemble_members_3000_steps_equal_oracle(fp64_members):
    for (nx, ny), (failed, out) in fp64_members.items():
        check_digest(f"ens_{nx}x{ny}", failed, out)


def test_batched_ensemble_members_3000_steps_within_tolerance(fp64_members):
    """One member of each ensemble size stepped together in one batch for
    the ensemble's 3000 steps (graph-replayed step), each against its FP64
    run."""
    b = Batch(capacity=4)
    ids = []
    for p, bumps in ensemble_members():
        w, _ = cases.make_world(p, bumps)
        ids.append((b.add(w, 3000), (p.nx, p.ny)))
    b.step(3000)
    b.sync()
    bad = []
    for mid, sz in ids:
        st = b.member(mid)
        assert st["done"] and st["steps"] == 3000 and st["failed_step"] == 0, st
        w = b.download(mid)
        ref = fp64_members[sz][1]
        for f in FIELDS:
            l2, mx = rel(w.interior(f), ref[f][2:-2, 2:-2])
            print(f"batch member {sz[0]}x{sz[1]} 3000 steps {f}: L2 {l2:.3e} max {mx:.3e}", flush=True)
            if not (l2 <= TOL and mx <= TOL_MAX):
                bad.append((sz, f, l2, mx))
    b.close()
    assert not bad, bad


def test_served_stream_sizes_within_tolerance():
    """Requests of every served size, 4096x2048 included, arriving while
    others are on the device (Poisson, seeded), with shortened step counts;
    each result against the FP64 run of the same request."""
    rng = np.random.default_rng(404)
    sizes = [(512, 256), (4096, 2048), (1024, 512), (2048, 1024), (512, 256), (1024, 512)]
    reqs, t = [], 0.0
    for nx, ny in sizes:
        t += float(rng.exponential(0.25))
        reqs.append((t, nx, ny, int(rng.integers(500, 1500)), cases.storm_bumps(rng)))
    got = {}

    def keep(k, batch, mid):
        w = batch.download(mid)
        got[k] = {f: w.interior(f).copy() for f in FIELDS}

    srv = Server(capacity=8, tick_steps=40)
    st = srv.run(reqs, on_result=keep)
    rec = srv.records
    srv.close()
    print(st, flush=True)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    spans = sorted((r["admitted"], r["completed"]) for r in rec.values())
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    bad = []
    for k, (_, nx, ny, steps, bumps) in enumerate(reqs):
        assert rec[k]["steps"] == steps
        failed, ref = fp64_run(cases.member_params(nx, ny, steps), bumps, steps)
        assert failed == 0
        for f in FIELDS:
            l2, mx = rel(got[k][f], ref[f][2:-2, 2:-2])
            print(f"served request {k} {nx}x{ny} {steps} steps {f}: L2 {l2:.3e} max {mx:.3e}", flush=True)
            if not (l2 <= TOL and mx <= TOL_MAX):
                bad.append((k, f, l2, mx))
    assert not bad, bad

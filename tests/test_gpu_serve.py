"""The served stream: requests arriving on their own schedule join a running
batch (continuous batching) and leave when done.  Each result must be the one
the same request gives alone -- bit for bit, since a member runs the same
arithmetic in the mix as alone -- and the runs must overlap on the device."""
import numpy as np
import pytest

from sw2_amd import FORM_MIXED, Session, cases
from sw2_amd.serve import Server

pytestmark = pytest.mark.gpu


def solo(nx, ny, steps, bumps):
    w, _ = cases.make_world(cases.member_params(nx, ny, steps), bumps)
    s = Session(w, FORM_MIXED)
    s.begin(0, w.params.t0)
    s.step(steps)
    assert s.failed_step()[0] == 0
    s.download()
    out = {f: w.field(f).copy() for f in ("eta", "u", "v")}
    s.close()
    w.close()
    return out


@pytest.mark.parametrize("policy", ["srpt", "all"])
def test_served_results_equal_solo_runs(policy):
    rng = np.random.default_rng(11)
    sizes = [(128, 64), (256, 128), (192, 40), (512, 256)]
    reqs = []
    t = 0.0
    for k in range(9):
        t += float(rng.exponential(0.05))
        nx, ny = sizes[k % len(sizes)]
        reqs.append((t, nx, ny, int(rng.integers(30, 120)), cases.storm_bumps(rng)))
    got = {}

    def keep(k, batch, mid):
        w = batch.download(mid)
        got[k] = {f: w.field(f).copy() for f in ("eta", "u", "v")}

    srv = Server(capacity=3, tick_steps=9, policy=policy)
    st = srv.run(reqs, on_result=keep)
    rec = srv.records
    srv.close()
    print(st, flush=True)
    assert st["completed"] == len(reqs) and st["failed"] == 0 and st["policy"] == policy
    assert st["p50_s"] <= st["p99_s"]
    # continuous batching: some request was admitted while another was still running
    spans = sorted((r["admitted"], r["completed"]) for r in rec.values())
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for k, (_, nx, ny, steps, bumps) in enumerate(reqs):
        assert rec[k]["steps"] == steps
        ref = solo(nx, ny, steps, bumps)
        for f in ("eta", "u", "v"):
            assert np.array_equal(got[k][f], ref[f]), (k, f)


def test_srpt_holds_long_runs_and_results_stay_solo():
    """With a tile budget smaller than the admitted runs, the scheduler
    advances the shortest remaining work first and holds the others; the
    held runs resume exactly where they stopped."""
    rng = np.random.default_rng(12)
    reqs = [(0.0, 512, 256, 90, cases.storm_bumps(rng)), (0.0, 512, 256, 30, cases.storm_bumps(rng)),
            (0.0, 256, 128, 60, cases.storm_bumps(rng))]
    got = {}

    def keep(k, batch, mid):
        w = batch.download(mid)
        got[k] = {f: w.field(f).copy() for f in ("eta", "u", "v")}

    srv = Server(capacity=3, tick_steps=10, policy="srpt", fill=0)
    srv.fill_tiles = 1  # one run at a time: the shortest remaining work
    st = srv.run(reqs, on_result=keep)
    rec = srv.records
    srv.close()
    assert st["completed"] == 3
    order = sorted(rec, key=lambda k: rec[k]["completed"])
    assert order == [2, 1, 0], order  # cell-steps: 60 x 256x128 < 30 x 512x256 < 90 x 512x256
    for k, (_, nx, ny, steps, bumps) in enumerate(reqs):
        ref = solo(nx, ny, steps, bumps)
        for f in ("eta", "u", "v"):
            assert np.array_equal(got[k][f], ref[f]), (k, f)


def test_served_requests_with_in_loop_storms_equal_solo_runs():
    """Requests carrying in-loop storms (one injecting tracer, one height
    only) beside requests without: the storm members step without the graph
    and inject after every step; every result, tracer and storm marks
    included, equals its solo run."""
    storm = np.array([[0.0, 1e5, 150.0, 359.0, 15.0, 1, 0.5, 2e12, 1, 3.0, 8.0, 1.0]])
    sink = np.array([[50.0, 2500.0, 100.0, 120.0, -25.0, 0, -30.0, -8e11, 1, 2.5, 6.0, 0.0]])
    rng = np.random.default_rng(13)
    reqs = [(0.0, 256, 128, 80, cases.storm_bumps(rng), storm), (0.02, 512, 256, 60, cases.storm_bumps(rng)),
            (0.04, 384, 192, 70, cases.storm_bumps(rng), sink), (0.06, 256, 128, 50, cases.storm_bumps(rng))]
    got = {}

    def keep(k, batch, mid):
        w = batch.download(mid)
        got[k] = {f: w.field(f).copy() for f in ("eta", "u", "v", "tracer", "perturbs")}

    srv = Server(capacity=4, tick_steps=10, policy="srpt")
    st = srv.run(reqs, on_result=keep)
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for k, req in enumerate(reqs):
        _, nx, ny, steps, bumps = req[:5]
        kw = {} if len(req) < 6 else {"Gaussians": req[5]}
        w, _ = cases.make_world(cases.member_params(nx, ny, steps, **kw), bumps)
        s = Session(w, FORM_MIXED)
        s.begin(0, w.params.t0)
        s.step(steps)
        assert s.failed_step()[0] == 0
        s.download()
        for f in ("eta", "u", "v", "tracer", "perturbs"):
            assert np.array_equal(got[k][f], w.field(f)), (k, f)
        s.close()
        w.close()
    assert got[0]["perturbs"].max() == 1.0
    assert got[0]["tracer"].max() > 0.5

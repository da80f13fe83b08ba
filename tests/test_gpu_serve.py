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

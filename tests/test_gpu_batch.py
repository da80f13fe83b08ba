"""Batches: worlds of uneven sizes stepped together (one persistent launch
per phase, the step replayed from a captured hipGraph, members joining and
leaving between steps).  Every member's result must be the one the same
world gives alone in a mixed-precision session -- bit for bit, since the
member runs the same arithmetic -- and, through that, within the mixed
form's tolerance of the FP64 oracle (tests/test_gpu_mixed.py)."""
import numpy as np
import pytest

from sw2_amd import Batch, FORM_MIXED, Session, SW2Error, cases

pytestmark = pytest.mark.gpu

FIELDS = ("eta", "u", "v")


def solo(p, bumps, n):
    w, _ = cases.make_world(p, bumps)
    s = Session(w, FORM_MIXED)
    s.begin(0, p.t0)
    s.step(n)
    assert s.failed_step()[0] == 0
    s.download()
    return w


def same(a, b, label):
    for f in FIELDS:
        x, y = a.field(f), b.field(f)
        assert np.array_equal(x, y), (label, f, int(np.sum(x != y)))


def small_members(seed, n):
    rng = np.random.default_rng(seed)
    sizes = [(128, 64), (256, 128), (192, 40), (512, 256)]
    out = []
    for k in range(n):
        nx, ny = sizes[k % len(sizes)]
        steps = int(rng.integers(20, 60))
        out.append((cases.member_params(nx, ny, steps), cases.storm_bumps(rng), steps))
    return out


def test_batch_members_equal_solo_runs():
    members = small_members(1, 6)
    b = Batch(capacity=8)
    ids = []
    for p, bumps, steps in members:
        w, _ = cases.make_world(p, bumps)
        ids.append(b.add(w, steps))
    print("launch", b.launch(), flush=True)
    b.step(max(m[2] for m in members) + 5)  # members stop at their own last step
    b.sync()
    for mid, (p, bumps, steps) in zip(ids, members):
        st = b.member(mid)
        assert st["done"] and st["steps"] == steps and st["failed_step"] == 0, st
        same(b.download(mid), solo(p, bumps, steps), f"member {mid}")
    assert sorted(b.finished()) == sorted(ids)
    b.close()


def test_graph_and_direct_launches_agree():
    p = cases.member_params(256, 128, 40)
    bumps = cases.storm_bumps(np.random.default_rng(3))
    ref = solo(p, bumps, 37)
    b = Batch(capacity=2)
    b.set_launch(97, 131, 5)  # odd persistent grids; 37 = 7 graphs of 5 + 2 direct steps
    w, _ = cases.make_world(p, bumps)
    mid = b.add(w, 37)
    b.step(37)
    b.sync()
    assert b.member(mid)["steps"] == 37
    same(b.download(mid), ref, "graph+direct")
    b.close()


def test_continuous_batching_join_and_leave():
    """Members join while others run and leave when done; each result is
    the solo result."""
    members = small_members(7, 7)
    b = Batch(capacity=4)
    pending = list(enumerate(members))
    live = {}
    results = {}
    ticks = 0
    while pending or live:
        while pending:
            k, (p, bumps, steps) = pending[0]
            w, _ = cases.make_world(p, bumps)
            mid = b.add(w, steps)
            if mid is None:
                break
            live[mid] = k
            pending.pop(0)
        b.step(7)
        ticks += 1
        for mid in b.finished():
            k = live.pop(mid)
            results[k] = b.download(mid)
            b.remove(mid)
        assert ticks < 200
    for k, (p, bumps, steps) in enumerate(members):
        same(results[k], solo(p, bumps, steps), f"request {k}")
    b.close()


def test_members_admitted_while_steps_run():
    """sw2_batch_add does not wait for the batch's queued steps: members
    admitted right after a sw2_batch_step (a replayed graph still running,
    or not yet started) join at the next step, and every member, early or
    late, ends on its solo result."""
    members = small_members(5, 6)
    b = Batch(capacity=8)
    b.set_launch(0, 0, 8)
    ids = []
    for p, bumps, steps in members[:3]:
        w, _ = cases.make_world(p, bumps)
        ids.append(b.add(w, steps))
    b.step(16)  # returns at once; the graph runs while the next members are admitted
    late = []
    for p, bumps, steps in members[3:]:
        w, _ = cases.make_world(p, bumps)
        late.append(b.add(w, steps))
    assert all(b.member(mid)["steps"] == 0 for mid in late)  # joined, not yet stepped
    assert all(b.member(mid)["steps"] == 16 for mid in ids)
    b.step(max(m[2] for m in members) + 5)
    b.sync()
    for mid, (p, bumps, steps) in zip(ids + late, members):
        st = b.member(mid)
        assert st["done"] and st["steps"] == steps and st["failed_step"] == 0, st
        same(b.download(mid), solo(p, bumps, steps), f"member {mid}")
    b.close()


def test_admission_by_hbm_footprint():
    p = cases.member_params(512, 256, 10)
    w, _ = cases.make_world(p, [])
    foot = Batch.footprint(w)
    b = Batch(capacity=8, hbm_budget=2.5 * foot)
    ids = [b.add(cases.make_world(p, [])[0], 10) for _ in range(3)]
    assert ids[0] is not None and ids[1] is not None and ids[2] is None
    assert b.count() == (2, 2 * foot)
    b.remove(ids[0])
    assert b.add(cases.make_world(p, [])[0], 10) is not None
    b.close()


def test_batch_members_with_storms_equal_solo_runs():
    """Members with height-only storms (one crossing the seam and running
    past the member's last step, one a sink whose window closes mid-run) of
    different sizes, beside a member without storms.  One storm member is
    held for 20 steps and released.  Each member's result, storm marks
    included, equals the same world run alone in a mixed session."""
    G1 = np.array([[0.0, 1e5, 200.0, 358.0, 20.0, 1, 0.4, 3e12, 1, 3.0, 8.0, 0.0]])
    G2 = np.array([[100.0, 3000.0, 100.0, 150.0, -30.0, 0, -40.0, -1e12, 1, 2.5, 6.0, 0.0]])
    specs = [(256, 128, G1, 120), (512, 256, G2, 90), (384, 192, None, 100)]
    b = Batch(capacity=4)
    members = []
    for nx, ny, G, n in specs:
        kw = {} if G is None else {"Gaussians": G}
        p = cases.scaled_params(nx, ny, steps=n, lat=70.0, **kw)
        bumps = cases.single_bump(nx)
        w, _ = cases.make_world(p, bumps)
        members.append((b.add(w, n), p, bumps, n))
    b.step(30)
    b.sync()
    b.pause(members[0][0])
    b.step(20)
    b.sync()
    assert b.member(members[0][0])["steps"] == 30
    b.pause(members[0][0], False)
    b.step(200)  # members stop at their own last step
    b.sync()
    for mid, p, bumps, n in members:
        st = b.member(mid)
        assert st["done"] and st["steps"] == n and st["failed_step"] == 0, st
        w = b.download(mid)
        ref = solo(p, bumps, n)
        same(w, ref, f"storm member {mid}")
        assert np.array_equal(w.field("perturbs"), ref.field("perturbs"))
    assert b.download(members[0][0]).field("perturbs").max() == 1.0
    b.close()


def test_batch_members_with_tracer_storms_equal_solo_runs():
    """A member whose storms inject tracer (advected, dissipated, written by
    the injections) beside a height-only storm member and a quiet one, on
    the default and on 64 x 16 tiles: every member's eta, u, v, tracer and
    storm marks equal its solo mixed run bit for bit."""
    G1 = np.array([[0.0, 1e5, 200.0, 358.0, 20.0, 1, 0.4, 3e12, 1, 3.0, 8.0, 1.0],
                   [50.0, 3000.0, 100.0, 150.0, -30.0, 0, -40.0, -1e12, 1, 2.5, 6.0, 1.0]])
    G2 = np.array([[100.0, 3000.0, 100.0, 150.0, -30.0, 0, -40.0, -1e12, 1, 2.5, 6.0, 0.0]])
    specs = [(256, 128, G1, 110), (512, 256, G2, 90), (384, 192, None, 100)]
    for tiles in ((32, 16), (64, 16)):
        b = Batch(capacity=4)
        b.set_tiles(*tiles)
        members = []
        for nx, ny, G, n in specs:
            kw = {} if G is None else {"Gaussians": G, "TracerDis": 1e-6}
            p = cases.scaled_params(nx, ny, steps=n, lat=70.0, sch=0, **kw)  # one limiter per batch
            bumps = cases.single_bump(nx + 1)
            w, _ = cases.make_world(p, bumps)
            members.append((b.add(w, n), p, bumps, n))
        b.step(200)
        b.sync()
        for mid, p, bumps, n in members:
            st = b.member(mid)
            assert st["done"] and st["steps"] == n and st["failed_step"] == 0, st
            w = b.download(mid)
            ref = solo(p, bumps, n)
            same(w, ref, f"member {mid} on {tiles} tiles")
            for f in ("tracer", "perturbs"):
                assert np.array_equal(w.field(f), ref.field(f)), (f, mid, tiles)
        assert b.download(members[0][0]).interior("tracer").max() > 0.5
        b.close()


def test_all_members_held_then_released():
    """Every member held: the launches number no tiles, and the steps
    change nothing.  Released, one member mid-table held for a while (its
    tiles out of the numbering, its neighbours' renumbered), every member
    ends on its solo result, through graph replays and direct launches."""
    members = small_members(11, 5)
    b = Batch(capacity=6)
    b.set_launch(0, 0, 4)
    ids = []
    for p, bumps, steps in members:
        w, _ = cases.make_world(p, bumps)
        ids.append(b.add(w, steps))
    b.step(6)
    b.sync()
    for mid in ids:
        b.pause(mid)
    b.step(9)
    b.sync()
    assert all(b.member(mid)["steps"] == 6 for mid in ids)
    for mid in ids:
        b.pause(mid, False)
    b.pause(ids[2])
    b.step(10)
    b.sync()
    assert b.member(ids[2])["steps"] == 6 and b.member(ids[0])["steps"] == 16
    b.pause(ids[2], False)
    b.step(max(m[2] for m in members) + 5)
    b.sync()
    for mid, (p, bumps, steps) in zip(ids, members):
        st = b.member(mid)
        assert st["done"] and st["steps"] == steps and st["failed_step"] == 0, st
        same(b.download(mid), solo(p, bumps, steps), f"member {mid}")
    b.close()


def test_storm_member_failing_inside_a_replayed_window():
    """A storm member whose storm injects a height beyond SW2's range check
    (check_stats_scaf, sppde_core.c:377-395) at step 21, in the middle of
    the batch's second replayed graph (16 steps each): the late height check
    flags that step, the member stops there with its state of that step, and
    a quiet member and a healthy storm member beside it run on.  Every
    member's failed step, fields and storm marks equal its solo run's."""
    def run_solo(p, bumps, n):
        w, _ = cases.make_world(p, bumps)
        s = Session(w, FORM_MIXED)
        s.begin(0, p.t0)
        s.step(n)
        s.download()
        return w, s.failed_step()[0]

    n = 48
    base = cases.scaled_params(256, 128, steps=n, lat=70.0)
    t_on = base.t0 + 20.5 * base.Dt  # active from step 21 (t = t0 + 21 Dt), half a step in
    G_bad = np.array([[t_on, t_on + 1e6, 1.0, 120.0, 20.0, 1, 0.0, 1e30, 1, 3.0, 8.0, 0.0]])
    G_ok = np.array([[0.0, 1e6, 200.0, 300.0, -20.0, 1, 0.4, 3e12, 1, 3.0, 8.0, 0.0]])
    specs = [("bad", dict(Gaussians=G_bad)), ("ok", dict(Gaussians=G_ok)), ("quiet", {})]
    b = Batch(capacity=4)
    b.set_launch(0, 0, 16)
    members = []
    for k, (name, kw) in enumerate(specs):
        p = cases.scaled_params(256, 128, steps=n, lat=70.0, **kw)
        bumps = cases.single_bump(7 + k)
        w, _ = cases.make_world(p, bumps)
        members.append((name, b.add(w, n), p, bumps))
    b.step(n)
    b.sync()
    for name, mid, p, bumps in members:
        ref, f_ref = run_solo(p, bumps, n)
        st = b.member(mid)
        assert st["done"] and st["failed_step"] == f_ref, (name, st, f_ref)
        if name == "bad":
            assert f_ref == 21, f_ref
        else:
            assert f_ref == 0 and st["steps"] == n, (name, st)
        w = b.download(mid)
        same(w, ref, f"{name} member")
        assert np.array_equal(w.field("perturbs"), ref.field("perturbs")), name
    b.close()


@pytest.mark.parametrize("tiles", [(64, 8), (64, 16)])
def test_tile_shapes_equal_solo_runs(tiles):
    """The batch kernels on taller tiles: every member (ragged sizes, storms,
    one held for a while, the shape switched mid-run) equals its solo run
    bit for bit -- a cell's arithmetic does not depend on the tile that
    computes it."""
    members = small_members(7, 6)
    G = np.array([[0.0, 1e9, 100.0, 40.0, 20.0, 0, 30.0, 2e11, 1, 3.0, 8.0, 0.0]])
    p_storm = cases.member_params(256, 128, 50, Gaussians=G)
    members.append((p_storm, cases.storm_bumps(np.random.default_rng(3)), 50))
    b = Batch(capacity=8)
    default = b.tiles()
    assert default == (32, 16)
    b.set_tiles(*tiles)
    assert b.tiles() == tiles
    ids = []
    for p, bumps, steps in members:
        w, _ = cases.make_world(p, bumps)
        ids.append(b.add(w, steps))
    b.pause(ids[1])
    b.step(10)
    b.pause(ids[1], False)
    b.set_tiles(*default)  # between steps, members included
    b.step(7)
    b.set_tiles(*tiles)
    b.step(max(m[2] for m in members) + 5)
    b.sync()
    for mid, (p, bumps, steps) in zip(ids, members):
        st = b.member(mid)
        assert st["done"] and st["steps"] == steps and st["failed_step"] == 0, st
        same(b.download(mid), solo(p, bumps, steps), f"member {mid} on {tiles} tiles")
    with pytest.raises(SW2Error, match="persistent workgroups"):
        b.set_launch(256, 0, 8)
    with pytest.raises(SW2Error, match="no batch kernels"):
        b.set_tiles(48, 8)
    b.close()

"""Serving SW2 runs: a continuous-batching request loop over sw2_batch_*.

Requests (arrival time, grid, steps, initial storm bumps, and optionally a
table of in-loop storms in SW2's Gaussian-table columns) arrive on their own
schedule.  The loop admits a waiting request as soon as a member slot and
the HBM budget allow (sw2_batch_add checks the member's footprint), keeps
every admitted run's state on the device, advances all of them together in
ticks of a few steps (one persistent launch per phase for every member,
replayed from the captured step graph), and retires a run the tick it ends.
Requests that arrive while others are on the GPU join the running batch at
the next tick; their work overlaps with the runs already there.  A request's
world (geometry, wind table, storm bumps) is built on the host by a small
pool of builder threads as soon as it arrives, while the device runs the
current tick, so admission only uploads it.

Scheduling.  With policy "srpt" (shortest remaining processing time, the
default) every tick advances the admitted runs with the least remaining
work (steps left x cells), as many as it takes to give the persistent
launches `fill` tiles per resident workgroup, and holds the rest
(sw2_batch_pause): under overload short runs finish early instead of all
runs finishing late.  Policy "all" advances every admitted run every tick
(processor sharing).

Reported: time-to-completion (arrival to retirement) p50 / p99, runs per
hour, cell-updates per second and the fraction of time the device had work.
"""
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import Batch, cases


def _build(req):
    _, nx, ny, steps, bumps = req[:5]
    storms = req[5] if len(req) > 5 else None
    kw = {} if storms is None else {"Gaussians": np.asarray(storms, dtype=float)}
    w, _ = cases.make_world(cases.member_params(nx, ny, steps, **kw), bumps)
    return w


class Server:
    def __init__(self, capacity=64, hbm_budget=0.0, tick_steps=40, launch=None, builders=2, prefetch=8,
                 policy="srpt", fill=4):
        if policy not in ("srpt", "all"):
            raise ValueError(f"policy {policy!r}: 'srpt' or 'all'")
        self.batch = Batch(capacity=capacity, hbm_budget=hbm_budget)
        if launch:  # (blocks_eta, blocks_uv, steps_per_graph[, tile_x, tile_y]), as the autotuner gives it
            from . import tune
            if len(launch) == 5:
                tune.apply(self.batch, launch)
            else:
                self.batch.set_launch(*launch)
        self.tick = tick_steps
        self.pool = ThreadPoolExecutor(max_workers=builders)
        self.prefetch = prefetch
        self.policy = policy
        # tiles that keep the device busy: `fill` per workgroup that can be
        # resident at once (four of each batch kernel per compute unit), or
        # per persistent workgroup when the launch shape sets them
        from . import tune
        resident = max(self.batch.launch()[:2]) or 4 * tune.device_cus()
        self.fill_tiles = fill * resident

    def close(self):
        self.pool.shutdown(wait=True)
        self.batch.close()

    def _prefetch(self, waiting):
        """Start building the worlds of the first `prefetch` waiting requests."""
        for wq in list(waiting)[: self.prefetch]:
            if wq[2] is None:
                wq[2] = self.pool.submit(_build, wq[1])

    def _admit(self, waiting, live, left, rejected, t0):
        """Admit waiting requests in arrival order.  With runs on the device
        only worlds already built are taken; with none the head is waited
        for.  A request that does not fit even an empty batch is rejected."""
        while waiting:
            self._prefetch(waiting)
            k, req, fut = waiting[0]
            ta, nx, ny, steps = req[:4]
            if live and not fut.done():
                break
            w = fut.result()
            mid = self.batch.add(w, steps)
            if mid is None and not live:
                # no room even in an empty batch: it can never run here
                waiting.popleft()
                w.close()
                rejected.append(k)
                continue
            if mid is None:
                break
            waiting.popleft()
            live[mid] = (k, ta, time.perf_counter() - t0)
            left[mid] = (steps, nx * ny, ((nx + 63) // 64 + 1) * ((ny + 7) // 8))

    def run(self, requests, on_result=None, max_wall=None, log=None):
        """Serve `requests` = [(t_arrival_s, nx, ny, steps, bumps[, storms]), ...] in
        real time; returns the statistics (and calls on_result(k, batch,
        member_id) before each run is retired, for callers that keep the
        results).  Waiting requests are admitted in arrival order; the
        first `prefetch` of them have their worlds built ahead."""
        reqs = sorted(enumerate(requests), key=lambda r: r[1][0])
        waiting = deque()  # [request index, request, future or None]
        live = {}  # member id -> (request index, arrival, admitted)
        left = {}  # member id -> (steps left, cells, tiles)
        held = set()
        done = {}
        rejected = []
        cells_steps = 0.0
        busy = 0.0
        t0 = time.perf_counter()
        nxt = 0
        while nxt < len(reqs) or waiting or live:
            now = time.perf_counter() - t0
            if max_wall and now > max_wall:
                break
            while nxt < len(reqs) and reqs[nxt][1][0] <= now:
                waiting.append([reqs[nxt][0], reqs[nxt][1], None])
                nxt += 1
            self._prefetch(waiting)
            if not live:
                # nothing on the device: admit in arrival order, waiting for
                # the head's world to be built
                self._admit(waiting, live, left, rejected, t0)
                self._prefetch(waiting)
                if not live:
                    if nxt < len(reqs):
                        time.sleep(max(0.0, min(0.05, reqs[nxt][1][0] - (time.perf_counter() - t0))))
                    continue
            active = self._schedule(left, held)
            tb = time.perf_counter()
            self.batch.step(self.tick)
            # while the tick runs: admit the waiting requests whose worlds
            # are built (their uploads overlap the tick; they step from the
            # next one)
            self._admit(waiting, live, left, rejected, t0)
            self._prefetch(waiting)
            fin = self.batch.finished()  # waits for the tick
            busy += time.perf_counter() - tb
            tnow = time.perf_counter() - t0
            if log and fin:
                log(f"t={tnow:.1f}s: {len(fin)} done, {len(done) + len(fin)}/{len(requests)} completed, "
                    f"{len(live) - len(fin)} admitted ({len(active)} advancing), {len(waiting)} waiting")
            for mid in active:
                st_, c_, t_ = left[mid]
                left[mid] = (max(0, st_ - self.tick), c_, t_)
            for mid in fin:
                cells = left.pop(mid, (0, 0, 0))[1]
                held.discard(mid)
                k, ta, tadm = live.pop(mid)
                st = self.batch.member(mid)
                cells_steps += float(cells) * st["steps"]  # the steps it ran, not the ones it asked for
                if on_result:
                    on_result(k, self.batch, mid)
                done[k] = {"arrival": ta, "admitted": tadm, "completed": tnow, "steps": st["steps"],
                           "failed_step": st["failed_step"]}
                self.batch.remove(mid)
        for mid in live:  # runs still admitted when max_wall stopped the loop: the steps they ran
            cells_steps += float(left[mid][1]) * self.batch.member(mid)["steps"]
        for wq in waiting:  # worlds built for requests never admitted (max_wall)
            if wq[2] is not None:
                wq[2].result().close()
        wall = time.perf_counter() - t0
        self.records = done
        lat = np.array([d["completed"] - d["arrival"] for d in done.values()]) if done else np.array([np.nan])
        wait = np.array([d["admitted"] - d["arrival"] for d in done.values()]) if done else np.array([np.nan])
        return {
            "requests": len(requests), "completed": len(done), "rejected": len(rejected),
            "failed": sum(1 for d in done.values() if d["failed_step"]),
            "wall_s": wall, "runs_per_hour": len(done) / wall * 3600.0,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "queue_wait_p50_s": float(np.percentile(wait, 50)),
            "cell_updates_per_s": cells_steps / wall, "device_busy_frac": busy / wall,
            "policy": self.policy,
        }

    def _schedule(self, left, held):
        """Choose the runs the next tick advances and hold the others;
        returns the advancing member ids."""
        ids = sorted(left, key=lambda m: (left[m][0] * left[m][1], m))
        if self.policy == "all":
            active = ids
        else:
            active, tiles = [], 0
            for m in ids:
                if tiles >= self.fill_tiles:
                    break
                active.append(m)
                tiles += left[m][2]
        on = set(active)
        for m in ids:
            if m in on and m in held:
                self.batch.pause(m, False)
                held.discard(m)
            elif m not in on and m not in held:
                self.batch.pause(m, True)
                held.add(m)
        return active

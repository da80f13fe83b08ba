"""Benchmark of the GPU time step on the 4096x2048 FP64 workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps C]
    python bench.py --copy-calib        (a device copy of known size, for
                                         scaling the profiler's byte counters)

A step is one pass of SW2's per-step update (core.c:134-338) over the
4096x2048 channel, lat -85..85, Dt 10 s, del-4 5.97e10 m^4/s, minmod, jet
as background wind plus one storm bump; the fields stay resident in HBM.

Prints one JSON line:
  value         whole-job cell-updates/s over the K timed steps (wall clock
                between synchronisations);
  roofline      the dominant kernel, the momentum kernel of the fused step:
                88 algorithmic bytes per cell (reads eta, u, v and four
                Adams-Bashforth histories, writes u, v and two tendencies)
                over its average launch time from HIP events around each
                launch on the session's stream; the whole step at 120
                algorithmic bytes per cell-update beside it; `traffic` is the
                measured HBM bytes per launch from the profiler summary in
                profiles/ (byte counters scaled by a device copy), if present;
  cpu_baseline  SW2 itself (its sources built unchanged into oracle/_ref)
                on a bounded sample of the same workload under mpiexec, one
                rank per host CPU (at most 16, npx = ranks, npy = 1), timed
                by SW2's own step timer; beside it ("port") the oracle, the
                CPU restatement of SW2's path, with its row sweeps on the same
                number of threads and on one thread;
  mixed         the same grid in the mixed-precision form (anomalies about an
                FP64 mean held as FP32 value + FP32 residual, FP32 fluxes,
                84 algorithmic bytes per cell-update) for
                5000 steps, its speed-up over the FP64 form and its relative
                L2 / max errors against the FP64 run of the same steps;
  ensemble      64 members of uneven sizes (256x128 .. 2048x1024, 1-5 seeded
                storms each) stepped 3000 steps together in the batched mixed
                form (graph-replayed step, autotuned launch shape): runs/hour,
                against every member run alone in the FP64 form (measured),
                and every member's relative L2 / max error of eta, u, v
                against its FP64 run;
  served        Poisson arrivals at 4 requests/s of storm runs (512x256 ..
                4096x2048, 500..20000 steps) through the continuous-batching
                loop: p50 / p99 time-to-completion and runs/hour, against
                the FP64 form serving the same stream one run at a time
                (first come, first served; every request run in the FP64
                form, measured), and every result's relative L2 / max error
                against its FP64 run.  The offered load exceeds what one GPU
                sustains, so p50 / p99 grow with the number of requests.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
STEP_BYTES = 120       # per cell-update: eta,u,v read+written (48) + 6 histories read (48) + 3 tendencies written (24)
UV_BYTES = 88          # momentum kernel per cell: eta,u,v + 4 histories read (56), u,v + 2 tendencies written (32)
ETA_BYTES = 56         # height kernel per cell: eta,u,v + 2 histories read (40), eta + 1 tendency written (16)
MIXED_BYTES = 84       # mixed form per cell-update: eta,u,v as FP32 value + residual read and written (48),
                       # 6 FP32 histories read (24), 3 FP32 tendencies written (12)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_4096x2048_fp64.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=10, help="steps of the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="MPI ranks of SW2 and threads of the oracle in the CPU baseline (0: the host's CPUs, at most 16)")
    ap.add_argument("--ref-steps", type=int, default=8, help="steps of SW2 itself in the CPU baseline (0: skip)")
    ap.add_argument("--kernel-steps", type=int, default=100, help="steps timed per kernel after the timed region")
    ap.add_argument("--workload", default="c1", choices=["c0", "c1"])
    ap.add_argument("--copy-calib", action="store_true", help="only copy 3 x 512 MiB device to device and exit")
    ap.add_argument("--mixed-steps", type=int, default=5000, help="steps of the mixed-form run (0: skip)")
    ap.add_argument("--ensemble-members", type=int, default=64, help="members of the batched ensemble (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=3000)
    ap.add_argument("--no-tune", action="store_true", help="the batch's default launch shape, no autotuning")
    ap.add_argument("--served-requests", type=int, default=40, help="requests of the served stream (0: skip)")
    ap.add_argument("--storm-steps", type=int, default=500,
                    help="steps of the 4096x2048 run with a moving storm, fused and phased (0: skip)")
    return ap.parse_args()


def copy_calib():
    """Three device-to-device copies of 512 MiB through the HIP runtime: the
    reference transfer the profiler's FETCH_SIZE / WRITE_SIZE are scaled by."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = 512 * 1024 * 1024
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, ctypes.c_size_t(n)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    if hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calib_bytes": 3 * n}), flush=True)


def cpu_baseline_port(p_fn, bumps, steps, threads):
    """The oracle on the same workload: `steps` steps with its row sweeps on
    `threads` threads (bit-equal to its serial run), and 2 steps on one
    thread beside it."""
    from oracle import Oracle
    from sw2_amd import cases
    p = p_fn(steps)
    w, (eta, u, v) = cases.make_world(p, bumps)
    cells = p.nx * p.ny
    out = {}
    for th, n in ((threads, steps), (1, min(2, steps))):
        o = Oracle(p)
        o.load(eta, u, v)
        o.set_threads(th)
        t0 = time.perf_counter()
        o.step(n)
        out[th] = (n, time.perf_counter() - t0)
        o.close()
    w.close()
    n, dt = out[threads]
    n1, dt1 = out[1]
    return {"value": cells * n / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{p.nx}x{p.ny}, {n} steps of the same workload, {threads} threads, {dt:.1f} s "
                      f"(host: {os.cpu_count()} CPUs)",
            "single_thread": {"value": cells * n1 / dt1, "cores": 1, "sample": f"{n1} steps, {dt1:.1f} s"}}


MPIEXEC = os.environ.get("SW2_MPIEXEC", "/opt/conda/bin/mpiexec")


def cpu_baseline_reference(p_fn, bumps, steps, ranks):
    """SW2 itself (oracle/_ref/sw2ref: SW2's sources built unchanged, its
    loop() driven as main/Main.c does) on the same workload: `steps` steps
    on `ranks` MPI ranks of this host, npx = ranks, npy = 1 (longitude
    blocks of nx/ranks columns).  The rate is the grid's cells over the mean
    wall time per step of SW2's own step timer (time_step00 of cr_info).
    None when the binary or mpiexec is missing."""
    import subprocess
    import tempfile
    from oracle import ref
    from sw2_amd import cases
    if not (os.path.exists(ref.RUN_PATH) and os.path.exists(MPIEXEC)):
        return None
    p = p_fn(steps)
    w, state = cases.make_world(p, bumps)
    w.close()
    with tempfile.TemporaryDirectory(prefix="sw2ref_") as d:
        ref.write_case(d, "bench", p, state)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        cmd = [MPIEXEC, "-n", str(ranks), ref.RUN_PATH, str(ranks), "1", "bench.sw", "bench", d]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, cwd=d, env=env, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            progress(f"SW2 reference run failed ({r.returncode}): {r.stdout[-400:]} {r.stderr[-400:]}")
            return None
    step_s = None
    for line in r.stdout.splitlines():
        f = line.split()
        if len(f) >= 10 and f[0] == "cr_info:" and f[1] == "time_step00":
            step_s = float(f[9])  # tavg
    if step_s is None:
        return None
    cells = p.nx * p.ny
    return {"value": cells / step_s, "unit": "cell-updates/s", "cores": ranks, "kind": "reference",
            "ranks": ranks, "npx": ranks, "npy": 1,
            "sample": f"SW2 (its sources built unchanged) under mpiexec, {p.nx}x{p.ny}, {steps} steps, "
                      f"{ranks} ranks (npx {ranks}, npy 1), mean {step_s:.3f} s per step from SW2's own step "
                      f"timer, {wall:.1f} s with set-up (host: {os.cpu_count()} CPUs)"}


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel_errors(a, b):
    import numpy as np
    d = a - b
    nb = float(np.linalg.norm(b))
    mb = float(np.max(np.abs(b)))
    return {"l2": float(np.linalg.norm(d)) / nb if nb else float(np.linalg.norm(d)),
            "max": float(np.max(np.abs(d))) / mb if mb else float(np.max(np.abs(d)))}


def bench_mixed(steps):
    """The 4096x2048 workload in both forms for `steps` steps."""
    from sw2_amd import FORM_FP64, FORM_MIXED, Session, cases
    p = cases.config_params("c2", steps=steps)
    bumps = cases.single_bump(2024)
    res, worlds = {}, {}
    for form in (FORM_FP64, FORM_MIXED):
        progress(f"mixed-form comparison: form {form}, {steps} steps")
        w, _ = cases.make_world(p, bumps)
        s = Session(w, form)
        s.begin(0, p.t0)
        s.step(2)
        s.timed_steps(0)
        t0 = time.perf_counter()
        dev_ms = s.timed_steps(steps - 2)
        wall = time.perf_counter() - t0
        res[form] = {"wall": wall, "dev_ms": dev_ms, "failed": s.failed_step()[0], "energy": s.energy()}
        s.download()
        s.close()
        worlds[form] = w
    cells = p.nx * p.ny
    r64, r32 = res[FORM_FP64], res[FORM_MIXED]
    gbs = MIXED_BYTES * cells / (r32["dev_ms"] / (steps - 2) / 1e3) / 1e9
    out = {"workload": f"{p.nx}x{p.ny} jet + storm (the FP64 workload's case), {steps} steps",
           "value": cells * (steps - 2) / r32["wall"], "unit": "cell-updates/s", "dtype": "f32 anomalies / f64 mean",
           "fp64_value": cells * (steps - 2) / r64["wall"], "speedup_vs_fp64": r64["wall"] / r32["wall"],
           "bytes_per_cell": MIXED_BYTES, "achieved_gbs": gbs, "roofline_frac": gbs / HBM_PEAK_GBS,
           "failed_step": r32["failed"], "fp64_failed_step": r64["failed"],
           "errors_vs_fp64": {f: rel_errors(worlds[FORM_MIXED].interior(f), worlds[FORM_FP64].interior(f)) for f in ("eta", "u", "v")},
           "mass_rel_diff": abs(r32["energy"][2] - r64["energy"][2]) / abs(r64["energy"][2]),
           "energy_rel_diff": abs((r32["energy"][0] + r32["energy"][1]) - (r64["energy"][0] + r64["energy"][1])) /
                              abs(r64["energy"][0] + r64["energy"][1])}
    for w in worlds.values():
        w.close()
    return out


def bench_storm(steps):
    """The FP64 workload's case with a moving storm that injects tracer
    (perturbs.c's Gaussians): the fused step (injections after the momentum
    kernel, tracer advected every step), then a shorter phased run of the
    same case for comparison."""
    import numpy as np
    from sw2_amd import Session, cases
    G = np.array([[0.0, 1e9, 100.0, 30.0, 15.0, 0, 20.0, 3e11, 1, 3.0, 6.0, 1.0]])
    res = {}
    for label, form, n in (("fused", 0, steps), ("phased", 0x100, max(20, steps // 5))):
        progress(f"storm case: {label} step, {n} steps")
        p = cases.config_params("c1", steps=n + 10, Gaussians=G)
        w, _ = cases.make_world(p, cases.single_bump(2024))
        s = Session(w, form)
        s.begin(0, p.t0)
        s.step(5)
        s.timed_steps(0)
        t0 = time.perf_counter()
        dev_ms = s.timed_steps(n)
        wall = time.perf_counter() - t0
        kt = s.kernel_times(20) if s.fused else None
        res[label] = {"ms_per_step": wall * 1e3 / n, "device_ms_per_step": dev_ms / n, "fused": s.fused,
                      "failed_step": s.failed_step()[0],
                      "kernel_us": {"kf_eta": kt[0] / 20 * 1e3, "kf_uv": kt[1] / 20 * 1e3} if kt else None}
        s.close()
        w.close()
        cells = p.nx * p.ny
    f, ph = res["fused"], res["phased"]
    return {"workload": f"{p.nx}x{p.ny} FP64 jet + moving 6-degree storm injecting tracer, {steps} steps",
            "value": cells / (f["ms_per_step"] / 1e3), "unit": "cell-updates/s", "dtype": "f64",
            "ms_per_step": f["ms_per_step"], "kernel_us": f["kernel_us"], "failed_step": f["failed_step"],
            "phased_ms_per_step": ph["ms_per_step"], "speedup_vs_phased": ph["ms_per_step"] / f["ms_per_step"]}


def fp64_step_ms(nx, ny, steps=40):
    """Device ms per step of the FP64 form on an nx x ny member case."""
    import numpy as np
    from sw2_amd import FORM_FP64, Session, cases
    p = cases.member_params(nx, ny, steps + 10)
    w, _ = cases.make_world(p, cases.storm_bumps(np.random.default_rng(nx)))
    s = Session(w, FORM_FP64)
    s.begin(0, p.t0)
    s.step(10)
    ms = s.timed_steps(steps) / steps
    s.close()
    w.close()
    return ms


def fp64_run(p, bumps, steps):
    """One run in the FP64 form, which equals the oracle bit for bit: its
    wall time (world upload to the result's download into the world's host
    arrays, the world built beforehand, as a served or batched result is
    timed), failed step and final eta, u, v interiors."""
    from sw2_amd import FORM_FP64, Session, cases
    w, _ = cases.make_world(p, bumps)
    t0 = time.perf_counter()
    s = Session(w, FORM_FP64)
    s.begin(0, p.t0)
    s.timed_steps(steps)  # returns once the last step has completed
    failed = s.failed_step()[0]
    s.download()
    wall = time.perf_counter() - t0
    s.close()
    out = {f: w.interior(f).copy() for f in FIELDS}
    w.close()
    return wall, failed, out


FIELDS = ("eta", "u", "v")


def error_summary(errs):
    """errs: [(label, {field: {"l2", "max"}})] -> the worst L2 and max per
    field over all results, and every result's errors."""
    worst = {f: {"l2": max(e[f]["l2"] for _, e in errs), "max": max(e[f]["max"] for _, e in errs)} for f in FIELDS}
    return {"results": len(errs), "worst": worst,
            "per_result": [[lab] + [round(e[f][k], 9) for f in FIELDS for k in ("l2", "max")] for lab, e in errs],
            "per_result_columns": ["result"] + [f"{f}_{k}" for f in FIELDS for k in ("l2", "max")]}


def bench_ensemble(n, steps, tune_launch=True):
    """The batched ensemble, then every member run alone in the FP64 form
    (measured): runs/hour of both and every member's relative L2 / max
    error of eta, u, v against its FP64 run."""
    from sw2_amd import Batch, cases, tune
    specs = cases.ensemble(n, steps)
    sizes = [(p.nx, p.ny) for p, _ in specs]
    if tune_launch:
        progress(f"ensemble: tuning the launch shape for {n} members")
        shape, rec = tune.tune(sizes)
    else:
        shape, rec = None, None
    progress(f"ensemble: launch {shape}; admitting {n} members")
    b = Batch(capacity=n)
    if shape:
        tune.apply(b, shape)
    shape = tuple(b.launch()) + tuple(b.tiles())
    progress(f"ensemble: building {n} worlds")
    worlds = [cases.make_world(p, bumps)[0] for p, bumps in specs]
    # timed as the FP64 runs are: from the members' uploads to their results
    # downloaded into the worlds' host arrays (the worlds built beforehand)
    progress(f"ensemble: admitting {n} members, {steps} steps, results downloaded")
    t0 = time.perf_counter()
    ids = [b.add(w, steps) for w in worlds]
    t_admit = time.perf_counter() - t0
    b.step(steps)  # the graph's capture included
    b.sync()
    t_steps = time.perf_counter() - t0 - t_admit
    fin = b.finished()
    failed = sum(1 for m in fin if b.member(m)["failed_step"])
    got = [b.download(mid) for mid in ids]
    wall = time.perf_counter() - t0
    got = [{f: w.interior(f).copy() for f in FIELDS} for w in got]
    b.close()
    for w in worlds:
        w.close()
    cells_steps = sum(float(nx) * ny for nx, ny in sizes) * steps
    progress(f"ensemble: every member alone in the FP64 form, {steps} steps")
    t64, errs, f64_failed = 0.0, [], 0
    for (p, bumps), res in zip(specs, got):
        w64, fl, ref = fp64_run(p, bumps, steps)
        t64 += w64
        f64_failed += fl != 0
        errs.append((f"{p.nx}x{p.ny}", {f: rel_errors(res[f], ref[f]) for f in FIELDS}))
    return {"members": n, "steps": steps, "sizes": {f"{a}x{b_}": sizes.count((a, b_)) for a, b_ in sorted(set(sizes))},
            "runs_per_hour": n / wall * 3600.0, "cell_updates_per_s": cells_steps / t_steps, "wall_s": wall,
            "steps_s": t_steps, "download_s": wall - t_admit - t_steps,
            "completed": len(fin), "failed": failed,
            "launch": {"blocks_eta": shape[0], "blocks_uv": shape[1], "steps_per_graph": shape[2],
                       "tiles": [shape[3], shape[4]], "tuning": rec},
            "admit_s": t_admit, "fp64_wall_s": t64, "fp64_failed": f64_failed,
            "fp64_runs_per_hour": n / t64 * 3600.0, "vs_fp64_runs_per_hour": (n / wall) / (n / t64),
            "fp64_note": "every member run alone in the FP64 form, one after another (measured wall time, upload to "
                         "download, as the batch's wall_s)",
            "errors_vs_fp64": error_summary(errs)}


def bench_served(n):
    """The served stream, then the FP64 form serving the same stream one run
    at a time, first come first served: every request run in the FP64 form
    (measured wall time each), the queue replayed on the arrival times.
    Every result's relative L2 / max error against its FP64 run."""
    from sw2_amd import cases
    from sw2_amd.serve import Server
    reqs = cases.served_requests(n)
    got = {}

    def keep(k, batch, mid):
        got[k] = batch.download(mid)  # the result in its world's host arrays (kept; read after the run)

    # the launch and tile shape tuned on a batch of the stream's size mix
    from sw2_amd import tune
    mix = [(512, 256)] * 4 + [(1024, 512)] * 3 + [(2048, 1024)] * 2 + [(4096, 2048)]
    shape, rec = tune.tune(mix)
    progress(f"served stream: {n} requests, launch {shape}")
    srv = Server(capacity=48, hbm_budget=64e9, tick_steps=40, launch=shape)
    st = srv.run(reqs, on_result=keep, log=progress)
    srv.close()
    cells = [float(nx) * ny for _, nx, ny, *_ in reqs]
    offered = sum(c * r[3] for c, r in zip(cells, reqs)) / max(r[0] for r in reqs)
    st.update({"arrivals_per_s": 4.0, "capacity": 48, "hbm_budget_bytes": 64e9, "tick_steps": 40,
               "sizes": "512x256/1024x512/2048x1024/4096x2048 at 0.4/0.3/0.2/0.1", "steps": "uniform 500..20000",
               "offered_cell_updates_per_s": offered, "launch": list(shape), "tuning": rec})
    progress("served stream: every request alone in the FP64 form")
    t, errs, durations = 0.0, [], {}
    for k, (ta, nx, ny, steps, bumps) in enumerate(reqs):
        w64, fl, ref = fp64_run(cases.member_params(nx, ny, steps), bumps, steps)
        durations[k] = w64
        if k in got:
            errs.append((f"{k}:{nx}x{ny}:{steps}", {f: rel_errors(got[k].interior(f), ref[f]) for f in FIELDS}))
    for k, r in sorted(enumerate(reqs), key=lambda kr: kr[1][0]):
        t = max(t, r[0]) + durations[k]
    st.update({"fp64_fcfs_wall_s": t, "fp64_runs_per_hour": len(reqs) / t * 3600.0,
               "vs_fp64_runs_per_hour": st["runs_per_hour"] / (len(reqs) / t * 3600.0),
               "fp64_note": "every request run alone in the FP64 form (measured wall time), first come first "
                            "served on the same arrival times",
               "errors_vs_fp64": error_summary(errs)})
    return st


def measured_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus})", file=sys.stderr)
        sys.exit(2)
    if a.copy_calib:
        copy_calib()
        return
    from sw2_amd import Session, cases

    wl = a.workload
    total = a.warmup + a.steps + a.kernel_steps
    p_fn = lambda n: cases.config_params(wl, steps=n)
    bumps = cases.single_bump(2024)
    # the CPU baseline first, before this process touches the GPU (SW2 runs
    # in child processes under mpiexec)
    threads = a.cpu_threads if a.cpu_threads > 0 else max(1, min(16, len(os.sched_getaffinity(0))))
    cpu = None
    if a.ref_steps > 0:
        progress(f"CPU baseline: SW2 on {threads} MPI ranks, {a.ref_steps} steps")
        cpu = cpu_baseline_reference(p_fn, bumps, a.ref_steps, threads)
    if a.cpu_steps > 0:
        progress(f"CPU baseline: {a.cpu_steps} oracle steps on {threads} threads")
        port = cpu_baseline_port(p_fn, bumps, a.cpu_steps, threads)
        if cpu is None:
            cpu = port
        else:
            cpu["port"] = port
            cpu["reference_over_port"] = cpu["value"] / port["value"]
    progress("FP64 workload")
    p = p_fn(total)
    w, _ = cases.make_world(p, bumps)
    s = Session(w)
    s.begin(0, p.t0)
    if a.warmup:
        s.step(a.warmup)
        s.timed_steps(0)  # drains the stream
    t0 = time.perf_counter()
    dev_ms = s.timed_steps(a.steps)  # returns once the last step has completed
    wall = time.perf_counter() - t0
    kt = s.kernel_times(a.kernel_steps) if (a.kernel_steps > 0 and s.fused) else None
    failed, msg = s.failed_step()
    if failed:
        print(f"step {failed} failed SW2's guard: {msg}", file=sys.stderr)
        sys.exit(1)
    cells = p.nx * p.ny
    value = cells * a.steps / wall
    step_ms = dev_ms / a.steps
    step_gbs = STEP_BYTES * cells / (step_ms / 1e3) / 1e9
    fused = s.fused
    s.close()
    w.close()
    names = {"c1": "4096x2048 FP64 jet + storm, lat -85..85, Dt 10 s, del-4 5.97e10, minmod",
             "c0": "360x180 FP64 jet + storm, lat -80..80, Dt 60 s, del-4 1e15, minmod"}
    if kt:
        eta_ms, uv_ms = kt[0] / a.kernel_steps, kt[1] / a.kernel_steps
        uv_gbs = UV_BYTES * cells / (uv_ms / 1e3) / 1e9
        tr = measured_traffic() if wl == "c1" else None
        roof = {"bound": "hbm", "achieved": uv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": uv_gbs / HBM_PEAK_GBS,
                "traffic": tr["kf_uv"]["bytes"] if tr else None,
                "kernel": "kf_uv (momentum kernel of the fused step)", "bytes_per_cell": UV_BYTES,
                "kernel_ms": uv_ms, "eta_kernel_ms": eta_ms,
                "eta_kernel_gbs": ETA_BYTES * cells / (eta_ms / 1e3) / 1e9,
                "step": {"bytes_per_cell": STEP_BYTES, "device_ms": step_ms, "achieved": step_gbs,
                         "frac": step_gbs / HBM_PEAK_GBS}}
        if tr:
            roof["traffic_source"] = tr.get("source")
    else:
        roof = {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": step_gbs / HBM_PEAK_GBS, "traffic": None,
                "kernel": "whole step (one kernel per phase)", "bytes_per_cell": STEP_BYTES, "device_ms": step_ms}
    extras = {}
    if a.mixed_steps > 0 and wl == "c1":
        extras["mixed"] = bench_mixed(a.mixed_steps)
    if a.ensemble_members > 0:
        extras["ensemble"] = bench_ensemble(a.ensemble_members, a.ensemble_steps, tune_launch=not a.no_tune)
    if a.served_requests > 0:
        extras["served"] = bench_served(a.served_requests)
    if a.storm_steps > 0 and wl == "c1":
        extras["storm"] = bench_storm(a.storm_steps)
    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (balanced background jet + seeded storm bump)",
        "config": {"workload": names[wl], "grid": [p.nx, p.ny], "cells": cells, "path": "fused" if fused else "phased"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

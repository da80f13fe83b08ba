"""Benchmark of the GPU time step on the 4096x2048 FP64 workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps C]

A step is one pass of SW2's per-step update (core.c:134-338) over the
4096x2048 channel, lat -85..85, Dt 10 s, del-4 5.97e10 m^4/s, minmod, jet
as background wind plus one storm bump; the fields stay resident in HBM.
Prints one JSON line: whole-job cell-updates/s, the HBM roofline of the
step (120 algorithmic bytes per cell-update) from HIP events on the
session's stream, and the oracle (the CPU restatement of SW2's path) timed
on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
BYTES_PER_CELL = {"f64": 120}  # eta,u,v read+write (48) + 6 AB histories read (48) + 3 tendencies written (24)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps of the CPU baseline sample (0: skip)")
    ap.add_argument("--workload", default="c1", choices=["c0", "c1"])
    return ap.parse_args()


def cpu_baseline(p_fn, bumps, steps):
    """The oracle on the same workload, `steps` steps, one core."""
    from oracle import Oracle
    from sw2_amd import cases
    p = p_fn(steps)
    w, (eta, u, v) = cases.make_world(p, bumps)
    o = Oracle(p)
    o.load(eta, u, v)
    w.close()
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    cells = p.nx * p.ny
    return {"value": cells * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{p.nx}x{p.ny}, {steps} steps of the same workload, single thread, {dt:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus})", file=sys.stderr)
        sys.exit(2)
    from sw2_amd import Session, cases

    wl = a.workload
    total = a.warmup + a.steps
    p_fn = lambda n: cases.config_params(wl, steps=n)
    bumps = cases.single_bump(2024)
    p = p_fn(total)
    w, _ = cases.make_world(p, bumps)
    s = Session(w)
    s.begin(0, p.t0)
    if a.warmup:
        s.step(a.warmup)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None
    sync()
    t0 = time.perf_counter()
    dev_ms = s.timed_steps(a.steps)
    sync()
    wall = time.perf_counter() - t0
    failed, msg = s.failed_step()
    if failed:
        print(f"step {failed} failed SW2's guard: {msg}", file=sys.stderr)
        sys.exit(1)
    cells = p.nx * p.ny
    value = cells * a.steps / wall
    step_s = dev_ms / 1e3 / a.steps
    achieved = BYTES_PER_CELL["f64"] * cells / step_s / 1e9
    s.close()
    w.close()
    cpu = cpu_baseline(p_fn, bumps, a.cpu_steps) if a.cpu_steps > 0 else None
    names = {"c1": "4096x2048 FP64 jet + storm, lat -85..85, Dt 10 s, del-4 5.97e10, minmod",
             "c0": "360x180 FP64 jet + storm, lat -80..80, Dt 60 s, del-4 1e15, minmod"}
    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (balanced background jet + seeded storm bump)",
        "config": {"workload": names[wl], "grid": [p.nx, p.ny], "cells": cells},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole step (all kernels of one step, HIP events on the session stream)",
                     "device_ms_per_step": dev_ms / a.steps},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

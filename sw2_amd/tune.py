"""Autotuner of the batched step's launch shape and tile shape.

For a mix of member grid sizes it builds a scratch batch of the same
composition (members at rest: the mixed-form kernels take the same path
whatever the data), times whole steps for candidate shapes and keeps the
fastest.  The knobs are the tile shape of the batch kernels (32 x 16,
64 x 8 or 64 x 16 cells), the persistent workgroups of the height kernel
and of the momentum kernel (multiples of the device's compute-unit count;
32 x 16 tiles only) and the steps per captured graph.  The search is
coordinate descent from the batch's default (one workgroup per 32 x 16
tile, 8 steps per graph): tile shapes first, then the workgroup counts,
then the steps per graph.  A shape is (blocks_eta, blocks_uv,
steps_per_graph, tile_x, tile_y).  Results are cached per device and mix in
a JSON file (default: tuned_launch.json beside this module, or
$SW2_TUNE_CACHE).
"""
import json
import os
import time
from collections import Counter

from . import Batch, cases
from ._ffi import lib

CACHE = os.environ.get("SW2_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_launch.json"))


def device_cus():
    """Compute units of the device."""
    return lib().sw2_device_cus()


def mix_key(sizes, ncu):
    c = Counter(sizes)
    return f"cu{ncu}|mixed|" + ",".join(f"{nx}x{ny}:{k}" for (nx, ny), k in sorted(c.items()))


TILE_SHAPES = [(32, 16), (64, 8), (64, 16)]  # the first is a new batch's


def apply(b, shape):
    """Set a batch to a shape (blocks_eta, blocks_uv, steps_per_graph,
    tile_x, tile_y)."""
    b.set_launch(0, 0, shape[2])  # persistent workgroups exist for the default tiles only
    b.set_tiles(shape[3], shape[4])
    b.set_launch(*shape[:3])


def _time_steps(b, shape, steps, repeats=3):
    """Seconds per step of a shape: the fastest of `repeats` timings of
    `steps` steps (a single short timing picked shapes by noise)."""
    apply(b, shape)
    b.step(shape[2])  # capture and warm up
    b.sync()
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        best = min(best, (time.perf_counter() - t0) / steps)
    return best


def tune(sizes, steps=32, cache=CACHE, verbose=False):
    """Best (blocks_eta, blocks_uv, steps_per_graph, tile_x, tile_y) for
    members of `sizes` ([(nx, ny), ...]); returns (shape, record)."""
    ncu = device_cus()
    key = mix_key(sizes, ncu)
    db = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                db = json.load(f)
        except (OSError, ValueError):
            db = {}
    if key in db and len(db[key]["shape"]) == 5:
        r = db[key]
        return tuple(r["shape"]), r
    b = Batch(capacity=len(sizes))
    worlds = []
    for nx, ny in sizes:
        w, _ = cases.make_world(cases.member_params(nx, ny, 1 << 30), [])
        worlds.append(w)
        b.add(w, 1 << 30)
    t_start = time.perf_counter()
    default = tuple(b.launch()) + tuple(b.tiles())  # one workgroup per 64 x 8 tile, 8 steps per graph
    best = default
    best_t = _time_steps(b, best, steps)
    tried = {best: best_t}
    axes = [(3, TILE_SHAPES), (0, [1, 2, 3, 4, 6]), (1, [1, 2, 3, 4, 6]), (2, [4, 8, 16])]
    for axis, options in axes:
        for o in options:
            shape = list(best)
            if axis == 3:
                shape[3:5] = o
            else:
                shape[axis] = o * ncu if axis < 2 else o
            shape = tuple(shape)
            if (shape[0] or shape[1]) and shape[3:5] != TILE_SHAPES[0]:
                continue  # persistent workgroups run the default tiles only
            if shape in tried:
                continue
            t = _time_steps(b, shape, steps)
            tried[shape] = t
            if verbose:
                print("tune", shape, f"{t * 1e3:.3f} ms/step", flush=True)
            if t < best_t:
                best, best_t = shape, t
    b.close()
    for w in worlds:
        w.close()
    rec = {"shape": list(best), "ms_per_step": best_t * 1e3, "candidates": len(tried),
           "tuning_s": time.perf_counter() - t_start,
           "default": list(default), "default_ms_per_step": tried[default] * 1e3}
    db[key] = rec
    if cache:
        try:
            with open(cache, "w") as f:
                json.dump(db, f, indent=1)
        except OSError:
            pass
    return best, rec

"""The sparse storm injection (sw2_gaussians_sparse: only the cells inside
each truncation radius are evaluated) against the full-grid one
(sw2_gaussians_at, which mirrors perturbs.c:22-395 cell by cell): the same
cells, the same increments and tracer values bit for bit, for storms on the
periodic seam, overlapping copies, near the pole, degree and metre sizes, a
zero-volume moment and a sink.  CPU only."""
import ctypes

import numpy as np
import pytest

from sw2_amd import World, cases
from sw2_amd._ffi import SW, lib

P = ctypes.POINTER(ctypes.c_double)


class Inj(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("cap", ctypes.c_int), ("has_inc", ctypes.c_int),
                ("idx", ctypes.POINTER(ctypes.c_int)), ("inc", P), ("trv", P)]


def both(G, t, nx=256, ny=128, lat=80.0):
    p = cases.jet_params(nx, ny, lat, 60.0 * 360.0 / nx, 100, 0.0, Gaussians=np.asarray(G, dtype=float))
    w = World(p)
    L = lib()
    L.sw2_gaussians_at.argtypes = [ctypes.POINTER(SW), ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.sw2_gaussians_sparse.argtypes = [ctypes.POINTER(SW), ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    L.sw2_inj_free.argtypes = [ctypes.c_void_p, ctypes.c_int]
    S = w.S
    dense = [(np.zeros(S), np.zeros(S), np.zeros(S, dtype=np.uint8)) for _ in range(8)]
    inc = (P * 8)(*[d[0].ctypes.data_as(P) for d in dense])
    trv = (P * 8)(*[d[1].ctypes.data_as(P) for d in dense])
    msk = (ctypes.POINTER(ctypes.c_ubyte) * 8)(*[d[2].ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)) for d in dense])
    K = L.sw2_gaussians_at(ctypes.byref(w.sw), p.Dt, t, 8, inc, trv, msk)
    inj = (Inj * 8)()
    K2 = L.sw2_gaussians_sparse(ctypes.byref(w.sw), p.Dt, t, 8, inj)
    out = []
    for k in range(max(K, 0)):
        n = inj[k].n
        idx = np.ctypeslib.as_array(inj[k].idx, shape=(n,)).copy() if n else np.zeros(0, dtype=np.int32)
        sv = np.ctypeslib.as_array(inj[k].inc, shape=(n,)).copy() if n else np.zeros(0)
        st = np.ctypeslib.as_array(inj[k].trv, shape=(n,)).copy() if n else np.zeros(0)
        out.append((dense[k], (idx, sv, st, inj[k].has_inc)))
    L.sw2_inj_free(inj, 8)
    w.close()
    return K, K2, out


CASES = {
    # t0, t1, tau, lon, lat, isdegvel, vel, Q, isdegsize, sigma, radius, inject
    "seam": [[0.0, 5000.0, 100.0, 359.0, 10.0, 1, 0.5, 2e11, 1, 4.0, 12.0, 1.0]],
    "two overlapping": [[0.0, 5000.0, 100.0, 359.0, 10.0, 1, 0.5, 2e11, 1, 4.0, 12.0, 1.0],
                        [0.0, 5000.0, 150.0, 5.0, 14.0, 0, -30.0, 1e11, 1, 3.0, 9.0, 0.5]],
    "high latitude": [[0.0, 5000.0, 100.0, 120.0, 74.0, 1, 2.0, 1e11, 1, 3.0, 15.0, 1.0]],
    "metres": [[0.0, 5000.0, 100.0, 200.0, -30.0, 0, 50.0, 5e10, 0, 2e6, 5e6, 0.3]],
    "sink": [[0.0, 5000.0, 100.0, 90.0, 0.0, 1, 0.0, -1e11, 1, 2.0, 6.0, 0.0]],
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("t", [60.0, 1234.5])
def test_sparse_injection_equals_full_grid(name, t):
    K, K2, out = both(CASES[name], t)
    assert K == K2 == len(CASES[name])
    for (inc, trv, mask), (idx, sv, st, has_inc) in out:
        cells = np.flatnonzero(mask)
        assert len(cells) > 0
        assert np.array_equal(cells, idx)
        assert np.array_equal(trv[idx], st)
        if has_inc:
            assert np.array_equal(inc[idx], sv)
            assert not np.any(np.delete(inc, idx))  # nothing outside the radius
        else:
            assert np.isnan(inc.ravel()[0])


def test_sparse_injection_split_over_threads():
    """A band of more than 32 rows is split over the host's OpenMP threads;
    the joined lists are the full-grid cells in SW2's order all the same."""
    G = [[0.0, 5000.0, 100.0, 359.5, 20.0, 1, 0.5, 2e11, 1, 5.0, 14.0, 1.0]]
    K, K2, out = both(G, 600.0, nx=1024, ny=512)
    assert K == K2 == 1
    (inc, trv, mask), (idx, sv, st, has_inc) = out[0]
    assert len(idx) > 3000
    assert np.array_equal(np.flatnonzero(mask), idx)
    assert np.array_equal(inc[idx], sv) and np.array_equal(trv[idx], st)


def test_zero_volume_moment_leaves_eta_alone():
    # Q(t) at t0 is zero: SW2 marks the tracer but adds nothing to eta
    K, K2, out = both([[100.0, 5000.0, 100.0, 30.0, 10.0, 1, 0.0, 2e11, 1, 4.0, 8.0, 1.0]], 100.0)
    assert K == K2 == 1
    (inc, trv, mask), (idx, sv, st, has_inc) = out[0]
    assert has_inc == 0 and np.isnan(inc.ravel()[0])
    assert np.array_equal(np.flatnonzero(mask), idx) and np.array_equal(trv[idx], st)


@pytest.mark.parametrize("env", [{"OMP_THREAD_LIMIT": "1"}, {"OMP_DYNAMIC": "true", "OMP_THREAD_LIMIT": "2"}])
def test_sparse_injection_with_a_smaller_thread_team(env):
    """The band is cut into as many chunks as the host has threads (up to
    8); when the runtime grants a smaller team (a thread limit, dynamic
    teams, a caller's enclosing parallel region) every chunk is still
    evaluated.  Run in a fresh process, where the OpenMP settings apply."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path.insert(0, %r); "
            "from tests.test_injection import test_sparse_injection_split_over_threads as t; t(); print('chunks ok')") % root
    e = dict(os.environ, OMP_NUM_THREADS="8", **env)
    r = subprocess.run([sys.executable, "-c", code], env=e, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "chunks ok" in r.stdout, r.stderr[-2000:]

"""sw2_session_energy_noalloc (the allocation-free energy reduction) against
sw2_session_energy, through the C-ABI: the same three sums bit for bit, on
two calls in a row, the second on a smaller world (and in the other form),
with the results left in device memory the caller owns."""
import ctypes
from ctypes import byref, c_double, c_size_t, c_void_p

import numpy as np
import pytest

from sw2_amd import FORM_FP64, FORM_MIXED, Session, cases, lib

pytestmark = pytest.mark.gpu


def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
    h.hipFree.argtypes = [c_void_p]
    h.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [c_void_p, ctypes.c_int, c_size_t]
    return h


def energies(s, h, d_out):
    ke, ape, m = c_double(), c_double(), c_double()
    assert lib().sw2_session_energy(s.h, byref(ke), byref(ape), byref(m)) == 0
    assert h.hipMemset(d_out, 0xFF, 3 * 8) == 0  # NaN bytes: the call must write all three
    assert lib().sw2_session_energy_noalloc(s.h, d_out, None) == 0
    assert h.hipDeviceSynchronize() == 0
    out = np.zeros(3)
    assert h.hipMemcpy(out.ctypes.data, d_out, 3 * 8, 2) == 0  # hipMemcpyDeviceToHost
    return np.array([ke.value, ape.value, m.value]), out


def test_noalloc_energy_equals_session_energy():
    h = hip()
    d_out = c_void_p()
    assert h.hipMalloc(byref(d_out), c_size_t(3 * 8)) == 0
    try:
        for nx, ny, form, n in ((512, 256, FORM_FP64, 30), (256, 128, FORM_MIXED, 20), (160, 64, FORM_FP64, 10)):
            p = cases.scaled_params(nx, ny, steps=n, lat=70.0)
            w, _ = cases.make_world(p, cases.single_bump(nx))
            s = Session(w, form)
            s.begin(0, p.t0)
            s.step(n)
            ref, got = energies(s, h, d_out)
            print(f"{nx}x{ny} form {form}: session {ref.tolist()} noalloc {got.tolist()}", flush=True)
            assert np.all(np.isfinite(ref)) and np.array_equal(ref, got), (nx, ny, form, ref, got)
            s.close()
            w.close()
    finally:
        h.hipFree(d_out)


def test_noalloc_energy_rejects_missing_output():
    p = cases.scaled_params(160, 64, steps=2, lat=70.0)
    w, _ = cases.make_world(p, cases.single_bump(1))
    s = Session(w, FORM_FP64)
    assert lib().sw2_session_energy_noalloc(s.h, None, None) == -1
    s.close()
    w.close()

"""The GPU step's division (sw2_amd/csrc/sw2_xdiv.h), checked on the CPU
against the IEEE correctly rounded quotient: ddiv for any operands and xdiv
for divisors known ahead, fed a candidate quotient that is only faithful (the
worst a hardware sequence within one ulp can return), over the whole exponent
range and every metric divisor of the single-run and ensemble grids.  The
same header is compiled into the GPU kernels, whose results the GPU parity
tests compare with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from sw2_amd import World, cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cpu_has_fma():
    try:
        return " fma " in open("/proc/cpuinfo").read().replace("\n", " ")
    except OSError:
        return False


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("xdiv")
    exe = d / "xdiv_check"
    flags = ["-O2", "-ffp-contract=off"] + (["-mfma"] if cpu_has_fma() else [])
    subprocess.run(["gcc", *flags, "-o", str(exe), os.path.join(ROOT, "tests", "xdiv_check.c"), "-lm"], check=True)
    return exe


def grid_divisors():
    out = []
    for nx, ny, lat in [(360, 180, 80.0), (4096, 2048, 85.0), (256, 128, 80.0), (2048, 1024, 80.0)]:
        w = World(cases.scaled_params(nx, ny, steps=1, lat=lat))
        for name in ("Dxe", "Dxc", "Dxv", "Dyn", "Dyc", "Dyv"):
            out.append(w.rows(name))
        w.close()
    b = np.unique(np.concatenate(out))
    return b[b > 0]


def test_xdiv_matches_ieee_division(checker, tmp_path):
    b = grid_divisors()
    assert len(b) > 1000
    f = tmp_path / "div.bin"
    b.astype(np.float64).tofile(f)
    r = subprocess.run([str(checker), "7", "150000", str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    n = int(r.stdout.split()[1])
    assert n > 8_000_000
    assert "mismatches 0" in r.stdout

"""The library's per-lane search code (mpmc_hnu_amd/csrc/pr76_device.hpp, flash_device.hpp)
compiled as plain host C++ (tests/host_build, with a stand-in for the HIP attribute header):

* the device Rachford-Rice with its certified-sign fast path (Newton enclosure, dyadic
  jump over the certified bisection levels, certified interior test) is bit-identical to
  plain bisection in status, beta, x, y and iteration count;
* the device's serial max-3-phase flash (solve_max3) reproduces the oracle's statuses and
  evaluation counts cell for cell on the three-phase sour-gas sweep;
* the same serial path runs clean under MemorySanitizer (no read of an unwritten value in
  any branch or store), where clang's MSan runtime is available.

No GPU is involved: this checks the arithmetic and control flow of the device code."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_bindings as ob
from mpmc_hnu_amd import workloads as W

HERE = Path(__file__).resolve().parent / "host_build"
FLAGS = ["-std=c++20", "-O2", "-ffp-contract=off", "-Wno-unknown-pragmas", "-I", str(HERE)]
CLANG = Path("/opt/rocm/llvm/bin/clang++")


def _build(tmp_path, src, out, cxx="g++", extra=()):
    exe = tmp_path / out
    subprocess.run([cxx, *FLAGS, *extra, "-o", str(exe), str(HERE / src)], check=True, capture_output=True)
    return exe


def _write_cells(path, mix, p, t, z):
    with open(path, "wb") as f:
        f.write(np.int32(z.shape[0]).tobytes())
        f.write(np.int64(p.size).tobytes())
        for a in (mix["tc"], mix["pc"], mix["omega"], np.asarray(mix["kij"]).ravel(), p, t, np.ascontiguousarray(z).ravel()):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


def test_rachford_rice_fast_path_is_plain_bisection(tmp_path):
    exe = _build(tmp_path, "rr_check.cpp", "rr_check")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "diffs=0" in r.stdout and "interior=0" not in r.stdout


def _sour(n):
    mix, p, t, z = W.sour_gas_sweep(2048)
    idx = np.linspace(0, 2047, n).astype(int)
    return mix, p[idx], t[idx], np.ascontiguousarray(z[:, idx])


def _run_max3(exe, cells, n):
    r = subprocess.run([str(exe), str(cells), "0", str(n)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return np.loadtxt(r.stdout.splitlines(), dtype=np.int64, ndmin=2), r.stderr


def test_host_built_serial_max3_matches_the_oracle(tmp_path):
    mix, p, t, z = _sour(192)
    cells = tmp_path / "cells.bin"
    _write_cells(cells, mix, p, t, z)
    h, _ = _run_max3(_build(tmp_path, "max3_host.cpp", "max3_host"), cells, p.size)
    o = ob.flash(mix, p, t, z, 1e-11, 3)
    assert np.array_equal(h[:, 1], (o["status"] >> 4) & 0xF)
    assert np.array_equal(h[:, 2], (o["status"] >> 8) & 0xF)
    assert np.array_equal(h[:, 4], o["counters"][0])
    assert np.sum(h[:, 1] == 2) > 10  # three-phase cells present


@pytest.mark.skipif(not CLANG.exists(), reason="clang with the MemorySanitizer runtime is not installed")
def test_serial_max3_under_memory_sanitizer(tmp_path):
    mix, p, t, z = _sour(96)
    cells = tmp_path / "cells.bin"
    _write_cells(cells, mix, p, t, z)
    exe = _build(tmp_path, "max3_host.cpp", "max3_msan", cxx=str(CLANG),
                 extra=("-O1", "-g", "-fsanitize=memory", "-fsanitize-memory-track-origins", "-fno-omit-frame-pointer"))
    h, err = _run_max3(exe, cells, p.size)
    assert "MemorySanitizer" not in err, err[-4000:]
    plain, _ = _run_max3(_build(tmp_path, "max3_host.cpp", "max3_host"), cells, p.size)
    assert np.array_equal(h, plain)

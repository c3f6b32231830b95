"""The reference-side drop-in (include/mpmc_pr76_backend.hpp: GpuPr76PtFlashBackend, a
PtFlashBackend on this library's C ABI) against the reference's own Pr76PtFlashBackend,
through oracle/_ref/adapter_check (built by `make -C oracle adapter` from the reference's
headers; the check is skipped where that build is absent):

* CPU: results rebuilt from committed device flash records (tests/golden/
  device_flash_records.npz, written by tests/golden/make_device_records.py on an MI355X)
  and from the oracle's records must be structurally valid and equal to the reference
  backend's results on the same requests: phase-set status, candidate phases (fractions,
  compositions, Z within 1e-9, ln phi within 1e-9 absolute plus relative, branch labels),
  transition evidence entry by entry, capability and conventions (oracle/adapter_check.cpp).
* GPU: GpuPr76PtFlashBackend::solve_batch on the device against the reference backend run
  on the host, at the same bar.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_bindings as ob
from mpmc_hnu_amd import workloads as W

ROOT = Path(__file__).resolve().parent.parent
CHECK = ROOT / "oracle" / "_ref" / "adapter_check"
RECORDS = Path(__file__).resolve().parent / "golden" / "device_flash_records.npz"
KEYS = (("status", np.int32), ("phase_count", np.int32), ("beta", np.float64), ("zfactor", np.float64),
        ("branch", np.int32), ("x", np.float64), ("ln_phi", np.float64))

needs_check = pytest.mark.skipif(not CHECK.exists(), reason="oracle/_ref/adapter_check not built (needs the reference)")


def write_cells(path, mix, p, t, z, rec=None):
    with open(path, "wb") as f:
        f.write(np.int32(z.shape[0]).tobytes())
        f.write(np.int64(p.size).tobytes())
        for a in (mix["tc"], mix["pc"], mix["omega"], np.asarray(mix["kij"]).ravel(), p, t,
                  np.ascontiguousarray(z).ravel()):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        if rec is not None:
            for k, dt in KEYS:
                f.write(np.ascontiguousarray(rec[k], dtype=dt).ravel().tobytes())


def run_check(mode, path, timeout=900):
    r = subprocess.run([str(CHECK), mode, str(path)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:]
    return r.stdout


def _sour():
    mix, p, t, z = W.sour_gas_sweep(2048)
    idx = np.arange(0, 2048, 4)
    return mix, p[idx], t[idx], np.ascontiguousarray(z[:, idx])


def _boundary():
    mix = W.mixture_c1()
    p, t, z = W.cells_c1()
    p, t, z = W.boundary_feeds(p, t, ob.flash(mix, p, t, z, 1e-12, 3), 5e-11)
    return mix, p, t, z


@needs_check
@pytest.mark.skipif(not RECORDS.exists(), reason="device records not committed")
@pytest.mark.parametrize("case", ["sour", "boundary", "c2"])
def test_rebuild_from_device_records(tmp_path, case):
    d = np.load(RECORDS)
    mix = {k: d[f"{case}_{k}"] for k in ("tc", "pc", "omega", "kij")}
    rec = {k: d[f"{case}_{k}"] for k, _ in KEYS}
    path = tmp_path / "cells.bin"
    write_cells(path, mix, d[f"{case}_p"], d[f"{case}_t"], d[f"{case}_z"], rec)
    out = run_check("rebuild", path)
    assert "mismatches=0" in out


@needs_check
@pytest.mark.parametrize("case", ["sour", "boundary"])
def test_rebuild_from_oracle_records(tmp_path, case):
    """The mapping itself, on the oracle's records (bit-identical to the reference's flash):
    every transition route occurs here, the phase-disappearance evidence of the boundary feeds
    included."""
    mix, p, t, z = _sour() if case == "sour" else _boundary()
    o = ob.flash(mix, p, t, z, 1e-11, 3)
    path = tmp_path / "cells.bin"
    write_cells(path, mix, p, t, z, o)
    out = run_check("rebuild", path)
    assert "mismatches=0" in out


@pytest.mark.gpu
@needs_check
@pytest.mark.parametrize("case", ["sour", "boundary"])
def test_gpu_backend_solve_batch(gpu, tmp_path, case):
    mix, p, t, z = _sour() if case == "sour" else _boundary()
    print(f"[case] adapter solve_batch {case} cells={p.size}", flush=True)
    path = tmp_path / "cells.bin"
    write_cells(path, mix, p, t, z)
    out = run_check("solve", path)
    assert "mismatches=0" in out

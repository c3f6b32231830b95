"""Write tests/golden/device_flash_records.npz: inputs and the flash records this library's
gfx950 flash publishes for them (max 3 phases, FP64, fugacity tolerance 1e-11), for the
reference-side adapter check (oracle/adapter_check.cpp, tests/test_adapter.py):

  sour      every 4th cell of the sour-gas three-phase sweep (512 cells)
  boundary  feeds 5e-11 inside a two-phase boundary of the nc=3 workload (split attempts
            that end at a phase-fraction endpoint; 512 cells)
  c2        the first 512 cells of the nc=6 grid workload

Run on a GPU host:  python tests/golden/make_device_records.py [out.npz]
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import oracle_bindings as ob  # noqa: E402
from mpmc_hnu_amd import flash, workloads as W  # noqa: E402

KEYS = ("status", "phase_count", "beta", "zfactor", "branch", "x", "ln_phi")


def cases():
    mix, p, t, z = W.sour_gas_sweep(2048)
    idx = np.arange(0, 2048, 4)
    yield "sour", mix, p[idx], t[idx], np.ascontiguousarray(z[:, idx])
    mix = W.mixture_c1()
    p, t, z = W.cells_c1()
    # the boundary feeds are built from the oracle's record of the nc=3 workload
    p, t, z = W.boundary_feeds(p, t, ob.flash(mix, p, t, z, 1e-12, 3), 5e-11)
    yield "boundary", mix, p, t, z
    mix = W.mixture_linspaced(6)
    p, t, z = W.cells_c2(ncell=512)
    yield "c2", mix, p, t, z


def main(out: str):
    arrays = {}
    for name, mix, p, t, z in cases():
        m = flash.Mixture.from_dict(mix)
        try:
            rec = flash.to_host(flash.flash_batch(m, p, t, z, fugacity_tolerance=1e-11, max_phases=3))
        finally:
            m.close()
        for k in ("tc", "pc", "omega", "kij"):
            arrays[f"{name}_{k}"] = np.asarray(mix[k], dtype=np.float64)
        arrays[f"{name}_p"], arrays[f"{name}_t"], arrays[f"{name}_z"] = p, t, z
        for k in KEYS:
            arrays[f"{name}_{k}"] = rec[k]
        print(name, p.size, "cells", flush=True)
    np.savez_compressed(out, **arrays)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else str(HERE / "device_flash_records.npz"))

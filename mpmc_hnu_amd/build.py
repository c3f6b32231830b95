"""Build recipe for the gfx950 library (and the CPU checkers under oracle/).

    python -m mpmc_hnu_amd.build            # library + oracle (+ reference build if present)
    python -m mpmc_hnu_amd.build --lib-only

The library is built in-tree as mpmc_hnu_amd/libmpmc_hnu_amd.so with hipcc for gfx950.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmpmc_hnu_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: only explicit fma() calls fuse, as in the reference's x86-64 build.
# -packed-fp32-ops disabled: no v_pk_*_f32 generation.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-ffp-contract=off"]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def hip_sources():
    return sorted(CSRC.glob("*.hip"))


def host_sources():
    return sorted(CSRC.glob("*.cpp"))


def build_library(jobs: int = 8) -> Path:
    """Compile every .hip (device + C ABI) and .cpp (host adapters) into one .so."""
    objdir = ROOT / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    procs = []
    for src in hip_sources():
        obj = objdir / (src.stem + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
        print("+", " ".join(str(c) for c in cmd), flush=True)
        procs.append((subprocess.Popen([str(c) for c in cmd]), src))
        objs.append(obj)
        if len(procs) >= jobs:
            p, s = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError(f"hipcc failed on {s}")
    for p, s in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {s}")
    for src in host_sources():
        obj = objdir / (src.stem + ".host.o")
        _run([HIPCC, *HOST_FLAGS, "-I", ROOT / "include", "-c", src, "-o", obj])
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_oracle(with_reference: bool = True) -> None:
    _run(["make", "-C", ROOT / "oracle", "cpuref"])
    ref = os.environ.get("MPMC_REFERENCE", "/root/reference")
    if with_reference and Path(ref, "modules").is_dir():
        _run(["make", "-C", ROOT / "oracle", "ref", f"REF={ref}"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args(argv)
    build_library(a.jobs)
    if not a.lib_only:
        build_oracle()


if __name__ == "__main__":
    main()

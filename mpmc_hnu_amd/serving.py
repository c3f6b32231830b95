"""Continuous-batching flash server (mpmc_server_* in include/mpmc_hnu_amd.h) and the
serving workload: Poisson arrivals of flash jobs of log-uniform size and mixed
component counts.

The server plays the part of the reference's PT service loop
(modules/runtime/include/mpmc/runtime/pt_service.hpp): requests are PtFlashRequest
batches, answered with the same per-cell records as flash.flash_batch.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import native
from . import workloads as W

_P = ctypes.c_void_p


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


class Job:
    """One submitted job: its host inputs and result arrays (kept alive until done)."""

    def __init__(self, server, nc, p, t, z, hints=None):
        n = p.size
        self.server = server
        self.nc, self.n = nc, n
        self.inputs = (p, t, z, hints)
        self.out = {
            "status": np.zeros(n, np.int32),
            "phase_count": np.zeros(n, np.int32),
            "beta": np.zeros((3, n)),
            "zfactor": np.zeros((3, n)),
            "branch": np.zeros((3, n), np.int32),
            "x": np.zeros((3, nc, n)),
            "ln_phi": np.zeros((3, nc, n)),
        }
        self.id = None
        self.latency_s = None

    def done(self) -> bool:
        flag, lat = ctypes.c_int32(0), ctypes.c_double()
        self.server._check(self.server.lib.mpmc_server_poll(self.server.h, self.id, ctypes.byref(flag),
                                                            ctypes.byref(lat)), "poll")
        return bool(flag.value)

    def wait(self) -> dict:
        lat = ctypes.c_double()
        self.server._check(self.server.lib.mpmc_server_wait(self.server.h, self.id, ctypes.byref(lat)), "wait")
        self.latency_s = lat.value
        return self.out


class Server:
    def __init__(self, slots=2, precision=0, max_phases=3, fugacity_tolerance=1e-11, chunk_cells=262144,
                 small_job_cells=65536, hbm_cap_bytes=200 << 30, stream=None):
        import torch  # torch's HIP runtime first, see native.library

        self.lib = native.library()
        o = native.ServerOptions()
        self.lib.mpmc_server_default_options(ctypes.byref(o))
        o.slots, o.precision, o.max_phases = slots, precision, max_phases
        o.fugacity_tolerance, o.chunk_cells, o.small_job_cells = fugacity_tolerance, chunk_cells, small_job_cells
        o.hbm_cap_bytes = hbm_cap_bytes
        s = stream if stream is not None else torch.cuda.current_stream()
        o.stream = s.cuda_stream
        self.h = _P()
        self._check(self.lib.mpmc_server_create(ctypes.byref(o), ctypes.byref(self.h)), "create")
        self.slots = self.lib.mpmc_server_slots(self.h)
        self.device_bytes = self.lib.mpmc_server_device_bytes(self.h)
        self._mix_arrays = {}

    def _check(self, rc, what):
        if rc != 0:
            raise native.MpmcError(f"server {what} failed (rc={rc}): "
                                   f"{self.lib.mpmc_server_last_error().decode(errors='replace')}")

    def submit(self, mix: dict, p, t, z, hints=None) -> Job:
        """Queue one job; ``hints`` ([nhint <= 2, nc, ncell], NaN in component 0 = absent)
        are warm-start compositions (PtSolveHints initial stability starts)."""
        z = np.ascontiguousarray(z, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        t = np.ascontiguousarray(t, dtype=np.float64)
        nc = z.shape[0]
        key = id(mix)
        if key not in self._mix_arrays:
            self._mix_arrays[key] = (mix, [np.ascontiguousarray(mix[k], dtype=np.float64)
                                           for k in ("tc", "pc", "omega", "kij")])
        tc, pc, om, kij = self._mix_arrays[key][1]
        if hints is not None:
            hints = np.ascontiguousarray(hints, dtype=np.float64)
            if hints.ndim != 3 or hints.shape[1:] != z.shape or hints.shape[0] > 2:
                raise ValueError("submit: expected hints[nhint <= 2, nc, ncell]")
        job = Job(self, nc, p, t, z, hints)
        o = native.HostOutputs(*[job.out[k].ctypes.data for k in
                                 ("status", "phase_count", "beta", "zfactor", "branch", "x", "ln_phi")])
        jid = ctypes.c_int64()
        nh = 0 if hints is None else hints.shape[0]
        self._check(self.lib.mpmc_server_submit_hinted(self.h, nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), p.size,
                                                       _ptr(p), _ptr(t), _ptr(z), nh,
                                                       _ptr(hints) if nh else None, ctypes.byref(o),
                                                       ctypes.byref(jid)), "submit")
        job.id = jid.value
        return job

    def close(self):
        if self.h:
            self.lib.mpmc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serving_jobs(duration_s: float, rate_per_s: float = 200.0, seed: int = 7, max_cells: int = 1_000_000,
                 nc_range=(2, 10), warm_fraction: float = 0.2):
    """The serving workload: Poisson arrivals at `rate_per_s` for `duration_s`, cells per
    job log-uniform on [1, max_cells], nc uniform in nc_range, P/T from the grid fields of
    the nc=6 grid workload (a random contiguous run of cells), z ~ Dirichlet(2).
    A `warm_fraction` of the jobs repeat an earlier job's cells at pressures shifted by
    U[-2%, 2%] and name it in "warm_from": they are warm-started from its results."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p_field, t_field, _ = W.cells_c2()
    jobs = []
    at = 0.0
    while True:
        at += rng.exponential(1.0 / rate_per_s)
        if at >= duration_s:
            break
        if jobs and rng.uniform() < warm_fraction:
            k = int(rng.integers(0, len(jobs)))
            prior = jobs[k]
            p = prior["p"] * (1.0 + rng.uniform(-0.02, 0.02, prior["p"].size))
            jobs.append({"at": at, "nc": prior["nc"], "p": p, "t": prior["t"].copy(), "z": prior["z"],
                         "warm_from": k})
            continue
        n = int(np.exp(rng.uniform(0.0, np.log(max_cells))))
        n = max(1, min(n, max_cells))
        nc = int(rng.integers(nc_range[0], nc_range[1] + 1))
        start = int(rng.integers(0, p_field.size - n + 1))
        z = W._dirichlet(rng, 2.0, nc, n)
        jobs.append({"at": at, "nc": nc, "p": p_field[start:start + n].copy(), "t": t_field[start:start + n].copy(),
                     "z": z})
    return jobs


def run_load(server: Server, jobs, mixtures: dict) -> list:
    """Submit `jobs` at their arrival times (open loop) and wait for all; returns
    (cells, latency_s, warm_started) per job. A job naming a "warm_from" job is submitted
    with that job's published phase compositions as warm-start hints when it has finished
    by then (else without hints)."""
    t0 = time.perf_counter()
    handles = []
    for j in jobs:
        delay = j["at"] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        hints = None
        k = j.get("warm_from")
        if k is not None and handles[k][1].done():
            hints = W.warm_start_hints(handles[k][1].out, j["z"])
        handles.append((j, server.submit(mixtures[j["nc"]], j["p"], j["t"], j["z"], hints=hints), hints is not None))
    res = []
    for j, h, warm in handles:
        h.wait()
        res.append((j["p"].size, h.latency_s, warm))
    return res

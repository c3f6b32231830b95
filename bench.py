"""Benchmark: batched PR76 max-3-phase PT flash at nc=6 on 1,048,576 cells (one GPU).

A step is one flash pass over the whole batch (inputs already resident in HBM). Prints
one JSON line with cells/s, the roofline of the flash kernel and the CPU baseline
(the reference's own flash, built from its headers into oracle/_ref, or the oracle
restatement when that build is absent), timed on this host on a bounded sample.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix), dense
PEAK_HBM_GBS = 8000.0


def algorithmic_flops(counters: np.ndarray, nc: int) -> float:
    """Frozen per-cell work model: W = E_prop*F_prop(nc) + I_rr*(6nc+4) + I_rr3*(12nc+20)
    + E_tpd*6nc + 15nc + 3nc(nc-1)/2 with F_prop = 2nc^2 + 12nc + 90."""
    c = counters.astype(np.float64)
    e_prop, e_tpd, i_rr, i_rr3 = c[0].sum(), c[1].sum(), c[2].sum(), c[3].sum()
    ncell = counters.shape[1]
    f_prop = 2 * nc * nc + 12 * nc + 90
    return (e_prop * f_prop + i_rr * (6 * nc + 4) + i_rr3 * (12 * nc + 20) + e_tpd * 6 * nc
            + ncell * (15 * nc + 3 * nc * (nc - 1) / 2))


def algorithmic_bytes(nc: int, ncell: int) -> float:
    """HBM bytes per flash: inputs 8(nc+2), outputs status/count 8 + 3 phases x
    (beta, Z, branch 20 B + x, ln_phi 16nc) + counters 16, per cell."""
    return ncell * (8 * (nc + 2) + 8 + 3 * (20 + 16 * nc) + 16)


def host_cores():
    """(cores this process may use, std::thread::hardware_concurrency-like count, lscpu CPUs)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    lscpu = None
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("CPU(s):"):
                lscpu = int(line.split(":")[1])
    except Exception:
        pass
    return usable, os.cpu_count(), lscpu


def cpu_baseline(mix, p, t, z, sample: int, sample_1core: int = 8192, repeats: int = 5):
    """The reference's own flash (oracle/_ref, built from its headers) on the first
    ``sample`` cells of the workload with every core this process may use (at most 16, a
    GPU box's CPU share), and on the first ``sample_1core`` cells with one thread: warm-up,
    then the median of ``repeats`` runs each. Falls back to the oracle restatement when the
    reference build is absent."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_bindings as ob

    which = "ref" if ob.have_ref() else "cpuref"
    usable, hw, lscpu = host_cores()
    cores = max(1, min(usable, 16))

    def timed(n, threads):
        ps, ts, zs = p[:n].copy(), t[:n].copy(), np.ascontiguousarray(z[:, :n])
        ob.flash(mix, ps[:256], ts[:256], zs[:, :256], 1e-11, 3, threads=threads, which=which)  # warm-up
        runs = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            ob.flash(mix, ps, ts, zs, 1e-11, 3, threads=threads, which=which)
            runs.append(time.perf_counter() - t0)
        return float(np.median(runs)), runs

    n = min(sample, p.size)
    dt, runs = timed(n, cores)
    n1 = min(sample_1core, p.size)
    dt1, runs1 = timed(n1, 1)
    return {
        "value": n / dt,
        "unit": "cells/s",
        "cores": cores,
        "kind": "reference" if which == "ref" else "port",
        "sample": f"first {n} cells of the same nc=6 grid workload, max3 flash, {cores} threads, median of "
                  f"{repeats} runs {dt:.2f} s (runs {', '.join(f'{r:.2f}' for r in runs)})",
        "one_core": {"value": n1 / dt1, "unit": "cells/s", "cores": 1,
                     "sample": f"first {n1} cells, 1 thread, median of {repeats} runs {dt1:.2f} s"},
        "host": {"usable_cpus": usable, "os_cpu_count": hw, "lscpu_cpus": lscpu},
    }


# HBM bytes of one Newton evaluation at nc=6 on 100x100x50, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE counters (k_asm_cell + k_asm_rows), each scaled by what the counter reads for
# a plain 1 GiB streaming kernel (profiles/assembly_2p_nc6_100x100x50_kernel_stats.md).
# HBM bytes of one flash of the 1M-cell nc=6 grid: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# summed over the launch sequence (profiles/flash_nc6_1M_hbm_traffic.md), FP64 and
# decision-exact mixed precision.
FLASH_FP64_TRAFFIC_BYTES = 2329324 * 2.0 * 1024 + 1766400 * 1.0 * 1024
FLASH_EXACT_TRAFFIC_BYTES = 3380342 * 2.0 * 1024 + 1914209 * 1.0 * 1024
ASSEMBLY_TRAFFIC_BYTES = (36615 + 1072624) * 2.0 * 1024 + (728516 + 2578166) * 1.0 * 1024


def assembly_leg(dims, steps: int, warmup: int) -> dict:
    """One Newton evaluation of the two-phase natural-variable system (residual + block
    Jacobian) on a structured grid, nc=6, inputs resident in HBM."""
    import torch

    from mpmc_hnu_amd import assembly as A, flash, workloads as W

    nc = 6
    mix = W.mixture_linspaced(nc)
    grid = W.tpfa_grid(*dims)
    m = flash.Mixture.from_dict(mix)

    def roots(p, t, x):
        r = flash.to_host(flash.phase_batch(m, p, t, x))
        return r["root_count"], r["root_flags"]

    print(f"[bench] assembly: building state on {dims}", file=sys.stderr, flush=True)
    q, br = W.two_phase_state(nc, grid["ncell"], roots, seed=8, dims=dims)
    g = A.Grid(grid)
    md = A.default_model(nc)
    inv, st, flag = A.inventory(m, g, md, q, br)
    prev = inv * 0.99
    dt = 86400.0
    out = A.assemble(m, g, md, q, br, prev, dt)
    assert out["domain_error"] == 0, "assembly state raised a domain error"
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        A.assemble(m, g, md, q, br, prev, dt, out=out, check_domain=False)
    torch.cuda.synchronize()
    e0 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    e1 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    qd, bd, pd = out["_inputs"]
    for k in range(steps):
        e0[k].record(stream)
        A.assemble(m, g, md, qd, bd, pd, dt, out=out, check_domain=False)
        e1[k].record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([e0[k].elapsed_time(e1[k]) for k in range(steps)]))
    # the same evaluation captured once into a HIP graph and replayed (one launch per Newton step)
    plan = A.NewtonPlan(m, g, md, qd, bd, pd, dt)
    for _ in range(warmup):
        plan.launch(stream)
    torch.cuda.synchronize()
    for k in range(steps):
        e0[k].record(stream)
        plan.launch(stream)
        e1[k].record(stream)
    torch.cuda.synchronize()
    plan_ms = float(np.mean([e0[k].elapsed_time(e1[k]) for k in range(steps)]))
    plan.close()
    # Schur reduction of the assembled system to (nc+1) x (nc+1) blocks
    red = A.reduce(g, nc, out)
    for _ in range(warmup):
        A.reduce(g, nc, out, out=red)
    torch.cuda.synchronize()
    for k in range(steps):
        e0[k].record(stream)
        A.reduce(g, nc, out, out=red)
        e1[k].record(stream)
    torch.cuda.synchronize()
    red_ms = float(np.mean([e0[k].elapsed_time(e1[k]) for k in range(steps)]))
    red_singular = int(torch.sum(red["status"] != 0).item())
    del red
    # the post-SNES phase-transition scan of the same state: inventory -> feed, flash, decision
    from mpmc_hnu_amd import scan as S

    scan = {}
    for prec, dmode, name in ((0, 0, "fp64"), (1, 0, "mixed")):
        res = S.transition_scan_2p(m, g, md, qd, bd, precision=prec, density_mode=dmode)
        times = []
        for _ in range(max(1, min(steps, 3))):
            t0 = time.perf_counter()
            res = S.transition_scan_2p(m, g, md, qd, bd, precision=prec, density_mode=dmode)
            times.append((time.perf_counter() - t0) * 1e3)
        codes = np.bincount(res["code"].cpu().numpy(), minlength=4)
        scan[name] = {"ms": float(np.mean(times)), "cells_per_s": grid["ncell"] / (np.mean(times) * 1e-3),
                      "same": int(codes[0]), "proposals": int(codes[1]), "indeterminate": int(codes[2]),
                      "errors": int(codes[3]), "status": res["status"], "first_stop": res["first_stop"]}
        print(f"[bench] transition scan ({name}): {scan[name]['ms']:.1f} ms, {codes[1]} proposals, "
              f"{codes[2]} indeterminate", file=sys.stderr, flush=True)
    ncell, nface, Q = grid["ncell"], grid["owner"].size, 2 * nc + 1
    nnz = ncell * Q * Q + 2 * nface * (nc + 1) * Q
    algo_bytes = 8.0 * (ncell * (Q + nc + 3) + nface * 2 + nnz + ncell * Q)
    gbs = algo_bytes / (ms * 1e-3) / 1e9
    # reduction: reads the unreduced system (and each neighbour's M once per face side),
    # writes M, the reduced residual and blocks
    P = nc + 1
    red_bytes = 8.0 * (ncell * (Q * Q + Q) + 2 * nface * (P * Q + nc * (P + 1))
                       + ncell * (nc * (P + 1) + P + P * P) + 2 * nface * P * P)
    red_gbs = red_bytes / (red_ms * 1e-3) / 1e9
    print(f"[bench] assembly: {ms:.3f} ms per Newton evaluation, {gbs:.0f} GB/s algorithmic "
          f"({ncell} cells, {nface} faces); graph replay {plan_ms:.3f} ms; Schur reduction to "
          f"{P}x{P} blocks {red_ms:.3f} ms ({red_gbs:.0f} GB/s, {red_singular} singular cells)",
          file=sys.stderr, flush=True)
    g.close()
    m.close()
    out = {"grid": list(dims), "cells": ncell, "faces": int(nface), "unknowns_per_cell": Q, "ms": ms,
           "graph_replay_ms": plan_ms,
           "algorithmic_bytes": algo_bytes, "achieved_gbs": gbs, "peak_gbs": PEAK_HBM_GBS,
           "frac": gbs / PEAK_HBM_GBS,
           "what": "residual + block Jacobian (diagonal and off-diagonal Q x Q blocks), two-phase natural variables",
           "reduced": {"ms": red_ms, "block": P, "algorithmic_bytes": red_bytes, "achieved_gbs": red_gbs,
                       "frac": red_gbs / PEAK_HBM_GBS, "singular_cells": red_singular,
                       "what": "Schur reduction of the assembled system to (nc+1) x (nc+1) blocks "
                               "(mpmc_assembly_reduce_2p)"},
           "transition_scan": dict(scan, what="post-SNES phase-transition scan of the same state (inventory, feed, "
                                               "flash, per-cell decision; host clock, includes the summary sync)")}
    if tuple(dims) == (100, 100, 50):
        out["roofline"] = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                           "frac": gbs / PEAK_HBM_GBS, "traffic": ASSEMBLY_TRAFFIC_BYTES,
                           "traffic_source": "profiles/assembly_2p_nc6_100x100x50_kernel_stats.md"}
    return out


def uneven_workload(scale: float = 1.0):
    """The uneven multi-mixture batch: 256 mixtures, nc ~ U{2..12}, cells per mixture
    log-uniform on [1e3, 2e5] times ``scale`` (about 9.7M cells at scale 1), 30% of each
    mixture's cells within +-2% in P of a phase boundary found by scan and bisection on P
    with the GPU flash. Returns (layout, mixtures, [(p, t, z)] per mixture, cells,
    near-boundary cells)."""
    from mpmc_hnu_amd import flash, workloads as W

    layout = W.uneven_layout(scale=scale)
    mixes = [flash.Mixture.from_dict(m["mixture"]) for m in layout]
    base = W.uneven_base_points(layout)

    def phase_count(reqs):
        res = [None] * len(reqs)
        by_nc = {}
        for i, (k, _, _, _) in enumerate(reqs):
            by_nc.setdefault(layout[k]["nc"], []).append(i)
        for nc, idx in by_nc.items():
            cap = flash.multi_max_mixtures(nc)
            for s0 in range(0, len(idx), cap):
                chunk = idx[s0:s0 + cap]
                sizes = [reqs[i][1].size for i in chunk]
                off = np.concatenate([[0], np.cumsum(sizes)])
                out = flash.flash_multi([mixes[reqs[i][0]] for i in chunk], off,
                                        np.concatenate([reqs[i][1] for i in chunk]),
                                        np.concatenate([reqs[i][2] for i in chunk]),
                                        np.concatenate([reqs[i][3] for i in chunk], axis=1), counters=False)
                pc = out["phase_count"].cpu().numpy()
                for j, i in enumerate(chunk):
                    res[i] = pc[off[j]:off[j + 1]]
        return res

    pb = W.boundary_pressures(phase_count, base)
    cells = W.uneven_cells(layout, base, pb)
    ncell = int(sum(m["ncell"] for m in layout))
    near = int(sum(round(0.3 * m["ncell"]) if np.isfinite(b).any() else 0 for m, b in zip(layout, pb)))
    return layout, mixes, cells, ncell, near


def uneven_leg(steps: int, warmup: int, scale: float = 1.0) -> dict:
    """The uneven multi-mixture batch (uneven_workload). Timed as one flash plan (ragged
    launches grouped by nc, captured in a HIP graph) against the same segments flashed one
    mixture at a time; both must give identical records."""
    import torch

    from mpmc_hnu_amd import flash

    t0 = time.perf_counter()
    layout, mixes, cells, ncell, near = uneven_workload(scale)
    dev = torch.device("cuda", 0)
    segments = [(mixes[k], torch.as_tensor(p, device=dev), torch.as_tensor(t, device=dev),
                 torch.as_tensor(z, device=dev)) for k, (p, t, z) in enumerate(cells)]
    print(f"[bench] uneven: {len(layout)} mixtures, {ncell} cells ({near} near a boundary), "
          f"built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    # launch-group autotuner: the mixtures per launch group chosen by timing graph replays
    plan, tune = flash.autotune_plan(segments, candidates=(0, 16, 8), replays=1, stream=stream, counters=False)
    print(f"[bench] uneven: autotuned mixtures per group {plan.group_mixtures} "
          f"(ms per replay by candidate: {', '.join(f'{k}: {v:.0f}' for k, v in tune.items())})",
          file=sys.stderr, flush=True)
    for _ in range(warmup):
        plan.launch(stream)
    torch.cuda.synchronize()
    e0 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    e1 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    for k in range(steps):
        e0[k].record(stream)
        plan.launch(stream)
        e1[k].record(stream)
    torch.cuda.synchronize()
    plan_ms = float(np.mean([e0[k].elapsed_time(e1[k]) for k in range(steps)]))
    # the same segments one mixture at a time, sharing one workspace
    nmax = max(m["ncell"] for m in layout)
    ws = torch.empty(max(flash.workspace_bytes(nc, nmax) for nc in range(2, 13)), dtype=torch.uint8, device=dev)
    alone = [flash.allocate_outputs(mix.nc, p.numel(), dev, counters=False) for mix, p, _, _ in segments]
    a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a0.record(stream)
    for (mix, p, t, z), out in zip(segments, alone):
        flash.flash_batch(mix, p, t, z, out=out, counters=False, workspace=ws, stream=stream)
    a1.record(stream)
    torch.cuda.synchronize()
    alone_ms = a0.elapsed_time(a1)
    identical = all(torch.equal(o1[k], o2[k]) if k not in ("beta", "zfactor", "x", "ln_phi") else
                    torch.equal(torch.nan_to_num(o1[k], nan=7.0), torch.nan_to_num(o2[k], nan=7.0))
                    for o1, o2 in zip(plan.outputs, alone)
                    for k in ("status", "phase_count", "beta", "zfactor", "branch", "x", "ln_phi"))
    counts = {}
    for o in plan.outputs:
        v, c = np.unique(o["phase_count"].cpu().numpy(), return_counts=True)
        for a, b in zip(v, c):
            counts[int(a)] = counts.get(int(a), 0) + int(b)
    groups = plan.groups
    # the same plan in mixed precision (same launch groups)
    mplan = flash.Plan(segments, precision=1, group_mixtures=plan.group_mixtures, counters=False)
    for _ in range(warmup):
        mplan.launch(stream)
    torch.cuda.synchronize()
    for k in range(steps):
        e0[k].record(stream)
        mplan.launch(stream)
        e1[k].record(stream)
    torch.cuda.synchronize()
    mixed_ms = float(np.mean([e0[k].elapsed_time(e1[k]) for k in range(steps)]))
    mixed_count_diff = int(sum(int(torch.sum(o1["phase_count"] != o2["phase_count"]).item())
                               for o1, o2 in zip(plan.outputs, mplan.outputs)))
    mplan.close()
    plan.close()
    for m in mixes:
        m.close()
    print(f"[bench] uneven: plan {plan_ms:.1f} ms ({ncell / (plan_ms * 1e-3) / 1e6:.2f} M cells/s, {groups} launch "
          f"groups); one mixture at a time {alone_ms:.1f} ms; identical records: {identical}; mixed-precision plan "
          f"{mixed_ms:.1f} ms ({mixed_count_diff} phase-count differences)", file=sys.stderr, flush=True)
    return {"mixtures": len(layout), "cells": ncell, "near_boundary_cells": near, "launch_groups": groups,
            "plan_ms": plan_ms, "cells_per_s": ncell / (plan_ms * 1e-3), "one_mixture_at_a_time_ms": alone_ms,
            "group_mixtures": plan.group_mixtures, "autotune_ms_by_group_mixtures": tune,
            "speedup_vs_one_at_a_time": alone_ms / plan_ms, "identical_records": bool(identical),
            "phase_counts": counts, "precision": "fp64",
            "mixed_plan_ms": mixed_ms, "mixed_cells_per_s": ncell / (mixed_ms * 1e-3),
            "mixed_phase_count_differences": mixed_count_diff,
            "what": "max-3-phase flash of every cell; ragged launches grouped by nc, one HIP graph"}


def copy_leg(steps: int) -> dict:
    """Plain streaming kernel over 1 GiB (2^27 doubles, dst = -src: one elementwise read
    and one write per element): the calibration kernel for HBM byte counters
    (FETCH_SIZE / WRITE_SIZE read 1 GiB each per launch when exact)."""
    import torch

    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        torch.neg(src, out=dst)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    gbs = 2.0 * n * 8 / (ms * 1e-3) / 1e9
    print(f"[bench] copy: {ms:.3f} ms per 1 GiB copy, {gbs:.0f} GB/s read+write", file=sys.stderr, flush=True)
    return {"bytes_each_way": n * 8, "ms": ms, "gbs_read_plus_write": gbs}


def serve_leg(seconds: float) -> dict:
    """Serving workload: Poisson arrivals at 200 jobs/s, cells per job log-uniform on
    [1, 1e6], nc uniform in 2..10, through the continuous-batching server, FP64 and mixed."""
    from mpmc_hnu_amd import serving as S, workloads as W

    print(f"[bench] serving: generating {seconds} s of arrivals", file=sys.stderr, flush=True)
    jobs = S.serving_jobs(seconds)
    mixtures = {nc: W.mixture_linspaced(nc) for nc in range(2, 11)}
    out = {"arrival_rate_per_s": 200.0, "arrival_seconds": seconds, "jobs": len(jobs),
           "cells": int(sum(j["p"].size for j in jobs)), "small_job_cells": 65536,
           "warm_start_jobs": int(sum("warm_from" in j for j in jobs))}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        srv = S.Server(slots=2, precision=prec)
        t0 = time.perf_counter()
        res = S.run_load(srv, jobs, mixtures)
        makespan = time.perf_counter() - t0
        srv.close()
        n = np.array([r[0] for r in res])
        lat = np.array([r[1] for r in res]) * 1e3
        warm = np.array([r[2] for r in res])
        small = n <= 65536
        out[name] = {
            "p50_ms_small": float(np.percentile(lat[small], 50)),
            "p99_ms_small": float(np.percentile(lat[small], 99)),
            "p50_ms_all": float(np.percentile(lat, 50)),
            "p99_ms_all": float(np.percentile(lat, 99)),
            "cells_per_s": float(n.sum() / makespan),
            "makespan_s": makespan,
            "warm_started_jobs": int(warm.sum()),
            "slots": srv.slots,
        }
        print(f"[bench] serving {name}: p50/p99 small {out[name]['p50_ms_small']:.2f}/{out[name]['p99_ms_small']:.2f} ms, "
              f"all {out[name]['p50_ms_all']:.1f}/{out[name]['p99_ms_all']:.1f} ms, {out[name]['cells_per_s']:.3e} cells/s",
              file=sys.stderr, flush=True)
    # the latency floor: the same job mix at a tenth of the rate (20 jobs/s), mixed precision
    light = S.serving_jobs(seconds, rate_per_s=20.0, seed=8)
    srv = S.Server(slots=2, precision=1)
    res = S.run_load(srv, light, mixtures)
    srv.close()
    n = np.array([r[0] for r in res])
    lat = np.array([r[1] for r in res]) * 1e3
    small = n <= 65536
    out["mixed_at_20_jobs_per_s"] = {
        "jobs": len(light), "small_jobs": int(small.sum()),
        "p50_ms_small": float(np.percentile(lat[small], 50)) if small.any() else None,
        "p99_ms_small": float(np.percentile(lat[small], 99)) if small.any() else None,
        "p50_ms_all": float(np.percentile(lat, 50)), "p99_ms_all": float(np.percentile(lat, 99))}
    print(f"[bench] serving mixed at 20 jobs/s: p50/p99 small {out['mixed_at_20_jobs_per_s']['p50_ms_small']:.2f}/"
          f"{out['mixed_at_20_jobs_per_s']['p99_ms_small']:.2f} ms ({int(small.sum())} small of {len(light)} jobs)",
          file=sys.stderr, flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=None, help="prefix of the 1,048,576-cell grid (default: all)")
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--assembly-dims", type=str, default="100,100,50",
                    help="grid of the assembly leg (nx,ny,nz; empty: skip)")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="seconds of Poisson arrivals for the serving leg (0: skip)")
    ap.add_argument("--uneven-scale", type=float, default=1.0,
                    help="cells per mixture of the uneven leg, as a multiple of the 9.7M-cell batch")
    ap.add_argument("--legs", type=str, default="flash,assembly,uneven,serving",
                    help="comma list of legs to run (flash, assembly, uneven, serving; copy: a 1 GiB streaming "
                         "kernel that calibrates HBM byte counters); the JSON line is complete only with the "
                         "first four")
    a = ap.parse_args(argv)
    legs = set(a.legs.split(","))
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              "(cells are independent, so N GPUs would shard cell ranges with no exchange).", file=sys.stderr)
        sys.exit(2)

    import torch

    from mpmc_hnu_amd import flash, workloads as W

    torch.cuda.set_device(0)
    if "flash" not in legs:
        # a partial run (profiling one leg): print that leg's object only
        part = {}
        if "copy" in legs:
            part["copy"] = copy_leg(a.steps)
        if "assembly" in legs and a.assembly_dims:
            part["assembly"] = assembly_leg(tuple(int(v) for v in a.assembly_dims.split(",")), a.steps, a.warmup)
        if "uneven" in legs:
            part["uneven"] = uneven_leg(a.steps, a.warmup, a.uneven_scale)
        if "serving" in legs and a.serve_seconds > 0:
            part["serving"] = serve_leg(a.serve_seconds)
        print(json.dumps(part), flush=True)
        return
    mix = W.mixture_linspaced(6)
    p, t, z = W.cells_c2(ncell=a.cells)
    nc, ncell = z.shape
    dev = torch.device("cuda", 0)
    pd = torch.as_tensor(p, device=dev)
    td = torch.as_tensor(t, device=dev)
    zd = torch.as_tensor(z, device=dev)
    m = flash.Mixture.from_dict(mix)
    stream = torch.cuda.current_stream()

    def run(precision):
        out = flash.allocate_outputs(nc, ncell, dev, counters=True)
        print(f"[bench] precision={precision} nc={nc} cells={ncell} warmup={a.warmup} steps={a.steps}",
              file=sys.stderr, flush=True)
        for _ in range(a.warmup):
            flash.flash_batch(m, pd, td, zd, out=out, precision=precision)
        torch.cuda.synchronize()
        ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
        ev1 = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
        w0 = time.perf_counter()
        for k in range(a.steps):
            ev0[k].record(stream)
            flash.flash_batch(m, pd, td, zd, out=out, stream=stream, precision=precision)
            ev1[k].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        gpu_ms = float(np.mean([ev0[k].elapsed_time(ev1[k]) for k in range(a.steps)]))
        host = flash.to_host(out)
        print(f"[bench] precision={precision}: {gpu_ms:.3f} ms/flash (events), wall {wall * 1e3 / a.steps:.3f} ms",
              file=sys.stderr, flush=True)
        return gpu_ms, wall * 1e3 / a.steps, host, out

    def run_closure(fl):
        """Dual closure of every two-phase cell of a flash result (the same cells)."""
        res = flash.closure_batch(m, pd, td, zd, fl, stream=stream)
        torch.cuda.synchronize()
        e0 = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
        e1 = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
        for k in range(a.steps):
            e0[k].record(stream)
            flash.closure_batch(m, pd, td, zd, fl, stream=stream, out=res)
            e1[k].record(stream)
        torch.cuda.synchronize()
        ms = float(np.mean([e0[k].elapsed_time(e1[k]) for k in range(a.steps)]))
        st = flash.to_host(res)["status"]
        lin = int(np.sum(st == 0))
        print(f"[bench] closure: {ms:.3f} ms for {ncell} cells, {lin} linearized", file=sys.stderr, flush=True)
        return ms, lin

    fp64_ms, fp64_wall, h64, d64 = run(0)
    closure_ms, closure_lin = run_closure(d64)
    del d64
    exact_ms, exact_wall, hex_, _ = run(1)
    fast_ms, fast_wall, hfs, _ = run(2)

    def agreement(h):
        status_diff = int(np.sum((h64["status"] & 0xFFF) != (h["status"] & 0xFFF)))
        count_diff = int(np.sum(h64["phase_count"] != h["phase_count"]))
        same = ((h64["status"] & 0xFFF) == (h["status"] & 0xFFF)) & (h64["phase_count"] == h["phase_count"])
        rel = 0.0
        for ph in range(3):
            have = same & (h64["phase_count"] > ph)
            if have.any():
                rel = max(rel, float(np.max(np.abs(h["beta"][ph, have] - h64["beta"][ph, have]) /
                                           np.abs(h64["beta"][ph, have]))))
                rel = max(rel, float(np.max(np.abs(h["x"][ph][:, have] - h64["x"][ph][:, have]) /
                                           np.abs(h64["x"][ph][:, have]))))
        return {"status_mismatches": status_diff, "phase_count_mismatches": count_diff,
                "max_rel_diff_beta_x": rel}

    exact_vs = agreement(hex_)
    fast_vs = agreement(hfs)
    # the headline is the mixed-precision rate when its phase counts (and statuses) equal
    # FP64's on every cell of the workload, else the FP64 rate
    exact_ok = exact_vs["status_mismatches"] == 0 and exact_vs["phase_count_mismatches"] == 0
    host = hex_ if exact_ok else h64
    flops = algorithmic_flops(h64["counters"], nc)
    achieved = flops / (fp64_ms * 1e-3) / 1e12
    flops_mixed = algorithmic_flops(hex_["counters"], nc)
    achieved_mixed = flops_mixed / (exact_ms * 1e-3) / 1e12
    status = host["status"]
    accepted = int(np.sum((status & 0xF) == 0))
    counts = {int(k): int(v) for k, v in zip(*np.unique(host["phase_count"][(status & 0xF) == 0], return_counts=True))}
    print(f"[bench] exact mixed vs fp64: {exact_vs}; fast mixed vs fp64: {fast_vs}; accepted {accepted}, "
          f"phase counts {counts}", file=sys.stderr, flush=True)
    asm = None
    if a.assembly_dims and "assembly" in legs:
        asm = assembly_leg(tuple(int(v) for v in a.assembly_dims.split(",")), a.steps, a.warmup)
    uneven = None
    if "uneven" in legs:
        uneven = uneven_leg(a.steps, a.warmup, a.uneven_scale)
    serving = None
    if a.serve_seconds > 0 and "serving" in legs:
        serving = serve_leg(a.serve_seconds)
    cpu = None
    if not a.no_cpu:
        print("[bench] timing CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(mix, p, t, z, a.cpu_sample)
    m.close()
    ms_per_step = exact_wall if exact_ok else fp64_wall
    line = {
        "metric": "flash cells/s at nc=6 (mixed vs FP64 GPU, same phase count); serving p99",
        "value": ncell / (ms_per_step * 1e-3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (mixed)" if exact_ok else "f64",
        "data": "synthetic",
        "config": {
            "workload": "PR76 max-3-phase PT flash, nc=6 linspaced mixture, 128x128x64 grid, "
                        "Gaussian-random-field P 5-40 MPa / T 320-420 K, z~Dirichlet(2)",
            "cells": ncell,
            "nc": nc,
            "precision": ("decision-exact mixed (FP32 screening of stability trials; every decision-bearing "
                          "search in FP64)") if exact_ok else "fp64",
            "fugacity_tolerance": 1e-11,
        },
        "fp64_cells_per_s": ncell / (fp64_ms * 1e-3),
        "mixed_cells_per_s": ncell / (exact_ms * 1e-3),
        "mixed_speedup_vs_fp64": fp64_ms / exact_ms,
        "mixed_vs_fp64": exact_vs,
        "fast_mixed": {"cells_per_s": ncell / (fast_ms * 1e-3), "speedup_vs_fp64": fp64_ms / fast_ms,
                       "vs_fp64": fast_vs,
                       "what": "precision 2: FP32 split attempts and Newton polish; decisions differ where the "
                               "reference's own split search ends on a line-search failure"},
        "accepted_phase_counts": counts,
        "closure": {"ms": closure_ms, "cells": ncell, "linearized_cells": closure_lin,
                    "what": "Dual<nc+1> closure + implicit sensitivity of every accepted two-phase cell "
                            "(primal Z, rho; d(beta, x, y, Z, rho)/d(p, T, z_red)) on the FP64 flash result"},
        "roofline": {
            "bound": "valu",
            "pipe": "vector ALU (FP64 and FP32; no MFMA: no dense contraction in the flash); peak = FP64 vector",
            "achieved": achieved_mixed if exact_ok else achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": (achieved_mixed if exact_ok else achieved) / PEAK_FP64_TFLOPS,
            "traffic": ((FLASH_EXACT_TRAFFIC_BYTES if exact_ok else FLASH_FP64_TRAFFIC_BYTES)
                        if ncell == 1048576 and nc == 6 else None),
            "traffic_source": "profiles/flash_nc6_1M_hbm_traffic.md",
            "kernel": ("decision-exact mixed flash, whole launch sequence" if exact_ok else
                       "staged FP64 flash, whole launch sequence"),
            "kernel_ms": exact_ms if exact_ok else fp64_ms,
            "algorithmic_flop_per_launch": flops_mixed if exact_ok else flops,
            "algorithmic_bytes_per_launch": algorithmic_bytes(nc, ncell),
        },
        "fp64_roofline": {
            "bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS, "kernel_ms": fp64_ms, "algorithmic_flop_per_launch": flops,
            "traffic": FLASH_FP64_TRAFFIC_BYTES if ncell == 1048576 and nc == 6 else None,
            "traffic_source": "profiles/flash_nc6_1M_hbm_traffic.md",
        },
        "assembly": asm,
        "uneven": uneven,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

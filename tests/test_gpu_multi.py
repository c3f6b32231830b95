"""Ragged multi-mixture batches (mpmc_flash_multi) and flash plans (mpmc_flash_plan_*):
every cell's record is bit-identical to mpmc_flash_batch on its own mixture's cells, and a
plan's graph replays reproduce the eager results on updated inputs."""
import numpy as np
import pytest

from mpmc_hnu_amd import workloads as W

KEYS = ("status", "phase_count", "beta", "zfactor", "branch", "x", "ln_phi", "counters")


def _host(out):
    from mpmc_hnu_amd import flash
    return flash.to_host(out)


def _alone(mix, p, t, z):
    from mpmc_hnu_amd import flash
    return _host(flash.flash_batch(mix, p, t, z, max_phases=3))


def _same(a, b, what):
    for k in KEYS:
        assert np.array_equal(a[k], b[k], equal_nan=True), (what, k)


@pytest.mark.gpu
@pytest.mark.parametrize("nc,sizes", [(4, [300, 0, 1, 777, 64, 2049]), (2, [5, 1000, 3]), (9, [200, 150])])
def test_multi_equals_each_mixture_alone(gpu, nc, sizes):
    from mpmc_hnu_amd import flash

    mixes = [flash.Mixture.from_dict(W.mixture_linspaced(nc, seed=10 + k)) for k in range(len(sizes))]
    try:
        cells = [W.cells_uniform(nc, max(n, 1), seed=50 + k) for k, n in enumerate(sizes)]
        cells = [(p[:n], t[:n], z[:, :n]) for (p, t, z), n in zip(cells, sizes)]
        off = np.concatenate([[0], np.cumsum(sizes)])
        p = np.concatenate([c[0] for c in cells])
        t = np.concatenate([c[1] for c in cells])
        z = np.ascontiguousarray(np.concatenate([c[2] for c in cells], axis=1))
        print(f"[case] flash_multi nc={nc} sizes={sizes}", flush=True)
        got = _host(flash.flash_multi(mixes, off, p, t, z))
        for k, n in enumerate(sizes):
            if n == 0:
                continue
            a = _alone(mixes[k], *cells[k])
            sl = slice(off[k], off[k + 1])
            part = {key: got[key][..., sl] for key in KEYS}
            _same(part, a, (nc, k))
    finally:
        for m in mixes:
            m.close()


@pytest.mark.gpu
def test_mixed_precision_multi_equals_each_mixture_alone(gpu):
    """The mixed-precision schedule on a ragged batch: the reduced-precision per-cell
    constants and every search read the cell's own mixture, so records are bit-identical
    to mixed-precision mpmc_flash_batch on each mixture's cells."""
    from mpmc_hnu_amd import flash

    nc, sizes = 6, [500, 1, 1200, 64]
    mixes = [flash.Mixture.from_dict(W.mixture_linspaced(nc, seed=30 + k)) for k in range(len(sizes))]
    try:
        cells = [W.cells_uniform(nc, n, seed=70 + k) for k, n in enumerate(sizes)]
        off = np.concatenate([[0], np.cumsum(sizes)])
        p = np.concatenate([c[0] for c in cells])
        t = np.concatenate([c[1] for c in cells])
        z = np.ascontiguousarray(np.concatenate([c[2] for c in cells], axis=1))
        print(f"[case] mixed-precision flash_multi nc={nc} sizes={sizes}", flush=True)
        got = _host(flash.flash_multi(mixes, off, p, t, z, precision=1))
        for k in range(len(sizes)):
            a = _host(flash.flash_batch(mixes[k], *cells[k], max_phases=3, precision=1))
            sl = slice(off[k], off[k + 1])
            _same({key: got[key][..., sl] for key in KEYS}, a, ("mixed", k))
    finally:
        for m in mixes:
            m.close()


@pytest.mark.gpu
def test_multi_many_mixtures_and_limits(gpu):
    from mpmc_hnu_amd import flash, native

    nc = 3
    cap = flash.multi_max_mixtures(nc)
    assert cap >= 64
    mixes = [flash.Mixture.from_dict(W.mixture_linspaced(nc, seed=100 + k)) for k in range(48)]
    try:
        rng = np.random.Generator(np.random.PCG64(7))
        sizes = [int(v) for v in rng.integers(1, 40, size=len(mixes))]
        off = np.concatenate([[0], np.cumsum(sizes)])
        p, t, z = W.cells_uniform(nc, int(off[-1]), seed=8)
        print(f"[case] flash_multi {len(mixes)} mixtures, {off[-1]} cells", flush=True)
        got = _host(flash.flash_multi(mixes, off, p, t, z))
        for k in range(len(mixes)):
            sl = slice(off[k], off[k + 1])
            a = _alone(mixes[k], p[sl], t[sl], np.ascontiguousarray(z[:, sl]))
            _same({key: got[key][..., sl] for key in KEYS}, a, k)
        other = flash.Mixture.from_dict(W.mixture_linspaced(4))
        try:
            with pytest.raises(native.MpmcError):
                flash.flash_multi([mixes[0], other], [0, 1, 2], p[:2], t[:2], z[:, :2])
        finally:
            other.close()
    finally:
        for m in mixes:
            m.close()


@pytest.mark.gpu
def test_plan_groups_replays_and_updates(gpu):
    import torch

    from mpmc_hnu_amd import flash

    rng = np.random.Generator(np.random.PCG64(3))
    mixes = {nc: [flash.Mixture.from_dict(W.mixture_linspaced(nc, seed=200 + k)) for k in range(2)]
             for nc in (2, 3, 6, 12)}
    try:
        segments = []
        for j in range(14):
            nc = (2, 3, 6, 12)[j % 4]
            mix = mixes[nc][int(rng.integers(0, 2))]
            n = int(rng.integers(1, 600))
            p, t, z = W.cells_uniform(nc, n, seed=300 + j)
            segments.append((mix, torch.as_tensor(p, device="cuda"), torch.as_tensor(t, device="cuda"),
                             torch.as_tensor(z, device="cuda")))
        print(f"[case] plan over {len(segments)} segments", flush=True)
        plan = flash.Plan(segments)
        try:
            assert plan.groups == 4
            torch.cuda.synchronize()
            for (mix, p, t, z), out in zip(segments, plan.outputs):
                _same(_host(out), _alone(mix, p, t, z), "creation run")
            for rep in range(3):
                # new input contents, eager work between replays
                for mix, p, t, z in segments:
                    p.mul_(1.0 + 0.05 * (rep + 1))
                _ = torch.ones(1024, device="cuda").sum().item()
                plan.launch()
                torch.cuda.synchronize()
                for (mix, p, t, z), out in zip(segments, plan.outputs):
                    _same(_host(out), _alone(mix, p, t, z), ("replay", rep))
        finally:
            plan.close()
    finally:
        for ms in mixes.values():
            for m in ms:
                m.close()


@pytest.mark.gpu
def test_autotuned_group_size_keeps_every_record(gpu):
    """The launch-group autotuner: plans with 1, 2 and the maximum mixtures per launch group
    split the same segments into different launch groups and give identical records; the
    autotuner returns one of them with a timing per candidate."""
    import torch

    from mpmc_hnu_amd import flash

    mixes = [flash.Mixture.from_dict(W.mixture_linspaced(6, seed=400 + k)) for k in range(5)]
    try:
        segments = []
        for j in range(10):
            p, t, z = W.cells_uniform(6, 200 + 37 * j, seed=500 + j)
            segments.append((mixes[j % 5], torch.as_tensor(p, device="cuda"), torch.as_tensor(t, device="cuda"),
                             torch.as_tensor(z, device="cuda")))
        print("[case] autotuned plan over 10 segments of 5 mixtures", flush=True)
        plans = {g: flash.Plan(segments, group_mixtures=g) for g in (0, 2, 1)}
        try:
            # segments in the given order, a group closing when its table is full
            assert plans[0].groups == 1 and plans[2].groups == 5 and plans[1].groups == 10
            torch.cuda.synchronize()
            for k in range(len(segments)):
                ref = _host(plans[0].outputs[k])
                for g in (2, 1):
                    _same(_host(plans[g].outputs[k]), ref, ("group_mixtures", g, k))
        finally:
            for pl in plans.values():
                pl.close()
        best, tune = flash.autotune_plan(segments, candidates=(0, 2, 1), replays=1)
        try:
            assert set(tune) == {0, 2, 1} and best.group_mixtures in tune
            assert all(v > 0 for v in tune.values())
        finally:
            best.close()
    finally:
        for m in mixes:
            m.close()


@pytest.mark.gpu
def test_uneven_batch_plan_equals_one_mixture_at_a_time(gpu):
    """The uneven-batch bench leg at 1% of its cells: 256 mixtures of nc 2..12, 30% of the
    cells near a phase boundary; the plan's records equal the per-mixture flashes."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    import bench

    print("[case] uneven leg at scale 0.01", flush=True)
    r = bench.uneven_leg(steps=1, warmup=0, scale=0.01)
    assert r["mixtures"] == 256 and r["launch_groups"] >= 11
    assert r["near_boundary_cells"] > 0.25 * r["cells"]
    assert r["identical_records"]
    assert set(r["phase_counts"]) >= {1, 2}


@pytest.mark.gpu
def test_uneven_batch_against_the_oracle(gpu):
    """The uneven batch at 1% of its cells (256 mixtures of nc 2..12, 30% of the cells within
    2% in P of a phase boundary): every cell of every mixture, near-boundary cells included,
    flashed by one FP64 plan has the oracle's decisions and values (the flash's bar: 1e-9, or
    within the oracle's own stopping error at tolerance 1e-11), and
    the decision-exact mixed plan has the FP64 plan's decisions."""
    import sys
    from pathlib import Path

    import torch

    import oracle_bindings as ob
    import parity
    from mpmc_hnu_amd import flash

    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    import bench

    layout, mixes, cells, ncell, near = bench.uneven_workload(scale=0.01)
    print(f"[case] uneven batch vs oracle: {len(layout)} mixtures, {ncell} cells, {near} near a boundary",
          flush=True)
    try:
        dev = torch.device("cuda", 0)
        segments = [(mixes[k], torch.as_tensor(p, device=dev), torch.as_tensor(t, device=dev),
                     torch.as_tensor(z, device=dev)) for k, (p, t, z) in enumerate(cells)]
        plan = flash.Plan(segments)
        exact = flash.Plan(segments, precision=1)
        try:
            torch.cuda.synchronize()
            flips = stalled = 0
            for k, (m, (p, t, z)) in enumerate(zip(layout, cells)):
                g = _host(plan.outputs[k])
                o = ob.flash(m["mixture"], p, t, z, 1e-11, 3)
                rep, bad = parity.compare(g, o)
                mix = m["mixture"]
                loose, errs = parity.within_stopping_error(
                    g, o, np.nonzero(bad)[0],
                    lambda c, f: ob.flash(mix, p[c], t[c], np.ascontiguousarray(z[:, c]), 1e-11 * f, 3))
                if rep["decision_cells"] or rep["value_mismatch"]:
                    print(f"  mixture {k} nc={m['nc']} n={p.size}", rep, np.nonzero(bad)[0][:8].tolist(), errs[:8],
                          flush=True)
                flips += rep["decision_cells"] + loose.size
                e = _host(exact.outputs[k])
                dec = ((e["status"] & 0xFFF) != (g["status"] & 0xFFF)) | (e["phase_count"] != g["phase_count"])
                assert np.all((g["status"][dec] & 0xF) == 2), (k, np.nonzero(dec)[0][:8].tolist())
                stalled += int(dec.sum())
            print(f"  FP64 plan vs oracle: {flips} cells differ; exact mixed plan: {stalled} stalled-descent cells",
                  flush=True)
            assert flips == 0
            assert stalled <= 1e-4 * ncell
        finally:
            exact.close()
            plan.close()
    finally:
        for m in mixes:
            m.close()

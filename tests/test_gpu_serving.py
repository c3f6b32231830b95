"""The continuous-batching server (mpmc_server_*): results of every job in a concurrent
mix are identical to the same job flashed alone, and small jobs are not held behind
large ones."""
import numpy as np
import pytest

import oracle_bindings as ob
from mpmc_hnu_amd import workloads as W

KEYS = ("status", "phase_count", "beta", "zfactor", "branch", "x", "ln_phi")


def _alone(mix, p, t, z):
    from mpmc_hnu_amd import flash

    m = flash.Mixture.from_dict(mix)
    try:
        return flash.to_host(flash.flash_batch(m, p, t, z, max_phases=3))
    finally:
        m.close()


@pytest.mark.gpu
def test_server_results_alone_and_in_the_mix(gpu):
    from mpmc_hnu_amd import serving as S

    rng = np.random.Generator(np.random.PCG64(21))
    mixtures = {nc: W.mixture_linspaced(nc) for nc in range(2, 11)}
    jobs = []
    for k in range(24):
        nc = int(rng.integers(2, 11))
        n = int(np.exp(rng.uniform(0, np.log(150_000))))
        p, t, z = W.cells_uniform(nc, max(n, 1), seed=100 + k)
        jobs.append((nc, p, t, z))
    print(f"[case] server mix of {len(jobs)} jobs, {sum(j[1].size for j in jobs)} cells", flush=True)
    srv = S.Server(slots=2, chunk_cells=32768, small_job_cells=16384)
    try:
        handles = [srv.submit(mixtures[nc], p, t, z) for nc, p, t, z in jobs]
        results = [h.wait() for h in handles]
    finally:
        srv.close()
    for (nc, p, t, z), r in zip(jobs, results):
        a = _alone(mixtures[nc], p, t, z)
        for k in KEYS:
            assert np.array_equal(r[k], a[k], equal_nan=True), (nc, p.size, k)
    # the served records against the oracle: every cell of the jobs up to 20,000 cells
    import parity

    checked = 0
    for (nc, p, t, z), r in zip(jobs, results):
        if p.size > 20_000:
            continue
        o = ob.flash(mixtures[nc], p, t, z, 1e-11, 3)
        rep, bad = parity.compare(r, o)
        assert rep["decision_cells"] == 0 and rep["value_mismatch"] == 0, (nc, p.size, rep)
        checked += p.size
    print(f"  {checked} served cells match the oracle", flush=True)
    assert checked > 10_000


@pytest.mark.gpu
def test_small_jobs_overtake_a_large_one(gpu):
    """A 1M-cell job is in flight; small jobs submitted after it finish long before it."""
    from mpmc_hnu_amd import serving as S

    mix = W.mixture_linspaced(6)
    p, t, z = W.cells_c2()
    srv = S.Server(slots=2, chunk_cells=32768, small_job_cells=16384)
    print("[case] small jobs behind a 1M-cell job", flush=True)
    try:
        big = srv.submit(mix, p, t, z)
        small = [srv.submit(mix, p[k * 1000:(k + 1) * 1000], t[k * 1000:(k + 1) * 1000],
                            np.ascontiguousarray(z[:, k * 1000:(k + 1) * 1000])) for k in range(20)]
        for s in small:
            s.wait()
        big.wait()
    finally:
        srv.close()
    worst_small = max(s.latency_s for s in small)
    print(f"  big {big.latency_s * 1e3:.1f} ms, worst small {worst_small * 1e3:.1f} ms", flush=True)
    assert worst_small < 0.5 * big.latency_s


@pytest.mark.gpu
def test_warm_started_jobs_alone_and_in_the_mix(gpu):
    """Jobs with warm-start hints (the phase compositions of the same cells at a shifted
    pressure) and jobs without share launches; every job's records equal the hinted (or
    plain) flash of that job alone."""
    import torch

    from mpmc_hnu_amd import flash, serving as S

    mixtures = {nc: W.mixture_linspaced(nc) for nc in (4, 6)}
    jobs = []
    for k in range(12):
        nc = (4, 6)[k % 2]
        p, t, z = W.cells_uniform(nc, 700 + 311 * k, seed=600 + k)
        hints = W.warm_start_hints(ob.flash(mixtures[nc], p * 0.98, t, z, 1e-11, 3), z) if k % 3 else None
        jobs.append((nc, p, t, z, hints))
    print(f"[case] server mix of {len(jobs)} jobs, {sum(j[4] is not None for j in jobs)} warm-started", flush=True)
    srv = S.Server(slots=2, chunk_cells=8192, small_job_cells=4096)
    try:
        handles = [srv.submit(mixtures[nc], p, t, z, hints=h) for nc, p, t, z, h in jobs]
        results = [h.wait() for h in handles]
    finally:
        srv.close()
    for (nc, p, t, z, hints), r in zip(jobs, results):
        m = flash.Mixture.from_dict(mixtures[nc])
        try:
            a = flash.to_host(flash.flash_batch(m, p, t, z, hints=hints, counters=False))
        finally:
            m.close()
        torch.cuda.synchronize()
        for k in KEYS:
            assert np.array_equal(r[k], a[k], equal_nan=True), (nc, p.size, k, hints is not None)

"""GPU parity: the HIP flash against the oracle (and the reference build when present)
on identical seeded inputs.

Bar: status word, phase count and branch labels identical; fractions, compositions and
Z within 1e-9 relative; ln(phi) within 1e-9 absolute + relative; fugacity equality of
the published phases within 1e-10.

The device libm (log, exp, log1p, expm1) and glibc can differ in the last ulp, so the
successive-substitution iterates of the two sides are not bit-identical; both stop
inside the same fugacity tolerance (1e-11), which bounds the difference of converged
fractions by tolerance x conditioning. Decisions must nevertheless agree on every cell
(a mismatch is printed with its cell index), every value within 1e-9, and evaluation
counters (which can differ by an iteration) agree in total within 1%.

Mixed precision: precision 1 (decision-exact) runs every decision-bearing search in FP64
from its start, so its statuses, phase counts and labels must equal precision 0's on every
cell, except cells where precision 0's own search ends indeterminate because an FP64 descent
stalls where the FP32 screen reaches a stationary point (a descent passing a saddle of the
tangent-plane distance or converging too slowly for the 512-step limit); such cells are
listed and counted. Precision 2 (fast) is held to values within 1e-7 and a decision
share, since its FP32 split attempts converge where the reference's line search fails.
"""
import numpy as np
import pytest

import oracle_bindings as ob
import parity
from mpmc_hnu_amd import workloads as W

# share of cells whose decision may differ from FP64 in the fast mixed schedule (precision 2)
FAST_MAX_FLIP = 2e-3


def _run_gpu(mix, p, t, z, tol, maxph):
    from mpmc_hnu_amd import flash

    m = flash.Mixture.from_dict(mix)
    try:
        res = flash.to_host(flash.flash_batch(m, p, t, z, fugacity_tolerance=tol, max_phases=maxph))
    finally:
        m.close()
    return res


def _values_within_stopping_error(name, g, o, bad, mix, p, t, z, tol, maxph):
    """Every value within 1e-9 relative of the oracle's, or (cells listed) within the
    oracle's own stopping error at this fugacity tolerance (parity.within_stopping_error)."""
    cells = np.nonzero(bad)[0]
    if cells.size == 0:
        return
    bad_cells, errs = parity.within_stopping_error(
        g, o, cells, lambda c, f: ob.flash(mix, p[c], t[c], np.ascontiguousarray(z[:, c]), tol * f, maxph))
    print(f"  {name}: values beyond 1e-9 at cells {cells[:16].tolist()}, (device, oracle) error against the "
          f"oracle at a smaller tolerance: {[(f'{a:.1e}', f'{b:.1e}') for a, b in errs[:16]]}", flush=True)
    assert bad_cells.size == 0, bad_cells.tolist()


def _check(name, mix, p, t, z, tol, maxph):
    print(f"[case] {name} nc={z.shape[0]} cells={p.size} max_phases={maxph}", flush=True)
    g = _run_gpu(mix, p, t, z, tol, maxph)
    o = ob.flash(mix, p, t, z, tol, maxph)
    rep, bad = parity.compare(g, o)
    print(name, rep, flush=True)
    nbad_dec = rep["status_mismatch"] + rep["count_mismatch"] + rep["branch_mismatch"]
    if nbad_dec:
        idx = np.nonzero(bad)[0][:8]
        print("  decision mismatches at cells", idx.tolist(), "gpu", g["status"][idx].tolist(), "oracle",
              o["status"][idx].tolist(), flush=True)
    assert nbad_dec == 0, rep
    # search flags of the status word (MPMC_FLAG_SPLIT_*, MPMC_FLAG_MAX3_ATTEMPTED)
    fdiff = np.nonzero(g["status"] != o["status"])[0]
    print(f"  {name}: status words (flags included) differ at {fdiff.size} cells {fdiff[:8].tolist()}", flush=True)
    assert fdiff.size == 0
    _values_within_stopping_error(name, g, o, bad, mix, p, t, z, tol, maxph)
    res = parity.fugacity_residual(g)
    acc = (g["status"] & 0xF) == 0
    assert np.all(res[acc] <= 1e-10), float(res[acc].max())
    # evaluation counts: same totals within 1%
    same = ~bad
    for k in range(2):
        a, b = g["counters"][k, same].sum(dtype=np.float64), o["counters"][k, same].sum(dtype=np.float64)
        assert abs(a - b) <= 0.01 * b, (k, a, b)
    return g, o


@pytest.mark.gpu
def test_c1_vle_4096(gpu):
    mix = W.mixture_c1()
    p, t, z = W.cells_c1()
    _check("C1", mix, p, t, z, 1e-12, 2)


@pytest.mark.gpu
def test_c2_max3_prefix(gpu):
    mix = W.mixture_linspaced(6)
    p, t, z = W.cells_c2(ncell=131072)
    _check("C2", mix, p, t, z, 1e-11, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [1, 2, 3, 4, 5, 7, 8, 9, 10, 11, 12])
def test_uniform_nc(gpu, nc):
    mix = W.mixture_linspaced(nc)
    p, t, z = W.cells_uniform(nc, 1024, seed=100 + nc)
    _check(f"U{nc}", mix, p, t, z, 1e-11, 3)


@pytest.mark.gpu
def test_three_phase_sour_gas_sweep(gpu):
    mix, p, t, z = W.sour_gas_sweep(2048)
    g, o = _check("sour", mix, p, t, z, 1e-11, 3)
    assert np.sum((g["status"] >> 4 & 0xF) == 2) > 100  # accepted three-phase cells present


@pytest.mark.gpu
def test_phase_batch_roots(gpu):
    from mpmc_hnu_amd import flash

    mix = W.mixture_linspaced(6)
    p, t, z = W.cells_uniform(6, 4096, seed=7)
    m = flash.Mixture.from_dict(mix)
    g = flash.to_host(flash.phase_batch(m, p, t, z))
    m.close()
    o = ob.phase(mix, p, t, z)
    assert np.array_equal(g["root_status"], o["root_status"])
    assert np.array_equal(g["root_count"], o["root_count"])
    assert np.array_equal(g["root_flags"], o["root_flags"])
    fin = np.isfinite(o["zroot"])
    assert np.array_equal(fin, np.isfinite(g["zroot"]))
    # bisection on identical coefficients: roots bit-identical
    assert np.array_equal(g["zroot"][fin], o["zroot"][fin])
    fl = np.isfinite(o["ln_phi"])
    assert np.allclose(g["ln_phi"][fl], o["ln_phi"][fl], rtol=1e-13, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("maxph", [2, 3])
def test_staged_and_serial_schedules_agree(gpu, maxph):
    """The staged batch (lane per trial / attempt) and the serial one-lane-per-cell path
    run the same searches: identical decisions and values."""
    from mpmc_hnu_amd import flash

    mix = W.mixture_linspaced(5)
    p, t, z = W.cells_uniform(5, 2048, seed=77)
    m = flash.Mixture.from_dict(mix)
    a = flash.to_host(flash.flash_batch(m, p, t, z, max_phases=maxph, schedule=0))
    b = flash.to_host(flash.flash_batch(m, p, t, z, max_phases=maxph, schedule=1))
    m.close()
    rep, bad = parity.compare(a, b, rel=0.0, abs_lnphi=0.0)
    print("staged vs serial", rep, flush=True)
    assert not bad.any(), rep
    assert np.array_equal(a["counters"], b["counters"])


@pytest.mark.gpu
def test_caller_workspace(gpu):
    import torch

    from mpmc_hnu_amd import flash

    mix = W.mixture_linspaced(4)
    p, t, z = W.cells_uniform(4, 1000, seed=5)
    m = flash.Mixture.from_dict(mix)
    ws = torch.empty(flash.workspace_bytes(4, 1000), dtype=torch.uint8, device=gpu)
    a = flash.to_host(flash.flash_batch(m, p, t, z, workspace=ws))
    b = flash.to_host(flash.flash_batch(m, p, t, z))
    m.close()
    rep, bad = parity.compare(a, b, rel=0.0, abs_lnphi=0.0)
    assert not bad.any(), rep


def _record_bits(out):
    import torch

    return {k: (v.view(torch.int64) if v.dtype == torch.float64 else v).clone()
            for k, v in out.items() if isinstance(v, torch.Tensor)}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_record_is_independent_of_buffer_history(gpu, precision):
    """A cell's record is a function of its inputs alone: the same batch on zeroed
    workspace and output buffers, on the buffers left by another batch, and again on
    zeroed buffers gives bit-identical records (every word, NaN payloads included), so no
    stage reads workspace or output memory it did not write in the same call."""
    import torch

    from mpmc_hnu_amd import flash

    mix, pb, tb, zb = W.sour_gas_sweep(2048, seed=3)
    _, pa, ta, za = W.sour_gas_sweep(2048, seed=8)
    n = pb.size
    print(f"[case] buffer history, sour-gas sweep n={n}, precision {precision}", flush=True)
    m = flash.Mixture.from_dict(mix)
    try:
        ws = torch.zeros(flash.workspace_bytes(m.nc, n), dtype=torch.uint8, device=gpu)
        out = flash.allocate_outputs(m.nc, n, gpu)

        def fresh():
            ws.zero_()
            for v in out.values():
                if isinstance(v, torch.Tensor):
                    v.zero_()

        fresh()
        flash.flash_batch(m, pb, tb, zb, precision=precision, out=out, workspace=ws)
        ref = _record_bits(out)
        flash.flash_batch(m, pa, ta, za, precision=precision, out=out, workspace=ws)
        flash.flash_batch(m, pb, tb, zb, precision=precision, out=out, workspace=ws)
        after = _record_bits(out)
        fresh()
        flash.flash_batch(m, pb, tb, zb, precision=precision, out=out, workspace=ws)
        again = _record_bits(out)
        torch.cuda.synchronize()
    finally:
        m.close()
    for k in ref:
        assert torch.equal(ref[k], after[k]), f"{k} depends on the previous batch"
        assert torch.equal(ref[k], again[k]), f"{k} differs between identical calls"
    three = ((ref["status"] & 0xF) == 0) & (ref["phase_count"] == 3)
    assert int(three.sum()) > 100


def _mixed_vs_fp64(mix, p, t, z, maxph=3, precision=1):
    from mpmc_hnu_amd import flash

    m = flash.Mixture.from_dict(mix)
    a = flash.to_host(flash.flash_batch(m, p, t, z, max_phases=maxph, precision=precision))
    b = flash.to_host(flash.flash_batch(m, p, t, z, max_phases=maxph, precision=0))
    m.close()
    return a, b


def _exact_bar(name, a, b, max_stalled=0):
    """Precision 1 against precision 0: identical decisions, except cells that precision 0
    leaves indeterminate (status set 2) because one of its FP64 descents stalls; those are
    listed and at most ``max_stalled``. Cells with identical decisions have identical
    values whenever the same searches ran, and within 1e-9 otherwise."""
    rep, bad = parity.compare(a, b, rel=1e-9, abs_lnphi=1e-9)
    dec = ((a["status"] & 0xFFF) != (b["status"] & 0xFFF)) | (a["phase_count"] != b["phase_count"])
    idx = np.nonzero(dec)[0]
    print(f"{name} exact mixed vs fp64", rep, "decision cells", idx[:16].tolist(),
          [hex(v) for v in b["status"][idx[:16]]], [hex(v) for v in a["status"][idx[:16]]], flush=True)
    assert np.all((b["status"][idx] & 0xF) == 2), "a decision differs on a cell FP64 decides"
    assert idx.size <= max_stalled, idx.size
    assert rep["value_mismatch"] == 0, rep


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [3, 6, 9, 12])
def test_mixed_precision_matches_fp64(gpu, nc):
    """Decision-exact mixed precision (precision 1): the same statuses, phase counts and
    labels as FP64 on every cell, values within 1e-9 of FP64 and within 1e-8 of the FP64
    oracle, fugacity equality 1e-10."""
    mix = W.mixture_linspaced(nc)
    p, t, z = W.cells_uniform(nc, 16384, seed=300 + nc)
    a, b = _mixed_vs_fp64(mix, p, t, z)
    _exact_bar(f"U{nc}", a, b)
    o = ob.flash(mix, p, t, z, 1e-11, 3)
    rep, bad = parity.compare(a, o, rel=1e-8, abs_lnphi=1e-8)
    print(f"exact mixed nc={nc} vs oracle", rep, flush=True)
    assert rep["decision_cells"] == 0 and rep["value_mismatch"] == 0, rep
    res = parity.fugacity_residual(a)
    acc = (a["status"] & 0xF) == 0
    assert np.all(res[acc] <= 1e-10), float(res[acc].max())


@pytest.mark.gpu
def test_mixed_precision_c2_grid(gpu):
    """Decision-exact mixed precision on the first 131,072 cells of the nc=6 grid workload:
    no decision differs from FP64 or from the FP64 oracle."""
    mix = W.mixture_linspaced(6)
    p, t, z = W.cells_c2(ncell=131072)
    a, b = _mixed_vs_fp64(mix, p, t, z)
    _exact_bar("C2", a, b)
    o = ob.flash(mix, p, t, z, 1e-11, 3)
    rep, bad = parity.compare(a, o, rel=1e-8, abs_lnphi=1e-8)
    print("exact mixed C2 vs oracle", rep, flush=True)
    assert rep["decision_cells"] == 0 and rep["value_mismatch"] == 0, rep


@pytest.mark.gpu
def test_mixed_precision_three_phase_sweep(gpu):
    """The sour-gas sweep crosses the three-phase region, where FP64 descents pass saddles of
    the tangent-plane distance: precision 1 may differ only on cells FP64 leaves
    indeterminate (at most 1 of 2,048 here, listed)."""
    mix, p, t, z = W.sour_gas_sweep(2048)
    a, b = _mixed_vs_fp64(mix, p, t, z)
    _exact_bar("sour", a, b, max_stalled=2)
    assert np.sum((a["status"] >> 4 & 0xF) == 2) > 100


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [3, 6])
def test_fast_mixed_precision(gpu, nc):
    """Fast mixed precision (precision 2): values within 1e-7 of FP64 on cells with the same
    decision; decision differences (cells where the reference's split search ends on a
    line-search failure that FP32 attempts do not meet) are printed and bounded."""
    mix = W.mixture_linspaced(nc)
    p, t, z = W.cells_uniform(nc, 16384, seed=300 + nc)
    a, b = _mixed_vs_fp64(mix, p, t, z, precision=2)
    rep, bad = parity.compare(a, b, rel=1e-7, abs_lnphi=1e-7)
    print(f"fast mixed nc={nc}", rep, flush=True)
    assert rep["decision_cells"] <= FAST_MAX_FLIP * p.size, rep
    assert rep["value_mismatch"] == 0, rep
    res = parity.fugacity_residual(a)
    acc = (a["status"] & 0xF) == 0
    assert np.all(res[acc] <= 1e-10), float(res[acc].max())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_warm_start_hints(gpu, precision):
    """mpmc_flash_batch_hinted: the phase compositions of a flash at 3% lower pressure as
    warm-start hints. FP64 matches the oracle's hinted flash at the usual bar (hints add
    stability trials, counted identically); mixed precision at the mixed bar."""
    from mpmc_hnu_amd import flash

    nc = 6
    mix = W.mixture_linspaced(nc)
    p, t, z = W.cells_uniform(nc, 2048, seed=77)
    prev = ob.flash(mix, p * 0.97, t, z, 1e-11, 3)
    hints = W.warm_start_hints(prev, z)
    print(f"[case] hinted flash precision={precision}, {int(np.isfinite(hints[0, 0]).sum())} hinted cells", flush=True)
    m = flash.Mixture.from_dict(mix)
    try:
        g = flash.to_host(flash.flash_batch(m, p, t, z, precision=precision, hints=hints))
    finally:
        m.close()
    o = ob.flash(mix, p, t, z, 1e-11, 3, hints=hints)
    rel = 1e-7 if precision == 2 else 1e-9
    rep, bad = parity.compare(g, o, rel=rel, abs_lnphi=rel)
    print("hinted", rep, flush=True)
    assert rep["decision_cells"] <= (FAST_MAX_FLIP * p.size if precision == 2 else 0), rep
    assert rep["value_mismatch"] == 0, rep
    res = parity.fugacity_residual(g)
    acc = (g["status"] & 0xF) == 0
    assert np.all(res[acc] <= 1e-10), float(res[acc].max())
    if precision == 0:
        same = ~bad
        for k in range(2):
            a, b = g["counters"][k, same].sum(dtype=np.float64), o["counters"][k, same].sum(dtype=np.float64)
            assert abs(a - b) <= 0.01 * b, (k, a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,max_phases", [(0, 3), (0, 2), (1, 3), (2, 3)])
def test_final_review_hints(gpu, precision, max_phases):
    """mpmc_flash_batch_warm: initial and final-review hints (the phase compositions of a
    flash at 3% lower pressure), a final hint with the wrong support rejecting its cell.
    FP64 matches the oracle's warm flash at the usual bar, with the review's extra trials
    counted identically; mixed precision at the mixed bar."""
    from mpmc_hnu_amd import flash

    nc = 6
    mix = W.mixture_linspaced(nc)
    p, t, z = W.cells_uniform(nc, 2048, seed=79)
    prev = ob.flash(mix, p * 0.97, t, z, 1e-11, max_phases)
    fh = W.warm_start_hints(prev, z)
    ih = W.warm_start_hints(prev, z, nhint=1)
    bad_cell = int(np.nonzero(np.isfinite(fh[1, 0]))[0][0])
    fh[1, :, bad_cell] = 0.0
    fh[1, 0, bad_cell] = 1.0
    print(f"[case] warm flash precision={precision} max_phases={max_phases}, "
          f"{int(np.isfinite(fh[0, 0]).sum())} cells with final hints", flush=True)
    m = flash.Mixture.from_dict(mix)
    try:
        g = flash.to_host(flash.flash_batch(m, p, t, z, max_phases=max_phases, precision=precision, hints=ih,
                                            final_hints=fh))
    finally:
        m.close()
    o = ob.flash(mix, p, t, z, 1e-11, max_phases, hints=ih, final_hints=fh)
    assert g["status"][bad_cell] & 0x2000 and o["status"][bad_cell] & 0x2000
    rel = 1e-7 if precision == 2 else 1e-9
    rep, bad = parity.compare(g, o, rel=rel, abs_lnphi=rel)
    print("warm", rep, flush=True)
    assert rep["decision_cells"] <= (FAST_MAX_FLIP * p.size if precision == 2 else 0), rep
    assert rep["value_mismatch"] == 0, rep
    res = parity.fugacity_residual(g)
    acc = (g["status"] & 0xF) == 0
    assert np.all(res[acc] <= 1e-10), float(res[acc].max())
    if precision == 0:
        same = ~bad
        for k in range(2):
            a, b = g["counters"][k, same].sum(dtype=np.float64), o["counters"][k, same].sum(dtype=np.float64)
            assert abs(a - b) <= 0.01 * b, (k, a, b)

"""GPU parity of the post-SNES phase-transition scan (mpmc_transition_scan,
mpmc_transition_scan_2p) with the oracle's restatement of the reference's scanner
(oracle/transition_scan.py).

Bar: the scan rules applied to the device's own flash record give bit-identical codes,
target counts and target densities (the decision kernel is integer logic plus one IEEE
quotient); the flash record itself meets the flash's parity bar against the checker
flash; the two-phase feeds are within one ulp of the long double normalisation of the
device inventory, and the inventory within the assembly's 1e-8 of the oracle's. Density
mode 1 takes a single target phase's Z from the oracle's phase routine at the feed (bit-
identical to the device's, as the phase batch test shows) at the record's branch."""
import numpy as np
import pytest

import oracle_bindings as ob
import parity
from mpmc_hnu_amd import workloads as W
from oracle import transition_scan as S



def _host(t):
    return t.cpu().numpy()


def _check_rules(g, p, t, cnt, code0, single_z=None):
    from mpmc_hnu_amd import flash

    rec = flash.to_host(g["flash"])
    code, target, dens = S.decide(cnt, p, t, rec["status"], rec["phase_count"], rec["zfactor"], code0,
                                  single_z=single_z(rec) if single_z else None)
    assert np.array_equal(_host(g["code"]), code)
    assert np.array_equal(_host(g["target_count"]), target)
    gd = _host(g["target_density"])
    assert np.array_equal(np.isnan(gd), np.isnan(dens))
    fin = ~np.isnan(dens)
    assert np.array_equal(gd[fin], dens[fin])
    st, first, err, props = S.summary(code)
    assert (g["status"], g["first_stop"], g["rc"] == 1, g["proposals"]) == (st, first, err, props)
    return rec, code


def _check_record(rec, o, n, tight):
    """The flash's own bar (test_gpu_flash.py): every decision identical, every value within
    1e-9 relative or within the oracle's own stopping error (``tight(cells, factor)``: the
    oracle at a smaller fugacity tolerance on those cells)."""
    rep, bad = parity.compare(rec, o)
    flips = rep["status_mismatch"] + rep["count_mismatch"] + rep["branch_mismatch"]
    print("  flash record vs checker:", rep, flush=True)
    assert flips == 0, rep
    cells = np.nonzero(bad)[0]
    bad_cells, errs = parity.within_stopping_error(rec, o, cells, tight)
    if cells.size:
        print("  values beyond 1e-9 at", cells[:16].tolist(), errs[:16], flush=True)
    assert bad_cells.size == 0, bad_cells.tolist()


@pytest.mark.gpu
def test_scan_of_given_sources_matches_oracle(gpu):
    from mpmc_hnu_amd import flash, scan

    mix, p, t, z = W.sour_gas_sweep(1024)
    rng = np.random.Generator(np.random.PCG64(21))
    cnt = rng.integers(1, 4, p.size).astype(np.int32)
    z, p = z.copy(), p.copy()
    # malformed sources (the reference's invalid_argument): a negative fraction, zero
    # pressure, phase counts 0 and 4, a composition whose sum is off one
    z[0, 300] = -z[0, 300]
    p[317] = 0.0
    cnt[340], cnt[341] = 0, 4
    z[1, 360] *= 1.01
    # a sum 1000 eps off one passes source validation (4096 nc eps) but not the flash's
    # 64 eps check: the flash's domain_error, caught by the scanner as indeterminate
    z[0, 380] += 1000 * np.finfo(float).eps
    print("[case] scan of 1024 sour-gas sources, nc=6", flush=True)
    m = flash.Mixture.from_dict(mix)
    try:
        g = scan.transition_scan(m, p, t, z, cnt)
    finally:
        m.close()
    code0 = np.where(S.validate_source(p, t, cnt, z), S.SAME, S.ERROR).astype(np.int32)
    assert np.sum(code0 == S.ERROR) == 5
    rec, code = _check_rules(g, p, t, cnt, code0)
    assert np.sum(code == S.PROPOSAL) > 100 and np.sum(code == S.SAME) > 100
    assert code[380] == S.INDETERMINATE and rec["status"][380] & S.FLAG_REJECTED
    target = _host(g["target_count"])
    # 2 -> 3, the reference's regression case, is a proposal; a lower target without the
    # backend's disappearance evidence (3 -> 2 from a base split accepted directly) is an error
    assert np.any((code == S.PROPOSAL) & (cnt == 2) & (target == 3))
    direct = ((rec["status"] >> 8) & 0xF) == 1
    assert np.any((code == S.ERROR) & (cnt == 3) & (target == 2) & direct)
    assert not np.any((code == S.PROPOSAL) & (target < cnt) & direct)
    _check_record(rec, ob.flash(mix, p, t, z, 1e-11, 3), p.size,
                  lambda c, f: ob.flash(mix, p[c], t[c], np.ascontiguousarray(z[:, c]), 1e-11 * f, 3))


@pytest.mark.gpu
def test_scan_without_stops_is_complete(gpu):
    from mpmc_hnu_amd import flash, scan

    mix = W.mixture_linspaced(6)
    p, t, z = W.cells_uniform(6, 512, seed=33)
    cnt = np.full(p.size, 2, np.int32)
    print("[case] scan of 512 uniform sources, nc=6, two-phase backend", flush=True)
    m = flash.Mixture.from_dict(mix)
    try:
        g = scan.transition_scan(m, p, t, z, cnt, max_phases=2)
    finally:
        m.close()
    rec, code = _check_rules(g, p, t, cnt, np.zeros(p.size, np.int32))
    o = ob.flash(mix, p, t, z, 1e-11, 2)
    _check_record(rec, o, p.size, lambda c, f: ob.flash(mix, p[c], t[c], np.ascontiguousarray(z[:, c]), 1e-11 * f, 2))
    if not np.any((code == S.INDETERMINATE) | (code == S.ERROR)):
        assert g["status"] == S.COMPLETE and g["first_stop"] == -1 and g["rc"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("nc,dims", [(3, (4, 4, 3)), (6, (5, 4, 3))])
def test_scan_of_a_two_phase_state_matches_oracle(gpu, nc, dims):
    from mpmc_hnu_amd import assembly as A, flash, scan

    mix = W.mixture_linspaced(nc)
    grid = W.tpfa_grid(*dims, seed=nc)
    q, br = W.two_phase_state(nc, grid["ncell"], ob.phase_roots(mix), seed=20 + nc, dims=dims)
    model = A.default_model(nc)
    print(f"[case] two-phase state scan nc={nc} grid={dims}", flush=True)
    m = flash.Mixture.from_dict(mix)
    gr = A.Grid(grid)
    try:
        g = scan.transition_scan_2p(m, gr, model, q, br)
    finally:
        gr.close()
        m.close()
    inv, feed = _host(g["inventory"]), _host(g["feed"])
    oinv = ob.inventory(mix, A.model_array(model, nc), q, br, grid["poro"])
    assert np.all(np.abs(inv - oinv) <= 1e-8 * np.abs(oinv))
    ref_feed, ok = S.source_feed(inv)
    assert np.all(ok)
    assert np.all(np.abs(feed - ref_feed) <= np.spacing(ref_feed))
    p, t = q[0], q[1]
    code0 = np.where(S.validate_source(p, t, np.full(p.size, 2), feed), S.SAME, S.INDETERMINATE).astype(np.int32)
    rec, code = _check_rules(g, p, t, 2, code0)
    _check_record(rec, ob.flash(mix, p, t, feed, 1e-11, 3), p.size,
                  lambda c, f: ob.flash(mix, p[c], t[c], np.ascontiguousarray(feed[:, c]), 1e-11 * f, 3))


@pytest.mark.gpu
def test_mixed_precision_scan_applies_the_same_rules(gpu):
    """precision 1: the scan's decisions are the rules applied to the mixed flash record
    (whose own agreement with FP64 is tested with the flash)."""
    from mpmc_hnu_amd import flash, scan

    mix, p, t, z = W.sour_gas_sweep(1024, seed=12)
    cnt = np.full(p.size, 2, np.int32)
    print("[case] mixed-precision scan of 1024 sour-gas sources", flush=True)
    m = flash.Mixture.from_dict(mix)
    try:
        g = scan.transition_scan(m, p, t, z, cnt, precision=1)
    finally:
        m.close()
    _, code = _check_rules(g, p, t, cnt, np.zeros(p.size, np.int32))
    assert np.any(code == S.PROPOSAL)


def _provider_z(mix, p, t, z):
    """Z of the provider root (the branch word of phase 0) of each cell's feed, from the
    oracle's phase routine."""
    roots = ob.phase(mix, p, t, z)["zroot"]

    def pick(rec):
        idx = rec["branch"][0] & 0xFF
        ok = (rec["branch"][0] >= 0) & (idx < 3)
        return np.where(ok, roots[np.minimum(idx, 2), np.arange(p.size)], np.nan)
    return pick


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_single_phase_density_mode(gpu, precision):
    """density_mode 1: a 2 -> 1 candidate gets p / (Z R T) of the feed's provider root and
    then fails the evidence rule (no disappearance evidence for a single-phase target): an
    error where density mode 0 leaves it indeterminate; every other decision is the one of
    density mode 0."""
    from mpmc_hnu_amd import flash, scan

    mix, p, t, z = W.sour_gas_sweep(1024, seed=4)
    cnt = np.full(p.size, 2, np.int32)
    print(f"[case] density mode 1 scan of 1024 sour-gas sources, precision {precision}", flush=True)
    m = flash.Mixture.from_dict(mix)
    try:
        g1 = scan.transition_scan(m, p, t, z, cnt, precision=precision, density_mode=1)
        g0 = scan.transition_scan(m, p, t, z, cnt, precision=precision)
    finally:
        m.close()
    zero = np.zeros(p.size, np.int32)
    _, code1 = _check_rules(g1, p, t, cnt, zero, single_z=_provider_z(mix, p, t, z))
    _, code0 = _check_rules(g0, p, t, cnt, zero)
    one = _host(g1["target_count"]) == 1
    assert np.sum((code1 == S.ERROR) & one) > 10 and not np.any((code1 == S.PROPOSAL) & one)
    if precision == 0:
        # the FP64 flash is deterministic: the two calls share their record
        assert np.array_equal(code0[~one], code1[~one])
        assert np.all(code0[one] == S.INDETERMINATE)


@pytest.mark.gpu
def test_density_mode_outside_0_1_is_an_argument_error(gpu):
    """An unknown density_mode is rejected before any launch, and the summary is left as
    the caller initialised it."""
    from mpmc_hnu_amd import flash, native, scan

    mix = W.mixture_linspaced(3)
    p, t, z = W.cells_uniform(3, 64, seed=2)
    print("[case] density_mode 2 rejected, 64 cells nc=3", flush=True)
    m = flash.Mixture.from_dict(mix)
    try:
        with pytest.raises(native.MpmcError, match="density_mode"):
            scan.transition_scan(m, p, t, z, np.full(p.size, 2, np.int32), density_mode=2)
        g = scan.transition_scan(m, p, t, z, np.full(p.size, 2, np.int32), density_mode=1)
    finally:
        m.close()
    _check_rules(g, p, t, 2, np.zeros(p.size, np.int32), single_z=_provider_z(mix, p, t, z))

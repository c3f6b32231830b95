"""The oracle of the phase-transition scan (oracle/transition_scan.py), pinned without the
GPU: the three cases of the reference's scanner regression
(tests/flow_discretization/petsc/post_snes_pt_flash_phase_transition_scanner_test.cpp:139-215)
on synthetic flash records and on the PR76 checker flash, the ordered stop rule of the cell
loop, source validation at its tolerance edge, and the long double feed normalisation
against exact rational arithmetic."""
from fractions import Fraction

import numpy as np

import oracle_bindings as ob
from mpmc_hnu_amd import workloads as W
from oracle import transition_scan as S

R = 8.31446261815324


def _accepted(count):
    # status word: set accepted (0) | max3 status (count - 1) << 4
    return (count - 1) << 4


def test_reference_regression_cases_on_synthetic_records():
    p = np.array([1.0e7, 1.0e7, 1.0e7])
    t = np.array([350.0, 350.0, 350.0])
    status = np.array([_accepted(3), _accepted(3), 2 | (5 << 4)])
    count = np.array([3, 3, 0])
    zf = np.array([[0.8, 0.8, np.nan], [0.31, 0.31, np.nan], [0.052, 0.052, np.nan]])
    code, target, dens = S.decide([2, 3, 2], p, t, status, count, zf)
    # 2 -> 3: a proposal with positive target densities p / (Z R T)
    assert code[0] == S.PROPOSAL and target[0] == 3
    for ph in range(3):
        assert dens[ph, 0] > 0
        assert abs(dens[ph, 0] * zf[ph, 0] * R * t[0] / p[0] - 1.0) <= 4 * np.finfo(float).eps
    # same cardinality: no proposal; indeterminate flash: indeterminate, no proposal
    assert code[1] == S.SAME and target[1] == 3 and np.all(np.isnan(dens[:, 1]))
    assert code[2] == S.INDETERMINATE and target[2] == 0 and np.all(np.isnan(dens[:, 2]))
    # the indeterminate cell ends the scan and clears the proposals collected before it
    assert S.summary(code) == (S.STATUS_INDETERMINATE, 2, False, 0)


def test_target_density_that_is_not_positive_is_indeterminate():
    code, target, dens = S.decide([1], [1.0e6], [300.0], [_accepted(2)], [2], [[0.9], [-1.0], [np.nan]])
    assert code[0] == S.INDETERMINATE and target[0] == 2 and np.all(np.isnan(dens[:, 0]))


def test_rejected_feed_is_indeterminate():
    """A validated source whose feed the flash rejects: the flash's std::domain_error is
    caught by the scanner's catch-all as indeterminate (:292-299)."""
    code, target, _ = S.decide([2], [1.0e6], [300.0], [2 | S.FLAG_REJECTED], [0], np.full((3, 1), np.nan))
    assert code[0] == S.INDETERMINATE and target[0] == 0


def test_lower_targets_need_disappearance_evidence():
    """make_phase_set_transition_candidate_from_flash (phase_set_transition_flash_adapter.hpp:
    171-183) raises invalid_argument for a lower target without accepted phase-disappearance
    evidence; the PR76 max3 report carries it only for 3 -> 2 through the tail's fresh
    two-phase neighbour (max3 two-phase after an unstable base split, pr76_pt_flash_backend.hpp:
    138-146). Higher targets always carry source-challenge and accepted-target evidence."""
    p, t = np.full(4, 1.0e7), np.full(4, 350.0)
    zf = np.array([[0.8] * 4, [0.3] * 4, [np.nan] * 4])
    direct = 1 << 4 | 1 << 8            # accepted, max3 two-phase, base split two-phase
    neighbour = 1 << 4 | 2 << 8         # accepted, max3 two-phase, base split unstable
    code, target, dens = S.decide([3, 3, 1, 1], p, t, [direct, neighbour, direct, neighbour], [2] * 4, zf)
    assert code.tolist() == [S.ERROR, S.PROPOSAL, S.PROPOSAL, S.PROPOSAL]
    assert np.all(np.isnan(dens[:, 0])) and np.all(dens[:2, 1] > 0)
    assert S.summary(code) == (S.COMPLETE, 0, True, 0)


def test_first_stopping_cell_decides_the_scan():
    assert S.summary([0, 1, 0]) == (S.COMPLETE, -1, False, 1)
    # an error returns with the proposals before it; an indeterminate cell clears them
    assert S.summary([0, 1, 3, 2]) == (S.COMPLETE, 2, True, 1)
    assert S.summary([1, 1, 3, 1]) == (S.COMPLETE, 2, True, 2)
    assert S.summary([1, 2, 3, 1]) == (S.STATUS_INDETERMINATE, 1, False, 0)


def test_source_validation():
    nc, n = 3, 8
    z = np.full((nc, n), 1.0 / 3.0)
    p = np.full(n, 1e6)
    t = np.full(n, 300.0)
    cnt = np.full(n, 2)
    tol = 4096 * np.finfo(float).eps * nc
    z[0, 1] = -1e-3            # a non-positive fraction
    p[2] = 0.0                 # pressure
    t[3] = np.inf              # temperature
    cnt[4] = 0                 # phase count outside [1, 3]
    cnt[5] = 4
    z[0, 6] += 2.0 * tol       # sum off by twice the tolerance
    z[0, 7] += 0.25 * tol      # within the tolerance
    ok = S.validate_source(p, t, cnt, z)
    assert ok.tolist() == [True, False, False, False, False, False, False, True]
    # a three-phase source against a two-phase backend is unsupported
    assert not S.validate_source(p[:1], t[:1], [3], z[:, :1], max_phases=2)[0]
    # one component: malformed (component_ids.size() >= 2)
    assert not S.validate_source([1e6], [300.0], [1], np.ones((1, 1)))[0]


def test_feed_normalisation_against_exact_quotients():
    rng = np.random.Generator(np.random.PCG64(5))
    nc, n = 6, 400
    inv = np.vstack([rng.lognormal(3.0, 2.0, (nc, n)), rng.uniform(1e6, 1e8, (1, n))])
    inv[2, 7] = 0.0            # a vanished component cannot define a PT feed
    inv[4, 9] = np.nan
    feed, ok = S.source_feed(inv)
    assert not ok[7] and not ok[9] and np.all(np.isnan(feed[:, [7, 9]]))
    exact_hits = 0
    for c in np.nonzero(ok)[0]:
        total = sum(Fraction(float(inv[i, c])) for i in range(nc))
        for i in range(nc):
            q = float(Fraction(float(inv[i, c])) / total)   # correctly rounded quotient
            d = abs(feed[i, c] - q)
            assert d <= np.spacing(q), (c, i)
            exact_hits += d == 0
    # the long double path rounds twice: rarely one ulp off the correctly rounded quotient
    assert exact_hits >= 0.99 * nc * (n - 2)


def test_pr76_two_to_three_proposal_with_the_checker_flash():
    mix, p, t, z = W.sour_gas_sweep(64)
    o = ob.flash(mix, p, t, z, 1e-11, 3)
    three = np.nonzero(((o["status"] & 0xF) == 0) & (o["phase_count"] == 3))[0]
    assert three.size > 0
    c = int(three[0])
    code, target, dens = S.decide([2], p[c:c + 1], t[c:c + 1], o["status"][c:c + 1], o["phase_count"][c:c + 1],
                                  o["zfactor"][:, c:c + 1])
    assert code[0] == S.PROPOSAL and target[0] == 3 and np.all(dens[:, 0] > 0)
    code, _, _ = S.decide([3], p[c:c + 1], t[c:c + 1], o["status"][c:c + 1], o["phase_count"][c:c + 1],
                          o["zfactor"][:, c:c + 1])
    assert code[0] == S.SAME


def test_single_phase_density_from_the_supplied_root():
    """density mode 1: a 2 -> 1 candidate takes p / (Z R T) of the supplied Z and then meets
    the evidence rule, which a single-phase target never passes (an error, as the
    reference's adapter raises); without a Z (the record publishes no single-phase Z) the
    cell stays indeterminate; a supplied Z that is not positive is indeterminate too;
    multi-phase targets ignore it."""
    p, t = np.array([2.0e6, 2.0e6, 2.0e6]), np.array([320.0, 320.0, 320.0])
    status = np.array([_accepted(1), _accepted(1), _accepted(2)])
    zf = np.array([[np.nan, np.nan, 0.9], [np.nan, np.nan, 0.1], [np.nan] * 3])
    code, _, _ = S.decide([2, 2, 1], p, t, status, [1, 1, 2], zf)
    assert code.tolist() == [S.INDETERMINATE, S.INDETERMINATE, S.PROPOSAL]
    code, target, dens = S.decide([2, 2, 1], p, t, status, [1, 1, 2], zf, single_z=[0.85, -0.2, 5.0])
    assert code.tolist() == [S.ERROR, S.INDETERMINATE, S.PROPOSAL] and target.tolist() == [1, 1, 2]
    assert np.all(np.isnan(dens[:, 0])) and np.all(np.isnan(dens[:, 1]))
    assert dens[0, 2] == 2.0e6 / (0.9 * R * 320.0)

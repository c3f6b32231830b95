"""CPU restatement of the reference's post-SNES PT-flash phase-transition scan.

Test infrastructure only: the tests import it as the checker of mpmc_transition_scan and
mpmc_transition_scan_2p. The product never imports it.

Follows modules/flow/discretization/petsc/include/mpmc/flow_discretization_petsc/
post_snes_pt_flash_phase_transition_scanner.hpp:
  validate_source   validate_source_snapshot (:76-115)
  source_feed       make_source_snapshot (:160-200), the feed from the current inventory
  decide            scan_post_snes_pt_flash_source_cell_3d (:203-301) on a flash record,
                    with the transition-evidence rule of
                    make_phase_set_transition_candidate_from_flash
                    (modules/flow/include/mpmc/flow/phase_set_transition_flash_adapter.hpp:132-205)
                    and the density resolver of the reference's scanner test
                    (tests/flow_discretization/petsc/post_snes_pt_flash_phase_transition_scanner_test.cpp:26-57),
                    or (single_z) with a single phase's Z supplied by the caller
  summary           the cell loop of scan_post_snes_pt_flash_phase_transitions_3d (:330-420):
                    the first indeterminate cell ends the scan as indeterminate, the first
                    error cell ends it with the error

The reference accumulates and divides in long double; numpy's longdouble is the same x87
80-bit type on x86-64, and the sums here run in the reference's sequential order.
"""
from __future__ import annotations

import numpy as np

SAME, PROPOSAL, INDETERMINATE, ERROR = 0, 1, 2, 3
COMPLETE, STATUS_INDETERMINATE = 0, 1
EPS = np.finfo(np.float64).eps
GAS_CONSTANT = float(np.longdouble("8.31446261815324"))  # the PR76 provider's R as a double
FLAG_REJECTED = 0x2000
# status word fields (set | max3 << 4 | split << 8): Pr76PtMax3Status::two_phase and
# PtSplitStatus::phase_set_unstable (pr76_three_phase.hpp:133-140, pt_split.hpp:71-74)
MAX3_TWO = 1
SPLIT_UNSTABLE = 2


def validate_source(p, t, count, z, max_phases: int = 3) -> np.ndarray:
    """True where the source is well formed (:76-115) and its phase count is supported."""
    p = np.asarray(p, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    count = np.asarray(count)
    z = np.asarray(z, dtype=np.float64)
    nc = z.shape[0]
    ok = (count >= 1) & (count <= 3) & (count <= max_phases) & np.isfinite(p) & (p > 0) & np.isfinite(t) & (t > 0)
    total = np.zeros(z.shape[1], dtype=np.longdouble)
    for i in range(nc):
        ok &= np.isfinite(z[i]) & (z[i] > 0)
        total += z[i].astype(np.longdouble)
    tol = np.longdouble(4096.0) * np.longdouble(EPS) * np.longdouble(nc)
    with np.errstate(invalid="ignore"):
        ok &= np.isfinite(total) & (np.abs(total - np.longdouble(1.0)) <= tol)
    return ok & (nc >= 2)


def source_feed(inventory):
    """z_i = N_i / sum_j N_j in long double, rounded to double (:166-190); the second value
    marks the cells whose inventory is finite and positive (else NaN feed)."""
    inv = np.asarray(inventory, dtype=np.float64)
    nc = inv.shape[0] - 1   # the last row is the energy inventory
    n = inv.shape[1]
    ok = np.ones(n, dtype=bool)
    total = np.zeros(n, dtype=np.longdouble)
    for i in range(nc):
        ok &= np.isfinite(inv[i]) & (inv[i] > 0)
        total += inv[i].astype(np.longdouble)
    ok &= np.isfinite(total) & (total > 0)
    feed = np.full((nc, n), np.nan)
    with np.errstate(invalid="ignore", divide="ignore"):
        for i in range(nc):
            feed[i] = np.where(ok, (inv[i].astype(np.longdouble) / total).astype(np.float64), np.nan)
    return feed, ok


def decide(source_count, p, t, status, phase_count, zfactor, code_in=None, single_z=None):
    """Per-cell code, target count (the accepted phase count, 0 without an accepted set)
    and target densities (proposals only) from a flash record (:229-299).

    A flash that rejects the feed of a validated source raises std::domain_error, which the
    scanner catches as indeterminate (:292-299). Once the target densities are resolved, the
    candidate needs the backend's transition evidence
    (make_phase_set_transition_candidate_from_flash,
    phase_set_transition_flash_adapter.hpp:171-205): every higher target has it; a lower
    target only through a fresh disappearance solve, which the PR76 backend reports for
    3 -> 2 when its three-phase tail resolved a two-phase neighbour (max3 two-phase after an
    unstable base split, pr76_pt_flash_backend.hpp:138-146); any other lower target is the
    adapter's std::invalid_argument, an error.

    single_z (density mode 1, an extension with no reference counterpart): per cell, the Z
    that a single accepted phase takes when the record publishes none (the Z of its
    provider root at the feed)."""
    status = np.asarray(status)
    n = status.size
    src = np.broadcast_to(np.asarray(source_count), (n,))
    code = np.zeros(n, np.int32) if code_in is None else np.array(code_in, dtype=np.int32)
    target = np.zeros(n, np.int32)
    dens = np.full((3, n), np.nan)
    for c in range(n):
        if code[c] != SAME:
            continue
        st = int(status[c])
        if st & FLAG_REJECTED:
            code[c] = INDETERMINATE
            continue
        if st & 0xF != 0:          # no accepted phase set
            code[c] = INDETERMINATE
            continue
        target[c] = int(phase_count[c])
        if target[c] == int(src[c]):
            continue
        code[c] = PROPOSAL
        rho = np.full(3, np.nan)
        for ph in range(target[c]):
            zf = float(zfactor[ph][c])
            if single_z is not None and target[c] == 1 and not np.isfinite(zf):
                zf = float(single_z[c])
            den = zf * GAS_CONSTANT * float(t[c])
            r = float(p[c]) / den if den != 0 else np.inf
            if not (np.isfinite(zf) and zf > 0 and np.isfinite(r) and r > 0):
                code[c] = INDETERMINATE
            rho[ph] = r
        neighbour = int(src[c]) == 3 and target[c] == 2 and (st >> 4) & 0xF == MAX3_TWO and \
            (st >> 8) & 0xF == SPLIT_UNSTABLE
        if code[c] == PROPOSAL and target[c] < int(src[c]) and not neighbour:
            code[c] = ERROR
        if code[c] == PROPOSAL:
            dens[:, c] = rho
    return code, target, dens


def summary(code):
    """(status, first_stop, error_first, proposals) of the reference's ordered cell loop
    (:366-407): it returns at the first indeterminate cell with its proposals cleared, or at
    the first error with the proposals collected before that cell."""
    code = np.asarray(code)
    stops = np.nonzero((code == INDETERMINATE) | (code == ERROR))[0]
    if stops.size == 0:
        return COMPLETE, -1, False, int(np.sum(code == PROPOSAL))
    first = int(stops[0])
    err = bool(code[first] == ERROR)
    return (COMPLETE if err else STATUS_INDETERMINATE), first, err, \
        (int(np.sum(code[:first] == PROPOSAL)) if err else 0)

"""Shared cases of the ordered-path PT continuation tests (oracle and GPU).

REFERENCE_SCENARIOS restate the paths of the reference's own continuation tests
(tests/flash/pr76_pt_continuation/pr76_pt_continuation_test.cpp:33-215) on the synthetic
three-component fixture (tests/flash/pr76_three_phase/synthetic_fixture.hpp:32-96) with the
test's structural options (:24-30): the fixture's three reference phase compositions as the
fixed three-phase start (seed 1/3, 1/3) and as both fallback start lists. Each scenario lists
what the reference test requires: per point the accepted phase count (None: unresolved), and
for chosen points the incoming hint kind, carried start count, carried three-phase start and
its consumption; ``rejected`` marks a path the reference rejects before any solve.
"""
from __future__ import annotations

import numpy as np

# synthetic_fixture.hpp:32-66 (symmetric max3 fixture) and :68-77 (reference phases)
FIXTURE = {
    "tc": np.array([190.6, 500.0, 500.0]),
    "pc": np.array([4.6e6, 5.0e6, 5.0e6]),
    "omega": np.array([0.01, 0.10, 0.10]),
    "kij": np.array([[0.0, 0.05, 0.05], [0.05, 0.0, 0.20], [0.05, 0.20, 0.0]]),
}
REFERENCE_PHASES = np.array([
    [0.04844395156186141, 0.9076712804842868, 0.0438847679538518],
    [0.9696550397025219, 0.015172480148739028, 0.015172480148739028],
    [0.048443951561861175, 0.04388476795385164, 0.9076712804842871],
])
EQUAL_FEED = np.full((3, 1), 1.0 / 3.0)

# expected: list of accepted phase counts; checks: {point: (incoming, carried_count or None,
# carried_three, consumed)}
REFERENCE_SCENARIOS = {
    # repeated_three_phase_is_fresh_continuation (:33-58)
    "repeated_three_phase": dict(path=[(1.0e6, 250.0)] * 3, counts=[3, 3, 3],
                                 checks={0: (0, None, False, None), 1: (3, None, True, True),
                                         2: (None, None, None, True)}),
    # moving_three_phase_is_fresh_continuation (:92-125)
    "moving_three_phase": dict(path=[(1.90e6, 325.0), (1.99e6, 325.6)], counts=[3, 3],
                               checks={1: (3, 3, True, True)}),
    # invalid_path_is_rejected_before_solve (:60-72)
    "invalid_path": dict(path=[(-1.0, 250.0)], rejected=True),
    # bidirectional_phase_sequence (:146-184), forward and reverse
    "forward": dict(path=[(1.90e6, 325.0), (2.80e6, 331.0), (1.00e6, 347.0), (1.00e6, 348.0)],
                    counts=[3, 2, 2, 1], checks={1: (3, None, True, False), 2: (2, None, False, None)}),
    "reverse": dict(path=[(1.00e6, 348.0), (1.00e6, 347.0), (2.80e6, 331.0), (1.90e6, 325.0)],
                    counts=[1, 2, 2, 3], checks={}),
    # unresolved_point_resets_continuation (:186-215)
    "unresolved_resets": dict(path=[(1.0e6, 330.0), (1.0e6, 335.0), (1.0e6, 340.0)], counts=[3, None, 2],
                              checks={2: (0, 0, False, None)}),
}


def scenario_arrays(path):
    p = np.array([[a] for a, _ in path], dtype=np.float64)
    t = np.array([[b] for _, b in path], dtype=np.float64)
    return p, t


def check_scenario(name, case, rec):
    """Assert the reference test's requirements on a continuation record of one path
    (arrays with a leading point axis, path axis of length 1)."""
    status = rec["status"][:, 0].astype(np.int64)
    info = rec["info"][:, 0].astype(np.int64)
    if case.get("rejected"):
        assert np.all(status & 0x2000), (name, status)
        return
    assert not np.any(status & 0x2000), (name, status)
    for k, want in enumerate(case["counts"]):
        accepted = (status[k] & 0xF) == 0
        if want is None:
            assert not accepted, (name, k, hex(status[k]))
        else:
            assert accepted and rec["phase_count"][k, 0] == want, (name, k, hex(status[k]))
    # transition brackets: adjacent accepted points with different counts, each an adjacent step
    counts = case["counts"]
    for k in range(1, len(counts)):
        if counts[k - 1] is not None and counts[k] is not None and counts[k - 1] != counts[k]:
            assert abs(counts[k - 1] - counts[k]) == 1, (name, k)
    for k, (incoming, carried, three, consumed) in case["checks"].items():
        if incoming is not None:
            assert info[k] & 0xF == incoming, (name, k, hex(info[k]))
        if carried is not None:
            assert (info[k] >> 4) & 0xF == carried, (name, k, hex(info[k]))
        if three is not None:
            assert bool(info[k] & 0x100) == three, (name, k, hex(info[k]))
        if consumed is not None:
            assert bool(status[k] & 0x20000) == consumed, (name, k, hex(status[k]))


def flatten(rec):
    """A continuation record [npoint, ..., npath] as one flash record of npoint*npath cells
    (point-major), for the flash comparison helpers."""
    m, n = rec["status"].shape
    out = {"status": rec["status"].reshape(m * n), "phase_count": rec["phase_count"].reshape(m * n)}
    for k in ("beta", "zfactor", "branch"):
        out[k] = np.ascontiguousarray(np.moveaxis(rec[k], 0, 1).reshape(3, m * n))
    for k in ("x", "ln_phi"):
        a = rec[k]
        out[k] = np.ascontiguousarray(np.moveaxis(a, 0, 2).reshape(3, a.shape[2], m * n))
    if rec.get("counters") is not None:
        out["counters"] = np.ascontiguousarray(np.moveaxis(rec["counters"], 0, 1).reshape(4, m * n))
    return out

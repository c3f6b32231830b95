"""Pins the oracle (oracle/cpuref.cpp) on the CPU:

* against the Decimal(80) goldens the reference's own tests hold (PR76 phase values,
  two binary VLE splits, a ternary split);
* against closed forms computed independently here in long double (the PR76 cubic in
  its original Z form with textbook mixing, the binary Rachford-Rice solution);
* against vectors written by the reference's own flash (tests/golden/*.npz), and live
  against the reference build when it is present.
"""
import math
from pathlib import Path

import numpy as np
import pytest

import oracle_bindings as ob
import parity
from golden import reference_goldens as G
from mpmc_hnu_amd import workloads as W

GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_flash_vectors.npz"
LD = np.longdouble
R_LD = LD("8.31446261815324")


def _flash1(mix, p, t, feed, maxph=2, tol=1e-11):
    z = np.asarray(feed, dtype=np.float64).reshape(-1, 1)
    return ob.flash(mix, [p], [t], z, tol, maxph)


def test_pt_phase_decimal_golden():
    s = G.PT_STATE
    r = ob.phase(G.PT_MIXTURE, [s["p"]], [s["t"]], s["w"].reshape(3, 1))
    assert r["root_status"][0] == 0 and r["root_count"][0] == 3
    for k in range(3):
        assert abs(r["zroot"][k, 0] - G.PT_GOLDEN[k, 0]) <= G.PT_TOL * abs(G.PT_GOLDEN[k, 0])
        for i in range(3):
            g = G.PT_GOLDEN[k, i + 1]
            assert abs(r["ln_phi"][k, i, 0] - g) <= G.PT_TOL * abs(g), (k, i)


def _textbook_phase_ld(mix, p, t, w):
    """Roots and ln(phi) from the original PR76 Z cubic with textbook vdW mixing, in
    long double: a_ij = sqrt(a_i a_j)(1-k_ij), a = sum_ij w_i w_j a_ij."""
    tc, pc, om, kij = (np.asarray(mix[k], dtype=LD) for k in ("tc", "pc", "omega", "kij"))
    p, t, w = LD(p), LD(t), np.asarray(w, dtype=LD)
    kappa = LD("0.37464") + LD("1.54226") * om - LD("0.26992") * om * om
    ai = LD("0.45724") * (R_LD * tc) ** 2 / pc * (1 + kappa * (1 - np.sqrt(t / tc))) ** 2
    bi = LD("0.07780") * R_LD * tc / pc
    aij = np.sqrt(np.outer(ai, ai)) * (1 - kij)
    a = w @ aij @ w
    b = w @ bi
    A = a * p / (R_LD * t) ** 2
    B = b * p / (R_LD * t)
    c = [LD(1), -(1 - B), A - 3 * B * B - 2 * B, -(A * B - B * B - B ** 3)]
    approx = np.roots(np.array(c, dtype=np.float64))
    roots = []
    for z0 in approx:
        if abs(z0.imag) > 1e-7:
            continue
        z = LD(z0.real)
        for _ in range(60):  # Newton polish in long double
            f = ((c[0] * z + c[1]) * z + c[2]) * z + c[3]
            d = (3 * c[0] * z + 2 * c[1]) * z + c[2]
            z = z - f / d
        if z > B:
            roots.append(z)
    roots = sorted(roots)
    s2 = np.sqrt(LD(2))
    out = []
    for z in roots:
        lphi = bi / b * (z - 1) - np.log(z - B) - A / (2 * s2 * B) * (2 * (aij @ w) / a - bi / b) * \
            np.log((z + (1 + s2) * B) / (z + (1 - s2) * B))
        out.append((z, lphi))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_phase_against_textbook_long_double(seed):
    rng = np.random.default_rng(seed)
    nc = int(rng.integers(2, 9))
    mix = W.mixture_linspaced(nc, seed=seed)
    p, t, x = W.cells_uniform(nc, 64, seed=seed)
    r = ob.phase(mix, p, t, x)
    checked = 0
    for c in range(p.size):
        if r["root_status"][c] != 0:
            continue
        ref = _textbook_phase_ld(mix, p[c], t[c], x[:, c])
        if len(ref) != r["root_count"][c]:
            # a near-degenerate pair can be resolved differently by the two methods
            assert len(ref) >= 1
            continue
        for k, (z, lphi) in enumerate(ref):
            assert abs(r["zroot"][k, c] - float(z)) <= 1e-12 * float(z)
            flags = r["root_flags"][k, c]
            if (flags & 0xF) == 2:  # admissible root evaluated
                got = r["ln_phi"][k, :, c]
                assert np.allclose(got, np.asarray(lphi, dtype=np.float64), rtol=1e-10, atol=1e-11)
        checked += 1
    assert checked >= 48


def test_rachford_rice_binary_closed_form():
    rng = np.random.default_rng(0)
    for _ in range(200):
        z1 = rng.uniform(0.05, 0.95)
        z = np.array([z1, 1 - z1])
        lk = np.array([rng.uniform(0.2, 3.0), -rng.uniform(0.2, 3.0)])
        K = np.exp(lk.astype(LD))
        beta_exact = -(z[0] * (K[0] - 1) + z[1] * (K[1] - 1)) / ((K[0] - 1) * (K[1] - 1))
        st, beta, x, y, it = ob.rachford_rice(z, lk)
        if not (0 < beta_exact < 1):
            assert st != 0
            continue
        assert st == 0
        assert abs(beta - float(beta_exact)) <= 1e-13
        assert np.allclose(beta * y + (1 - beta) * x, z, rtol=0, atol=1e-15)
        assert np.allclose(y / x, np.exp(lk), rtol=1e-12)


@pytest.mark.parametrize("system", sorted(G.SPLIT_SYSTEMS))
def test_binary_split_decimal_goldens(system):
    s = G.SPLIT_SYSTEMS[system]
    for k, f in enumerate(s["feeds"]):
        r = _flash1(s["mixture"], s["p"], s["t"], [f, 1 - f], maxph=2)
        assert (r["status"][0] & 0xFFF) == 0 | (0xF << 4) | (1 << 8)  # accepted, two-phase
        assert r["phase_count"][0] == 2
        g = s["golden"][k]
        got = [r["x"][0, 0, 0], r["x"][1, 0, 0], r["beta"][1, 0], r["zfactor"][0, 0], r["zfactor"][1, 0]]
        for a, b in zip(got, g):
            assert abs(a - b) <= G.SPLIT_TOL * (1 + abs(b)), (k, a, b)


def test_ternary_split_decimal_goldens():
    for case in G.TERNARY_CASES:
        r = _flash1(G.TERNARY_MIXTURE, G.TERNARY_STATE["p"], G.TERNARY_STATE["t"], case["feed"], maxph=2)
        assert r["phase_count"][0] == 2 and (r["status"][0] & 0xF) == 0
        near = lambda a, b: abs(a - b) <= G.TERNARY_TOL * (1 + abs(b))  # noqa: E731
        assert near(r["beta"][1, 0], case["beta"])
        for i in range(3):
            assert near(r["x"][0, i, 0], case["liquid"][i]) and near(r["x"][1, i, 0], case["vapor"][i])
        assert near(r["zfactor"][0, 0], G.TERNARY_Z[0]) and near(r["zfactor"][1, 0], G.TERNARY_Z[1])


def _golden_cases():
    if not GOLDEN.exists():
        return []
    with np.load(GOLDEN) as d:
        return sorted({k.split("/")[0] for k in d.files})


@pytest.mark.parametrize("case", _golden_cases())
def test_oracle_matches_reference_vectors(case):
    with np.load(GOLDEN) as d:
        mix = {k: d[f"{case}/{k}"] for k in ("tc", "pc", "omega", "kij")}
        p, t, z = d[f"{case}/p"], d[f"{case}/t"], d[f"{case}/z"]
        tol, maxph = float(d[f"{case}/tol"]), int(d[f"{case}/max_phases"])
        ref = {k: d[f"{case}/out/{k}"] for k in ("status", "phase_count", "beta", "zfactor", "branch", "x", "ln_phi")}
        e_prop = d[f"{case}/out/e_prop"]
    o = ob.flash(mix, p, t, z, tol, maxph)
    rep, bad = parity.compare(o, ref, rel=0.0, abs_lnphi=0.0)
    # same arithmetic in the same order on the same host: bit-identical
    assert not bad.any(), rep
    assert np.array_equal(o["counters"][0], e_prop)


@pytest.mark.skipif(not ob.have_ref(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("nc", [2, 3, 5, 6, 8, 11])
def test_oracle_matches_reference_live(nc):
    mix = W.mixture_linspaced(nc, seed=nc + 20)
    p, t, z = W.cells_uniform(nc, 512, seed=nc + 40)
    o = ob.flash(mix, p, t, z, 1e-11, 3)
    r = ob.flash(mix, p, t, z, 1e-11, 3, which="ref")
    rep, bad = parity.compare(o, r, rel=0.0, abs_lnphi=0.0)
    assert not bad.any(), rep
    assert np.array_equal(o["counters"][:2], r["counters"][:2])
    # the whole status word, search flags included (MPMC_FLAG_SPLIT_*, MPMC_FLAG_MAX3_ATTEMPTED)
    assert np.array_equal(o["status"], r["status"])


@pytest.mark.skipif(not ob.have_ref(), reason="reference build (oracle/_ref) not present")
def test_search_flags_match_reference():
    """The status-word search flags (a disappearance attempt, a converged candidate, a
    three-phase attempt) equal the reference's, on the sour-gas sweep (three-phase attempts,
    indeterminate splits with a candidate) and on feeds 5e-11 inside a two-phase boundary
    (split attempts that end at a phase-fraction endpoint)."""
    mix, p, t, z = W.sour_gas_sweep(2048)
    idx = np.arange(0, 2048, 2)
    p, t, z = p[idx], t[idx], np.ascontiguousarray(z[:, idx])
    o = ob.flash(mix, p, t, z, 1e-11, 3)
    r = ob.flash(mix, p, t, z, 1e-11, 3, which="ref")
    assert np.array_equal(o["status"], r["status"])
    assert np.any(o["status"] & 0x8000) and np.any(o["status"] & 0x10000)
    indet = (o["status"] >> 8 & 0xF) == 3
    assert np.any(indet & (o["status"] & 0x8000 != 0))
    mix = W.mixture_c1()
    p, t, z = W.cells_c1()
    p, t, z = W.boundary_feeds(p, t, ob.flash(mix, p, t, z, 1e-12, 3), 5e-11)
    o = ob.flash(mix, p, t, z, 1e-12, 3)
    r = ob.flash(mix, p, t, z, 1e-12, 3, which="ref")
    assert np.array_equal(o["status"], r["status"])
    assert np.sum(o["status"] & 0x4000 != 0) > 50


def test_planted_mistake_is_caught():
    """A dropped term in the oracle's tangent-plane distance would flip decisions: emulate
    it by flashing a perturbed mixture (kij of the first pair shifted by 1e-3) and check
    that the comparison used by the parity tests reports it."""
    mix = W.mixture_linspaced(4)
    p, t, z = W.cells_uniform(4, 256, seed=3)
    good = ob.flash(mix, p, t, z, 1e-11, 3)
    bad_mix = dict(mix)
    kij = mix["kij"].copy()
    kij[0, 1] += 1e-3
    kij[1, 0] += 1e-3
    bad_mix["kij"] = kij
    wrong = ob.flash(bad_mix, p, t, z, 1e-11, 3)
    rep, bad = parity.compare(wrong, good)
    assert bad.any() and rep["value_mismatch"] + rep["status_mismatch"] > 0


@pytest.mark.skipif(not ob.have_ref(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("nc", [3, 6])
def test_warm_start_hints_match_reference(nc):
    """Warm starts (PtSolveHints::initial_stability_starts through the backend's
    initial_starts): the phase compositions of a flash at 3% lower pressure, handed to the
    flash of the same feeds. The oracle and the compiled reference agree bit for bit, the
    hints add stability trials, and a hint with the wrong support rejects the cell."""
    mix = W.mixture_linspaced(nc, seed=nc + 60)
    p, t, z = W.cells_uniform(nc, 384, seed=nc + 70)
    prev = ob.flash(mix, p * 0.97, t, z, 1e-11, 3, which="ref")
    hints = W.warm_start_hints(prev, z)
    assert np.isfinite(hints[1, 0]).sum() > 10
    bad_cell = int(np.nonzero(np.isfinite(hints[0, 0]))[0][0])
    hints[0, :, bad_cell] = 0.0
    hints[0, 0, bad_cell] = 1.0  # support differs from the feed's
    o = ob.flash(mix, p, t, z, 1e-11, 3, hints=hints)
    r = ob.flash(mix, p, t, z, 1e-11, 3, which="ref", hints=hints)
    rep, bad = parity.compare(o, r, rel=0.0, abs_lnphi=0.0)
    assert not bad.any(), rep
    assert np.array_equal(o["counters"][:2], r["counters"][:2])
    assert o["status"][bad_cell] & 0x2000 and r["status"][bad_cell] & 0x2000
    plain = ob.flash(mix, p, t, z, 1e-11, 3)
    hinted = np.isfinite(hints[0, 0])
    hinted[bad_cell] = False
    assert np.all(o["counters"][1][hinted] > plain["counters"][1][hinted])


@pytest.mark.parametrize("max_phases", [2, 3])
def test_final_review_hints_match_reference(max_phases):
    """Final-review warm starts (PtSolveHints::final_two_phase_stability_starts through
    final_starts): the phase compositions of a flash at 3% lower pressure handed to the
    common-tangent review of the same feeds, together with initial hints. The oracle and
    the compiled reference agree bit for bit, the review runs the extra trials, and a final
    hint with the wrong support rejects the cell."""
    nc = 6
    mix = W.mixture_linspaced(nc, seed=77)
    p, t, z = W.cells_uniform(nc, 384, seed=78)
    prev = ob.flash(mix, p * 0.97, t, z, 1e-11, max_phases, which="ref")
    fh = W.warm_start_hints(prev, z)
    ih = W.warm_start_hints(prev, z, nhint=1)
    assert np.isfinite(fh[0, 0]).sum() > 10
    bad_cell = int(np.nonzero(np.isfinite(fh[1, 0]))[0][0])
    fh[1, :, bad_cell] = 0.0
    fh[1, 0, bad_cell] = 1.0  # support differs from the feed's
    o = ob.flash(mix, p, t, z, 1e-11, max_phases, hints=ih, final_hints=fh)
    r = ob.flash(mix, p, t, z, 1e-11, max_phases, which="ref", hints=ih, final_hints=fh)
    rep, bad = parity.compare(o, r, rel=0.0, abs_lnphi=0.0)
    assert not bad.any(), rep
    assert np.array_equal(o["counters"][:2], r["counters"][:2])
    assert o["status"][bad_cell] & 0x2000 and r["status"][bad_cell] & 0x2000
    # the review of two-phase cells with a final hint runs more trials than without
    plain = ob.flash(mix, p, t, z, 1e-11, max_phases, hints=ih)
    two = (plain["phase_count"] == 2) & np.isfinite(fh[0, 0])
    two[bad_cell] = False
    assert two.sum() > 0
    assert np.all(o["counters"][1][two] > plain["counters"][1][two])

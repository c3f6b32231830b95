"""The oracle's closure and implicit sensitivity (oracle/cpuref.cpp, build_closure and
differentiate) pinned on the CPU:

* against the Decimal goldens of the reference's own tests
  (tests/flash/pt_split/sensitivity_references.hpp, tests/physics/thermodynamic_closure/
  closure_references.hpp) with the tolerances those tests use;
* against the reference itself (oracle/_ref, built from its headers), bit for bit;
* against what the mathematics fixes: central differences of the oracle's own flash,
  sum(dx) = 0 in each column, and the differentiated material balance;
* against the reference's status policy tests (closure_test.cpp:264-317).
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_bindings as ob
from golden import reference_goldens as G
from mpmc_hnu_amd import workloads as W

Q_NAMES = ("p", "T")


def near(actual, expected, rel, floor):
    actual = np.asarray(actual, dtype=np.float64)
    expected = np.asarray(expected, dtype=np.float64)
    tol = rel * np.maximum(floor, np.abs(expected))
    return np.all(np.isfinite(actual)) and np.all(np.abs(actual - expected) <= tol)


def run_case(name, which="cpuref", **kw):
    case = G.CLOSURE_CASES[name]
    z = np.array(case["feed"], dtype=np.float64)[:, None]
    return ob.closure(case["mixture"], np.array([case["p"]]), np.array([case["t"]]), z, which=which, **kw), \
        z.shape[0]


@pytest.mark.parametrize("name", ["binary", "ternary"])
def test_sensitivity_matches_reference_goldens(name):
    r, nc = run_case(name)
    assert r["status"][0] == 0, hex(r["status"][0])
    S = G.SENSITIVITY_GOLDEN[name]
    rel, floor = G.SENSITIVITY_TOL
    assert near(r["jac"][0, :, 0], S["beta"], rel, floor)
    assert near(r["log_k_jac"][:, :, 0], S["log_k"], rel, floor)
    assert near(r["jac"][1:1 + nc, :, 0], S["liquid"], rel, floor)
    assert near(r["jac"][1 + nc:1 + 2 * nc, :, 0], S["vapor"], rel, floor)
    # sensitivity_test.cpp:191-200: well conditioned, tiny backward error and residual
    assert r["diag"][0, 0] > 0.1 and r["diag"][1, 0] < 1e-12 and r["diag"][2, 0] <= 1e-11 + 1e-11


@pytest.mark.parametrize("name", ["binary", "ternary"])
def test_closure_matches_reference_goldens(name):
    r, nc = run_case(name)
    C = G.CLOSURE_GOLDEN[name]
    prim = r["primal"][:, 0]
    assert np.all(np.abs(prim - np.array(C["primal"])) <= G.CLOSURE_PRIMAL_TOL * (1 + np.abs(C["primal"])))
    rel, floor = G.CLOSURE_DERIV_TOL
    for row, key in ((2 * nc + 1, "z_l"), (2 * nc + 2, "z_v"), (2 * nc + 3, "rho_l"), (2 * nc + 4, "rho_v")):
        assert near(r["jac"][row, :, 0], C[key], rel, floor), key


def test_closure_status_policies():
    """closure_test.cpp:264-317: a derivative guard keeps the primal; an indeterminate flash
    publishes nothing."""
    r, _ = run_case("binary", min_fraction=0.45)
    assert r["status"][0] == 0 | (1 << 4) | (2 << 8)  # valid, unavailable, phase_boundary
    assert np.all(np.isfinite(r["primal"][:, 0])) and np.all(np.isnan(r["jac"][:, :, 0]))
    r, _ = run_case("binary", min_rcond=0.90)
    assert r["status"][0] == 0 | (1 << 4) | (3 << 8)  # ill_conditioned_equilibrium
    mix = G.SPLIT_SYSTEMS["carbon_dioxide_methane"]["mixture"]
    r = ob.closure(mix, np.array([2339623.8796839858]), np.array([220.0]), np.array([[0.30], [0.70]]))
    assert r["status"][0] == 2 | (1 << 4)  # indeterminate primal, no payload
    assert np.all(np.isnan(r["primal"][:, 0]))


def test_closure_against_central_differences():
    """d beta/dq, dx/dq and dZ/dq against central differences of the oracle's own VLE
    flash: a dropped term or a wrong column fails this at the 1e-4 level."""
    case = G.CLOSURE_CASES["ternary"]
    mix, p0, t0 = case["mixture"], case["p"], case["t"]
    z0 = np.array(case["feed"])
    r = ob.closure(mix, np.array([p0]), np.array([t0]), z0[:, None])
    assert r["status"][0] == 0
    nc = z0.size
    steps = [(0, p0 * 1e-5), (1, t0 * 1e-5)] + [(2 + k, 1e-5) for k in range(nc - 1)]
    for col, h in steps:
        ps, ts, zs = np.full(2, p0), np.full(2, t0), np.repeat(z0[:, None], 2, axis=1)
        if col == 0:
            ps += (h, -h)
        elif col == 1:
            ts += (h, -h)
        else:
            zs[col - 2] += (h, -h)
            zs[nc - 1] -= (h, -h)
        f = ob.flash(mix, ps, ts, zs, max_phases=2)
        assert np.all((f["status"] & 0xF) == 0) and np.all(f["phase_count"] == 2)
        fd_beta = (f["beta"][1, 0] - f["beta"][1, 1]) / (2 * h)
        fd_x = (f["x"][0, :, 0] - f["x"][0, :, 1]) / (2 * h)
        fd_zl = (f["zfactor"][0, 0] - f["zfactor"][0, 1]) / (2 * h)
        scale = max(1e-6, np.max(np.abs(r["jac"][:, col, 0])))
        assert abs(fd_beta - r["jac"][0, col, 0]) <= 1e-4 * scale
        assert np.all(np.abs(fd_x - r["jac"][1:1 + nc, col, 0]) <= 1e-4 * scale)
        assert abs(fd_zl - r["jac"][2 * nc + 1, col, 0]) <= 1e-4 * scale


@pytest.mark.parametrize("nc", [2, 3, 6, 12])
def test_closure_invariants(nc):
    """sum_i dx_i = sum_i dy_i = 0 and (1-b) dx + b dy + (y-x) db = dz in every column."""
    mix = W.mixture_linspaced(nc)
    p, t, z = W.cells_uniform(nc, 512, seed=11 + nc)
    r = ob.closure(mix, p, t, z)
    f = ob.flash(mix, p, t, z, max_phases=2)
    ok = r["status"] == 0
    assert ok.sum() > 5
    J = r["jac"][:, :, ok]
    beta = f["beta"][1, ok]
    x, y = f["x"][0][:, ok], f["x"][1][:, ok]
    q = nc + 1
    for col in range(q):
        dx, dy, db = J[1:1 + nc, col], J[1 + nc:1 + 2 * nc, col], J[0, col]
        scale = 1 + np.abs(dx).sum(0) + np.abs(dy).sum(0) + np.abs(db)
        assert np.all(np.abs(dx.sum(0)) <= 1e-10 * scale)
        assert np.all(np.abs(dy.sum(0)) <= 1e-10 * scale)
        dz = np.zeros((nc, 1))
        if col >= 2:
            dz[col - 2] = 1.0
            dz[nc - 1] = -1.0
        bal = (1 - beta) * dx + beta * dy + (y - x) * db - dz
        assert np.all(np.abs(bal) <= 1e-9 * scale)


@pytest.mark.skipif(not ob.have_ref(), reason="reference build absent (oracle/_ref)")
@pytest.mark.parametrize("nc", [2, 3, 6, 12])
def test_oracle_closure_is_the_reference_closure(nc):
    """The restatement against build_pr76_pt_vle_closure compiled from the reference, bit for bit."""
    mix = W.mixture_linspaced(nc)
    p, t, z = W.cells_uniform(nc, 1024, seed=40 + nc)
    a = ob.closure(mix, p, t, z, which="cpuref")
    b = ob.closure(mix, p, t, z, which="ref")
    assert np.array_equal(a["status"], b["status"])
    assert np.array_equal(a["primal"], b["primal"], equal_nan=True)
    assert np.array_equal(a["jac"], b["jac"], equal_nan=True)
    ok = b["status"] == 0
    assert ok.sum() > 5
    assert np.array_equal(a["log_k_jac"][:, :, ok], b["log_k_jac"][:, :, ok])
    assert np.array_equal(a["diag"][:, ok], b["diag"][:, ok])


CLOSURE_VECTORS = Path(__file__).resolve().parent / "golden" / "ref_closure_vectors.npz"


@pytest.mark.parametrize("name", ["cl2", "cl3", "cl6", "cl12"])
def test_oracle_closure_matches_reference_vectors(name):
    """Against closures the reference computed for seeded cells (tests/golden/make_golden.py):
    bit for bit, without needing the reference build."""
    d = np.load(CLOSURE_VECTORS)
    mix = {k: d[f"{name}/{k}"] for k in ("tc", "pc", "omega", "kij")}
    r = ob.closure(mix, d[f"{name}/p"], d[f"{name}/t"], d[f"{name}/z"])
    assert np.array_equal(r["status"], d[f"{name}/out/status"])
    assert np.array_equal(r["primal"], d[f"{name}/out/primal"], equal_nan=True)
    assert np.array_equal(r["jac"], d[f"{name}/out/jac"], equal_nan=True)
    ok = r["status"] == 0
    assert ok.any()
    assert np.array_equal(r["log_k_jac"][:, :, ok], d[f"{name}/out/log_k_jac"][:, :, ok])

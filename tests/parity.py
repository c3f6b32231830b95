"""Result comparison shared by the oracle tests and the GPU parity tests.

Bar (per cell): status word, phase count and per-phase branch word (root index and
smooth flag) identical; phase fractions and compositions within ``rel`` relative;
ln(phi) within ``abs_lnphi`` absolute plus ``rel`` relative.
"""
from __future__ import annotations

import numpy as np


def compare(a: dict, b: dict, rel: float = 1e-9, abs_lnphi: float = 1e-9, ignore_flags: bool = True):
    """Returns (report dict, mismatching-cell mask)."""
    sa, sb = a["status"], b["status"]
    if ignore_flags:
        sa, sb = sa & 0xFFF, sb & 0xFFF
    status_bad = sa != sb
    count_bad = a["phase_count"] != b["phase_count"]
    np_ = np.minimum(a["phase_count"], b["phase_count"])
    n = sa.size
    branch_bad = np.zeros(n, bool)
    beta_err = np.zeros(n)
    x_err = np.zeros(n)
    lp_err = np.zeros(n)
    z_err = np.zeros(n)
    for ph in range(3):
        have = np_ > ph
        branch_bad |= have & (a["branch"][ph] != b["branch"][ph])
        ba, bb = a["beta"][ph], b["beta"][ph]
        e = np.where(have, np.abs(ba - bb) / np.maximum(np.abs(bb), 1e-300), 0.0)
        beta_err = np.maximum(beta_err, e)
        xa, xb = a["x"][ph], b["x"][ph]
        e = np.where(have[None, :], np.abs(xa - xb) / np.maximum(np.abs(xb), 1e-300), 0.0).max(axis=0)
        x_err = np.maximum(x_err, e)
        la, lb = a["ln_phi"][ph], b["ln_phi"][ph]
        if rel > 0:
            e = np.where(have[None, :], np.abs(la - lb) / (abs_lnphi / rel + np.abs(lb)), 0.0).max(axis=0)
        else:
            e = np.where(have[None, :] & (la != lb), np.inf, 0.0).max(axis=0)
        lp_err = np.maximum(lp_err, e)
        za, zb = a["zfactor"][ph], b["zfactor"][ph]
        both = have & np.isfinite(za) & np.isfinite(zb)
        e = np.where(both, np.abs(za - zb) / np.maximum(np.abs(zb), 1e-300), 0.0)
        z_err = np.maximum(z_err, e)
        zmiss = have & (np.isfinite(za) != np.isfinite(zb))
        z_err = np.where(zmiss, np.inf, z_err)
    value_bad = (beta_err > rel) | (x_err > rel) | (lp_err > rel) | (z_err > rel)
    decision_bad = status_bad | count_bad | branch_bad
    bad = decision_bad | value_bad
    report = {
        "cells": int(n),
        "status_mismatch": int(status_bad.sum()),
        "count_mismatch": int(count_bad.sum()),
        "branch_mismatch": int(branch_bad.sum()),
        "value_mismatch": int((value_bad & ~decision_bad).sum()),
        "decision_cells": int(decision_bad.sum()),
        "max_beta_rel": float(beta_err[~decision_bad].max(initial=0.0)),
        "max_x_rel": float(x_err[~decision_bad].max(initial=0.0)),
        "max_lnphi_err": float(lp_err[~decision_bad].max(initial=0.0) * rel),
        "max_z_rel": float(z_err[~decision_bad].max(initial=0.0)),
    }
    return report, bad


def fugacity_residual(res: dict) -> np.ndarray:
    """max_i |ln(x_i phi_i)_0 - ln(x_i phi_i)_k| over the published phases of each cell."""
    n = res["status"].size
    out = np.zeros(n)
    npc = res["phase_count"]
    mu0 = np.log(res["x"][0]) + res["ln_phi"][0]
    for ph in (1, 2):
        have = npc > ph
        mu = np.log(res["x"][ph]) + res["ln_phi"][ph]
        d = np.where(have[None, :], np.abs(mu - mu0), 0.0).max(axis=0)
        out = np.maximum(out, d)
    return out


def _value_err(a: dict, b: dict, cells: np.ndarray) -> np.ndarray:
    """Largest relative difference of beta, x and Z of a and b at ``cells`` (same decisions)."""
    out = np.zeros(cells.size)
    npc = np.minimum(a["phase_count"][cells], b["phase_count"][cells])
    for ph in range(3):
        have = npc > ph
        for key in ("beta", "zfactor"):
            va, vb = a[key][ph][cells], b[key][ph][cells]
            e = np.abs(va - vb) / np.maximum(np.abs(vb), 1e-300)
            out = np.maximum(out, np.where(have & np.isfinite(e), e, 0.0))
        xa, xb = a["x"][ph][:, cells], b["x"][ph][:, cells]
        e = (np.abs(xa - xb) / np.maximum(np.abs(xb), 1e-300)).max(axis=0)
        out = np.maximum(out, np.where(have, e, 0.0))
    return out


def within_stopping_error(g: dict, o: dict, cells: np.ndarray, tight) -> tuple:
    """Cells whose values differ beyond 1e-9 between the device (g) and the oracle (o), both
    stopped at the same fugacity tolerance tol, judged against ``tight(cells, factor)``: the
    oracle at tolerance tol * factor on those cells, with factor 1e-2, or 1e-1 where the
    tighter solve reaches another decision. The device value is accepted when it is within
    1e-9 of the tight solution, or no farther from it than twice the oracle's own stopping
    error. A phase fraction of order 1e-4 is determined only to about 1e-12 by a fugacity
    residual of 1e-11, so 1e-9 relative can be below the reference's own determinacy there.
    Returns (bad cells, per-cell (device error, oracle error))."""
    if cells.size == 0:
        return cells, []
    sub = {k: (v[..., cells] if isinstance(v, np.ndarray) else v) for k, v in g.items()}
    subo = {k: (v[..., cells] if isinstance(v, np.ndarray) else v) for k, v in o.items()}
    idx = np.arange(cells.size)
    ok = np.zeros(cells.size, bool)
    errs = [(np.inf, np.inf)] * cells.size
    for factor in (1e-2, 1e-1):
        t = tight(cells, factor)
        same = ((t["status"] & 0xFFF) == (subo["status"] & 0xFFF)) & (t["phase_count"] == subo["phase_count"])
        eg = _value_err(sub, t, idx)
        eo = _value_err(subo, t, idx)
        now = ~ok & same
        ok |= now & ((eg <= 1e-9) | (eg <= 2.0 * eo))
        for i in np.nonzero(now)[0]:
            errs[i] = (float(eg[i]), float(eo[i]))
    return cells[~ok], errs

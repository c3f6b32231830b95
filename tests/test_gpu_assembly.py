"""GPU parity of the two-phase assembly (mpmc_assembly_2p) against the oracle's
restatement of the reference's assembly on the same grid and state.

Bar: every residual and Jacobian entry within 1e-8 relative of the oracle's, relative to
max(|value|, 1e-9 x the largest entry of its row) so that entries that cancel to ~0 are
judged on their row's scale. The device evaluates each phase once in Dual<nc+1> over its
own sub-chart, the oracle re-seeds the whole 2nc+1 chart in Dual<4> blocks as the
reference does; they differ by rounding and the device libm only.
"""
import numpy as np
import pytest

import oracle_bindings as ob
from mpmc_hnu_amd import workloads as W

REL = 1e-8


def _gpu_assembly(mix, grid, q, br, prev, dt, model):
    from mpmc_hnu_amd import assembly as A, flash

    m = flash.Mixture.from_dict(mix)
    g = A.Grid(grid)
    try:
        out = A.assemble(m, g, model, q, br, prev, dt)
        res = {k: out[k].cpu().numpy() for k in ("cell_status", "residual", "diag", "offdiag")}
        res["domain_error"] = out["domain_error"]
        inv, st, flag = A.inventory(m, g, model, q, br)
        res["inventory"] = inv.cpu().numpy()
    finally:
        g.close()
        m.close()
    return res


def _close(a, b, rel=REL):
    """Row-scaled relative agreement of a[rows, ...] with b."""
    b2 = b.reshape(b.shape[0], -1)
    a2 = a.reshape(a.shape[0], -1)
    row = np.max(np.abs(b2), axis=1, keepdims=True)
    scale = np.maximum(np.abs(b2), 1e-9 * row)
    err = np.abs(a2 - b2) / np.where(scale > 0, scale, 1.0)
    return float(err.max()) <= rel, float(err.max())


@pytest.mark.gpu
@pytest.mark.parametrize("nc,dims,low_p", [(2, (4, 3, 2), False), (3, (3, 3, 3), False), (6, (5, 4, 3), False),
                                           (12, (3, 2, 2), False), (3, (4, 4, 3), True),
                                           # 4,800 cells = 75 record blocks: the strided block order and
                                           # the XCD renumbering of several blocks per XCD are exercised
                                           (6, (24, 20, 10), False)])
def test_gpu_assembly_matches_oracle(gpu, nc, dims, low_p):
    """low_p: 0.5-3 MPa, where some vapour-like cells take the largest of three roots."""
    from mpmc_hnu_amd import assembly as A

    mix = W.mixture_linspaced(nc)
    grid = W.tpfa_grid(*dims, seed=nc + 3 if low_p else nc)
    kw = dict(p_range=(0.5 * W.MPA, 3.0 * W.MPA), t_range=(290.0, 390.0)) if low_p else {}
    q, br = W.two_phase_state(nc, grid["ncell"], ob.phase_roots(mix), seed=(40 if low_p else 20) + nc, dims=dims,
                              **kw)
    if low_p:
        assert np.any(br[1] == 2)
    model = A.model_array(A.default_model(nc), nc)
    prev = ob.inventory(mix, model, q, br, grid["poro"]) * 0.98
    dt = 86400.0
    print(f"[case] assembly nc={nc} grid={dims}", flush=True)
    o = ob.assembly(mix, model, q, br, grid, prev, dt)
    assert o["rc"] == 0
    g = _gpu_assembly(mix, grid, q, br, prev, dt, A.default_model(nc))
    assert g["domain_error"] == 0 and np.all(g["cell_status"] == 0)
    Q = 2 * nc + 1
    n = grid["ncell"]
    ok, worst = _close(g["residual"].T, o["residual"].T)
    print(f"  residual worst rel {worst:.3e}", flush=True)
    assert ok, worst
    # diagonal blocks, row by row of every cell
    gd = g["diag"].reshape(Q, Q, n).transpose(2, 0, 1).reshape(n * Q, Q)
    od = o["diag"].reshape(Q, Q, n).transpose(2, 0, 1).reshape(n * Q, Q)
    ok, worst = _close(gd, od)
    print(f"  diagonal blocks worst rel {worst:.3e}", flush=True)
    assert ok, worst
    # off-diagonal blocks: slot j of cell c is CSR position a = off[c] + j <-> (face, side);
    # padding slots past a cell's degree are zero
    off, adj = A.adjacency_host(grid["owner"], grid["neighbour"], n)
    rows = nc + 1
    deg = np.diff(off)
    assert g["offdiag"].shape[1:] == (deg.max(), n)
    for j in range(deg.max()):
        assert np.all(g["offdiag"][:, j, deg <= j] == 0.0)
    go = A.ell_to_csr(g["offdiag"], off, grid["owner"].size).reshape(rows, Q, -1)
    worst_all = 0.0
    for c in range(n):
        for a in range(off[c], off[c + 1]):
            f = adj[a]
            side = 0 if grid["owner"][f] == c else 1
            ob_blk = o["offdiag"][side].reshape(rows, Q, -1)[:, :, f]
            ok, worst = _close(go[:, :, a], ob_blk)
            worst_all = max(worst_all, worst)
            assert ok, (c, f, side, worst)
    print(f"  off-diagonal blocks worst rel {worst_all:.3e}", flush=True)
    # inventories (the previous-time input)
    ok, worst = _close(g["inventory"].T, ob.inventory(mix, model, q, br, grid["poro"]).T)
    assert ok, worst


@pytest.mark.gpu
def test_newton_plan_replays_match_eager_assembly(gpu):
    """A captured Newton evaluation (mpmc_assembly_plan_*) replayed over iterates written in
    place gives, on every replay, the records the eager mpmc_assembly_2p gives on the same
    iterate; eager work and synchronisations run between the replays."""
    import torch

    from mpmc_hnu_amd import assembly as A, flash

    nc, dims = 6, (6, 5, 4)
    mix = W.mixture_linspaced(nc)
    grid = W.tpfa_grid(*dims, seed=11)
    roots = ob.phase_roots(mix)
    states = [W.two_phase_state(nc, grid["ncell"], roots, seed=70 + k, dims=dims) for k in range(3)]
    m = flash.Mixture.from_dict(mix)
    g = A.Grid(grid)
    md = A.default_model(nc)
    try:
        q0, br0 = states[0]
        inv, _, _ = A.inventory(m, g, md, q0, br0)
        prev = (inv * 0.99).contiguous()
        dev = torch.device("cuda", 0)
        qbuf = torch.as_tensor(q0, device=dev).clone()
        bbuf = torch.as_tensor(br0, device=dev).clone()
        plan = A.NewtonPlan(m, g, md, qbuf, bbuf, prev, 3600.0)
        try:
            for rep in range(6):
                q, br = states[rep % len(states)]
                qbuf.copy_(torch.as_tensor(q, device=dev))
                bbuf.copy_(torch.as_tensor(br, device=dev))
                out = plan.launch()
                torch.cuda.synchronize()
                got = {k: out[k].cpu().numpy().copy() for k in ("cell_status", "residual", "diag", "offdiag")}
                assert plan.domain_error() == 0
                ref = A.assemble(m, g, md, q, br, prev, 3600.0)
                assert ref["domain_error"] == 0
                torch.cuda.synchronize()
                for k in ("cell_status", "residual", "diag", "offdiag"):
                    assert np.array_equal(got[k], ref[k].cpu().numpy()), (rep, k)
        finally:
            plan.close()
    finally:
        g.close()
        m.close()

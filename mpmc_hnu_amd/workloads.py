"""Seeded synthetic workloads for the PR76 per-cell kernel.

Every generator is deterministic (numpy PCG64 with a fixed seed) and returns host
arrays in the library's structure-of-arrays layout: ``p[ncell]``, ``t[ncell]``,
``z[nc, ncell]``. Mixtures are plain dicts ``{tc, pc, omega, kij}`` in SI units.
"""
from __future__ import annotations

import math

import numpy as np

MPA = 1.0e6


def _dirichlet(rng: np.random.Generator, alpha: float, nc: int, ncell: int) -> np.ndarray:
    """Feed fractions z ~ Dirichlet(alpha), normalised so |sum(z)-1| stays within a few ulps."""
    if alpha == 1.0:
        g = -np.log(rng.random((nc, ncell)))
    else:
        g = rng.gamma(alpha, 1.0, size=(nc, ncell))
    g = np.maximum(g, np.finfo(np.float64).tiny)
    z = g / g.sum(axis=0, keepdims=True)
    # One correction pass on the largest entry keeps the rounded sum at 1 within 64 eps.
    resid = 1.0 - z.sum(axis=0)
    j = np.argmax(z, axis=0)
    z[j, np.arange(ncell)] += resid
    return np.ascontiguousarray(z)


def mixture_c1() -> dict:
    """nc=3 hydrocarbon-like mixture: Tc 190/425/620 K, Pc 4.6/3.8/2.1 MPa,
    omega 0.011/0.20/0.49, kij = 0 except 0.02 between components 1 and 3."""
    kij = np.zeros((3, 3))
    kij[0, 2] = kij[2, 0] = 0.02
    return {
        "tc": np.array([190.0, 425.0, 620.0]),
        "pc": np.array([4.6, 3.8, 2.1]) * MPA,
        "omega": np.array([0.011, 0.20, 0.49]),
        "kij": kij,
    }


def mixture_linspaced(nc: int, seed: int = 2) -> dict:
    """Linspaced mixture: Tc ascending over 190-800 K, Pc descending over 4.6-1.5 MPa,
    omega over 0.01-0.9; symmetric kij ~ U[0, 0.15] in the first row/column, else 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kij = np.zeros((nc, nc))
    row = rng.uniform(0.0, 0.15, size=nc)
    for j in range(1, nc):
        kij[0, j] = kij[j, 0] = row[j]
    if nc == 1:
        tc, pc, om = np.array([190.0]), np.array([4.6 * MPA]), np.array([0.01])
    else:
        tc = np.linspace(190.0, 800.0, nc)
        pc = np.linspace(4.6, 1.5, nc) * MPA
        om = np.linspace(0.01, 0.9, nc)
    return {"tc": tc, "pc": pc, "omega": om, "kij": kij}


def cells_c1(ncell: int = 4096, seed: int = 1):
    """T ~ U[300,450] K, P ~ U[2,30] MPa, z ~ Dirichlet(1,1,1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = rng.uniform(300.0, 450.0, size=ncell)
    p = rng.uniform(2.0, 30.0, size=ncell) * MPA
    z = _dirichlet(rng, 1.0, 3, ncell)
    return p, t, z


def gaussian_field(shape, sigma: float, rng: np.random.Generator) -> np.ndarray:
    """White noise convolved with a periodic Gaussian kernel (FFT), standard scale."""
    noise = rng.standard_normal(shape)
    freqs = [np.fft.fftfreq(n) for n in shape]
    k2 = np.zeros(shape)
    for axis, f in enumerate(freqs):
        s = [1] * len(shape)
        s[axis] = len(f)
        k2 = k2 + (f.reshape(s) ** 2)
    kernel = np.exp(-2.0 * (math.pi ** 2) * (sigma ** 2) * k2)
    field = np.real(np.fft.ifftn(np.fft.fftn(noise) * kernel))
    return field


def _minmax(field: np.ndarray, lo: float, hi: float) -> np.ndarray:
    f0, f1 = field.min(), field.max()
    return lo + (hi - lo) * (field - f0) / (f1 - f0)


def cells_grid(nc: int, grid=(128, 128, 64), sigma: float = 8.0, seed: int = 3, alpha: float = 2.0):
    """P, T as seeded smooth Gaussian random fields on a structured grid
    (P 5-40 MPa, T 320-420 K) and z ~ Dirichlet(alpha) per cell. Cell order is
    C-order over the grid (x slowest)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = _minmax(gaussian_field(grid, sigma, rng), 5.0 * MPA, 40.0 * MPA).ravel()
    t = _minmax(gaussian_field(grid, sigma, rng), 320.0, 420.0).ravel()
    ncell = p.size
    z = _dirichlet(rng, alpha, nc, ncell)
    return np.ascontiguousarray(p), np.ascontiguousarray(t), z


def cells_c2(ncell: int | None = None, seed: int = 3):
    """nc=6 grid workload on 128x128x64 = 1,048,576 cells; ``ncell`` takes a prefix."""
    p, t, z = cells_grid(6, seed=seed)
    if ncell is not None:
        p, t, z = p[:ncell].copy(), t[:ncell].copy(), np.ascontiguousarray(z[:, :ncell])
    return p, t, z


def cells_uniform(nc: int, ncell: int, seed: int = 5, alpha: float = 2.0):
    """P ~ U[5,40] MPa, T ~ U[320,420] K, z ~ Dirichlet(alpha): small parity batches."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = rng.uniform(5.0, 40.0, size=ncell) * MPA
    t = rng.uniform(320.0, 420.0, size=ncell)
    z = _dirichlet(rng, alpha, nc, ncell)
    return p, t, z


def mixture_sour_gas() -> dict:
    """Six-component acid gas (CO2, N2, H2S, C1, C2, C3) of Li & Firoozabadi (2012),
    Table 3, as held by the reference's three-phase fixture
    (tests/flash/pr76_three_phase/sour_gas_fixture.hpp:29-82)."""
    tc = np.array([304.211, 126.2, 373.2, 190.564, 305.322, 369.825])
    pc = np.array([73.819e5, 33.9e5, 89.4e5, 45.992e5, 48.718e5, 42.462e5])
    om = np.array([0.225, 0.039, 0.081, 0.01141, 0.10574, 0.15813])
    pairs = {(0, 1): -0.02, (0, 2): 0.12, (1, 2): 0.20, (0, 3): 0.125, (1, 3): 0.031, (2, 3): 0.10,
             (0, 4): 0.135, (1, 4): 0.042, (2, 4): 0.08, (0, 5): 0.150, (1, 5): 0.091, (2, 5): 0.08}
    kij = np.zeros((6, 6))
    for (i, j), v in pairs.items():
        kij[i, j] = kij[j, i] = v
    return {"tc": tc, "pc": pc, "omega": om, "kij": kij}


SOUR_GAS_FEED = np.array([0.5, 0.1195792770100074885969092518210906, 0.03346041255361154605487099189869971,
                          0.1167540336306079379127238069303560, 0.1797093062836135883994826060317244,
                          0.05049697052215943903601334331812921])


def sour_gas_sweep(ncell: int = 2048, seed: int = 11):
    """Cells scattered around the acid-gas three-phase point (2 MPa, 178.8 K):
    P ~ U[1,4] MPa, T ~ U[160,200] K, feed perturbed by 30% lognormal-ish noise."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = rng.uniform(1.0e6, 4.0e6, ncell)
    t = rng.uniform(160.0, 200.0, ncell)
    z = np.abs(SOUR_GAS_FEED[:, None] * (1.0 + 0.3 * rng.standard_normal((6, ncell))))
    z = z / z.sum(axis=0, keepdims=True)
    resid = 1.0 - z.sum(axis=0)
    j = np.argmax(z, axis=0)
    z[j, np.arange(ncell)] += resid
    return mixture_sour_gas(), p, t, np.ascontiguousarray(z)


def sour_gas_paths(npath: int = 64, npoint: int = 16, seed: int = 13):
    """Ordered isobaric cooling paths through the acid-gas three-phase region: path j holds
    p_j ~ U[1.5,3] MPa and T falling from 200 K to 160 K over npoint states (uneven steps),
    feed perturbed by 10% noise. Returns (mixture, p [npoint, npath], t [npoint, npath], z)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pj = rng.uniform(1.5e6, 3.0e6, npath)
    steps = rng.uniform(0.5, 1.5, (npoint, npath))
    frac = np.cumsum(steps, axis=0) - steps[0]
    frac = frac / np.maximum(frac[-1], 1e-300)
    t = 200.0 - 40.0 * frac
    p = np.broadcast_to(pj, (npoint, npath)).copy()
    z = np.abs(SOUR_GAS_FEED[:, None] * (1.0 + 0.1 * rng.standard_normal((6, npath))))
    z = z / z.sum(axis=0, keepdims=True)
    resid = 1.0 - z.sum(axis=0)
    j = np.argmax(z, axis=0)
    z[j, np.arange(npath)] += resid
    return mixture_sour_gas(), np.ascontiguousarray(p), np.ascontiguousarray(t), np.ascontiguousarray(z)


def boundary_feeds(p, t, record: dict, eps: float, ncell: int = 512):
    """Feeds just inside a two-phase boundary: for the first ``ncell`` cells of ``record``
    (a flash record: status, phase_count, x) that were accepted with two phases, the feed
    x_L + eps (x_V - x_L) at the same state. The split of such a feed converges to a
    phase fraction of order eps, the reference's phase-disappearance endpoint for eps
    below its 1e-10 minimum fraction. Returns (p, t, z)."""
    two = np.nonzero(((record["status"] & 0xF) == 0) & (record["phase_count"] == 2))[0][:ncell]
    x0 = record["x"][0][:, two]
    y0 = record["x"][1][:, two]
    z = x0 + eps * (y0 - x0)
    z = z / z.sum(axis=0, keepdims=True)
    return np.ascontiguousarray(p[two]), np.ascontiguousarray(t[two]), np.ascontiguousarray(z)


# ---------------------------------------------------------------- two-phase assembly grid
MILLIDARCY = 9.869233e-16  # m2
GRAVITY = 9.80665          # m/s2, acting along -z


def tpfa_grid(nx: int = 100, ny: int = 100, nz: int = 50, dx: float = 10.0, dy: float = 10.0, dz: float = 2.0,
              seed: int = 4, log_k_sigma: float = 1.5, thermal_conductivity: float = 2.0) -> dict:
    """Structured TPFA grid: cells c = i + nx (j + ny k); internal faces along x, y, z with
    owner the lower-index cell. ln k ~ N(ln 100 mD, sigma^2) i.i.d. per cell, porosity
    ~ U[0.1, 0.3]. Half transmissibility T_h = A (n.K.d)/(d.d) = 2 k A / h for an
    axis-aligned face (tpfa_half_transmissibility_3d.hpp:11-17), T_f = T_o T_n / (T_o + T_n);
    gdot = g . (x_n - x_o) with g = -9.80665 e_z; thermal conductance lambda A / h."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ncell = nx * ny * nz
    k = np.exp(np.log(100.0 * MILLIDARCY) + log_k_sigma * rng.standard_normal(ncell))
    poro = rng.uniform(0.1, 0.3, size=ncell)
    idx = np.arange(ncell).reshape(nz, ny, nx)
    owners, neighbours, trans, gdot, gth = [], [], [], [], []
    for axis, (h, area, step) in enumerate(((dx, dy * dz, (0, 0, 1)), (dy, dx * dz, (0, 1, 0)),
                                            (dz, dx * dy, (1, 0, 0)))):
        sl_o = tuple(slice(0, idx.shape[a] - step[a]) for a in range(3))
        sl_n = tuple(slice(step[a], idx.shape[a]) for a in range(3))
        o = idx[sl_o].ravel()
        nb = idx[sl_n].ravel()
        to = 2.0 * k[o] * area / h
        tn = 2.0 * k[nb] * area / h
        owners.append(o)
        neighbours.append(nb)
        trans.append((to * tn) / (to + tn))
        gdot.append(np.full(o.size, -GRAVITY * h if axis == 2 else 0.0))
        gth.append(np.full(o.size, thermal_conductivity * area / h))
    return {
        "dims": (nx, ny, nz), "ncell": ncell,
        "bulk": np.full(ncell, dx * dy * dz), "poro": poro,
        "owner": np.concatenate(owners).astype(np.int32), "neighbour": np.concatenate(neighbours).astype(np.int32),
        "trans": np.concatenate(trans), "gdot": np.concatenate(gdot), "gth": np.concatenate(gth),
    }


def two_phase_state(nc: int, ncell: int, roots, seed: int = 6, dims=None, p_range=(15.0 * MPA, 25.0 * MPA),
                    t_range=(330.0, 390.0)):
    """A two-phase natural-variable state per cell: p, T smooth fields over p_range / t_range
    (on `dims` when given), S_0 ~ U[0.2, 0.8], liquid-like x_0 ~ Dirichlet(3 .. 1) and
    vapour-like x_1 ~ Dirichlet(1 .. 3) over the components. ``roots(p, t, x)`` returns
    (root_count[n], root_flags[3][n]) for compositions x[nc][n]; phase 0 takes the smallest
    and phase 1 the largest root with a valid derivative. Cells without one are redrawn.
    Returns q[2nc+1][ncell] and branch[2][ncell]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if dims is not None:
        shape = (dims[2], dims[1], dims[0])
        sigma = max(0.5, min(6.0, min(shape) / 3.0))
        p = _minmax(gaussian_field(shape, sigma, rng), *p_range).ravel()
        t = _minmax(gaussian_field(shape, sigma, rng), *t_range).ravel()
    else:
        p = rng.uniform(*p_range, size=ncell)
        t = rng.uniform(*t_range, size=ncell)
    q = np.zeros((2 * nc + 1, ncell))
    branch = np.zeros((2, ncell), np.int32)
    todo = np.arange(ncell)
    a_liq = np.linspace(3.0, 1.0, nc)[:, None]
    a_vap = np.linspace(1.0, 3.0, nc)[:, None]
    for _ in range(50):
        m = todo.size
        if m == 0:
            break
        x0 = rng.gamma(np.broadcast_to(a_liq, (nc, m)))
        x0 /= x0.sum(axis=0, keepdims=True)
        x1 = rng.gamma(np.broadcast_to(a_vap, (nc, m)))
        x1 /= x1.sum(axis=0, keepdims=True)
        x0 = np.maximum(x0, 1e-6)
        x1 = np.maximum(x1, 1e-6)
        # dependent last component as the chart defines it
        x0[nc - 1] = 1.0 - x0[:nc - 1].sum(axis=0)
        x1[nc - 1] = 1.0 - x1[:nc - 1].sum(axis=0)
        good = (x0[nc - 1] > 1e-6) & (x1[nc - 1] > 1e-6)
        br = np.full((2, m), -1, np.int32)
        for ph, xx in enumerate((x0, x1)):
            cnt, flags = roots(p[todo], t[todo], np.ascontiguousarray(xx))
            valid = ((flags & 0x10) != 0) & (np.arange(3)[:, None] < np.asarray(cnt)[None, :])
            anyv = valid.any(axis=0)
            first = np.argmax(valid, axis=0)
            last = 2 - np.argmax(valid[::-1], axis=0)
            br[ph] = np.where(anyv, first if ph == 0 else last, -1)
        good &= (br >= 0).all(axis=0)
        sel = todo[good]
        q[0, sel] = p[sel]
        q[1, sel] = t[sel]
        q[2, sel] = rng.uniform(0.2, 0.8, size=sel.size)
        q[3:3 + nc - 1, sel] = x0[:nc - 1, good]
        q[3 + nc - 1:, sel] = x1[:nc - 1, good]
        branch[:, sel] = br[:, good]
        todo = todo[~good]
    if todo.size:
        raise RuntimeError(f"two_phase_state: {todo.size} cells without a valid root")
    return np.ascontiguousarray(q), branch


# ---------------------------------------------------------------- uneven multi-mixture batch
def uneven_layout(seed: int = 9, nmix: int = 256, scale: float = 1.0, nc_range=(2, 12),
                  cells_range=(1.0e3, 2.0e5)) -> list:
    """Mixtures and cell counts of the uneven batch: nc ~ U{2..12}, each mixture the
    linspaced family of its nc with its own kij row (seed 1000 + k), cells per mixture
    log-uniform on [1e3, 2e5] (times ``scale``), about 9.6M cells in all at scale 1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(nmix):
        nc = int(rng.integers(nc_range[0], nc_range[1] + 1))
        n = int(math.exp(rng.uniform(math.log(cells_range[0]), math.log(cells_range[1]))) * scale)
        out.append({"nc": nc, "mixture": mixture_linspaced(nc, seed=1000 + k), "ncell": max(n, 1)})
    return out


def uneven_base_points(layout: list, seed: int = 9, points: int = 32) -> list:
    """Per mixture, the (T, z) points whose boundary pressures seed the near-boundary
    cells: T ~ U[320, 420] K, z ~ Dirichlet(2)."""
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    return [(rng.uniform(320.0, 420.0, size=points), _dirichlet(rng, 2.0, m["nc"], points)) for m in layout]


def uneven_cells(layout: list, base: list, boundary: list, seed: int = 9, near_fraction: float = 0.3) -> list:
    """Cells of each mixture: a ``near_fraction`` share sits within +-2% in P of the
    boundary pressure of one of the mixture's base points (``boundary[k][j]``, NaN where
    none was found), at that point's T and z; the rest are P ~ U[5, 40] MPa,
    T ~ U[320, 420] K, z ~ Dirichlet(2). Returns [(p, t, z)] per mixture, cells shuffled."""
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    out = []
    for m, (tb, zb), pb in zip(layout, base, boundary):
        n, nc = m["ncell"], m["nc"]
        ok = np.nonzero(np.isfinite(pb))[0]
        n_near = int(round(near_fraction * n)) if ok.size else 0
        p = rng.uniform(5.0, 40.0, size=n) * MPA
        t = rng.uniform(320.0, 420.0, size=n)
        z = _dirichlet(rng, 2.0, nc, n)
        if n_near:
            j = ok[rng.integers(0, ok.size, size=n_near)]
            p[:n_near] = pb[j] * (1.0 + rng.uniform(-0.02, 0.02, size=n_near))
            t[:n_near] = tb[j]
            z[:, :n_near] = zb[:, j]
        perm = rng.permutation(n)
        out.append((np.ascontiguousarray(p[perm]), np.ascontiguousarray(t[perm]), np.ascontiguousarray(z[:, perm])))
    return out


def boundary_pressures(phase_count, items: list, lo: float = 0.5 * MPA, hi: float = 60.0 * MPA, grid: int = 24,
                       steps: int = 12) -> list:
    """Highest pressure in [lo, hi] at which the phase count changes, per (T, z) point, by a
    log-spaced scan and bisection on P. ``phase_count(requests)`` flashes a list of
    (item index, p[m], t[m], z[nc, m]) requests at once and returns the phase counts per
    request; ``items`` is a list of (T[m], z[nc, m]). NaN where the count never changes."""
    pgrid = np.exp(np.linspace(np.log(lo), np.log(hi), grid))
    reqs = []
    for k, (tb, zb) in enumerate(items):
        m = tb.size
        reqs.append((k, np.repeat(pgrid, m), np.tile(tb, grid), np.tile(zb, (1, grid))))
    counts = phase_count(reqs)
    lo_p, hi_p, lo_c, found = [], [], [], []
    for (k, _, _, _), c in zip(reqs, counts):
        m = items[k][0].size
        c = c.reshape(grid, m)
        change = c[1:] != c[:-1]
        has = change.any(axis=0)
        last = np.where(has, grid - 2 - np.argmax(change[::-1], axis=0), 0)
        lo_p.append(pgrid[last])
        hi_p.append(pgrid[last + 1])
        lo_c.append(c[last, np.arange(m)])
        found.append(has)
    for _ in range(steps):
        mid = [np.sqrt(a * b) for a, b in zip(lo_p, hi_p)]
        counts = phase_count([(k, mid[k], items[k][0], items[k][1]) for k in range(len(items))])
        for k, c in enumerate(counts):
            same = c == lo_c[k]
            lo_p[k] = np.where(same, mid[k], lo_p[k])
            hi_p[k] = np.where(same, hi_p[k], mid[k])
    return [np.where(f, np.sqrt(a * b), np.nan) for f, a, b in zip(found, lo_p, hi_p)]


def warm_start_hints(prev: dict, z: np.ndarray, nhint: int = 2) -> np.ndarray:
    """Warm-start hints [nhint, nc, ncell] from a previous flash result of the same cells
    (host dict of SoA arrays): hint h is the composition of published phase h. A hint is
    absent (NaN) where that phase was not published, or where its composition does not
    have exactly the feed's support."""
    nc, n = z.shape
    hints = np.full((nhint, nc, n), np.nan)
    for h in range(nhint):
        have = prev["phase_count"] > h
        x = prev["x"][h]
        ok = have & np.all(np.isfinite(x), axis=0) & np.all((x > 0) == (z > 0), axis=0)
        hints[h][:, ok] = x[:, ok]
    return hints

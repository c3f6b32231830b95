// Two-phase natural-variable residual and Jacobian on gfx950: the Newton-iteration
// assembly of TwoPhaseSnesAssemblyContext3D::evaluate_complete_assembly
// (modules/flow/discretization/petsc/include/mpmc/flow_discretization_petsc/
// two_phase_snes_assembly.hpp:550-1090) for one rank, without PETSc.
//
// Per cell, q = (p, T, S_0, x_0 independent [nc-1], x_1 independent [nc-1]), Q = 2nc+1
// unknowns (two_phase_cell_state.hpp:120-231). Rows per cell: nc component balances,
// one energy balance, nc fugacity-equilibrium rows.
//
// k_asm_cell, one lane per (cell, phase), the two phases of a cell in the two halves of
// one wave:
//   * the phase's selected-branch properties over its own sub-chart (p, T, x_p independent)
//     in one Dual<nc+1> pass (the closure's phase evaluation, then molar mass, molar
//     density, mass density, viscosity, enthalpy, internal energy: the selected-phase
//     property closure of pr76_selected_phase_property_closure.hpp:336-465 and its chart,
//     :533-820);
//   * Corey relative permeability and mobility (two_phase_transport.hpp:38-137);
//   * a record of the phase for the face pass;
//   * the equilibrium rows ln(x0/x1) + ln phi0 - ln phi1 and their Jacobian, the other
//     phase's ln phi and its p, T derivatives taken across the wave by shuffles.
// k_asm_rows, one wave per (64 cells, column chunk): the cell's columns are split into
// chunks {p, T, S_0}, then runs of x_0 columns and of x_1 columns. Per lane, one column at
// a time: the component and energy accumulation with backward Euler
// (two_phase_accumulation.hpp:18-338, component_accumulation_time.hpp:189-300,
// energy_accumulation.hpp:511-580), then every face of the cell in slot order: potential
// upwind and TPFA linearization (two_phase_transport.hpp:173-321,
// two_phase_tpfa.hpp:129-667) from both cells' records, added to this cell's rows, and
// the off-diagonal entries of that face written to the cell's slot; the first chunk also
// forms the residual. Each face is evaluated by both its cells, so there are no atomics
// and the sums are deterministic. Off-diagonal blocks are stored by (slot, cell), cell
// fastest, so every store of a wave is one contiguous 512-byte run.
#pragma once

#include "closure_device.hpp"

namespace mpmc::dev {

// Build-defined transport, caloric and rock models (the reference leaves them to
// caller-owned providers): mu = a0 + a1 rho^2, h = cp (T - T0), U_rock = cv (T - T0),
// kr_p = S_p^2.
struct AsmModelDev {
    double mu_a0, mu_a1, cp, t0, cv_rock;
    double molar_mass[12];
};

struct AsmGrid {
    int64_t ncell, nface;
    const double* bulk;        // [ncell] m3
    const double* poro;        // [ncell]
    const int32_t* owner;      // [nface]
    const int32_t* neighbour;  // [nface]
    const double* trans;       // [nface] T_f m3
    const double* gdot;        // [nface] g . (x_n - x_o), m2/s2
    const double* gth;         // [nface] thermal conductance W/K
    const int32_t* adj_off;    // [ncell + 1] CSR of cell -> adjacent faces
    const int32_t* adj_face;   // [2 nface]
    const int32_t* slot_face;  // [max_degree][ncell]: face of slot j of cell c (CSR order), -1 past the degree
    const int32_t* slot_other; // [max_degree][ncell]: the cell across that face, + kSlotNeighbour when c is
                               // the face's neighbour; -1 past the degree
    int32_t max_degree;
    int32_t block_stride;      // largest |owner - neighbour| of a face, in 64-cell blocks (>= 1)
};

constexpr int32_t kSlotNeighbour = 1 << 30;  // cell counts stay below 2^30

struct AsmIo {
    const double* q;        // [Q][ncell]
    const int32_t* branch;  // [2][ncell]
    const double* prev;     // [nc+1][ncell]
    double dt;
    int32_t* status;        // [ncell] 0 ok, 1 domain error
    double* residual;       // [Q][ncell]
    double* diag;           // [Q*Q][ncell]
    double* offdiag;        // [(nc+1)*Q][max_degree][ncell], by (slot, cell)
    double* rec;            // workspace: per-cell face records
    unsigned long long* flag; // domain-error flag (set to 1)
};

// Record per cell for the face pass, SoA over cells: per phase p, with the phase's
// sub-chart columns s = 0 (p), 1 (T), 2 + k (x_p,k):
//   lam, dlam[nc+1], dlam_dS, rho, drho[nc+1], c, dc[nc+1], h, dh[nc+1], x[nc], u, du[nc+1]
// then p, T.
template <int NC>
struct Rec {
    static constexpr int S = NC + 1;
    static constexpr int kLam = 0, kDlam = 1, kDlamS = 1 + S, kRho = 2 + S, kDrho = 3 + S, kC = 3 + 2 * S,
                         kDc = 4 + 2 * S, kH = 4 + 3 * S, kDh = 5 + 3 * S, kX = 5 + 4 * S, kU = 5 + 4 * S + NC,
                         kDu = 6 + 4 * S + NC;
    static constexpr int kPhase = kDu + S;
    static constexpr int kP = 2 * kPhase, kT = 2 * kPhase + 1, kSize = 2 * kPhase + 2;
};

// Column chunks of the face pass: the p, T, S_0 columns in runs of C0, then the x_0
// columns and the x_1 columns in runs of CW. One wave works on one chunk of 64 cells; a
// workgroup holds up to kWaves chunks of the same 64 cells. (Measured at nc = 6 on
// 500,000 cells: single-column chunks are 15% slower, three-column x chunks 40%.)
template <int NC>
struct Chunks {
    static constexpr int NX = NC - 1;
    static constexpr int C0 = 3;
    static constexpr int CW = NX <= 6 ? NX : 4;
    static constexpr int N0 = 3 / C0;
    static constexpr int PER = (NX + CW - 1) / CW;
    static constexpr int N = N0 + 2 * PER;
    static constexpr int W = CW > C0 ? CW : C0;
    static constexpr int kWaves = N < 8 ? N : 8;
    static constexpr int kGroups = (N + kWaves - 1) / kWaves;
    // chunk k -> first column c0, width cnt, phase cph (-1: both phases)
    __device__ static void decode(int k, int& c0, int& cnt, int& cph) {
        if (k < N0) {
            c0 = k * C0;
            cnt = C0;
            cph = -1;
            return;
        }
        k -= N0;
        cph = k / PER;
        const int j = k % PER;
        c0 = 3 + cph * NX + j * CW;
        cnt = NX - j * CW < CW ? NX - j * CW : CW;
    }
};
// Visiting order of the 64-cell blocks in the face pass: the block indices read as rows of
// `stride` blocks, walked column by column. Cells a face apart are at most `stride` blocks
// apart, so the blocks visited close together in time include each other's far
// neighbours (the cells one layer up and down on a structured grid), whose records are
// then still in L2 when they are read. A bijection of [0, nblk) for any stride >= 1.
__device__ __forceinline__ int64_t strided_block(int64_t t, int64_t nblk, int64_t stride) {
    if (stride <= 1 || stride >= nblk) { return t; }
    const int64_t full = nblk / stride, rem = nblk % stride;
    int64_t r, k;
    if (t < rem * (full + 1)) {
        r = t / (full + 1);
        k = t % (full + 1);
    } else {
        const int64_t u = t - rem * (full + 1);
        r = rem + u / full;
        k = u % full;
    }
    return k * stride + r;
}

// Streaming (non-temporal) store of an output that no kernel of this pass reads back, so
// that the Jacobian stream does not evict the records the face pass re-reads from L2.
__device__ __forceinline__ void put(double* p, double v) { __builtin_nontemporal_store(v, p); }

// Workgroups are dispatched round-robin over the 8 XCDs; renumber them so that each XCD
// walks one contiguous range of cells and the neighbours' records it reads stay in its L2.
__device__ __forceinline__ unsigned xcd_contiguous_block(unsigned b, unsigned nb) {
    const unsigned per = nb / 8u, rem = nb % 8u, x = b % 8u, l = b / 8u;
    return x < rem ? x * (per + 1u) + l : rem * (per + 1u) + (x - rem) * per + l;
}

// Records are stored in blocks of 64 cells, field-major inside a block: field f of cell c
// sits at ((c / 64) * kSize + f) * 64 + c % 64. One wave's accesses to a field are one
// contiguous 512-byte run, and for a given cell every field is a constant offset from
// one base address.
template <int NC>
__device__ __forceinline__ int64_t rix(int field, int64_t cell) {
    return ((cell >> 6) * Rec<NC>::kSize + field) * 64 + (cell & 63);
}

// column of q for sub-chart column s of phase p
template <int NC>
__device__ __forceinline__ int qcol(int phase, int s) {
    return s < 2 ? s : 3 + phase * (NC - 1) + (s - 2);
}

// One wave = 32 cells x 2 phases: lane l evaluates phase l / 32 of cell l % 32, so each lane
// holds one phase's Dual state. The two lanes of a cell meet through cross-lane shuffles
// for what needs both phases (status, inventories, equilibrium rows).
template <int NC, bool kInventory>
__global__ __launch_bounds__(64) void k_asm_cell(const double* __restrict__ mixc, AsmModelDev md, AsmGrid G, AsmIo io) {
    constexpr int Q = 2 * NC + 1;
    constexpr int S = NC + 1;
    using R = Rec<NC>;
    using D = Dual<S>;
    __shared__ Mix<NC> smix;
    load_mix<NC>(mixc, smix);
    __syncthreads();
    const int64_t n = G.ncell;
    const int ph = static_cast<int>(threadIdx.x) >> 5;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * 32 + (threadIdx.x & 31);
    const bool live = c < n;
    // both phases' compositions (the equilibrium rows need the other phase's x as well)
    double P = 1.0, T = 1.0, s0 = 0.0;
    double x[2][NC];
    bool ok = live;
    for (int q = 0; q < 2; ++q) {
        for (int k = 0; k < NC; ++k) { x[q][k] = 0.5; }
    }
    if (live) {
        P = io.q[c];
        T = io.q[n + c];
        s0 = io.q[2 * n + c];
        // reconstruct_composition (pr76_selected_phase_property_closure.hpp:115-151)
        for (int q = 0; q < 2; ++q) {
            double sum = 0;
            for (int k = 0; k + 1 < NC; ++k) {
                x[q][k] = io.q[(3 + q * (NC - 1) + k) * n + c];
                ok = ok && finite(x[q][k]) && x[q][k] > 0;
                sum += x[q][k];
            }
            x[q][NC - 1] = 1.0 - sum;
            ok = ok && finite(x[q][NC - 1]) && x[q][NC - 1] > 0;
        }
        ok = ok && finite(P) && P > 0 && finite(T) && T > 0;
    }
    const double sat = ph == 0 ? s0 : 1.0 - s0;
    // this lane's phase composition, selected without a run-time array index
    double xp[NC];
    for (int k = 0; k < NC; ++k) { xp[k] = ph == 0 ? x[0][k] : x[1][k]; }
    double zv = 0, zg[S], lnp[NC], J[NC * S];
    double cval = 0, rhov = 0, muv = 0, hv = 0, uv = 0;
    double dcv[S], drho[S], dmu[S], dhv[S], duv[S];
    if (ok) {
        ok = phase_dual_body<NC>(smix, P, T, xp, io.branch[ph * n + c], zv, zg, lnp, J) == PF_OK;
    }
    if (ok) {
        // molar mass, molar density p / (Z R T), mass density, mu, h, u in dual arithmetic
        D Pd = D::seed(P, 0), Td = D::seed(T, 1);
        D mm(0.0), isum(0.0);
        for (int k = 0; k + 1 < NC; ++k) {
            const D xk = D::seed(xp[k], 2 + k);
            isum = isum + xk;
            mm = mm + xk * md.molar_mass[k];
        }
        const D xl = D(1.0) - isum;
        mm = mm + xl * md.molar_mass[NC - 1];
        D z;
        z.v = zv;
        for (int k = 0; k < S; ++k) { z.d[k] = zg[k]; }
        const D den = (z * smix.r_gas) * Td;
        const D cd = Pd / den;
        const D rho = cd * mm;
        const D mu = md.mu_a0 + md.mu_a1 * (rho * rho);
        const D h = md.cp * (Td - md.t0);
        const D u = h - Pd / rho;
        ok = mm.finite() && mm.v > 0 && cd.finite() && cd.v > 0 && rho.finite() && rho.v > 0 && mu.finite() &&
             mu.v > 0 && h.finite() && u.finite();
        cval = cd.v; rhov = rho.v; muv = mu.v; hv = h.v; uv = u.v;
        for (int k = 0; k < S; ++k) {
            dcv[k] = cd.d[k]; drho[k] = rho.d[k]; dmu[k] = mu.d[k]; dhv[k] = h.d[k]; duv[k] = u.d[k];
        }
    }
    // a cell is evaluated when both of its phases are
    const int ok_other = __shfl_xor(ok ? 1 : 0, 32);
    const bool cell_ok = ok && ok_other != 0;
    if (live && ph == 0) {
        io.status[c] = cell_ok ? 0 : 1;
        if (!cell_ok) { atomicOr(io.flag, 1ull); }
    }
    if constexpr (kInventory) {
        // N_i and E at this state (previous-time inventories), into residual rows 0..nc
        const double phi = live ? G.poro[c] : 0.0;
        for (int i = 0; i < NC; ++i) {
            const double mine = phi * sat * cval * xp[i];
            const double theirs = __shfl_down(mine, 32);
            if (live && cell_ok && ph == 0) { io.residual[i * n + c] = (0.0 + mine) + theirs; }
        }
        const double mine = phi * sat * rhov * uv;
        const double theirs = __shfl_down(mine, 32);
        if (live && cell_ok && ph == 0) {
            io.residual[NC * n + c] = ((0.0 + mine) + theirs) + (1.0 - phi) * (md.cv_rock * (T - md.t0));
        }
        return;
    }
    double* rec = io.rec;
    if (live && cell_ok) {
        // mobility with Corey kr = S^2 (two_phase_transport.hpp:92-125)
        const double kr = sat * sat;
        const double dkr = ph == 0 ? 2.0 * sat : -2.0 * sat;
        const double mu = muv, mu2 = mu * mu;
        const int b = ph * R::kPhase;
        rec[rix<NC>(b + R::kLam, c)] = kr / mu;
        for (int k = 0; k < S; ++k) { rec[rix<NC>(b + R::kDlam + k, c)] = (0.0 * mu - kr * dmu[k]) / mu2; }
        rec[rix<NC>(b + R::kDlamS, c)] = (dkr * mu - kr * 0.0) / mu2;
        rec[rix<NC>(b + R::kRho, c)] = rhov;
        rec[rix<NC>(b + R::kC, c)] = cval;
        rec[rix<NC>(b + R::kH, c)] = hv;
        rec[rix<NC>(b + R::kU, c)] = uv;
        for (int k = 0; k < S; ++k) {
            rec[rix<NC>(b + R::kDrho + k, c)] = drho[k];
            rec[rix<NC>(b + R::kDc + k, c)] = dcv[k];
            rec[rix<NC>(b + R::kDh + k, c)] = dhv[k];
            rec[rix<NC>(b + R::kDu + k, c)] = duv[k];
        }
        for (int i = 0; i < NC; ++i) { rec[rix<NC>(b + R::kX + i, c)] = xp[i]; }
        if (ph == 0) {
            rec[rix<NC>(R::kP, c)] = P;
            rec[rix<NC>(R::kT, c)] = T;
        }
    }
    // equilibrium rows ln(x0_i / x1_i) + ln phi0_i - ln phi1_i (+ ln(p/p) = 0): phase-0
    // lanes write the residual and the p, T, S_0 and x_0 columns (with phase 1's ln phi
    // and its p, T derivatives from the partner lane), phase-1 lanes the x_1 columns.
    // The component and energy rows are formed by the face pass.
    for (int i = 0; i < NC; ++i) {
        const double lnp1 = __shfl_down(lnp[i], 32);
        const double j1p = __shfl_down(J[i * S + 0], 32);
        const double j1t = __shfl_down(J[i * S + 1], 32);
        if (!(live && cell_ok)) { continue; }
        const double r = x[0][i] / x[1][i];
        const int first = 3 + ph * (NC - 1);
        if (ph == 0) {
            io.residual[(NC + 1 + i) * n + c] = ((log(r) + lnp[i]) - lnp1) + 0.0;
            // p and T: both phases depend on them; S_0: neither
            io.diag[((NC + 1 + i) * Q + 0) * n + c] = ((0.0 / r + J[i * S + 0]) - j1p) + 0.0;
            io.diag[((NC + 1 + i) * Q + 1) * n + c] = ((0.0 / r + J[i * S + 1]) - j1t) + 0.0;
            io.diag[((NC + 1 + i) * Q + 2) * n + c] = ((0.0 / r + 0.0) - 0.0) + 0.0;
        }
        for (int k = 0; k + 1 < NC; ++k) {
            // column first + k: x_ph,k (and x_ph,last = 1 - sum through the dependent entry)
            const double dcomp = i == NC - 1 ? -1.0 : (k == i ? 1.0 : 0.0);
            const double dr = ph == 0 ? (dcomp - r * 0.0) / x[1][i] : (0.0 - r * dcomp) / x[1][i];
            const double g = ph == 0 ? ((dr / r + J[i * S + 2 + k]) - 0.0) + 0.0
                                     : ((dr / r + 0.0) - J[i * S + 2 + k]) + 0.0;
            io.diag[((NC + 1 + i) * Q + first + k) * n + c] = g;
        }
    }
}

template <int NC>
__global__ __launch_bounds__(64 * Chunks<NC>::kWaves) void k_asm_rows(AsmModelDev md, AsmGrid G, AsmIo io) {
    constexpr int Q = 2 * NC + 1;
    constexpr int ROWS = NC + 1;
    using R = Rec<NC>;
    using C = Chunks<NC>;
    constexpr int W = C::W;
    const int64_t n = G.ncell;
    const double* rec = io.rec;
    // wave-uniform column chunk: first column c0, width cnt, phase cph (-1: both phases);
    // consecutive workgroups (after the XCD renumbering) are the chunk groups of one block
    // of 64 cells
    const int64_t w = xcd_contiguous_block(blockIdx.x, gridDim.x);
    const int chunk = static_cast<int>(w % C::kGroups) * C::kWaves + static_cast<int>(threadIdx.x) / 64;
    if (chunk >= C::N) { return; }
    int c0 = 0, cnt = 0, cph = -1;
    C::decode(chunk, c0, cnt, cph);
    const int64_t blk = strided_block(w / C::kGroups, (n + 63) / 64, G.block_stride);
    const int64_t c = blk * 64 + static_cast<int64_t>(threadIdx.x % 64);
    if (c >= n || io.status[c] != 0) { return; }
    // sub-chart index of column col in phase ph (-1: the phase does not depend on it)
    auto sub = [](int ph, int col) {
        if (col < 2) { return col; }
        const int first = 3 + ph * (NC - 1);
        return (col >= first && col < first + NC - 1) ? 2 + (col - first) : -1;
    };
    auto dcomp = [](int ph, int i, int col) {
        const int first = 3 + ph * (NC - 1);
        if (col < first || col >= first + NC - 1) { return 0.0; }
        if (i == NC - 1) { return -1.0; }
        return col - first == i ? 1.0 : 0.0;
    };
    const int ph_lo = cph < 0 ? 0 : cph, ph_hi = cph < 0 ? 1 : cph;

    const double phi = G.poro[c];
    const double s0 = io.q[2 * n + c];
    const double sat[2] = {s0, 1.0 - s0};
    const double inv_dt = 1.0 / io.dt;
    const int64_t ostride = static_cast<int64_t>(G.max_degree) * n;
    const double inv_v = 1.0 / G.bulk[c];
    double res[ROWS];
    if (chunk == 0) {
        // accumulation residual and backward Euler (two_phase_accumulation.hpp:18-338,
        // component_accumulation_time.hpp:189-300)
        for (int i = 0; i < NC; ++i) {
            double a = 0;
            for (int ph = 0; ph < 2; ++ph) {
                const int b = ph * R::kPhase;
                a += phi * sat[ph] * rec[rix<NC>(b + R::kC, c)] * rec[rix<NC>(b + R::kX + i, c)];
            }
            res[i] = (a - io.prev[i * n + c]) * inv_dt;
        }
        double fluid = 0;
        for (int ph = 0; ph < 2; ++ph) {
            const int b = ph * R::kPhase;
            fluid += phi * sat[ph] * rec[rix<NC>(b + R::kRho, c)] * rec[rix<NC>(b + R::kU, c)];
        }
        const double e = fluid + (1.0 - phi) * (md.cv_rock * (rec[rix<NC>(R::kT, c)] - md.t0));
        res[NC] = (e - io.prev[NC * n + c]) / io.dt;
    }
    // one column at a time: its accumulation derivative, then every face in slot order
#pragma unroll 1
    for (int cc = 0; cc < W; ++cc) {
        if (cc >= cnt) { break; }
        const int col = c0 + cc;
        const bool with_residual = chunk == 0 && cc == 0;
        double acc[ROWS];
        {
            const double dsat[2] = {col == 2 ? 1.0 : 0.0, col == 2 ? -1.0 : 0.0};
            for (int i = 0; i < NC; ++i) {
                double g = 0;
                for (int ph = ph_lo; ph <= ph_hi; ++ph) {
                    const int b = ph * R::kPhase, s = sub(ph, col);
                    const double cv = rec[rix<NC>(b + R::kC, c)], xi = rec[rix<NC>(b + R::kX + i, c)];
                    const double dc = s < 0 ? 0.0 : rec[rix<NC>(b + R::kDc + s, c)];
                    g += phi * (dsat[ph] * cv * xi + sat[ph] * dc * xi + sat[ph] * cv * dcomp(ph, i, col));
                }
                acc[i] = g * inv_dt;
            }
            double g = (1.0 - phi) * (col == 1 ? md.cv_rock : 0.0);
            for (int ph = ph_lo; ph <= ph_hi; ++ph) {
                const int b = ph * R::kPhase, s = sub(ph, col);
                const double rho = rec[rix<NC>(b + R::kRho, c)], u = rec[rix<NC>(b + R::kU, c)];
                const double dr = s < 0 ? 0.0 : rec[rix<NC>(b + R::kDrho + s, c)];
                const double du = s < 0 ? 0.0 : rec[rix<NC>(b + R::kDu + s, c)];
                g += phi * (dsat[ph] * rho * u + sat[ph] * dr * u + sat[ph] * rho * du);
            }
            acc[NC] = g / io.dt;
        }
#pragma unroll 1
        for (int j = 0; j < G.max_degree; ++j) {
            const int32_t v = G.slot_other[static_cast<int64_t>(j) * n + c];
            const int32_t f = G.slot_face[static_cast<int64_t>(j) * n + c];
            double* od = io.offdiag + static_cast<int64_t>(j) * n + c + static_cast<int64_t>(col) * ostride;
            if (v < 0) {
                // padding slot past the cell's degree
                for (int r = 0; r < ROWS; ++r) { put(od + static_cast<int64_t>(r * Q) * ostride, 0.0); }
                continue;
            }
            const bool side_owner = (v & kSlotNeighbour) == 0;
            const int64_t other = v & (kSlotNeighbour - 1);
            const int64_t o = side_owner ? c : other, nb = side_owner ? other : c;
            if (io.status[other] != 0) { continue; }
            const double tf = G.trans[f], gd = G.gdot[f], gth = G.gth[f];
            // rows are scaled by this cell's bulk volume: +1/V_o as owner, -1/V_n as neighbour
            const double sv = side_owner ? inv_v : -inv_v;
            double jo[ROWS], jn[ROWS], rate[ROWS];
            for (int r = 0; r < ROWS; ++r) { jo[r] = 0; jn[r] = 0; rate[r] = 0; }
            for (int ph = ph_lo; ph <= ph_hi; ++ph) {
                const int b = ph * R::kPhase;
                // potential, upwind and flux (two_phase_transport.hpp:173-321,
                // two_phase_tpfa.hpp:231-300)
                // both cells' fields are loaded, and the upwind ones selected afterwards, so
                // that no load waits for the upwind decision
                const double rho_o = rec[rix<NC>(b + R::kRho, o)], rho_n = rec[rix<NC>(b + R::kRho, nb)];
                const double p_o = rec[rix<NC>(R::kP, o)], p_n = rec[rix<NC>(R::kP, nb)];
                const double lam_o = rec[rix<NC>(b + R::kLam, o)], lam_n = rec[rix<NC>(b + R::kLam, nb)];
                const double c_o = rec[rix<NC>(b + R::kC, o)], c_n = rec[rix<NC>(b + R::kC, nb)];
                const double h_o = rec[rix<NC>(b + R::kH, o)], h_n = rec[rix<NC>(b + R::kH, nb)];
                const int s = sub(ph, col);
                auto sd = [&](int field, int64_t cell) { return s < 0 ? 0.0 : rec[rix<NC>(b + field + s, cell)]; };
                const double dc_o = sd(R::kDc, o), dc_n = sd(R::kDc, nb);
                const double dh_o = sd(R::kDh, o), dh_n = sd(R::kDh, nb);
                const double rf = 0.5 * (rho_o + rho_n);
                const double delta = (p_n - p_o) - rf * gd;
                const bool own = !(delta > 0.0);
                const int64_t up = own ? o : nb;
                const double lamu = own ? lam_o : lam_n;
                const double flux = -tf * lamu * delta;
                const double cu = own ? c_o : c_n;
                const double rhou = own ? rho_o : rho_n;
                const double hu = own ? h_o : h_n;
                // derivatives wrt column col of the owner's and the neighbour's q
                const double dlo = (col == 2 ? rec[rix<NC>(b + R::kDlamS, o)] : 0.0) + sd(R::kDlam, o);
                const double dln = (col == 2 ? rec[rix<NC>(b + R::kDlamS, nb)] : 0.0) + sd(R::kDlam, nb);
                const double dro = sd(R::kDrho, o), drn = sd(R::kDrho, nb);
                const double dphio = -(col == 0 ? 1.0 : 0.0) - 0.5 * dro * gd;
                const double dphin = (col == 0 ? 1.0 : 0.0) - 0.5 * drn * gd;
                const double fo = -tf * (lamu * dphio + delta * (own ? dlo : 0.0));
                const double fn = -tf * (lamu * dphin + delta * (own ? 0.0 : dln));
                const double dcu = own ? dc_o : dc_n;
                const int first = 3 + ph * (NC - 1);
                for (int i = 0; i < NC; ++i) {
                    const double xi = rec[rix<NC>(b + R::kX + i, up)];
                    const double content = cu * xi;
                    double dx = 0;
                    if (col >= first && col < first + NC - 1) { dx = i == NC - 1 ? -1.0 : (col - first == i ? 1.0 : 0.0); }
                    const double dcontent = dcu * xi + cu * dx;
                    jo[i] += (own ? dcontent : 0.0) * flux + content * fo;
                    jn[i] += (own ? 0.0 : dcontent) * flux + content * fn;
                    if (with_residual) { rate[i] += content * flux; }
                }
                const double ed = rhou * hu;
                const double ded = (own ? dro : drn) * hu + rhou * (own ? dh_o : dh_n);
                jo[NC] += (own ? ded : 0.0) * flux + ed * fo;
                jn[NC] += (own ? 0.0 : ded) * flux + ed * fn;
                if (with_residual) { rate[NC] += ed * flux; }
            }
            if (with_residual) {
                // this cell's residual rows: owner +rate / V_o, neighbour -rate / V_n
                for (int i = 0; i < NC; ++i) { res[i] += rate[i] * sv; }
                const double etot = rate[NC] + gth * (rec[rix<NC>(R::kT, o)] - rec[rix<NC>(R::kT, nb)]);
                res[NC] += etot * sv;
            }
            if (col == 1) {
                jo[NC] += gth;
                jn[NC] -= gth;
            }
            // this cell's diagonal gets d/d(own q), its slot d/d(other q)
            for (int r = 0; r < ROWS; ++r) {
                acc[r] += (side_owner ? jo[r] : jn[r]) * sv;
                put(od + static_cast<int64_t>(r * Q) * ostride, (side_owner ? jn[r] : jo[r]) * sv);
            }
        }
        for (int r = 0; r < ROWS; ++r) { put(io.diag + (r * Q + col) * n + c, acc[r]); }
    }
    if (chunk == 0) {
        for (int r = 0; r < ROWS; ++r) { io.residual[r * n + c] = res[r]; }
    }
}

template <int NC>
size_t asm_record_bytes(int64_t ncell) {
    return static_cast<size_t>(Rec<NC>::kSize) * static_cast<size_t>((ncell + 63) / 64 * 64) * sizeof(double);
}

// The domain-error flag is cleared by a kernel rather than a memset so that the launch
// sequence can be captured in a graph.
static __global__ void k_asm_clear(unsigned long long* flag) {
    if (threadIdx.x == 0) { *flag = 0ull; }
}

// inventory = true: only N_i and E of every cell into io.residual rows 0..nc.
template <int NC>
hipError_t launch_assembly(const double* mixc, const AsmModelDev& md, const AsmGrid& G, const AsmIo& io, bool inventory,
                           hipStream_t s) {
    const unsigned blocks = static_cast<unsigned>((G.ncell + 31) / 32);
    hipLaunchKernelGGL(k_asm_clear, dim3(1), dim3(64), 0, s, io.flag);
    if (inventory) {
        hipLaunchKernelGGL((k_asm_cell<NC, true>), dim3(blocks), dim3(64), 0, s, mixc, md, G, io);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_asm_cell<NC, false>), dim3(blocks), dim3(64), 0, s, mixc, md, G, io);
    const unsigned row_blocks = static_cast<unsigned>((G.ncell + 63) / 64 * Chunks<NC>::kGroups);
    hipLaunchKernelGGL(k_asm_rows<NC>, dim3(row_blocks), dim3(64 * Chunks<NC>::kWaves), 0, s, md, G, io);
    return hipGetLastError();
}

#define MPMC_ASM_SIG(N)                                                                                           \
    hipError_t launch_assembly<N>(const double*, const AsmModelDev&, const AsmGrid&, const AsmIo&, bool, hipStream_t)
#define MPMC_DECLARE_ASSEMBLY(N) extern template MPMC_ASM_SIG(N);                                                  \
    extern template size_t asm_record_bytes<N>(int64_t);
#define MPMC_INSTANTIATE_ASSEMBLY(N) template MPMC_ASM_SIG(N);                                                     \
    template size_t asm_record_bytes<N>(int64_t);

} // namespace mpmc::dev

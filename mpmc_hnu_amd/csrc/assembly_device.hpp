// Two-phase natural-variable residual and Jacobian on gfx950: the Newton-iteration
// assembly of TwoPhaseSnesAssemblyContext3D::evaluate_complete_assembly
// (modules/flow/discretization/petsc/include/mpmc/flow_discretization_petsc/
// two_phase_snes_assembly.hpp:550-1090) for one rank, without PETSc.
//
// Per cell, q = (p, T, S_0, x_0 independent [nc-1], x_1 independent [nc-1]), Q = 2nc+1
// unknowns (two_phase_cell_state.hpp:120-231). Rows per cell: nc component balances,
// one energy balance, nc fugacity-equilibrium rows.
//
// k_asm_cell, one lane per cell:
//   * each phase's selected-branch properties over its own sub-chart (p, T, x_p independent)
//     in one Dual<nc+1> pass (the closure's phase_dual, then molar mass, molar density,
//     mass density, viscosity, enthalpy, internal energy: the selected-phase property
//     closure of pr76_selected_phase_property_closure.hpp:336-465 and its chart, :533-820);
//   * equilibrium rows ln(x0/x1) + ln phi0 - ln phi1 and their Jacobian;
//   * Corey relative permeability and mobility (two_phase_transport.hpp:38-137);
//   * component and energy accumulation with backward Euler (two_phase_accumulation.hpp:
//     18-338, component_accumulation_time.hpp:189-300, energy_accumulation.hpp:511-580);
//   * a compact record per phase for the face pass.
// k_asm_rows, one lane per cell: walks the cell's faces (CSR adjacency), evaluates each
// face's potential upwind and TPFA linearization (two_phase_transport.hpp:173-321,
// two_phase_tpfa.hpp:129-667) from both cells' records, and adds this cell's side:
// residual rows and diagonal block, and writes its off-diagonal block for that face.
// Each face is evaluated by both its cells, so there are no atomics and the sums are
// deterministic; the extra flops are cheap next to the Jacobian bytes that bound the pass.
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
};

struct AsmIo {
    const double* q;        // [Q][ncell]
    const int32_t* branch;  // [2][ncell]
    const double* prev;     // [nc+1][ncell]
    double dt;
    int32_t* status;        // [ncell] 0 ok, 1 domain error
    double* residual;       // [Q][ncell]
    double* diag;           // [Q*Q][ncell]
    double* offdiag;        // [(nc+1)*Q][2 nface], by CSR position
    double* rec;            // workspace: per-cell face records
    unsigned long long* flag; // domain-error flag (set to 1)
};

// Record per cell for the face pass, SoA over cells: per phase p, with the phase's
// sub-chart columns s = 0 (p), 1 (T), 2 + k (x_p,k):
//   lam, dlam[nc+1], dlam_dS, rho, drho[nc+1], c, dc[nc+1], h, dh[nc+1], x[nc]
// then p, T.
template <int NC>
struct Rec {
    static constexpr int S = NC + 1;
    static constexpr int kPhase = 1 + S + 1 + 1 + S + 1 + S + 1 + S + NC;
    static constexpr int kLam = 0, kDlam = 1, kDlamS = 1 + S, kRho = 2 + S, kDrho = 3 + S, kC = 3 + 2 * S,
                         kDc = 4 + 2 * S, kH = 4 + 3 * S, kDh = 5 + 3 * S, kX = 5 + 4 * S;
    static constexpr int kP = 2 * kPhase, kT = 2 * kPhase + 1, kSize = 2 * kPhase + 2;
};

// column of q for sub-chart column s of phase p
template <int NC>
__device__ __forceinline__ int qcol(int phase, int s) {
    return s < 2 ? s : 3 + phase * (NC - 1) + (s - 2);
}

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
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double qv[Q];
        for (int k = 0; k < Q; ++k) { qv[k] = io.q[k * n + c]; }
        const double P = qv[0], T = qv[1];
        const double phi = G.poro[c];
        double sat[2] = {qv[2], 1.0 - qv[2]};
        double x[2][NC];
        double lnp[2][NC], J[2][NC * S];
        double cval[2], rhov[2], muv[2], hv[2], uv[2];
        double dcv[2][S], drho[2][S], dmu[2][S], dhv[2][S], duv[2][S];
        bool ok = true;
        for (int ph = 0; ph < 2 && ok; ++ph) {
            // reconstruct_composition (pr76_selected_phase_property_closure.hpp:115-151)
            double sum = 0;
            for (int k = 0; k + 1 < NC; ++k) {
                x[ph][k] = qv[3 + ph * (NC - 1) + k];
                ok = ok && finite(x[ph][k]) && x[ph][k] > 0;
                sum += x[ph][k];
            }
            x[ph][NC - 1] = 1.0 - sum;
            ok = ok && finite(x[ph][NC - 1]) && x[ph][NC - 1] > 0 && finite(P) && P > 0 && finite(T) && T > 0;
            if (!ok) { break; }
            double zv, zg[S];
            if (phase_dual<NC>(smix, P, T, x[ph], io.branch[ph * n + c], zv, zg, lnp[ph], J[ph]) != PF_OK) {
                ok = false;
                break;
            }
            // molar mass, molar density p / (Z R T), mass density, mu, h, u in dual arithmetic
            D Pd = D::seed(P, 0), Td = D::seed(T, 1);
            D mm(0.0), isum(0.0);
            for (int k = 0; k + 1 < NC; ++k) {
                const D xk = D::seed(x[ph][k], 2 + k);
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
            ok = ok && mm.finite() && mm.v > 0 && cd.finite() && cd.v > 0 && rho.finite() && rho.v > 0 &&
                 mu.finite() && mu.v > 0 && h.finite() && u.finite();
            cval[ph] = cd.v; rhov[ph] = rho.v; muv[ph] = mu.v; hv[ph] = h.v; uv[ph] = u.v;
            for (int k = 0; k < S; ++k) {
                dcv[ph][k] = cd.d[k]; drho[ph][k] = rho.d[k]; dmu[ph][k] = mu.d[k]; dhv[ph][k] = h.d[k];
                duv[ph][k] = u.d[k];
            }
        }
        if (!ok) {
            io.status[c] = 1;
            atomicOr(io.flag, 1ull);
            continue;
        }
        io.status[c] = 0;
        if constexpr (kInventory) {
            // N_i and E at this state (previous-time inventories), into residual rows 0..nc
            for (int i = 0; i < NC; ++i) {
                double acc = 0;
                for (int ph = 0; ph < 2; ++ph) { acc += phi * sat[ph] * cval[ph] * x[ph][i]; }
                io.residual[i * n + c] = acc;
            }
            double fluid = 0;
            for (int ph = 0; ph < 2; ++ph) { fluid += phi * sat[ph] * rhov[ph] * uv[ph]; }
            io.residual[NC * n + c] = fluid + (1.0 - phi) * (md.cv_rock * (T - md.t0));
            continue;
        }
        double* rec = io.rec;
        // mobility with Corey kr = S^2 (two_phase_transport.hpp:92-125)
        for (int ph = 0; ph < 2; ++ph) {
            const double kr = sat[ph] * sat[ph];
            const double dkr = ph == 0 ? 2.0 * sat[0] : -2.0 * sat[1];
            const double mu = muv[ph], mu2 = mu * mu;
            const int b = ph * R::kPhase;
            rec[(b + R::kLam) * n + c] = kr / mu;
            for (int k = 0; k < S; ++k) { rec[(b + R::kDlam + k) * n + c] = (0.0 * mu - kr * dmu[ph][k]) / mu2; }
            rec[(b + R::kDlamS) * n + c] = (dkr * mu - kr * 0.0) / mu2;
            rec[(b + R::kRho) * n + c] = rhov[ph];
            rec[(b + R::kC) * n + c] = cval[ph];
            rec[(b + R::kH) * n + c] = hv[ph];
            for (int k = 0; k < S; ++k) {
                rec[(b + R::kDrho + k) * n + c] = drho[ph][k];
                rec[(b + R::kDc + k) * n + c] = dcv[ph][k];
                rec[(b + R::kDh + k) * n + c] = dhv[ph][k];
            }
            for (int i = 0; i < NC; ++i) { rec[(b + R::kX + i) * n + c] = x[ph][i]; }
        }
        rec[R::kP * n + c] = P;
        rec[R::kT * n + c] = T;
        // accumulation (two_phase_accumulation.hpp:18-160) and backward Euler
        const double inv_dt = 1.0 / io.dt;
        auto dsat = [](int ph, int col) { return col != 2 ? 0.0 : (ph == 0 ? 1.0 : -1.0); };
        auto dcomp = [](int ph, int i, int col) {
            const int first = 3 + ph * (NC - 1);
            if (col < first || col >= first + NC - 1) { return 0.0; }
            if (i == NC - 1) { return -1.0; }
            return col - first == i ? 1.0 : 0.0;
        };
        auto subd = [](const double* g, int ph, int col) {
            // q column -> sub-chart derivative of phase ph (zero outside its chart)
            if (col < 2) { return g[col]; }
            const int first = 3 + ph * (NC - 1);
            if (col < first || col >= first + NC - 1) { return 0.0; }
            return g[2 + (col - first)];
        };
        for (int i = 0; i < NC; ++i) {
            double acc = 0;
            for (int ph = 0; ph < 2; ++ph) { acc += phi * sat[ph] * cval[ph] * x[ph][i]; }
            io.residual[i * n + c] = (acc - io.prev[i * n + c]) * inv_dt;
        }
        for (int col = 0; col < Q; ++col) {
            for (int i = 0; i < NC; ++i) {
                double g = 0;
                for (int ph = 0; ph < 2; ++ph) {
                    const double s = sat[ph], ds = dsat(ph, col), cc = cval[ph], dc = subd(dcv[ph], ph, col);
                    g += phi * (ds * cc * x[ph][i] + s * dc * x[ph][i] + s * cc * dcomp(ph, i, col));
                }
                io.diag[(i * Q + col) * n + c] = g * inv_dt;
            }
        }
        {
            double fluid = 0;
            for (int ph = 0; ph < 2; ++ph) { fluid += phi * sat[ph] * rhov[ph] * uv[ph]; }
            const double e = fluid + (1.0 - phi) * (md.cv_rock * (T - md.t0));
            io.residual[NC * n + c] = (e - io.prev[NC * n + c]) / io.dt;
            for (int col = 0; col < Q; ++col) {
                double g = (1.0 - phi) * (col == 1 ? md.cv_rock : 0.0);
                for (int ph = 0; ph < 2; ++ph) {
                    const double s = sat[ph], ds = dsat(ph, col), rho = rhov[ph], dr = subd(drho[ph], ph, col);
                    const double u = uv[ph], du = subd(duv[ph], ph, col);
                    g += phi * (ds * rho * u + s * dr * u + s * rho * du);
                }
                io.diag[(NC * Q + col) * n + c] = g / io.dt;
            }
        }
        // equilibrium rows: ln(x0_i / x1_i) + ln phi0_i - ln phi1_i (+ ln(p/p) = 0)
        for (int i = 0; i < NC; ++i) {
            const double r = x[0][i] / x[1][i];
            io.residual[(NC + 1 + i) * n + c] = ((log(r) + lnp[0][i]) - lnp[1][i]) + 0.0;
            for (int col = 0; col < Q; ++col) {
                const double dr = (dcomp(0, i, col) - r * dcomp(1, i, col)) / x[1][i];
                const double g = ((dr / r + subd(&J[0][i * S], 0, col)) - subd(&J[1][i * S], 1, col)) + 0.0;
                io.diag[((NC + 1 + i) * Q + col) * n + c] = g;
            }
        }
    }
}

template <int NC>
__global__ __launch_bounds__(64) void k_asm_rows(AsmGrid G, AsmIo io) {
    constexpr int Q = 2 * NC + 1;
    constexpr int S = NC + 1;
    constexpr int ROWS = NC + 1;
    using R = Rec<NC>;
    const int64_t n = G.ncell;
    const int64_t nadj = 2 * G.nface;
    const double* rec = io.rec;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (io.status[c] != 0) { continue; }
        double res[ROWS];
        double dg[ROWS * Q];
        for (int r = 0; r < ROWS; ++r) {
            res[r] = io.residual[r * n + c];
            for (int col = 0; col < Q; ++col) { dg[r * Q + col] = io.diag[(r * Q + col) * n + c]; }
        }
        const double vc = G.bulk[c];
        for (int32_t a = G.adj_off[c]; a < G.adj_off[c + 1]; ++a) {
            const int32_t f = G.adj_face[a];
            const int64_t o = G.owner[f], nb = G.neighbour[f];
            const bool side_owner = o == c;
            const int64_t other = side_owner ? nb : o;
            if (io.status[other] != 0) { continue; }
            const double tf = G.trans[f], gd = G.gdot[f], gth = G.gth[f];
            const double vo = G.bulk[o], vn = G.bulk[nb];
            // per phase: potential, upwind, flux (two_phase_transport.hpp:173-321,
            // two_phase_tpfa.hpp:231-300)
            double flux[2], lamu[2], delta[2];
            bool own[2];
            for (int ph = 0; ph < 2; ++ph) {
                const int b = ph * R::kPhase;
                const double rf = 0.5 * (rec[(b + R::kRho) * n + o] + rec[(b + R::kRho) * n + nb]);
                delta[ph] = (rec[R::kP * n + nb] - rec[R::kP * n + o]) - rf * gd;
                own[ph] = !(delta[ph] > 0.0);
                lamu[ph] = rec[(b + R::kLam) * n + (own[ph] ? o : nb)];
                flux[ph] = -tf * lamu[ph] * delta[ph];
            }
            double rate[NC];
            for (int i = 0; i < NC; ++i) {
                double r = 0;
                for (int ph = 0; ph < 2; ++ph) {
                    const int b = ph * R::kPhase;
                    const int64_t up = own[ph] ? o : nb;
                    r += (rec[(b + R::kC) * n + up] * rec[(b + R::kX + i) * n + up]) * flux[ph];
                }
                rate[i] = r;
            }
            double erate = 0;
            for (int ph = 0; ph < 2; ++ph) {
                const int b = ph * R::kPhase;
                const int64_t up = own[ph] ? o : nb;
                erate += (rec[(b + R::kRho) * n + up] * rec[(b + R::kH) * n + up]) * flux[ph];
            }
            const double etot = erate + gth * (rec[R::kT * n + o] - rec[R::kT * n + nb]);
            // this cell's rows: owner +rate / V_o, neighbour -rate / V_n
            for (int i = 0; i < NC; ++i) { res[i] += side_owner ? rate[i] / vo : -rate[i] / vn; }
            res[NC] += side_owner ? etot / vo : -etot / vn;
            // Jacobian columns of both cells: derivative of the face rates wrt the owner's
            // q (jo) and the neighbour's q (jn); this cell's diagonal gets d/d(own q) and its
            // off-diagonal block d/d(other q)
            for (int col = 0; col < Q; ++col) {
                double jo[ROWS], jn[ROWS];
                for (int r = 0; r < ROWS; ++r) { jo[r] = 0; jn[r] = 0; }
                for (int ph = 0; ph < 2; ++ph) {
                    const int b = ph * R::kPhase;
                    // sub-chart derivative of the phase's record fields wrt q column col
                    const int first = 3 + ph * (NC - 1);
                    const int s = col < 2 ? col : ((col >= first && col < first + NC - 1) ? 2 + (col - first) : -1);
                    auto sd = [&](int field, int64_t cell) { return s < 0 ? 0.0 : rec[(b + field + s) * n + cell]; };
                    const double dlo = (col == 2 ? rec[(b + R::kDlamS) * n + o] : 0.0) + (s < 0 ? 0.0 : rec[(b + R::kDlam + s) * n + o]);
                    const double dln = (col == 2 ? rec[(b + R::kDlamS) * n + nb] : 0.0) + (s < 0 ? 0.0 : rec[(b + R::kDlam + s) * n + nb]);
                    const double dphio = -(col == 0 ? 1.0 : 0.0) - 0.5 * sd(R::kDrho, o) * gd;
                    const double dphin = (col == 0 ? 1.0 : 0.0) - 0.5 * sd(R::kDrho, nb) * gd;
                    const double fo = -tf * (lamu[ph] * dphio + delta[ph] * (own[ph] ? dlo : 0.0));
                    const double fn = -tf * (lamu[ph] * dphin + delta[ph] * (own[ph] ? 0.0 : dln));
                    const int64_t up = own[ph] ? o : nb;
                    const double cc = rec[(b + R::kC) * n + up];
                    const double dcu = sd(R::kDc, up);
                    for (int i = 0; i < NC; ++i) {
                        const double xi = rec[(b + R::kX + i) * n + up];
                        const double content = cc * xi;
                        double dx = 0;
                        if (col >= first && col < first + NC - 1) { dx = i == NC - 1 ? -1.0 : (col - first == i ? 1.0 : 0.0); }
                        const double dcontent = dcu * xi + cc * dx;
                        jo[i] += (own[ph] ? dcontent : 0.0) * flux[ph] + content * fo;
                        jn[i] += (own[ph] ? 0.0 : dcontent) * flux[ph] + content * fn;
                    }
                    const double rho = rec[(b + R::kRho) * n + up], h = rec[(b + R::kH) * n + up];
                    const double ed = rho * h;
                    const double ded = sd(R::kDrho, up) * h + rho * sd(R::kDh, up);
                    jo[NC] += (own[ph] ? ded : 0.0) * flux[ph] + ed * fo;
                    jn[NC] += (own[ph] ? 0.0 : ded) * flux[ph] + ed * fn;
                }
                if (col == 1) {
                    jo[NC] += gth;
                    jn[NC] -= gth;
                }
                for (int r = 0; r < ROWS; ++r) {
                    double dself, doff;
                    if (side_owner) {
                        dself = jo[r] / vo;
                        doff = jn[r] / vo;
                    } else if (r < NC) {
                        dself = -jn[r] / vn;
                        doff = -jo[r] / vn;
                    } else {
                        dself = jn[r] * (-1.0 / vn);
                        doff = jo[r] * (-1.0 / vn);
                    }
                    dg[r * Q + col] += dself;
                    io.offdiag[static_cast<int64_t>(r * Q + col) * nadj + a] = doff;
                }
            }
        }
        (void)vc;
        for (int r = 0; r < ROWS; ++r) {
            io.residual[r * n + c] = res[r];
            for (int col = 0; col < Q; ++col) { io.diag[(r * Q + col) * n + c] = dg[r * Q + col]; }
        }
    }
}

template <int NC>
size_t asm_record_bytes(int64_t ncell) {
    return static_cast<size_t>(Rec<NC>::kSize) * static_cast<size_t>(ncell) * sizeof(double);
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
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((G.ncell + 63) / 64, 65536));
    hipLaunchKernelGGL(k_asm_clear, dim3(1), dim3(64), 0, s, io.flag);
    if (inventory) {
        hipLaunchKernelGGL((k_asm_cell<NC, true>), dim3(blocks), dim3(64), 0, s, mixc, md, G, io);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_asm_cell<NC, false>), dim3(blocks), dim3(64), 0, s, mixc, md, G, io);
    hipLaunchKernelGGL(k_asm_rows<NC>, dim3(blocks), dim3(64), 0, s, G, io);
    return hipGetLastError();
}

#define MPMC_ASM_SIG(N)                                                                                           \
    hipError_t launch_assembly<N>(const double*, const AsmModelDev&, const AsmGrid&, const AsmIo&, bool, hipStream_t)
#define MPMC_DECLARE_ASSEMBLY(N) extern template MPMC_ASM_SIG(N);                                                  \
    extern template size_t asm_record_bytes<N>(int64_t);
#define MPMC_INSTANTIATE_ASSEMBLY(N) template MPMC_ASM_SIG(N);                                                     \
    template size_t asm_record_bytes<N>(int64_t);

} // namespace mpmc::dev

// Schur reduction of the two-phase Newton system to (nc+1) x (nc+1) blocks.
//
// The unreduced system (launch_assembly) has Q = 2nc+1 unknowns and rows per cell: nc+1
// conservation rows (components, energy) that couple to the neighbours across faces, and nc
// fugacity-equilibrium rows that involve the cell's own unknowns only
// (two_phase_snes_assembly.hpp:550-1090 assembles exactly these). Splitting each cell's
// unknowns into nc+1 primaries P and nc secondaries S, the equilibrium rows
//     J_EP dx_P + J_ES dx_S = -r_E
// give dx_S = -(M_P dx_P + M_r) with M = J_ES^-1 [J_EP | r_E] (nc x (nc+2) per cell), and the
// conservation rows become the reduced block system
//     A_cc = J_CP(c) - J_CS(c) M_P(c),   A_cn = O_CP(c,n) - O_CS(c,n) M_P(n),
//     b_c  = r_C(c) - J_CS(c) M_r(c) - sum_n O_CS(c,n) M_r(n),      A dx_P = -b,
// which is what a solver on (nc+1)^2 blocks takes; the secondaries follow cell by cell from
// M. No approximation is made: the reduced and unreduced systems have the same solution.
//
// Layouts (cell index fastest): M in back[S][P+1][n] (column P is M_r), the reduced
// residual[P][n], diag[P*P][n] (row-major), offdiag[P*P][max_degree][n] with the slots of
// the unreduced ELL layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mpmc::dev {

// Column partition of a cell's q: the nc secondary columns, then the nc+1 primary ones,
// each list in increasing q order.
struct RedColumns {
    int32_t sec[12];
    int32_t prim[13];
};

struct RedIo {
    int64_t n;
    int32_t max_degree;
    const int32_t* slot_other; // [max_degree][n]: cell across the slot's face (+ 2^30 when the
                               // cell is the face's owner), -1 past the cell's degree
    const double* residual;    // unreduced [Q][n]
    const double* diag;        // [Q*Q][n]
    const double* offdiag;     // [(nc+1)*Q][max_degree][n]
    double* back;              // [nc][nc+2][n]
    double* rres;              // [nc+1][n]
    double* rdiag;             // [(nc+1)^2][n]
    double* roff;              // [(nc+1)^2][max_degree][n]
    int32_t* status;           // [n] 0, or 1 where J_ES is singular to working precision
};

// Per cell: LU of J_ES with partial pivoting (row swaps as value selects, so every array
// index is a compile-time constant and nothing is kept in private memory), M, the reduced
// diagonal block and the cell's own part of the reduced residual.
template <int NC>
__global__ __launch_bounds__(64) void k_red_cell(RedColumns cols, RedIo io) {
    constexpr int Q = 2 * NC + 1, P = NC + 1, S = NC;
    const int64_t n = io.n;
    const double nan = __builtin_nan("");
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        // a = J_ES (rows: equilibrium rows P.., columns: secondaries)
        double a[S][S];
        double scale = 0;
#pragma unroll
        for (int i = 0; i < S; ++i) {
#pragma unroll
            for (int j = 0; j < S; ++j) {
                a[i][j] = io.diag[((P + i) * Q + cols.sec[j]) * n + c];
                scale = __builtin_fmax(scale, __builtin_fabs(a[i][j]));
            }
        }
        int perm[S]; // perm[k]: the row swapped with row k at step k
        bool singular = !(scale > 0) || !__builtin_isfinite(scale);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            int p = k;
            double best = __builtin_fabs(a[k][k]);
#pragma unroll
            for (int i = k + 1; i < S; ++i) {
                const double v = __builtin_fabs(a[i][k]);
                const bool take = v > best;
                best = take ? v : best;
                p = take ? i : p;
            }
            perm[k] = p;
#pragma unroll
            for (int i = k + 1; i < S; ++i) {
                const bool sw = i == p;
#pragma unroll
                for (int j = 0; j < S; ++j) {
                    const double t = a[k][j];
                    a[k][j] = sw ? a[i][j] : a[k][j];
                    a[i][j] = sw ? t : a[i][j];
                }
            }
            // pivot guard of the reference's LU (64 eps n |A|, pt_sensitivity_detail.hpp:152-226)
            singular = singular || !(__builtin_fabs(a[k][k]) > 64.0 * 2.220446049250313e-16 * S * scale);
            const double inv = 1.0 / a[k][k];
#pragma unroll
            for (int i = k + 1; i < S; ++i) {
                const double f = a[i][k] * inv;
                a[i][k] = f;
#pragma unroll
                for (int j = k + 1; j < S; ++j) { a[i][j] = __builtin_fma(-f, a[k][j], a[i][j]); }
            }
        }
        io.status[c] = singular ? 1 : 0;
        // J_CS of the cell, kept while the columns of M are formed
        double jcs[P][S];
#pragma unroll
        for (int r = 0; r < P; ++r) {
#pragma unroll
            for (int s = 0; s < S; ++s) { jcs[r][s] = io.diag[(r * Q + cols.sec[s]) * n + c]; }
        }
        // columns k < P: J_EP[:, prim[k]]; column P: r_E
#pragma unroll
        for (int k = 0; k <= P; ++k) {
            double m[S];
#pragma unroll
            for (int i = 0; i < S; ++i) {
                m[i] = k < P ? io.diag[((P + i) * Q + cols.prim[k < P ? k : 0]) * n + c] : io.residual[(P + i) * n + c];
            }
            // the row swaps of the factorisation, in order
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int i = s + 1; i < S; ++i) {
                    const bool sw = perm[s] == i;
                    const double t = m[s];
                    m[s] = sw ? m[i] : m[s];
                    m[i] = sw ? t : m[i];
                }
            }
#pragma unroll
            for (int i = 1; i < S; ++i) {
#pragma unroll
                for (int j = 0; j < i; ++j) { m[i] = __builtin_fma(-a[i][j], m[j], m[i]); }
            }
#pragma unroll
            for (int i = S - 1; i >= 0; --i) {
#pragma unroll
                for (int j = i + 1; j < S; ++j) { m[i] = __builtin_fma(-a[i][j], m[j], m[i]); }
                m[i] = m[i] / a[i][i];
            }
#pragma unroll
            for (int i = 0; i < S; ++i) {
                m[i] = singular ? nan : m[i];
                io.back[(i * (P + 1) + k) * n + c] = m[i];
            }
            // column k of the reduced diagonal block, or the cell's part of the residual
#pragma unroll
            for (int r = 0; r < P; ++r) {
                double v = k < P ? io.diag[(r * Q + cols.prim[k < P ? k : 0]) * n + c] : io.residual[r * n + c];
#pragma unroll
                for (int s = 0; s < S; ++s) { v = __builtin_fma(-jcs[r][s], m[s], v); }
                if (k < P) { io.rdiag[(r * P + k) * n + c] = v; }
                else { io.rres[r * n + c] = v; }
            }
        }
    }
}

// Per (cell, slot): the reduced off-diagonal block against the neighbour's M, and the
// neighbours' parts of the reduced residual. Runs after k_red_cell has written every M.
template <int NC>
__global__ __launch_bounds__(64) void k_red_rows(RedColumns cols, RedIo io) {
    constexpr int Q = 2 * NC + 1, P = NC + 1, S = NC;
    const int64_t n = io.n;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double b[P];
#pragma unroll
        for (int r = 0; r < P; ++r) { b[r] = io.rres[r * n + c]; }
        for (int j = 0; j < io.max_degree; ++j) {
            const int32_t other = io.slot_other[j * n + c];
            const int64_t base = static_cast<int64_t>(j) * n + c;
            if (other < 0) {
#pragma unroll
                for (int e = 0; e < P * P; ++e) { io.roff[e * io.max_degree * n + base] = 0.0; }
                continue;
            }
            const int64_t nb = other & 0x3FFFFFFF;
            double ocs[P][S];
#pragma unroll
            for (int r = 0; r < P; ++r) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    ocs[r][s] = io.offdiag[(static_cast<int64_t>(r * Q + cols.sec[s]) * io.max_degree) * n + base];
                }
            }
#pragma unroll
            for (int k = 0; k <= P; ++k) {
                double m[S];
#pragma unroll
                for (int s = 0; s < S; ++s) { m[s] = io.back[(s * (P + 1) + k) * n + nb]; }
#pragma unroll
                for (int r = 0; r < P; ++r) {
                    double v = k < P ? io.offdiag[(static_cast<int64_t>(r * Q + cols.prim[k < P ? k : 0]) *
                                                   io.max_degree) * n + base]
                                     : 0.0;
#pragma unroll
                    for (int s = 0; s < S; ++s) { v = __builtin_fma(-ocs[r][s], m[s], v); }
                    if (k < P) { io.roff[(static_cast<int64_t>(r * P + k) * io.max_degree) * n + base] = v; }
                    else { b[r] += v; }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < P; ++r) { io.rres[r * n + c] = b[r]; }
    }
}

} // namespace mpmc::dev

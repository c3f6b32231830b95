/*
 * cpuref — the oracle: an FP64 CPU restatement of the reference's PR76 flash path.
 *
 * Test infrastructure only. Loaded by tests/ and by bench.py's CPU-baseline leg;
 * the product library never links it. Result layout is the host-memory twin of
 * include/mpmc_hnu_amd.h (structure of arrays, cell index fastest).
 */
#ifndef MPMC_CPUREF_H
#define MPMC_CPUREF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* max_phases: 2 = VLE path (solve_pr76_pt_vle), 3 = max3 path (solve_pr76_pt_max3).
 * threads <= 0: all hardware threads. Returns 0 on success. */
int cpuref_flash_batch(int nc, const double* tc, const double* pc, const double* omega,
                       const double* kij, int64_t ncell, const double* p, const double* t,
                       const double* z, double fugacity_tolerance, int max_phases, int threads,
                       int32_t* status, int32_t* phase_count, double* beta, double* zfactor,
                       int32_t* branch, double* x, double* ln_phi, uint32_t* counters);

/* cpuref_flash_batch with warm-start hints: hints[(h*nc + i)*ncell + c] (h < nhint <= 6)
 * appended to each cell's initial stability starts; NaN in component 0: no hint. */
int cpuref_flash_batch_hinted(int nc, const double* tc, const double* pc, const double* omega,
                              const double* kij, int64_t ncell, const double* p, const double* t,
                              const double* z, int nhint, const double* hints, double fugacity_tolerance,
                              int max_phases, int threads, int32_t* status, int32_t* phase_count,
                              double* beta, double* zfactor, int32_t* branch, double* x, double* ln_phi,
                              uint32_t* counters);

/* cpuref_flash_batch_hinted plus final-review hints: final_hints[(h*nc + i)*ncell + c]
 * (h < nfinal <= 6) placed before the pair's compositions in the common-tangent review
 * (final_starts); NaN in component 0: no hint. */
int cpuref_flash_batch_warm(int nc, const double* tc, const double* pc, const double* omega,
                            const double* kij, int64_t ncell, const double* p, const double* t,
                            const double* z, int nhint, const double* hints, int nfinal, const double* final_hints,
                            double fugacity_tolerance, int max_phases, int threads, int32_t* status,
                            int32_t* phase_count, double* beta, double* zfactor, int32_t* branch, double* x,
                            double* ln_phi, uint32_t* counters);

/* solve_pr76_pt_continuation over npath ordered paths of npoint states each: p, t and the
 * outputs are npoint consecutive cell batches of npath cells (point-major); z[i*npath + j] is
 * path j's feed. carry bit 0: carry the previous accepted point's phase compositions as
 * stability starts; bit 1: carry its three-phase state. fi [nfi][nc] / ff [nff][nc]: fallback
 * initial / final starts (<= 3 each); has_start: fixed three-phase start start[3][nc] with
 * seed[2]. info[k*npath + j] = incoming hint phase count | carried start count << 4 |
 * carried three-phase start << 8. */
int cpuref_pt_continuation(int nc, const double* tc, const double* pc, const double* omega,
                           const double* kij, int64_t npath, int npoint, const double* p, const double* t,
                           const double* z, int carry, int nfi, const double* fi, int nff, const double* ff,
                           int has_start, const double* start, const double* seed, double fugacity_tolerance,
                           int threads, int32_t* status, int32_t* phase_count, double* beta, double* zfactor,
                           int32_t* branch, double* x, double* ln_phi, uint32_t* counters, int32_t* info);

int cpuref_phase_batch(int nc, const double* tc, const double* pc, const double* omega,
                       const double* kij, int64_t ncell, const double* p, const double* t,
                       const double* x, int32_t* root_status, int32_t* root_count, double* zroot,
                       int32_t* root_flags, double* ln_phi);

/* Rachford-Rice at fixed log K for one feed: returns the RachfordRiceStatus code and
 * writes beta, x[nc], y[nc], iterations. */
int cpuref_rachford_rice(int nc, const double* z, const double* log_k, double* beta, double* x,
                         double* y, int32_t* iterations);

/* Two-phase (VLE) flash of each cell followed by build_pr76_pt_vle_closure and
 * differentiate_pr76_pt_vle: status primal | linearization << 4 | reason << 8 (0x2000 when
 * the reference would throw), primal [4][n] (Z_L, Z_V, rho_L, rho_V), jac [2nc+5][nc+1][n]
 * (beta_V, x_L[nc], y_V[nc], Z_L, Z_V, rho_L, rho_V over q = p, T, z_0..z_{nc-2}),
 * log_k_jac [nc][nc+1][n] and diag [3][n] (rcond, backward error, residual) may be NULL. */
int cpuref_closure_batch(int nc, const double* tc, const double* pc, const double* omega,
                         const double* kij, int64_t ncell, const double* p, const double* t,
                         const double* z, double fugacity_tolerance, double min_fraction,
                         double min_rcond, int threads, int32_t* status, double* primal, double* jac,
                         double* log_k_jac, double* diag);

/* Two-phase natural-variable residual and Jacobian (TwoPhaseSnesAssemblyContext3D::
 * evaluate_complete_assembly) on a list of cells and internal faces. model = {mu_a0, mu_a1,
 * cp, T0, cv_rock, molar_mass[nc]}; q[2nc+1][ncell]; branch[2][ncell] selected root index
 * per phase; prev[nc+1][ncell] previous N_i and E; outputs residual[Q][ncell],
 * diag[Q*Q][ncell] (row-major block), offdiag[2][(nc+1)*Q][nface] (owner row x neighbour
 * column, neighbour row x owner column). Returns 0, 2 on a cell domain error. */
int cpuref_assembly(int nc, const double* tc, const double* pc, const double* omega, const double* kij,
                    const double* model, int64_t ncell, const double* q, const int32_t* branch,
                    const double* bulk, const double* poro, const double* prev, double dt, int64_t nface,
                    const int32_t* owner, const int32_t* neighbour, const double* trans, const double* gdot,
                    const double* gth, int threads, int32_t* cell_status, double* residual, double* diag,
                    double* offdiag);
int cpuref_accumulation(int nc, const double* tc, const double* pc, const double* omega, const double* kij,
                        const double* model, int64_t ncell, const double* q, const int32_t* branch,
                        const double* poro, int threads, double* acc);

#ifdef __cplusplus
}
#endif
#endif

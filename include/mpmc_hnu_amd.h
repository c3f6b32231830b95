/*
 * mpmc_hnu_amd — C ABI of the MI355X (gfx950) PR76 per-cell compositional kernel.
 *
 * Plain pointers and sizes only. Every function returns an mpmc_rc code and never
 * throws. Device buffers are caller owned; mixtures are library-owned device
 * residents addressed by opaque handles. Calls are ordered on the HIP stream that
 * is passed as `stream` (NULL = the null stream).
 *
 * Layouts (all device memory, structure of arrays, cell index fastest):
 *   p[ncell], t[ncell]                  Pa, K
 *   z[nc][ncell]                        feed mole fractions, z[i*ncell + c]
 *   per-phase scalars  v[3][ncell]      v[ph*ncell + c]
 *   per-phase vectors  x[3][nc][ncell]  x[(ph*nc + i)*ncell + c]
 */
#ifndef MPMC_HNU_AMD_H
#define MPMC_HNU_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPMC_MAX_COMPONENTS 12
#define MPMC_MAX_PHASES 3

/* Return codes. */
enum mpmc_rc {
    MPMC_OK = 0,
    MPMC_E_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument / length_error */
    MPMC_E_DOMAIN = 2,           /* reference: std::domain_error / range_error (input data) */
    MPMC_E_DEVICE = 3,           /* HIP runtime failure */
    MPMC_E_NO_MEMORY = 4,
    MPMC_E_NOT_BUILT = 5         /* no gfx950 code object for this device */
};

/* Per-cell outcome codes, mirroring the reference enums value by value. */
/* PtPhaseSetStatus (modules/flash/include/mpmc/flash/pt_phase_set.hpp:36-40) */
enum mpmc_phase_set_status { MPMC_SET_ACCEPTED = 0, MPMC_SET_PHASE_SET_UNSTABLE = 1,
                             MPMC_SET_INDETERMINATE = 2 };
/* Pr76PtMax3Status (modules/flash/include/mpmc/flash/pr76_three_phase.hpp:133-140) */
enum mpmc_max3_status { MPMC_MAX3_SINGLE = 0, MPMC_MAX3_TWO = 1, MPMC_MAX3_THREE = 2,
                        MPMC_MAX3_BOUNDARY_UNRESOLVED = 3, MPMC_MAX3_HIGHER = 4,
                        MPMC_MAX3_INDETERMINATE = 5 };
/* PtSplitStatus (modules/flash/include/mpmc/flash/pt_split.hpp:71-74) */
enum mpmc_split_status { MPMC_SPLIT_SINGLE = 0, MPMC_SPLIT_TWO = 1, MPMC_SPLIT_UNSTABLE = 2,
                         MPMC_SPLIT_INDETERMINATE = 3 };

/* status word per cell: set | max3 << 4 | split << 8 | flags << 12 */
#define MPMC_STATUS_SET(s) ((s) & 0xF)
#define MPMC_STATUS_MAX3(s) (((s) >> 4) & 0xF)
#define MPMC_STATUS_SPLIT(s) (((s) >> 8) & 0xF)
/* flag: the reference rejects this cell's input with an exception (invalid_argument /
 * domain_error); no phase set is published. */
#define MPMC_FLAG_REJECTED 0x2000
/* flags of the search behind the published set, the facts the reference's transition
 * evidence reads from its solve result (project_pr76_pt_max3_transition_report and
 * project_pr76_pt_transition_report, pr76_pt_flash_backend.hpp:42-164):
 *   SPLIT_DISAPPEARANCE  an attempt of the base two-phase split ended at a phase-fraction
 *                        endpoint (PtSplitAttemptStatus::phase_disappearance)
 *   SPLIT_CANDIDATE      the base split selected a converged pair (PtSplitResult::candidate())
 *   MAX3_ATTEMPTED       the three-phase stage ran at least one attempt
 *                        (Pr76PtMax3Result::attempts not empty)
 *   START_CONSUMED       an attempt ran from the three-phase start a continuation carried
 *                        from the previous path point (carried_three_phase_start_consumed,
 *                        pr76_pt_continuation.hpp:183-190,256-259) */
#define MPMC_FLAG_SPLIT_DISAPPEARANCE 0x4000
#define MPMC_FLAG_SPLIT_CANDIDATE 0x8000
#define MPMC_FLAG_MAX3_ATTEMPTED 0x10000
#define MPMC_FLAG_START_CONSUMED 0x20000

/* branch word per phase: provider root index | smooth << 8 ; -1 when the slot is empty */
#define MPMC_BRANCH_INDEX(b) ((b) & 0xFF)
#define MPMC_BRANCH_SMOOTH(b) (((b) >> 8) & 0x1)

/* counters per cell, counters[k*ncell + c]:
 *   k=0 E_prop : EOS phase evaluations (one mixing + cubic + ln(phi) at one composition)
 *   k=1 E_tpd  : tangent-plane-distance evaluations
 *   k=2 I_rr   : Rachford-Rice bisection iterations
 *   k=3 I_rr3  : three-phase generalized RR Newton iterations */
#define MPMC_NUM_COUNTERS 4

typedef struct mpmc_mixture* mpmc_mixture_t;

/*
 * Prepare one PR76 mixture and keep its constants resident on the current device.
 * Replaces PrParameterSet::create + Pr76Phase::from_parameters
 * (modules/thermodynamics/include/mpmc/thermodynamics/pr_parameters.hpp:53-154,
 *  pr76_mixture.hpp:229-251, pr76_pure.hpp:181-199). kij is dense row-major nc*nc and
 * must be symmetric with a zero diagonal.
 */
int mpmc_mixture_create(int nc, const double* tc_k, const double* pc_pa, const double* omega,
                        const double* kij, mpmc_mixture_t* out);
int mpmc_mixture_destroy(mpmc_mixture_t mixture);
int mpmc_mixture_components(mpmc_mixture_t mixture);

typedef struct mpmc_flash_options {
    /* PtSplitIterationOptions::fugacity_tolerance (pt_split.hpp:31); default 1e-11 */
    double fugacity_tolerance;
    /* 2: solve_pr76_pt_phase_set (VLE, pr76_phase_set.hpp:42-52)
     * 3: Pr76PtFlashBackend::solve (max3, pr76_pt_flash_backend.hpp:190-237) */
    int32_t max_phases;
    /* 0: every evaluation in FP64 with the reference's bisection roots.
     * 1: decision-exact mixed precision. FP32 descents (mixing, ln(phi), TPD and the
     *    line search in FP32; cubic brackets, roots and root selection in FP64) screen
     *    every stability trial. Every search whose outcome or end point bears on a
     *    decision is run in FP64 from its start: trials FP32 cannot classify beyond
     *    rounding reach, the negative trials that seed the splits, every split attempt,
     *    and the whole review of a cell whose review finds instability. Only trials FP32
     *    classifies as stationary beyond rounding reach keep their FP32 status. Phase
     *    counts, statuses and labels are those of precision 0, except on cells where
     *    precision 0 ends indeterminate because such a trial's FP64 descent reaches its
     *    512-step limit (none of the 1M cells of the nc=6 grid workload, 6 of 9.7M cells
     *    of the uneven 256-mixture batch); values are FP64.
     * 2: fast mixed precision. TPD descents and split attempts run in FP32 (Rachford-Rice
     *    in FP32 away from the ends of (0, 1)); searches within FP32 rounding of a
     *    threshold, or slow enough to meet the iteration limit, are continued in FP64 from
     *    their current point; the selected split is polished by FP64 Newton and every
     *    published value is FP64. Decisions can differ from precision 0 on cells where the
     *    reference's own split search ends on a line-search failure or the iteration
     *    limit (269 of 1,048,576 cells of the nc=6 grid workload); values of the other
     *    cells agree within 1e-7 relative.
     *    Precisions 1 and 2 run the staged schedule only. */
    int32_t precision;
    /* 0: staged batch (one lane per stability trial / split attempt, persistent kernels)
     * 1: serial (one lane per cell runs the whole search) */
    int32_t schedule;
    /* flash plans: the most mixtures in one launch group (0: as many as the group's LDS
     * mixture table holds, mpmc_flash_multi_max_mixtures). Smaller groups mean more
     * launches but smaller tables and more resident workgroups per CU; the autotuner in
     * mpmc_hnu_amd/flash.py (autotune_plan) picks it by timing. */
    int32_t group_mixtures;
    int32_t reserved[2];
} mpmc_flash_options;

typedef struct mpmc_flash_outputs {
    int32_t* status;      /* [ncell] status word */
    int32_t* phase_count; /* [ncell] candidate phase-set size (0 = no candidate set) */
    double* beta;         /* [3][ncell] mole phase fractions */
    double* zfactor;      /* [3][ncell] compressibility factor, NaN when not published */
    int32_t* branch;      /* [3][ncell] branch word */
    double* x;            /* [3][nc][ncell] phase compositions */
    double* ln_phi;       /* [3][nc][ncell] ln(fugacity coefficient) */
    uint32_t* counters;   /* [4][ncell] or NULL */
} mpmc_flash_outputs;

void mpmc_flash_default_options(mpmc_flash_options* options);

/*
 * Batched PT flash, one result per cell, identical in meaning to calling
 * PtFlashBackend::solve (modules/flash/include/mpmc/flash/pt_flash_backend.hpp:187-195)
 * once per cell with {p, T, feed}. Numerical failure is a per-cell status, never an
 * error code; only argument and runtime failures return non-zero.
 */
int mpmc_flash_batch(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t,
                     const double* z, const mpmc_flash_options* options,
                     const mpmc_flash_outputs* out, void* stream);

/* Device workspace bytes the staged schedule needs for ncell cells of an nc-component
 * mixture (0 when nc is out of range). */
size_t mpmc_flash_workspace_bytes(int nc, int64_t ncell);

/* mpmc_flash_batch with a caller-owned device workspace of at least
 * mpmc_flash_workspace_bytes(nc, ncell) bytes. mpmc_flash_batch itself keeps one
 * workspace per mixture handle (grown on demand), so concurrent calls on one handle
 * from different streams must use this entry point with separate workspaces. */
int mpmc_flash_batch_ws(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t,
                        const double* z, const mpmc_flash_options* options,
                        const mpmc_flash_outputs* out, void* workspace, size_t workspace_bytes,
                        void* stream);

/* Bytes of the static scratch of mpmc_flash_batch_static (256 MiB). */
#define MPMC_STATIC_SCRATCH_BYTES (256u * 1024u * 1024u)

/* mpmc_flash_batch without any allocation: the staged schedule's workspace is a static
 * device array of MPMC_STATIC_SCRATCH_BYTES that every device running the library holds
 * until the process ends; the call itself allocates no device memory. No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Limit: ncell with
 * mpmc_flash_workspace_bytes(nc, ncell) <= MPMC_STATIC_SCRATCH_BYTES (about 38,000 cells
 * at nc = 6, 19,000 at nc = 12); larger batches are rejected with MPMC_E_INVALID_ARGUMENT
 * before any device work. Every call on one device shares that device's array: calls
 * that overlap on one device (other streams, other host threads) race and must be
 * serialised by the caller. Each call writes every workspace word it reads, so no state
 * carries over between calls. Results are those of mpmc_flash_batch. */
int mpmc_flash_batch_static(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t,
                            const double* z, const mpmc_flash_options* options, const mpmc_flash_outputs* out,
                            void* stream);

/* Warm-started batch flash: up to 6 compositions per cell, hints[(h*nc + i)*ncell + c]
 * (device memory), appended in order to the initial stability test's starts after the
 * generated ones, as Pr76PtFlashBackendOptions::initial_starts does for the
 * initial_stability_starts of PtSolveHints (pt_solve_hints.hpp:26-43,
 * pr76_executable_model.hpp:134-140; pr76_three_phase.hpp:328-341). A hint whose component 0
 * is NaN is absent for that cell. A present hint must be a composition (sum 1 within 64 eps)
 * with exactly the feed's support, else the cell is rejected (MPMC_FLAG_REJECTED), as the
 * reference rejects the request (pt_stability.hpp:264-271). workspace NULL: the mixture's own.
 * Staged schedule only. */
int mpmc_flash_batch_hinted(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t,
                            const double* z, int nhint, const double* hints,
                            const mpmc_flash_options* options, const mpmc_flash_outputs* out,
                            void* workspace, size_t workspace_bytes, void* stream);

/* mpmc_flash_batch_hinted plus final-review warm starts: up to 6 compositions per cell,
 * final_hints[(h*nc + i)*ncell + c] (device memory), placed before the selected pair's two
 * compositions in the common-tangent review, as Pr76PtFlashBackendOptions::final_starts
 * does for the final_two_phase_stability_starts of PtSolveHints (pt_solve_hints.hpp:34-36,
 * pr76_executable_model.hpp:139; pt_split.hpp:508-510). The same NaN, composition and
 * support rules as the initial hints (pt_split.hpp:440-446). Staged schedule only. */
int mpmc_flash_batch_warm(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t,
                          const double* z, int nhint, const double* hints, int nfinal, const double* final_hints,
                          const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* workspace,
                          size_t workspace_bytes, void* stream);

/*
 * Ordered-path PT continuation (solve_pr76_pt_continuation, pr76_pt_continuation.hpp:194-310)
 * of npath paths at once. Path j has the feed z[i*npath + j] and npoint ordered states
 * p[k*npath + j], t[k*npath + j] (device memory); point k of every path is a max3 flash
 * (options->max_phases must be 3; staged schedule) of the path feed normalised once, and
 * the points of one path run in order while the paths run side by side.
 * continuation options (NULL: mpmc_continuation_default_options):
 *   carry_compositions (carry_previous_phase_compositions): the previous point's accepted
 *     phase compositions (a single phase carries its normalised feed) lead the next point's
 *     initial and final-review stability starts;
 *   carry_three_phase (carry_previous_three_phase_state): a three-phase previous point's
 *     compositions and phase-1/2 fractions are the next point's first three-phase start;
 *   fallback_initial / fallback_final (fallback_initial_starts / _final_starts, at most 3
 *     each): appended after the carried compositions; every start list is kept free of
 *     repeats (exact equality);
 *   three_phase_start (point_options.three_phase_starts with one entry): a fixed start
 *     tried after the carried one, with seed start_seed.
 * After a point that accepts no phase set nothing is carried.
 * out holds npoint consecutive flash records of npath cells (the mpmc_flash_outputs layout
 * with ncell = npath, point k at offset k times one record); MPMC_FLAG_START_CONSUMED marks a
 * point whose three-phase stage ran from the carried start. point_info (device, [npoint]
 * [npath], may be NULL) = incoming hint phase count (0: none) | carried stability start
 * count << 4 | carried three-phase start << 8. A path whose feed or states fail the
 * reference's path check, or one of whose points is rejected (including by a fallback or
 * fixed start the feed does not support), is published as rejected (MPMC_FLAG_REJECTED) at
 * every point with point_info 0, as the reference rejects the whole path. Transition
 * brackets follow from consecutive accepted phase counts. workspace NULL: the mixture's
 * own, else mpmc_pt_continuation_workspace_bytes(nc, npath) bytes of device memory.
 */
typedef struct mpmc_continuation_options {
    int carry_compositions;  /* default 1 */
    int carry_three_phase;   /* default 1 */
    int nfallback_initial;   /* 0..3 */
    int nfallback_final;     /* 0..3 */
    int three_phase_start;   /* 0 or 1 */
    double fallback_initial[3][MPMC_MAX_COMPONENTS]; /* [k][i], i < nc */
    double fallback_final[3][MPMC_MAX_COMPONENTS];
    double start[3][MPMC_MAX_COMPONENTS];            /* [phase][i] */
    double start_seed[2];    /* fractions of phases 1, 2; default 1/3, 1/3 */
} mpmc_continuation_options;

void mpmc_continuation_default_options(mpmc_continuation_options* options);
size_t mpmc_pt_continuation_workspace_bytes(int nc, int64_t npath);
int mpmc_pt_continuation(mpmc_mixture_t mixture, int64_t npath, int npoint, const double* p, const double* t,
                         const double* z, const mpmc_continuation_options* continuation,
                         const mpmc_flash_options* options, const mpmc_flash_outputs* out, int32_t* point_info,
                         void* workspace, size_t workspace_bytes, void* stream);

/*
 * Ragged multi-mixture batch: the cells of several mixtures with the same component count
 * in one staged launch sequence. Mixture k owns cells [cell_offset[k], cell_offset[k+1])
 * of the batch (cell_offset: host array of nmix + 1 entries, cell_offset[0] = 0,
 * cell_offset[nmix] = ncell); p, t, z and the outputs cover all ncell cells in the
 * layouts of mpmc_flash_batch. The mixture table is staged in LDS and every cell reads its
 * own mixture's constants, so one persistent pipeline serves the whole uneven batch.
 * Results are bit-identical to mpmc_flash_batch on each mixture's cells. The workspace is
 * caller-owned (mpmc_flash_multi_workspace_bytes), so the call allocates nothing and can
 * be captured in a graph. Staged schedule only (options->schedule = 0).
 * The reference solves one request at a time (PtFlashBackend::solve,
 * modules/flash/include/mpmc/flash/pt_flash_backend.hpp:187-195); this is the batch form of
 * that call for many mixtures at once. */
int mpmc_flash_multi_max_mixtures(int nc);
size_t mpmc_flash_multi_workspace_bytes(int nc, int nmix, int64_t ncell);
int mpmc_flash_multi(int nmix, const mpmc_mixture_t* mixtures, const int64_t* cell_offset, const double* p,
                     const double* t, const double* z, const mpmc_flash_options* options,
                     const mpmc_flash_outputs* out, void* workspace, size_t workspace_bytes, void* stream);

/*
 * Flash plan: a fixed uneven batch (segments of any mixtures and component counts, each
 * with its own device inputs and outputs in the layouts of mpmc_flash_batch) prepared
 * once and replayed. Creation groups the segments by component count into ragged
 * launches (mpmc_flash_multi: one persistent pipeline per group of up to
 * mpmc_flash_multi_max_mixtures(nc) mixtures), allocates every buffer, uploads the
 * mixture tables, runs the batch once and captures the whole launch sequence (gather,
 * pipeline, scatter per group) into one HIP graph. mpmc_flash_plan_launch replays the
 * graph on the caller's stream: the segments' current input contents are flashed and their
 * outputs written, with no allocation and one launch call. Results are those of
 * mpmc_flash_batch on each segment. A plan is bound to the device pointers it was made
 * with; replays of one plan must not overlap. No counterpart in the reference: it is the
 * repeated batch caller of PtFlashBackend::solve (pt_flash_backend.hpp:187-195), e.g. a
 * scan of every cell after each Newton solve. */
typedef struct mpmc_flash_segment {
    mpmc_mixture_t mixture;
    int64_t ncell;
    const double* p;
    const double* t;
    const double* z;          /* [nc][ncell] */
    mpmc_flash_outputs out;   /* counters: given for every segment or for none */
} mpmc_flash_segment;

typedef struct mpmc_flash_plan* mpmc_flash_plan_t;

int mpmc_flash_plan_create(int nseg, const mpmc_flash_segment* segments, const mpmc_flash_options* options,
                           mpmc_flash_plan_t* out);
int mpmc_flash_plan_launch(mpmc_flash_plan_t plan, void* stream);
/* Launch groups and device bytes held by the plan. */
int mpmc_flash_plan_info(mpmc_flash_plan_t plan, int32_t* groups, size_t* device_bytes);
/* Waits for the device, then frees the plan. */
int mpmc_flash_plan_destroy(mpmc_flash_plan_t plan);

/*
 * Candidate-phase PR76 evaluation, one composition per cell: all roots, then ln(phi) of
 * every root (admissibility is in root_flags). Replaces Pr76Phase::roots_full + evaluate_full
 * (modules/thermodynamics/include/mpmc/thermodynamics/pr76_phase.hpp:125-148).
 *   root_status[ncell]            Pr76RootStatus (success=0, near_multiple=1,
 *                                 iteration_limit=2, unrepresentable=3)
 *   root_count[ncell]
 *   zroot[3][ncell]               Z of each root (increasing), NaN past count
 *   root_flags[3][ncell]          slope_sign+1 | derivative_valid << 4
 *   ln_phi[3][nc][ncell]          ln(phi) of each root, NaN past count
 */
int mpmc_phase_batch(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t,
                     const double* x, int32_t* root_status, int32_t* root_count, double* zroot,
                     int32_t* root_flags, double* ln_phi, void* stream);

/*
 * Rachford-Rice at fixed log K, one feed per cell: solve_rachford_rice
 * (modules/flash/include/mpmc/flash/rachford_rice.hpp:70-147).
 *   z[nc][n], log_k[nc][n]; guess[n] (optional, NULL = 0.5) only steers the root enclosure
 *   that lets the bisection skip evaluations whose sign is certified, and never changes the
 *   result; a negative guess runs the plain bisection.
 *   status[n]: RachfordRiceStatus (interior=0, no_resolved_interior_root=1, degenerate=2,
 *              iteration_limit=3, unrepresentable=4; -2 = input rejected)
 *   beta[n]; x[nc][n], y[nc][n] (NaN unless interior); iterations[n] bisection steps
 */
int mpmc_rr_batch(int nc, int64_t n, const double* z, const double* log_k, const double* guess,
                  int32_t* status, double* beta, double* x, double* y, int32_t* iterations,
                  void* stream);

/* ---------------------------------------------------------------- closure
 * Thermodynamic closure of accepted two-phase flash results: build_pr76_pt_vle_closure
 * (modules/physics/include/mpmc/physics/pr76_thermodynamic_closure.hpp:148-381) with its
 * implicit sensitivity differentiate_pr76_pt_vle (modules/flash/include/mpmc/flash/
 * pr76_sensitivity.hpp:51-390). Columns q = (p, T, z_0..z_{nc-2}), nq = nc + 1.
 *
 * Per-cell status word: primal | linearization << 4 | reason << 8, with
 *   primal        ThermodynamicClosurePrimalStatus (valid=0, phase_regime_not_implemented=1,
 *                 indeterminate=2), thermodynamic_closure.hpp:18-22
 *   linearization ThermodynamicClosureLinearizationStatus (available=0, unavailable=1)
 *   reason        ThermodynamicClosureLinearizationReason (none=0 .. solution_not_accepted=7),
 *                 thermodynamic_closure.hpp:26-35
 *   0x2000        the cell is one the reference rejects with an exception
 * Two-phase sets (VLE status two-phase, or max3 status two-phase) are closed; single-phase
 * and three-phase sets are phase_regime_not_implemented; others are indeterminate. */
typedef struct mpmc_closure_options {
    double minimum_derivative_phase_fraction; /* PtSensitivityOptions (pt_sensitivity.hpp:26-34), 2^-26 */
    double minimum_reciprocal_condition;      /* default 0 */
    double fugacity_tolerance;                /* tolerance of the flash that produced the inputs, 1e-11 */
    int32_t reserved[2];
} mpmc_closure_options;

typedef struct mpmc_closure_outputs {
    int32_t* status;   /* [ncell] status word */
    double* primal;    /* [4][ncell] Z_L, Z_V, rho_L, rho_V (mol/m^3); NaN unless the primal is valid */
    double* jac;       /* [2nc+5][nq][ncell] d/dq of beta_V, x_L[nc], y_V[nc], Z_L, Z_V, rho_L, rho_V,
                        * jac[(row*nq + col)*ncell + c]; NaN unless the linearization is available */
    double* log_k_jac; /* [nc][nq][ncell] dlogK/dq, or NULL */
    double* diag;      /* [3][ncell] equilibrium Jacobian rcond, solve backward error, residual norm,
                        * or NULL */
} mpmc_closure_outputs;

void mpmc_closure_default_options(mpmc_closure_options* options);

/* Closure of every cell of a flash result (the mpmc_flash_batch outputs status, phase_count,
 * beta, zfactor, branch and x on the same p, t, z). nc >= 2. Uses a workspace held by the
 * mixture handle (one closure at a time per handle). */
int mpmc_closure_batch(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t,
                       const double* z, const mpmc_flash_outputs* flash,
                       const mpmc_closure_options* options, const mpmc_closure_outputs* out,
                       void* stream);

/* ---------------------------------------------------------------- two-phase assembly
 * Newton-iteration residual and Jacobian of the two-phase natural-variable system:
 * TwoPhaseSnesAssemblyContext3D::evaluate_complete_assembly
 * (modules/flow/discretization/petsc/include/mpmc/flow_discretization_petsc/
 *  two_phase_snes_assembly.hpp:550-1090) on one rank; PETSc SNES/KSP stay on the host.
 *
 * Per cell q = (p, T, S_0, x_0 independent [nc-1], x_1 independent [nc-1]), Q = 2nc+1,
 * the last component of each phase dependent (two_phase_cell_state.hpp:120-231). Rows per
 * cell: nc component balances, energy, nc fugacity-equilibrium rows
 * ln(x0_i/x1_i) + ln phi0_i - ln phi1_i. Phase properties come from the PR76 selected
 * branch of each phase (root index per cell and phase, the flash's branch).
 *
 * Transport, caloric, rock and relative-permeability models are caller-owned in the
 * reference (pr76_selected_phase_property_closure.hpp:32-45,
 * selected_phase_production_cell_evaluator.hpp:24-64); here they are these closed forms:
 *   mu = mu_a0 + mu_a1 rho_mass^2, h = cp (T - t0), U_rock = cv_rock (T - t0), kr_p = S_p^2.
 *
 * Layouts (device memory, cell or face index fastest):
 *   q[Q][ncell], branch[2][ncell], prev[nc+1][ncell] (N_i and E of the previous time),
 *   residual[Q][ncell], diag[Q*Q][ncell] (row-major Q x Q block of the cell's own columns),
 *   offdiag[(nc+1)*Q][max_degree][ncell]: for slot j < degree(c) of cell c, the block (rows
 *   of c, columns of the cell across face adj_face[adj_off[c] + j]) at [.][j][c]; slots past
 *   a cell's degree hold zeros. The slots follow the grid's cell -> face CSR
 *   (mpmc_tpfa_grid_adjacency), so (j, c) is CSR position adj_off[c] + j, and the matrix is
 *   block-CSR with Q x Q blocks stored ELL-style, cell fastest, so that every store is
 *   coalesced. Equilibrium rows have no off-diagonal entries. */
typedef struct mpmc_tpfa_grid* mpmc_tpfa_grid_t;

/* Internal faces of a mesh: owner/neighbour cells, T_f [m3] (harmonic combination of the
 * half transmissibilities, tpfa_static_face_transmissibility_3d.hpp:125-182), gdot = g . (x_n
 * - x_o) [m2/s2], thermal conductance [W/K]; per cell bulk volume [m3] and porosity. Device
 * arrays, referenced (not copied) until destroy. */
int mpmc_tpfa_grid_create(int64_t ncell, int64_t nface, const double* bulk_volume, const double* porosity,
                          const int32_t* owner, const int32_t* neighbour, const double* transmissibility,
                          const double* gdot, const double* thermal_conductance, mpmc_tpfa_grid_t* out);
int mpmc_tpfa_grid_destroy(mpmc_tpfa_grid_t grid);
/* The grid's cell -> face CSR: adj_off[ncell+1], adj_face[2 nface] (device memory); a
 * cell's faces are in face order. */
int mpmc_tpfa_grid_adjacency(mpmc_tpfa_grid_t grid, const int32_t** adj_off, const int32_t** adj_face);
/* Number of cells of the grid. */
int mpmc_tpfa_grid_cells(mpmc_tpfa_grid_t grid, int64_t* ncell);
/* The largest number of faces of one cell: the slot count of the off-diagonal blocks. The
 * reference's DMPlex adjacency (two_phase_snes_assembly.hpp:1719-1784 preallocation) has no
 * such call; it sizes the ELL layout here. */
int mpmc_tpfa_grid_max_degree(mpmc_tpfa_grid_t grid, int32_t* max_degree);

typedef struct mpmc_assembly_model {
    double mu_a0, mu_a1; /* Pa s, Pa s m6/kg2 */
    double cp;           /* J/(kg K) */
    double t0;           /* K */
    double cv_rock;      /* J/(m3 K) */
    double molar_mass[MPMC_MAX_COMPONENTS]; /* kg/mol */
} mpmc_assembly_model;

typedef struct mpmc_assembly_outputs {
    int32_t* cell_status; /* [ncell] 0, or 1 where the cell's evaluation raised a domain error */
    double* residual;     /* [Q][ncell] */
    double* diag;         /* [Q*Q][ncell] */
    double* offdiag;      /* [(nc+1)*Q][max_degree][ncell] */
} mpmc_assembly_outputs;

void mpmc_assembly_default_model(int nc, mpmc_assembly_model* model);
/* One Newton evaluation. domain_error (optional host pointer): synchronises the stream and
 * reports whether any cell raised a domain error (the reference's domain_error status; the
 * outputs are then not a valid assembly). */
int mpmc_assembly_2p(mpmc_mixture_t mixture, mpmc_tpfa_grid_t grid, const mpmc_assembly_model* model,
                     const double* q, const int32_t* branch, const double* prev, double dt,
                     const mpmc_assembly_outputs* out, int32_t* domain_error, void* stream);
/* Pore-volume inventories N_i and E at a state into inventory[nc+1][ncell] (the previous-time
 * input of mpmc_assembly_2p). */
int mpmc_inventory_2p(mpmc_mixture_t mixture, mpmc_tpfa_grid_t grid, const mpmc_assembly_model* model,
                      const double* q, const int32_t* branch, int32_t* cell_status, double* inventory,
                      int32_t* domain_error, void* stream);

/* Schur-reduced system with (nc+1) x (nc+1) blocks (the block-CSR Jacobian a solver on
 * nc+1 unknowns per cell takes). No reference counterpart: the reference hands the
 * unreduced Q x Q system to PETSc (two_phase_snes_assembly.hpp:1719-1784); this is an exact
 * algebraic transform of that system, not an approximation. Each cell's unknowns split into
 * nc secondary columns of q (secondary[k], host array, NULL: x_0 of component 0 and the
 * nc-1 independent compositions of phase 1) and the nc+1 others, the primaries, in q order.
 * The cell's equilibrium rows give dx_S = -(M_P dx_P + M_r) with
 * M = J_ES^-1 [J_EP | r_E]; eliminating dx_S from the conservation rows leaves
 *   A_cc = J_CP - J_CS M_P(c), A_cn = O_CP - O_CS M_P(n),
 *   b_c = r_C - J_CS M_r(c) - sum_n O_CS M_r(n),       A dx_P = -b.
 * Inputs: the outputs of mpmc_assembly_2p on this grid. Outputs (device, cell fastest):
 *   residual[nc+1][ncell] = b, diag[(nc+1)^2][ncell] = A_cc (row-major),
 *   offdiag[(nc+1)^2][max_degree][ncell] = A_cn in the slots of the unreduced layout (zeros
 *   past a cell's degree), back[nc][nc+2][ncell] = M (column nc+1 is M_r), status[ncell]:
 *   1 where J_ES is singular to working precision (pivot below 64 eps nc |J_ES|; that
 *   cell's M and reduced entries are NaN). No allocation; two kernels on the stream. */
typedef struct mpmc_reduced_outputs {
    int32_t* status;
    double* residual;
    double* diag;
    double* offdiag;
    double* back;
} mpmc_reduced_outputs;
int mpmc_assembly_reduce_2p(mpmc_tpfa_grid_t grid, int nc, const int32_t* secondary,
                            const mpmc_assembly_outputs* in, const mpmc_reduced_outputs* out, void* stream);

/* Newton-evaluation plan: one mpmc_assembly_2p evaluation on these exact device buffers
 * (the per-cell selected-phase property closure, then the TPFA rows), captured once into a
 * HIP graph. mpmc_assembly_plan_launch replays it on the caller's stream from the current
 * contents of q, branch and prev (dt fixed at creation), with no allocation and one launch
 * call; the outputs are those of mpmc_assembly_2p. The plan owns its record workspace and
 * error flag, so plans on one grid are independent; replays of one plan must not overlap.
 * Replaces the per-iteration evaluate_complete_assembly call of snes_function and
 * snes_jacobian (two_phase_snes_assembly.hpp:1698,1753). */
typedef struct mpmc_assembly_plan* mpmc_assembly_plan_t;
int mpmc_assembly_plan_create(mpmc_mixture_t mixture, mpmc_tpfa_grid_t grid, const mpmc_assembly_model* model,
                              const double* q, const int32_t* branch, const double* prev, double dt,
                              const mpmc_assembly_outputs* out, mpmc_assembly_plan_t* plan);
int mpmc_assembly_plan_launch(mpmc_assembly_plan_t plan, void* stream);
/* Synchronises the stream and reports whether the last replay raised a domain error. */
int mpmc_assembly_plan_domain_error(mpmc_assembly_plan_t plan, void* stream, int32_t* domain_error);
/* Waits for the device, then frees the plan. */
int mpmc_assembly_plan_destroy(mpmc_assembly_plan_t plan);

/* ---------------------------------------------------------------- phase-transition scan
 * The post-SNES PT-flash phase-transition scanner
 * (flow_discretization_petsc/post_snes_pt_flash_phase_transition_scanner.hpp): one flash
 * per cell at its current state, compared with the cell's current phase count.
 *
 * Per-cell code (scan_post_snes_pt_flash_source_cell_3d, :203-301):
 *   MPMC_SCAN_SAME           accepted phase set of the source's size: no proposal
 *   MPMC_SCAN_PROPOSAL       accepted set of another size: a transition proposal, with
 *                            target_count phases (the flash record holds their beta, x, Z,
 *                            ln phi) and target phase molar densities p / (Z R T)
 *   MPMC_SCAN_INDETERMINATE  no accepted phase set, a feed the flash rejects (its
 *                            std::domain_error, caught by the scanner, :292-299), a target
 *                            density that is not finite and positive, or (two-phase state
 *                            scan) a cell whose state or inventory cannot define a PT feed
 *   MPMC_SCAN_ERROR          a malformed source or an unsupported source phase count, or a
 *                            lower target without the backend's phase-disappearance evidence
 *                            (invalid_argument / PETSC_ERR_ARG_INCOMP; the PR76 backend
 *                            reports such evidence only for 3 -> 2 through a fresh two-phase
 *                            neighbour solve, pr76_pt_flash_backend.hpp:138-146;
 *                            phase_set_transition_flash_adapter.hpp:171-183)
 * The reference walks the cells in order and stops at the first indeterminate cell (scan
 * status indeterminate, proposals dropped) or the first error (the error returned). The
 * summary reports that outcome: first_stop = index of the first such cell (-1: none),
 * status = MPMC_SCAN_STATUS_INDETERMINATE when that cell is indeterminate (proposals 0),
 * and the call returns MPMC_E_INVALID_ARGUMENT when it is an error (proposals: those
 * before the error cell). Every cell's code is written either way. The transition evidence report and the trigger of the reference's candidate
 * (phase_set_transition_flash_adapter.hpp:132-220) are not produced; the caller builds the
 * candidate from the record. The call synchronises the stream to fill the summary.
 *
 * density_mode 0 takes each target phase's Z from the flash record, as the density
 * resolver of the reference's scanner test does (scanner_test.cpp:26-57); the record of a
 * single accepted phase publishes no Z (pt_vle_phase_set.hpp:68-73), so every proposal to
 * one phase is indeterminate. density_mode 1 (a resolver with no counterpart in the
 * reference) gives that phase the Z of its provider root at the feed (the root the branch
 * word names, from mpmc_phase_batch); the candidate then meets the evidence rule, and since
 * the PR76 backend reports no disappearance evidence for a single-phase target, such cells
 * are errors, as the reference's adapter raises for them. It allocates 8 (4 + 3 nc) bytes
 * per cell for the duration of the call. */
enum mpmc_scan_code { MPMC_SCAN_SAME = 0, MPMC_SCAN_PROPOSAL = 1, MPMC_SCAN_INDETERMINATE = 2,
                      MPMC_SCAN_ERROR = 3 };
/* PostSnesPhaseTransitionScanStatus3D (post_snes_phase_transition_controller.hpp:78-81) */
enum mpmc_scan_status { MPMC_SCAN_COMPLETE = 0, MPMC_SCAN_STATUS_INDETERMINATE = 1 };

typedef struct mpmc_scan_outputs {
    int32_t* code;           /* [ncell] mpmc_scan_code */
    int32_t* target_count;   /* [ncell] accepted phase count, 0 without an accepted set */
    double* target_density;  /* [3][ncell] mol/m3 of the target phases of a proposal, else NaN */
} mpmc_scan_outputs;

typedef struct mpmc_scan_summary {
    int32_t status;          /* mpmc_scan_status */
    int64_t first_stop;      /* first indeterminate or error cell, -1 when none */
    int64_t proposals;       /* proposals the reference's cell loop holds when it returns:
                                the MPMC_SCAN_PROPOSAL cells before first_stop, 0 when the
                                status is indeterminate */
} mpmc_scan_summary;

/* Scan of caller-given PT sources (PostSnesPtFlashSourceCellSnapshot3D, :30-37): p, t,
 * z[nc][ncell] and source_phase_count[ncell] (1..3) on the device; the flash runs with
 * `options` and writes its records to `out`. Source validation is
 * validate_source_snapshot (:76-115). Replaces the per-cell
 * scan_post_snes_pt_flash_source_cell_3d call. */
int mpmc_transition_scan(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t, const double* z,
                         const int32_t* source_phase_count, const mpmc_flash_options* options,
                         int density_mode, const mpmc_flash_outputs* out, const mpmc_scan_outputs* scan,
                         mpmc_scan_summary* summary, void* stream);

/* Scan of a converged two-phase state q[Q][ncell] (the assembly's natural variables) on
 * `grid`: each cell's source is its current pore-volume inventory (mpmc_inventory_2p, into
 * inventory[nc+1][ncell]) normalised to z_i = N_i / sum N (into feed[nc][ncell]) at the
 * cell's pressure q[0] and temperature q[1], with source phase count 2
 * (make_source_snapshot, :160-200). Replaces scan_post_snes_pt_flash_phase_transitions_3d
 * (:303-423) for the owned cells of a two-phase system; the caller checks that SNES
 * converged. */
int mpmc_transition_scan_2p(mpmc_mixture_t mixture, mpmc_tpfa_grid_t grid, const mpmc_assembly_model* model,
                            const double* q, const int32_t* branch, double* inventory, double* feed,
                            const mpmc_flash_options* options, int density_mode, const mpmc_flash_outputs* out,
                            const mpmc_scan_outputs* scan, mpmc_scan_summary* summary, void* stream);

/* ---------------------------------------------------------------- serving
 * Continuous-batching flash server: the PT service request loop (reference
 * modules/runtime/include/mpmc/runtime/pt_service.hpp, PtFlashBackend::solve per request)
 * on one device. Jobs are admitted as they arrive; each launch packs cells of all pending
 * jobs on one mixture (small jobs first, then large ones in arrival order); `slots`
 * launches are in flight at once on the one launch stream. Device memory is
 * allocated once at creation (one flash workspace and per-slot chunk buffers, nc <= 12)
 * and the slot count is reduced until it fits hbm_cap_bytes. Mixtures are kept resident
 * per parameter set. Results are those of mpmc_flash_batch on the job's cells. */
typedef struct mpmc_server* mpmc_server_t;

typedef struct mpmc_server_options {
    int32_t slots;             /* chunks in flight, default 2 */
    int32_t precision;         /* as mpmc_flash_options.precision */
    int32_t max_phases;        /* 2 or 3, default 3 */
    int32_t reserved0;
    double fugacity_tolerance; /* default 1e-11 */
    int64_t chunk_cells;       /* cells per launch, default 262144 */
    int64_t small_job_cells;   /* jobs up to this size are served first, and launches are capped
                                * at this size while such jobs wait; default 65536 */
    size_t hbm_cap_bytes;      /* device memory the server may hold, default 200 GiB */
    void* stream;              /* launch stream (hipStream_t), NULL = the null stream */
} mpmc_server_options;

/* Host result buffers of one job, same SoA layout as mpmc_flash_outputs with n = ncell. */
typedef struct mpmc_host_outputs {
    int32_t* status;
    int32_t* phase_count;
    double* beta;
    double* zfactor;
    int32_t* branch;
    double* x;
    double* ln_phi;
} mpmc_host_outputs;

void mpmc_server_default_options(mpmc_server_options* options);
int mpmc_server_create(const mpmc_server_options* options, mpmc_server_t* out);
/* Drains every submitted job, then releases the device memory. */
int mpmc_server_destroy(mpmc_server_t server);
int mpmc_server_slots(mpmc_server_t server);
size_t mpmc_server_device_bytes(mpmc_server_t server);
/* Queue one job. p[ncell], t[ncell], z[nc][ncell] and the outputs are host memory that must
 * stay valid until the job is done. Returns at once with the job id. */
int mpmc_server_submit(mpmc_server_t server, int nc, const double* tc_k, const double* pc_pa,
                       const double* omega, const double* kij, int64_t ncell, const double* p,
                       const double* t, const double* z, const mpmc_host_outputs* out,
                       int64_t* job_id);
/* mpmc_server_submit with warm-start hints (host, [nhint][nc][ncell], nhint <= 2, read until
 * the job is done): the per-request PtSolveHints initial_stability_starts, as in
 * mpmc_flash_batch_hinted. */
int mpmc_server_submit_hinted(mpmc_server_t server, int nc, const double* tc_k, const double* pc_pa,
                              const double* omega, const double* kij, int64_t ncell, const double* p,
                              const double* t, const double* z, int nhint, const double* hints,
                              const mpmc_host_outputs* out, int64_t* job_id);
/* Block until the job is done; latency_s = completion - submission (seconds). */
int mpmc_server_wait(mpmc_server_t server, int64_t job_id, double* latency_s);
int mpmc_server_poll(mpmc_server_t server, int64_t job_id, int32_t* done, double* latency_s);
/* Last error text of a server call on this thread (never NULL). */
const char* mpmc_server_last_error(void);

/* Last error text for this thread (never NULL). */
const char* mpmc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MPMC_HNU_AMD_H */

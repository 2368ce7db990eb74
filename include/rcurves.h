/*
 * rcurves.h — C ABI of the MI355X Sobolev-step library (librcurves.so).
 *
 * Drop-in boundary for the reference's src/core energy / gradient /
 * Sobolev-step API and its worker-solver messages. Plain pointers and sizes
 * only; every entry point is synchronous on the calling thread and returns an
 * rc_status. All device memory is owned by an rc_ctx (one HIP stream each);
 * host buffers passed in are read/written only during the call.
 *
 * Layouts: vertices are AoS double[V][3] (the reference's Vec3[]), edges are
 * int32[E][2] (Edge[]), flattened fields use the coordinate-block layout
 * [x0..x{V-1}, y0.., z0..] (src/core/sobolev/layout.ts:23-58).
 */
#ifndef RCURVES_H
#define RCURVES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RC_OK = 0,
    RC_SINGULAR = 1,     /* linsolve.ts /singular/ throw contract */
    RC_NONFINITE = 2,    /* linsolve.ts "treating as singular" (non-finite input) */
    RC_BAD_ARG = 3,
    RC_HIP_ERROR = 4,
    RC_UNSUPPORTED = 5,
    RC_DIVERGED = 6,     /* multigrid: MGDivergenceError (tpe_stage2_oracle.py:1061-1081) */
} rc_status;

typedef struct rc_ctx rc_ctx;

/* Tangent-point exponents and regulariser (optimizer.ts:18 DEFAULTS). */
typedef struct {
    double alpha;
    double beta;
    double epsilon;
} rc_params;

/* One ConstraintBlock as POD (constraintSet.ts:39-52); blocks stack in array
 * order exactly like a ConstraintSet. */
enum { RC_BLOCK_BARYCENTER = 0, RC_BLOCK_TOTAL_LENGTH = 1, RC_BLOCK_EDGE_LENGTHS = 2, RC_BLOCK_POINT = 3 };
typedef struct {
    int32_t kind;
    int32_t vertex;       /* RC_BLOCK_POINT: pinned vertex index */
    double x0[3];         /* RC_BLOCK_BARYCENTER: frozen target */
    double L0;            /* RC_BLOCK_TOTAL_LENGTH: frozen target */
    double target[3];     /* RC_BLOCK_POINT: frozen pin target */
    const double* ell0;   /* RC_BLOCK_EDGE_LENGTHS: frozen per-edge targets (host) */
    int32_t ell0_len;
    int32_t _pad;
} rc_block;

/* PenaltyConfig (penalties.ts:33-49): zero weight = off. */
typedef struct {
    double w_length;
    double w_diff;
    double w_field;
    double X[3];
} rc_penalties;

enum { RC_PRECISION_FP64 = 0, RC_PRECISION_MIXED = 1 };

/* Differential of the energy: StepOptions.mode 'analytical' | 'finiteDiff'
 * (optimizer.ts:21,96). */
enum { RC_MODE_ANALYTICAL = 0, RC_MODE_FINITE_DIFF = 1 };

/* SobolevStepOptions (optimizer.ts:95-175) + LineSearchOptions
 * (lineSearch.ts:333-371) + ProjectConstraintSetOptions (:120-145). */
typedef struct {
    rc_params params;
    int32_t has_energy_before;
    int32_t precision;          /* RC_PRECISION_FP64 (plain) | RC_PRECISION_MIXED */
    double energy_before;
    int32_t projection_frozen;  /* 1: 'frozen' (store.ts:359 app setting, default here);
                                 * 0: 'reassemble' (lineSearch.ts:263-276, FP64 form only) */
    int32_t max_iter;           /* projection maxIter, default 8 */
    double tol_abs, tol_rel;    /* projection tolerances, default 1e-4 */
    double c1, shrink, tau0, tau_min; /* Armijo: 1e-4, 0.5, 1, 1e-12 */
    double refine_tol;          /* mixed form: PCG relative residual target, default 1e-12 */
    int32_t refine_max_iter;    /* mixed form: PCG iteration cap, default 200 */
    int32_t collect_timings;    /* fill rc_timings (hipEvent per phase) */
    int32_t gradient_mode;      /* dE by RC_MODE_ANALYTICAL (default) | RC_MODE_FINITE_DIFF
                                 * (SobolevStepOptions.mode, optimizer.ts:96,278-282) */
    double fd_h;                /* forward-difference step h, default 1e-6 (optimizer.ts:18) */
    int32_t saddle_refine;      /* refinement steps of the saddle solution [g~; lambda] with
                                 * residuals formed in double-double (compensated) arithmetic,
                                 * default 1; 0 = off. No counterpart in the reference, whose
                                 * FP64 LDL^T stops at its own rounding: kappa(A) ~ 1e9 at
                                 * V = 8192 leaves an FP64 solve ~2e-8 from the exact solution
                                 * of the assembled system, the refined one at its rounding. */
    int32_t refine_columns;     /* refinement steps, before the above, of all 3(k+1) solve
                                 * columns A_hat^-1 [dE, C^T] (compensated residuals), default
                                 * 0: makes the frozen projection's corrections exact too (they
                                 * use A_hat^-1 C^T), which matters when they are large (length
                                 * targets that grow, many pins) */
    double saddle_refine_tol;   /* mixed form: PCG relative residual of each refinement
                                 * correction, default 1e-6 */
    int32_t report_ir_residual; /* 1: also form saddle_ir_residual, the compensated residual of
                                 * the returned (refined) solution: a diagnostic, one more
                                 * compensated pass over A beside the line search; default 0
                                 * (saddle_ir_residual reports 0) */
} rc_step_opts;

/* SobolevStepStats (optimizer.ts:196-202) + line-search diagnostics. */
enum {
    RC_REASON_NONE = 0,
    RC_REASON_ZERO_OR_NONFINITE_GRADIENT_NORM = 1,
    RC_REASON_NOT_A_DESCENT_DIRECTION = 2,
    RC_REASON_ARMIJO_FAILED = 3,
    RC_REASON_PROJECTION_FAILED = 4,
    RC_REASON_TAU_BELOW_MIN = 5,
    RC_REASON_SINGULAR_SYSTEM = 6,
};
typedef struct {
    double energy;
    double energy_before;
    double tau;
    double residual;          /* ||K z - rhs|| / max(1, ||rhs||), FP64, structured */
    double gradient_l2_norm;
    double slope;
    double phi_norm;
    double refine_residual;   /* final PCG relative residual (mixed form) */
    int32_t accepted;
    int32_t converged;
    int32_t projection_iterations; /* -1 = null */
    int32_t reason;
    int32_t trials;
    int32_t refine_iters;
    double saddle_ir_residual;   /* compensated ||rhs - K z|| / max(1, ||rhs||) of the returned
                                  * solution (0 when saddle_refine = 0 or report_ir_residual = 0) */
    double saddle_ir_correction; /* relative size of the last refinement correction */
    int32_t saddle_ir_steps;     /* refinement steps applied (columns + saddle) */
    int32_t fp64_fallback;       /* mixed form: 1 when the PCG missed refine_tol within
                                  * refine_max_iter and the solve was redone with the FP64
                                  * factorisation (refine_residual / refine_iters keep the PCG's) */
} rc_step_stats;

/* PhaseSample (phaseTimings.ts:22-31): accumulated device time of one phase
 * over a step (HIP events on the stream the phase runs on) and how often it
 * fired; calls = 0 means the phase did not fire (absent from the reference's
 * Partial ledger). */
typedef struct {
    double ms;
    int32_t calls;
    int32_t _pad;
} rc_phase;

/* SobolevStepTimings (phaseTimings.ts:9-20, 32-36); sub-phases overlap their
 * parents as in the reference: factor inside saddle, energy and projection
 * inside lineSearch, the projection's correction solves are counted in saddle
 * (and, reassembled, in assembleA and factor) as the reference's wraps count
 * them. bHigh and bLow are one fused kernel here (both Gram matrices in one
 * pass over the pair tiles), reported as assembleA; expand never fires (nor
 * does it in the reference's flat path, gradient.ts:95-99). The frozen
 * projection's correction solves run inside the projection kernel: their
 * count is in saddle.calls, their time in projection.ms only. A singular
 * system records neither factor nor saddle (the reference's factor throws
 * inside both wraps). */
typedef struct {
    rc_phase dE, energy, assembleA, saddle, factor, projection, lineSearch, step;
} rc_timings;

/* Fills defaults (optimizer.ts:18, lineSearch.ts:207-209,435-438). */
void rc_default_step_opts(rc_step_opts* o);

/* Topology message / calculateDisjointPairs (solverWorker.ts:37-43,
 * tangentPointEnergy.ts:33-47): builds the device topology once per
 * graphVersion. Disjointness is derived on the device from endpoints. */
rc_status rc_ctx_create(const int32_t* edges, int32_t E, int32_t V, rc_ctx** out);
void rc_ctx_destroy(rc_ctx* ctx);
/* Message of the last failing call on this context (never NULL). */
const char* rc_last_error(const rc_ctx* ctx);

/* calculateEnergy (tangentPointEnergy.ts:50-111). */
rc_status rc_energy(rc_ctx* ctx, const double* xyz, const rc_params* p, double* out);
/* The energy of rc_energy without any allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Device pointers only:
 * d_xyz = SoA x[V], y[V], z[V]; d_edges = int32[E][2]; *d_energy receives the
 * energy; work is enqueued on `stream` (a hipStream_t, NULL = the default
 * stream) and the call returns without waiting. The edge table and partial
 * sums live in a static device array of the library (2 MiB per GPU, held
 * until the process ends): E <= 16384 (RC_UNSUPPORTED beyond). Calls that
 * overlap on one GPU share that array and race. Endpoints outside [0, V) or
 * self-loops give a NaN energy (they are not read). Same kernels and partial
 * order as rc_energy: the same bits. */
rc_status rc_energy_noalloc(const double* d_xyz, int32_t V, const int32_t* d_edges, int32_t E,
                            const rc_params* p, double* d_energy, void* stream);
/* gradientAnalytical (tangentPointEnergy.ts:139-392): adjoint of rc_energy,
 * AoS double[V][3] out, ascent orientation. Optionally also the energy. */
rc_status rc_gradient(rc_ctx* ctx, const double* xyz, const rc_params* p, double* grad,
                      double* energy_out);
/* gradientFiniteDiff (tangentPointEnergy.ts:114-136): forward difference
 * (E(x + h e_vd) - E(x)) / h for every vertex v and coordinate d, AoS out.
 * h must be finite and nonzero. */
rc_status rc_gradient_fd(rc_ctx* ctx, const double* xyz, const rc_params* p, double h,
                         double* grad);
/* step (optimizer.ts:40-68): the raw L2 step. dE at xyz_in by mode
 * (RC_MODE_ANALYTICAL | RC_MODE_FINITE_DIFF, h used by the latter),
 * xyz_out = xyz_in - step_size * dE per coordinate (product rounded, then the
 * difference), *energy = the energy at xyz_out. field (nullable) receives dE
 * (StepOptions.collectField). xyz_in is not modified. */
rc_status rc_step(rc_ctx* ctx, const double* xyz_in, const rc_params* p, int32_t mode,
                  double step_size, double h, double* xyz_out, double* energy, double* field);
/* Barnes-Hut tangent-point energy: the reference's stage-2 oracle bh_energy
 * (oracle/tpe_stage2_oracle.py:456-596) — edge BVH with leaf_size edges per
 * leaf, a cluster replaces its edges for a target edge when it has >= 2
 * edges, none sharing a vertex with the target, and rx / dist <= theta,
 * rT <= theta_t (the oracle's default theta_t is 1.5 theta). theta = 0 gives
 * the exact energy. stats (nullable, 3 entries): visits, approximated edge
 * references, direct pairs. The tree is built on the host from xyz. */
rc_status rc_energy_bh(rc_ctx* ctx, const double* xyz, const rc_params* p, double theta,
                       double theta_t, int32_t leaf_size, double* energy, int64_t* stats);
/* Barnes-Hut differential as the stage-2 oracle forms it: the forward
 * difference (E_bh(x + h e_vd) - E_bh(x)) / h, a fresh tree per moved
 * coordinate (tpe_stage2_oracle.py:133-141, 1289-1290). AoS out. */
rc_status rc_gradient_bh(rc_ctx* ctx, const double* xyz, const rc_params* p, double theta,
                         double theta_t, int32_t leaf_size, double h, double* grad);
/* Block-cluster-tree products of the fractional-Sobolev edge kernels: the
 * reference's stage-2 oracle bct_matvec_kernel (oracle/tpe_stage2_oracle.py:
 * 649-709): phi = K psi with psi, phi double[E][ncomp] (ncomp 1 or 3), kind
 * 0 'high' / 1 'low'; stats (nullable, 4): leaves, admissible blocks, direct
 * blocks, direct pairs. rc_bct_apply_b is bct_apply_B_scalar (:763-795):
 * out[V] = B u through two products. The tree is built on the host from xyz. */
rc_status rc_bct_matvec(rc_ctx* ctx, const double* xyz, double eps, const double* psi,
                        int32_t ncomp, int32_t kind, double theta, double theta_t,
                        int32_t leaf_size, double* phi, int64_t* stats);
rc_status rc_bct_apply_b(rc_ctx* ctx, const double* xyz, double eps, const double* u,
                         int32_t kind, double theta, double theta_t, int32_t leaf_size,
                         double* out);
/* assembleAFlat (innerProduct.ts:325-350): V x V row-major A = B + B0. */
rc_status rc_assemble_a(rc_ctx* ctx, const double* xyz, const rc_params* p, double* A);

/* penaltyEnergy / penaltyGradient (penalties.ts:139-241). grad nullable. */
rc_status rc_penalties_eval(rc_ctx* ctx, const double* xyz, const rc_penalties* pen,
                            double* energy, double* grad);

/* solveConstrainedGradientSetFrozen (gradient.ts:84-115): g~ and lambda for
 * a given dE (AoS). lambda has rc_constraint_rows() entries. */
rc_status rc_solve_constrained_gradient(rc_ctx* ctx, const double* xyz, const double* dE,
                                        const rc_block* blocks, int32_t nblocks,
                                        const rc_step_opts* o, double* gtilde, double* lambda,
                                        double* residual);
int32_t rc_constraint_rows(const rc_ctx* ctx, const rc_block* blocks, int32_t nblocks);

/* sobolevStepSet (optimizer.ts:231-440), projectionMode by
 * o->projection_frozen, dE by o->gradient_mode. Positions in/out are host AoS; gtilde_out
 * (collectField -> descentField) and timings are nullable. Never fails on a
 * singular saddle: that folds into stats->reason like the reference. */
rc_status rc_sobolev_step(rc_ctx* ctx, const double* xyz_in, const rc_block* blocks,
                          int32_t nblocks, const rc_penalties* pen, const rc_step_opts* o,
                          double* xyz_out, rc_step_stats* stats, double* gtilde_out,
                          rc_timings* timings);

/* Device-resident form for callers that keep the curve in HBM across steps:
 * rc_set_positions uploads once, rc_sobolev_step_resident advances the
 * context's device positions in place (an accepted step replaces them) with
 * one host round trip (the stats), rc_get_positions downloads. */
rc_status rc_set_positions(rc_ctx* ctx, const double* xyz);
rc_status rc_get_positions(rc_ctx* ctx, double* xyz);
rc_status rc_sobolev_step_resident(rc_ctx* ctx, const rc_block* blocks, int32_t nblocks,
                                   const rc_penalties* pen, const rc_step_opts* o,
                                   rc_step_stats* stats, rc_timings* timings);

/* Multigrid projected solver of the saddle system: the reference's stage-2
 * oracle mg_projected_solve / mg_saddles (oracle/tpe_stage2_oracle.py:796-1198).
 * The caller builds the level hierarchy on the host (curve coarsening,
 * per-level constraint Jacobians and (C C^T)^+, as the reference does in
 * Python); the solve runs on the device. Level 0 is the context's curve at
 * xyz (its Gram matrix is assembled); level l >= 1 gives, for every vertex of
 * level l-1, its (up to two) parents in level l and their weights (-1 = none),
 * and its operator is the Galerkin product P^T A_{l-1} P unless A is given
 * (the reference's 'rediscretize' option). Vectors (rhs, x) are 3n doubles in
 * coordinate blocks [x.., y.., z..]. blocks: the level-0 set, used by the
 * dense cleanup (the library's saddle solve). RC_DIVERGED when the residual
 * grows past 10x the best one or a cleanup fails its certificate. */
#define RC_MG_MAXLEVELS 16
#define RC_MG_MAXHIST 64
typedef struct {
    int32_t n;             /* vertices of this level */
    int32_t k;             /* constraint rows (<= 256) */
    const double* C;       /* k x 3n, row-major */
    const double* Gp;      /* k x k: pinv(C C^T) */
    const int32_t* parent; /* levels >= 1: 2 per vertex of level l-1 */
    const double* weight;  /* levels >= 1: 2 per vertex of level l-1 */
    const double* A;       /* optional n x n operator of this level (NULL: Galerkin) */
} rc_mg_level;
typedef struct {
    double tol;            /* projected residual target (1e-10) */
    int32_t max_cycles;    /* V-cycles (50) */
    int32_t nu_pre, nu_post;  /* projected-CG smoothing steps (2, 2) */
    int32_t init_nested;   /* 1: nested-iteration initial guess, 0: zero */
} rc_mg_opts;
typedef struct {
    int32_t iterations;    /* residual evaluations - 1, as the reference counts */
    int32_t used_direct_cleanup;
    int32_t nhist;
    int32_t _pad;
    double residual;       /* final projected residual */
    double history[RC_MG_MAXHIST];
} rc_mg_stats;
rc_status rc_mg_solve(rc_ctx* ctx, const double* xyz, const rc_params* p, const rc_block* blocks,
                      int32_t nblocks, const rc_mg_level* levels, int32_t nlevels,
                      const double* rhs, const rc_mg_opts* o, double* x, rc_mg_stats* stats);
/* y = A u per coordinate block with the Gram matrix of the curve at xyz
 * (u, y: 3V doubles in coordinate blocks). */
rc_status rc_gram_apply(rc_ctx* ctx, const double* xyz, const rc_params* p, const double* u,
                        double* y);

/* Kernel-level profile of the dominant kernel of the FP64 step, the
 * trailing update of the Cholesky factorisation (k_syrk128, all of the
 * trailing matrix but the next panel's column): while enabled,
 * a pair of HIP events brackets each of its launches on the context's stream.
 * No counterpart in the reference (measurement support for benchmarks).
 * Enabling (or disabling) resets the totals; reading synchronises the
 * context's stream and returns the totals since the reset. */
typedef struct {
    double syrk_ms;         /* summed launch durations */
    double syrk_flop;       /* algorithmic flops: 128 m (m + 1) per launch, m rows updated */
    int32_t syrk_launches;
    int32_t _pad;
} rc_kernel_profile;

rc_status rc_kernel_profile_enable(rc_ctx* ctx, int32_t enable);
rc_status rc_kernel_profile_read(rc_ctx* ctx, rc_kernel_profile* out);

#ifdef __cplusplus
}
#endif
#endif

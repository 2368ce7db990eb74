/*
 * rc_oracle — CPU FP64 restatement of the reference Repulsive Curves Sobolev
 * step, used ONLY as the checker for the HIP library (tests, smoke check and
 * the CPU-baseline leg of the bench). Nothing in the product links it.
 *
 * Every function follows the reference TypeScript op order where the
 * reference's own goldens pin it bit-for-bit (energy and analytic gradient),
 * and its algorithm elsewhere. Citations are to the reference repository.
 *
 * Layout: vertices are AoS double[V][3] (the reference's Vec3[]), edges are
 * int32[E][2], flattened vectors use the coordinate-block layout
 * [x0..x{V-1}, y0.., z0..] (src/core/sobolev/layout.ts:23-58).
 */
#ifndef RC_ORACLE_H
#define RC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constraint block kinds, in the reference's ConstraintBlock.kind order
 * (src/core/sobolev/constraintSet.ts:39-43). */
enum { RCO_BARYCENTER = 0, RCO_TOTAL_LENGTH = 1, RCO_EDGE_LENGTHS = 2, RCO_POINT = 3 };

/* Differential modes of step() and sobolevStepSet (optimizer.ts:21,96). */
enum { RCO_MODE_ANALYTICAL = 0, RCO_MODE_FINITE_DIFF = 1 };

/* Failure reasons (src/core/sobolev/lineSearch.ts:374-379, optimizer.ts:186). */
enum {
    RCO_REASON_NONE = 0,
    RCO_REASON_ZERO_OR_NONFINITE_GRADIENT_NORM = 1,
    RCO_REASON_NOT_A_DESCENT_DIRECTION = 2,
    RCO_REASON_ARMIJO_FAILED = 3,
    RCO_REASON_PROJECTION_FAILED = 4,
    RCO_REASON_TAU_BELOW_MIN = 5,
    RCO_REASON_SINGULAR_SYSTEM = 6,
};

typedef struct {
    int32_t kind;
    int32_t vertex;       /* RCO_POINT: pinned vertex index */
    double x0[3];         /* RCO_BARYCENTER: frozen target x0 */
    double L0;            /* RCO_TOTAL_LENGTH: frozen target L0 */
    double target[3];     /* RCO_POINT: frozen pin target */
    const double* ell0;   /* RCO_EDGE_LENGTHS: frozen per-edge targets */
    int32_t ell0_len;
    int32_t _pad;
} rco_block;

typedef struct {
    double w_length;
    double w_diff;
    double w_field;
    double X[3];
} rco_penalties;

typedef struct {
    double alpha, beta, epsilon;
    int32_t has_energy_before;
    int32_t projection_frozen;  /* 1 = 'frozen', 0 = 'reassemble' */
    double energy_before;
    int32_t factor_lu;          /* 1 = 'lu', 0 = 'ldlt' (default) */
    int32_t max_iter;           /* projection maxIter (default 8) */
    double tol_abs, tol_rel;    /* projection tolerances (default 1e-4) */
    double c1, shrink, tau0, tau_min;
    const rco_penalties* penalties; /* nullable */
    int32_t gradient_mode;      /* RCO_MODE_ANALYTICAL (default) | RCO_MODE_FINITE_DIFF */
    int32_t _pad2;
    double fd_h;                /* forward-difference step (default 1e-6, optimizer.ts:18) */
} rco_step_opts;

typedef struct {
    double energy;
    double energy_before;
    double tau;
    double residual;
    double gradient_l2_norm;
    double slope;
    double phi_norm;
    int32_t accepted;
    int32_t converged;
    int32_t projection_iterations; /* -1 = null */
    int32_t reason;
    int32_t trials;
    int32_t _pad;
} rco_step_stats;

/* Fills defaults matching the reference (optimizer.ts:18, lineSearch.ts:207-209,435-438). */
void rco_default_opts(rco_step_opts* o);

/* Threads of the pair loops (energy, gradient, Gram matrices); 0 = the
 * OpenMP default. Results are bit-identical for every thread count: terms
 * are computed in parallel and accumulated in the reference's order. */
void rco_set_threads(int n);
int rco_threads(void);

/* ---- energy / gradient (src/core/tangentPointEnergy.ts) ---- */
double rco_energy(int V, int E, const double* pos, const int32_t* edges,
                  double alpha, double beta, double eps);
void rco_gradient(int V, int E, const double* pos, const int32_t* edges,
                  double alpha, double beta, double eps, double* grad);
/* Forward-difference gradient (tangentPointEnergy.ts:114-136), AoS out. */
void rco_gradient_fd(int V, int E, const double* pos, const int32_t* edges,
                     double alpha, double beta, double eps, double h, double* grad);
/* Raw L2 step (optimizer.ts:40-68): out = pos - step_size * dE, energy at out;
 * field (nullable) receives dE. */
void rco_step(int V, int E, const double* pos, const int32_t* edges, double alpha, double beta,
              double eps, int mode, double step_size, double h, double* out, double* energy,
              double* field);
/* Barnes-Hut energy (the reference's stage-2 oracle, tpe_stage2_oracle.py:
 * 456-596): edge BVH with leaf_size, spatial ratio theta, tangent radius
 * theta_t. stats (nullable): visits, approximated edge refs, direct pairs. */
double rco_energy_bh(int V, int E, const double* pos, const int32_t* edges, double alpha,
                     double beta, double eps, double theta, double theta_t, int leaf_size,
                     int64_t* stats);
/* Forward difference of the Barnes-Hut energy, tree rebuilt per move. */
void rco_gradient_bh_fd(int V, int E, const double* pos, const int32_t* edges, double alpha,
                        double beta, double eps, double theta, double theta_t, int leaf_size,
                        double h, double* grad);
/* Block-cluster-tree products (the reference's stage-2 oracle,
 * tpe_stage2_oracle.py:598-795): phi = K psi for the edge kernel kind
 * (0 'high', 1 'low'), psi/phi [E][ncomp] (ncomp 1 or 3); stats (nullable, 4):
 * leaves, admissible blocks, direct blocks, direct pairs. rco_bct_apply_b is
 * bct_apply_B_scalar: out[V] = B u through two products. */
void rco_bct_matvec(int V, int E, const double* pos, const int32_t* edges, double eps,
                    const double* psi, int ncomp, int kind, double theta, double theta_t,
                    int leaf_size, double* phi, int64_t* stats);
void rco_bct_apply_b(int V, int E, const double* pos, const int32_t* edges, double eps,
                     const double* u, int kind, double theta, double theta_t, int leaf_size,
                     double* out);
/* Same energy with a compensated (Neumaier) outer sum: separates the
 * reference's left-to-right rounding from device error at large E. */
double rco_energy_compensated(int V, int E, const double* pos, const int32_t* edges,
                              double alpha, double beta, double eps);

/* ---- Gram matrices (src/core/sobolev/innerProduct.ts) ---- */
void rco_assemble_b_high(int V, int E, const double* pos, const int32_t* edges,
                         double alpha, double beta, double eps, double* B);
void rco_assemble_b_low(int V, int E, const double* pos, const int32_t* edges,
                        double alpha, double beta, double eps, double* B0);
void rco_assemble_a(int V, int E, const double* pos, const int32_t* edges,
                    double alpha, double beta, double eps, double* A);

/* Checker-only: A with each entry accumulated in long double and rounded
 * once (same FP64 pair weights; not a reference function). */
void rco_assemble_a_ld(int V, int E, const double* pos, const int32_t* edges,
                       double alpha, double beta, double eps, double* A);

/* ---- constraints (constraints.ts, constraintSet.ts, lineSearch.ts:86-109) ---- */
void rco_barycenter_target(int V, int E, const double* pos, const int32_t* edges, double* x0);
double rco_total_length(int V, int E, const double* pos, const int32_t* edges);
int rco_constraint_rows(int V, int E, int nblocks, const rco_block* blocks);
/* phi[k], C[k*3V] (row-major), scales[nblocks] (nullable), rows[nblocks] (nullable) */
void rco_eval_constraints(int V, int E, const double* pos, const int32_t* edges,
                          int nblocks, const rco_block* blocks,
                          double* phi, double* C, double* scales, int32_t* rows);

/* ---- penalties (src/core/sobolev/penalties.ts) ---- */
int rco_penalties_active(const rco_penalties* p);
double rco_penalty_energy(int V, int E, const double* pos, const int32_t* edges,
                          const rco_penalties* p);
void rco_penalty_gradient(int V, int E, const double* pos, const int32_t* edges,
                          const rco_penalties* p, double* g);

/* ---- dense solvers (src/core/sobolev/linsolve.ts) ---- */
int rco_ldlt_factor(double* k, int n, int32_t* ipiv);
void rco_ldlt_solve(const double* m, const int32_t* ipiv, int n, double* b);
int rco_lu_factor(double* k, int n, int32_t* piv);
void rco_lu_solve(const double* m, const int32_t* piv, int n, const double* rhs, double* x);

/* Saddle solve from the scalar A (linsolve.ts:742-803). x[3V], lambda[k].
 * Returns 0, or 1 on a singular/non-finite system. */
int rco_solve_saddle_from_a(const double* A, int V, const double* C, int k,
                            const double* rhs_top, const double* rhs_bottom,
                            int factor_lu, double* x, double* lambda, double* residual);

/* rho = rhs - K z for K = [[I3 (x) A, C^T],[C, 0]], z = [x; lambda], in
 * double-double arithmetic, rounded once per entry (the checker's
 * extra-precise residual for refining saddle solutions; rhs parts nullable
 * = 0). A V x V row-major, C k x 3V, x/rho_top coordinate-block 3V. */
void rco_saddle_residual_dd(const double* A, int V, const double* C, int k,
                            const double* r_top, const double* r_bot, const double* x,
                            const double* lambda, double* rho_top, double* rho_bot);

/* ---- L2h norm (lineSearch.ts:53-75) ---- */
double rco_l2_curve_norm(int V, int E, const double* values, const double* pos,
                         const int32_t* edges);

/* Constrained gradient solve (gradient.ts:84-115): dE AoS[V][3] in, gtilde AoS out. */
int rco_solve_constrained_gradient(int V, int E, const double* pos, const int32_t* edges,
                                   double alpha, double beta, double eps,
                                   const double* dE, int nblocks, const rco_block* blocks,
                                   int factor_lu, double* gtilde, double* lambda,
                                   double* residual);

/* Line search (lineSearch.ts:423-562) with dE and gtilde given (AoS).
 * Frozen projection uses K(gamma0) built from pos. */
int rco_line_search(int V, int E, const double* pos, const int32_t* edges,
                    const double* dE, const double* gtilde,
                    int nblocks, const rco_block* blocks,
                    const rco_step_opts* o, double* pos_out, rco_step_stats* st);

/* One full Sobolev step (optimizer.ts:231-440); dE by o->gradient_mode. */
int rco_sobolev_step(int V, int E, const double* pos, const int32_t* edges,
                     int nblocks, const rco_block* blocks, const rco_step_opts* o,
                     double* pos_out, double* gtilde_out, rco_step_stats* st);

#ifdef __cplusplus
}
#endif
#endif

// Device-side descriptors shared by the step kernels (rc_step.hip) and the
// C-ABI layer (rc_api.hip).
#pragma once

#include <cstdint>

#define RC_MAXPIN 64
#define RC_LSG_NW 64                 // workgroups of the grid-wide line-search kernels
#define RC_LSG_PART (8 + RC_MAXPIN)  // doubles of one workgroup's partial record
#define RC_KMAX 1024  // constraint rows supported on the device
#define RC_KSMALL 60  // up to here the (k+3)^2 Schur system is factored in LDS; the
                      // context's k-sized buffers start at this capacity and grow

// ConstraintSet as POD (constraintSet.ts:39-52), in set order. At most one
// barycenter block and one length block (total or per-edge), plus pins —
// the shapes dispatchDescentStep builds (src/core/dispatch.ts:213-242).
struct DevSet {
    int32_t nblocks, k;
    int32_t kind[RC_MAXPIN + 4];
    int32_t row0[RC_MAXPIN + 4];
    int32_t rows[RC_MAXPIN + 4];
    int32_t pin_of_block[RC_MAXPIN + 4];   // pin slot for point blocks
    int32_t has_bary, has_len, len_per_edge, npin;
    int32_t len_row0;
    int32_t _pad;
    int32_t pin_vertex[RC_MAXPIN];
    double x0[3];
    double L0;
    double pin_target[RC_MAXPIN][3];
};

struct DevPen {
    double wL, wD, wF, X[3];   // resolved (X normalised, wF = 0 if |X| < 1e-14)
    int32_t active;
    int32_t _pad;
};

// Line-search control block (device memory). Doubles first.
struct LSCtl {
    double tau;
    double e0;            // objective at gamma0 (tpe + penalties)
    double slope;
    double gnorm;
    double e1;            // objective of the accepted trial
    double penE;          // penalty energy of the current trial
    double phi_norm;
    double residual;
    double c1, shrink, tau_min, tol_abs, tol_rel;
    double tau_acc;
    double refine_res;
    int32_t done;
    int32_t accepted;
    int32_t converged;
    int32_t reason;
    int32_t last_reason;
    int32_t trials;
    int32_t proj_iters;
    int32_t proj_ok;
    int32_t skip_energy;
    int32_t counter;      // last-block ticket of the energy kernel
    int32_t max_iter;
    int32_t refine_iters;
    int32_t proj_iters_acc;
    int32_t need_solve;   // reassemble mode: the trial waits for a solve at cur
    int32_t proj_it;      // reassemble mode: corrections applied to the trial so far
    int32_t energy_evals; // trial energies evaluated (phase ledger)
    int32_t proj_solves;  // projection correction solves over all trials (phase ledger)
};

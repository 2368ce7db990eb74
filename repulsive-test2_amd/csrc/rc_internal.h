// Internal declarations of librcurves: the device context, edge/vertex
// tables and the launchers each translation unit provides.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "rcurves.h"

#define RC_TILE 64          // edges per tile of the pair kernels (one wave64 lane per edge)
#define RC_VTILE 32         // vertices per ownership tile of the Gram assembly
#define RC_RMAX 36          // incident edges of one vertex tile staged in LDS by the assembly
                            // (larger tiles are walked in chunks of this many; 36 keeps
                            // the assembly at 39.7 KiB of LDS, four workgroups per CU)
#define RC_NB 64            // Cholesky / triangular-solve block size
#define RC_BWD_MAXCH 32     // most right-hand-side chunks in one backward-sweep launch
#define RC_SWEEP_G 4        // workgroups sharing one 128-row block of a triangular sweep
#define RC_SCAL_RESID 8     // d_scal slot of the saddle residual
#define RC_NOALLOC_EMAX 16384  // largest E of rc_energy_noalloc (its static scratch)
#define RED_T 1024          // threads of the one-workgroup-per-sum reductions (few long sums)

// Structure-of-arrays per-edge geometry refreshed by k_edge_prep for the
// current positions. Index = edge id; padded entries (id >= E) are unused.
struct EdgeSoA {
    double *ax, *ay, *az;   // position of endpoint i1
    double *bx, *by, *bz;   // position of endpoint i2
    double *ell;            // ||e|| + eps
    double *rlen;           // ||e|| (raw)
    double *ux, *uy, *uz;   // unit tangent, 0 when ||e|| < 1e-14
    double *lpow;           // ell^(1-alpha)
    double *dpc;            // (1-alpha) ell^(-alpha)
    int32_t *i1, *i2;
};

struct DevSet;
struct LSCtl;

struct rc_ctx {
    int V = 0, E = 0, Epad = 0, ntiles = 0, npairs = 0;
    int nvt = 0, ngram = 0;   // vertex tiles, assembly tile pairs
    int nblk = 0;             // Cholesky blocks of RC_NB rows (always even)
    int bwd_chunks = 1;       // right-hand-side chunks per backward-sweep launch
    int maxdeg = 0;
    int wide_tiles = 0;       // some vertex tile touches more than RC_RMAX edges
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // factorisation lookahead (trailing updates)
    std::vector<hipEvent_t> dep_ev; // factorisation stream dependencies
    hipEvent_t asm_go = nullptr;    // step: edge table ready (second stream may assemble)
    hipEvent_t asm_done = nullptr;  // step: Gram assembly done
    hipEvent_t res_go = nullptr;    // step: solution ready (second stream may form the residual)
    hipEvent_t res_done = nullptr;  // step: residual done
    std::string err;

    // topology (device)
    int32_t* d_i1 = nullptr;
    int32_t* d_i2 = nullptr;
    int32_t* d_vOff = nullptr;    // vertex -> incident (edge, slot) CSR
    int32_t* d_vEdge = nullptr;
    int32_t* d_vSlot = nullptr;   // 0 if vertex is i1 of the edge, 1 if i2
    int32_t* d_vLocal = nullptr;  // local index of the edge in its vertex tile's edge list
    int32_t* d_rOff = nullptr;    // vertex tile -> incident edge list CSR
    int32_t* d_rEdge = nullptr;
    int32_t* d_owner = nullptr;   // owner vertex tile of each edge (tile of i1)
    int32_t* d_interior = nullptr; // (v, I_v, J_v) for degree-2 vertices, ascending v
    int ninterior = 0;
    std::vector<int32_t> h_edges;  // host copy (E x 2)

    // geometry (device)
    double* d_pos = nullptr;      // SoA x[V], y[V], z[V] current curve
    double* d_work = nullptr;     // SoA scratch positions (trial / uploads)
    EdgeSoA ed{};
    double* d_edge_block = nullptr;
    int32_t* d_edge_iblock = nullptr;

    // pair-kernel reductions
    double* d_epart = nullptr;    // energy partial per tile pair
    double* d_gpart = nullptr;    // gradient partials [ntiles][6][Epad]
    double* d_grad = nullptr;     // SoA gradient (3V)
    double* d_scal = nullptr;     // small FP64 scalars (energy results etc.)

    // Gram assembly
    double* d_rsp = nullptr;      // row-sum partials [nvt][2][E]
    double* d_rs = nullptr;       // row sums W_K, Q_K (2E)
    double* d_F = nullptr;        // Vp x Vp FP64: A (Gram assembly output), then its
                                  // factor (lower) of A + mu 11^T with A kept in the
                                  // strict upper triangle, the blocks below the 128 x 128
                                  // block diagonal row-scaled (L~_RC = L_RR^-1 L_RC);
                                  // below, the right-hand-side rows (B^T, then (L^-1 B)^T)
    double* d_Adiag = nullptr;    // diagonal of A (Vp), saved before the factorisation
    double* d_symv = nullptr;     // residual product partials [nblk][nblk][3][NB]
    float* d_F32 = nullptr;       // V x V FP32 scaled factor (mixed form), row-scaled below
                                  // the 128 x 128 block diagonal like d_F
    double* d_Linv128 = nullptr;  // inverted 128-column diagonal blocks [nblk/2][128][128]
    float* d_Linv32 = nullptr;    // inverted 64-column diagonal blocks of d_F32 [nblk][64][64]
    double* d_X128m = nullptr;    // inverted 128-row diagonal blocks of d_F32 [nblk/2][128][128]
    double* d_dscale = nullptr;   // symmetric diagonal scaling (mixed form)
    double mixed_shift = 0.0;     // diagonal shift of the last FP32 factorisation
    double mixed_shift_start = 0.0; // first shift the next FP32 factorisation tries (0: 2^-23)
    int32_t fp64_fallback = 0;    // the last mixed solve was redone with the FP64 factor
    int32_t* d_flags = nullptr;   // triangular-solve block flags
    int32_t* d_fcnt = nullptr;    // factorisation completion counters [2][nblk]
    int32_t epoch = 0;
    uint32_t* d_ticket = nullptr; // triangular-solve ticket counter (only grows)
    uint32_t tickets = 0;         // tickets issued so far (mod 2^32)
    int32_t* d_status = nullptr;  // device status words (factor breakdown, spin timeouts)

    // solve workspace
    int nrhs_cap = 0;
    double* d_rhs = nullptr;      // [nrhs][V]
    double* d_sol = nullptr;      // [nrhs][V]
    double* d_tmp = nullptr;      // [nrhs][V]
    double* d_tmp2 = nullptr;
    double* d_tmp3 = nullptr;
    float* d_tmp32 = nullptr;

    // step state (device)
    int kcap = 0;
    double* d_C = nullptr;        // k x 3V constraint Jacobian at gamma0
    double* d_phi = nullptr;
    double* d_blockdata = nullptr; // per-block targets
    double* d_ell0 = nullptr;      // edge-length targets
    double* d_schur = nullptr;     // (k+3)^2 factored Schur matrix + pivots
    double* d_ls = nullptr;        // line-search state scalars
    double* d_dir = nullptr;       // SoA direction (3V)
    double* d_dE = nullptr;        // SoA total differential (3V)
    double* d_gt = nullptr;        // SoA g~ (3V)
    double* d_Y = nullptr;         // A^-1 C^T blocks [3k][V] (per coordinate block)
    double* d_stage = nullptr;     // 3V AoS staging for uploads/downloads
    double* d_host_stage = nullptr; // pinned staging
    size_t host_stage_bytes = 0;
    bool has_isolated = false;

    // step scratch
    DevSet* d_set = nullptr;
    LSCtl* d_ctl = nullptr;
    double* d_dots = nullptr;
    double* d_bsum = nullptr;
    double* d_ssol = nullptr;
    double* d_Minv = nullptr;
    double* d_Ax = nullptr;
    double* d_Cg = nullptr;
    double* d_stats = nullptr;
    double* d_epart_ls = nullptr;
    double* d_lsp = nullptr;       // grid line-search partial records [RC_LSG_NW][RC_LSG_PART]
    uint32_t* d_gsync = nullptr;   // grid line-search counters [4], zero between launches
    double* d_refine = nullptr;   // [0..1] PCG (residual, iterations); [2..7] saddle refinement
    double* d_ir = nullptr;       // saddle refinement workspace (rc_refine.hip), on first use
    size_t ir_doubles = 0;

    // Barnes-Hut tree on the device (rc_energy_bh), grown on demand
    double* d_bh_f = nullptr;       // 6 doubles per node
    int32_t* d_bh_i = nullptr;      // 4 ints per node + node order + positions
    unsigned long long* d_bh_stats = nullptr;
    size_t bh_cap_d = 0, bh_cap_i = 0;
    int32_t* d_bh_front = nullptr;  // traversal work items (frontier of the tree)
    size_t bh_front_cap = 0;
    double* d_bh_part = nullptr;    // work-item sums when the gradient partials are too small
    size_t bh_part_cap = 0;
    // block-cluster-tree products (rc_bct_matvec, rc_bct_apply_b), grown on demand
    double* d_bct_f = nullptr;
    int32_t* d_bct_i = nullptr;
    size_t bct_cap_d = 0, bct_cap_i = 0;

    // kernel-level profile (rc_kernel_profile_enable / _read)
    int kprof = 0;
    std::vector<hipEvent_t> kev;   // event pool, pairs around trailing-update launches
    std::vector<double> kev_flop;  // algorithmic flops of each recorded launch
    size_t kev_used = 0;           // events in use since the last read
    std::vector<hipEvent_t> led_ev; // event pool of the step's phase ledger
    double kprof_ms = 0.0, kprof_flop = 0.0;
    int kprof_launches = 0;
};

// ---- helpers --------------------------------------------------------------
#define RC_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) {                                             \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(_e);   \
            return RC_HIP_ERROR;                                            \
        }                                                                   \
    } while (0)

#define RC_CHECK_LAUNCH()                                                   \
    do {                                                                    \
        hipError_t _e = hipGetLastError();                                  \
        if (_e != hipSuccess) {                                             \
            ctx->err = std::string("kernel launch: ") + hipGetErrorString(_e); \
            return RC_HIP_ERROR;                                            \
        }                                                                   \
    } while (0)

// launchers (defined in the .hip files)
rc_status rc_launch_edge_prep(rc_ctx* ctx, const double* pos, const rc_params* p);
// triangular solves with the row-scaled factors (rc_sweep.hip)
rc_status rc_ltilde_fp64(rc_ctx* ctx);
rc_status rc_ltilde_f32(rc_ctx* ctx);
rc_status rc_solve_fp64_forward(rc_ctx* ctx, double* d_B, int nrhs);
rc_status rc_solve_fp64_backward(rc_ctx* ctx, double* d_Y, int nrhs);
rc_status rc_precond_f32(rc_ctx* ctx, const double* R, double* W, int nrhs,
                         const int32_t* skip);
int rc_trsm_max_coresident(void);
int rc_sweep_chunks(int nblk);
rc_status rc_launch_energy(rc_ctx* ctx, const rc_params* p, double* d_out);
rc_status rc_launch_gradient(rc_ctx* ctx, const rc_params* p, double* d_grad_soa);
rc_status rc_launch_gradient_fd(rc_ctx* ctx, const rc_params* p, double h, double* d_grad_soa);
rc_status rc_launch_l2_update(rc_ctx* ctx, const double* pos, const double* grad,
                              double step_size, double* out);
rc_status rc_launch_gram(rc_ctx* ctx, const rc_params* p);
rc_status rc_launch_gram(rc_ctx* ctx, const rc_params* p, hipStream_t st);
rc_status rc_upload_positions_aos(rc_ctx* ctx, const double* xyz, double* d_dst);
rc_status rc_download_positions_aos(rc_ctx* ctx, const double* d_src, double* xyz);

// C ABI of librcurves: context lifecycle, topology tables and the step
// orchestration (one HIP stream per context, synchronous entry points).
#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "rc_internal.h"
#include "rc_step.h"

rc_status rc_factor_fp64(rc_ctx* ctx, int nx);
rc_status rc_factor_mixed(rc_ctx* ctx);
rc_status rc_solve_mixed(rc_ctx* ctx, const double* d_B, double* d_X, int nrhs, double tol,
                         int max_iter, double* d_refine_out);
rc_status rc_launch_constraint_jacobian(rc_ctx* ctx, const DevSet* d_set, int k);
rc_status rc_launch_penalty_grad(rc_ctx* ctx, const DevPen& pen, const double* gsrc, int Vsrc,
                                 double* gdst, int Vdst);
rc_status rc_launch_penalty_energy(rc_ctx* ctx, const DevPen& pen, double* d_out);
rc_status rc_launch_build_rhs(rc_ctx* ctx, const double* dE, int k, double* B, int ncols);
rc_status rc_launch_schur(rc_ctx* ctx, const double* X, const double* dE, int k, double* gt,
                          double* d_dots, double* d_bsum, double* d_sol, double* d_Minv);
rc_status rc_launch_residual(rc_ctx* ctx, hipStream_t st, const double* gt, const double* dE,
                             int k, const double* d_sol, double* d_Ax, double* d_Cg,
                             double* d_out);
rc_status rc_launch_ls_setup(rc_ctx* ctx, const double* gt, const double* dE, double* dir,
                             const double* e0tpe, const double* e0pen, int has_e0,
                             double e0_given, double tau0, LSCtl* d_ctl);
rc_status rc_launch_ls_project(rc_ctx* ctx, int k, LSCtl* d_ctl, const DevSet* d_set,
                               const double* pos0, const double* dir, double* cur,
                               const double* Y, const double* Minv);
rc_status rc_launch_ls_finalize(rc_ctx* ctx, LSCtl* d_ctl, double* pos, const double* cur,
                                int update, double* d_stats);
rc_status rc_launch_ls_project_re(rc_ctx* ctx, int phase, LSCtl* d_ctl, const DevSet* d_set,
                                  const DevPen& pen, const rc_params* p, const double* pos0,
                                  const double* dir, double* cur, const double* Y,
                                  const double* Minv);
rc_status rc_launch_ls_energy(rc_ctx* ctx, LSCtl* d_ctl, const rc_params* p, const DevPen& pen,
                              const double* cur, double* d_epart_ls);
rc_status rc_saddle_refine(rc_ctx* ctx, int k, int precision, int nsteps, double pcg_tol,
                           int pcg_max_iter, double* d_out);
rc_status rc_refine_columns(rc_ctx* ctx, int k, int precision, int nsteps, double pcg_tol,
                            int pcg_max_iter, double* d_out);
rc_status rc_refine_zero(rc_ctx* ctx, double* d_out);
rc_status rc_saddle_final_residual(rc_ctx* ctx, int k, hipStream_t st, double* d_out);

namespace {

template <class T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
}

// PCG diagnostics into the control block: residual of the gradient solve,
// iterations of it and of the refinement corrections (mixed form).
__global__ void k_copy_refine(LSCtl* c, const double* r) {
    c->refine_res = r[0];
    c->refine_iters = (int)(r[1] + r[5]);
}

__global__ void k_pick_output(const double* __restrict__ pos, const double* __restrict__ cur,
                              const LSCtl* __restrict__ c, int V, double* __restrict__ o) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double* s = c->accepted ? cur : pos;
    o[3 * v] = s[v];
    o[3 * v + 1] = s[V + v];
    o[3 * v + 2] = s[2 * V + v];
}

}  // namespace

extern "C" void rc_default_step_opts(rc_step_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->params.alpha = 3.0;
    o->params.beta = 6.0;
    o->params.epsilon = 1e-10;
    o->has_energy_before = 0;
    o->precision = RC_PRECISION_FP64;
    o->projection_frozen = 1;
    o->max_iter = 8;
    o->tol_abs = 1e-4;
    o->tol_rel = 1e-4;
    o->c1 = 1e-4;
    o->shrink = 0.5;
    o->tau0 = 1.0;
    o->tau_min = 1e-12;
    o->refine_tol = 1e-12;
    o->refine_max_iter = 200;
    o->collect_timings = 0;
    o->gradient_mode = RC_MODE_ANALYTICAL;
    o->fd_h = 1e-6;
    o->saddle_refine = 1;
    o->refine_columns = 0;
    o->saddle_refine_tol = 1e-6;
    o->report_ir_residual = 0;
}

extern "C" const char* rc_last_error(const rc_ctx* ctx) {
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

static void free_all(rc_ctx* ctx) {
    void* ptrs[] = {ctx->d_i1, ctx->d_i2, ctx->d_vOff, ctx->d_vEdge, ctx->d_vSlot, ctx->d_vLocal,
                    ctx->d_rOff, ctx->d_rEdge, ctx->d_owner, ctx->d_interior, ctx->d_pos,
                    ctx->d_work, ctx->d_edge_block, ctx->d_edge_iblock, ctx->d_epart,
                    ctx->d_gpart, ctx->d_grad, ctx->d_scal, ctx->d_rsp, ctx->d_rs, ctx->d_Adiag,
                    ctx->d_symv,
                    ctx->d_F, ctx->d_F32, ctx->d_Linv128, ctx->d_Linv32, ctx->d_X128m, ctx->d_dscale,
                    ctx->d_flags, ctx->d_ticket, ctx->d_fcnt, ctx->d_status, ctx->d_rhs, ctx->d_sol, ctx->d_tmp,
                    ctx->d_tmp2, ctx->d_tmp3, ctx->d_tmp32, ctx->d_C, ctx->d_phi,
                    ctx->d_blockdata, ctx->d_ell0, ctx->d_schur, ctx->d_ls, ctx->d_dir,
                    ctx->d_dE, ctx->d_gt, ctx->d_Y};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    void* sp[] = {ctx->d_set, ctx->d_ctl, ctx->d_dots, ctx->d_bsum, ctx->d_ssol, ctx->d_Minv,
                  ctx->d_Ax, ctx->d_Cg, ctx->d_stats, ctx->d_epart_ls, ctx->d_lsp, ctx->d_gsync, ctx->d_refine, ctx->d_ir,
                  ctx->d_stage, ctx->d_bh_f, ctx->d_bh_i, ctx->d_bh_stats, ctx->d_bct_f,
                  ctx->d_bct_i, ctx->d_bh_front, ctx->d_bh_part};
    for (void* p : sp)
        if (p) (void)hipFree(p);
    if (ctx->d_host_stage) (void)hipHostFree(ctx->d_host_stage);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
}

extern "C" void rc_ctx_destroy(rc_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    for (hipEvent_t e : ctx->kev) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->dep_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->led_ev) (void)hipEventDestroy(e);
    if (ctx->asm_go) (void)hipEventDestroy(ctx->asm_go);
    if (ctx->asm_done) (void)hipEventDestroy(ctx->asm_done);
    if (ctx->res_go) (void)hipEventDestroy(ctx->res_go);
    if (ctx->res_done) (void)hipEventDestroy(ctx->res_done);
    free_all(ctx);
    delete ctx;
}

static rc_status build_topology(rc_ctx* ctx, const int32_t* edges) {
    const int V = ctx->V, E = ctx->E;
    std::vector<int32_t> i1(ctx->Epad, -1), i2(ctx->Epad, -2);
    std::vector<int32_t> deg(V, 0);
    for (int K = 0; K < E; ++K) {
        const int a = edges[2 * K], b = edges[2 * K + 1];
        if (a < 0 || a >= V || b < 0 || b >= V) {
            ctx->err = "edge " + std::to_string(K) + " has an endpoint outside [0, V)";
            return RC_BAD_ARG;
        }
        if (a == b) {
            ctx->err = "edge " + std::to_string(K) + " is a self-loop";
            return RC_BAD_ARG;
        }
        i1[K] = a;
        i2[K] = b;
        deg[a]++;
        deg[b]++;
    }
    std::vector<int32_t> vOff(V + 1, 0), vEdge(2 * (size_t)E), vSlot(2 * (size_t)E),
        vLocal(2 * (size_t)E);
    for (int v = 0; v < V; ++v) vOff[v + 1] = vOff[v] + deg[v];
    std::vector<int32_t> fill(vOff.begin(), vOff.end() - 1);
    for (int K = 0; K < E; ++K) {
        vEdge[fill[i1[K]]] = K;
        vSlot[fill[i1[K]]++] = 0;
        vEdge[fill[i2[K]]] = K;
        vSlot[fill[i2[K]]++] = 1;
    }
    int maxdeg = 0;
    bool iso = false;
    for (int v = 0; v < V; ++v) {
        maxdeg = std::max(maxdeg, deg[v]);
        if (deg[v] == 0) iso = true;
    }
    ctx->maxdeg = maxdeg;
    ctx->has_isolated = iso;
    // vertex tiles and their incident edge lists
    const int nvt = ctx->nvt;
    std::vector<int32_t> rOff(nvt + 1, 0), rEdge;
    for (int t = 0; t < nvt; ++t) {
        std::vector<int32_t> lst;
        for (int v = t * RC_VTILE; v < std::min(V, (t + 1) * RC_VTILE); ++v)
            for (int q = vOff[v]; q < vOff[v + 1]; ++q) lst.push_back(vEdge[q]);
        std::sort(lst.begin(), lst.end());
        lst.erase(std::unique(lst.begin(), lst.end()), lst.end());
        // tiles touching more edges than the assembly stages in LDS are
        // assembled in edge chunks (k_gram_tiles_wide)
        if ((int)lst.size() > RC_RMAX) ctx->wide_tiles = 1;
        for (int v = t * RC_VTILE; v < std::min(V, (t + 1) * RC_VTILE); ++v)
            for (int q = vOff[v]; q < vOff[v + 1]; ++q)
                vLocal[q] = (int32_t)(std::lower_bound(lst.begin(), lst.end(), vEdge[q]) -
                                      lst.begin());
        rEdge.insert(rEdge.end(), lst.begin(), lst.end());
        rOff[t + 1] = (int32_t)rEdge.size();
    }
    std::vector<int32_t> owner(E);
    for (int K = 0; K < E; ++K) owner[K] = i1[K] / RC_VTILE;

    RC_HIP(dalloc(&ctx->d_i1, ctx->Epad));
    RC_HIP(dalloc(&ctx->d_i2, ctx->Epad));
    RC_HIP(dalloc(&ctx->d_vOff, V + 1));
    RC_HIP(dalloc(&ctx->d_vEdge, 2 * (size_t)E));
    RC_HIP(dalloc(&ctx->d_vSlot, 2 * (size_t)E));
    RC_HIP(dalloc(&ctx->d_vLocal, 2 * (size_t)E));
    RC_HIP(dalloc(&ctx->d_rOff, nvt + 1));
    RC_HIP(dalloc(&ctx->d_rEdge, rEdge.size()));
    RC_HIP(dalloc(&ctx->d_owner, E));
    RC_HIP(hipMemcpy(ctx->d_i1, i1.data(), sizeof(int32_t) * ctx->Epad, hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(ctx->d_i2, i2.data(), sizeof(int32_t) * ctx->Epad, hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(ctx->d_vOff, vOff.data(), sizeof(int32_t) * (V + 1), hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(ctx->d_vEdge, vEdge.data(), sizeof(int32_t) * 2 * E, hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(ctx->d_vSlot, vSlot.data(), sizeof(int32_t) * 2 * E, hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(ctx->d_vLocal, vLocal.data(), sizeof(int32_t) * 2 * E, hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(ctx->d_rOff, rOff.data(), sizeof(int32_t) * (nvt + 1), hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(ctx->d_rEdge, rEdge.data(), sizeof(int32_t) * rEdge.size(),
                     hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(ctx->d_owner, owner.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice));
    ctx->h_edges.assign(edges, edges + 2 * (size_t)E);
    return RC_OK;
}

// Buffers sized by the number of constraint rows k: the factor storage with
// the 3 + 3k right-hand-side rows it carries below A, the solve columns, C,
// the Schur system and its workspace. Allocated for RC_KSMALL rows at context
// creation and grown (never shrunk) when a constraint set needs more; called
// between steps only (nothing in flight uses them).
static rc_status ensure_k_capacity(rc_ctx* ctx, int k) {
    if (k <= ctx->kcap) return RC_OK;
    const int kc = std::max(k, RC_KSMALL);
    const int Vp = ctx->nblk * RC_NB;
    const size_t nrhs = 3 + 3 * (size_t)kc;
    const size_t frows = (size_t)Vp + RC_NB * ((nrhs + RC_NB - 1) / RC_NB);
    const size_t n = (size_t)kc + 3;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    double** bufs[] = {&ctx->d_F, &ctx->d_rhs, &ctx->d_sol, &ctx->d_C, &ctx->d_dots, &ctx->d_bsum,
                       &ctx->d_ssol, &ctx->d_Minv, &ctx->d_Cg, &ctx->d_phi, &ctx->d_schur};
    const size_t sizes[] = {frows * Vp, nrhs * Vp, nrhs * Vp, (size_t)kc * 3 * Vp,
                            (size_t)kc * (kc + 1), 3 * ((size_t)kc + 1), n, n * n, (size_t)kc,
                            (size_t)kc + 8, kc > RC_KSMALL ? 3 * n * n + 2 * n : 1};
    ctx->kcap = 0;
    for (int i = 0; i < 11; ++i) {
        if (*bufs[i]) (void)hipFree(*bufs[i]);
        *bufs[i] = nullptr;
    }
    // the mixed form's PCG workspace is sized by the right-hand-side count:
    // dropped here, reallocated at its next use
    if (ctx->d_tmp2) (void)hipFree(ctx->d_tmp2);
    if (ctx->d_tmp3) (void)hipFree(ctx->d_tmp3);
    ctx->d_tmp2 = nullptr;
    ctx->d_tmp3 = nullptr;
    // so is the saddle refinement's
    if (ctx->d_ir) (void)hipFree(ctx->d_ir);
    ctx->d_ir = nullptr;
    for (int i = 0; i < 11; ++i) {
        hipError_t he = dalloc(bufs[i], sizes[i]);
        if (he != hipSuccess) {
            ctx->err = "hipMalloc (buffers for " + std::to_string(kc) +
                       " constraint rows): " + hipGetErrorString(he);
            return RC_HIP_ERROR;
        }
    }
    // the factor storage starts finite (its padding outside V x V is read by
    // the factorisation's tiles)
    RC_HIP(hipMemset(ctx->d_F, 0, sizeof(double) * frows * Vp));
    ctx->kcap = kc;
    ctx->nrhs_cap = (int)nrhs;
    return RC_OK;
}

extern "C" rc_status rc_ctx_create(const int32_t* edges, int32_t E, int32_t V, rc_ctx** out) {
    if (!out) return RC_BAD_ARG;
    *out = nullptr;
    if (!edges || E < 1 || V < 2) return RC_BAD_ARG;
    rc_ctx* ctx = new (std::nothrow) rc_ctx();
    if (!ctx) return RC_BAD_ARG;
    ctx->V = V;
    ctx->E = E;
    ctx->Epad = ((E + RC_TILE - 1) / RC_TILE) * RC_TILE;
    ctx->ntiles = ctx->Epad / RC_TILE;
    ctx->npairs = ctx->ntiles * (ctx->ntiles + 1) / 2;
    ctx->nvt = (V + RC_VTILE - 1) / RC_VTILE;
    ctx->ngram = ctx->nvt * (ctx->nvt + 1) / 2;
    // an even number of 64-row blocks: the FP64 factor works on 128-column panels
    ctx->nblk = 2 * ((V + 2 * RC_NB - 1) / (2 * RC_NB));
    auto fail = [&](rc_status s) {
        *out = ctx;  // keep the context so rc_last_error can report; caller destroys
        return s;
    };
    hipError_t he = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
    if (he != hipSuccess) {
        ctx->err = std::string("hipStreamCreate: ") + hipGetErrorString(he);
        return fail(RC_HIP_ERROR);
    }
    rc_status s = build_topology(ctx, edges);
    if (s != RC_OK) return fail(s);
    const int Vp = ctx->nblk * RC_NB;
    const int Epad = ctx->Epad;
#define ALLOC(ptr, n)                                                        \
    do {                                                                     \
        hipError_t _e = dalloc(&(ptr), (n));                                 \
        if (_e != hipSuccess) {                                              \
            ctx->err = std::string("hipMalloc(" #ptr "): ") + hipGetErrorString(_e); \
            return fail(RC_HIP_ERROR);                                       \
        }                                                                    \
    } while (0)
    ALLOC(ctx->d_pos, 3 * (size_t)V);
    ALLOC(ctx->d_work, 3 * (size_t)V);
    ALLOC(ctx->d_edge_block, 14 * (size_t)Epad);
    double* eb = ctx->d_edge_block;
    double** fields[] = {&ctx->ed.ax, &ctx->ed.ay, &ctx->ed.az, &ctx->ed.bx, &ctx->ed.by,
                         &ctx->ed.bz, &ctx->ed.ell, &ctx->ed.rlen, &ctx->ed.ux, &ctx->ed.uy,
                         &ctx->ed.uz, &ctx->ed.lpow, &ctx->ed.dpc};
    for (int i = 0; i < 13; ++i) *fields[i] = eb + (size_t)i * Epad;
    ALLOC(ctx->d_edge_iblock, 2 * (size_t)Epad);
    ctx->ed.i1 = ctx->d_edge_iblock;
    ctx->ed.i2 = ctx->d_edge_iblock + Epad;
    ALLOC(ctx->d_epart, ctx->npairs);
    ALLOC(ctx->d_gpart, (size_t)ctx->ntiles * 6 * Epad);
    ALLOC(ctx->d_grad, 3 * (size_t)Vp);
    ALLOC(ctx->d_scal, 16);
    ALLOC(ctx->d_rsp, (size_t)ctx->nvt * 2 * ctx->E);
    ALLOC(ctx->d_rs, 2 * (size_t)ctx->E);
    ALLOC(ctx->d_Adiag, (size_t)Vp);
    ALLOC(ctx->d_symv, (size_t)ctx->nblk * ctx->nblk * 3 * RC_NB);
    ALLOC(ctx->d_Linv128, (size_t)ctx->nblk * RC_NB * RC_NB * 2);
    ALLOC(ctx->d_flags, (size_t)ctx->nblk / 2 * RC_SWEEP_G * RC_BWD_MAXCH);
    ALLOC(ctx->d_ticket, 1);
    ALLOC(ctx->d_fcnt, 2 * (size_t)ctx->nblk);
    ALLOC(ctx->d_status, 4);
    ALLOC(ctx->d_tmp, 3 * (size_t)Vp);
    ALLOC(ctx->d_ell0, ctx->E);
    ALLOC(ctx->d_dir, 3 * (size_t)Vp);
    ALLOC(ctx->d_dE, 3 * (size_t)Vp);
    ALLOC(ctx->d_gt, 3 * (size_t)Vp);
    ALLOC(ctx->d_set, 1);
    ALLOC(ctx->d_ctl, 1);
    ALLOC(ctx->d_Ax, 3 * (size_t)Vp);
    ALLOC(ctx->d_stats, 24);
    ALLOC(ctx->d_epart_ls, 1024);
    ALLOC(ctx->d_lsp, (size_t)RC_LSG_NW * RC_LSG_PART);
    ALLOC(ctx->d_gsync, 4);
    ALLOC(ctx->d_refine, 8);
    ALLOC(ctx->d_stage, 3 * (size_t)V);
#undef ALLOC
    he = hipHostMalloc((void**)&ctx->d_host_stage, 4096, hipHostMallocDefault);
    if (he != hipSuccess) {
        ctx->err = std::string("hipHostMalloc: ") + hipGetErrorString(he);
        return fail(RC_HIP_ERROR);
    }
    ctx->host_stage_bytes = 4096;
    he = hipMemset(ctx->d_flags, 0, sizeof(int32_t) * ctx->nblk / 2 * RC_SWEEP_G * RC_BWD_MAXCH);
    if (he == hipSuccess) he = hipMemset(ctx->d_status, 0, sizeof(int32_t) * 4);
    if (he == hipSuccess) he = hipMemset(ctx->d_ticket, 0, sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(ctx->d_gsync, 0, sizeof(uint32_t) * 4);
    if (he == hipSuccess) he = hipMemset(ctx->d_Adiag, 0, sizeof(double) * (size_t)Vp);
    // the inverse's zero blocks above the diagonal are never written
    if (he == hipSuccess)
        he = hipMemset(ctx->d_Linv128, 0, sizeof(double) * (size_t)ctx->nblk * RC_NB * RC_NB * 2);
    if (he != hipSuccess) {
        ctx->err = std::string("hipMemset: ") + hipGetErrorString(he);
        return fail(RC_HIP_ERROR);
    }
    s = ensure_k_capacity(ctx, RC_KSMALL);
    if (s != RC_OK) return fail(s);
    // edge endpoint tables for the pair kernels
    he = hipMemcpy(ctx->ed.i1, ctx->d_i1, sizeof(int32_t) * Epad, hipMemcpyDeviceToDevice);
    if (he == hipSuccess)
        he = hipMemcpy(ctx->ed.i2, ctx->d_i2, sizeof(int32_t) * Epad, hipMemcpyDeviceToDevice);
    if (he != hipSuccess) {
        ctx->err = std::string("hipMemcpy: ") + hipGetErrorString(he);
        return fail(RC_HIP_ERROR);
    }
    ctx->bwd_chunks = rc_sweep_chunks(ctx->nblk);
    *out = ctx;
    return RC_OK;
}

// ---------------------------------------------------------------------------

static bool params_ok(const rc_params* p) {
    return p && std::isfinite(p->alpha) && std::isfinite(p->beta) && std::isfinite(p->epsilon) &&
           p->alpha != 0.0;
}

extern "C" rc_status rc_set_positions(rc_ctx* ctx, const double* xyz) {
    if (!ctx || !xyz) return RC_BAD_ARG;
    rc_status s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s != RC_OK) return s;
    RC_HIP(hipStreamSynchronize(ctx->stream));
    return RC_OK;
}

extern "C" rc_status rc_get_positions(rc_ctx* ctx, double* xyz) {
    if (!ctx || !xyz) return RC_BAD_ARG;
    return rc_download_positions_aos(ctx, ctx->d_pos, xyz);
}

extern "C" rc_status rc_energy(rc_ctx* ctx, const double* xyz, const rc_params* p, double* out) {
    if (!ctx || !xyz || !out || !params_ok(p)) return RC_BAD_ARG;
    rc_status s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s == RC_OK) s = rc_launch_energy(ctx, p, ctx->d_scal);
    if (s != RC_OK) return s;
    RC_HIP(hipMemcpyAsync(ctx->d_host_stage, ctx->d_scal, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->d_host_stage[0];
    return RC_OK;
}

extern "C" rc_status rc_gradient(rc_ctx* ctx, const double* xyz, const rc_params* p, double* grad,
                                 double* energy_out) {
    if (!ctx || !xyz || !grad || !params_ok(p)) return RC_BAD_ARG;
    rc_status s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s == RC_OK && energy_out) s = rc_launch_energy(ctx, p, ctx->d_scal);
    if (s == RC_OK) s = rc_launch_gradient(ctx, p, ctx->d_dE);
    if (s != RC_OK) return s;
    if (energy_out) {
        RC_HIP(hipMemcpyAsync(ctx->d_host_stage, ctx->d_scal, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
    }
    s = rc_download_positions_aos(ctx, ctx->d_dE, grad);
    if (s != RC_OK) return s;
    if (energy_out) *energy_out = ctx->d_host_stage[0];
    return RC_OK;
}

static bool fd_h_ok(double h) { return std::isfinite(h) && h != 0.0; }

extern "C" rc_status rc_gradient_fd(rc_ctx* ctx, const double* xyz, const rc_params* p, double h,
                                    double* grad) {
    if (!ctx || !xyz || !grad || !params_ok(p) || !fd_h_ok(h)) return RC_BAD_ARG;
    rc_status s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s == RC_OK) s = rc_launch_gradient_fd(ctx, p, h, ctx->d_dE);
    if (s != RC_OK) return s;
    return rc_download_positions_aos(ctx, ctx->d_dE, grad);
}

extern "C" rc_status rc_step(rc_ctx* ctx, const double* xyz_in, const rc_params* p, int32_t mode,
                             double step_size, double h, double* xyz_out, double* energy,
                             double* field) {
    if (!ctx || !xyz_in || !xyz_out || !energy || !params_ok(p)) return RC_BAD_ARG;
    if (mode != RC_MODE_ANALYTICAL && mode != RC_MODE_FINITE_DIFF) return RC_BAD_ARG;
    if (mode == RC_MODE_FINITE_DIFF && !fd_h_ok(h)) return RC_BAD_ARG;
    rc_status s = rc_upload_positions_aos(ctx, xyz_in, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s == RC_OK)
        s = mode == RC_MODE_ANALYTICAL ? rc_launch_gradient(ctx, p, ctx->d_dE)
                                       : rc_launch_gradient_fd(ctx, p, h, ctx->d_dE);
    if (s == RC_OK) s = rc_launch_l2_update(ctx, ctx->d_pos, ctx->d_dE, step_size, ctx->d_work);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_work, p);
    if (s == RC_OK) s = rc_launch_energy(ctx, p, ctx->d_scal);
    if (s != RC_OK) return s;
    RC_HIP(hipMemcpyAsync(ctx->d_host_stage, ctx->d_scal, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    *energy = ctx->d_host_stage[0];
    s = rc_download_positions_aos(ctx, ctx->d_work, xyz_out);
    if (s == RC_OK && field) s = rc_download_positions_aos(ctx, ctx->d_dE, field);
    return s;
}

extern "C" rc_status rc_assemble_a(rc_ctx* ctx, const double* xyz, const rc_params* p, double* A) {
    if (!ctx || !xyz || !A || !params_ok(p)) return RC_BAD_ARG;
    rc_status s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s == RC_OK) s = rc_launch_gram(ctx, p);
    if (s != RC_OK) return s;
    // A sits in the factor storage (leading dimension Vp)
    const size_t ld = sizeof(double) * (size_t)ctx->nblk * RC_NB;
    RC_HIP(hipMemcpy2DAsync(A, sizeof(double) * (size_t)ctx->V, ctx->d_F, ld,
                            sizeof(double) * (size_t)ctx->V, ctx->V, hipMemcpyDeviceToHost,
                            ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    return RC_OK;
}

static DevPen resolve_pen(const rc_penalties* pen) {
    DevPen d{};
    if (!pen) return d;
    d.wL = pen->w_length;
    d.wD = pen->w_diff;
    d.wF = pen->w_field;
    if (d.wF != 0.0) {
        const double n = std::sqrt(pen->X[0] * pen->X[0] + pen->X[1] * pen->X[1] + pen->X[2] * pen->X[2]);
        if (n < 1e-14) {
            d.wF = 0.0;
        } else {
            d.X[0] = pen->X[0] / n;
            d.X[1] = pen->X[1] / n;
            d.X[2] = pen->X[2] / n;
        }
    }
    d.active = (d.wL != 0.0 || d.wD != 0.0 || d.wF != 0.0) ? 1 : 0;
    return d;
}

extern "C" rc_status rc_penalties_eval(rc_ctx* ctx, const double* xyz, const rc_penalties* pen,
                                       double* energy, double* grad) {
    if (!ctx || !xyz || !energy) return RC_BAD_ARG;
    rc_params p{3.0, 6.0, 1e-10};
    const DevPen dp = resolve_pen(pen);
    rc_status s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, &p);
    if (s == RC_OK) s = rc_launch_penalty_energy(ctx, dp, ctx->d_scal + 4);
    if (s != RC_OK) return s;
    if (grad) {
        const int Vp = ctx->nblk * RC_NB;
        RC_HIP(hipMemsetAsync(ctx->d_tmp, 0, sizeof(double) * 3 * Vp, ctx->stream));
        s = rc_launch_penalty_grad(ctx, dp, ctx->d_tmp, Vp, ctx->d_dE, ctx->V);
        if (s == RC_OK) s = rc_download_positions_aos(ctx, ctx->d_dE, grad);
        if (s != RC_OK) return s;
    }
    RC_HIP(hipMemcpyAsync(ctx->d_host_stage, ctx->d_scal + 4, sizeof(double),
                          hipMemcpyDeviceToHost, ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    *energy = ctx->d_host_stage[0];
    return RC_OK;
}

// Translate the POD block list into the device descriptor (set order kept).
static rc_status make_set(rc_ctx* ctx, const rc_block* blocks, int32_t nblocks, DevSet& ds) {
    std::memset(&ds, 0, sizeof(ds));
    if (nblocks < 0 || nblocks > RC_MAXPIN + 2 || (nblocks > 0 && !blocks)) {
        ctx->err = "constraint block list: bad count";
        return RC_BAD_ARG;
    }
    int row = 0;
    for (int b = 0; b < nblocks; ++b) {
        const rc_block& B = blocks[b];
        ds.kind[b] = B.kind;
        ds.row0[b] = row;
        switch (B.kind) {
            case RC_BLOCK_BARYCENTER:
                if (ds.has_bary) {
                    ctx->err = "more than one barycenter block";
                    return RC_UNSUPPORTED;
                }
                ds.has_bary = 1;
                std::memcpy(ds.x0, B.x0, sizeof(ds.x0));
                ds.rows[b] = 3;
                break;
            case RC_BLOCK_TOTAL_LENGTH:
                if (ds.has_len) {
                    ctx->err = "more than one length block";
                    return RC_UNSUPPORTED;
                }
                ds.has_len = 1;
                ds.L0 = B.L0;
                ds.len_row0 = row;
                ds.rows[b] = 1;
                break;
            case RC_BLOCK_EDGE_LENGTHS: {
                if (ds.has_len) {
                    ctx->err = "more than one length block";
                    return RC_UNSUPPORTED;
                }
                ds.has_len = 1;
                ds.len_per_edge = 1;
                ds.len_row0 = row;
                ds.rows[b] = ctx->E;
                std::vector<double> e0(ctx->E, NAN);
                for (int i = 0; i < ctx->E && B.ell0 && i < B.ell0_len; ++i) e0[i] = B.ell0[i];
                RC_HIP(hipMemcpy(ctx->d_ell0, e0.data(), sizeof(double) * ctx->E,
                                 hipMemcpyHostToDevice));
                break;
            }
            case RC_BLOCK_POINT:
                if (ds.npin >= RC_MAXPIN) {
                    ctx->err = "too many point blocks";
                    return RC_UNSUPPORTED;
                }
                ds.pin_of_block[b] = ds.npin;
                ds.pin_vertex[ds.npin] = B.vertex;
                std::memcpy(ds.pin_target[ds.npin], B.target, sizeof(double) * 3);
                ds.npin++;
                ds.rows[b] = 3;
                break;
            default:
                ctx->err = "unknown constraint block kind";
                return RC_BAD_ARG;
        }
        row += ds.rows[b];
    }
    ds.nblocks = nblocks;
    ds.k = row;
    if (row > RC_KMAX) {
        ctx->err = "constraint rows " + std::to_string(row) + " exceed the device limit " +
                   std::to_string(RC_KMAX);
        return RC_UNSUPPORTED;
    }
    return ensure_k_capacity(ctx, row);
}

extern "C" int32_t rc_constraint_rows(const rc_ctx* ctx, const rc_block* blocks, int32_t nblocks) {
    if (!ctx) return -1;
    int k = 0;
    for (int b = 0; b < nblocks; ++b) {
        switch (blocks[b].kind) {
            case RC_BLOCK_BARYCENTER: k += 3; break;
            case RC_BLOCK_TOTAL_LENGTH: k += 1; break;
            case RC_BLOCK_EDGE_LENGTHS: k += ctx->E; break;
            case RC_BLOCK_POINT: k += 3; break;
            default: break;
        }
    }
    return k;
}

__global__ void k_aos_to_blocks(const double* __restrict__ aos, int V, int S,
                                double* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= S) return;
    for (int c = 0; c < 3; ++c) out[(size_t)c * S + v] = (v < V) ? aos[3 * v + c] : 0.0;
}

__global__ void k_blocks_to_aos(const double* __restrict__ in, int V, int S,
                                double* __restrict__ aos) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    for (int c = 0; c < 3; ++c) aos[3 * v + c] = in[(size_t)c * S + v];
}

// Gradient-solve phase shared by rc_solve_constrained_gradient and the step:
// assembly, constraints, factor, solves, Schur system, g~ and residual.
// dE (stride Vp) must be resident in ctx->d_dE; edge table at gamma0.
// gram_done (nullable): the Gram assembly was launched on the second stream
// and records this event when done; otherwise it is launched here.
// resid_done (nullable): run the residual on the second stream and record
// this event there when done (the caller waits for it before reading the
// residual); otherwise it runs on the main stream.
// Phase ledger of one step (phaseTimings.ts:9-78): event pairs on the
// stream each phase runs on, summed per key after the step's final
// synchronisation. A pair tagged with a trial index counts only when the line
// search examined that trial (the frozen search launches trials in batches;
// those past the decision return at once). Call counts the host cannot know
// (trial energies, projection correction solves) come from device counters.
enum LedgerKey { LK_DE, LK_ENERGY, LK_ASM, LK_SADDLE, LK_FACTOR, LK_PROJ, LK_LS, LK_STEP, LK_N };

struct Ledger {
    struct Mark {
        int key, trial;
        hipEvent_t a, b;
    };
    rc_ctx* ctx = nullptr;
    std::vector<Mark> marks;
    size_t used = 0;
    int calls[LK_N] = {};

    explicit Ledger(rc_ctx* c) : ctx(c) {}
    hipError_t event(hipEvent_t& e) {
        if (used == ctx->led_ev.size()) {
            hipEvent_t n;
            const hipError_t r = hipEventCreateWithFlags(&n, hipEventDisableSystemFence);
            if (r != hipSuccess) return r;
            ctx->led_ev.push_back(n);
        }
        e = ctx->led_ev[used++];
        return hipSuccess;
    }
    // Opens a pair on stream st; *id receives its index for close().
    hipError_t open(int key, hipStream_t st, int* id, int trial = -1) {
        Mark m{key, trial, nullptr, nullptr};
        hipError_t r = event(m.a);
        if (r == hipSuccess) r = hipEventRecord(m.a, st);
        if (r != hipSuccess) return r;
        marks.push_back(m);
        *id = (int)marks.size() - 1;
        return hipSuccess;
    }
    hipError_t close(int id, hipStream_t st) {
        hipError_t r = event(marks[id].b);
        return r == hipSuccess ? hipEventRecord(marks[id].b, st) : r;
    }
    // Sums the pairs into t (trials counted: 0 .. ntrials - 1).
    hipError_t sum(rc_timings* t, int ntrials) const {
        rc_phase* ph[LK_N] = {&t->dE, &t->energy, &t->assembleA, &t->saddle,
                              &t->factor, &t->projection, &t->lineSearch, &t->step};
        for (const Mark& m : marks) {
            if (m.trial >= ntrials || !m.b) continue;
            float ms = 0.0f;
            const hipError_t r = hipEventElapsedTime(&ms, m.a, m.b);
            if (r != hipSuccess) return r;
            ph[m.key]->ms += ms;
        }
        for (int k = 0; k < LK_N; ++k) ph[k]->calls += calls[k];
        return hipSuccess;
    }
};

// Records a ledger pair around a block of work when a ledger is armed.
#define LED_OPEN(led, key, st, id, ...)                                              \
    do {                                                                         \
        if (led) RC_HIP((led)->open((key), (st), &(id), ##__VA_ARGS__));         \
    } while (0)
#define LED_CLOSE(led, id, st)                                                   \
    do {                                                                         \
        if (led) RC_HIP((led)->close((id), (st)));                               \
    } while (0)

// The gradient solve (gradient.ts:88-120): assembly unless the caller ran it
// (gram_done), the constraint Jacobian, the factorisation, the solve and its
// refinement and the Schur system; the structured residual follows on the
// second stream when resid_go / resid_done are given. led: the step's ledger
// (saddle and factor; null = not timed).
static rc_status solve_phase(rc_ctx* ctx, const rc_params* p, const DevSet& ds, int precision,
                             const rc_step_opts* o, Ledger* led, hipEvent_t gram_done = nullptr,
                             hipEvent_t resid_go = nullptr, hipEvent_t resid_done = nullptr) {
    const int k = ds.k;
    ctx->fp64_fallback = 0;
    RC_HIP(hipMemcpyAsync(ctx->d_set, &ds, sizeof(DevSet), hipMemcpyHostToDevice, ctx->stream));
    rc_status s = RC_OK;
    if (gram_done) {
        RC_HIP(hipStreamWaitEvent(ctx->stream, gram_done, 0));
    } else {
        s = rc_launch_gram(ctx, p);
        if (s != RC_OK) return s;
    }
    int id_saddle = -1, id_factor = -1;
    LED_OPEN(led, LK_SADDLE, ctx->stream, id_saddle);
    if (k > 0) {
        s = rc_launch_constraint_jacobian(ctx, ctx->d_set, k);
        if (s != RC_OK) return s;
    }
    const int nrhs = 3 + 3 * k;
    const int Vp = ctx->nblk * RC_NB;
    // FP64: the right-hand sides [dE, C^T] ride below the factor as 64-row
    // blocks, so the factorisation leaves L^-1 B there
    const int nx = (nrhs + RC_NB - 1) / RC_NB;
    double* Frhs = ctx->d_F + (size_t)Vp * Vp;
    if (precision == RC_PRECISION_MIXED) {
        LED_OPEN(led, LK_FACTOR, ctx->stream, id_factor);
        s = rc_factor_mixed(ctx);
    } else {
        s = rc_launch_build_rhs(ctx, ctx->d_dE, k, Frhs, nx * RC_NB);
        if (s != RC_OK) return s;
        LED_OPEN(led, LK_FACTOR, ctx->stream, id_factor);
        s = rc_factor_fp64(ctx, nx);
    }
    if (s != RC_OK) return s;
    LED_CLOSE(led, id_factor, ctx->stream);
    if (ctx->has_isolated) {
        // isolated vertex: zero rows of A and C -> singular saddle (optimizer-sobolev.test.ts:83-105)
        int32_t one = 1;
        RC_HIP(hipMemcpyAsync(ctx->d_status, &one, sizeof(int32_t), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    if (precision == RC_PRECISION_MIXED) {
        s = rc_launch_build_rhs(ctx, ctx->d_dE, k, ctx->d_rhs, nrhs);
        if (s != RC_OK) return s;
        const double ptol = o ? o->refine_tol : 1e-12;
        s = rc_solve_mixed(ctx, ctx->d_rhs, ctx->d_sol, nrhs, ptol,
                           o ? o->refine_max_iter : 200, ctx->d_refine);
        if (s != RC_OK) return s;
        // a PCG that misses its tolerance within its iteration budget (the
        // FP32 preconditioner of a badly conditioned A_hat, kappa u_32 >> 1
        // on large curves) leaves an inaccurate solve: redo it with the
        // FP64 factorisation, on the device, and report it in the stats
        RC_HIP(hipMemcpyAsync(ctx->d_host_stage + 40, ctx->d_refine, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
        RC_HIP(hipStreamSynchronize(ctx->stream));
        if (!(ctx->d_host_stage[40] <= ptol)) {
            ctx->fp64_fallback = 1;
            precision = RC_PRECISION_FP64;
        }
    }
    if (precision != RC_PRECISION_MIXED && ctx->fp64_fallback) {
        s = rc_launch_build_rhs(ctx, ctx->d_dE, k, Frhs, nx * RC_NB);
        if (s != RC_OK) return s;
        int id_f64 = -1;
        LED_OPEN(led, LK_FACTOR, ctx->stream, id_f64);
        s = rc_factor_fp64(ctx, nx);
        if (s != RC_OK) return s;
        LED_CLOSE(led, id_f64, ctx->stream);
        if (led) led->calls[LK_FACTOR] += 1;
        RC_HIP(hipMemcpyAsync(ctx->d_sol, Frhs, sizeof(double) * (size_t)nrhs * Vp,
                              hipMemcpyDeviceToDevice, ctx->stream));
        s = rc_solve_fp64_backward(ctx, ctx->d_sol, nrhs);
        if (s != RC_OK) return s;
    } else if (precision != RC_PRECISION_MIXED) {
        RC_HIP(hipMemcpyAsync(ctx->d_sol, Frhs, sizeof(double) * (size_t)nrhs * Vp,
                              hipMemcpyDeviceToDevice, ctx->stream));
        s = rc_solve_fp64_backward(ctx, ctx->d_sol, nrhs);
        if (s != RC_OK) return s;
        RC_HIP(hipMemsetAsync(ctx->d_refine, 0, sizeof(double) * 2, ctx->stream));
    }
    // refinement with compensated residuals (k = 0 leaves A's null space
    // unconstrained: no Schur system to correct through): optionally of all
    // the solve's columns first (A_hat^-1 C^T exact for the frozen
    // projection), then of the saddle solution [g~; lambda]
    const bool refinable = k > 0 && !ctx->has_isolated;
    const int ncol = refinable ? (o ? o->refine_columns : 0) : 0;
    const int nref = refinable ? (o ? o->saddle_refine : 1) : 0;
    const double ptol = o ? o->saddle_refine_tol : 1e-6;
    const int pmax = o ? o->refine_max_iter : 200;
    s = rc_refine_zero(ctx, ctx->d_refine + 2);
    if (s == RC_OK && ncol > 0)
        s = rc_refine_columns(ctx, k, precision, ncol, ptol, pmax, ctx->d_refine + 2);
    if (s != RC_OK) return s;
    s = rc_launch_schur(ctx, ctx->d_sol, ctx->d_dE, k, ctx->d_gt, ctx->d_dots, ctx->d_bsum, ctx->d_ssol,
                        ctx->d_Minv);
    if (s != RC_OK) return s;
    if (nref > 0) {
        s = rc_saddle_refine(ctx, k, precision, nref, ptol, pmax, ctx->d_refine + 2);
        if (s != RC_OK) return s;
    }
    // the structured residual, a diagnostic the line search does not read,
    // runs beside it on the second stream: outside the saddle pair
    LED_CLOSE(led, id_saddle, ctx->stream);
    if (led) {
        led->calls[LK_SADDLE] += 1;
        led->calls[LK_FACTOR] += 1;
    }
    hipStream_t rst = ctx->stream;
    if (resid_go && resid_done) {
        RC_HIP(hipEventRecord(resid_go, ctx->stream));
        RC_HIP(hipStreamWaitEvent(ctx->stream2, resid_go, 0));
        rst = ctx->stream2;
    }
    s = rc_launch_residual(ctx, rst, ctx->d_gt, ctx->d_dE, k, ctx->d_ssol, ctx->d_Ax, ctx->d_Cg,
                           ctx->d_scal + RC_SCAL_RESID);
    // the refined solution's compensated residual: a diagnostic as well,
    // formed only on request
    if (s == RC_OK && nref > 0 && o && o->report_ir_residual)
        s = rc_saddle_final_residual(ctx, k, rst, ctx->d_refine + 2);
    if (s == RC_OK && rst == ctx->stream2) RC_HIP(hipEventRecord(resid_done, ctx->stream2));
    return s;
}

extern "C" rc_status rc_solve_constrained_gradient(rc_ctx* ctx, const double* xyz,
                                                   const double* dE, const rc_block* blocks,
                                                   int32_t nblocks, const rc_step_opts* o,
                                                   double* gtilde, double* lambda,
                                                   double* residual) {
    if (!ctx || !xyz || !dE || !gtilde) return RC_BAD_ARG;
    rc_step_opts def;
    if (!o) {
        rc_default_step_opts(&def);
        o = &def;
    }
    if (!params_ok(&o->params)) return RC_BAD_ARG;
    DevSet ds;
    rc_status s = make_set(ctx, blocks, nblocks, ds);
    if (s != RC_OK) return s;
    const int Vp = ctx->nblk * RC_NB;
    s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, &o->params);
    if (s != RC_OK) return s;
    RC_HIP(hipMemcpyAsync(ctx->d_work, dE, sizeof(double) * 3 * ctx->V, hipMemcpyHostToDevice,
                          ctx->stream));
    hipLaunchKernelGGL(k_aos_to_blocks, dim3((Vp + 255) / 256), dim3(256), 0, ctx->stream,
                       ctx->d_work, ctx->V, Vp, ctx->d_dE);
    RC_CHECK_LAUNCH();
    s = solve_phase(ctx, &o->params, ds, o->precision, o, nullptr);
    if (s != RC_OK) return s;
    hipLaunchKernelGGL(k_blocks_to_aos, dim3((ctx->V + 255) / 256), dim3(256), 0, ctx->stream,
                       ctx->d_gt, ctx->V, Vp, ctx->d_work);
    RC_CHECK_LAUNCH();
    RC_HIP(hipMemcpyAsync(gtilde, ctx->d_work, sizeof(double) * 3 * ctx->V, hipMemcpyDeviceToHost,
                          ctx->stream));
    std::vector<double> sol(ds.k + 3);
    RC_HIP(hipMemcpyAsync(sol.data(), ctx->d_ssol, sizeof(double) * (ds.k + 3), hipMemcpyDeviceToHost,
                          ctx->stream));
    double resid = 0.0;
    RC_HIP(hipMemcpyAsync(&resid, ctx->d_scal + RC_SCAL_RESID, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    int32_t st[4];
    RC_HIP(hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    if (st[0] & 2) {
        ctx->err = "device dependency wait timed out (triangular solve or factorisation)";
        return RC_HIP_ERROR;
    }
    if (lambda)
        for (int i = 0; i < ds.k; ++i) lambda[i] = sol[i];
    if (residual) *residual = resid;
    if (st[0] & (1 | 4)) {
        ctx->err = "saddle system is singular";
        return RC_SINGULAR;
    }
    return RC_OK;
}

static int trial_count(double tau0, double shrink, double tau_min) {
    int n = 0;
    double t = tau0;
    while (t >= tau_min && n < 4096) {
        ++n;
        t *= shrink;
    }
    return n;
}

// Line search with projectionMode 'reassemble' (lineSearch.ts:263-276): every
// projection correction rebuilds A and C at the current iterate and
// refactors the saddle there (the FP64 factorisation with the right-hand-side
// rows), then a single-workgroup launch applies it and checks the iterate.
// The host reads one flag per correction; the trials otherwise follow the
// frozen form (same candidate, energy and Armijo kernels). The factor storage
// is reused, so the residual of the gradient solve (second stream) must be
// done first.
static rc_status line_search_reassemble(rc_ctx* ctx, const rc_params* p, const DevSet& ds,
                                        const DevPen& dp, int ntr, int max_iter, Ledger* led) {
    const int Vp = ctx->nblk * RC_NB;
    const int k = ds.k;
    const int nrhs = 3 + 3 * k;
    const int nx = (nrhs + RC_NB - 1) / RC_NB;
    double* Frhs = ctx->d_F + (size_t)Vp * Vp;
    const double* Y = ctx->d_sol + 3 * (size_t)Vp;
    RC_HIP(hipStreamWaitEvent(ctx->stream, ctx->res_done, 0));
    int32_t* flag = (int32_t*)ctx->d_host_stage;
    rc_status s = RC_OK;
    for (int trial = 0; trial < ntr; ++trial) {
        int id_proj = -1, id_e = -1;
        LED_OPEN(led, LK_PROJ, ctx->stream, id_proj, trial);
        s = rc_launch_ls_project_re(ctx, 0, ctx->d_ctl, ctx->d_set, dp, p, ctx->d_pos, ctx->d_dir,
                                    ctx->d_work, Y, ctx->d_Minv);
        if (s != RC_OK) return s;
        for (int corr = 0; corr <= max_iter; ++corr) {
            RC_HIP(hipMemcpyAsync(flag, (char*)ctx->d_ctl + offsetof(LSCtl, need_solve),
                                  sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            RC_HIP(hipStreamSynchronize(ctx->stream));
            if (!*flag) break;
            // A(cur), C(cur) and the factored saddle at the iterate
            int id_asm = -1, id_sad = -1, id_fac = -1;
            LED_OPEN(led, LK_ASM, ctx->stream, id_asm, trial);
            s = rc_launch_edge_prep(ctx, ctx->d_work, p);
            if (s == RC_OK) s = rc_launch_gram(ctx, p);
            if (s != RC_OK) return s;
            LED_CLOSE(led, id_asm, ctx->stream);
            if (led) led->calls[LK_ASM] += 1;
            LED_OPEN(led, LK_SADDLE, ctx->stream, id_sad, trial);
            if (k > 0) s = rc_launch_constraint_jacobian(ctx, ctx->d_set, k);
            if (s == RC_OK) s = rc_launch_build_rhs(ctx, ctx->d_dE, k, Frhs, nx * RC_NB);
            if (s != RC_OK) return s;
            LED_OPEN(led, LK_FACTOR, ctx->stream, id_fac, trial);
            s = rc_factor_fp64(ctx, nx);
            if (s != RC_OK) return s;
            LED_CLOSE(led, id_fac, ctx->stream);
            RC_HIP(hipMemcpyAsync(ctx->d_sol, Frhs, sizeof(double) * (size_t)nrhs * Vp,
                                  hipMemcpyDeviceToDevice, ctx->stream));
            s = rc_solve_fp64_backward(ctx, ctx->d_sol, nrhs);
            // g~ of the iterate is not needed: scratch output
            if (s == RC_OK)
                s = rc_launch_schur(ctx, ctx->d_sol, ctx->d_dE, k, ctx->d_tmp, ctx->d_dots,
                                    ctx->d_bsum, ctx->d_ssol, ctx->d_Minv);
            if (s != RC_OK) return s;
            LED_CLOSE(led, id_sad, ctx->stream);
            s = rc_launch_ls_project_re(ctx, 1, ctx->d_ctl, ctx->d_set, dp, p, ctx->d_pos,
                                        ctx->d_dir, ctx->d_work, Y, ctx->d_Minv);
            if (s != RC_OK) return s;
        }
        LED_CLOSE(led, id_proj, ctx->stream);
        LED_OPEN(led, LK_ENERGY, ctx->stream, id_e, trial);
        s = rc_launch_ls_energy(ctx, ctx->d_ctl, p, dp, ctx->d_work, ctx->d_epart_ls);
        if (s != RC_OK) return s;
        LED_CLOSE(led, id_e, ctx->stream);
        RC_HIP(hipMemcpyAsync(flag, (char*)ctx->d_ctl + offsetof(LSCtl, done), sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
        RC_HIP(hipStreamSynchronize(ctx->stream));
        if (*flag) break;
    }
    return RC_OK;
}

// Device-resident step on ctx->d_pos. update: replace d_pos with the accepted
// curve. Stats (and g~ / positions when requested) are the single D2H copy.
static rc_status step_impl(rc_ctx* ctx, const rc_block* blocks, int32_t nblocks,
                           const rc_penalties* pen, const rc_step_opts* o, rc_step_stats* stats,
                           rc_timings* timings, double* gtilde_out, int update,
                           double* xyz_out) {
    if (!ctx || !o || !stats || !params_ok(&o->params)) return RC_BAD_ARG;
    if (!o->projection_frozen && o->precision != RC_PRECISION_FP64) {
        ctx->err = "projectionMode 'reassemble' is provided for the FP64 form only";
        return RC_UNSUPPORTED;
    }
    if (!(o->shrink > 0.0 && o->shrink < 1.0) || !(o->tau0 > 0.0) || o->max_iter < 0) {
        ctx->err = "line-search options out of range";
        return RC_BAD_ARG;
    }
    if ((o->gradient_mode != RC_MODE_ANALYTICAL && o->gradient_mode != RC_MODE_FINITE_DIFF) ||
        (o->gradient_mode == RC_MODE_FINITE_DIFF && !fd_h_ok(o->fd_h))) {
        ctx->err = "gradient mode or finite-difference step out of range";
        return RC_BAD_ARG;
    }
    DevSet ds;
    rc_status s = make_set(ctx, blocks, nblocks, ds);
    if (s != RC_OK) return s;
    const DevPen dp = resolve_pen(pen);
    const rc_params* p = &o->params;
    const int Vp = ctx->nblk * RC_NB;
    Ledger ledger(ctx);
    Ledger* led = (o->collect_timings && timings) ? &ledger : nullptr;
    int id_step = -1, id_asm = -1, id_e0 = -1, id_de = -1, id_ls = -1;
    LED_OPEN(led, LK_STEP, ctx->stream, id_step);

    s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s != RC_OK) return s;
    // The Gram assembly needs only the edge table: it runs on the second
    // stream beside the energy and the adjoint gradient (all earlier users of
    // the factor storage are ordered before asm_go on this stream).
    if (!ctx->asm_go) {
        RC_HIP(hipEventCreateWithFlags(&ctx->asm_go, hipEventDisableTiming | hipEventDisableSystemFence));
        RC_HIP(hipEventCreateWithFlags(&ctx->asm_done, hipEventDisableTiming | hipEventDisableSystemFence));
        RC_HIP(hipEventCreateWithFlags(&ctx->res_go, hipEventDisableTiming | hipEventDisableSystemFence));
        RC_HIP(hipEventCreateWithFlags(&ctx->res_done, hipEventDisableTiming | hipEventDisableSystemFence));
    }
    RC_HIP(hipEventRecord(ctx->asm_go, ctx->stream));
    RC_HIP(hipStreamWaitEvent(ctx->stream2, ctx->asm_go, 0));
    LED_OPEN(led, LK_ASM, ctx->stream2, id_asm);
    s = rc_launch_gram(ctx, p, ctx->stream2);
    if (s != RC_OK) return s;
    LED_CLOSE(led, id_asm, ctx->stream2);
    if (led) led->calls[LK_ASM] += 1;
    RC_HIP(hipEventRecord(ctx->asm_done, ctx->stream2));
    // E0 = objective at gamma0 unless the caller supplies it (optimizer.ts:117-132)
    if (!o->has_energy_before) {
        LED_OPEN(led, LK_ENERGY, ctx->stream, id_e0);
        s = rc_launch_energy(ctx, p, ctx->d_scal + 2);
        if (s == RC_OK) s = rc_launch_penalty_energy(ctx, dp, ctx->d_scal + 3);
        if (s != RC_OK) return s;
        LED_CLOSE(led, id_e0, ctx->stream);
        if (led) led->calls[LK_ENERGY] += 1;
    }
    // dE = gradient (analytic or forward difference) + penalty gradient
    // (optimizer.ts:278-294), stride Vp
    LED_OPEN(led, LK_DE, ctx->stream, id_de);
    s = o->gradient_mode == RC_MODE_FINITE_DIFF ? rc_launch_gradient_fd(ctx, p, o->fd_h, ctx->d_grad)
                                                : rc_launch_gradient(ctx, p, ctx->d_grad);
    if (s != RC_OK) return s;
    RC_HIP(hipMemsetAsync(ctx->d_dE, 0, sizeof(double) * 3 * Vp, ctx->stream));
    s = rc_launch_penalty_grad(ctx, dp, ctx->d_grad, ctx->V, ctx->d_dE, Vp);
    if (s != RC_OK) return s;
    LED_CLOSE(led, id_de, ctx->stream);
    if (led) led->calls[LK_DE] += 1;

    // control block
    LSCtl hc;
    std::memset(&hc, 0, sizeof(hc));
    hc.c1 = o->c1;
    hc.shrink = o->shrink;
    hc.tau_min = o->tau_min;
    hc.tol_abs = o->tol_abs;
    hc.tol_rel = o->tol_rel;
    hc.max_iter = o->max_iter;
    RC_HIP(hipMemcpyAsync(ctx->d_ctl, &hc, sizeof(LSCtl), hipMemcpyHostToDevice, ctx->stream));

    // the residual (a diagnostic the line search does not read) runs on the
    // second stream beside the line search
    s = solve_phase(ctx, p, ds, o->precision, o, led, ctx->asm_done, ctx->res_go, ctx->res_done);
    if (s != RC_OK) return s;
    LED_OPEN(led, LK_LS, ctx->stream, id_ls);
    s = rc_launch_ls_setup(ctx, ctx->d_gt, ctx->d_dE, ctx->d_dir, ctx->d_scal + 2,
                           ctx->d_scal + 3, o->has_energy_before, o->energy_before, o->tau0,
                           ctx->d_ctl);
    if (s != RC_OK) return s;
    // refine diagnostics into the control block
    hipLaunchKernelGGL(k_copy_refine, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctl,
                       (const double*)ctx->d_refine);
    RC_CHECK_LAUNCH();

    const int ntr = trial_count(o->tau0, o->shrink, o->tau_min);
    // trials are enqueued in batches, the host reading the done flag after
    // each: the first trial alone (an accepted full step is the common case,
    // and the trials launched past the decision still cost their launches),
    // then 4, 12 and the rest
    const int batches[4] = {1, 4, 12, 1 << 20};
    int done_trials = o->projection_frozen ? 0 : ntr;
    if (!o->projection_frozen) {
        s = line_search_reassemble(ctx, p, ds, dp, ntr, o->max_iter, led);
        if (s != RC_OK) return s;
    }
    // The step's tail (results, outputs, one staging copy) is enqueued after
    // every batch and read back at the one synchronisation that follows: a
    // search decided in the batch ends there; otherwise the next batch runs
    // and the tail is enqueued again (the finalisation changes nothing until
    // the search has accepted a trial).
    bool waited_res = false;
    auto tail = [&](bool& done) -> rc_status {
        LED_CLOSE(led, id_ls, ctx->stream);
        if (!waited_res) {
            RC_HIP(hipStreamWaitEvent(ctx->stream, ctx->res_done, 0));
            waited_res = true;
        }
        rc_status ts = rc_launch_ls_finalize(ctx, ctx->d_ctl, ctx->d_pos, ctx->d_work, update,
                                             ctx->d_stats);
        if (ts != RC_OK) return ts;
        RC_HIP(hipMemcpyAsync(ctx->d_host_stage, ctx->d_stats, 24 * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
        if (gtilde_out) {
            hipLaunchKernelGGL(k_blocks_to_aos, dim3((ctx->V + 255) / 256), dim3(256), 0,
                               ctx->stream, ctx->d_gt, ctx->V, Vp, ctx->d_tmp);
            RC_CHECK_LAUNCH();
            RC_HIP(hipMemcpyAsync(gtilde_out, ctx->d_tmp, sizeof(double) * 3 * ctx->V,
                                  hipMemcpyDeviceToHost, ctx->stream));
        }
        if (xyz_out) {
            // accepted -> the projected trial curve, otherwise the input echo
            hipLaunchKernelGGL(k_pick_output, dim3((ctx->V + 255) / 256), dim3(256), 0,
                               ctx->stream, (const double*)ctx->d_pos,
                               (const double*)ctx->d_work, (const LSCtl*)ctx->d_ctl, ctx->V,
                               ctx->d_stage);
            RC_CHECK_LAUNCH();
            RC_HIP(hipMemcpyAsync(xyz_out, ctx->d_stage, sizeof(double) * 3 * ctx->V,
                                  hipMemcpyDeviceToHost, ctx->stream));
        }
        LED_CLOSE(led, id_step, ctx->stream);
        RC_HIP(hipStreamSynchronize(ctx->stream));
        done = *(const int32_t*)(ctx->d_host_stage + 14) != 0;
        return RC_OK;
    };
    bool done = false, tail_read = false;
    for (int bi = 0; bi < 4 && done_trials < ntr; ++bi) {
        const int nb = std::min(batches[bi], ntr - done_trials);
        for (int i = 0; i < nb; ++i) {
            int id_proj = -1, id_e = -1;
            LED_OPEN(led, LK_PROJ, ctx->stream, id_proj, done_trials + i);
            s = rc_launch_ls_project(ctx, ds.k, ctx->d_ctl, ctx->d_set, ctx->d_pos, ctx->d_dir,
                                     ctx->d_work, ctx->d_sol + 3 * (size_t)Vp, ctx->d_Minv);
            if (s != RC_OK) return s;
            LED_CLOSE(led, id_proj, ctx->stream);
            LED_OPEN(led, LK_ENERGY, ctx->stream, id_e, done_trials + i);
            s = rc_launch_ls_energy(ctx, ctx->d_ctl, p, dp, ctx->d_work, ctx->d_epart_ls);
            if (s != RC_OK) return s;
            LED_CLOSE(led, id_e, ctx->stream);
        }
        done_trials += nb;
        if ((s = tail(done)) != RC_OK) return s;
        tail_read = true;
        if (done) break;
    }
    if (!tail_read && (s = tail(done)) != RC_OK) return s;
    const int32_t* hst = (const int32_t*)(ctx->d_host_stage + 12);
    if (hst[0] & 2) {
        ctx->err = "device dependency wait timed out (triangular solve or factorisation)";
        return RC_HIP_ERROR;
    }
    const double* d = ctx->d_host_stage;
    stats->energy = d[0];
    stats->energy_before = d[1];
    stats->tau = d[2];
    stats->residual = d[3];
    stats->gradient_l2_norm = d[4];
    stats->slope = d[5];
    stats->phi_norm = d[6];
    stats->refine_residual = d[7];
    const int32_t* is = (const int32_t*)(d + 8);
    stats->accepted = is[0];
    stats->converged = is[1];
    stats->projection_iterations = is[2];
    stats->reason = is[3];
    stats->trials = is[4];
    stats->refine_iters = is[5];
    stats->saddle_ir_residual = d[16];
    stats->saddle_ir_steps = (int32_t)d[17];
    stats->saddle_ir_correction = d[18];
    stats->fp64_fallback = ctx->fp64_fallback;
    if (led) {
        // device-counted calls (trial energies, correction solves); the
        // phases that did not fire report calls = 0 and no time
        led->calls[LK_STEP] += 1;
        led->calls[LK_ENERGY] += is[6];
        led->calls[LK_SADDLE] += is[7];
        if (!o->projection_frozen) led->calls[LK_FACTOR] += is[7];
        led->calls[LK_PROJ] += stats->trials;
        const bool singular = stats->reason == RC_REASON_SINGULAR_SYSTEM;
        if (!singular && !stats->converged) led->calls[LK_LS] += 1;
        if (singular) led->calls[LK_SADDLE] = led->calls[LK_FACTOR] = 0;
        std::memset(timings, 0, sizeof(*timings));
        RC_HIP(led->sum(timings, stats->trials));
        rc_phase* ph[LK_N] = {&timings->dE, &timings->energy, &timings->assembleA,
                              &timings->saddle, &timings->factor, &timings->projection,
                              &timings->lineSearch, &timings->step};
        for (rc_phase* q : ph)
            if (q->calls == 0) q->ms = 0.0;
    }
    return RC_OK;
}

extern "C" rc_status rc_sobolev_step_resident(rc_ctx* ctx, const rc_block* blocks,
                                              int32_t nblocks, const rc_penalties* pen,
                                              const rc_step_opts* o, rc_step_stats* stats,
                                              rc_timings* timings) {
    return step_impl(ctx, blocks, nblocks, pen, o, stats, timings, nullptr, 1, nullptr);
}

extern "C" rc_status rc_sobolev_step(rc_ctx* ctx, const double* xyz_in, const rc_block* blocks,
                                     int32_t nblocks, const rc_penalties* pen,
                                     const rc_step_opts* o, double* xyz_out, rc_step_stats* stats,
                                     double* gtilde_out, rc_timings* timings) {
    if (!ctx || !xyz_in || !xyz_out) return RC_BAD_ARG;
    rc_status s = rc_upload_positions_aos(ctx, xyz_in, ctx->d_pos);
    if (s != RC_OK) return s;
    return step_impl(ctx, blocks, nblocks, pen, o, stats, timings, gtilde_out, 0, xyz_out);
}

// ---------------------------------------------------------------------------
// Kernel-level profile of the trailing update (measurement support).

static rc_status kprof_collect(rc_ctx* ctx) {
    if (ctx->kev_used == 0) return RC_OK;
    RC_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i + 1 < ctx->kev_used; i += 2) {
        float ms = 0.f;
        RC_HIP(hipEventElapsedTime(&ms, ctx->kev[i], ctx->kev[i + 1]));
        ctx->kprof_ms += ms;
        ctx->kprof_flop += ctx->kev_flop[i / 2];
        ++ctx->kprof_launches;
    }
    ctx->kev_used = 0;
    ctx->kev_flop.clear();
    return RC_OK;
}

extern "C" rc_status rc_kernel_profile_enable(rc_ctx* ctx, int32_t enable) {
    if (!ctx) return RC_BAD_ARG;
    rc_status s = kprof_collect(ctx);
    if (s != RC_OK) return s;
    ctx->kprof = enable ? 1 : 0;
    ctx->kprof_ms = 0.0;
    ctx->kprof_flop = 0.0;
    ctx->kprof_launches = 0;
    return RC_OK;
}

extern "C" rc_status rc_kernel_profile_read(rc_ctx* ctx, rc_kernel_profile* out) {
    if (!ctx || !out) return RC_BAD_ARG;
    rc_status s = kprof_collect(ctx);
    if (s != RC_OK) return s;
    out->syrk_ms = ctx->kprof_ms;
    out->syrk_flop = ctx->kprof_flop;
    out->syrk_launches = ctx->kprof_launches;
    out->_pad = 0;
    return RC_OK;
}

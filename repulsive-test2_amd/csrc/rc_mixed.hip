// Mixed-precision solve of A_hat X = B: preconditioned conjugate gradients in
// FP64 (matrix products with the FP64 A, FP64 vectors and scalars) with the
// FP32 Cholesky factor of S = D^-1/2 A_hat D^-1/2 (FP16x3 MFMA trailing
// updates) as preconditioner: M^-1 r = D^-1/2 S~^-1 D^-1/2 r.
//
// Plain iterative refinement x += M^-1 (b - A x) converges only while
// kappa(A_hat) * u_factor < 1; for this fractional operator kappa grows like
// V^2.4 and crosses that bound around V ~ 4096 with an FP32-quality factor,
// so the refinement is run as PCG, which converges for any SPD
// preconditioner and needs only a few iterations when M^-1 A_hat is close to
// the identity. Every right-hand side column has its own CG scalars; a
// column stops once ||r||/||b|| <= tol, and all kernels of an iteration
// return at once when every column has stopped.
#include "rc_internal.h"

namespace {

constexpr int GEMM_CHUNK = 16;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// out[c] = sum_v X[c][v] (v < V); one RED_T-thread workgroup per column
__global__ __launch_bounds__(RED_T) void k_colsum(const double* __restrict__ X, int V, int Vp,
                                                  double* __restrict__ out,
                                                  const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[RED_T / 64];
    const int c = blockIdx.x;
    double s = 0.0;
#pragma unroll 8
    for (int v = threadIdx.x; v < V; v += RED_T) s += X[(size_t)c * Vp + v];
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int q = 0; q < RED_T / 64; ++q) a += sh[q];
        out[c] = a;
    }
}

typedef double d4 __attribute__((ext_vector_type(4)));

// Q = A_hat P for nc <= 16 columns of P (ld Vp): Q[c][u] = sum_j A[u][j] P[c][j]
// + mu colsum[c]. One workgroup per 16 rows u, read along the rows; its
// four waves split j in four and add their partials in LDS in a fixed
// order. v_mfma_f64_16x16x4f64 with A operand P[c = lane % 16][k] and B
// operand A[u0 + lane % 16][k], where lane group g = lane / 16 takes
// k = j + 4 g + m in the m-th of four products: each lane loads 32
// consecutive bytes of its row of A (and of P) per four products, so a wave
// reads 16 rows x 128 contiguous bytes per load, A once per product.
__global__ __launch_bounds__(256) void k_ahat_mfma(const double* __restrict__ A, int V, int Vp,
                                                   const double* __restrict__ mu,
                                                   const double* __restrict__ colsum,
                                                   const double* __restrict__ P, int nc,
                                                   double* __restrict__ Q,
                                                   const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double red[4][16][17];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int u0 = blockIdx.x * 16;
    const int span = ((V + 63) / 64) * 16;      // a quarter of V, a multiple of 16
    const int j0 = w * span, j1 = min(V, j0 + span);
    const bool cval = lr < nc;
    const double* Pc = P + (size_t)(cval ? lr : 0) * Vp;
    const bool uval = u0 + lr < V;
    const double* Au = A + (size_t)(uval ? u0 + lr : 0) * Vp;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    // 16 j per step, four steps' loads issued together (the last group's
    // steps past j1 read the first step's addresses and add zeros)
    const int nit = j1 > j0 ? (j1 - j0 + 15) / 16 : 0;
    for (int it = 0; it < nit; it += 4) {
        double2 a[4][2], p[4][2];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int jj = j0 + 16 * (it + q < nit ? it + q : it) + 4 * lk;   // jj + 3 < Vp
            a[q][0] = *reinterpret_cast<const double2*>(Au + jj);
            a[q][1] = *reinterpret_cast<const double2*>(Au + jj + 2);
            p[q][0] = *reinterpret_cast<const double2*>(Pc + jj);
            p[q][1] = *reinterpret_cast<const double2*>(Pc + jj + 2);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int jj = j0 + 16 * (it + q) + 4 * lk;
            const double av[4] = {a[q][0].x, a[q][0].y, a[q][1].x, a[q][1].y};
            const double pv[4] = {p[q][0].x, p[q][0].y, p[q][1].x, p[q][1].y};
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const bool ok = it + q < nit && jj + m < j1;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64((cval && ok) ? pv[m] : 0.0,
                                                           (uval && ok) ? av[m] : 0.0, acc, 0,
                                                           0, 0);
            }
        }
    }
    // acc[q] = partial Q[c = 4 q + lk][u = u0 + lr]
#pragma unroll
    for (int q = 0; q < 4; ++q) red[w][4 * q + lk][lr] = acc[q];
    __syncthreads();
    const int c = t >> 4, uu = t & 15;
    if (c < nc && u0 + uu < V)
        Q[(size_t)c * Vp + u0 + uu] =
            ((red[0][c][uu] + red[1][c][uu]) + red[2][c][uu]) + red[3][c][uu] + mu[0] * colsum[c];
}

// out[c] = <X[c], Y[c]> over v < V; one RED_T-thread workgroup per column
__global__ __launch_bounds__(RED_T) void k_coldot(const double* __restrict__ X,
                                                  const double* __restrict__ Y, int V, int Vp,
                                                  double* __restrict__ out,
                                                  const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[RED_T / 64];
    const int c = blockIdx.x;
    double s = 0.0;
#pragma unroll 8
    for (int v = threadIdx.x; v < V; v += RED_T)
        s = fma(X[(size_t)c * Vp + v], Y[(size_t)c * Vp + v], s);
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int q = 0; q < RED_T / 64; ++q) a += sh[q];
        out[c] = a;
    }
}

// R = B - Q
__global__ void k_sub(const double* __restrict__ B, const double* __restrict__ Q, int Vp, int nrhs,
                      double* __restrict__ R) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)nrhs * Vp) return;
    R[j] = B[j] - Q[j];
}

// X += alpha P, R -= alpha Q with alpha = rz / pq per active column
__global__ void k_pcg_update(double* __restrict__ X, double* __restrict__ R,
                             const double* __restrict__ P, const double* __restrict__ Q,
                             const double* __restrict__ rz, const double* __restrict__ pq,
                             const int32_t* __restrict__ conv, int V, int Vp,
                             const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (v >= V || conv[c]) return;
    const double al = rz[c] / pq[c];
    const size_t j = (size_t)c * Vp + v;
    X[j] = fma(al, P[j], X[j]);
    R[j] = fma(-al, Q[j], R[j]);
}

// P = Z + (rz_new / rz_old) P per active column
__global__ void k_pcg_dir(const double* __restrict__ Z, double* __restrict__ P,
                          const double* __restrict__ rz_new, const double* __restrict__ rz_old,
                          const int32_t* __restrict__ conv, int V, int Vp, int first,
                          const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (v >= V || conv[c]) return;
    const size_t j = (size_t)c * Vp + v;
    P[j] = first ? Z[j] : fma(rz_new[c] / rz_old[c], P[j], Z[j]);
}

// Convergence bookkeeping: ctl[0] = all converged flag (int), ctl[1] = iteration count,
// res[0] = max relative residual, res[1] = iterations (as double, for the step stats).
__global__ void k_pcg_check(const double* __restrict__ rr, const double* __restrict__ bn,
                            int nrhs, double tol, int32_t* __restrict__ conv,
                            int32_t* __restrict__ ctl, double* __restrict__ res, int count_iter) {
    if (threadIdx.x != 0) return;
    if (ctl[0]) return;
    int all = 1;
    double mx = 0.0;
    for (int c = 0; c < nrhs; ++c) {
        const double r = (bn[c] > 0.0) ? sqrt(rr[c]) / bn[c] : 0.0;
        if (!(r > tol) || r != r) conv[c] = 1;  // NaN stops the column too
        if (!conv[c]) all = 0;
        mx = (r > mx) ? r : mx;
    }
    if (count_iter) ctl[1] += 1;
    ctl[0] = all;
    res[0] = mx;
    res[1] = (double)ctl[1];
}

__global__ void k_sqrt_inplace(double* __restrict__ x, int n) {
    const int i = threadIdx.x;
    if (i < n) x[i] = sqrt(x[i]);
}

}  // namespace

struct PcgBufs {
    double *R, *Z, *P, *Q;
    double *bn, *rr, *rzA, *rzB, *pq, *cs;
    int32_t *conv, *ctl;
    int cap;
};

static rc_status pcg_bufs(rc_ctx* ctx, PcgBufs& b) {
    const int Vp = ctx->nblk * RC_NB;
    const int cap = ctx->nrhs_cap;
    if (!ctx->d_tmp2) {
        RC_HIP(hipMalloc((void**)&ctx->d_tmp2, sizeof(double) * (size_t)4 * cap * Vp));
        RC_HIP(hipMalloc((void**)&ctx->d_tmp3, sizeof(double) * (size_t)8 * cap + 64));
    }
    b.R = ctx->d_tmp2;
    b.Z = b.R + (size_t)cap * Vp;
    b.P = b.Z + (size_t)cap * Vp;
    b.Q = b.P + (size_t)cap * Vp;
    b.bn = ctx->d_tmp3;
    b.rr = b.bn + cap;
    b.rzA = b.rr + cap;
    b.rzB = b.rzA + cap;
    b.pq = b.rzB + cap;
    b.cs = b.pq + cap;
    b.conv = (int32_t*)(b.cs + cap);
    b.ctl = b.conv + cap + 2;
    b.cap = cap;
    return RC_OK;
}

// W = M^-1 R for nrhs columns (scale, FP32-factor solve, scale); R may be W.
static rc_status precond(rc_ctx* ctx, const double* R, double* W, int nrhs,
                         const int32_t* skip) {
    return rc_precond_f32(ctx, R, W, nrhs, skip);
}

static rc_status ahat_mult(rc_ctx* ctx, const double* X, double* Q, double* cs, int nrhs,
                           const int32_t* skip) {
    const int Vp = ctx->nblk * RC_NB;
    hipLaunchKernelGGL(k_colsum, dim3(nrhs), dim3(RED_T), 0, ctx->stream, X, ctx->V, Vp, cs, skip);
    RC_CHECK_LAUNCH();
    for (int c0 = 0; c0 < nrhs; c0 += GEMM_CHUNK) {
        const int nc = (nrhs - c0) < GEMM_CHUNK ? (nrhs - c0) : GEMM_CHUNK;
        hipLaunchKernelGGL(k_ahat_mfma, dim3((ctx->V + 15) / 16), dim3(256), 0, ctx->stream,
                           (const double*)ctx->d_F, ctx->V, Vp, (const double*)(ctx->d_scal + 1),
                           (const double*)(cs + c0), (const double*)(X + (size_t)c0 * Vp), nc,
                           Q + (size_t)c0 * Vp, skip);
        RC_CHECK_LAUNCH();
    }
    return RC_OK;
}

// Solve A_hat X = B (nrhs columns, ld Vp). d_refine_out[0] = final max
// relative residual, [1] = PCG iterations.
rc_status rc_solve_mixed(rc_ctx* ctx, const double* d_B, double* d_X, int nrhs, double tol,
                         int max_iter, double* d_refine_out) {
    const int Vp = ctx->nblk * RC_NB;
    PcgBufs b;
    rc_status s = pcg_bufs(ctx, b);
    if (s != RC_OK) return s;
    if (nrhs > b.cap) {
        ctx->err = "too many right-hand sides";
        return RC_BAD_ARG;
    }
    const dim3 gv((ctx->V + 255) / 256, nrhs);
    RC_HIP(hipMemsetAsync(b.conv, 0, sizeof(int32_t) * (b.cap + 4), ctx->stream));
    RC_HIP(hipMemsetAsync(d_refine_out, 0, sizeof(double) * 2, ctx->stream));
    // X0 = M^-1 B
    if ((s = precond(ctx, d_B, d_X, nrhs, nullptr)) != RC_OK) return s;
    // R = B - A_hat X0, ||B||
    if ((s = ahat_mult(ctx, d_X, b.Q, b.cs, nrhs, nullptr)) != RC_OK) return s;
    hipLaunchKernelGGL(k_sub, dim3((unsigned)(((size_t)nrhs * Vp + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_B, b.Q, Vp, nrhs, b.R);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_coldot, dim3(nrhs), dim3(RED_T), 0, ctx->stream, d_B, d_B, ctx->V, Vp, b.bn,
                       (const int32_t*)nullptr);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sqrt_inplace, dim3(1), dim3(256), 0, ctx->stream, b.bn, nrhs);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_coldot, dim3(nrhs), dim3(RED_T), 0, ctx->stream, b.R, b.R, ctx->V, Vp, b.rr,
                       (const int32_t*)nullptr);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(64), 0, ctx->stream, b.rr, b.bn, nrhs, tol,
                       b.conv, b.ctl, d_refine_out, 0);
    RC_CHECK_LAUNCH();
    const int32_t* skip = b.ctl;
    // Z = M^-1 R, rz = <R, Z>, P = Z
    if ((s = precond(ctx, b.R, b.Z, nrhs, skip)) != RC_OK) return s;
    hipLaunchKernelGGL(k_coldot, dim3(nrhs), dim3(RED_T), 0, ctx->stream, b.R, b.Z, ctx->V, Vp, b.rzA,
                       skip);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pcg_dir, gv, dim3(256), 0, ctx->stream, b.Z, b.P, b.rzA, b.rzA, b.conv,
                       ctx->V, Vp, 1, skip);
    RC_CHECK_LAUNCH();
    double* rz_old = b.rzA;
    double* rz_new = b.rzB;
    const int batch = 6;
    for (int it = 0; it < max_iter; ++it) {
        if ((s = ahat_mult(ctx, b.P, b.Q, b.cs, nrhs, skip)) != RC_OK) return s;
        hipLaunchKernelGGL(k_coldot, dim3(nrhs), dim3(RED_T), 0, ctx->stream, b.P, b.Q, ctx->V, Vp,
                           b.pq, skip);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_pcg_update, gv, dim3(256), 0, ctx->stream, d_X, b.R, b.P, b.Q, rz_old,
                           b.pq, b.conv, ctx->V, Vp, skip);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_coldot, dim3(nrhs), dim3(RED_T), 0, ctx->stream, b.R, b.R, ctx->V, Vp,
                           b.rr, skip);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(64), 0, ctx->stream, b.rr, b.bn, nrhs, tol,
                           b.conv, b.ctl, d_refine_out, 1);
        RC_CHECK_LAUNCH();
        if (it + 1 == max_iter) break;
        if ((s = precond(ctx, b.R, b.Z, nrhs, skip)) != RC_OK) return s;
        hipLaunchKernelGGL(k_coldot, dim3(nrhs), dim3(RED_T), 0, ctx->stream, b.R, b.Z, ctx->V, Vp,
                           rz_new, skip);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_pcg_dir, gv, dim3(256), 0, ctx->stream, b.Z, b.P, rz_new, rz_old,
                           b.conv, ctx->V, Vp, 0, skip);
        RC_CHECK_LAUNCH();
        double* tswap = rz_old;
        rz_old = rz_new;
        rz_new = tswap;
        if ((it + 1) % batch == 0) {
            RC_HIP(hipMemcpyAsync(ctx->d_host_stage + 32, b.ctl, sizeof(int32_t) * 2,
                                  hipMemcpyDeviceToHost, ctx->stream));
            RC_HIP(hipStreamSynchronize(ctx->stream));
            if (((int32_t*)(ctx->d_host_stage + 32))[0]) break;
        }
    }
    return RC_OK;
}

// Multigrid projected solver of the constrained saddle system: the stage-2
// oracle's mg_projected_solve (oracle/tpe_stage2_oracle.py:957-1155 of the
// reference) on the device. The caller builds the level hierarchy (curve
// coarsening, per-level constraint Jacobians C_l and (C_l C_l^T)^+) on the
// host, as the reference does in Python; this file runs the solve:
//   - level 0's operator is the Gram matrix A of the context's curve (the
//     device assembly), copied to a dense n x n buffer; coarse operators are
//     the Galerkin products A_l = P_l^T A_{l-1} P_l (prepare_levels,
//     :969-999; P3 = I3 (x) P, so the scalar P suffices), or operators the
//     caller supplies (the reference's 'rediscretize' A/B option);
//   - smoothing: projected CG on Pi A Pi, Pi = I - C^T (C C^T)^+ C, a fixed
//     number of steps with the reference's two early exits
//     (smooth_projected_cg, :1013-1035);
//   - V-cycles (v_cycle, :1038-1058): pre-smooth, projected residual,
//     restriction Pi_c P^T r, recursion, prolongation Pi_f (x + P e), post-
//     smooth; the coarsest level (or one of <= 8 vertices) is solved directly
//     (the saddle matrix [[A3, C^T], [C, 0]] by LU with partial pivoting on
//     the host, a few dozen unknowns);
//   - nested-iteration initial guess (:1084-1106), the divergence guard
//     (residual growth > 10x the best so far, or not finite) and the dense
//     cleanup with its certificate (:1109-1155); the cleanup's dense solve is
//     the library's saddle solve of level 0.
// Vectors are coordinate blocks [x_0..x_{n-1}, y.., z..] of one level. The
// dense product is a grid kernel (one wave per row, A read once); every
// other operation of the iteration is a single-workgroup kernel over a
// level's vectors (3n doubles) with the CG scalars and exit flags in device
// memory, so the host reads back only the residual norm once per cycle and
// the coarsest level's right-hand side.
#include "rc_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int WG = 1024;
constexpr int MG_KMAX = 256;   // constraint rows per level handled by the projection

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// sum over the workgroup in a fixed order (sh: 32 doubles); every thread
// gets the result
__device__ double block_sum(double v, double* sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wsum(v);
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += sh[i];
    __syncthreads();
    return s;
}

struct Lev {
    int n, k;
    const double* C;    // k x 3n
    const double* Gp;   // k x k, (C C^T)^+
};

// v <- Pi v = v - C^T (Gp (C v)). t: MG_KMAX doubles of shared scratch.
__device__ void project(const Lev& L, double* v, double* t, double* u) {
    const int m = 3 * L.n, lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (L.k == 0) return;
    __syncthreads();
    for (int r = w; r < L.k; r += nw) {
        const double* c = L.C + (size_t)r * m;
        double s = 0.0;
        for (int i = lane; i < m; i += 64) s += c[i] * v[i];
        s = wsum(s);
        if (lane == 0) t[r] = s;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < L.k; r += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < L.k; ++q) s += L.Gp[(size_t)r * L.k + q] * t[q];
        u[r] = s;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        double s = 0.0;
        for (int r = 0; r < L.k; ++r) s += L.C[(size_t)r * m + i] * u[r];
        v[i] -= s;
    }
    __syncthreads();
}

__device__ double dot(const double* a, const double* b, int m, double* sh) {
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) s += a[i] * b[i];
    return block_sum(s, sh);
}

// y_c = A x_c for the three coordinate blocks; one wave per row. skip:
// nullable flag (a finished smoother).
__global__ __launch_bounds__(256) void k_mg_matvec(const double* __restrict__ A, int lda, int n,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ y,
                                                   const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    const double* a = A + (size_t)row * lda;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int j = lane; j < n; j += 64) {
        const double v = a[j];
        s0 += v * x[j];
        s1 += v * x[n + j];
        s2 += v * x[2 * n + j];
    }
    s0 = wsum(s0);
    s1 = wsum(s1);
    s2 = wsum(s2);
    if (lane == 0) {
        y[row] = s0;
        y[n + row] = s1;
        y[2 * n + row] = s2;
    }
}

// Scalars of one smoother run (per level): [0] rs, [1] done, [2] zero start
struct CgState {
    double rs;
    int32_t done;
    int32_t zero;
};

// x <- Pi x (the smoother's start, before the residual's product)
__global__ __launch_bounds__(WG) void k_mg_project(Lev L, double* x) {
    __shared__ double t[MG_KMAX], u[MG_KMAX];
    project(L, x, t, u);
}

// r = Pi (rhs - Ax); p = r; rs = r.r; a zero residual ends the smoother with
// x as is (smooth_projected_cg's early return)
__global__ __launch_bounds__(WG) void k_mg_cg_begin(Lev L, const double* rhs, const double* Ax,
                                                    double* r, double* p, CgState* st) {
    __shared__ double t[MG_KMAX], u[MG_KMAX], sh[32];
    const int m = 3 * L.n;
    for (int i = threadIdx.x; i < m; i += blockDim.x) r[i] = rhs[i] - Ax[i];
    project(L, r, t, u);
    for (int i = threadIdx.x; i < m; i += blockDim.x) p[i] = r[i];
    const double rs = dot(r, r, m, sh);
    if (threadIdx.x == 0) {
        st->rs = rs;
        st->zero = rs == 0.0 ? 1 : 0;
        st->done = st->zero;
    }
}

// tp = Pi p (the step's product operand)
__global__ __launch_bounds__(WG) void k_mg_cg_pre(Lev L, const double* p, double* tp,
                                                  const CgState* st) {
    __shared__ double t[MG_KMAX], u[MG_KMAX];
    if (st->done) return;
    const int m = 3 * L.n;
    for (int i = threadIdx.x; i < m; i += blockDim.x) tp[i] = p[i];
    project(L, tp, t, u);
}

// one CG step after Ap = A (Pi p): Ap = Pi Ap; alpha = rs / p.Ap; x = Pi (x +
// alpha p); r -= alpha Ap; rs' = r.r; p = r + (rs'/rs) p, with the reference's
// exits (|p.Ap| < 1e-300 before the update, rs' <= 1e-30 after it)
__global__ __launch_bounds__(WG) void k_mg_cg_step(Lev L, double* x, double* r, double* p,
                                                   double* Ap, CgState* st) {
    __shared__ double t[MG_KMAX], u[MG_KMAX], sh[32];
    __shared__ int stop;
    if (st->done) return;
    const int m = 3 * L.n;
    project(L, Ap, t, u);
    const double denom = dot(p, Ap, m, sh);
    if (threadIdx.x == 0) stop = fabs(denom) < 1e-300;
    __syncthreads();
    if (stop) {
        if (threadIdx.x == 0) st->done = 1;
        return;
    }
    const double a = st->rs / denom;
    for (int i = threadIdx.x; i < m; i += blockDim.x) x[i] = x[i] + a * p[i];
    project(L, x, t, u);
    for (int i = threadIdx.x; i < m; i += blockDim.x) r[i] = r[i] - a * Ap[i];
    const double rsn = dot(r, r, m, sh);
    if (rsn <= 1e-30) {
        if (threadIdx.x == 0) st->done = 1;
        return;
    }
    const double b = rsn / st->rs;
    for (int i = threadIdx.x; i < m; i += blockDim.x) p[i] = r[i] + b * p[i];
    __syncthreads();
    if (threadIdx.x == 0) st->rs = rsn;
}

// the smoother's result Pi x (not after a zero starting residual)
__global__ __launch_bounds__(WG) void k_mg_cg_end(Lev L, double* x, const CgState* st) {
    __shared__ double t[MG_KMAX], u[MG_KMAX];
    if (st->zero) return;
    project(L, x, t, u);
}

// r = Pi (rhs - Ax); out[0] = |r| / max(1, |Pi rhs|) (projected_residual,
// :1006-1010); w: scratch vector
__global__ __launch_bounds__(WG) void k_mg_residual(Lev L, const double* rhs, const double* Ax,
                                                    double* r, double* w, double* out) {
    __shared__ double t[MG_KMAX], u[MG_KMAX], sh[32];
    const int m = 3 * L.n;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        r[i] = rhs[i] - Ax[i];
        w[i] = rhs[i];
    }
    project(L, r, t, u);
    project(L, w, t, u);
    const double rn = sqrt(dot(r, r, m, sh));
    const double bn = sqrt(dot(w, w, m, sh));
    if (threadIdx.x == 0) out[0] = rn / fmax(1.0, bn);
}

// coarse rhs = Pi_c (P^T r): coarse vertex j sums its children (fine vertex,
// weight) in ascending fine order
__global__ __launch_bounds__(WG) void k_mg_restrict(Lev Lc, int nf, const int32_t* ch_off,
                                                    const int32_t* ch_idx, const double* ch_w,
                                                    const double* rf, double* rc) {
    __shared__ double t[MG_KMAX], u[MG_KMAX];
    const int nc = Lc.n;
    for (int q = threadIdx.x; q < 3 * nc; q += blockDim.x) {
        const int c = q / nc, j = q % nc;
        double s = 0.0;
        for (int e = ch_off[j]; e < ch_off[j + 1]; ++e) s += ch_w[e] * rf[(size_t)c * nf + ch_idx[e]];
        rc[q] = s;
    }
    project(Lc, rc, t, u);
}

// x_f = Pi_f (x_f + P e_c): fine vertex a takes its (up to two) parents
__global__ __launch_bounds__(WG) void k_mg_prolong(Lev Lf, int nc, const int32_t* par,
                                                   const double* pw, const double* ec, double* xf,
                                                   int add) {
    __shared__ double t[MG_KMAX], u[MG_KMAX];
    const int nf = Lf.n;
    for (int q = threadIdx.x; q < 3 * nf; q += blockDim.x) {
        const int c = q / nf, a = q % nf;
        double s = 0.0;
        for (int h = 0; h < 2; ++h) {
            const int pj = par[2 * a + h];
            if (pj >= 0) s += pw[2 * a + h] * ec[(size_t)c * nc + pj];
        }
        xf[q] = add ? xf[q] + s : s;
    }
    project(Lf, xf, t, u);
}

// Galerkin product, first half: T = A_f P (nf x nc), T[a][j] = sum over the
// children (b, w) of coarse j of w A_f[a][b]
__global__ void k_mg_gal_right(const double* __restrict__ Af, int ldf, int nf, int nc,
                               const int32_t* __restrict__ ch_off,
                               const int32_t* __restrict__ ch_idx,
                               const double* __restrict__ ch_w, double* __restrict__ T) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (j >= nc) return;
    const double* row = Af + (size_t)a * ldf;
    double s = 0.0;
    for (int e = ch_off[j]; e < ch_off[j + 1]; ++e) s += ch_w[e] * row[ch_idx[e]];
    T[(size_t)a * nc + j] = s;
}

// second half: A_c[i][j] = sum over the children (a, w) of coarse i of
// w T[a][j]
__global__ void k_mg_gal_left(const double* __restrict__ T, int nc,
                              const int32_t* __restrict__ ch_off,
                              const int32_t* __restrict__ ch_idx,
                              const double* __restrict__ ch_w, double* __restrict__ Ac) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= nc) return;
    double s = 0.0;
    for (int e = ch_off[i]; e < ch_off[i + 1]; ++e) s += ch_w[e] * T[(size_t)ch_idx[e] * nc + j];
    Ac[(size_t)i * nc + j] = s;
}

// copy of the assembled Gram matrix (ld Vp) into a dense n x n buffer
__global__ void k_mg_copy_a(const double* __restrict__ F, int ldF, int n, double* __restrict__ A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j < n) A[(size_t)i * n + j] = F[(size_t)i * ldF + j];
}

// x = Pi (x + dx) (the dense cleanup's update)
__global__ __launch_bounds__(WG) void k_mg_cleanup_add(Lev L, double* x, const double* dx) {
    __shared__ double t[MG_KMAX], u[MG_KMAX];
    for (int i = threadIdx.x; i < 3 * L.n; i += blockDim.x) x[i] = x[i] + dx[i];
    project(L, x, t, u);
}

__global__ void k_mg_fill(double* v, size_t m, double x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) v[i] = x;
}

// Dense LU with partial pivoting (largest magnitude, lowest row on ties) and
// the solve, for the coarsest level's saddle matrix.
bool lu_solve(std::vector<double>& K, int n, std::vector<double>& b) {
    std::vector<int> piv(n);
    for (int c = 0; c < n; ++c) {
        int p = c;
        double best = std::fabs(K[(size_t)c * n + c]);
        for (int r = c + 1; r < n; ++r) {
            const double v = std::fabs(K[(size_t)r * n + c]);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        if (!(best > 0.0)) return false;
        piv[c] = p;
        if (p != c) {
            for (int j = 0; j < n; ++j) std::swap(K[(size_t)c * n + j], K[(size_t)p * n + j]);
            std::swap(b[c], b[p]);
        }
        const double d = K[(size_t)c * n + c];
        for (int r = c + 1; r < n; ++r) {
            const double f = K[(size_t)r * n + c] / d;
            K[(size_t)r * n + c] = f;
            if (f != 0.0)
                for (int j = c + 1; j < n; ++j) K[(size_t)r * n + j] -= f * K[(size_t)c * n + j];
            b[r] -= f * b[c];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = b[r];
        for (int j = r + 1; j < n; ++j) s -= K[(size_t)r * n + j] * b[j];
        b[r] = s / K[(size_t)r * n + r];
    }
    return true;
}

struct DevLevel {
    int n = 0, k = 0;
    double* A = nullptr;     // n x n
    double* C = nullptr;     // k x 3n
    double* Gp = nullptr;    // k x k
    // prolongation to the finer level (levels >= 1)
    int32_t* par = nullptr;  // 2 per finer vertex
    double* pw = nullptr;
    int32_t* ch_off = nullptr, *ch_idx = nullptr;
    double* ch_w = nullptr;
    // vectors (3n each)
    double *x = nullptr, *rhs = nullptr, *r = nullptr, *p = nullptr, *t = nullptr, *Ap = nullptr;
    CgState* st = nullptr;
    std::vector<double> hC, hA;   // host copies for the direct solve
    Lev lev() const { return Lev{n, k, C, Gp}; }
};

struct Mg {
    rc_ctx* ctx;
    std::vector<DevLevel> L;
    std::vector<void*> allocs;
    double* d_scal = nullptr;   // [0] residual norm
    double* d_T = nullptr;      // Galerkin scratch
    double* d_w = nullptr;      // residual scratch (3 n0)
    hipStream_t s;
    ~Mg() {
        for (void* p : allocs) (void)hipFree(p);
    }
    template <class T>
    rc_status alloc(T** p, size_t count) {
        if (count == 0) count = 1;
        if (hipMalloc((void**)p, sizeof(T) * count) != hipSuccess) {
            ctx->err = "hipMalloc failed (multigrid buffers)";
            return RC_HIP_ERROR;
        }
        allocs.push_back((void*)*p);
        return RC_OK;
    }
};

#define MG_CHECK(call)                                                    \
    do {                                                                  \
        hipError_t _e = (call);                                           \
        if (_e != hipSuccess) {                                           \
            mg.ctx->err = std::string(#call) + ": " + hipGetErrorString(_e); \
            return RC_HIP_ERROR;                                          \
        }                                                                 \
    } while (0)
#define MG_LAUNCH()                                                       \
    do {                                                                  \
        hipError_t _e = hipGetLastError();                                \
        if (_e != hipSuccess) {                                           \
            mg.ctx->err = std::string("kernel launch: ") + hipGetErrorString(_e); \
            return RC_HIP_ERROR;                                          \
        }                                                                 \
    } while (0)

rc_status matvec(Mg& mg, const DevLevel& l, const double* x, double* y, const int32_t* skip) {
    hipLaunchKernelGGL(k_mg_matvec, dim3((l.n + 3) / 4), dim3(256), 0, mg.s, (const double*)l.A,
                       l.n, l.n, x, y, skip);
    MG_LAUNCH();
    return RC_OK;
}

// smooth_projected_cg on level li, x in place
rc_status smooth(Mg& mg, int li, double* x, const double* rhs, int steps) {
    DevLevel& l = mg.L[li];
    const Lev lv = l.lev();
    hipLaunchKernelGGL(k_mg_project, dim3(1), dim3(WG), 0, mg.s, lv, x);
    MG_LAUNCH();
    rc_status s = matvec(mg, l, x, l.Ap, nullptr);
    if (s != RC_OK) return s;
    hipLaunchKernelGGL(k_mg_cg_begin, dim3(1), dim3(WG), 0, mg.s, lv, rhs, (const double*)l.Ap,
                       l.r, l.p, l.st);
    MG_LAUNCH();
    const int32_t* done = &l.st->done;
    for (int it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(k_mg_cg_pre, dim3(1), dim3(WG), 0, mg.s, lv, (const double*)l.p, l.t,
                           (const CgState*)l.st);
        MG_LAUNCH();
        s = matvec(mg, l, l.t, l.Ap, done);
        if (s != RC_OK) return s;
        hipLaunchKernelGGL(k_mg_cg_step, dim3(1), dim3(WG), 0, mg.s, lv, x, l.r, l.p, l.Ap, l.st);
        MG_LAUNCH();
    }
    hipLaunchKernelGGL(k_mg_cg_end, dim3(1), dim3(WG), 0, mg.s, lv, x, (const CgState*)l.st);
    MG_LAUNCH();
    return RC_OK;
}

// direct_projected_solve on level li: x = saddle solution for [rhs; 0]
rc_status direct(Mg& mg, int li, const double* rhs, double* x) {
    DevLevel& l = mg.L[li];
    const int m = 3 * l.n, N = m + l.k;
    std::vector<double> b(N, 0.0);
    MG_CHECK(hipMemcpyAsync(b.data(), rhs, sizeof(double) * m, hipMemcpyDeviceToHost, mg.s));
    MG_CHECK(hipStreamSynchronize(mg.s));
    std::vector<double> K((size_t)N * N, 0.0);
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < l.n; ++i)
            for (int j = 0; j < l.n; ++j)
                K[(size_t)(c * l.n + i) * N + c * l.n + j] = l.hA[(size_t)i * l.n + j];
    for (int r = 0; r < l.k; ++r)
        for (int i = 0; i < m; ++i) {
            K[(size_t)(m + r) * N + i] = l.hC[(size_t)r * m + i];
            K[(size_t)i * N + m + r] = l.hC[(size_t)r * m + i];
        }
    if (!lu_solve(K, N, b)) {
        mg.ctx->err = "multigrid: the coarsest saddle system is singular";
        return RC_SINGULAR;
    }
    MG_CHECK(hipMemcpyAsync(x, b.data(), sizeof(double) * m, hipMemcpyHostToDevice, mg.s));
    MG_CHECK(hipStreamSynchronize(mg.s));
    return RC_OK;
}

rc_status vcycle(Mg& mg, int li, double* x, const double* rhs, int nu_pre, int nu_post) {
    DevLevel& l = mg.L[li];
    const int last = (int)mg.L.size() - 1;
    if (li == last || l.n <= 8) return direct(mg, li, rhs, x);
    rc_status s = smooth(mg, li, x, rhs, nu_pre);
    if (s != RC_OK) return s;
    // r = Pi (rhs - A x)
    s = matvec(mg, l, x, l.Ap, nullptr);
    if (s != RC_OK) return s;
    hipLaunchKernelGGL(k_mg_residual, dim3(1), dim3(WG), 0, mg.s, l.lev(), rhs, (const double*)l.Ap,
                       l.r, mg.d_w, mg.d_scal + 1);
    MG_LAUNCH();
    DevLevel& c = mg.L[li + 1];
    hipLaunchKernelGGL(k_mg_restrict, dim3(1), dim3(WG), 0, mg.s, c.lev(), l.n,
                       (const int32_t*)c.ch_off, (const int32_t*)c.ch_idx, (const double*)c.ch_w,
                       (const double*)l.r, c.rhs);
    MG_LAUNCH();
    hipLaunchKernelGGL(k_mg_fill, dim3((3 * c.n + 255) / 256), dim3(256), 0, mg.s, c.x,
                       (size_t)3 * c.n, 0.0);
    MG_LAUNCH();
    s = vcycle(mg, li + 1, c.x, c.rhs, nu_pre, nu_post);
    if (s != RC_OK) return s;
    hipLaunchKernelGGL(k_mg_prolong, dim3(1), dim3(WG), 0, mg.s, l.lev(), c.n,
                       (const int32_t*)c.par, (const double*)c.pw, (const double*)c.x, x, 1);
    MG_LAUNCH();
    return smooth(mg, li, x, rhs, nu_post);
}

// projected residual of level 0 -> host
rc_status residual0(Mg& mg, const double* x, const double* rhs, double& rn) {
    DevLevel& l = mg.L[0];
    rc_status s = matvec(mg, l, x, l.Ap, nullptr);
    if (s != RC_OK) return s;
    hipLaunchKernelGGL(k_mg_residual, dim3(1), dim3(WG), 0, mg.s, l.lev(), rhs, (const double*)l.Ap,
                       l.r, mg.d_w, mg.d_scal);
    MG_LAUNCH();
    MG_CHECK(hipMemcpyAsync(&rn, mg.d_scal, sizeof(double), hipMemcpyDeviceToHost, mg.s));
    MG_CHECK(hipStreamSynchronize(mg.s));
    return RC_OK;
}

}  // namespace

rc_status rc_launch_gram(rc_ctx* ctx, const rc_params* p, hipStream_t st);
rc_status rc_launch_edge_prep(rc_ctx* ctx, const double* pos, const rc_params* p);
rc_status rc_upload_positions_aos(rc_ctx* ctx, const double* xyz, double* d_soa);

extern "C" rc_status rc_mg_solve(rc_ctx* ctx, const double* xyz, const rc_params* p,
                                 const rc_block* blocks, int32_t nblocks,
                                 const rc_mg_level* levels, int32_t nlevels, const double* rhs,
                                 const rc_mg_opts* o, double* x_out, rc_mg_stats* stats) {
    if (!ctx || !xyz || !p || !levels || nlevels < 1 || nlevels > RC_MG_MAXLEVELS || !rhs ||
        !o || !x_out || !stats)
        return RC_BAD_ARG;
    if (levels[0].n != ctx->V || !(o->tol > 0.0) || o->max_cycles < 0 ||
        o->max_cycles + 2 > RC_MG_MAXHIST || o->nu_pre < 0 || o->nu_post < 0) {
        ctx->err = "multigrid: level 0 must be the context's curve; options out of range";
        return RC_BAD_ARG;
    }
    std::memset(stats, 0, sizeof(*stats));
    Mg mg;
    mg.ctx = ctx;
    mg.s = ctx->stream;
    mg.L.resize(nlevels);
    // ---- levels: sizes, constraints, prolongations
    for (int li = 0; li < nlevels; ++li) {
        const rc_mg_level& in = levels[li];
        DevLevel& l = mg.L[li];
        if (in.n < 1 || in.k < 0 || in.k > MG_KMAX || (in.k > 0 && (!in.C || !in.Gp)) ||
            (li > 0 && (!in.parent || !in.weight))) {
            ctx->err = "multigrid: level " + std::to_string(li) + " malformed";
            return RC_BAD_ARG;
        }
        l.n = in.n;
        l.k = in.k;
        const int m = 3 * l.n;
        rc_status s = mg.alloc(&l.A, (size_t)l.n * l.n);
        if (s == RC_OK) s = mg.alloc(&l.C, (size_t)l.k * m);
        if (s == RC_OK) s = mg.alloc(&l.Gp, (size_t)l.k * l.k);
        for (double** v : {&l.x, &l.rhs, &l.r, &l.p, &l.t, &l.Ap})
            if (s == RC_OK) s = mg.alloc(v, (size_t)m);
        if (s == RC_OK) s = mg.alloc(&l.st, 1);
        if (s != RC_OK) return s;
        if (l.k > 0) {
            MG_CHECK(hipMemcpyAsync(l.C, in.C, sizeof(double) * l.k * m, hipMemcpyHostToDevice, mg.s));
            MG_CHECK(hipMemcpyAsync(l.Gp, in.Gp, sizeof(double) * l.k * l.k, hipMemcpyHostToDevice, mg.s));
            l.hC.assign(in.C, in.C + (size_t)l.k * m);
        }
        if (li > 0) {
            const int nf = mg.L[li - 1].n;
            std::vector<std::vector<std::pair<int, double>>> ch(l.n);
            for (int a = 0; a < nf; ++a)
                for (int h = 0; h < 2; ++h) {
                    const int pj = in.parent[2 * a + h];
                    if (pj < -1 || pj >= l.n) {
                        ctx->err = "multigrid: prolongation index out of range";
                        return RC_BAD_ARG;
                    }
                    if (pj >= 0) ch[pj].push_back({a, in.weight[2 * a + h]});
                }
            std::vector<int32_t> off(l.n + 1, 0), idx;
            std::vector<double> wt;
            for (int j = 0; j < l.n; ++j) {
                for (auto& e : ch[j]) {
                    idx.push_back(e.first);
                    wt.push_back(e.second);
                }
                off[j + 1] = (int32_t)idx.size();
            }
            s = mg.alloc(&l.par, (size_t)2 * nf);
            if (s == RC_OK) s = mg.alloc(&l.pw, (size_t)2 * nf);
            if (s == RC_OK) s = mg.alloc(&l.ch_off, (size_t)l.n + 1);
            if (s == RC_OK) s = mg.alloc(&l.ch_idx, idx.size());
            if (s == RC_OK) s = mg.alloc(&l.ch_w, wt.size());
            if (s != RC_OK) return s;
            MG_CHECK(hipMemcpyAsync(l.par, in.parent, sizeof(int32_t) * 2 * nf, hipMemcpyHostToDevice, mg.s));
            MG_CHECK(hipMemcpyAsync(l.pw, in.weight, sizeof(double) * 2 * nf, hipMemcpyHostToDevice, mg.s));
            MG_CHECK(hipMemcpyAsync(l.ch_off, off.data(), sizeof(int32_t) * (l.n + 1), hipMemcpyHostToDevice, mg.s));
            if (!idx.empty()) {
                MG_CHECK(hipMemcpyAsync(l.ch_idx, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice, mg.s));
                MG_CHECK(hipMemcpyAsync(l.ch_w, wt.data(), sizeof(double) * wt.size(), hipMemcpyHostToDevice, mg.s));
            }
            MG_CHECK(hipStreamSynchronize(mg.s));   // the host vectors above go out of scope
        }
    }
    rc_status s = mg.alloc(&mg.d_scal, 4);
    if (s == RC_OK) s = mg.alloc(&mg.d_w, (size_t)3 * mg.L[0].n);
    size_t tmax = 0;   // the largest Galerkin intermediate A_f P (n_f x n_c)
    for (int li = 1; li < nlevels; ++li)
        tmax = std::max(tmax, (size_t)mg.L[li - 1].n * mg.L[li].n);
    if (s == RC_OK && tmax > 0) s = mg.alloc(&mg.d_T, tmax);
    if (s != RC_OK) return s;
    // ---- operators: level 0 = the Gram matrix at xyz, then Galerkin (or given)
    s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s == RC_OK) s = rc_launch_gram(ctx, p, ctx->stream);
    if (s != RC_OK) return s;
    const int Vp = ctx->nblk * RC_NB;
    {
        DevLevel& l0 = mg.L[0];
        hipLaunchKernelGGL(k_mg_copy_a, dim3((l0.n + 255) / 256, l0.n), dim3(256), 0, mg.s,
                           (const double*)ctx->d_F, Vp, l0.n, l0.A);
        MG_LAUNCH();
    }
    for (int li = 1; li < nlevels; ++li) {
        DevLevel& l = mg.L[li];
        const DevLevel& f = mg.L[li - 1];
        if (levels[li].A) {
            MG_CHECK(hipMemcpyAsync(l.A, levels[li].A, sizeof(double) * l.n * l.n,
                                    hipMemcpyHostToDevice, mg.s));
            continue;
        }
        hipLaunchKernelGGL(k_mg_gal_right, dim3((l.n + 255) / 256, f.n), dim3(256), 0, mg.s,
                           (const double*)f.A, f.n, f.n, l.n, (const int32_t*)l.ch_off,
                           (const int32_t*)l.ch_idx, (const double*)l.ch_w, mg.d_T);
        MG_LAUNCH();
        hipLaunchKernelGGL(k_mg_gal_left, dim3((l.n + 255) / 256, l.n), dim3(256), 0, mg.s,
                           (const double*)mg.d_T, l.n, (const int32_t*)l.ch_off,
                           (const int32_t*)l.ch_idx, (const double*)l.ch_w, l.A);
        MG_LAUNCH();
    }
    // host copies of the operators solved directly (the coarsest level, and
    // every level of <= 8 vertices)
    for (int li = 0; li < nlevels; ++li) {
        DevLevel& l = mg.L[li];
        if (li == nlevels - 1 || l.n <= 8) {
            l.hA.resize((size_t)l.n * l.n);
            MG_CHECK(hipMemcpyAsync(l.hA.data(), l.A, sizeof(double) * l.n * l.n,
                                    hipMemcpyDeviceToHost, mg.s));
        }
    }
    DevLevel& l0 = mg.L[0];
    const int m0 = 3 * l0.n;
    MG_CHECK(hipMemcpyAsync(l0.rhs, rhs, sizeof(double) * m0, hipMemcpyHostToDevice, mg.s));
    MG_CHECK(hipStreamSynchronize(mg.s));
    // ---- initial guess
    double* x = l0.x;
    const int nu_pre = o->nu_pre, nu_post = o->nu_post;
    if (o->init_nested && nlevels > 1) {
        // rhs_l = Pi_l P^T rhs_{l-1}; direct solve at the coarsest; prolong
        // and smooth level by level (nested_iteration_init)
        for (int li = 1; li < nlevels; ++li) {
            DevLevel& c = mg.L[li];
            hipLaunchKernelGGL(k_mg_restrict, dim3(1), dim3(WG), 0, mg.s, c.lev(), mg.L[li - 1].n,
                               (const int32_t*)c.ch_off, (const int32_t*)c.ch_idx,
                               (const double*)c.ch_w, (const double*)mg.L[li - 1].rhs, c.rhs);
            MG_LAUNCH();
        }
        s = direct(mg, nlevels - 1, mg.L[nlevels - 1].rhs, mg.L[nlevels - 1].x);
        if (s != RC_OK) return s;
        for (int li = nlevels - 2; li >= 0; --li) {
            DevLevel& f = mg.L[li];
            const DevLevel& c = mg.L[li + 1];
            hipLaunchKernelGGL(k_mg_prolong, dim3(1), dim3(WG), 0, mg.s, f.lev(), c.n,
                               (const int32_t*)c.par, (const double*)c.pw, (const double*)c.x, f.x, 0);
            MG_LAUNCH();
            s = smooth(mg, li, f.x, f.rhs, nu_pre);
            if (s != RC_OK) return s;
        }
    } else {
        hipLaunchKernelGGL(k_mg_fill, dim3((m0 + 255) / 256), dim3(256), 0, mg.s, x, (size_t)m0, 0.0);
        MG_LAUNCH();
    }
    // ---- cycles with the divergence guard
    double best = INFINITY, rn = 0.0;
    int nh = 0;
    for (int it = 0; it <= o->max_cycles; ++it) {
        s = residual0(mg, x, l0.rhs, rn);
        if (s != RC_OK) return s;
        stats->history[nh++] = rn;
        if (!std::isfinite(rn) || rn > 10.0 * best) {
            stats->iterations = nh - 1;
            stats->residual = rn;
            stats->nhist = nh;
            ctx->err = "multigrid diverged at cycle " + std::to_string(it) + ": residual " +
                       std::to_string(rn) + " against best " + std::to_string(best);
            return RC_DIVERGED;
        }
        best = std::min(best, rn);
        if (rn <= o->tol) break;
        s = vcycle(mg, 0, x, l0.rhs, nu_pre, nu_post);
        if (s != RC_OK) return s;
    }
    s = residual0(mg, x, l0.rhs, rn);
    if (s != RC_OK) return s;
    int cleanup = 0;
    if (rn > o->tol) {
        // dense cleanup: dx = saddle solution for [rhs - A0 x; 0] (the
        // library's FP64 solve of level 0), x = Pi (x + dx), certified
        s = matvec(mg, l0, x, l0.Ap, nullptr);
        if (s != RC_OK) return s;
        std::vector<double> hx(m0), hAx(m0), hr(m0), aos((size_t)m0), dxa((size_t)m0);
        MG_CHECK(hipMemcpyAsync(hAx.data(), l0.Ap, sizeof(double) * m0, hipMemcpyDeviceToHost, mg.s));
        MG_CHECK(hipStreamSynchronize(mg.s));
        for (int i = 0; i < m0; ++i) hr[i] = rhs[i] - hAx[i];
        const int n0 = l0.n;
        for (int v = 0; v < n0; ++v)
            for (int c = 0; c < 3; ++c) aos[3 * v + c] = hr[c * n0 + v];
        rc_step_opts so;
        rc_default_step_opts(&so);
        so.params = *p;
        s = rc_solve_constrained_gradient(ctx, xyz, aos.data(), blocks, nblocks, &so, dxa.data(),
                                          nullptr, nullptr);
        if (s != RC_OK) return s;
        for (int v = 0; v < n0; ++v)
            for (int c = 0; c < 3; ++c) hx[c * n0 + v] = dxa[3 * v + c];
        MG_CHECK(hipMemcpyAsync(l0.t, hx.data(), sizeof(double) * m0, hipMemcpyHostToDevice, mg.s));
        MG_CHECK(hipStreamSynchronize(mg.s));
        hipLaunchKernelGGL(k_mg_cleanup_add, dim3(1), dim3(WG), 0, mg.s, l0.lev(), x,
                           (const double*)l0.t);
        MG_LAUNCH();
        cleanup = 1;
        s = residual0(mg, x, l0.rhs, rn);
        if (s != RC_OK) return s;
        if (nh < RC_MG_MAXHIST) stats->history[nh++] = rn;
        if (!std::isfinite(rn) || rn > 100.0 * o->tol) {
            stats->iterations = nh - 1;
            stats->residual = rn;
            stats->nhist = nh;
            stats->used_direct_cleanup = 1;
            ctx->err = "multigrid: dense cleanup failed to certify (residual " +
                       std::to_string(rn) + ")";
            return RC_DIVERGED;
        }
    }
    MG_CHECK(hipMemcpyAsync(x_out, x, sizeof(double) * m0, hipMemcpyDeviceToHost, mg.s));
    MG_CHECK(hipStreamSynchronize(mg.s));
    stats->iterations = nh - 1;
    stats->residual = rn;
    stats->nhist = nh;
    stats->used_direct_cleanup = cleanup;
    return RC_OK;
}

// y = A u per coordinate block with the Gram matrix of the curve at xyz
// (u, y: 3V, coordinate blocks)
extern "C" rc_status rc_gram_apply(rc_ctx* ctx, const double* xyz, const rc_params* p,
                                   const double* u, double* y) {
    if (!ctx || !xyz || !p || !u || !y) return RC_BAD_ARG;
    Mg mg;
    mg.ctx = ctx;
    mg.s = ctx->stream;
    const int V = ctx->V, Vp = ctx->nblk * RC_NB;
    double *du = nullptr, *dy = nullptr;
    rc_status s = mg.alloc(&du, (size_t)3 * V);
    if (s == RC_OK) s = mg.alloc(&dy, (size_t)3 * V);
    if (s == RC_OK) s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s == RC_OK) s = rc_launch_gram(ctx, p, ctx->stream);
    if (s != RC_OK) return s;
    MG_CHECK(hipMemcpyAsync(du, u, sizeof(double) * 3 * V, hipMemcpyHostToDevice, mg.s));
    hipLaunchKernelGGL(k_mg_matvec, dim3((V + 3) / 4), dim3(256), 0, mg.s, (const double*)ctx->d_F,
                       Vp, V, (const double*)du, dy, (const int32_t*)nullptr);
    MG_LAUNCH();
    MG_CHECK(hipMemcpyAsync(y, dy, sizeof(double) * 3 * V, hipMemcpyDeviceToHost, mg.s));
    MG_CHECK(hipStreamSynchronize(mg.s));
    return RC_OK;
}

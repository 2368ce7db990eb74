// Iterative refinement of the saddle solution with double-double residuals.
//
// kappa(A_hat) grows like V^2.3 for the fractional metric (~1e9 at V = 8192),
// so a solve whose residual is formed in FP64 cannot get the direction below
// ~kappa * u_64 relative error: a blocked FP64 Cholesky lands at ~2e-8 there,
// and so does any method that stops on an FP64 residual. Each refinement step
// here forms the residual of the reference system
//     [[I3 (x) A, C^T],[C, 0]] [g~; lambda] = [dE; 0]     (linsolve.ts:771-790)
// with every product split exactly (TwoProduct by fma) and every sum carried
// as an unevaluated (sum, error) pair (compensated dot products: the result
// is as accurate as if formed in twice the working precision, then rounded
// once), and corrects [g~; lambda] through the same Schur reduction as the
// solve: dx = A_hat^-1 (rho - C^T dlambda) + N dt / V with
//     [dlambda; dt] = M^-1 [C A_hat^-1 rho - rho_lambda; N^T rho],
// A_hat^-1 by the FP64 factor's forward and backward sweeps (plain form) or
// by PCG with the FP32 factor (mixed form). The fixed point is the exact
// solution for the assembled A, C and dE, whatever the accuracy of the
// factor: the correction only has to contract the error.
#include "rc_internal.h"
#include "rc_step.h"

rc_status rc_solve_mixed(rc_ctx* ctx, const double* d_B, double* d_X, int nrhs, double tol,
                         int max_iter, double* d_refine_out);

namespace {

// (s, e): an unevaluated sum s + e, e collecting the rounding errors.
struct cpair {
    double s, e;
};

// a + b exactly as (s, t) (TwoSum). Its operands are never products the
// compiler could fuse into these additions: cp_fma pins its rounded product
// (the library builds with -ffp-contract=fast, which the backend applies
// regardless of source-level contraction pragmas).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& t) {
    s = a + b;
    const double bb = s - a;
    t = (a - (s - bb)) + (b - bb);
}

// acc += a * x (TwoProduct by fma, TwoSum into the pair). The empty asm
// makes the rounded product p opaque, so no later addition is fused with
// a * x (which would add the exact product where TwoSum assumes p).
__device__ __forceinline__ void cp_fma(cpair& acc, double a, double x) {
    double p = a * x;
    asm volatile("" : "+v"(p));
    const double pe = fma(a, x, -p);
    double s, t;
    two_sum(acc.s, p, s, t);
    acc.s = s;
    acc.e = acc.e + (t + pe);
}

__device__ __forceinline__ cpair cp_add(cpair a, cpair b) {
    cpair r;
    double t;
    two_sum(a.s, b.s, r.s, t);
    r.e = (a.e + b.e) + t;
    return r;
}

__device__ __forceinline__ cpair cp_shfl_xor(cpair a, int off) {
    cpair r;
    r.s = __shfl_xor(a.s, off, 64);
    r.e = __shfl_xor(a.e, off, 64);
    return r;
}

__device__ __forceinline__ void upper_tile_of(int p, int& I, int& J) {
    int j = (int)((__builtin_sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= p) ++j;
    while (j * (j + 1) / 2 > p) --j;
    J = j;
    I = p - j * (j + 1) / 2;
}

// Compensated A x for the three coordinate blocks, A from the strict upper
// triangle of the factor storage: one workgroup per 64 x 64 tile (I, J),
// I <= J, read once; its row sums are block I's partial from J, its column
// sums block J's partial from I (diagonal tiles: strict upper part, both to
// one partial). part_s / part_e [I][J][c][i].
__global__ __launch_bounds__(256) void k_ir_tiles(const double* __restrict__ F, int V, int Vp,
                                                  int nblk, const double* __restrict__ x,
                                                  double* __restrict__ part_s,
                                                  double* __restrict__ part_e,
                                                  const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double rsh[RC_NB][3][2];
    __shared__ double csh[4][RC_NB][3][2];
    int I, J;
    upper_tile_of(blockIdx.x, I, J);
    const int t = threadIdx.x, w = t >> 6;
    const int ri = t >> 4, cj = t & 15;
    const int i0 = I * RC_NB + 4 * ri, j0 = J * RC_NB + 4 * cj;
    double a[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double* p = F + (size_t)(i0 + r) * Vp + j0;
        const double2 v0 = *reinterpret_cast<const double2*>(p);
        const double2 v1 = *reinterpret_cast<const double2*>(p + 2);
        a[r][0] = v0.x;
        a[r][1] = v0.y;
        a[r][2] = v1.x;
        a[r][3] = v1.y;
    }
    double xi[4][3], xj[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xi[q][c] = x[(size_t)c * Vp + i0 + q];
            xj[q][c] = x[(size_t)c * Vp + j0 + q];
        }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool on = i0 + r < V && j0 + k < V && (I < J || j0 + k > i0 + r);
            a[r][k] = on ? a[r][k] : 0.0;
        }
    cpair rs[4][3], cs[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cpair u = {0.0, 0.0}, v = {0.0, 0.0};
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                cp_fma(u, a[q][m], xj[m][c]);
                cp_fma(v, a[m][q], xi[m][c]);
            }
            rs[q][c] = u;
            cs[q][c] = v;
        }
#pragma unroll
    for (int off = 1; off < 16; off <<= 1)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) rs[q][c] = cp_add(rs[q][c], cp_shfl_xor(rs[q][c], off));
#pragma unroll
    for (int off = 16; off < 64; off <<= 1)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) cs[q][c] = cp_add(cs[q][c], cp_shfl_xor(cs[q][c], off));
    if (cj == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                rsh[4 * ri + q][c][0] = rs[q][c].s;
                rsh[4 * ri + q][c][1] = rs[q][c].e;
            }
    if (((t >> 4) & 3) == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                csh[w][4 * cj + q][c][0] = cs[q][c].s;
                csh[w][4 * cj + q][c][1] = cs[q][c].e;
            }
    __syncthreads();
    if (t < 3 * RC_NB) {
        const int c = t >> 6, i = t & 63;
        cpair col = {csh[0][i][c][0], csh[0][i][c][1]};
#pragma unroll
        for (int ww = 1; ww < 4; ++ww) col = cp_add(col, cpair{csh[ww][i][c][0], csh[ww][i][c][1]});
        const cpair row = {rsh[i][c][0], rsh[i][c][1]};
        if (I == J) {
            const cpair both = cp_add(row, col);
            const size_t o = (((size_t)I * nblk + I) * 3 + c) * RC_NB + i;
            part_s[o] = both.s;
            part_e[o] = both.e;
        } else {
            const size_t o1 = (((size_t)I * nblk + J) * 3 + c) * RC_NB + i;
            const size_t o2 = (((size_t)J * nblk + I) * 3 + c) * RC_NB + i;
            part_s[o1] = row.s;
            part_e[o1] = row.e;
            part_s[o2] = col.s;
            part_e[o2] = col.e;
        }
    }
}

// rho_c[v] = dE_c[v] - (A x)_c[v] - sum_p C[p][c Vp + v] lambda_p, the sum
// compensated and rounded once; also into Z (the correction solve's right-
// hand side, zero on padding). rn[(I * 3 + c) * 2 + 0 / 1]: the block's
// sums of rho^2 and dE^2. One wave per (64-row block I, component c).
__global__ __launch_bounds__(64) void k_ir_rho(const double* __restrict__ part_s,
                                               const double* __restrict__ part_e,
                                               const double* __restrict__ d,
                                               const double* __restrict__ x,
                                               const double* __restrict__ C,
                                               const double* __restrict__ lam, int k,
                                               const double* __restrict__ dE, int V, int Vp,
                                               int nblk, double* __restrict__ rho,
                                               double* __restrict__ Z, double* __restrict__ rn,
                                               const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int I = blockIdx.x, c = blockIdx.y, i = threadIdx.x, v = I * RC_NB + i;
    const size_t j = (size_t)c * Vp + v;
    double r = 0.0, b = 0.0;
    if (v < V) {
        cpair s = {0.0, 0.0};
        cp_fma(s, d[v], x[j]);
        const size_t base = ((size_t)I * nblk * 3 + c) * RC_NB + i;
#pragma unroll 8
        for (int J = 0; J < nblk; ++J) {
            const size_t o = base + (size_t)J * 3 * RC_NB;
            s = cp_add(s, cpair{part_s[o], part_e[o]});
        }
        for (int p = 0; p < k; ++p) cp_fma(s, C[(size_t)p * 3 * Vp + j], lam[p]);
        b = dE[j];
        double h, t;
        two_sum(b, -s.s, h, t);
        r = h + (t - s.e);
    }
    rho[j] = r;
    Z[j] = r;
    double r2 = r * r, b2 = b * b;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        r2 += __shfl_xor(r2, off, 64);
        b2 += __shfl_xor(b2, off, 64);
    }
    if (i == 0) {
        rn[((size_t)I * 3 + c) * 2] = r2;
        rn[((size_t)I * 3 + c) * 2 + 1] = b2;
    }
}

// rho_lambda[p] = 0 - C_p . x (compensated), one workgroup per row p.
__global__ __launch_bounds__(RED_T) void k_ir_rhob(const double* __restrict__ C,
                                                   const double* __restrict__ x, int V, int Vp,
                                                   double* __restrict__ rhob,
                                                   const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[RED_T / 64][2];
    const int p = blockIdx.x;
    const double* row = C + (size_t)p * 3 * Vp;
    cpair s = {0.0, 0.0};
#pragma unroll 8
    for (int q = threadIdx.x; q < 3 * V; q += RED_T) {
        const int c = q / V, v = q - c * V;
        const size_t j = (size_t)c * Vp + v;
        cp_fma(s, row[j], x[j]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s = cp_add(s, cp_shfl_xor(s, off));
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6][0] = s.s;
        sh[threadIdx.x >> 6][1] = s.e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        cpair a = {sh[0][0], sh[0][1]};
        for (int ww = 1; ww < RED_T / 64; ++ww) a = cp_add(a, cpair{sh[ww][0], sh[ww][1]});
        rhob[p] = -(a.s + a.e);
    }
}

// Compensated R = B - A_hat X for columns 0 .. nc-1 (nc <= 32) of X, B, R
// (ld Vp), A_hat = A + mu 1 1^T with A from the strict upper triangle of the
// factor storage and its diagonal d; mu (1^T X) is added in FP64 (the 1
// direction is where A_hat is well conditioned). One workgroup per 16-row
// block walks the 64-column tiles of its rows: upper entries read along the
// stored rows, lower ones as the stored transpose (16 consecutive rows of a
// stored column), both into LDS; thread (row t % 16, column pair t / 16)
// accumulates its two (sum, error) pairs over every column of A.
__global__ __launch_bounds__(256) void k_ir_cols(const double* __restrict__ F,
                                                 const double* __restrict__ d, int V, int Vp,
                                                 const double* __restrict__ X,
                                                 const double* __restrict__ B,
                                                 const double* __restrict__ colsum,
                                                 const double* __restrict__ mu, int nc,
                                                 double* __restrict__ R) {
    __shared__ double As[16][65];
    __shared__ double Xs[64][33];
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * 16;
    const int ii = t & 15, cg = t >> 4;
    cpair acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
    const int ntile = (V + 63) / 64;
    for (int J = 0; J < ntile; ++J) {
        const int jb = J * 64;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            // upper entries (j > i): columns fastest
            {
                const int r = e >> 6, c = e & 63, i = i0 + r, j = jb + c;
                if (i < V && j < V && j > i) As[r][c] = F[(size_t)i * Vp + j];
            }
            // lower entries (j < i) from the stored transpose: rows fastest
            {
                const int r = e & 15, c = e >> 4, i = i0 + r, j = jb + c;
                if (i < V && j < V && j < i) As[r][c] = F[(size_t)j * Vp + i];
                if (j == i && i < V) As[r][c] = d[i];
                if (i >= V || j >= V) As[r][c] = 0.0;
            }
        }
        for (int e = t; e < 64 * 32; e += 256) {
            const int c = e >> 6, jj = e & 63, j = jb + jj;
            Xs[jj][c] = (c < nc && j < V) ? X[(size_t)c * Vp + j] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int jj = 0; jj < 64; ++jj) {
            const double a = As[ii][jj];
            cp_fma(acc[0], a, Xs[jj][2 * cg]);
            cp_fma(acc[1], a, Xs[jj][2 * cg + 1]);
        }
    }
    const int i = i0 + ii;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int c = 2 * cg + q;
        if (c >= nc || i >= Vp) continue;
        double r = 0.0;
        if (i < V) {
            double h, tt;
            two_sum(B[(size_t)c * Vp + i], -acc[q].s, h, tt);
            r = h + ((tt - acc[q].e) - mu[0] * colsum[c]);
        }
        R[(size_t)c * Vp + i] = r;
    }
}

// out[c] = sum_v X[c][v] over v < V, one workgroup per column.
__global__ __launch_bounds__(256) void k_ir_colsum(const double* __restrict__ X, int V, int Vp,
                                                   double* __restrict__ out) {
    __shared__ double sh[4];
    const int c = blockIdx.x;
    double s = 0.0;
#pragma unroll 8
    for (int v = threadIdx.x; v < V; v += 256) s += X[(size_t)c * Vp + v];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[c] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// X += Z over every column (v < V); dn[blockIdx.x * 2 + 0 / 1] = this
// block's sums of Z^2 and X^2 over the first three columns (dE's solution).
__global__ __launch_bounds__(256) void k_ir_addcols(double* __restrict__ X,
                                                    const double* __restrict__ Z, int V, int Vp,
                                                    int ncols, double* __restrict__ dn) {
    __shared__ double sh[2][4];
    const int v = blockIdx.x * 256 + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (v < V) {
        for (int c = 0; c < ncols; ++c) {
            const size_t j = (size_t)c * Vp + v;
            const double z = Z[j];
            const double x = X[j] + z;
            X[j] = x;
            if (c < 3) {
                a += z * z;
                b += x * x;
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        dn[blockIdx.x * 2] = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
        dn[blockIdx.x * 2 + 1] = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
    }
}

// Relative residual ||rho|| / max(1, ||[dE; 0]||) of the current iterate
// (block partials rn, rho_lambda) into out[0].
__global__ __launch_bounds__(256) void k_ir_resnorm(const double* __restrict__ rn, int nparts,
                                                    const double* __restrict__ rhob, int k,
                                                    double* __restrict__ out,
                                                    const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[2][256];
    double r2 = 0.0, b2 = 0.0;
    for (int q = threadIdx.x; q < nparts; q += 256) {
        r2 += rn[2 * q];
        b2 += rn[2 * q + 1];
    }
    for (int p = threadIdx.x; p < k; p += 256) r2 += rhob[p] * rhob[p];
    sh[0][threadIdx.x] = r2;
    sh[1][threadIdx.x] = b2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(sh[0][0]) / fmax(1.0, sqrt(sh[1][0]));
}

// w[p] = <C_p, Z> for p < k; w[k + c] = sum_v rho_c[v]. One workgroup each.
__global__ __launch_bounds__(RED_T) void k_ir_dots(const double* __restrict__ C,
                                                   const double* __restrict__ Z,
                                                   const double* __restrict__ rho, int k, int V,
                                                   int Vp, double* __restrict__ w,
                                                   const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[RED_T / 64];
    const int p = blockIdx.x;
    double s = 0.0;
    if (p < k) {
        const double* row = C + (size_t)p * 3 * Vp;
#pragma unroll 8
        for (int q = threadIdx.x; q < 3 * V; q += RED_T) {
            const int c = q / V, v = q - c * V;
            const size_t j = (size_t)c * Vp + v;
            s = fma(row[j], Z[j], s);
        }
    } else {
        const double* col = rho + (size_t)(p - k) * Vp;
#pragma unroll 8
        for (int v = threadIdx.x; v < V; v += RED_T) s += col[v];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int ww = 0; ww < RED_T / 64; ++ww) a += sh[ww];
        w[p] = a;
    }
}

// [dlambda; dt] = M^-1 [w_{<k} - rho_lambda; w_{>=k}], one wave per row i;
// lambda / t (sol) advance by it.
__global__ __launch_bounds__(64) void k_ir_small(const double* __restrict__ Minv,
                                                 const double* __restrict__ w,
                                                 const double* __restrict__ rhob, int k,
                                                 double* __restrict__ delta,
                                                 double* __restrict__ sol,
                                                 const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int n = k + 3, i = blockIdx.x;
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += 64) {
        const double rj = (j < k) ? w[j] - rhob[j] : w[j];
        s = fma(Minv[(size_t)i * n + j], rj, s);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) {
        delta[i] = s;
        sol[i] += s;
    }
}

// dx_c[v] = Z_c[v] - sum_q dlambda_q Y_{q,c}[v] + dt_c / V; g~ += dx.
// dn[(blockIdx.y * gridDim.x + blockIdx.x) * 2 + 0 / 1]: sums of dx^2, g~^2.
__global__ __launch_bounds__(256) void k_ir_update(const double* __restrict__ Z,
                                                   const double* __restrict__ Y,
                                                   const double* __restrict__ delta, int k,
                                                   int V, int Vp, double* __restrict__ gt,
                                                   double* __restrict__ dn,
                                                   const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[2][4];
    const int v = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    double dx = 0.0, g = 0.0;
    if (v < V) {
        const size_t j = (size_t)c * Vp + v;
        dx = Z[j];
        for (int q = 0; q < k; ++q) dx -= delta[q] * Y[(size_t)(3 * q + c) * Vp + v];
        dx += delta[k + c] / V;
        g = gt[j] + dx;
        gt[j] = g;
    }
    double a = dx * dx, b = g * g;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const size_t o = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 2;
        dn[o] = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
        dn[o + 1] = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
    }
}

// ratio[0] = ||dx|| / ||g~|| of the correction just applied; count[0] += 1.
__global__ __launch_bounds__(256) void k_ir_corrnorm(const double* __restrict__ dn, int nparts,
                                                     double* __restrict__ ratio,
                                                     double* __restrict__ count,
                                                     const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[2][256];
    double a = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < nparts; q += 256) {
        a += dn[2 * q];
        b += dn[2 * q + 1];
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ratio[0] = (sh[1][0] > 0.0) ? sqrt(sh[0][0] / sh[1][0]) : 0.0;
        count[0] += 1.0;
    }
}

__global__ void k_ir_zero(double* __restrict__ out, int n) {
    if ((int)threadIdx.x < n) out[threadIdx.x] = 0.0;
}

// PCG iterations of a correction (out[5], written by the mixed solve) into
// the running total out[3].
__global__ void k_ir_accum_iters(double* __restrict__ out) { out[3] += out[5]; }

}  // namespace

// Workspace of the refinement (allocated on first use, sized for the
// context's constraint-row capacity; dropped when that grows).
static size_t ir_cols_offset(const rc_ctx* ctx) {
    const size_t Vp = (size_t)ctx->nblk * RC_NB;
    const size_t nb = (size_t)ctx->nblk;
    const size_t n = (size_t)ctx->kcap + 3;
    return 2 * nb * nb * 3 * RC_NB + 2 * 3 * Vp + 3 * n + 2 * 3 * nb +
           2 * 3 * ((Vp + 255) / 256) + 64;
}

static rc_status ir_bufs(rc_ctx* ctx) {
    if (ctx->d_ir) return RC_OK;
    const size_t Vp = (size_t)ctx->nblk * RC_NB;
    const size_t nr = (size_t)ctx->nrhs_cap;
    const size_t total = ir_cols_offset(ctx)       // the saddle residual's workspace
                         + 2 * nr * Vp             // column residuals R, corrections Z
                         + nr                      // column sums
                         + 2 * ((Vp + 255) / 256); // correction norm partials
    RC_HIP(hipMalloc((void**)&ctx->d_ir, sizeof(double) * total));
    ctx->ir_doubles = total;
    return RC_OK;
}

// Refinement steps on the solution of solve_phase: g~ (ctx->d_gt) and
// [lambda; t] (ctx->d_ssol), with Y = X columns 3.. (ctx->d_sol + 3 Vp) and
// M^-1 (ctx->d_Minv) of the Schur reduction. A in the strict upper triangle
// of the factor storage and ctx->d_Adiag (both forms keep it there). nsteps
// corrections; out[0] = relative compensated residual (of the returned
// iterate once rc_saddle_final_residual has run), out[1] = corrections applied, out[2] = ||dx|| / ||g~|| of the
// last one, out[3] = PCG iterations of the corrections (mixed form),
// out[4..5] the mixed solve's own (residual, iterations) slots.
rc_status rc_refine_zero(rc_ctx* ctx, double* d_out) {
    hipLaunchKernelGGL(k_ir_zero, dim3(1), dim3(64), 0, ctx->stream, d_out, 6);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

namespace {

struct IrBufs {
    double *part_s, *part_e, *rho, *Z, *rhob, *wv, *delta, *rn, *dn;
};

IrBufs ir_layout(rc_ctx* ctx) {
    const int Vp = ctx->nblk * RC_NB;
    const int nb = ctx->nblk;
    const size_t np = (size_t)nb * nb * 3 * RC_NB;
    IrBufs b;
    b.part_s = ctx->d_ir;
    b.part_e = b.part_s + np;
    b.rho = b.part_e + np;
    b.Z = b.rho + 3 * (size_t)Vp;
    b.rhob = b.Z + 3 * (size_t)Vp;
    b.wv = b.rhob + (ctx->kcap + 3);
    b.delta = b.wv + (ctx->kcap + 3);
    b.rn = b.delta + (ctx->kcap + 3);
    b.dn = b.rn + 2 * 3 * (size_t)nb;
    return b;
}

// Compensated saddle residual of the current [g~; lambda] on stream st:
// rho (top, coordinate blocks), rhob (constraint rows), d_out[0] = relative
// norm.
rc_status ir_residual(rc_ctx* ctx, int k, hipStream_t st, const IrBufs& b, double* d_out) {
    const int Vp = ctx->nblk * RC_NB;
    const int nb = ctx->nblk;
    hipLaunchKernelGGL(k_ir_tiles, dim3(nb * (nb + 1) / 2), dim3(256), 0, st,
                       (const double*)ctx->d_F, ctx->V, Vp, nb, (const double*)ctx->d_gt,
                       b.part_s, b.part_e, (const int32_t*)nullptr);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ir_rho, dim3(nb, 3), dim3(64), 0, st, (const double*)b.part_s,
                       (const double*)b.part_e, (const double*)ctx->d_Adiag,
                       (const double*)ctx->d_gt, (const double*)ctx->d_C,
                       (const double*)ctx->d_ssol, k, (const double*)ctx->d_dE, ctx->V, Vp, nb,
                       b.rho, b.Z, b.rn, (const int32_t*)nullptr);
    RC_CHECK_LAUNCH();
    if (k > 0) {
        hipLaunchKernelGGL(k_ir_rhob, dim3(k), dim3(RED_T), 0, st, (const double*)ctx->d_C,
                           (const double*)ctx->d_gt, ctx->V, Vp, b.rhob,
                           (const int32_t*)nullptr);
        RC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ir_resnorm, dim3(1), dim3(256), 0, st, (const double*)b.rn, 3 * nb,
                       (const double*)b.rhob, k, d_out, (const int32_t*)nullptr);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

}  // namespace

rc_status rc_saddle_refine(rc_ctx* ctx, int k, int precision, int nsteps, double pcg_tol,
                           int pcg_max_iter, double* d_out) {
    rc_status s = ir_bufs(ctx);
    if (s != RC_OK) return s;
    const int Vp = ctx->nblk * RC_NB;
    const IrBufs b = ir_layout(ctx);
    const int ngx = (Vp + 255) / 256;
    const double* Y = ctx->d_sol + 3 * (size_t)Vp;
    for (int it = 0; it < nsteps; ++it) {
        s = ir_residual(ctx, k, ctx->stream, b, d_out);
        if (s != RC_OK) return s;
        // Z = A_hat^-1 rho (three columns)
        if (precision == RC_PRECISION_MIXED) {
            // PCG writes its (max residual, iterations) pair into d_out[4..5]
            s = rc_solve_mixed(ctx, b.rho, b.Z, 3, pcg_tol, pcg_max_iter, d_out + 4);
            if (s != RC_OK) return s;
        } else {
            s = rc_solve_fp64_forward(ctx, b.Z, 3);
            if (s == RC_OK) s = rc_solve_fp64_backward(ctx, b.Z, 3);
            if (s != RC_OK) return s;
        }
        hipLaunchKernelGGL(k_ir_dots, dim3(k + 3), dim3(RED_T), 0, ctx->stream,
                           (const double*)ctx->d_C, (const double*)b.Z, (const double*)b.rho, k,
                           ctx->V, Vp, b.wv, (const int32_t*)nullptr);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ir_small, dim3(k + 3), dim3(64), 0, ctx->stream,
                           (const double*)ctx->d_Minv, (const double*)b.wv, (const double*)b.rhob,
                           k, b.delta, ctx->d_ssol, (const int32_t*)nullptr);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ir_update, dim3(ngx, 3), dim3(256), 0, ctx->stream,
                           (const double*)b.Z, Y, (const double*)b.delta, k, ctx->V, Vp, ctx->d_gt,
                           b.dn, (const int32_t*)nullptr);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ir_corrnorm, dim3(1), dim3(256), 0, ctx->stream, (const double*)b.dn,
                           3 * ngx, d_out + 2, d_out + 1, (const int32_t*)nullptr);
        RC_CHECK_LAUNCH();
        if (precision == RC_PRECISION_MIXED) {
            hipLaunchKernelGGL(k_ir_accum_iters, dim3(1), dim3(1), 0, ctx->stream, d_out);
            RC_CHECK_LAUNCH();
        }
    }
    return RC_OK;
}

// The compensated residual of the refined solution (d_out[0]), a diagnostic
// nothing on the step reads: launched on stream st after the refinement
// (the caller orders st after it), beside the line search.
rc_status rc_saddle_final_residual(rc_ctx* ctx, int k, hipStream_t st, double* d_out) {
    rc_status s = ir_bufs(ctx);
    if (s != RC_OK) return s;
    return ir_residual(ctx, k, st, ir_layout(ctx), d_out);
}

// Refinement of the solve's columns X = A_hat^-1 B (ctx->d_sol, the 3 + 3k
// columns [dE, C^T], ld Vp) with compensated residuals R = B - A_hat X
// (B rebuilt in ctx->d_rhs) and corrections Z = A_hat^-1 R: the FP64
// factor's forward and backward sweeps (plain form) or PCG with the FP32
// factor to a relative residual pcg_tol (mixed form); X += Z. With exact
// columns the Schur step that follows forms M, lambda, g~ and every frozen
// projection correction from the exact A_hat^-1 C^T. out[1] = steps,
// out[2] = ||Z|| / ||X|| over dE's three columns in the last step, out[3] =
// PCG iterations of the corrections (out[4..5] the mixed solve's slots).
rc_status rc_launch_build_rhs(rc_ctx* ctx, const double* dE, int k, double* B, int ncols);

rc_status rc_refine_columns(rc_ctx* ctx, int k, int precision, int nsteps, double pcg_tol,
                            int pcg_max_iter, double* d_out) {
    rc_status s = ir_bufs(ctx);
    if (s != RC_OK) return s;
    const int Vp = ctx->nblk * RC_NB;
    const int nrhs = 3 + 3 * k;
    double* R = ctx->d_ir + ir_cols_offset(ctx);
    double* Z = R + (size_t)ctx->nrhs_cap * Vp;
    double* cs = Z + (size_t)ctx->nrhs_cap * Vp;
    double* dn = cs + ctx->nrhs_cap;
    const int ngx = (Vp + 255) / 256;
    s = rc_launch_build_rhs(ctx, ctx->d_dE, k, ctx->d_rhs, nrhs);
    if (s != RC_OK) return s;
    for (int it = 0; it < nsteps; ++it) {
        hipLaunchKernelGGL(k_ir_colsum, dim3(nrhs), dim3(256), 0, ctx->stream,
                           (const double*)ctx->d_sol, ctx->V, Vp, cs);
        RC_CHECK_LAUNCH();
        for (int c0 = 0; c0 < nrhs; c0 += 32) {
            const int nc = nrhs - c0 < 32 ? nrhs - c0 : 32;
            hipLaunchKernelGGL(k_ir_cols, dim3(Vp / 16), dim3(256), 0, ctx->stream,
                               (const double*)ctx->d_F, (const double*)ctx->d_Adiag, ctx->V, Vp,
                               (const double*)(ctx->d_sol + (size_t)c0 * Vp),
                               (const double*)(ctx->d_rhs + (size_t)c0 * Vp),
                               (const double*)(cs + c0), (const double*)(ctx->d_scal + 1), nc,
                               R + (size_t)c0 * Vp);
            RC_CHECK_LAUNCH();
        }
        double* corr = R;
        if (precision == RC_PRECISION_MIXED) {
            s = rc_solve_mixed(ctx, R, Z, nrhs, pcg_tol, pcg_max_iter, d_out + 4);
            if (s != RC_OK) return s;
            corr = Z;
        } else {
            s = rc_solve_fp64_forward(ctx, R, nrhs);
            if (s == RC_OK) s = rc_solve_fp64_backward(ctx, R, nrhs);
            if (s != RC_OK) return s;
        }
        hipLaunchKernelGGL(k_ir_addcols, dim3(ngx), dim3(256), 0, ctx->stream, ctx->d_sol,
                           (const double*)corr, ctx->V, Vp, nrhs, dn);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ir_corrnorm, dim3(1), dim3(256), 0, ctx->stream, (const double*)dn,
                           ngx, d_out + 2, d_out + 1, (const int32_t*)nullptr);
        RC_CHECK_LAUNCH();
        if (precision == RC_PRECISION_MIXED) {
            hipLaunchKernelGGL(k_ir_accum_iters, dim3(1), dim3(1), 0, ctx->stream, d_out);
            RC_CHECK_LAUNCH();
        }
    }
    return RC_OK;
}

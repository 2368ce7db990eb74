// Constraint rows, penalties, the (k+3)^2 Schur system, the FP64 residual and
// the device-resident Armijo line search with frozen back-projection.
//
// Saddle reduction. With A_hat = A + mu 1 1^T (SPD; A 1 = 0), N = [1 (x) e_c]
// and Y = (I3 (x) A_hat)^{-1} C^T, the reference system
//   [[I3 (x) A, C^T],[C, 0]] [x; lambda] = [r; s]     (linsolve.ts:742-803)
// is equivalent to x = (I3 (x) A_hat)^{-1}(r - C^T lambda) + N t / V with
//   [[C Y, -C N / V],[(C N)^T, 0]] [lambda; t] = [C (I3 (x) A_hat)^{-1} r - s; N^T r].
// The gradient solve uses r = dE, s = 0; every frozen projection solve
// (lineSearch.ts:253-262) uses r = 0, s = -Phi and needs only Y and the
// inverse of the small matrix: O(V k) per projection iterate, no solve.
//
// Line search (lineSearch.ts:423-562): each trial is two kernels, a
// single-workgroup controller (candidate, projection loop, penalty energy,
// edge geometry of the projected curve) and the grid-wide energy kernel whose
// last workgroup applies the Armijo test. Trials are enqueued in batches; a
// finished line search turns the remaining kernels into immediate returns.
#include <algorithm>
#include <cmath>

#include "rc_internal.h"
#include "rc_handoff.h"
#include "rc_math.h"
#include "rc_pair_energy.h"
#include "rc_step.h"

namespace {

constexpr int CT = 1024;  // controller workgroup size

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v || o != o) ? o : v;  // NaN propagates like Math.max
    }
    return v;
}

// Block-wide sum over blockDim.x threads (fixed order). sh needs 32 doubles.
__device__ double block_sum(double v, double* sh) {
    v = wsum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += sh[i];
    __syncthreads();
    return s;
}
__device__ double block_max(double v, double* sh) {
    v = wmax(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = sh[0];
    for (int i = 1; i < nw; ++i) s = (sh[i] > s || sh[i] != sh[i]) ? sh[i] : s;
    __syncthreads();
    return s;
}

__device__ __forceinline__ double js_max(double a, double b) {
    if (a != a || b != b) return NAN;
    return a > b ? a : b;
}

// ---------------------------------------------------------------------------
// constraint Jacobian C (k x 3Vp, row-major, coordinate blocks of stride Vp),
// evaluated at the step base point from the edge table (raw lengths, guarded
// unit tangents). One thread per vertex gathers its incident edges in edge
// order (constraints.ts:67-120, constraintSet.ts:119-301). C pre-zeroed.
__global__ void k_constraint_jacobian(EdgeSoA ed, int V, int Vp, const int32_t* __restrict__ vOff,
                                      const int32_t* __restrict__ vEdge,
                                      const int32_t* __restrict__ vSlot,
                                      const DevSet* __restrict__ set, double* __restrict__ C) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const size_t m = 3 * (size_t)Vp;
    for (int b = 0; b < set->nblocks; ++b) {
        const int kind = set->kind[b];
        const int row = set->row0[b];
        if (kind == RC_BLOCK_BARYCENTER) {
            double acc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int q = vOff[v]; q < vOff[v + 1]; ++q) {
                const int K = vEdge[q];
                const double sg = vSlot[q] ? 1.0 : -1.0;
                const double T[3] = {ed.ux[K], ed.uy[K], ed.uz[K]};
                const double ell = ed.rlen[K];
                const double rv[3] = {0.5 * (ed.ax[K] + ed.bx[K]) - set->x0[0],
                                      0.5 * (ed.ay[K] + ed.by[K]) - set->x0[1],
                                      0.5 * (ed.az[K] + ed.bz[K]) - set->x0[2]};
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c)
                        acc[r][c] += sg * T[c] * rv[r] + 0.5 * ell * (r == c ? 1.0 : 0.0);
            }
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) C[(row + r) * m + (size_t)c * Vp + v] = acc[r][c];
        } else if (kind == RC_BLOCK_TOTAL_LENGTH) {
            double acc[3] = {0, 0, 0};
            for (int q = vOff[v]; q < vOff[v + 1]; ++q) {
                const int K = vEdge[q];
                const double sg = vSlot[q] ? -1.0 : 1.0;
                acc[0] += sg * ed.ux[K];
                acc[1] += sg * ed.uy[K];
                acc[2] += sg * ed.uz[K];
            }
            for (int c = 0; c < 3; ++c) C[row * m + (size_t)c * Vp + v] = acc[c];
        } else if (kind == RC_BLOCK_EDGE_LENGTHS) {
            for (int q = vOff[v]; q < vOff[v + 1]; ++q) {
                const int K = vEdge[q];
                const double sg = vSlot[q] ? -1.0 : 1.0;
                C[(row + K) * m + v] = sg * ed.ux[K];
                C[(row + K) * m + (size_t)Vp + v] = sg * ed.uy[K];
                C[(row + K) * m + 2 * (size_t)Vp + v] = sg * ed.uz[K];
            }
        } else if (kind == RC_BLOCK_POINT) {
            const int pv = set->pin_vertex[set->pin_of_block[b]];
            if (pv == v)
                for (int r = 0; r < 3; ++r) C[(row + r) * m + (size_t)r * Vp + v] = 1.0;
        }
    }
}

// ---------------------------------------------------------------------------
// penalties (penalties.ts:139-241), per-vertex gather of the edge stencils.
__global__ void k_penalty_grad_add(EdgeSoA ed, int V, int Vsrc, int Vdst,
                                   const int32_t* __restrict__ vOff,
                                   const int32_t* __restrict__ vEdge,
                                   const int32_t* __restrict__ vSlot, DevPen pen,
                                   const double* __restrict__ gsrc, double* __restrict__ gdst) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    double g[3] = {gsrc[v], gsrc[Vsrc + v], gsrc[2 * (size_t)Vsrc + v]};
    if (pen.active) {
        double p[3] = {0, 0, 0};
        for (int q = vOff[v]; q < vOff[v + 1]; ++q) {
            const int K = vEdge[q];
            const double sg = vSlot[q] ? 1.0 : -1.0;  // -at a, +at b
            const double T[3] = {ed.ux[K], ed.uy[K], ed.uz[K]};
            if (pen.wL != 0.0)
                for (int c = 0; c < 3; ++c) p[c] += sg * pen.wL * T[c];
            if (pen.wF != 0.0 && !(ed.rlen[K] < 1e-14)) {
                const double u = T[0] * pen.X[0] + T[1] * pen.X[1] + T[2] * pen.X[2];
                const double cc = 1.0 + u * u;
                for (int c = 0; c < 3; ++c) p[c] += sg * pen.wF * (cc * T[c] - 2.0 * u * pen.X[c]);
            }
            if (pen.wD != 0.0) {
                // interior (degree-2) endpoints w of K: K is w's first or second edge
                for (int e2 = 0; e2 < 2; ++e2) {
                    const int w = e2 ? ed.i2[K] : ed.i1[K];
                    if (vOff[w + 1] - vOff[w] != 2) continue;
                    const int rI = vEdge[vOff[w]], rJ = vEdge[vOff[w] + 1];
                    const double s = 2.0 * pen.wD * (ed.rlen[rI] - ed.rlen[rJ]);
                    const double sgn = (K == rI) ? 1.0 : -1.0;  // rI: -a +b ; rJ: +a -b
                    for (int c = 0; c < 3; ++c) p[c] += sgn * sg * s * T[c];
                }
            }
        }
        g[0] += p[0];
        g[1] += p[1];
        g[2] += p[2];
    }
    gdst[v] = g[0];
    gdst[Vdst + v] = g[1];
    gdst[2 * (size_t)Vdst + v] = g[2];
}

// Penalty energy of positions held in the edge table (one workgroup).
__device__ double penalty_energy_block(const EdgeSoA& ed, int E, const int32_t* vOff,
                                       const int32_t* vEdge, int V, const DevPen& pen,
                                       double* sh) {
    double sL = 0.0, sF = 0.0, sD = 0.0;
#pragma unroll 4
    for (int K = threadIdx.x; K < E; K += blockDim.x) {
        const double l = ed.rlen[K];
        sL += l;
        if (pen.wF != 0.0) {
            const double T[3] = {ed.ux[K], ed.uy[K], ed.uz[K]};
            const double cx = T[1] * pen.X[2] - T[2] * pen.X[1];
            const double cy = T[2] * pen.X[0] - T[0] * pen.X[2];
            const double cz = T[0] * pen.X[1] - T[1] * pen.X[0];
            sF += l * (cx * cx + cy * cy + cz * cz);
        }
    }
    if (pen.wD != 0.0)
#pragma unroll 8
        for (int w = threadIdx.x; w < V; w += blockDim.x)
            if (vOff[w + 1] - vOff[w] == 2) {
                const double d = ed.rlen[vEdge[vOff[w]]] - ed.rlen[vEdge[vOff[w] + 1]];
                sD += d * d;
            }
    sL = block_sum(sL, sh);
    sF = block_sum(sF, sh);
    sD = block_sum(sD, sh);
    double tot = 0.0;
    if (pen.wL != 0.0) tot += pen.wL * sL;
    if (pen.wD != 0.0) tot += pen.wD * sD;
    if (pen.wF != 0.0) tot += pen.wF * sF;
    return tot;
}

__global__ __launch_bounds__(CT) void k_penalty_energy(EdgeSoA ed, int E, int V,
                                                       const int32_t* __restrict__ vOff,
                                                       const int32_t* __restrict__ vEdge,
                                                       DevPen pen, double* __restrict__ out) {
    __shared__ double sh[32];
    const double e = pen.active ? penalty_energy_block(ed, E, vOff, vEdge, V, pen, sh) : 0.0;
    if (threadIdx.x == 0) out[0] = e;
}

// ---------------------------------------------------------------------------
// right-hand sides [dE_x, dE_y, dE_z, C_0 blocks, C_1 blocks, ...] (ld = Vp)
__global__ void k_build_rhs(const double* __restrict__ dE, const double* __restrict__ C, int V,
                            int Vp, int k, double* __restrict__ B) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (v >= Vp) return;
    double val = 0.0;
    if (v < V && col < 3 + 3 * k) {
        if (col < 3)
            val = dE[(size_t)col * Vp + v];
        else
            val = C[(size_t)(col - 3) * Vp + v];  // row q = (col-3)/3, block c = (col-3)%3
    }
    B[(size_t)col * Vp + v] = val;
}

// out[p * nr + q] = <L_p, R_q> over len entries (L stride ls, R stride rs).
// One workgroup (RED_T threads) per product: few products, each long.
__global__ __launch_bounds__(RED_T) void k_dots(const double* __restrict__ L, size_t ls,
                                                const double* __restrict__ R, size_t rs, int nr,
                                                size_t len, double* __restrict__ out) {
    __shared__ double sh[32];
    const int p = blockIdx.x / nr, q = blockIdx.x % nr;
    const double* a = L + p * ls;
    const double* b = R + q * rs;
    double s = 0.0;
#pragma unroll 8
    for (size_t j = threadIdx.x; j < len; j += RED_T) s = fma(a[j], b[j], s);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// The same products for many rows (sets beyond the LDS Schur bound): a
// workgroup per 8 x 8 tile of (p, q), each loaded entry used eight times;
// per-thread partial sums over j = t mod 256, then wave and workgroup sums.
__global__ __launch_bounds__(256) void k_dots_tile(const double* __restrict__ L, size_t ls,
                                                   int np, const double* __restrict__ R,
                                                   size_t rs, int nr, size_t len,
                                                   double* __restrict__ out) {
    __shared__ double red[4][64];
    const int ntq = (nr + 7) / 8;
    const int p0 = ((int)blockIdx.x / ntq) * 8, q0 = ((int)blockIdx.x % ntq) * 8;
    const double* a[8];
    const double* b[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        // rows past the end read the last row; their sums are not stored
        a[i] = L + (size_t)min(p0 + i, np - 1) * ls;
        b[i] = R + (size_t)min(q0 + i, nr - 1) * rs;
    }
    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int m = 0; m < 8; ++m) acc[i][m] = 0.0;
    for (size_t j = threadIdx.x; j < len; j += 256) {
        double av[8], bv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            av[i] = a[i][j];
            bv[i] = b[i][j];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int m = 0; m < 8; ++m) acc[i][m] = fma(av[i], bv[m], acc[i][m]);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const double v = wsum(acc[i][m]);
            if (lane == 0) red[w][i * 8 + m] = v;
        }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int i = threadIdx.x >> 3, m = threadIdx.x & 7;
        const double v = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                         (red[2][threadIdx.x] + red[3][threadIdx.x]);
        if (p0 + i < np && q0 + m < nr) out[(size_t)(p0 + i) * nr + q0 + m] = v;
    }
}

// Column sums of the coordinate blocks of C rows: CN[p][c] = sum_v C[p][c*Vp+v];
// also Nr[c] = sum_v dE[c*Vp + v] at CN[k*3 + c].
__global__ __launch_bounds__(RED_T) void k_block_sums(const double* __restrict__ C,
                                                      const double* __restrict__ dE, int k, int V,
                                                      int Vp, double* __restrict__ out) {
    __shared__ double sh[32];
    const int p = blockIdx.x / 3, c = blockIdx.x % 3;
    const double* src = (p < k) ? C + (size_t)p * 3 * Vp + (size_t)c * Vp : dE + (size_t)c * Vp;
    double s = 0.0;
#pragma unroll 8
    for (int v = threadIdx.x; v < V; v += RED_T) s += src[v];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// Small system M = [[S, -CN/V],[CN^T, 0]], n = k + 3. Partial-pivot LU in
// LDS; solves the gradient right-hand side and forms M^{-1} for the
// projection solves. dots[p*(k+1) + 0] = <C_p, X_r>, dots[p*(k+1)+1+q] = S_pq.
// out: sol[n] (lambda, t), Minv[n*n]. status bit 4 on a zero pivot.
__global__ __launch_bounds__(256) void k_schur(const double* __restrict__ dots,
                                               const double* __restrict__ bsum, int k, int V,
                                               double* __restrict__ sol,
                                               double* __restrict__ Minv,
                                               int32_t* __restrict__ status) {
    extern __shared__ double smem[];
    const int n = k + 3;
    double* M = smem;              // n x n
    double* Z = smem + n * n;      // n x (n + 1) right-hand sides
    int* piv = (int*)(Z + n * (n + 1));
    const int t = threadIdx.x;
    const double* CN = bsum;       // [k][3]
    const double* Nr = bsum + 3 * k;
    for (int q = t; q < n * n; q += blockDim.x) {
        const int i = q / n, j = q % n;
        double v = 0.0;
        if (i < k && j < k) v = dots[i * (k + 1) + 1 + j];
        else if (i < k) v = -CN[i * 3 + (j - k)] / V;
        else if (j < k) v = CN[j * 3 + (i - k)];
        M[q] = v;
    }
    for (int q = t; q < n * (n + 1); q += blockDim.x) {
        const int i = q / (n + 1), j = q % (n + 1);
        double v;
        if (j == 0) v = (i < k) ? dots[i * (k + 1)] : Nr[i - k];
        else v = (i == j - 1) ? 1.0 : 0.0;
        Z[q] = v;
    }
    __syncthreads();
    __shared__ int bad;
    if (t == 0) bad = 0;
    for (int c = 0; c < n; ++c) {
        __syncthreads();
        if (t == 0) {
            int pr = c;
            double mx = fabs(M[c * n + c]);
            for (int r = c + 1; r < n; ++r)
                if (fabs(M[r * n + c]) > mx) {
                    mx = fabs(M[r * n + c]);
                    pr = r;
                }
            piv[c] = pr;
            if (!(mx > 0.0)) bad = 1;
        }
        __syncthreads();
        const int pr = piv[c];
        if (pr != c) {
            for (int j = t; j < n; j += blockDim.x) {
                const double tmp = M[c * n + j];
                M[c * n + j] = M[pr * n + j];
                M[pr * n + j] = tmp;
            }
            for (int j = t; j < n + 1; j += blockDim.x) {
                const double tmp = Z[c * (n + 1) + j];
                Z[c * (n + 1) + j] = Z[pr * (n + 1) + j];
                Z[pr * (n + 1) + j] = tmp;
            }
        }
        __syncthreads();
        const double pv = M[c * n + c];
        for (int q = t; q < (n - c - 1) * (n + 1 + n - c - 1); q += blockDim.x) {
            const int w = n + 1 + n - c - 1;
            const int r = c + 1 + q / w, j = q % w;
            const double f = (pv != 0.0) ? M[r * n + c] / pv : 0.0;
            if (j < n - c - 1)
                M[r * n + c + 1 + j] -= f * M[c * n + c + 1 + j];
            else
                Z[r * (n + 1) + (j - (n - c - 1))] -= f * Z[c * (n + 1) + (j - (n - c - 1))];
        }
        __syncthreads();
    }
    // back substitution for all n+1 columns
    for (int j = t; j < n + 1; j += blockDim.x) {
        for (int i = n - 1; i >= 0; --i) {
            double s = Z[i * (n + 1) + j];
            for (int m = i + 1; m < n; ++m) s -= M[i * n + m] * Z[m * (n + 1) + j];
            Z[i * (n + 1) + j] = (M[i * n + i] != 0.0) ? s / M[i * n + i] : 0.0;
        }
    }
    __syncthreads();
    for (int q = t; q < n; q += blockDim.x) sol[q] = Z[q * (n + 1)];
    for (int q = t; q < n * n; q += blockDim.x) {
        const int i = q / n, j = q % n;
        Minv[q] = Z[i * (n + 1) + 1 + j];
    }
    if (t == 0 && bad) atomicOr(status, 4);
}

// The same system for n = k + 3 beyond the LDS bound, in a global workspace
// (M n x n, the right-hand sides Z n x (n + 1), U^T in an n x (n + 1) area):
// filled by k_schur_fill, then per column c a pivot step (one workgroup: the
// largest |M[r][c]|, r >= c, lowest row on ties, and the row swap) and an
// elimination step over the rows below (all workgroups), then the back
// substitution of every right-hand-side column in one launch — the
// partial-pivot LU of k_schur with the work of each step spread over the
// device.
__global__ void k_schur_fill(const double* __restrict__ dots, const double* __restrict__ bsum,
                             int k, int V, double* __restrict__ ws) {
    const int n = k + 3, w = n + 1;
    double* M = ws;
    double* Z = ws + (size_t)n * n;
    const double* CN = bsum;
    const double* Nr = bsum + 3 * k;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < (size_t)n * n) {
        const int i = (int)(q / n), j = (int)(q % n);
        double v = 0.0;
        if (i < k && j < k) v = dots[(size_t)i * (k + 1) + 1 + j];
        else if (i < k) v = -CN[i * 3 + (j - k)] / V;
        else if (j < k) v = CN[j * 3 + (i - k)];
        M[q] = v;
    } else if (q < (size_t)n * n + (size_t)n * w) {
        const size_t e = q - (size_t)n * n;
        const int i = (int)(e / w), j = (int)(e % w);
        double v;
        if (j == 0) v = (i < k) ? dots[(size_t)i * (k + 1)] : Nr[i - k];
        else v = (i == j - 1) ? 1.0 : 0.0;
        Z[e] = v;
    }
}

__global__ __launch_bounds__(1024) void k_schur_pivot(int n, int c, double* __restrict__ ws,
                                                      int32_t* __restrict__ status) {
    __shared__ double bv[1024];
    __shared__ int bi[1024];
    const int w = n + 1, t = threadIdx.x;
    double* M = ws;
    double* Z = ws + (size_t)n * n;
    double best = -1.0;
    int bidx = c;
    for (int r = c + t; r < n; r += 1024) {
        const double a = fabs(M[(size_t)r * n + c]);
        if (a > best) {
            best = a;
            bidx = r;
        }
    }
    bv[t] = best;
    bi[t] = bidx;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            const double o = bv[t + s];
            const int oi = bi[t + s];
            if (o > bv[t] || (o == bv[t] && oi < bi[t])) {
                bv[t] = o;
                bi[t] = oi;
            }
        }
        __syncthreads();
    }
    const int pr = bi[0];
    if (t == 0 && !(bv[0] > 0.0)) atomicOr(status, 4);
    if (pr != c) {
        for (int j = t; j < n; j += 1024) {
            const double tmp = M[(size_t)c * n + j];
            M[(size_t)c * n + j] = M[(size_t)pr * n + j];
            M[(size_t)pr * n + j] = tmp;
        }
        for (int j = t; j < w; j += 1024) {
            const double tmp = Z[(size_t)c * w + j];
            Z[(size_t)c * w + j] = Z[(size_t)pr * w + j];
            Z[(size_t)pr * w + j] = tmp;
        }
    }
}

// rows r = c + 1 + blockIdx.y; the threads stride over the trailing columns
// of M and all columns of Z
__global__ __launch_bounds__(256) void k_schur_elim(int n, int c, double* __restrict__ ws) {
    const int w = n + 1;
    double* M = ws;
    double* Z = ws + (size_t)n * n;
    const int r = c + 1 + (int)blockIdx.y;
    const double pv = M[(size_t)c * n + c];
    const double f = (pv != 0.0) ? M[(size_t)r * n + c] / pv : 0.0;
    const int wm = n - c - 1;
    for (int j = (int)(blockIdx.x * blockDim.x + threadIdx.x); j < wm + w;
         j += (int)(gridDim.x * blockDim.x)) {
        if (j < wm)
            M[(size_t)r * n + c + 1 + j] -= f * M[(size_t)c * n + c + 1 + j];
        else
            Z[(size_t)r * w + (j - wm)] -= f * Z[(size_t)c * w + (j - wm)];
    }
}

// back substitution: U^T (the upper triangle of the eliminated M) into the
// X area, then one workgroup per right-hand-side column j of Z, bottom up:
// x_i = z_i / U_ii, z_r -= U_ri x_i for r < i (the column in LDS, each row
// owned by one thread); the solution replaces the column of Z.
__global__ void k_schur_ut(int n, double* __restrict__ ws) {
    const double* M = ws;
    double* Ut = ws + (size_t)n * n + (size_t)n * (n + 1);
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)n * n) return;
    const size_t i = q / n, r = q % n;
    if (r <= i) Ut[q] = M[r * n + i];
}

__global__ __launch_bounds__(256) void k_schur_back_cols(int n, double* __restrict__ ws) {
    __shared__ double z[RC_KMAX + 3];
    __shared__ double xs[RC_KMAX + 3];
    const int w = n + 1, t = threadIdx.x, j = blockIdx.x;
    double* Z = ws + (size_t)n * n;
    const double* Ut = Z + (size_t)n * w;
    for (int r = t; r < n; r += 256) z[r] = Z[(size_t)r * w + j];
    __syncthreads();
    for (int i = n - 1; i >= 0; --i) {
        const double d = Ut[(size_t)i * n + i];
        const double x = (d != 0.0) ? z[i] / d : 0.0;
        for (int r = t; r < i; r += 256) z[r] -= Ut[(size_t)i * n + r] * x;
        if (t == 0) xs[i] = x;
        __syncthreads();
    }
    for (int r = t; r < n; r += 256) Z[(size_t)r * w + j] = xs[r];
}

__global__ void k_schur_out(int n, const double* __restrict__ ws, double* __restrict__ sol,
                            double* __restrict__ Minv) {
    const int w = n + 1;
    const double* X = ws + (size_t)n * n;  // the solved columns of Z
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < (size_t)n) sol[q] = X[q * w];
    if (q < (size_t)n * n) {
        const size_t i = q / n, j = q % n;
        Minv[q] = X[i * w + 1 + j];
    }
}

// g~_c[v] = X_c[v] - sum_q lambda_q Y_{q,c}[v] + t_c / V   (stride Vp)
__global__ void k_gtilde(const double* __restrict__ X, const double* __restrict__ sol, int k,
                         int V, int Vp, double* __restrict__ gt) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (v >= Vp) return;
    double s = 0.0;
    if (v < V) {
        s = X[(size_t)c * Vp + v];
        for (int q = 0; q < k; ++q) s -= sol[q] * X[(size_t)(3 + 3 * q + c) * Vp + v];
        s += sol[k + c] / V;
    }
    gt[(size_t)c * Vp + v] = s;
}

// y_c = A x_c for the three coordinate blocks (x, y stride Vp), from the
// strict upper triangle of the FP64 factor storage (which keeps A there:
// the factorisation writes only the lower triangle) and the saved diagonal
// d. One workgroup per 64 x 64 tile (I, J), I <= J, of the upper triangle,
// each tile read once: its row sums are block I's partial from J and its
// column sums (A is symmetric) block J's partial from I; a diagonal tile's
// strict upper part gives both to one partial. part[I][J][c][i] is added up
// by k_symv_reduce in a fixed order.
__device__ __forceinline__ void upper_tile(int p, int& I, int& J) {
    int j = (int)((__builtin_sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= p) ++j;
    while (j * (j + 1) / 2 > p) --j;
    J = j;
    I = p - j * (j + 1) / 2;
}

__global__ __launch_bounds__(256) void k_symv_tiles(const double* __restrict__ F, int V, int Vp,
                                                    int nblk, const double* __restrict__ x,
                                                    double* __restrict__ part) {
    __shared__ double rsh[RC_NB][3];
    __shared__ double csh[4][RC_NB][3];
    int I, J;
    upper_tile(blockIdx.x, I, J);
    const int t = threadIdx.x, w = t >> 6;
    const int ri = t >> 4, cj = t & 15;   // tile rows 4 ri .. 4 ri + 3, columns 4 cj .. 4 cj + 3
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
    // only the V x V part, and on a diagonal tile only above the diagonal
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool on = i0 + r < V && j0 + k < V && (I < J || j0 + k > i0 + r);
            a[r][k] = on ? a[r][k] : 0.0;
        }
    double rs[4][3], cs[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double u = 0.0, v = 0.0;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                u = fma(a[q][m], xj[m][c], u);
                v = fma(a[m][q], xi[m][c], v);
            }
            rs[q][c] = u;
            cs[q][c] = v;
        }
    // row sums over the 16 column groups (lanes 1, 2, 4, 8 apart), column
    // sums over the wave's 4 row groups (lanes 16, 32 apart); a butterfly
    // leaves the same bits in every lane
#pragma unroll
    for (int off = 1; off < 16; off <<= 1)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) rs[q][c] += __shfl_xor(rs[q][c], off, 64);
#pragma unroll
    for (int off = 16; off < 64; off <<= 1)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) cs[q][c] += __shfl_xor(cs[q][c], off, 64);
    if (cj == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) rsh[4 * ri + q][c] = rs[q][c];
    if (((t >> 4) & 3) == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) csh[w][4 * cj + q][c] = cs[q][c];
    __syncthreads();
    if (t < 3 * RC_NB) {
        const int c = t >> 6, i = t & 63;
        const double col = ((csh[0][i][c] + csh[1][i][c]) + csh[2][i][c]) + csh[3][i][c];
        if (I == J) {
            part[(((size_t)I * nblk + I) * 3 + c) * RC_NB + i] = rsh[i][c] + col;
        } else {
            part[(((size_t)I * nblk + J) * 3 + c) * RC_NB + i] = rsh[i][c];
            part[(((size_t)J * nblk + I) * 3 + c) * RC_NB + i] = col;
        }
    }
}

// y_c[v] = d_v x_c[v] + sum over J (in order) of part[I][J][c][v - 64 I].
__global__ __launch_bounds__(64) void k_symv_reduce(const double* __restrict__ part,
                                                    const double* __restrict__ d,
                                                    const double* __restrict__ x, int V, int Vp,
                                                    int nblk, double* __restrict__ y) {
    const int I = blockIdx.x, c = blockIdx.y, i = threadIdx.x, v = I * RC_NB + i;
    const double* p = part + ((size_t)I * nblk * 3 + c) * RC_NB + i;
    double s = (v < V) ? d[v] * x[(size_t)c * Vp + v] : 0.0;
#pragma unroll 8
    for (int J = 0; J < nblk; ++J) s += p[(size_t)J * 3 * RC_NB];
    y[(size_t)c * Vp + v] = s;
}

// ||K z - rhs|| / max(1, ||rhs||) (linsolve.ts:810-849) with z = [g~; lambda],
// rhs = [dE; 0]; Ax = A g~ blocks, Cg = C g~. Written to *out (its own slot:
// it may run beside the line search, which owns the control block).
__global__ __launch_bounds__(CT) void k_residual(const double* __restrict__ Ax,
                                                 const double* __restrict__ C,
                                                 const double* __restrict__ sol,
                                                 const double* __restrict__ dE,
                                                 const double* __restrict__ Cg, int k, int V,
                                                 int Vp, double* __restrict__ out) {
    __shared__ double sh[32];
    double r2 = 0.0, b2 = 0.0;
#pragma unroll 4
    for (int q = threadIdx.x; q < 3 * V; q += CT) {
        const int c = q / V, v = q % V;
        const size_t j = (size_t)c * Vp + v;
        double s = Ax[j];
        for (int p = 0; p < k; ++p) s += C[(size_t)p * 3 * Vp + j] * sol[p];
        const double d = s - dE[j];
        r2 += d * d;
        b2 += dE[j] * dE[j];
    }
    for (int p = threadIdx.x; p < k; p += CT) r2 += Cg[p] * Cg[p];
    r2 = block_sum(r2, sh);
    b2 = block_sum(b2, sh);
    if (threadIdx.x == 0) out[0] = sqrt(r2) / js_max(1.0, sqrt(b2));
}

// ---------------------------------------------------------------------------
// line search

// Grid-wide line-search kernels: up to RC_LSG_NW workgroups of LSG_T
// threads. Workgroups hand their partial sums over as write-through records
// (one per workgroup, RC_LSG_PART doubles), which are reduced in one fixed
// order (lane w of one wave holds workgroup w), so every workgroup that
// reduces them gets the same sums.
constexpr int LSG_T = 256;

// True in the workgroup that arrives last, after every workgroup's
// write-through stores are complete; it resets the counter for the next
// launch. Workgroup-collective.
__device__ bool last_workgroup(uint32_t* cnt) {
    __shared__ int last;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tk =
            __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = tk == gridDim.x - 1;
        if (last) __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return last != 0;
}

// Sum of field f of the records part[w][RC_LSG_PART], w < gridDim.x, in
// wave 0 (fixed order; every lane gets it). max: NaN-propagating maximum.
__device__ __forceinline__ double record_reduce(const double* part, int f, bool max) {
    const int lane = threadIdx.x & 63;
    const double v =
        lane < (int)gridDim.x ? load_wt(part + (size_t)lane * RC_LSG_PART + f) : 0.0;
    return max ? wmax(v) : wsum(v);
}

// Control block after the solve, part 1 (optimizer.ts:331-367,
// lineSearch.ts:459-497): the L2h norm of g~, its finite check and the
// failure reasons that do not need the slope; part 2 (k_ls_setup_dir)
// forms the direction and the slope.
__global__ __launch_bounds__(LSG_T) void k_ls_setup_norm(EdgeSoA ed, int E, int V, int Vp,
                                                         const double* __restrict__ gt,
                                                         const double* __restrict__ e0tpe,
                                                         const double* __restrict__ e0pen,
                                                         const int32_t* __restrict__ status,
                                                         int has_e0, double e0_given,
                                                         double tau0, LSCtl* ctl,
                                                         double* part, uint32_t* cnt) {
    __shared__ double sh[32];
    const int t = threadIdx.x, gid = blockIdx.x * LSG_T + t, gsz = gridDim.x * LSG_T;
    double s = 0.0, nf = 0.0;
    for (int K = gid; K < E; K += gsz) {
        const int a = ed.i1[K], b = ed.i2[K];
        const double l = ed.rlen[K];
        double n1 = 0.0, n2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double x1 = gt[(size_t)c * Vp + a], x2 = gt[(size_t)c * Vp + b];
            n1 += x1 * x1;
            n2 += x2 * x2;
        }
        s += 0.5 * l * (n1 + n2);
    }
    // finite check of g~ (a non-finite solve is the reference's "treating as singular")
    for (int c = 0; c < 3; ++c)
        for (int v = gid; v < V; v += gsz) {
            const double x = gt[(size_t)c * Vp + v];
            if (!(x - x == 0.0)) nf = 1.0;
        }
    s = block_sum(s, sh);
    nf = block_sum(nf, sh);
    if (t == 0) {
        store_wt(part + (size_t)blockIdx.x * RC_LSG_PART, s);
        store_wt(part + (size_t)blockIdx.x * RC_LSG_PART + 1, nf);
    }
    if (!last_workgroup(cnt) || t >= 64) return;
    s = record_reduce(part, 0, false);
    nf = record_reduce(part, 1, false);
    if (t != 0) return;
    const double e0 = has_e0 ? e0_given : (e0tpe[0] + e0pen[0]);
    const double gn = sqrt(s > 0.0 ? s : 0.0);
    const bool singular = (status[0] & (1 | 4)) != 0 || nf > 0.0;
    ctl->e0 = e0;
    ctl->e1 = e0;
    ctl->tau = tau0;
    ctl->gnorm = gn;
    ctl->slope = 0.0;
    ctl->accepted = 0;
    ctl->converged = 0;
    ctl->trials = 0;
    ctl->energy_evals = 0;
    ctl->proj_solves = 0;
    ctl->proj_iters = 0;
    ctl->proj_iters_acc = -1;
    ctl->tau_acc = 0.0;
    ctl->phi_norm = NAN;
    ctl->skip_energy = 1;
    ctl->counter = 0;
    ctl->last_reason = RC_REASON_TAU_BELOW_MIN;
    ctl->reason = RC_REASON_NONE;
    ctl->done = 0;
    if (singular) {
        ctl->done = 1;
        ctl->reason = RC_REASON_SINGULAR_SYSTEM;
        ctl->residual = NAN;
        ctl->gnorm = NAN;
    } else if (gn < 1e-4) {
        // optimizer.ts:355-367 (a NaN norm fails this test and goes on)
        ctl->done = 1;
        ctl->converged = 1;
        ctl->proj_iters_acc = 0;
    } else if (!(gn - gn == 0.0) || gn <= 0.0) {
        ctl->done = 1;
        ctl->reason = RC_REASON_ZERO_OR_NONFINITE_GRADIENT_NORM;
        ctl->proj_iters_acc = 0;
    }
}

// Control block, part 2: direction g~ / |g~|_L2h and slope dE . dir; not a
// descent direction ends the search. Nothing to do once part 1 ended it.
__global__ __launch_bounds__(LSG_T) void k_ls_setup_dir(int V, int Vp,
                                                        const double* __restrict__ gt,
                                                        const double* __restrict__ dE,
                                                        double* __restrict__ dir, LSCtl* ctl,
                                                        double* part, uint32_t* cnt) {
    __shared__ double sh[32];
    if (ctl->done) return;
    const int t = threadIdx.x, gid = blockIdx.x * LSG_T + t, gsz = gridDim.x * LSG_T;
    const double gn = ctl->gnorm;
    double sl = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int v = gid; v < V; v += gsz) {
            const size_t j = (size_t)c * Vp + v;
            const double d = gt[j] / gn;
            dir[j] = d;
            sl += dE[j] * d;
        }
    sl = block_sum(sl, sh);
    if (t == 0) store_wt(part + (size_t)blockIdx.x * RC_LSG_PART, sl);
    if (!last_workgroup(cnt) || t >= 64) return;
    sl = record_reduce(part, 0, false);
    if (t != 0) return;
    ctl->slope = sl;
    if (!(sl - sl == 0.0) || sl <= 0.0) {
        ctl->done = 1;
        ctl->reason = RC_REASON_NOT_A_DESCENT_DIRECTION;
        ctl->proj_iters_acc = 0;
    }
}

// |e| of edge K at positions pos (SoA x[V], y[V], z[V]).
__device__ __forceinline__ double edge_len(const EdgeSoA& ed, const double* __restrict__ pos,
                                           int V, int K) {
    const int a = ed.i1[K], b = ed.i2[K];
    const double ex = pos[b] - pos[a], ey = pos[V + b] - pos[V + a];
    const double ez = pos[2 * V + b] - pos[2 * V + a];
    return __builtin_sqrt(ex * ex + ey * ey + ez * ez);
}

// Row K of the edge table at positions pos, as k_edge_prep forms it;
// returns |e| and the guarded unit tangent in u.
__device__ __forceinline__ double edge_row(EdgeSoA& ed, const double* __restrict__ pos, int V,
                                           int K, double alpha, double eps, double (&u)[3]) {
    const int a = ed.i1[K], b = ed.i2[K];
    const double ax = pos[a], ay = pos[V + a], az = pos[2 * V + a];
    const double bx = pos[b], by = pos[V + b], bz = pos[2 * V + b];
    const double ex = bx - ax, ey = by - ay, ez = bz - az;
    const double r = __builtin_sqrt(ex * ex + ey * ey + ez * ez);
    const double ell = r + eps;
    double ux = 0, uy = 0, uz = 0;
    if (!(r < 1e-14)) {
        const double inv = 1.0 / r;
        ux = inv * ex;
        uy = inv * ey;
        uz = inv * ez;
    }
    ed.ax[K] = ax; ed.ay[K] = ay; ed.az[K] = az;
    ed.bx[K] = bx; ed.by[K] = by; ed.bz[K] = bz;
    ed.ell[K] = ell;
    ed.rlen[K] = r;
    ed.ux[K] = ux; ed.uy[K] = uy; ed.uz[K] = uz;
    ed.lpow[K] = pow(ell, 1.0 - alpha);
    ed.dpc[K] = (1.0 - alpha) * pow(ell, -alpha);
    u[0] = ux;
    u[1] = uy;
    u[2] = uz;
    return r;
}

// Block tests of the projection iterate (lineSearch.ts:221-249) from its
// sums over the edges and vertices (L, the barycenter moment m, the
// per-edge residual norm pe2, the radii Rb and Rpin): stacks Phi in set
// order into phi (a per-edge length block's rows are already there) and
// tests every block's scaled tolerance; returns whether all pass, |Phi| in
// phin. One thread. pin(pn, c): coordinate c of the vertex of pin pn (its
// index checked here).
template <class PinPos>
__device__ bool check_blocks(const LSCtl* ctl, const DevSet* __restrict__ set, double L,
                             double m0, double m1, double m2, double pe2, double Rb,
                             const double* Rpin, PinPos pin, int V, double* phi, double& phin) {
    bool conv = true;
    double tot = 0.0;
    for (int b = 0; b < set->nblocks; ++b) {
        const int kind = set->kind[b], r0 = set->row0[b];
        double bn2 = 0.0, scale = 1.0;
        if (kind == RC_BLOCK_BARYCENTER) {
            phi[r0] = m0; phi[r0 + 1] = m1; phi[r0 + 2] = m2;
            bn2 = m0 * m0 + m1 * m1 + m2 * m2;
            scale = js_max(1.0, L * js_max(1.0, Rb));
        } else if (kind == RC_BLOCK_TOTAL_LENGTH) {
            phi[r0] = set->L0 - L;
            bn2 = phi[r0] * phi[r0];
            scale = js_max(1.0, L);
        } else if (kind == RC_BLOCK_EDGE_LENGTHS) {
            bn2 = pe2;
            scale = js_max(1.0, L);
        } else {
            const int pn = set->pin_of_block[b];
            const int pv = set->pin_vertex[pn];
            const double* tg = set->pin_target[pn];
            if (pv >= 0 && pv < V) {
                phi[r0] = pin(pn, 0) - tg[0];
                phi[r0 + 1] = pin(pn, 1) - tg[1];
                phi[r0 + 2] = pin(pn, 2) - tg[2];
            } else {
                phi[r0] = phi[r0 + 1] = phi[r0 + 2] = NAN;
            }
            bn2 = phi[r0] * phi[r0] + phi[r0 + 1] * phi[r0 + 1] + phi[r0 + 2] * phi[r0 + 2];
            scale = js_max(1.0, Rpin[pn]);
        }
        tot += bn2;
        const double bn = sqrt(bn2);
        if (conv && bn > js_max(ctl->tol_abs, ctl->tol_rel * scale)) conv = false;
    }
    phin = sqrt(tot);
    return conv;
}

// Constraint check of the projection iterate cur (lineSearch.ts:221-249):
// evaluates Phi (stacked in set order into phi) and every block's scaled
// tolerance. Returns, in every thread, whether all blocks are within
// tolerance; phin receives |Phi|. Workgroup-collective.
__device__ __forceinline__ bool project_check(const LSCtl* ctl, const DevSet* __restrict__ set,
                                              const double* __restrict__ ell0,
                                              const double* __restrict__ cur, const EdgeSoA& ed,
                                              int V, int E, const int32_t* __restrict__ vOff,
                                              double* sh, double* phi, double* Rpin, int* flag_s,
                                              double& phin) {
    const int t = threadIdx.x;
    double L = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0, pe2 = 0.0;
#pragma unroll 4
    for (int K = t; K < E; K += CT) {
        const int a = ed.i1[K], b = ed.i2[K];
        const double ax = cur[a], ay = cur[V + a], az = cur[2 * V + a];
        const double bx = cur[b], by = cur[V + b], bz = cur[2 * V + b];
        const double ex = bx - ax, ey = by - ay, ez = bz - az;
        const double l = __builtin_sqrt(ex * ex + ey * ey + ez * ez);
        L += l;
        if (set->has_bary) {
            m0 += l * (0.5 * (ax + bx) - set->x0[0]);
            m1 += l * (0.5 * (ay + by) - set->x0[1]);
            m2 += l * (0.5 * (az + bz) - set->x0[2]);
        }
        if (set->len_per_edge) {
            const double ph = ell0[K] - l;
            phi[set->len_row0 + K] = ph;
            pe2 += ph * ph;
        }
    }
    L = block_sum(L, sh);
    m0 = block_sum(m0, sh);
    m1 = block_sum(m1, sh);
    m2 = block_sum(m2, sh);
    pe2 = block_sum(pe2, sh);
    double Rb = 0.0;
    if (set->has_bary)
#pragma unroll 4
        for (int v = t; v < V; v += CT) {
            if (vOff[v + 1] == vOff[v]) continue;  // endpoints of edges only
            const double dx = cur[v] - set->x0[0], dy = cur[V + v] - set->x0[1],
                         dz = cur[2 * V + v] - set->x0[2];
            Rb = js_max(Rb, __builtin_sqrt(dx * dx + dy * dy + dz * dz));
        }
    Rb = block_max(Rb, sh);
    for (int pn = 0; pn < set->npin; ++pn) {
        double R = 0.0;
        const double* tg = set->pin_target[pn];
#pragma unroll 4
        for (int v = t; v < V; v += CT) {
            const double dx = cur[v] - tg[0], dy = cur[V + v] - tg[1], dz = cur[2 * V + v] - tg[2];
            R = js_max(R, __builtin_sqrt(dx * dx + dy * dy + dz * dz));
        }
        R = block_max(R, sh);
        if (t == 0) Rpin[pn] = R;
    }
    if (t == 0) {
        double pn2 = 0.0;
        auto pin = [&](int pn, int c) { return cur[(size_t)c * V + set->pin_vertex[pn]]; };
        *flag_s = check_blocks(ctl, set, L, m0, m1, m2, pe2, Rb, Rpin, pin, V, phi, pn2) ? 1 : 0;
        sh[31] = pn2;
    }
    __syncthreads();
    phin = sh[31];
    const bool conv = *flag_s != 0;
    __syncthreads();
    return conv;
}

// Projection correction (lineSearch.ts:250-288) from the factored saddle:
// [lambda; t] = M^{-1} [Phi; 0], x = -Y lambda + N t / V, cur += x.
// Returns, in every thread, whether a component of x is not finite.
// Workgroup-collective.
__device__ __forceinline__ bool project_correct(int k, const double* __restrict__ Y,
                                                const double* __restrict__ Minv, int V, int Vp,
                                                double* __restrict__ cur, const double* phi,
                                                double* solv, double* sh) {
    const int t = threadIdx.x;
    const int n = k + 3;
    if (k > RC_KSMALL) {
        // one wave per row of M^{-1}, its lanes over the columns (coalesced)
        const int lane = t & 63;
        for (int i = t >> 6; i < n; i += CT / 64) {
            double s = 0.0;
            for (int j = lane; j < k; j += 64) s += Minv[(size_t)i * n + j] * phi[j];
            s = wsum(s);
            if (lane == 0) solv[i] = s;
        }
    } else {
        for (int i = t; i < n; i += CT) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += Minv[i * n + j] * phi[j];
            solv[i] = s;
        }
    }
    __syncthreads();
    double nf = 0.0;
    if (k > RC_KSMALL) {
        // a vertex's three components per thread, four partial sums over
        // the rows each: twelve independent loads in flight per thread
#pragma unroll 4
        for (int v = t; v < V; v += CT) {
            double s[3][4] = {};
            int p = 0;
            for (; p + 4 <= k; p += 4)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double lam = solv[p + u];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        s[c][u] += lam * Y[(size_t)(3 * (p + u) + c) * Vp + v];
                }
            for (; p < k; ++p)
#pragma unroll
                for (int c = 0; c < 3; ++c) s[c][0] += solv[p] * Y[(size_t)(3 * p + c) * Vp + v];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double x = solv[k + c] / V - ((s[c][0] + s[c][1]) + (s[c][2] + s[c][3]));
                if (!(x - x == 0.0)) nf = 1.0;
                cur[(size_t)c * V + v] += x;
            }
        }
        nf = block_sum(nf, sh);
        __syncthreads();
        return nf > 0.0;
    }
    for (int c = 0; c < 3; ++c) {
#pragma unroll 4
        for (int v = t; v < V; v += CT) {
            double x = solv[k + c] / V;
            for (int p = 0; p < k; ++p) x -= solv[p] * Y[(size_t)(3 * p + c) * Vp + v];
            if (!(x - x == 0.0)) nf = 1.0;
            cur[(size_t)c * V + v] += x;
        }
    }
    nf = block_sum(nf, sh);
    __syncthreads();
    return nf > 0.0;
}

// Candidate of the current trial: cur = gamma0 - tau dir (lineSearch.ts:501-503).
__device__ __forceinline__ void trial_candidate(const double* __restrict__ pos0,
                                                const double* __restrict__ dir, double tau, int V,
                                                int Vp, double* __restrict__ cur) {
    for (int c = 0; c < 3; ++c) {
#pragma unroll 4
        for (int v = threadIdx.x; v < V; v += CT)
            cur[(size_t)c * V + v] = pos0[(size_t)c * V + v] - tau * dir[(size_t)c * Vp + v];
    }
    __syncthreads();
}

// End of a trial's projection: on success (and a finite candidate,
// lineSearch.ts:525-527) the trial goes on to its energy (k_ls_edges forms
// the edge table and the penalty energy at cur over the device, then
// k_energy_ls); otherwise it fails with 'projection_failed' and tau
// shrinks. Workgroup-collective.
__device__ __forceinline__ void project_finish(LSCtl* ctl, int ok, int it, double phin,
                                               const double* __restrict__ cur, int V,
                                               double* sh) {
    const int t = threadIdx.x;
    double nfc = 0.0;
#pragma unroll 4
    for (int q = t; q < 3 * V; q += CT)
        if (!(cur[q] - cur[q] == 0.0)) nfc = 1.0;
    nfc = block_sum(nfc, sh);
    if (ok && nfc == 0.0) {
        if (t == 0) {
            ctl->proj_iters = it;
            ctl->phi_norm = phin;
            ctl->proj_ok = 1;
            ctl->skip_energy = 0;
        }
    } else if (t == 0) {
        ctl->proj_ok = 0;
        ctl->skip_energy = 1;
        ctl->trials += 1;
        ctl->last_reason = RC_REASON_PROJECTION_FAILED;
        ctl->tau = ctl->tau * ctl->shrink;
        if (!(ctl->tau >= ctl->tau_min)) {
            ctl->done = 1;
            ctl->reason = ctl->last_reason;
        }
    }
}

constexpr int WVL = 16;   // vertex lanes of the grid projection's wide correction

// Grid barrier of the launch's gridDim.x workgroups (all co-resident: at
// most RC_LSG_NW): each workgroup's write-through stores
// complete, then one thread adds 1 and waits for target. Workgroup-collective.
__device__ __forceinline__ void grid_barrier(uint32_t* bar, uint32_t target, int32_t* status) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1 << 24)) {
                atomicOr(status, 2);
                break;
            }
        }
    }
    __syncthreads();
}

// One trial's candidate and frozen projection (lineSearch.ts:196-295,
// 501-548) over the grid. Each check: every workgroup
// sums its edges and vertices (cur read write-through) into its record
// (L, m0..m2, pe2, Rb, non-finite cur, non-finite previous correction,
// Rpin) and the per-edge rows of Phi into phig; after a grid barrier every
// workgroup reduces the records in the same order and runs the same block
// tests, so all take the same branch; the correction [lambda; t] =
// M^-1 [Phi; 0], cur += -Y lambda + N t / V is applied by each workgroup to
// its own vertices, and a grid barrier precedes the next check (beyond
// RC_KSMALL rows, [lambda; t] is formed a wave per row over the grid and
// handed over through solvg behind one more barrier). A
// non-finite correction (lineSearch.ts:283-288) is seen in the next check's
// records and fails the trial, as the reference's break does. Counters:
// gcnt[0] the barrier, gcnt[1] the exits (the last workgroup to exit resets
// both).
template <int T>
__global__ __launch_bounds__(T) void k_ls_project_grid(
    LSCtl* ctl, const DevSet* __restrict__ set, const double* __restrict__ ell0,
    const double* __restrict__ pos0, const double* __restrict__ dir, double* cur, EdgeSoA ed,
    int V, int E, int Vp, const double* __restrict__ Y, const double* __restrict__ Minv,
    const int32_t* __restrict__ vOff, double* part, double* phig, double* solvg, uint32_t* gcnt,
    int32_t* status) {
    __shared__ double sh[32];
    __shared__ double phi[RC_KMAX + 8];
    __shared__ double solv[RC_KMAX + 8];
    __shared__ double rec[RC_LSG_PART];
    __shared__ double pinx[RC_MAXPIN][3];
    __shared__ double wred[3][T / WVL][WVL];   // correction partial sums (k > RC_KSMALL)
    __shared__ int conv_s;
    __shared__ double phin_s;
    if (ctl->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->skip_energy = 1;
        return;
    }
    const int t = threadIdx.x, nwg = (int)gridDim.x;
    const int gid = blockIdx.x * T + t, gsz = nwg * T;
    const int k = set->k, n = k + 3, npin = set->npin;
    const int maxit = ctl->max_iter;
    double* myrec = part + (size_t)blockIdx.x * RC_LSG_PART;
    // candidate cur = gamma0 - tau dir (lineSearch.ts:501-503)
    {
        const double tau = ctl->tau;
        for (int c = 0; c < 3; ++c)
            for (int v = gid; v < V; v += gsz)
                store_wt(cur + (size_t)c * V + v,
                         pos0[(size_t)c * V + v] - tau * dir[(size_t)c * Vp + v]);
    }
    uint32_t gen = 0;
    double nf_corr = 0.0;
    int ok = 0, it = 0, nsolve = 0;
    double phin = INFINITY, nfc = 0.0;
    for (it = 0; it <= maxit; ++it) {
        grid_barrier(gcnt, (++gen) * (uint32_t)nwg, status);   // cur is complete
        double L = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0, pe2 = 0.0, Rb = 0.0, nfw = 0.0;
        for (int K = gid; K < E; K += gsz) {
            const int a = ed.i1[K], b = ed.i2[K];
            const double ax = load_wt(cur + a), ay = load_wt(cur + V + a),
                         az = load_wt(cur + 2 * V + a);
            const double bx = load_wt(cur + b), by = load_wt(cur + V + b),
                         bz = load_wt(cur + 2 * V + b);
            const double ex = bx - ax, ey = by - ay, ez = bz - az;
            const double l = __builtin_sqrt(ex * ex + ey * ey + ez * ez);
            L += l;
            if (set->has_bary) {
                m0 += l * (0.5 * (ax + bx) - set->x0[0]);
                m1 += l * (0.5 * (ay + by) - set->x0[1]);
                m2 += l * (0.5 * (az + bz) - set->x0[2]);
            }
            if (set->len_per_edge) {
                const double ph = ell0[K] - l;
                store_wt(phig + set->len_row0 + K, ph);
                pe2 += ph * ph;
            }
        }
        for (int v = gid; v < V; v += gsz) {
            const double x = load_wt(cur + v), y = load_wt(cur + V + v),
                         z = load_wt(cur + 2 * V + v);
            if (!(x - x == 0.0) || !(y - y == 0.0) || !(z - z == 0.0)) nfw = 1.0;
            if (set->has_bary && vOff[v + 1] != vOff[v]) {
                const double dx = x - set->x0[0], dy = y - set->x0[1], dz = z - set->x0[2];
                Rb = js_max(Rb, __builtin_sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
        L = block_sum(L, sh);
        m0 = block_sum(m0, sh);
        m1 = block_sum(m1, sh);
        m2 = block_sum(m2, sh);
        pe2 = block_sum(pe2, sh);
        nfw = block_sum(nfw, sh);
        Rb = block_max(Rb, sh);
        if (t == 0) {
            store_wt(myrec + 0, L);
            store_wt(myrec + 1, m0);
            store_wt(myrec + 2, m1);
            store_wt(myrec + 3, m2);
            store_wt(myrec + 4, pe2);
            store_wt(myrec + 5, Rb);
            store_wt(myrec + 6, nfw);
            store_wt(myrec + 7, nf_corr);
        }
        for (int pn = 0; pn < npin; ++pn) {
            double R = 0.0;
            const double* tg = set->pin_target[pn];
            for (int v = gid; v < V; v += gsz) {
                const double dx = load_wt(cur + v) - tg[0], dy = load_wt(cur + V + v) - tg[1],
                             dz = load_wt(cur + 2 * V + v) - tg[2];
                R = js_max(R, __builtin_sqrt(dx * dx + dy * dy + dz * dz));
            }
            R = block_max(R, sh);
            if (t == 0) store_wt(myrec + 8 + pn, R);
        }
        // the pins' positions, read before any workgroup moves on to its
        // correction
        for (int q = t; q < 3 * npin; q += T) {
            const int pn = q / 3, c = q - 3 * pn, pv = set->pin_vertex[pn];
            pinx[pn][c] = (pv >= 0 && pv < V) ? load_wt(cur + (size_t)c * V + pv) : 0.0;
        }
        grid_barrier(gcnt, (++gen) * (uint32_t)nwg, status);   // the records are complete
        if (t < 64) {
            for (int f = 0; f < 8 + npin; ++f) {
                const double r = record_reduce(part, f, f == 5 || f >= 8);
                if (t == 0) rec[f] = r;
            }
        }
        if (set->len_per_edge)
            for (int K = t; K < E; K += T)
                phi[set->len_row0 + K] = load_wt(phig + set->len_row0 + K);
        __syncthreads();
        if (t == 0) {
            double pn2 = 0.0;
            auto pin = [&](int pn, int c) { return pinx[pn][c]; };
            conv_s = check_blocks(ctl, set, rec[0], rec[1], rec[2], rec[3], rec[4], rec[5],
                                  rec + 8, pin, V, phi, pn2) ? 1 : 0;
            phin_s = pn2;
        }
        __syncthreads();
        nfc = rec[6];
        if (rec[7] > 0.0) break;   // the last correction was not finite: ok = false
        phin = phin_s;
        if (conv_s) {
            ok = 1;
            break;
        }
        if (it == maxit) break;
        // frozen correction with Y and M^-1 of gamma0
        ++nsolve;
        if (k <= RC_KSMALL) {
            for (int i = t; i < n; i += T) {
                double sv = 0.0;
                for (int j = 0; j < k; ++j) sv += Minv[i * n + j] * phi[j];
                solv[i] = sv;
            }
        } else {
            // a wave per row of M^-1, its lanes over the columns (coalesced)
            const int lane = t & 63, nwv = nwg * (T / 64);
            for (int i = blockIdx.x * (T / 64) + (t >> 6); i < n; i += nwv) {
                double sv = 0.0;
                for (int j = lane; j < k; j += 64) sv += Minv[(size_t)i * n + j] * phi[j];
                sv = wsum(sv);
                if (lane == 0) store_wt(solvg + i, sv);
            }
            grid_barrier(gcnt, (++gen) * (uint32_t)nwg, status);   // [lambda; t] is complete
            for (int i = t; i < n; i += T) solv[i] = load_wt(solvg + i);
        }
        __syncthreads();
        double nf = 0.0;
        if (k <= RC_KSMALL) {
            for (int c = 0; c < 3; ++c)
                for (int v = gid; v < V; v += gsz) {
                    double x = solv[k + c] / V;
                    for (int p = 0; p < k; ++p) x -= solv[p] * Y[(size_t)(3 * p + c) * Vp + v];
                    if (!(x - x == 0.0)) nf = 1.0;
                    double* q = cur + (size_t)c * V + v;
                    store_wt(q, load_wt(q) + x);
                }
        } else {
            // this workgroup's vertices, VL at a time; lanes (vertex vl, row
            // slice pl) sum rows p = pl mod PL, combined in LDS in that order
            const int per = (V + nwg - 1) / nwg;
            const int vb = blockIdx.x * per, ve = min(V, vb + per);
            const int vl = t % WVL, pl = t / WVL;
            for (int v0 = vb; v0 < ve; v0 += WVL) {
                const int v = v0 + vl;
                double sc[3] = {0.0, 0.0, 0.0};
                if (v < ve) {
#pragma unroll 4
                    for (int p = pl; p < k; p += T / WVL) {
                        const double lam = solv[p];
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            sc[c] += lam * Y[(size_t)(3 * p + c) * Vp + v];
                    }
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) wred[c][pl][vl] = sc[c];
                __syncthreads();
                if (t < 3 * WVL && v0 + t % WVL < ve) {
                    const int c = t / WVL, vv = v0 + t % WVL;
                    double sum = 0.0;
                    for (int q = 0; q < T / WVL; ++q) sum += wred[c][q][t % WVL];
                    const double x = solv[k + c] / V - sum;
                    if (!(x - x == 0.0)) nf = 1.0;
                    double* q = cur + (size_t)c * V + vv;
                    store_wt(q, load_wt(q) + x);
                }
                __syncthreads();
            }
        }
        nf_corr = block_sum(nf, sh);
    }
    if (it > maxit) it = maxit;
    // end of the trial (project_finish): workgroup 0 writes the control block
    if (blockIdx.x == 0 && t == 0) {
        ctl->proj_solves += nsolve;
        if (ok && nfc == 0.0) {
            ctl->proj_iters = it;
            ctl->phi_norm = phin;
            ctl->proj_ok = 1;
            ctl->skip_energy = 0;
        } else {
            ctl->proj_ok = 0;
            ctl->skip_energy = 1;
            ctl->trials += 1;
            ctl->last_reason = RC_REASON_PROJECTION_FAILED;
            ctl->tau = ctl->tau * ctl->shrink;
            if (!(ctl->tau >= ctl->tau_min)) {
                ctl->done = 1;
                ctl->reason = ctl->last_reason;
            }
        }
    }
    // every workgroup is past its last barrier once it gets here: the last
    // one to arrive resets the counters for the next launch
    if (t == 0) {
        const uint32_t tk =
            __hip_atomic_fetch_add(gcnt + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tk == (uint32_t)nwg - 1) {
            __hip_atomic_store(gcnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(gcnt + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Reassemble-mode projection (lineSearch.ts:263-276): every correction uses
// A and C rebuilt and refactored at the current iterate, which the host does
// between launches. phase 0 starts the trial (candidate, first check);
// phase 1 applies the correction with the Y and M^{-1} just factored at cur
// (a singular system there fails the trial, as the reference's solve throws)
// and checks again. Not converged with corrections left: need_solve = 1 and
// Phi is kept in phi_g for the next launch. Single workgroup.
__global__ __launch_bounds__(CT) void k_ls_project_re(
    int phase, LSCtl* ctl, const DevSet* __restrict__ set, const double* __restrict__ ell0,
    DevPen pen, const double* __restrict__ pos0, const double* __restrict__ dir,
    double* __restrict__ cur, EdgeSoA ed, int V, int E, int Vp, const double* __restrict__ Y,
    const double* __restrict__ Minv, const int32_t* __restrict__ vOff,
    const int32_t* __restrict__ vEdge, double alpha, double eps, double* __restrict__ phi_g,
    const int32_t* __restrict__ status) {
    __shared__ double sh[32];
    __shared__ double phi[RC_KMAX + 8];
    __shared__ double solv[RC_KMAX + 8];
    __shared__ double Rpin[RC_MAXPIN];
    __shared__ int flag_s;
    const int t = threadIdx.x;
    if (ctl->done) {
        if (t == 0) {
            ctl->skip_energy = 1;
            ctl->need_solve = 0;
        }
        return;
    }
    const int k = set->k;
    const int maxit = ctl->max_iter;
    int it = 0;
    if (phase == 0) {
        trial_candidate(pos0, dir, ctl->tau, V, Vp, cur);
    } else {
        it = ctl->proj_it;
        bool bad = (status[0] & ~2) != 0;  // breakdown or zero Schur pivot at cur
        if (!bad) {
            // a solve that completed (the reference's throws on a singular system)
            if (t == 0) ctl->proj_solves += 1;
            for (int j = t; j < k; j += CT) phi[j] = phi_g[j];
            __syncthreads();
            bad = project_correct(k, Y, Minv, V, Vp, cur, phi, solv, sh);
        }
        if (bad) {
            if (t == 0) ctl->need_solve = 0;
            project_finish(ctl, 0, it, INFINITY, cur, V, sh);
            return;
        }
        ++it;
    }
    double phin = INFINITY;
    const bool conv = project_check(ctl, set, ell0, cur, ed, V, E, vOff, sh, phi, Rpin, &flag_s,
                                    phin);
    if (conv || it == maxit) {
        if (t == 0) ctl->need_solve = 0;
        project_finish(ctl, conv ? 1 : 0, it, phin, cur, V, sh);
        return;
    }
    for (int j = t; j < k; j += CT) phi_g[j] = phi[j];
    if (t == 0) {
        ctl->proj_it = it;
        ctl->need_solve = 1;
        ctl->skip_energy = 1;
    }
}


// Edge table of the projected trial (the energy kernel's input) and its
// penalty energy (penalties.ts:139-190), over the device: a thread per edge
// (table row, length and field terms) and per vertex (length-difference
// term, its two edge lengths recomputed from the positions). Per-workgroup
// partials; the last workgroup to finish (integer ticket, no float atomics)
// sums them in workgroup order into ctl->penE. Skipped when the trial's
// projection failed or the search is done.
__global__ __launch_bounds__(256) void k_ls_edges(LSCtl* ctl, const double* __restrict__ cur,
                                                  EdgeSoA ed, int V, int E, DevPen pen,
                                                  const int32_t* __restrict__ vOff,
                                                  const int32_t* __restrict__ vEdge,
                                                  double alpha, double eps,
                                                  double* __restrict__ part) {
    __shared__ double sh[32];
    __shared__ int last;
    if (ctl->skip_energy) return;
    const int t = threadIdx.x;
    const int stride = (int)gridDim.x * 256;
    double sL = 0.0, sF = 0.0, sD = 0.0;
    for (int K = (int)blockIdx.x * 256 + t; K < E; K += stride) {
        double u[3];
        const double l = edge_row(ed, cur, V, K, alpha, eps, u);
        sL += l;
        if (pen.wF != 0.0) {
            const double cx = u[1] * pen.X[2] - u[2] * pen.X[1];
            const double cy = u[2] * pen.X[0] - u[0] * pen.X[2];
            const double cz = u[0] * pen.X[1] - u[1] * pen.X[0];
            sF += l * (cx * cx + cy * cy + cz * cz);
        }
    }
    if (!pen.active) {
        if (blockIdx.x == 0 && t == 0) ctl->penE = 0.0;
        return;
    }
    if (pen.wD != 0.0)
        for (int w = (int)blockIdx.x * 256 + t; w < V; w += stride)
            if (vOff[w + 1] - vOff[w] == 2) {
                const double d = edge_len(ed, cur, V, vEdge[vOff[w]]) -
                                 edge_len(ed, cur, V, vEdge[vOff[w] + 1]);
                sD += d * d;
            }
    sL = block_sum(sL, sh);
    sF = block_sum(sF, sh);
    sD = block_sum(sD, sh);
    if (t == 0) {
        part[3 * blockIdx.x] = sL;
        part[3 * blockIdx.x + 1] = sF;
        part[3 * blockIdx.x + 2] = sD;
        __threadfence();
        last = atomicAdd(&ctl->counter, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last || t != 0) return;
    __threadfence();
    double L = 0.0, F = 0.0, D = 0.0;
    for (int i = 0; i < (int)gridDim.x; ++i) {
        L += __hip_atomic_load(part + 3 * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        F += __hip_atomic_load(part + 3 * i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        D += __hip_atomic_load(part + 3 * i + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    double tot = 0.0;
    if (pen.wL != 0.0) tot += pen.wL * L;
    if (pen.wD != 0.0) tot += pen.wD * D;
    if (pen.wF != 0.0) tot += pen.wF * F;
    ctl->penE = tot;
    ctl->counter = 0;
}

// Energy of the projected trial (edge table) with the Armijo decision in the
// last workgroup to finish (integer ticket; no float atomics).
template <class P>
__global__ __launch_bounds__(256) void k_energy_ls(EdgeSoA ed, int E, int npairs, P pw,
                                                   double eps, double* __restrict__ part,
                                                   LSCtl* ctl);

// Step results into one staging block the host reads with a single copy:
// doubles 0-7 and ints at double 8 (rc_step_stats' order), the device status
// words at double 12, the search's done flag at double 14, the refinement
// diagnostics at doubles 16-18.
__global__ __launch_bounds__(CT) void k_ls_finalize(LSCtl* ctl, double* __restrict__ pos,
                                                    const double* __restrict__ cur, int V,
                                                    int update, const double* __restrict__ resid,
                                                    const int32_t* __restrict__ status,
                                                    const double* __restrict__ refine,
                                                    double* __restrict__ stats) {
    const int accepted = ctl->accepted;
    if (update && accepted)
        for (int q = threadIdx.x; q < 3 * V; q += CT) pos[q] = cur[q];
    if (threadIdx.x == 0) {
        // layout mirrors rc_step_stats' leading doubles then ints
        stats[0] = accepted ? ctl->e1 : ctl->e0;
        stats[1] = ctl->e0;
        stats[2] = accepted ? ctl->tau_acc : 0.0;
        // NaN for a singular system (set by the line-search setup)
        stats[3] = (ctl->residual == ctl->residual) ? resid[0] : ctl->residual;
        stats[4] = ctl->gnorm;
        stats[5] = ctl->slope;
        stats[6] = accepted ? ctl->phi_norm : NAN;
        stats[7] = ctl->refine_res;
        int32_t* is = (int32_t*)(stats + 8);
        is[0] = accepted;
        is[1] = ctl->converged;
        is[2] = accepted ? ctl->proj_iters_acc : (ctl->converged ? 0 : ctl->proj_iters_acc);
        is[3] = ctl->reason;
        is[4] = ctl->trials;
        is[5] = ctl->refine_iters;
        is[6] = ctl->energy_evals;
        is[7] = ctl->proj_solves;
        int32_t* st = (int32_t*)(stats + 12);
        for (int q = 0; q < 4; ++q) st[q] = status[q];
        *(int32_t*)(stats + 14) = ctl->done;
        for (int q = 0; q < 3; ++q) stats[16 + q] = refine[q];
    }
}

}  // namespace

// The energy kernel with Armijo decision (pair arithmetic shared with the
// forward energy through rc_pair_energy.h).
namespace {

__device__ __forceinline__ void tri_decode_e(int p, int& a, int& b) {
    int bb = (int)((__builtin_sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((bb + 1) * (bb + 2) / 2 <= p) ++bb;
    while (bb * (bb + 1) / 2 > p) --bb;
    b = bb;
    a = p - bb * (bb + 1) / 2;
}


template <class P>
__global__ __launch_bounds__(256) void k_energy_ls(EdgeSoA ed, int E, int npairs, P pw,
                                                   double eps, double* __restrict__ part,
                                                   LSCtl* ctl) {
    __shared__ double wsums[4];
    __shared__ int last;
    __shared__ rcpair::ETile jt[4];
    if (ctl->skip_energy) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double sum = 0.0;
    for (int base = blockIdx.x * 4; base < npairs; base += gridDim.x * 4) {
        const int p = __builtin_amdgcn_readfirstlane(base + wave);
        if (p >= npairs) break;
        int a, b;
        tri_decode_e(p, a, b);
        sum += rcpair::tile_pair_energy(ed, E, pw, eps, a, b, lane, jt[wave]);
    }
    sum = wsum(sum);
    if (lane == 0) wsums[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ((wsums[0] + wsums[1]) + wsums[2]) + wsums[3];
        __threadfence();
        const int ticket = atomicAdd(&ctl->counter, 1);
        last = (ticket == (int)gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    __shared__ double sh[32];
    double s = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += 256)
        s += __hip_atomic_load(part + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        const double e1 = 0.5 * s + ctl->penE;
        ctl->trials += 1;
        ctl->energy_evals += 1;
        ctl->counter = 0;
        const double tau = ctl->tau;
        if (e1 - e1 == 0.0 && e1 <= ctl->e0 - ctl->c1 * tau * ctl->slope) {
            ctl->accepted = 1;
            ctl->done = 1;
            ctl->tau_acc = tau;
            ctl->e1 = e1;
            ctl->proj_iters_acc = ctl->proj_iters;
            ctl->reason = RC_REASON_NONE;
        } else {
            ctl->last_reason = RC_REASON_ARMIJO_FAILED;
            ctl->tau = tau * ctl->shrink;
            if (!(ctl->tau >= ctl->tau_min)) {
                ctl->done = 1;
                ctl->reason = ctl->last_reason;
            }
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host launchers

rc_status rc_launch_constraint_jacobian(rc_ctx* ctx, const DevSet* d_set, int k) {
    const int Vp = ctx->nblk * RC_NB;
    RC_HIP(hipMemsetAsync(ctx->d_C, 0, sizeof(double) * (size_t)k * 3 * Vp, ctx->stream));
    hipLaunchKernelGGL(k_constraint_jacobian, dim3((ctx->V + 255) / 256), dim3(256), 0,
                       ctx->stream, ctx->ed, ctx->V, Vp, ctx->d_vOff, ctx->d_vEdge, ctx->d_vSlot,
                       d_set, ctx->d_C);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_penalty_grad(rc_ctx* ctx, const DevPen& pen, const double* gsrc, int Vsrc,
                                 double* gdst, int Vdst) {
    hipLaunchKernelGGL(k_penalty_grad_add, dim3((ctx->V + 255) / 256), dim3(256), 0, ctx->stream,
                       ctx->ed, ctx->V, Vsrc, Vdst, ctx->d_vOff, ctx->d_vEdge, ctx->d_vSlot, pen,
                       gsrc, gdst);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_penalty_energy(rc_ctx* ctx, const DevPen& pen, double* d_out) {
    hipLaunchKernelGGL(k_penalty_energy, dim3(1), dim3(CT), 0, ctx->stream, ctx->ed, ctx->E,
                       ctx->V, ctx->d_vOff, ctx->d_vEdge, pen, d_out);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

// B = [dE, C^T blocks] as ncols >= 3 + 3k columns of length Vp (column
// stride Vp); columns past 3 + 3k are zero.
rc_status rc_launch_build_rhs(rc_ctx* ctx, const double* dE, int k, double* B, int ncols) {
    const int Vp = ctx->nblk * RC_NB;
    hipLaunchKernelGGL(k_build_rhs, dim3((Vp + 255) / 256, ncols), dim3(256), 0, ctx->stream,
                       dE, ctx->d_C, ctx->V, Vp, k, B);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

// Schur phase after X = A_hat^{-1} B: dots, block sums, small system, g~.
rc_status rc_launch_schur(rc_ctx* ctx, const double* X, const double* dE, int k, double* gt,
                          double* d_dots, double* d_bsum, double* d_sol, double* d_Minv) {
    const int Vp = ctx->nblk * RC_NB;
    const size_t len = 3 * (size_t)Vp;
    if (k > 0) {
        // dots[p][0] = <C_p, X_r>, dots[p][1+q] = <C_p, Y_q>; X columns are
        // consecutive Vp blocks, so X_r and Y_q are contiguous 3Vp vectors.
        if (k > RC_KSMALL) {
            const unsigned tiles = (unsigned)(((k + 7) / 8) * ((k + 1 + 7) / 8));
            hipLaunchKernelGGL(k_dots_tile, dim3(tiles), dim3(256), 0, ctx->stream,
                               (const double*)ctx->d_C, len, k, X, len, k + 1, len, d_dots);
        } else {
            hipLaunchKernelGGL(k_dots, dim3(k * (k + 1)), dim3(RED_T), 0, ctx->stream, ctx->d_C, len,
                               X, len, k + 1, len, d_dots);
        }
        RC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_block_sums, dim3(3 * (k + 1)), dim3(RED_T), 0, ctx->stream, ctx->d_C, dE, k,
                       ctx->V, Vp, d_bsum);
    RC_CHECK_LAUNCH();
    const int n = k + 3;
    const size_t shm = sizeof(double) * ((size_t)n * n + (size_t)n * (n + 1)) + sizeof(int) * n;
    if (k > RC_KSMALL) {
        const size_t nfill = (size_t)n * n + (size_t)n * (n + 1);
        hipLaunchKernelGGL(k_schur_fill, dim3((unsigned)((nfill + 255) / 256)), dim3(256), 0,
                           ctx->stream, d_dots, d_bsum, k, ctx->V, ctx->d_schur);
        RC_CHECK_LAUNCH();
        for (int c = 0; c < n; ++c) {
            hipLaunchKernelGGL(k_schur_pivot, dim3(1), dim3(1024), 0, ctx->stream, n, c,
                               ctx->d_schur, ctx->d_status);
            if (c + 1 < n) {
                const int cols = (n - c - 1) + (n + 1);
                hipLaunchKernelGGL(k_schur_elim, dim3((cols + 255) / 256, n - c - 1), dim3(256), 0,
                                   ctx->stream, n, c, ctx->d_schur);
            }
        }
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_schur_ut, dim3((unsigned)(((size_t)n * n + 255) / 256)), dim3(256), 0,
                           ctx->stream, n, ctx->d_schur);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_schur_back_cols, dim3(n + 1), dim3(256), 0, ctx->stream, n,
                           ctx->d_schur);
        RC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_schur_out, dim3((unsigned)(((size_t)n * n + 255) / 256)), dim3(256), 0,
                           ctx->stream, n, (const double*)ctx->d_schur, d_sol, d_Minv);
        RC_CHECK_LAUNCH();
    } else if (k > 0) {
        hipLaunchKernelGGL(k_schur, dim3(1), dim3(256), shm, ctx->stream, d_dots, d_bsum, k, ctx->V,
                           d_sol, d_Minv, ctx->d_status);
        RC_CHECK_LAUNCH();
    } else {
        RC_HIP(hipMemsetAsync(d_sol, 0, sizeof(double) * 3, ctx->stream));
    }
    hipLaunchKernelGGL(k_gtilde, dim3((Vp + 255) / 256, 3), dim3(256), 0, ctx->stream, X, d_sol, k,
                       ctx->V, Vp, gt);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_residual(rc_ctx* ctx, hipStream_t st, const double* gt, const double* dE,
                             int k, const double* d_sol, double* d_Ax, double* d_Cg,
                             double* d_out) {
    const int Vp = ctx->nblk * RC_NB;
    const int nb = ctx->nblk;
    hipLaunchKernelGGL(k_symv_tiles, dim3(nb * (nb + 1) / 2), dim3(256), 0, st,
                       (const double*)ctx->d_F, ctx->V, Vp, nb, gt, ctx->d_symv);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_symv_reduce, dim3(nb, 3), dim3(RC_NB), 0, st,
                       (const double*)ctx->d_symv, (const double*)ctx->d_Adiag, gt, ctx->V, Vp, nb,
                       d_Ax);
    RC_CHECK_LAUNCH();
    if (k > 0) {
        hipLaunchKernelGGL(k_dots, dim3(k), dim3(RED_T), 0, st, ctx->d_C, 3 * (size_t)Vp, gt,
                           (size_t)0, 1, 3 * (size_t)Vp, d_Cg);
        RC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_residual, dim3(1), dim3(CT), 0, st, d_Ax, ctx->d_C, d_sol, dE, d_Cg, k,
                       ctx->V, Vp, d_out);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

// workgroups of the grid-wide line-search kernels (co-resident: <= RC_LSG_NW)
static int lsg_grid(const rc_ctx* ctx) {
    const int n = std::max(ctx->E, ctx->V);
    return std::max(1, std::min(RC_LSG_NW, (n + LSG_T - 1) / LSG_T));
}

rc_status rc_launch_ls_setup(rc_ctx* ctx, const double* gt, const double* dE, double* dir,
                             const double* e0tpe, const double* e0pen, int has_e0,
                             double e0_given, double tau0, LSCtl* d_ctl) {
    const int Vp = ctx->nblk * RC_NB;
    const dim3 grid(lsg_grid(ctx));
    hipLaunchKernelGGL(k_ls_setup_norm, grid, dim3(LSG_T), 0, ctx->stream, ctx->ed, ctx->E,
                       ctx->V, Vp, gt, e0tpe, e0pen, (const int32_t*)ctx->d_status, has_e0,
                       e0_given, tau0, d_ctl, ctx->d_lsp, ctx->d_gsync + 2);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ls_setup_dir, grid, dim3(LSG_T), 0, ctx->stream, ctx->V, Vp, gt, dE, dir,
                       d_ctl, ctx->d_lsp, ctx->d_gsync + 3);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_ls_energy(rc_ctx* ctx, LSCtl* d_ctl, const rc_params* p, const DevPen& pen,
                              const double* cur, double* d_epart_ls);

// One trial's candidate and frozen projection over the grid (beyond
// RC_KSMALL rows: RC_LSG_NW workgroups of 1024 threads, for the rows of
// M^-1 Phi and the O(Vk) correction);
// rc_launch_ls_energy follows. solvg (k + 3 doubles): the Schur workspace,
// idle once M^-1 is formed.
rc_status rc_launch_ls_project(rc_ctx* ctx, int k, LSCtl* d_ctl, const DevSet* d_set,
                               const double* pos0, const double* dir, double* cur,
                               const double* Y, const double* Minv) {
    const int Vp = ctx->nblk * RC_NB;
    if (k <= RC_KSMALL)
        hipLaunchKernelGGL(k_ls_project_grid<LSG_T>, dim3(lsg_grid(ctx)), dim3(LSG_T), 0,
                           ctx->stream, d_ctl, d_set, ctx->d_ell0, pos0, dir, cur, ctx->ed,
                           ctx->V, ctx->E, Vp, Y, Minv, ctx->d_vOff, ctx->d_lsp, ctx->d_phi,
                           ctx->d_schur, ctx->d_gsync, ctx->d_status);
    else
        hipLaunchKernelGGL(k_ls_project_grid<1024>, dim3(RC_LSG_NW), dim3(1024), 0, ctx->stream,
                           d_ctl, d_set, ctx->d_ell0, pos0, dir, cur, ctx->ed, ctx->V, ctx->E,
                           Vp, Y, Minv, ctx->d_vOff, ctx->d_lsp, ctx->d_phi, ctx->d_schur,
                           ctx->d_gsync, ctx->d_status);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_ls_project_re(rc_ctx* ctx, int phase, LSCtl* d_ctl, const DevSet* d_set,
                                  const DevPen& pen, const rc_params* p, const double* pos0,
                                  const double* dir, double* cur, const double* Y,
                                  const double* Minv) {
    const int Vp = ctx->nblk * RC_NB;
    hipLaunchKernelGGL(k_ls_project_re, dim3(1), dim3(CT), 0, ctx->stream, phase, d_ctl, d_set,
                       ctx->d_ell0, pen, pos0, dir, cur, ctx->ed, ctx->V, ctx->E, Vp, Y, Minv,
                       ctx->d_vOff, ctx->d_vEdge, p->alpha, p->epsilon, ctx->d_phi,
                       (const int32_t*)ctx->d_status);
    RC_CHECK_LAUNCH();
    return RC_OK;
}


// Energy of the projected trial cur: its edge table and penalty energy,
// then the pair energy with the Armijo decision (both skipped on the device
// when the trial's projection failed or the search is done). The penalty
// partials use d_epart_ls before the pair energy's partials do.
rc_status rc_launch_ls_energy(rc_ctx* ctx, LSCtl* d_ctl, const rc_params* p, const DevPen& pen,
                              const double* cur, double* d_epart_ls) {
    const int nwg = std::min((ctx->E + 255) / 256, 256);
    hipLaunchKernelGGL(k_ls_edges, dim3(nwg), dim3(256), 0, ctx->stream, d_ctl, cur, ctx->ed,
                       ctx->V, ctx->E, pen, (const int32_t*)ctx->d_vOff,
                       (const int32_t*)ctx->d_vEdge, p->alpha, p->epsilon, d_epart_ls);
    RC_CHECK_LAUNCH();
    int grid = (ctx->npairs + 3) / 4;
    if (grid > 1024) grid = 1024;
    if (p->alpha == 3.0 && p->beta == 6.0) {
        hipLaunchKernelGGL(k_energy_ls<Pow36>, dim3(grid), dim3(256), 0, ctx->stream, ctx->ed,
                           ctx->E, ctx->npairs, Pow36{}, p->epsilon, d_epart_ls, d_ctl);
    } else {
        const double s = (p->beta - 1) / p->alpha;
        PowGeneric pw{p->alpha, p->beta, 2 * (s - 1) + 1};
        hipLaunchKernelGGL(k_energy_ls<PowGeneric>, dim3(grid), dim3(256), 0, ctx->stream, ctx->ed,
                           ctx->E, ctx->npairs, pw, p->epsilon, d_epart_ls, d_ctl);
    }
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_ls_finalize(rc_ctx* ctx, LSCtl* d_ctl, double* pos, const double* cur,
                                int update, double* d_stats) {
    hipLaunchKernelGGL(k_ls_finalize, dim3(1), dim3(CT), 0, ctx->stream, d_ctl, pos, cur, ctx->V,
                       update, (const double*)(ctx->d_scal + RC_SCAL_RESID),
                       (const int32_t*)ctx->d_status, (const double*)(ctx->d_refine + 2), d_stats);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

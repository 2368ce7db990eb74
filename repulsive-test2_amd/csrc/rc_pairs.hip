// Tangent-point energy and its adjoint gradient over all disjoint edge pairs.
//
// Reference: calculateEnergy (src/core/tangentPointEnergy.ts:50-111) and
// gradientAnalytical with kernelDerivs (:139-392). The reference visits every
// ORDERED disjoint pair (I,J) and halves the sum; here every UNORDERED pair
// {I,J} is visited once and evaluates both ordered terms E_IJ + E_JI, which
// share the four endpoint distances. E = 1/2 * sum_{I<J} (E_IJ + E_JI).
//
// Work decomposition: edges are cut into tiles of 64; a wave64 owns one
// unordered tile pair (a <= b), lane t owns edge I = 64a + t. The energy
// kernel walks the 64 edges of tile b with wave-uniform J. The gradient
// kernel walks them in a lane rotation (lane t visits J = 64b + (t+s)%64 at
// step s) so the J-side accumulators can travel with their edge through a
// lane shuffle: every accumulator has one owner at every step, so there are
// no atomics and the sum order is fixed (bit-reproducible run to run).
#include <cmath>
#include <mutex>

#include "rc_internal.h"
#include "rc_math.h"
#include "rc_pair_energy.h"

namespace {

__device__ __forceinline__ void tri_decode(int p, int& a, int& b) {
    // p = b(b+1)/2 + a, 0 <= a <= b
    int bb = (int)((__builtin_sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((bb + 1) * (bb + 2) / 2 <= p) ++bb;
    while (bb * (bb + 1) / 2 > p) --bb;
    b = bb;
    a = p - bb * (bb + 1) / 2;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Edge table of the current positions. Endpoints i1[K * is], i2[K * is]
// (is = 1 for the context's tables, 2 for a caller's int32[E][2]). check:
// endpoints are not trusted (a caller's device table): an endpoint outside
// [0, V) or a self-loop reads no position and turns the edge's geometry into
// NaN, so every energy of it is NaN.
__global__ void k_edge_prep(const double* __restrict__ pos, int V, const int32_t* __restrict__ i1,
                            const int32_t* __restrict__ i2, int is, int check, int E, int Epad,
                            double alpha, double eps, EdgeSoA ed) {
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    if (K >= Epad) return;
    if (K < E) {
        int a = i1[(size_t)K * is], b = i2[(size_t)K * is];
        const bool bad = check && (a < 0 || a >= V || b < 0 || b >= V || a == b);
        if (bad) a = b = 0;
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
        ed.i1[K] = a;
        ed.i2[K] = b;
        if (bad) {
            const double nan = __builtin_nan("");
            ed.ax[K] = ed.ay[K] = ed.az[K] = ed.bx[K] = ed.by[K] = ed.bz[K] = nan;
            ed.ell[K] = ed.rlen[K] = ed.lpow[K] = ed.dpc[K] = nan;
            ed.i1[K] = -3;
            ed.i2[K] = -4;
        }
    } else {
        ed.ax[K] = ed.ay[K] = ed.az[K] = 0;
        ed.bx[K] = ed.by[K] = ed.bz[K] = 0;
        ed.ell[K] = 1;
        ed.rlen[K] = 0;
        ed.ux[K] = ed.uy[K] = ed.uz[K] = 0;
        ed.lpow[K] = 0;
        ed.dpc[K] = 0;
        ed.i1[K] = -1;
        ed.i2[K] = -2;
    }
}

// ---------------------------------------------------------------------------
// energy


template <class P>
__global__ __launch_bounds__(256) void k_energy(EdgeSoA ed, int E, int npairs, P pw, double eps,
                                                double* __restrict__ part) {
    __shared__ rcpair::ETile jt[4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int base = blockIdx.x * 4; base < npairs; base += gridDim.x * 4) {
        const int p = __builtin_amdgcn_readfirstlane(base + wave);
        if (p >= npairs) break;
        int a, b;
        tri_decode(p, a, b);
        double sum = rcpair::tile_pair_energy(ed, E, pw, eps, a, b, lane, jt[wave]);
        sum = wave_sum(sum);
        if (lane == 0) part[p] = sum;
    }
}

// Fixed-order reduction of per-tile-pair partials: thread t sums
// part[t], part[t+1024], ... then a shared-memory tree. out = scale * sum.
__global__ __launch_bounds__(1024) void k_reduce_sum(const double* __restrict__ part, int n,
                                                     double scale, double* __restrict__ out) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = scale * sh[0];
}

// ---------------------------------------------------------------------------
// gradient

// Cross-product part of kernelDerivs (tangentPointEnergy.ts:165-244) for
// f(e,d) = (|e x d|+eps)^a / (|d|+eps)^b: c = e x d, cA = (|c|+eps)^a, the
// factor K = a (|c|+eps)^(a-1) / |c| (0 when |c| < 1e-14) of the c-derivative,
// and cce = c x e. With the distance parts invB = deps^-b and
// invB1 = deps^-(b+1) the derivatives are
//   df/dd = K invB cce - b cA invB1 d/|d|,    df/de = K invB (d x c).
struct CrossPart {
    double c[3], cce[3], cA, K;
};

template <class P>
__device__ __forceinline__ CrossPart cross_part(const P& pw, double eps, double alpha, double ex,
                                                double ey, double ez, double dx, double dy,
                                                double dz) {
    CrossPart r;
    r.c[0] = ey * dz - ez * dy;
    r.c[1] = ez * dx - ex * dz;
    r.c[2] = ex * dy - ey * dx;
    double rc, invRc;
    rcm::sqrt_rsqrt(r.c[0] * r.c[0] + r.c[1] * r.c[1] + r.c[2] * r.c[2], rc, invRc);
    const double ceps = rc + eps;
    r.cA = pw.c_alpha(ceps);
    r.K = (rc >= 1e-14) ? alpha * pw.c_alpha_m1(ceps) * invRc : 0.0;
    r.cce[0] = r.c[1] * ez - r.c[2] * ey;
    r.cce[1] = r.c[2] * ex - r.c[0] * ez;
    r.cce[2] = r.c[0] * ey - r.c[1] * ex;
    return r;
}

struct EdgeRegs {
    double P[6];   // endpoint positions a, b
    double e[3];   // b - a
    double u[3];   // unit tangent (guarded)
    double ell, lp, dpc;
    int i1, i2;
};

// Contributions of one unordered disjoint pair {I,J} (both ordered terms)
// to the four endpoint slots: cI[0..2] -> I.i1, cI[3..5] -> I.i2, same for J.
// e_I x (p_i - q_j) does not depend on the endpoint i of I (the two
// differences differ by e_I), so the I-ordered terms share one cross part
// per endpoint j of J, and the J-ordered terms one per endpoint i of I; their
// e-derivatives over both endpoints sum to one cross product each:
// sum_i K invB_ij (d_ij x c_j) = K (sum_i invB_ij d_ij) x c_j.
template <class P>
__device__ __forceinline__ void pair_grad(const P& pw, double eps, double alpha, double beta,
                                          const EdgeRegs& I, const EdgeRegs& J, double* cI,
                                          double* cJ) {
    // Contributions accumulate straight into the four endpoint slots, scaled
    // by the ordered-pair bases b_IJ = ell_I^(1-a) ell_J / 4, b_JI likewise.
    const double bIJ = 0.25 * I.lp * J.ell;
    const double bJI = 0.25 * J.lp * I.ell;
    double d[2][2][3], dh[2][2][3], invB[2][2], invB1[2][2];
#pragma unroll
    for (int ia = 0; ia < 2; ++ia) {
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
#pragma unroll
            for (int c = 0; c < 3; ++c) d[ia][jb][c] = I.P[3 * ia + c] - J.P[3 * jb + c];
            double rd, invRd;
            rcm::sqrt_rsqrt(d[ia][jb][0] * d[ia][jb][0] + d[ia][jb][1] * d[ia][jb][1] +
                                d[ia][jb][2] * d[ia][jb][2],
                            rd, invRd);
            const double hs = (rd < 1e-14) ? 0.0 : invRd;
#pragma unroll
            for (int c = 0; c < 3; ++c) dh[ia][jb][c] = hs * d[ia][jb][c];
            const double deps = rd + eps;
            const double inv = rcm::rcp(deps);
            invB[ia][jb] = pw.inv_beta(deps, inv);
            invB1[ia][jb] = invB[ia][jb] * inv;
        }
    }
    // I-ordered terms, tangent e_I, d = p_i - q_j: one cross part per j
    CrossPart xi[2];
#pragma unroll
    for (int jb = 0; jb < 2; ++jb)
        xi[jb] = cross_part(pw, eps, alpha, I.e[0], I.e[1], I.e[2], d[0][jb][0], d[0][jb][1],
                            d[0][jb][2]);
    // J-ordered terms, tangent e_J, d' = q_j - p_i = -d: one cross part per i
    CrossPart xj[2];
#pragma unroll
    for (int ia = 0; ia < 2; ++ia)
        xj[ia] = cross_part(pw, eps, alpha, J.e[0], J.e[1], J.e[2], -d[ia][0][0], -d[ia][0][1],
                            -d[ia][0][2]);
    double sFI = 0.0, sFJ = 0.0;
    double deI[3] = {0.0, 0.0, 0.0}, deJ[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        cI[c] = 0.0;
        cJ[c] = 0.0;
    }
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const CrossPart& x = xi[jb];
        double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int ia = 0; ia < 2; ++ia) {
            sFI += x.cA * invB[ia][jb];
            const double kk = x.K * invB[ia][jb];
            const double cd = -beta * x.cA * invB1[ia][jb];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double dd = kk * x.cce[c] + cd * dh[ia][jb][c];
                cI[3 * ia + c] += bIJ * dd;
                cJ[3 * jb + c] -= bIJ * dd;
                s[c] += invB[ia][jb] * d[ia][jb][c];
            }
        }
        // K (sum_i invB_ij d_ij) x c_j
        deI[0] += x.K * (s[1] * x.c[2] - s[2] * x.c[1]);
        deI[1] += x.K * (s[2] * x.c[0] - s[0] * x.c[2]);
        deI[2] += x.K * (s[0] * x.c[1] - s[1] * x.c[0]);
    }
#pragma unroll
    for (int ia = 0; ia < 2; ++ia) {
        const CrossPart& x = xj[ia];
        double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
            sFJ += x.cA * invB[ia][jb];
            const double kk = x.K * invB[ia][jb];
            const double cd = -beta * x.cA * invB1[ia][jb];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double dd = kk * x.cce[c] - cd * dh[ia][jb][c];
                cJ[3 * jb + c] += bJI * dd;
                cI[3 * ia + c] -= bJI * dd;
                s[c] -= invB[ia][jb] * d[ia][jb][c];
            }
        }
        deJ[0] += x.K * (s[1] * x.c[2] - s[2] * x.c[1]);
        deJ[1] += x.K * (s[2] * x.c[0] - s[0] * x.c[2]);
        deJ[2] += x.K * (s[0] * x.c[1] - s[1] * x.c[0]);
    }
    // e = b - a: de/da = -1, de/db = +1; length derivatives d(ell_I^(1-a))
    // and d(ell_I) carry -u_I at i1, +u_I at i2
    const double gIs = 0.25 * J.ell * sFI * I.dpc + 0.25 * J.lp * sFJ;
    const double gJs = 0.25 * I.lp * sFI + 0.25 * I.ell * sFJ * J.dpc;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double gI = gIs * I.u[c] + bIJ * deI[c];
        const double gJ = gJs * J.u[c] + bJI * deJ[c];
        cI[c] -= gI;
        cI[3 + c] += gI;
        cJ[c] -= gJ;
        cJ[3 + c] += gJ;
    }
}

__device__ __forceinline__ void load_edge(const EdgeSoA& ed, int K, EdgeRegs& r) {
    r.P[0] = ed.ax[K]; r.P[1] = ed.ay[K]; r.P[2] = ed.az[K];
    r.P[3] = ed.bx[K]; r.P[4] = ed.by[K]; r.P[5] = ed.bz[K];
    r.e[0] = r.P[3] - r.P[0]; r.e[1] = r.P[4] - r.P[1]; r.e[2] = r.P[5] - r.P[2];
    r.u[0] = ed.ux[K]; r.u[1] = ed.uy[K]; r.u[2] = ed.uz[K];
    r.ell = ed.ell[K]; r.lp = ed.lpow[K]; r.dpc = ed.dpc[K];
    r.i1 = ed.i1[K]; r.i2 = ed.i2[K];
}

// LDS copy of one J tile (SoA, 64 entries per field).
struct JTile {
    double P[6][RC_TILE];
    double u[3][RC_TILE];
    double ell[RC_TILE], lp[RC_TILE], dpc[RC_TILE];
    int i1[RC_TILE], i2[RC_TILE];
};

template <class P>
__global__ __launch_bounds__(256) void k_gradient(EdgeSoA ed, int E, int Epad, int npairs, P pw,
                                                  double eps, double alpha, double beta,
                                                  double* __restrict__ gpart) {
    __shared__ JTile jt[4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    JTile& T = jt[wave];
    for (int base = blockIdx.x * 4; base < npairs; base += gridDim.x * 4) {
        const int p = __builtin_amdgcn_readfirstlane(base + wave);
        if (p >= npairs) break;
        int a, b;
        tri_decode(p, a, b);
        const bool diag = (a == b);
        const int I = a * RC_TILE + lane;
        EdgeRegs eI;
        load_edge(ed, I, eI);  // padded slots hold benign values
        {
            const int J = b * RC_TILE + lane;
            T.P[0][lane] = ed.ax[J]; T.P[1][lane] = ed.ay[J]; T.P[2][lane] = ed.az[J];
            T.P[3][lane] = ed.bx[J]; T.P[4][lane] = ed.by[J]; T.P[5][lane] = ed.bz[J];
            T.u[0][lane] = ed.ux[J]; T.u[1][lane] = ed.uy[J]; T.u[2][lane] = ed.uz[J];
            T.ell[lane] = ed.ell[J]; T.lp[lane] = ed.lpow[J]; T.dpc[lane] = ed.dpc[J];
            T.i1[lane] = ed.i1[J]; T.i2[lane] = ed.i2[J];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double accI[6] = {0, 0, 0, 0, 0, 0};
        double accJ[6] = {0, 0, 0, 0, 0, 0};  // held for J-lane (lane + s) % 64
        const int s0 = diag ? 1 : 0;
        const int s1 = diag ? 32 : 63;
        const bool Iok = I < E;
        for (int s = s0; s <= s1; ++s) {
            const int jl = (lane + s) & (RC_TILE - 1);
            const int J = b * RC_TILE + jl;
            EdgeRegs eJ;
#pragma unroll
            for (int c = 0; c < 6; ++c) eJ.P[c] = T.P[c][jl];
            eJ.e[0] = eJ.P[3] - eJ.P[0]; eJ.e[1] = eJ.P[4] - eJ.P[1]; eJ.e[2] = eJ.P[5] - eJ.P[2];
            eJ.u[0] = T.u[0][jl]; eJ.u[1] = T.u[1][jl]; eJ.u[2] = T.u[2][jl];
            eJ.ell = T.ell[jl]; eJ.lp = T.lp[jl]; eJ.dpc = T.dpc[jl];
            eJ.i1 = T.i1[jl]; eJ.i2 = T.i2[jl];
            const bool valid = Iok && J < E && (!diag || s < 32 || lane < 32) &&
                               eI.i1 != eJ.i1 && eI.i1 != eJ.i2 && eI.i2 != eJ.i1 &&
                               eI.i2 != eJ.i2;
            double cI[6], cJ[6];
            pair_grad(pw, eps, alpha, beta, eI, eJ, cI, cJ);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                // addToGrad's `x || 0` (tangentPointEnergy.ts:152-156): NaN adds 0
                accI[c] += (valid && cI[c] == cI[c]) ? cI[c] : 0.0;
                accJ[c] += (valid && cJ[c] == cJ[c]) ? cJ[c] : 0.0;
            }
            // pass every J accumulator one lane down so lane t holds J-lane (t+s+1)%64
#pragma unroll
            for (int c = 0; c < 6; ++c) accJ[c] = __shfl(accJ[c], (lane + 1) & 63, 64);
        }
        // lane t now holds J-lane (t + s0 + nsteps) % 64 = (t + s1 + 1) % 64;
        // realign so that lane t holds J-lane t
        const int src = (lane - s1 - 1) & 63;
#pragma unroll
        for (int c = 0; c < 6; ++c) accJ[c] = __shfl(accJ[c], src, 64);
        // slot writes: gpart[(t * 6 + c) * Epad + K]
        if (diag) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
                gpart[((size_t)a * 6 + c) * Epad + I] = accI[c] + accJ[c];
        } else {
            const int J = b * RC_TILE + lane;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                gpart[((size_t)b * 6 + c) * Epad + I] = accI[c];
                gpart[((size_t)a * 6 + c) * Epad + J] = accJ[c];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Per-vertex gather of the per-edge-endpoint partials in fixed order, times
// the reference's final 1/2 (tangentPointEnergy.ts:385-389), divided by div
// (1 for the analytic gradient, exact; h for the forward difference). Output
// SoA. A block owns 64 vertices (lane = vertex, so the slot reads of
// neighbouring vertices coalesce); its GATHER_WAVES waves split the partner
// tiles into contiguous ranges and the wave sums are added in wave order, so
// the result does not depend on scheduling.
constexpr int GATHER_WAVES = 16;

__global__ __launch_bounds__(64 * GATHER_WAVES) void k_grad_gather(
    const double* __restrict__ gpart, int ntiles, int Epad, int V,
    const int32_t* __restrict__ vOff, const int32_t* __restrict__ vEdge,
    const int32_t* __restrict__ vSlot, double div, double* __restrict__ g) {
    __shared__ double sh[GATHER_WAVES][3][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int v = blockIdx.x * 64 + lane;
    const int t0 = (w * ntiles) / GATHER_WAVES, t1 = ((w + 1) * ntiles) / GATHER_WAVES;
    double gx = 0, gy = 0, gz = 0;
    if (v < V) {
        for (int q = vOff[v]; q < vOff[v + 1]; ++q) {
            const int K = vEdge[q];
            const int sl = vSlot[q] * 3;
            for (int t = t0; t < t1; ++t) {
                const size_t base = ((size_t)t * 6 + sl) * Epad + K;
                gx += gpart[base];
                gy += gpart[base + Epad];
                gz += gpart[base + 2 * (size_t)Epad];
            }
        }
    }
    sh[w][0][lane] = gx;
    sh[w][1][lane] = gy;
    sh[w][2][lane] = gz;
    __syncthreads();
    if (w < 3 && v < V) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < GATHER_WAVES; ++k) s += sh[k][w][lane];
        g[(size_t)w * V + v] = (0.5 * s) / div;
    }
}

// ---------------------------------------------------------------------------
// forward-difference gradient
//
// gradientFiniteDiff (tangentPointEnergy.ts:114-136) forms, for each vertex v
// and coordinate d, (E(gamma + h e_vd) - E(gamma)) / h from two full
// energies. Moving v changes only the pairs {I, J} in which v is an endpoint
// of I (a J disjoint from I never contains v, and two edges through v are
// never disjoint), so
//   E(gamma + h e_vd) - E(gamma) = 1/2 sum_{I with endpoint v} sum_{J disjoint
//                                  from I} [pe(I', J) - pe(I, J)]
// with pe the unordered-pair energy (both ordered terms) and I' the edge with
// that endpoint coordinate moved to x + h (one rounding, as the reference's
// `+= h`). The per-pair differences are summed instead of differencing two
// O(E^2) sums, so the result is the same quotient with less cancellation.
//
// A wave owns one ordered tile pair (a, b) of the full E x E square; lane t
// owns edge I = 64a + t and walks tile b's 64 edges from LDS. For each J it
// evaluates pe(I, J) once and pe(I', J) for the six moves of I's endpoints
// (m = 3 * endpoint + coordinate, the gradient's slot layout), and writes
// the six differences summed over tile b to part[(b * 6 + m) * Epad + I].
// k_grad_gather then sums the slots of each vertex over all b in fixed order.

template <class P>
__global__ __launch_bounds__(256) void k_fd_delta(EdgeSoA ed, int E, int Epad, int ntiles, P pw,
                                                  double eps, double alpha, double h,
                                                  double* __restrict__ part) {
    __shared__ rcpair::ETile jt[4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    rcpair::ETile& T = jt[wave];
    const int ntask = ntiles * ntiles;
    for (int base = blockIdx.x * 4; base < ntask; base += gridDim.x * 4) {
        const int task = __builtin_amdgcn_readfirstlane(base + wave);
        if (task >= ntask) break;
        const int a = task / ntiles, b = task - a * ntiles;
        {
            const int J = b * RC_TILE + lane;
            const int Jc = J < E ? J : E - 1;
            T.P[0][lane] = ed.ax[Jc]; T.P[1][lane] = ed.ay[Jc]; T.P[2][lane] = ed.az[Jc];
            T.P[3][lane] = ed.bx[Jc]; T.P[4][lane] = ed.by[Jc]; T.P[5][lane] = ed.bz[Jc];
            T.ell[lane] = ed.ell[Jc];
            T.lp[lane] = ed.lpow[Jc];
            T.i1[lane] = ed.i1[Jc];
            T.i2[lane] = ed.i2[Jc];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int I = a * RC_TILE + lane;
        const bool Iok = I < E;
        const int Ic = Iok ? I : 0;
        const double PI[6] = {ed.ax[Ic], ed.ay[Ic], ed.az[Ic], ed.bx[Ic], ed.by[Ic], ed.bz[Ic]};
        const double ellI = ed.ell[Ic], lpI = ed.lpow[Ic];
        const int i1 = ed.i1[Ic], i2 = ed.i2[Ic];
        // the six moved copies of edge I: moved coordinate, length, l^(1-alpha)
        double xm[6], ellM[6], lpM[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            double Q[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) Q[c] = PI[c];
            Q[m] = PI[m] + h;
            xm[m] = Q[m];
            const double ex = Q[3] - Q[0], ey = Q[4] - Q[1], ez = Q[5] - Q[2];
            ellM[m] = __builtin_sqrt(ex * ex + ey * ey + ez * ez) + eps;
            lpM[m] = pow(ellM[m], 1.0 - alpha);
        }
        const int ns = min(RC_TILE, E - b * RC_TILE);
        double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
        for (int s = 0; s < ns; ++s) {
            const double PJ[6] = {T.P[0][s], T.P[1][s], T.P[2][s], T.P[3][s], T.P[4][s], T.P[5][s]};
            const double eJx = PJ[3] - PJ[0], eJy = PJ[4] - PJ[1], eJz = PJ[5] - PJ[2];
            const double ellJ = T.ell[s], lpJ = T.lp[s];
            const int j1 = T.i1[s], j2 = T.i2[s];
            const bool valid = Iok && i1 != j1 && i1 != j2 && i2 != j1 && i2 != j2;
            rcpair::PairParts q;
            const double e0 = rcpair::pair_energy_parts(pw, eps, PI, PJ, eJx, eJy, eJz, ellI, ellJ,
                                                        lpI, lpJ, q);
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                double Q[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) Q[c] = PI[c];
                Q[m] = xm[m];
                const double em =
                    m < 3 ? rcpair::pair_energy_moved<0>(pw, eps, Q, PJ, eJx, eJy, eJz, ellM[m],
                                                         ellJ, lpM[m], lpJ, q)
                          : rcpair::pair_energy_moved<1>(pw, eps, Q, PJ, eJx, eJy, eJz, ellM[m],
                                                         ellJ, lpM[m], lpJ, q);
                acc[m] += valid ? (em - e0) : 0.0;
            }
        }
#pragma unroll
        for (int m = 0; m < 6; ++m) part[((size_t)b * 6 + m) * Epad + I] = acc[m];
        // the tile is rewritten by the wave's next task
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void k_aos_to_soa(const double* __restrict__ aos, int V, double* __restrict__ soa) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    soa[v] = aos[3 * v];
    soa[V + v] = aos[3 * v + 1];
    soa[2 * V + v] = aos[3 * v + 2];
}

__global__ void k_soa_to_aos(const double* __restrict__ soa, int V, double* __restrict__ aos) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    aos[3 * v] = soa[v];
    aos[3 * v + 1] = soa[V + v];
    aos[3 * v + 2] = soa[2 * V + v];
}

// step()'s update (optimizer.ts:56-60): out = pos - step * grad with the
// product rounded before the difference. The library is built with
// -ffp-contract=fast, under which the backend fuses a multiply into a
// following add whatever the source pragmas say; the empty asm makes the
// rounded product an opaque register value, so no fused multiply-add forms.
__global__ void k_l2_update(const double* __restrict__ pos, const double* __restrict__ grad,
                            int n, double step, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t = step * grad[i];
    asm volatile("" : "+v"(t));
    out[i] = pos[i] - t;
}

bool is36(const rc_params* p) { return p->alpha == 3.0 && p->beta == 6.0; }

int pair_grid(int npairs) {
    int blocks = (npairs + 3) / 4;
    return blocks < 4096 ? blocks : 4096;
}

}  // namespace

rc_status rc_launch_edge_prep(rc_ctx* ctx, const double* pos, const rc_params* p) {
    hipLaunchKernelGGL(k_edge_prep, dim3((ctx->Epad + 255) / 256), dim3(256), 0, ctx->stream, pos,
                       ctx->V, (const int32_t*)ctx->d_i1, (const int32_t*)ctx->d_i2, 1, 0, ctx->E,
                       ctx->Epad, p->alpha, p->epsilon, ctx->ed);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_energy(rc_ctx* ctx, const rc_params* p, double* d_out) {
    const int grid = pair_grid(ctx->npairs);
    if (is36(p)) {
        hipLaunchKernelGGL(k_energy<Pow36>, dim3(grid), dim3(256), 0, ctx->stream, ctx->ed, ctx->E,
                           ctx->npairs, Pow36{}, p->epsilon, ctx->d_epart);
    } else {
        const double s = (p->beta - 1) / p->alpha;
        PowGeneric pw{p->alpha, p->beta, 2 * (s - 1) + 1};
        hipLaunchKernelGGL(k_energy<PowGeneric>, dim3(grid), dim3(256), 0, ctx->stream, ctx->ed,
                           ctx->E, ctx->npairs, pw, p->epsilon, ctx->d_epart);
    }
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(1024), 0, ctx->stream, ctx->d_epart, ctx->npairs,
                       0.5, d_out);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_gradient(rc_ctx* ctx, const rc_params* p, double* d_grad_soa) {
    const int grid = pair_grid(ctx->npairs);
    if (is36(p)) {
        hipLaunchKernelGGL(k_gradient<Pow36>, dim3(grid), dim3(256), 0, ctx->stream, ctx->ed,
                           ctx->E, ctx->Epad, ctx->npairs, Pow36{}, p->epsilon, p->alpha, p->beta,
                           ctx->d_gpart);
    } else {
        const double s = (p->beta - 1) / p->alpha;
        PowGeneric pw{p->alpha, p->beta, 2 * (s - 1) + 1};
        hipLaunchKernelGGL(k_gradient<PowGeneric>, dim3(grid), dim3(256), 0, ctx->stream, ctx->ed,
                           ctx->E, ctx->Epad, ctx->npairs, pw, p->epsilon, p->alpha, p->beta,
                           ctx->d_gpart);
    }
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_grad_gather, dim3((ctx->V + 63) / 64), dim3(64 * GATHER_WAVES), 0, ctx->stream,
                       ctx->d_gpart, ctx->ntiles, ctx->Epad, ctx->V, ctx->d_vOff, ctx->d_vEdge,
                       ctx->d_vSlot, 1.0, d_grad_soa);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_gradient_fd(rc_ctx* ctx, const rc_params* p, double h, double* d_grad_soa) {
    const int ntask = ctx->ntiles * ctx->ntiles;
    const int grid = pair_grid(ntask);
    if (is36(p)) {
        hipLaunchKernelGGL(k_fd_delta<Pow36>, dim3(grid), dim3(256), 0, ctx->stream, ctx->ed,
                           ctx->E, ctx->Epad, ctx->ntiles, Pow36{}, p->epsilon, p->alpha, h,
                           ctx->d_gpart);
    } else {
        const double s = (p->beta - 1) / p->alpha;
        PowGeneric pw{p->alpha, p->beta, 2 * (s - 1) + 1};
        hipLaunchKernelGGL(k_fd_delta<PowGeneric>, dim3(grid), dim3(256), 0, ctx->stream, ctx->ed,
                           ctx->E, ctx->Epad, ctx->ntiles, pw, p->epsilon, p->alpha, h,
                           ctx->d_gpart);
    }
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_grad_gather, dim3((ctx->V + 63) / 64), dim3(64 * GATHER_WAVES), 0, ctx->stream,
                       ctx->d_gpart, ctx->ntiles, ctx->Epad, ctx->V, ctx->d_vOff, ctx->d_vEdge,
                       ctx->d_vSlot, h, d_grad_soa);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_l2_update(rc_ctx* ctx, const double* pos, const double* grad,
                              double step_size, double* out) {
    const int n = 3 * ctx->V;
    hipLaunchKernelGGL(k_l2_update, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, pos, grad, n,
                       step_size, out);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_upload_positions_aos(rc_ctx* ctx, const double* xyz, double* d_dst) {
    // stage AoS into d_tmp then transpose on device
    RC_HIP(hipMemcpyAsync(ctx->d_stage, xyz, sizeof(double) * 3 * ctx->V, hipMemcpyHostToDevice,
                          ctx->stream));
    hipLaunchKernelGGL(k_aos_to_soa, dim3((ctx->V + 255) / 256), dim3(256), 0, ctx->stream,
                       ctx->d_stage, ctx->V, d_dst);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_download_positions_aos(rc_ctx* ctx, const double* d_src, double* xyz) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3((ctx->V + 255) / 256), dim3(256), 0, ctx->stream, d_src,
                       ctx->V, ctx->d_stage);
    RC_CHECK_LAUNCH();
    RC_HIP(hipMemcpyAsync(xyz, ctx->d_stage, sizeof(double) * 3 * ctx->V, hipMemcpyDeviceToHost,
                          ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    return RC_OK;
}

// ---------------------------------------------------------------------------
// Energy without allocation (rc_energy_noalloc): the edge table and the
// tile-pair partials of rc_energy live in this file-scope device array
// instead of context buffers, sized for E <= RC_NOALLOC_EMAX. Each GPU holds
// its own copy until the process ends; its device address is looked up once
// per device (first call with that device current) and kept.

namespace {

constexpr int NA_EPAD = RC_NOALLOC_EMAX;                  // a multiple of RC_TILE
constexpr int NA_TILES = NA_EPAD / RC_TILE;
constexpr int NA_PAIRS = NA_TILES * (NA_TILES + 1) / 2;
constexpr size_t NA_DOUBLES = 14 * (size_t)NA_EPAD + NA_PAIRS;

}  // namespace

static __device__ __align__(256) double scratch[NA_DOUBLES];

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
double* na_ptr[NA_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (the array's name is a
// host-side stand-in and must never reach a kernel).
rc_status scratch_on_current_device(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return RC_HIP_ERROR;
    std::lock_guard<std::mutex> lock(na_mu);
    if (!na_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RC_HIP_ERROR;
        na_ptr[dev] = static_cast<double*>(p);
    }
    *out = na_ptr[dev];
    return RC_OK;
}

}  // namespace

extern "C" rc_status rc_energy_noalloc(const double* d_xyz, int32_t V, const int32_t* d_edges,
                                       int32_t E, const rc_params* p, double* d_energy,
                                       void* stream) {
    if (!d_xyz || !d_edges || !d_energy || !p || V < 2 || E < 1) return RC_BAD_ARG;
    if (!std::isfinite(p->alpha) || !std::isfinite(p->beta) || !std::isfinite(p->epsilon) ||
        p->alpha == 0.0)
        return RC_BAD_ARG;
    if (E > RC_NOALLOC_EMAX) return RC_UNSUPPORTED;
    double* base = nullptr;
    rc_status s = scratch_on_current_device(&base);
    if (s != RC_OK) return s;
    const int Epad = ((E + RC_TILE - 1) / RC_TILE) * RC_TILE;
    const int ntiles = Epad / RC_TILE;
    const int npairs = ntiles * (ntiles + 1) / 2;
    EdgeSoA ed;
    double** fields[] = {&ed.ax, &ed.ay, &ed.az, &ed.bx, &ed.by, &ed.bz, &ed.ell,
                         &ed.rlen, &ed.ux, &ed.uy, &ed.uz, &ed.lpow, &ed.dpc};
    for (int i = 0; i < 13; ++i) *fields[i] = base + (size_t)i * Epad;
    ed.i1 = reinterpret_cast<int32_t*>(base + 13 * (size_t)Epad);
    ed.i2 = ed.i1 + Epad;
    double* part = base + 14 * (size_t)Epad;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_edge_prep, dim3((Epad + 255) / 256), dim3(256), 0, st, d_xyz, V, d_edges,
                       d_edges + 1, 2, 1, E, Epad, p->alpha, p->epsilon, ed);
    const int grid = pair_grid(npairs);
    if (is36(p)) {
        hipLaunchKernelGGL(k_energy<Pow36>, dim3(grid), dim3(256), 0, st, ed, E, npairs, Pow36{},
                           p->epsilon, part);
    } else {
        const double sx = (p->beta - 1) / p->alpha;
        PowGeneric pw{p->alpha, p->beta, 2 * (sx - 1) + 1};
        hipLaunchKernelGGL(k_energy<PowGeneric>, dim3(grid), dim3(256), 0, st, ed, E, npairs, pw,
                           p->epsilon, part);
    }
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(1024), 0, st, (const double*)part, npairs, 0.5,
                       d_energy);
    return hipGetLastError() == hipSuccess ? RC_OK : RC_HIP_ERROR;
}

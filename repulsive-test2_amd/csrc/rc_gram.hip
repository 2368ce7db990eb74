// Fractional Sobolev Gram matrix A = B + B0 (V x V, FP64).
//
// Reference: assembleBHighFlat (src/core/sobolev/innerProduct.ts:38-140),
// assembleBLowFlat (:177-285), assembleAFlat (:325-350). The reference scatters
// 16 + 16 signed updates per ordered disjoint pair and symmetrises. Summing
// both orders of each unordered pair {I,J} gives, with g_I = (d_i1 - d_i2)/ell_I
// and h_I = d_i1 + d_i2 (vertex indicator vectors),
//
//   A = sum_K [ W_K g_K g_K^T + Q_K h_K h_K^T ]
//       - sum_{I != J disjoint} [ X_IJ ell_I ell_J g_I g_J^T + Y_IJ h_I h_J^T ]
//
//   w_IJ = 1/4 ell_I ell_J sum_ab (|p_a - q_b| + eps)^-(2 sigma + 1)      (symmetric)
//   X_IJ = 2 w_IJ (T_I . T_J) / (ell_I ell_J),  W_K = 2 sum_J w_KJ
//   q_IJ = 1/16 ell_I ell_J sum_ab (|T_I x (p_a - q_b)| + eps)^2 (|.|+eps)^-4 (|.|+eps)^-(2s+1)
//   Y_IJ = q_IJ + q_JI,                         Q_K = sum_J Y_KJ
//
// Ownership tiles: a block owns the A tile (U x V') of 32 x 32 vertices
// (U <= V'), evaluates X, Y, 2w for every pair of the edges incident to U and
// to V' into LDS, and gathers each A entry from its <= deg^2 incident edge
// pairs (tiles touching more edges than LDS stages at once are walked in
// chunks by k_gram_tiles_wide). Every A entry has exactly one writer (no atomics); both orientations
// are written from one value, so A is exactly symmetric. Row sums W_K, Q_K
// come from per-tile partials [tile][K] (each slot has one writer) reduced in
// fixed order, then k_gram_band adds the edge-own terms on the diagonal and
// the adjacent entries.
#include "rc_internal.h"
#include "rc_math.h"

namespace {

struct GEdge {
    double P[RC_RMAX][6];
    double u[RC_RMAX][3];
    double ell[RC_RMAX];
    int id[RC_RMAX], i1[RC_RMAX], i2[RC_RMAX], own[RC_RMAX];
};

__device__ __forceinline__ void tri_decode(int p, int& a, int& b) {
    int bb = (int)((__builtin_sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((bb + 1) * (bb + 2) / 2 <= p) ++bb;
    while (bb * (bb + 1) / 2 > p) --bb;
    b = bb;
    a = p - bb * (bb + 1) / 2;
}

// T_I x (p_a - q_b) does not depend on the endpoint a (the two differences
// differ by ell_I T_I), and T_J x (p_a - q_b) not on b: 2 + 2 cross-product
// norms per pair instead of 8 (equal to the per-endpoint form to rounding).
template <class P>
__device__ __forceinline__ void pair_weights(const P& pw, double eps, const double* PI,
                                             const double* uI, double ellI, const double* PJ,
                                             const double* uJ, double ellJ, double& w2,
                                             double& X, double& Y) {
    double n2I[2], n2J[2];
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const double dx = PI[0] - PJ[3 * jb];
        const double dy = PI[1] - PJ[3 * jb + 1];
        const double dz = PI[2] - PJ[3 * jb + 2];
        const double cx = uI[1] * dz - uI[2] * dy;
        const double cy = uI[2] * dx - uI[0] * dz;
        const double cz = uI[0] * dy - uI[1] * dx;
        const double n = rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps;
        n2I[jb] = n * n;
    }
#pragma unroll
    for (int ia = 0; ia < 2; ++ia) {
        const double dx = PI[3 * ia] - PJ[0];
        const double dy = PI[3 * ia + 1] - PJ[1];
        const double dz = PI[3 * ia + 2] - PJ[2];
        const double cx = uJ[1] * dz - uJ[2] * dy;
        const double cy = uJ[2] * dx - uJ[0] * dz;
        const double cz = uJ[0] * dy - uJ[1] * dx;
        const double n = rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps;
        n2J[ia] = n * n;
    }
    double sumD = 0.0, sumQI = 0.0, sumQJ = 0.0;
#pragma unroll
    for (int ia = 0; ia < 2; ++ia) {
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
            const double dx = PI[3 * ia] - PJ[3 * jb];
            const double dy = PI[3 * ia + 1] - PJ[3 * jb + 1];
            const double dz = PI[3 * ia + 2] - PJ[3 * jb + 2];
            const double deps = rcm::sqrt_(dx * dx + dy * dy + dz * dz) + eps;
            const double pd = pw.inv_dexp(deps);
            const double inv = rcm::rcp(deps);
            const double i2 = inv * inv;
            const double inv4pd = i2 * i2 * pd;
            sumD += pd;
            sumQI += n2I[jb] * inv4pd;
            sumQJ += n2J[ia] * inv4pd;
        }
    }
    const double ll = ellI * ellJ;
    const double w = 0.25 * ll * sumD;
    const double c = uI[0] * uJ[0] + uI[1] * uJ[1] + uI[2] * uJ[2];
    w2 = 2.0 * w;
    X = w2 * c / ll;
    Y = 0.0625 * ll * (sumQI + sumQJ);
}

__device__ __forceinline__ void load_tile_edges(GEdge& G, const int32_t* __restrict__ rOff,
                                                const int32_t* __restrict__ rEdge, int t,
                                                const EdgeSoA& ed,
                                                const int32_t* __restrict__ owner, int& n) {
    const int o = rOff[t];
    n = rOff[t + 1] - o;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int K = rEdge[o + q];
        G.P[q][0] = ed.ax[K]; G.P[q][1] = ed.ay[K]; G.P[q][2] = ed.az[K];
        G.P[q][3] = ed.bx[K]; G.P[q][4] = ed.by[K]; G.P[q][5] = ed.bz[K];
        G.u[q][0] = ed.ux[K]; G.u[q][1] = ed.uy[K]; G.u[q][2] = ed.uz[K];
        G.ell[q] = ed.ell[K];
        G.id[q] = K;
        G.i1[q] = ed.i1[K];
        G.i2[q] = ed.i2[K];
        G.own[q] = owner[K];
    }
}

template <class P>
__global__ __launch_bounds__(256) void k_gram_tiles(EdgeSoA ed, P pw, double eps, int V, int E,
                                                    int nvt, int ngram,
                                                    const int32_t* __restrict__ rOff,
                                                    const int32_t* __restrict__ rEdge,
                                                    const int32_t* __restrict__ owner,
                                                    const int32_t* __restrict__ vOff,
                                                    const int32_t* __restrict__ vSlot,
                                                    const int32_t* __restrict__ vLocal,
                                                    double* __restrict__ A, int ldA,
                                                    double* __restrict__ rsp) {
    __shared__ GEdge GU, GV;
    __shared__ double Xt[RC_RMAX][RC_RMAX + 1];
    __shared__ double Yt[RC_RMAX][RC_RMAX + 1];
    __shared__ double Wt[RC_RMAX][RC_RMAX + 1];
    // per tile: incidences of its vertices (offsets relative to the tile's
    // first vertex, entries local edge index | slot << 16)
    __shared__ int begU[RC_VTILE + 1], begW[RC_VTILE + 1];
    __shared__ int incU[2 * RC_RMAX], incW[2 * RC_RMAX];
    const int t = threadIdx.x;
    for (int p = blockIdx.x; p < ngram; p += gridDim.x) {
        int U, W;
        tri_decode(p, U, W);  // U <= W
        // tiles beyond the staging bound go to k_gram_tiles_wide
        if (rOff[U + 1] - rOff[U] > RC_RMAX || rOff[W + 1] - rOff[W] > RC_RMAX) continue;
        const bool diag = (U == W);
        int nU, nV;
        load_tile_edges(GU, rOff, rEdge, U, ed, owner, nU);
        load_tile_edges(GV, rOff, rEdge, W, ed, owner, nV);
        {
            // a tile's incidences number at most 2 RC_RMAX (each of its <=
            // RC_RMAX edges has two endpoints)
            const int u0 = vOff[U * RC_VTILE], w0 = vOff[W * RC_VTILE];
            if (t <= RC_VTILE) {
                begU[t] = vOff[min(U * RC_VTILE + t, V)] - u0;
                begW[t] = vOff[min(W * RC_VTILE + t, V)] - w0;
            }
            const int nu = vOff[min(U * RC_VTILE + RC_VTILE, V)] - u0;
            const int nw = vOff[min(W * RC_VTILE + RC_VTILE, V)] - w0;
            if (t < nu) incU[t] = vLocal[u0 + t] | (vSlot[u0 + t] << 16);
            if (t >= 128 && t - 128 < nw) incW[t - 128] = vLocal[w0 + t - 128] | (vSlot[w0 + t - 128] << 16);
        }
        __syncthreads();
        // pair phase
        for (int q = t; q < nU * nV; q += blockDim.x) {
            const int li = q / nV, lj = q - (q / nV) * nV;
            const int I = GU.id[li], J = GV.id[lj];
            const bool dis = GU.i1[li] != GV.i1[lj] && GU.i1[li] != GV.i2[lj] &&
                             GU.i2[li] != GV.i1[lj] && GU.i2[li] != GV.i2[lj];
            double w2 = 0.0, X = 0.0, Y = 0.0;
            if (dis) {
                // canonical role order (lower edge id first) so equal pairs give
                // equal bits, chosen by operand address (no divergent branch)
                const bool ord = I <= J;
                pair_weights(pw, eps, ord ? GU.P[li] : GV.P[lj], ord ? GU.u[li] : GV.u[lj],
                             ord ? GU.ell[li] : GV.ell[lj], ord ? GV.P[lj] : GU.P[li],
                             ord ? GV.u[lj] : GU.u[li], ord ? GV.ell[lj] : GU.ell[li], w2, X, Y);
            }
            Xt[li][lj] = X;
            Yt[li][lj] = Y;
            Wt[li][lj] = w2;
        }
        __syncthreads();
        // row-sum partials: slot [t][K] for K owned by one side, summed over
        // the edges owned by the other side (each ordered pair counted once).
        // Four consecutive lanes per row, each over every fourth entry, then
        // combined in fixed order ((0 + 1) + (2 + 3)).
        for (int side = 0; side < (diag ? 1 : 2); ++side) {
            const int nr = side ? nV : nU, nc = side ? nU : nV;
            const int row = t >> 2, part = t & 3;
            double sw = 0.0, sy = 0.0;
            bool live = false;
            if (row < nr) {
                live = side ? (GV.own[row] == W) : (GU.own[row] == U);
                for (int c = part; c < nc; c += 4) {
                    const bool on = side ? (GU.own[c] == U) : (GV.own[c] == W);
                    const double wv = side ? Wt[c][row] : Wt[row][c];
                    const double yv = side ? Yt[c][row] : Yt[row][c];
                    sw += on ? wv : 0.0;
                    sy += on ? yv : 0.0;
                }
            }
            sw += __shfl_xor(sw, 1, 64);
            sy += __shfl_xor(sy, 1, 64);
            sw += __shfl_xor(sw, 2, 64);
            sy += __shfl_xor(sy, 2, 64);
            if (row < nr && part == 0 && live) {
                const int K = side ? GV.id[row] : GU.id[row];
                const int other = side ? U : W;
                rsp[((size_t)other * 2) * E + K] = sw;
                rsp[((size_t)other * 2 + 1) * E + K] = sy;
            }
        }
        // gather phase: val(u, v), u in U row tile, v in W column tile,
        // outer loop over incidences of the lower vertex index.
        for (int pass = 0; pass < (diag ? 1 : 2); ++pass) {
            for (int q = threadIdx.x; q < RC_VTILE * RC_VTILE; q += blockDim.x) {
                int r, c;
                if (pass == 0) {  // consecutive threads -> consecutive columns v
                    r = q / RC_VTILE;
                    c = q % RC_VTILE;
                } else {          // consecutive threads -> consecutive rows u (for A[v][u])
                    c = q / RC_VTILE;
                    r = q % RC_VTILE;
                }
                int u = U * RC_VTILE + r, v = W * RC_VTILE + c;
                if (u >= V || v >= V) continue;
                bool swap = diag && v < u;   // canonical (min, max) order
                // tile-local indices: lo in the U tile, hi in the W tile
                const int rl = swap ? c : r, rh = swap ? r : c;
                double val = 0.0;
                const int q10 = begU[rl], n1 = begU[rl + 1] - q10;
                const int q20 = begW[rh], n2 = begW[rh + 1] - q20;
                if (n1 <= 2 && n2 <= 2) {
                    // vertices of degree <= 2 (curves, and all but the junctions of
                    // networks): the same terms in the same order, unrolled so the
                    // LDS reads issue together
#pragma unroll
                    for (int a = 0; a < 2; ++a) {
                        if (a >= n1) break;
                        const int e1 = incU[q10 + a];
                        const int li = e1 & 0xffff;
                        const double si = (e1 >> 16) ? -1.0 : 1.0;
#pragma unroll
                        for (int b = 0; b < 2; ++b) {
                            if (b >= n2) break;
                            const int e2 = incW[q20 + b];
                            const int lj = e2 & 0xffff;
                            const double sj = (e2 >> 16) ? -1.0 : 1.0;
                            val -= si * sj * Xt[li][lj] + Yt[li][lj];
                        }
                    }
                } else {
                    for (int q1 = q10; q1 < q10 + n1; ++q1) {
                        const int e1 = incU[q1];
                        const int li = e1 & 0xffff;
                        const double si = (e1 >> 16) ? -1.0 : 1.0;
                        for (int q2 = q20; q2 < q20 + n2; ++q2) {
                            const int e2 = incW[q2];
                            const int lj = e2 & 0xffff;
                            const double sj = (e2 >> 16) ? -1.0 : 1.0;
                            val -= si * sj * Xt[li][lj] + Yt[li][lj];
                        }
                    }
                }
                if (pass == 0)
                    A[(size_t)u * ldA + v] = val;
                else
                    A[(size_t)v * ldA + u] = val;
            }
        }
        __syncthreads();
    }
}

// Edges c0 .. c0+n-1 of tile t's incident-edge list into G (n <= RC_RMAX).
__device__ __forceinline__ void load_edge_chunk(GEdge& G, const int32_t* __restrict__ rOff,
                                                const int32_t* __restrict__ rEdge, int t, int c0,
                                                int n, const EdgeSoA& ed,
                                                const int32_t* __restrict__ owner) {
    const int o = rOff[t] + c0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int K = rEdge[o + q];
        G.P[q][0] = ed.ax[K]; G.P[q][1] = ed.ay[K]; G.P[q][2] = ed.az[K];
        G.P[q][3] = ed.bx[K]; G.P[q][4] = ed.by[K]; G.P[q][5] = ed.bz[K];
        G.u[q][0] = ed.ux[K]; G.u[q][1] = ed.uy[K]; G.u[q][2] = ed.uz[K];
        G.ell[q] = ed.ell[K];
        G.id[q] = K;
        G.i1[q] = ed.i1[K];
        G.i2[q] = ed.i2[K];
        G.own[q] = owner[K];
    }
}

// Tile pairs of which a tile touches more than RC_RMAX edges (dense
// networks: a 32-vertex tile of a 3-D grid touches ~100): the same pair
// weights, row-sum partials and gather as k_gram_tiles, over every pair of
// RC_RMAX-edge chunks of the two tiles' edge lists. Each thread keeps its
// four A entries in registers across the chunk pairs (the lower-index
// vertex's incidences always in the U role, so both orientations of a
// diagonal-tile entry add the same terms in the same order: A stays exactly
// symmetric) and writes them once; the row-sum slots are accumulated in
// place by their single writer (first chunk stores, later ones add).
// Incidences come from global memory (their count per vertex is unbounded).
template <class P>
__global__ __launch_bounds__(256) void k_gram_tiles_wide(
    EdgeSoA ed, P pw, double eps, int V, int E, int ngram, const int32_t* __restrict__ rOff,
    const int32_t* __restrict__ rEdge, const int32_t* __restrict__ owner,
    const int32_t* __restrict__ vOff, const int32_t* __restrict__ vSlot,
    const int32_t* __restrict__ vLocal, double* __restrict__ A, int ldA,
    double* __restrict__ rsp) {
    __shared__ GEdge GU, GV;
    __shared__ double Xt[RC_RMAX][RC_RMAX + 1];
    __shared__ double Yt[RC_RMAX][RC_RMAX + 1];
    __shared__ double Wt[RC_RMAX][RC_RMAX + 1];
    const int t = threadIdx.x;
    for (int p = blockIdx.x; p < ngram; p += gridDim.x) {
        int U, W;
        tri_decode(p, U, W);
        const int nUt = rOff[U + 1] - rOff[U], nWt = rOff[W + 1] - rOff[W];
        if (nUt <= RC_RMAX && nWt <= RC_RMAX) continue;   // k_gram_tiles' pair
        const bool diag = (U == W);
        double acc[RC_VTILE * RC_VTILE / 256];
#pragma unroll
        for (int j = 0; j < RC_VTILE * RC_VTILE / 256; ++j) acc[j] = 0.0;
        // an edgeless tile still writes its row-sum slots (zero): one empty chunk
        for (int cu0 = 0; cu0 < (nUt > 0 ? nUt : 1); cu0 += RC_RMAX) {
            const int nu = min(RC_RMAX, nUt - cu0);
            for (int cw0 = 0; cw0 < (nWt > 0 ? nWt : 1); cw0 += RC_RMAX) {
                const int nv = min(RC_RMAX, nWt - cw0);
                __syncthreads();   // the previous chunk pair's LDS reads are done
                load_edge_chunk(GU, rOff, rEdge, U, cu0, nu, ed, owner);
                load_edge_chunk(GV, rOff, rEdge, W, cw0, nv, ed, owner);
                __syncthreads();
                for (int q = t; q < nu * nv; q += blockDim.x) {
                    const int li = q / nv, lj = q - (q / nv) * nv;
                    const int I = GU.id[li], J = GV.id[lj];
                    const bool dis = GU.i1[li] != GV.i1[lj] && GU.i1[li] != GV.i2[lj] &&
                                     GU.i2[li] != GV.i1[lj] && GU.i2[li] != GV.i2[lj];
                    double w2 = 0.0, X = 0.0, Y = 0.0;
                    if (dis) {
                        const bool ord = I <= J;
                        pair_weights(pw, eps, ord ? GU.P[li] : GV.P[lj], ord ? GU.u[li] : GV.u[lj],
                                     ord ? GU.ell[li] : GV.ell[lj], ord ? GV.P[lj] : GU.P[li],
                                     ord ? GV.u[lj] : GU.u[li], ord ? GV.ell[lj] : GU.ell[li], w2,
                                     X, Y);
                    }
                    Xt[li][lj] = X;
                    Yt[li][lj] = Y;
                    Wt[li][lj] = w2;
                }
                __syncthreads();
                // row-sum partials of this chunk pair (as k_gram_tiles)
                for (int side = 0; side < (diag ? 1 : 2); ++side) {
                    const int nr = side ? nv : nu, nc = side ? nu : nv;
                    const int row = t >> 2, part = t & 3;
                    double sw = 0.0, sy = 0.0;
                    bool live = false;
                    if (row < nr) {
                        live = side ? (GV.own[row] == W) : (GU.own[row] == U);
                        for (int c = part; c < nc; c += 4) {
                            const bool on = side ? (GU.own[c] == U) : (GV.own[c] == W);
                            const double wv = side ? Wt[c][row] : Wt[row][c];
                            const double yv = side ? Yt[c][row] : Yt[row][c];
                            sw += on ? wv : 0.0;
                            sy += on ? yv : 0.0;
                        }
                    }
                    sw += __shfl_xor(sw, 1, 64);
                    sy += __shfl_xor(sy, 1, 64);
                    sw += __shfl_xor(sw, 2, 64);
                    sy += __shfl_xor(sy, 2, 64);
                    if (row < nr && part == 0 && live) {
                        const int K = side ? GV.id[row] : GU.id[row];
                        const int other = side ? U : W;
                        const bool first = side ? (cu0 == 0) : (cw0 == 0);
                        double* s0 = rsp + ((size_t)other * 2) * E + K;
                        double* s1 = rsp + ((size_t)other * 2 + 1) * E + K;
                        *s0 = first ? sw : *s0 + sw;
                        *s1 = first ? sy : *s1 + sy;
                    }
                }
                // gather this chunk pair's terms into the thread's entries
#pragma unroll
                for (int j = 0; j < RC_VTILE * RC_VTILE / 256; ++j) {
                    const int q = t + 256 * j;
                    const int u = U * RC_VTILE + q / RC_VTILE, v = W * RC_VTILE + q % RC_VTILE;
                    if (u >= V || v >= V) continue;
                    const bool swap = diag && v < u;
                    const int lo = swap ? v : u, hi = swap ? u : v;
                    double val = acc[j];
                    for (int q1 = vOff[lo]; q1 < vOff[lo + 1]; ++q1) {
                        const int li = vLocal[q1] - cu0;
                        if (li < 0 || li >= nu) continue;
                        const double si = vSlot[q1] ? -1.0 : 1.0;
                        for (int q2 = vOff[hi]; q2 < vOff[hi + 1]; ++q2) {
                            const int lj = vLocal[q2] - cw0;
                            if (lj < 0 || lj >= nv) continue;
                            const double sj = vSlot[q2] ? -1.0 : 1.0;
                            val -= si * sj * Xt[li][lj] + Yt[li][lj];
                        }
                    }
                    acc[j] = val;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < RC_VTILE * RC_VTILE / 256; ++j) {
            const int q = t + 256 * j;
            const int u = U * RC_VTILE + q / RC_VTILE, v = W * RC_VTILE + q % RC_VTILE;
            if (u >= V || v >= V) continue;
            A[(size_t)u * ldA + v] = acc[j];
            if (!diag) A[(size_t)v * ldA + u] = acc[j];
        }
        __syncthreads();
    }
}

// Sums in tile order t = 0 .. nvt-1 (fixed order); the loads of 16 tiles are
// issued ahead of their adds, and 64-thread blocks spread the edges over
// E/64 workgroups.
__global__ __launch_bounds__(64) void k_rowsum_reduce(const double* __restrict__ rsp, int nvt,
                                                      int E, double* __restrict__ rs) {
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    if (K >= E) return;
    double sw = 0.0, sy = 0.0;
    int t = 0;
    for (; t + 16 <= nvt; t += 16) {
        double a[16], b[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            a[q] = rsp[((size_t)(t + q) * 2) * E + K];
            b[q] = rsp[((size_t)(t + q) * 2 + 1) * E + K];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            sw += a[q];
            sy += b[q];
        }
    }
    for (; t < nvt; ++t) {
        sw += rsp[((size_t)t * 2) * E + K];
        sy += rsp[((size_t)t * 2 + 1) * E + K];
    }
    rs[K] = sw;
    rs[E + K] = sy;
}

// Edge-own terms: sum_K W_K g_K g_K^T + Q_K h_K h_K^T touches (u,u) and the
// adjacent (u,v). One thread per vertex; each off-diagonal entry receives the
// same value from both of its vertices' threads (exact symmetry).
__global__ void k_gram_band(EdgeSoA ed, const double* __restrict__ rs, int V, int E,
                            const int32_t* __restrict__ vOff, const int32_t* __restrict__ vEdge,
                            double* __restrict__ A, int ldA) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= V) return;
    double dsum = 0.0;
    for (int q = vOff[u]; q < vOff[u + 1]; ++q) {
        const int K = vEdge[q];
        const double l = ed.ell[K];
        const double wk = rs[K] / (l * l);
        const double qk = rs[E + K];
        dsum += wk + qk;
        const int v = (ed.i1[K] == u) ? ed.i2[K] : ed.i1[K];
        A[(size_t)u * ldA + v] += qk - wk;
    }
    A[(size_t)u * ldA + u] += dsum;
}

}  // namespace

rc_status rc_launch_gram(rc_ctx* ctx, const rc_params* p, hipStream_t st) {
    const int grid = ctx->ngram < 8192 ? ctx->ngram : 8192;
    if (p->alpha == 3.0 && p->beta == 6.0) {
        hipLaunchKernelGGL(k_gram_tiles<Pow36>, dim3(grid), dim3(256), 0, st, ctx->ed,
                           Pow36{}, p->epsilon, ctx->V, ctx->E, ctx->nvt, ctx->ngram, ctx->d_rOff,
                           ctx->d_rEdge, ctx->d_owner, ctx->d_vOff, ctx->d_vSlot, ctx->d_vLocal,
                           ctx->d_F, ctx->nblk * RC_NB, ctx->d_rsp);
    } else {
        const double s = (p->beta - 1) / p->alpha;
        PowGeneric pw{p->alpha, p->beta, 2 * (s - 1) + 1};
        hipLaunchKernelGGL(k_gram_tiles<PowGeneric>, dim3(grid), dim3(256), 0, st,
                           ctx->ed, pw, p->epsilon, ctx->V, ctx->E, ctx->nvt, ctx->ngram,
                           ctx->d_rOff, ctx->d_rEdge, ctx->d_owner, ctx->d_vOff, ctx->d_vSlot,
                           ctx->d_vLocal, ctx->d_F, ctx->nblk * RC_NB, ctx->d_rsp);
    }
    RC_CHECK_LAUNCH();
    if (ctx->wide_tiles) {
        if (p->alpha == 3.0 && p->beta == 6.0) {
            hipLaunchKernelGGL(k_gram_tiles_wide<Pow36>, dim3(grid), dim3(256), 0, st, ctx->ed,
                               Pow36{}, p->epsilon, ctx->V, ctx->E, ctx->ngram, ctx->d_rOff,
                               ctx->d_rEdge, ctx->d_owner, ctx->d_vOff, ctx->d_vSlot,
                               ctx->d_vLocal, ctx->d_F, ctx->nblk * RC_NB, ctx->d_rsp);
        } else {
            const double s = (p->beta - 1) / p->alpha;
            PowGeneric pw{p->alpha, p->beta, 2 * (s - 1) + 1};
            hipLaunchKernelGGL(k_gram_tiles_wide<PowGeneric>, dim3(grid), dim3(256), 0, st,
                               ctx->ed, pw, p->epsilon, ctx->V, ctx->E, ctx->ngram, ctx->d_rOff,
                               ctx->d_rEdge, ctx->d_owner, ctx->d_vOff, ctx->d_vSlot,
                               ctx->d_vLocal, ctx->d_F, ctx->nblk * RC_NB, ctx->d_rsp);
        }
        RC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_rowsum_reduce, dim3((ctx->E + 63) / 64), dim3(64), 0, st,
                       ctx->d_rsp, ctx->nvt, ctx->E, ctx->d_rs);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gram_band, dim3((ctx->V + 255) / 256), dim3(256), 0, st, ctx->ed,
                       ctx->d_rs, ctx->V, ctx->E, ctx->d_vOff, ctx->d_vEdge, ctx->d_F,
                       ctx->nblk * RC_NB);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

rc_status rc_launch_gram(rc_ctx* ctx, const rc_params* p) { return rc_launch_gram(ctx, p, ctx->stream); }

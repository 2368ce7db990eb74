// Barnes-Hut tangent-point energy over an edge BVH.
//
// Reference: the stage-2 oracle's bh_energy, build_edge_bvh and
// bh_ordered_cluster_term (oracle/tpe_stage2_oracle.py:456-596). The tree is
// built on the host per call (the curve changes every call): each node keeps
// its edges contiguous in a node-order array, its mass-weighted centre (mass
// = ell + eps), the radii rx of the edge midpoints and rT of the unit
// tangents around their mass-weighted means; a node of more than leaf_size
// edges is split at the median of the 6-D coordinate (midpoint, tangent) with
// the largest spread after a stable sort on it. The traversal runs on the
// device, one lane per target edge I with an explicit stack (left child
// first): a node of >= 2 edges none of which shares a vertex with I (tested
// through the vertex -> edge table: an edge through an endpoint of I lies in
// the node iff its node-order position is in the node's range), at a
// positive distance with rx / dist <= theta and rT <= theta_t, contributes
// one cluster term with its centre as the source point and its mass; a leaf
// otherwise adds the exact ordered terms of its disjoint edges. E = 1/2 sum_I.
#include <algorithm>
#include <cmath>
#include <vector>

#include "rc_internal.h"
#include "rc_math.h"

namespace {

constexpr int BH_STACK = 64;   // > depth + 1 of a median-split tree over 2^31 edges

struct BhTree {
    std::vector<int32_t> ord, pos;               // node order and its inverse
    std::vector<int32_t> start, end, left, right;
    std::vector<double> cx, cy, cz, mass, rx, rT;
};

struct BhBuild {
    const std::vector<double>& C;   // midpoints [E][3]
    const std::vector<double>& T;   // unit tangents [E][3]
    const std::vector<double>& m;   // ell + eps
    std::vector<double> key;
    int leaf;
    BhTree& t;

    int build(int s, int e) {
        const int n = e - s;
        const int id = (int)t.start.size();
        t.start.push_back(s);
        t.end.push_back(e);
        t.left.push_back(-1);
        t.right.push_back(-1);
        double ms = 0.0, c[3] = {0, 0, 0}, tg[3] = {0, 0, 0};
        for (int p = s; p < e; ++p) ms += m[t.ord[p]];
        if (ms > 0) {
            for (int p = s; p < e; ++p) {
                const int k = t.ord[p];
                for (int d = 0; d < 3; ++d) {
                    c[d] += m[k] * C[3 * k + d];
                    tg[d] += m[k] * T[3 * k + d];
                }
            }
            for (int d = 0; d < 3; ++d) {
                c[d] /= ms;
                tg[d] /= ms;
            }
        } else if (n > 0) {
            for (int p = s; p < e; ++p)
                for (int d = 0; d < 3; ++d) {
                    c[d] += C[3 * t.ord[p] + d];
                    tg[d] += T[3 * t.ord[p] + d];
                }
            for (int d = 0; d < 3; ++d) {
                c[d] /= n;
                tg[d] /= n;
            }
        }
        double rx = 0.0, rT = 0.0;
        for (int p = s; p < e; ++p) {
            const int k = t.ord[p];
            double sx = 0.0, st = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double u = C[3 * k + d] - c[d], w = T[3 * k + d] - tg[d];
                sx += u * u;
                st += w * w;
            }
            rx = std::max(rx, std::sqrt(sx));
            rT = std::max(rT, std::sqrt(st));
        }
        t.cx.push_back(c[0]);
        t.cy.push_back(c[1]);
        t.cz.push_back(c[2]);
        t.mass.push_back(ms);
        t.rx.push_back(rx);
        t.rT.push_back(rT);
        if (n > leaf) {
            int dim = 0;
            double best = -1.0;
            for (int d = 0; d < 6; ++d) {
                double lo = INFINITY, hi = -INFINITY;
                for (int p = s; p < e; ++p) {
                    const int k = t.ord[p];
                    const double x = d < 3 ? C[3 * k + d] : T[3 * k + d - 3];
                    lo = std::min(lo, x);
                    hi = std::max(hi, x);
                }
                if (hi - lo > best) {
                    best = hi - lo;
                    dim = d;
                }
            }
            for (int p = s; p < e; ++p) {
                const int k = t.ord[p];
                key[k] = dim < 3 ? C[3 * k + dim] : T[3 * k + dim - 3];
            }
            std::stable_sort(t.ord.begin() + s, t.ord.begin() + e,
                             [&](int a, int b) { return key[a] < key[b]; });
            const int mid = n / 2;
            if (mid > 0 && n - mid > 0) {
                const int l = build(s, s + mid);
                const int r = build(s + mid, e);
                t.left[id] = l;
                t.right[id] = r;
            }
        }
        return id;
    }
};

struct BhDev {
    const double *cx, *cy, *cz, *mass, *rx, *rT;
    const int32_t *start, *end, *left, *right, *ord, *pos;
};

// Ordered exact term of edge I against J (tangentPointEnergy.ts:60-105 for one
// order): 1/4 ell_I^(1-alpha) ell_J sum_ij (|e_I x (p_i - q_j)| + eps)^alpha
// / (|p_i - q_j| + eps)^beta.
template <class P>
__device__ __forceinline__ double bh_exact(const P& pw, double eps, const double* PI,
                                           const double* eI, double lpI, const double* PJ,
                                           double ellJ) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const double dx = PI[3 * a] - PJ[3 * b], dy = PI[3 * a + 1] - PJ[3 * b + 1],
                         dz = PI[3 * a + 2] - PJ[3 * b + 2];
            const double dn = rcm::sqrt_(dx * dx + dy * dy + dz * dz) + eps;
            const double cx = eI[1] * dz - eI[2] * dy, cy = eI[2] * dx - eI[0] * dz,
                         cz = eI[0] * dy - eI[1] * dx;
            const double cn = rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps;
            s += pw.c_alpha(cn) * pw.inv_beta(dn, rcm::rcp(dn));
        }
    return 0.25 * lpI * ellJ * s;
}

// Cluster term (bh_ordered_cluster_term): the node's centre q stands for the
// two endpoints of every source edge, so 1/2 ell_I^(1-alpha) mass sum_i.
template <class P>
__device__ __forceinline__ double bh_cluster(const P& pw, double eps, const double* PI,
                                             const double* eI, double lpI, double qx, double qy,
                                             double qz, double mass) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const double dx = PI[3 * a] - qx, dy = PI[3 * a + 1] - qy, dz = PI[3 * a + 2] - qz;
        const double dn = rcm::sqrt_(dx * dx + dy * dy + dz * dz) + eps;
        const double cx = eI[1] * dz - eI[2] * dy, cy = eI[2] * dx - eI[0] * dz,
                     cz = eI[0] * dy - eI[1] * dx;
        const double cn = rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps;
        s += pw.c_alpha(cn) * pw.inv_beta(dn, rcm::rcp(dn));
    }
    return 0.5 * lpI * mass * s;
}

template <class P>
__global__ __launch_bounds__(256) void k_energy_bh(EdgeSoA ed, int E, BhDev t,
                                                   const int32_t* __restrict__ vOff,
                                                   const int32_t* __restrict__ vEdge, P pw,
                                                   double eps, double theta, double theta_t,
                                                   double* __restrict__ part,
                                                   unsigned long long* __restrict__ stats) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= E) return;
    const int i1 = ed.i1[I], i2 = ed.i2[I];
    const double PI[6] = {ed.ax[I], ed.ay[I], ed.az[I], ed.bx[I], ed.by[I], ed.bz[I]};
    const double eI[3] = {PI[3] - PI[0], PI[4] - PI[1], PI[5] - PI[2]};
    const double mx = 0.5 * (PI[0] + PI[3]), my = 0.5 * (PI[1] + PI[4]),
                 mz = 0.5 * (PI[2] + PI[5]);
    const double lpI = ed.lpow[I];
    int stack[BH_STACK];
    int sp = 0;
    stack[sp++] = 0;
    double sum = 0.0;
    unsigned long long visits = 0, approx = 0, direct = 0;
    while (sp > 0) {
        const int ni = stack[--sp];
        ++visits;
        const int s = t.start[ni], e = t.end[ni];
        const int n = e - s;
        if (n == 0) continue;
        const double dx = mx - t.cx[ni], dy = my - t.cy[ni], dz = mz - t.cz[ni];
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        bool admiss = n >= 2 && dist > 0.0 && t.rx[ni] / dist <= theta && t.rT[ni] <= theta_t;
        if (admiss) {
            // no edge through an endpoint of I (I included) in the node
            for (int q = vOff[i1]; q < vOff[i1 + 1] && admiss; ++q) {
                const int p = t.pos[vEdge[q]];
                if (p >= s && p < e) admiss = false;
            }
            for (int q = vOff[i2]; q < vOff[i2 + 1] && admiss; ++q) {
                const int p = t.pos[vEdge[q]];
                if (p >= s && p < e) admiss = false;
            }
        }
        if (admiss) {
            approx += n;
            sum += bh_cluster(pw, eps, PI, eI, lpI, t.cx[ni], t.cy[ni], t.cz[ni], t.mass[ni]);
            continue;
        }
        if (t.left[ni] < 0) {
            for (int p = s; p < e; ++p) {
                const int J = t.ord[p];
                const int j1 = ed.i1[J], j2 = ed.i2[J];
                if (i1 == j1 || i1 == j2 || i2 == j1 || i2 == j2) continue;
                ++direct;
                const double PJ[6] = {ed.ax[J], ed.ay[J], ed.az[J], ed.bx[J], ed.by[J], ed.bz[J]};
                sum += bh_exact(pw, eps, PI, eI, lpI, PJ, ed.ell[J]);
            }
            continue;
        }
        if (sp + 2 > BH_STACK) {   // cannot happen for a median split; counted, not faulted
            atomicAdd(stats + 3, 1ull);
            continue;
        }
        stack[sp++] = t.right[ni];
        stack[sp++] = t.left[ni];
    }
    part[I] = sum;
    atomicAdd(stats + 0, visits);
    atomicAdd(stats + 1, approx);
    atomicAdd(stats + 2, direct);
}

// Fixed-order sum of the per-edge partials, times 1/2.
__global__ __launch_bounds__(1024) void k_bh_reduce(const double* __restrict__ part, int n,
                                                    double* __restrict__ out) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = 0.5 * sh[0];
}

}  // namespace

extern "C" rc_status rc_energy_bh(rc_ctx* ctx, const double* xyz, const rc_params* p,
                                  double theta, double theta_t, int32_t leaf_size,
                                  double* energy, int64_t* stats) {
    if (!ctx || !xyz || !p || !energy) return RC_BAD_ARG;
    if (!std::isfinite(p->alpha) || !std::isfinite(p->beta) || !std::isfinite(p->epsilon) ||
        !(theta >= 0.0) || !std::isfinite(theta) || !(theta_t >= 0.0) ||
        !std::isfinite(theta_t) || leaf_size < 1) {
        ctx->err = "Barnes-Hut parameters out of range";
        return RC_BAD_ARG;
    }
    const int V = ctx->V, E = ctx->E;
    // host geometry for the tree (midpoints, unit tangents, masses)
    std::vector<double> C(3 * (size_t)E), T(3 * (size_t)E), m(E);
    for (int K = 0; K < E; ++K) {
        const int a = ctx->h_edges[2 * K], b = ctx->h_edges[2 * K + 1];
        double e[3], r2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            e[d] = xyz[3 * b + d] - xyz[3 * a + d];
            r2 += e[d] * e[d];
            C[3 * K + d] = 0.5 * (xyz[3 * a + d] + xyz[3 * b + d]);
        }
        const double r = std::sqrt(r2);
        m[K] = r + p->epsilon;
        for (int d = 0; d < 3; ++d) T[3 * K + d] = (r >= 1e-14) ? e[d] / r : 0.0;
    }
    BhTree tree;
    tree.ord.resize(E);
    for (int K = 0; K < E; ++K) tree.ord[K] = K;
    BhBuild bld{C, T, m, std::vector<double>(E), leaf_size, tree};
    bld.build(0, E);
    tree.pos.resize(E);
    for (int q = 0; q < E; ++q) tree.pos[tree.ord[q]] = q;
    const int nn = (int)tree.start.size();
    // device copies: 6 doubles and 4 ints per node, order and position per edge
    const size_t nd = 6 * (size_t)nn, ni = 4 * (size_t)nn + 2 * (size_t)E;
    if (nd > ctx->bh_cap_d || ni > ctx->bh_cap_i) {
        if (ctx->d_bh_f) (void)hipFree(ctx->d_bh_f);
        if (ctx->d_bh_i) (void)hipFree(ctx->d_bh_i);
        ctx->d_bh_f = nullptr;
        ctx->d_bh_i = nullptr;
        ctx->bh_cap_d = ctx->bh_cap_i = 0;
        RC_HIP(hipMalloc((void**)&ctx->d_bh_f, sizeof(double) * nd));
        RC_HIP(hipMalloc((void**)&ctx->d_bh_i, sizeof(int32_t) * ni));
        if (!ctx->d_bh_stats) RC_HIP(hipMalloc((void**)&ctx->d_bh_stats, sizeof(unsigned long long) * 4));
        ctx->bh_cap_d = nd;
        ctx->bh_cap_i = ni;
    }
    std::vector<double> hf(nd);
    std::vector<int32_t> hi(ni);
    const std::vector<double>* fs[6] = {&tree.cx, &tree.cy, &tree.cz, &tree.mass, &tree.rx, &tree.rT};
    for (int f = 0; f < 6; ++f) std::copy(fs[f]->begin(), fs[f]->end(), hf.begin() + (size_t)f * nn);
    const std::vector<int32_t>* is[4] = {&tree.start, &tree.end, &tree.left, &tree.right};
    for (int f = 0; f < 4; ++f) std::copy(is[f]->begin(), is[f]->end(), hi.begin() + (size_t)f * nn);
    std::copy(tree.ord.begin(), tree.ord.end(), hi.begin() + 4 * (size_t)nn);
    std::copy(tree.pos.begin(), tree.pos.end(), hi.begin() + 4 * (size_t)nn + E);
    RC_HIP(hipMemcpyAsync(ctx->d_bh_f, hf.data(), sizeof(double) * nd, hipMemcpyHostToDevice,
                          ctx->stream));
    RC_HIP(hipMemcpyAsync(ctx->d_bh_i, hi.data(), sizeof(int32_t) * ni, hipMemcpyHostToDevice,
                          ctx->stream));
    RC_HIP(hipMemsetAsync(ctx->d_bh_stats, 0, sizeof(unsigned long long) * 4, ctx->stream));
    rc_status s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, p);
    if (s != RC_OK) return s;
    BhDev t;
    const double* f = ctx->d_bh_f;
    t.cx = f; t.cy = f + nn; t.cz = f + 2 * (size_t)nn; t.mass = f + 3 * (size_t)nn;
    t.rx = f + 4 * (size_t)nn; t.rT = f + 5 * (size_t)nn;
    const int32_t* g = ctx->d_bh_i;
    t.start = g; t.end = g + nn; t.left = g + 2 * (size_t)nn; t.right = g + 3 * (size_t)nn;
    t.ord = g + 4 * (size_t)nn; t.pos = g + 4 * (size_t)nn + E;
    double* part = ctx->d_gpart;   // >= E doubles
    const dim3 grid((E + 255) / 256);
    if (p->alpha == 3.0 && p->beta == 6.0) {
        hipLaunchKernelGGL(k_energy_bh<Pow36>, grid, dim3(256), 0, ctx->stream, ctx->ed, E, t,
                           ctx->d_vOff, ctx->d_vEdge, Pow36{}, p->epsilon, theta, theta_t, part,
                           ctx->d_bh_stats);
    } else {
        const double sg = (p->beta - 1) / p->alpha;
        PowGeneric pw{p->alpha, p->beta, 2 * (sg - 1) + 1};
        hipLaunchKernelGGL(k_energy_bh<PowGeneric>, grid, dim3(256), 0, ctx->stream, ctx->ed, E,
                           t, ctx->d_vOff, ctx->d_vEdge, pw, p->epsilon, theta, theta_t, part,
                           ctx->d_bh_stats);
    }
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bh_reduce, dim3(1), dim3(1024), 0, ctx->stream, part, E, ctx->d_scal);
    RC_CHECK_LAUNCH();
    RC_HIP(hipMemcpyAsync(ctx->d_host_stage, ctx->d_scal, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    unsigned long long* hst = (unsigned long long*)(ctx->d_host_stage + 8);
    RC_HIP(hipMemcpyAsync(hst, ctx->d_bh_stats, sizeof(unsigned long long) * 4,
                          hipMemcpyDeviceToHost, ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    if (hst[3]) {
        ctx->err = "Barnes-Hut traversal stack exceeded";
        return RC_HIP_ERROR;
    }
    (void)V;
    *energy = ctx->d_host_stage[0];
    if (stats)
        for (int q = 0; q < 3; ++q) stats[q] = (int64_t)hst[q];
    return RC_OK;
}

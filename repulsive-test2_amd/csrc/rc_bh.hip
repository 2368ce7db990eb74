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
constexpr int BH_FD = 3;       // work items: the subtrees at this depth

struct BhTree {
    std::vector<int32_t> ord, pos;               // node order and its inverse
    std::vector<int32_t> start, end, left, right;
    std::vector<double> cx, cy, cz, tx, ty, tz, mass, rx, rT;
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
        t.tx.push_back(tg[0]);
        t.ty.push_back(tg[1]);
        t.tz.push_back(tg[2]);
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

// Work item (f, I): target edge I against frontier node f (the nodes of depth
// BH_FD, and leaves above that depth, in left-first order). The nodes above f are
// walked first, as the traversal from the root meets them: each counts its
// visit in the item that is its leftmost frontier descendant, and the first
// admissible one contributes its cluster term in that item and ends every
// item below it. Otherwise the item traverses f's subtree with an explicit
// stack (left child first). The per-item sums are added per edge in frontier
// order.
struct BhFront {
    const int32_t* node;    // frontier node
    const int32_t* anc;     // [BH_FD] ancestors from the root down, -1 past the last
    const int32_t* first;   // [BH_FD] 1 if this item is the ancestor's leftmost frontier descendant
    int nf;
};

template <class P>
__device__ __forceinline__ bool bh_admissible(const BhDev& t, int ni, int i1, int i2,
                                              const int32_t* __restrict__ vOff,
                                              const int32_t* __restrict__ vEdge, double mx,
                                              double my, double mz, double theta, double theta_t) {
    const int s = t.start[ni], e = t.end[ni];
    if (e - s < 2) return false;
    const double dx = mx - t.cx[ni], dy = my - t.cy[ni], dz = mz - t.cz[ni];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(dist > 0.0 && t.rx[ni] / dist <= theta && t.rT[ni] <= theta_t)) return false;
    for (int q = vOff[i1]; q < vOff[i1 + 1]; ++q) {
        const int p = t.pos[vEdge[q]];
        if (p >= s && p < e) return false;
    }
    for (int q = vOff[i2]; q < vOff[i2 + 1]; ++q) {
        const int p = t.pos[vEdge[q]];
        if (p >= s && p < e) return false;
    }
    return true;
}

template <class P>
__global__ __launch_bounds__(256) void k_energy_bh(EdgeSoA ed, int E, BhDev t, BhFront fr,
                                                   const int32_t* __restrict__ vOff,
                                                   const int32_t* __restrict__ vEdge, P pw,
                                                   double eps, double theta, double theta_t,
                                                   double* __restrict__ part,
                                                   unsigned long long* __restrict__ stats) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= fr.nf * E) return;
    const int f = gid / E, I = gid - f * E;   // a wave: 64 consecutive edges, one subtree
    const int i1 = ed.i1[I], i2 = ed.i2[I];
    const double PI[6] = {ed.ax[I], ed.ay[I], ed.az[I], ed.bx[I], ed.by[I], ed.bz[I]};
    const double eI[3] = {PI[3] - PI[0], PI[4] - PI[1], PI[5] - PI[2]};
    const double mx = 0.5 * (PI[0] + PI[3]), my = 0.5 * (PI[1] + PI[4]),
                 mz = 0.5 * (PI[2] + PI[5]);
    const double lpI = ed.lpow[I];
    double sum = 0.0;
    unsigned long long visits = 0, approx = 0, direct = 0;
    bool stop = false;
    for (int k = 0; k < BH_FD && !stop; ++k) {
        const int a = fr.anc[BH_FD * f + k];
        if (a < 0) break;
        const bool first = fr.first[BH_FD * f + k] != 0;
        if (first) ++visits;
        if (bh_admissible<P>(t, a, i1, i2, vOff, vEdge, mx, my, mz, theta, theta_t)) {
            if (first) {
                approx += t.end[a] - t.start[a];
                sum += bh_cluster(pw, eps, PI, eI, lpI, t.cx[a], t.cy[a], t.cz[a], t.mass[a]);
            }
            stop = true;
        }
    }
    if (!stop) {
        int stack[BH_STACK];
        int sp = 0;
        stack[sp++] = fr.node[f];
        while (sp > 0) {
            const int ni = stack[--sp];
            ++visits;
            const int s = t.start[ni], e = t.end[ni];
            if (e - s == 0) continue;
            if (bh_admissible<P>(t, ni, i1, i2, vOff, vEdge, mx, my, mz, theta, theta_t)) {
                approx += e - s;
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
    }
    part[(size_t)f * E + I] = sum;
    atomicAdd(stats + 0, visits);
    atomicAdd(stats + 1, approx);
    atomicAdd(stats + 2, direct);
}

// Per edge, the item sums in frontier order.
__global__ void k_bh_items(const double* __restrict__ part, int E, int nf,
                           double* __restrict__ out) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= E) return;
    double s = 0.0;
    for (int f = 0; f < nf; ++f) s += part[(size_t)f * E + I];
    out[I] = s;
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
    // frontier: nodes of depth BH_FD and leaves above it, left-first, with
    // their ancestors and whether the item is each ancestor's leftmost
    // descendant
    std::vector<int32_t> fnode, fanc, ffirst;
    {
        struct It { int n, d; int a[BH_FD]; };
        It root{0, 0, {}};
        for (int k = 0; k < BH_FD; ++k) root.a[k] = -1;
        std::vector<It> st{root};
        while (!st.empty()) {
            const It it = st.back();
            st.pop_back();
            if (it.d == BH_FD || tree.left[it.n] < 0) {
                fnode.push_back(it.n);
                for (int k = 0; k < BH_FD; ++k) {
                    fanc.push_back(it.a[k]);
                    ffirst.push_back(0);
                }
                continue;
            }
            It ch = it;
            ch.a[it.d] = it.n;
            ch.d = it.d + 1;
            ch.n = tree.right[it.n];
            st.push_back(ch);
            ch.n = tree.left[it.n];
            st.push_back(ch);
        }
        // leftmost frontier descendant of each ancestor: its first item in order
        std::vector<char> seen(nn, 0);
        for (size_t f = 0; f < fnode.size(); ++f)
            for (int k = 0; k < BH_FD; ++k) {
                const int a = fanc[BH_FD * f + k];
                if (a >= 0 && !seen[a]) {
                    seen[a] = 1;
                    ffirst[BH_FD * f + k] = 1;
                }
            }
    }
    const int nf = (int)fnode.size();
    // device copies of the frontier: after the tree arrays in the int buffer
    const size_t nfr = (1 + 2 * (size_t)BH_FD) * nf;
    if (ctx->bh_front_cap < nfr) {
        if (ctx->d_bh_front) (void)hipFree(ctx->d_bh_front);
        ctx->d_bh_front = nullptr;
        ctx->bh_front_cap = 0;
        RC_HIP(hipMalloc((void**)&ctx->d_bh_front, sizeof(int32_t) * nfr));
        ctx->bh_front_cap = nfr;
    }
    std::vector<int32_t> hfr;
    hfr.reserve(nfr);
    hfr.insert(hfr.end(), fnode.begin(), fnode.end());
    hfr.insert(hfr.end(), fanc.begin(), fanc.end());
    hfr.insert(hfr.end(), ffirst.begin(), ffirst.end());
    RC_HIP(hipMemcpyAsync(ctx->d_bh_front, hfr.data(), sizeof(int32_t) * hfr.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    BhFront fr{ctx->d_bh_front, ctx->d_bh_front + nf,
                ctx->d_bh_front + (1 + (size_t)BH_FD) * nf, nf};
    // item sums [nf][E] and per-edge sums [E] (nf <= 2^BH_FD)
    const size_t need = ((size_t)nf + 1) * E;
    double* part = ctx->d_gpart;
    if (need > (size_t)ctx->ntiles * 6 * ctx->Epad) {
        if (ctx->bh_part_cap < need) {
            if (ctx->d_bh_part) (void)hipFree(ctx->d_bh_part);
            ctx->d_bh_part = nullptr;
            ctx->bh_part_cap = 0;
            RC_HIP(hipMalloc((void**)&ctx->d_bh_part, sizeof(double) * need));
            ctx->bh_part_cap = need;
        }
        part = ctx->d_bh_part;
    }
    double* per_edge = part + (size_t)nf * E;
    const dim3 grid(((size_t)nf * E + 255) / 256);
    if (p->alpha == 3.0 && p->beta == 6.0) {
        hipLaunchKernelGGL(k_energy_bh<Pow36>, grid, dim3(256), 0, ctx->stream, ctx->ed, E, t, fr,
                           ctx->d_vOff, ctx->d_vEdge, Pow36{}, p->epsilon, theta, theta_t, part,
                           ctx->d_bh_stats);
    } else {
        const double sg = (p->beta - 1) / p->alpha;
        PowGeneric pw{p->alpha, p->beta, 2 * (sg - 1) + 1};
        hipLaunchKernelGGL(k_energy_bh<PowGeneric>, grid, dim3(256), 0, ctx->stream, ctx->ed, E,
                           t, fr, ctx->d_vOff, ctx->d_vEdge, pw, p->epsilon, theta, theta_t, part,
                           ctx->d_bh_stats);
    }
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bh_items, dim3((E + 255) / 256), dim3(256), 0, ctx->stream, part, E, nf,
                       per_edge);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bh_reduce, dim3(1), dim3(1024), 0, ctx->stream, (const double*)per_edge,
                       E, ctx->d_scal);
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

// ---------------------------------------------------------------------------
// Block-cluster-tree products of the edge kernels (bct_matvec_kernel,
// bct_kernel_node, exact_edge_kernel_pair, bct_apply_B_scalar:
// oracle/tpe_stage2_oracle.py:598-795). The host runs the dual recursion from
// (root, root) over the same tree and records the admissible blocks (rank-one
// terms with the node kernel of the centres and mean tangents, evaluated on
// the host) by target node and the direct leaf-leaf blocks by target leaf.
// The device forms the source moments sum_J mass_J psi_J per node, then each
// target edge I adds mass_I * sum kval * moment over the admissible blocks of
// the nodes on its leaf-to-root path and the exact pair kernels of its direct
// blocks. Kernels: 'high' k_IJ = 2 * 1/4 ell_I ell_J sum_ij d^-p, 'low'
// 1/4 ell_I ell_J sum_ij [(|T_I x d| + eps)^2 + (|T_J x -d| + eps)^2] d^-(4+p),
// d = |p_i - q_j| + eps, p = 7/3, disjoint pairs only.

namespace {

struct BctDev {
    const int32_t *start, *end, *parent, *leaf_of, *ord;
    const int32_t *adm_off, *adm_src, *dir_off, *dir_src;   // dir_src: source edges
    const double* kval;
};

// One wave per node: the lanes stride over the node's edges, then a
// fixed-order butterfly sum.
__global__ __launch_bounds__(256) void k_bct_moments(EdgeSoA ed, BctDev t, int nn,
                                                     const double* __restrict__ psi, int nc,
                                                     double* __restrict__ mom) {
    const int lane = threadIdx.x & 63;
    const int nd = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (nd >= nn) return;
    double s[3] = {0, 0, 0};
    for (int p = t.start[nd] + lane; p < t.end[nd]; p += 64) {
        const int J = t.ord[p];
        const double mj = ed.ell[J];
        for (int c = 0; c < nc; ++c) s[c] += mj * psi[(size_t)nc * J + c];
    }
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s[c] += __shfl_xor(s[c], off, 64);
    if (lane == 0)
        for (int c = 0; c < nc; ++c) mom[(size_t)nc * nd + c] = s[c];
}

// One wave per target edge I: the lanes stride over the admissible blocks of
// the nodes on I's leaf-to-root path and over the source edges of I's direct
// blocks (concatenated per target leaf on the host), then a fixed-order
// butterfly sum.
__global__ __launch_bounds__(256) void k_bct_apply(EdgeSoA ed, int E, BctDev t, int kind,
                                                   double eps, const double* __restrict__ psi,
                                                   int nc, const double* __restrict__ mom,
                                                   double* __restrict__ phi) {
    const int lane = threadIdx.x & 63;
    const int I = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (I >= E) return;
    double far[3] = {0, 0, 0}, near[3] = {0, 0, 0};
    for (int nd = t.leaf_of[I]; nd >= 0; nd = t.parent[nd])
        for (int q = t.adm_off[nd] + lane; q < t.adm_off[nd + 1]; q += 64) {
            const double kv = t.kval[q];
            const int sn = t.adm_src[q];
            for (int c = 0; c < nc; ++c) far[c] += kv * mom[(size_t)nc * sn + c];
        }
    const int i1 = ed.i1[I], i2 = ed.i2[I];
    const double PI[6] = {ed.ax[I], ed.ay[I], ed.az[I], ed.bx[I], ed.by[I], ed.bz[I]};
    const double TI[3] = {ed.ux[I], ed.uy[I], ed.uz[I]};
    const double ellI = ed.ell[I];
    const int leaf = t.leaf_of[I];
    for (int q = t.dir_off[leaf] + lane; q < t.dir_off[leaf + 1]; q += 64) {
        const int J = t.dir_src[q];
        const int j1 = ed.i1[J], j2 = ed.i2[J];
        if (i1 == j1 || i1 == j2 || i2 == j1 || i2 == j2) continue;
        const double PJ[6] = {ed.ax[J], ed.ay[J], ed.az[J], ed.bx[J], ed.by[J], ed.bz[J]};
        const double TJ[3] = {ed.ux[J], ed.uy[J], ed.uz[J]};
        double sh = 0.0, sl = 0.0;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const double dx = PI[3 * a] - PJ[3 * b], dy = PI[3 * a + 1] - PJ[3 * b + 1],
                             dz = PI[3 * a + 2] - PJ[3 * b + 2];
                const double dn = rcm::sqrt_(dx * dx + dy * dy + dz * dz) + eps;
                const double y = rcm::rcbrt(dn);
                const double y2 = y * y, y4 = y2 * y2;
                const double dp = y4 * y2 * y;   // dn^(-7/3)
                if (kind == 0) {
                    sh += dp;
                } else {
                    const double c1x = TI[1] * dz - TI[2] * dy, c1y = TI[2] * dx - TI[0] * dz,
                                 c1z = TI[0] * dy - TI[1] * dx;
                    const double c2x = -(TJ[1] * dz - TJ[2] * dy), c2y = -(TJ[2] * dx - TJ[0] * dz),
                                 c2z = -(TJ[0] * dy - TJ[1] * dx);
                    const double n1 = rcm::sqrt_(c1x * c1x + c1y * c1y + c1z * c1z) + eps;
                    const double n2 = rcm::sqrt_(c2x * c2x + c2y * c2y + c2z * c2z) + eps;
                    const double r = rcm::rcp(dn), r2 = r * r;
                    sl += (n1 * n1 + n2 * n2) * (r2 * r2) * dp;
                }
            }
        const double k = kind == 0 ? 2.0 * 0.25 * ellI * ed.ell[J] * sh
                                   : 0.25 * ellI * ed.ell[J] * sl;
        for (int c = 0; c < nc; ++c) near[c] += k * psi[(size_t)nc * J + c];
    }
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            far[c] += __shfl_xor(far[c], off, 64);
            near[c] += __shfl_xor(near[c], off, 64);
        }
    }
    if (lane == 0)
        for (int c = 0; c < nc; ++c) phi[(size_t)nc * I + c] = ellI * far[c] + near[c];
}

// y of bct_apply_B_scalar: 'high' ((u_b - u_a) / ell) T per edge (3 comps),
// 'low' (u_a + u_b) / 2; ones for K 1.
__global__ void k_bct_density(EdgeSoA ed, int E, int kind, const double* __restrict__ u,
                              double* __restrict__ y, double* __restrict__ ones) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= E) return;
    const double ua = u[ed.i1[I]], ub = u[ed.i2[I]];
    if (kind == 0) {
        const double du = (ub - ua) / ed.ell[I];
        y[3 * (size_t)I] = du * ed.ux[I];
        y[3 * (size_t)I + 1] = du * ed.uy[I];
        y[3 * (size_t)I + 2] = du * ed.uz[I];
    } else {
        y[I] = 0.5 * (ua + ub);
    }
    ones[I] = 1.0;
}

// out[v] from z = K1 y - Ky over the incident edges in edge order ('low':
// +z/2 at both ends; 'high': apply_Dt, coeff = (z . T) / ell, -coeff at i1,
// +coeff at i2).
__global__ void k_bct_gather(EdgeSoA ed, int V, int kind, const int32_t* __restrict__ vOff,
                             const int32_t* __restrict__ vEdge, const int32_t* __restrict__ vSlot,
                             const double* __restrict__ y, const double* __restrict__ K1,
                             const double* __restrict__ Ky, double* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    double s = 0.0;
    for (int q = vOff[v]; q < vOff[v + 1]; ++q) {
        const int I = vEdge[q];
        if (kind == 0) {
            const double T[3] = {ed.ux[I], ed.uy[I], ed.uz[I]};
            double dot = 0.0;
            for (int d = 0; d < 3; ++d) dot += (K1[I] * y[3 * (size_t)I + d] - Ky[3 * (size_t)I + d]) * T[d];
            const double coeff = dot / ed.ell[I];
            s += vSlot[q] ? coeff : -coeff;
        } else {
            s += 0.5 * (K1[I] * y[I] - Ky[I]);
        }
    }
    out[v] = s;
}

// Host: tree over xyz, the dual recursion's block lists and the counts.
struct BctHost {
    BhTree tree;
    std::vector<int32_t> parent, leaf_of, adm_off, adm_src, dir_off, dir_src;
    std::vector<double> kval;
    int64_t leaves = 0, adm = 0, dir = 0, pairs = 0;
};

static void bct_build_host(const rc_ctx* ctx, const double* xyz, double eps, double theta,
                           double theta_t, int leaf, int kind, BctHost& H) {
    const int E = ctx->E, V = ctx->V;
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
        m[K] = r + eps;
        for (int d = 0; d < 3; ++d) T[3 * K + d] = (r >= 1e-14) ? e[d] / r : 0.0;
    }
    BhTree& t = H.tree;
    t.ord.resize(E);
    for (int K = 0; K < E; ++K) t.ord[K] = K;
    BhBuild bld{C, T, m, std::vector<double>(E), leaf, t};
    if (E > 0) bld.build(0, E);
    t.pos.resize(E);
    for (int q = 0; q < E; ++q) t.pos[t.ord[q]] = q;
    const int nn = (int)t.start.size();
    H.parent.assign(nn, -1);
    for (int i = 0; i < nn; ++i)
        if (t.left[i] >= 0) {
            H.parent[t.left[i]] = i;
            H.parent[t.right[i]] = i;
        }
    H.leaf_of.assign(E, -1);
    for (int i = 0; i < nn; ++i)
        if (t.left[i] < 0)
            for (int p = t.start[i]; p < t.end[i]; ++p) H.leaf_of[t.ord[p]] = i;
    // vertex -> incident edges
    std::vector<int32_t> voff(V + 1, 0), vedge(2 * (size_t)E);
    for (int K = 0; K < E; ++K) {
        voff[ctx->h_edges[2 * K] + 1]++;
        voff[ctx->h_edges[2 * K + 1] + 1]++;
    }
    for (int v = 0; v < V; ++v) voff[v + 1] += voff[v];
    std::vector<int32_t> fill(voff.begin(), voff.end() - 1);
    for (int K = 0; K < E; ++K) {
        vedge[fill[ctx->h_edges[2 * K]]++] = K;
        vedge[fill[ctx->h_edges[2 * K + 1]]++] = K;
    }
    auto disjoint_blocks = [&](int ti, int si) {
        for (int p = t.start[ti]; p < t.end[ti]; ++p) {
            const int I = t.ord[p];
            for (int w = 0; w < 2; ++w) {
                const int v = ctx->h_edges[2 * I + w];
                for (int q = voff[v]; q < voff[v + 1]; ++q) {
                    const int pj = t.pos[vedge[q]];
                    if (pj >= t.start[si] && pj < t.end[si]) return false;
                }
            }
        }
        return true;
    };
    auto node_kernel = [&](int ti, int si) {
        const double d[3] = {t.cx[ti] - t.cx[si], t.cy[ti] - t.cy[si], t.cz[ti] - t.cz[si]};
        const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + eps;
        if (kind == 0) return 2.0 / std::pow(dn, 7.0 / 3.0);
        const double tT[3] = {t.tx[ti], t.ty[ti], t.tz[ti]}, tS[3] = {t.tx[si], t.ty[si], t.tz[si]};
        const double c1x = tT[1] * d[2] - tT[2] * d[1], c1y = tT[2] * d[0] - tT[0] * d[2],
                     c1z = tT[0] * d[1] - tT[1] * d[0];
        const double c2x = -(tS[1] * d[2] - tS[2] * d[1]), c2y = -(tS[2] * d[0] - tS[0] * d[2]),
                     c2z = -(tS[0] * d[1] - tS[1] * d[0]);
        const double c1 = std::sqrt(c1x * c1x + c1y * c1y + c1z * c1z) + eps;
        const double c2 = std::sqrt(c2x * c2x + c2y * c2y + c2z * c2z) + eps;
        return (std::pow(c1, 2.0) + std::pow(c2, 2.0)) / std::pow(dn, 4.0 + 7.0 / 3.0);
    };
    std::vector<std::pair<int, int>> adm, dir;   // (target, source)
    std::vector<double> kv;
    std::vector<std::pair<int, int>> stack;
    if (nn > 0) stack.push_back({0, 0});
    while (!stack.empty()) {   // same visiting order as the recursion (left first)
        const auto [ti, si] = stack.back();
        stack.pop_back();
        const int nt = t.end[ti] - t.start[ti], ns = t.end[si] - t.start[si];
        const double dx = t.cx[ti] - t.cx[si], dy = t.cy[ti] - t.cy[si], dz = t.cz[ti] - t.cz[si];
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        const bool well = nt + ns >= 3 && dist > 0.0 &&
                          std::max(t.rx[ti], t.rx[si]) / dist <= theta &&
                          std::max(t.rT[ti], t.rT[si]) <= theta_t && disjoint_blocks(ti, si);
        if (well) {
            ++H.leaves;
            ++H.adm;
            adm.push_back({ti, si});
            kv.push_back(node_kernel(ti, si));
            continue;
        }
        if (t.left[ti] < 0 && t.left[si] < 0) {
            ++H.leaves;
            ++H.dir;
            dir.push_back({ti, si});
            for (int p = t.start[ti]; p < t.end[ti]; ++p) {
                const int I = t.ord[p];
                const int a1 = ctx->h_edges[2 * I], a2 = ctx->h_edges[2 * I + 1];
                for (int q = t.start[si]; q < t.end[si]; ++q) {
                    const int J = t.ord[q];
                    const int b1 = ctx->h_edges[2 * J], b2 = ctx->h_edges[2 * J + 1];
                    if (a1 != b1 && a1 != b2 && a2 != b1 && a2 != b2) ++H.pairs;
                }
            }
            continue;
        }
        const bool split_t = t.left[ti] >= 0 && (t.left[si] < 0 || nt >= ns);
        if (split_t) {
            stack.push_back({t.right[ti], si});
            stack.push_back({t.left[ti], si});
        } else {
            stack.push_back({ti, t.right[si]});
            stack.push_back({ti, t.left[si]});
        }
    }
    // CSR by target node (stable: recursion order within a node)
    H.adm_off.assign(nn + 1, 0);
    for (const auto& b : adm) H.adm_off[b.first + 1]++;
    for (int i = 0; i < nn; ++i) H.adm_off[i + 1] += H.adm_off[i];
    H.adm_src.resize(adm.size());
    H.kval.resize(adm.size());
    {
        std::vector<int32_t> f(H.adm_off.begin(), H.adm_off.end() - 1);
        for (size_t q = 0; q < adm.size(); ++q) {
            const int at = f[adm[q].first]++;
            H.adm_src[at] = adm[q].second;
            H.kval[at] = kv[q];
        }
    }
    // the source edges of each target leaf's direct blocks, in block order
    H.dir_off.assign(nn + 1, 0);
    for (const auto& b : dir) H.dir_off[b.first + 1] += t.end[b.second] - t.start[b.second];
    for (int i = 0; i < nn; ++i) H.dir_off[i + 1] += H.dir_off[i];
    H.dir_src.resize(H.dir_off[nn]);
    {
        std::vector<int32_t> f(H.dir_off.begin(), H.dir_off.end() - 1);
        for (const auto& b : dir)
            for (int p = t.start[b.second]; p < t.end[b.second]; ++p)
                H.dir_src[f[b.first]++] = t.ord[p];
    }
}

// Uploads the host lists into the context's growable buffers; returns the
// device view and the double scratch after the kernel values.
static rc_status bct_upload(rc_ctx* ctx, const BctHost& H, size_t extra_d, BctDev& t,
                            double** scratch) {
    const int nn = (int)H.tree.start.size(), E = ctx->E;
    const size_t na = H.adm_src.size(), ndr = H.dir_src.size();
    const size_t ni = 3 * (size_t)nn + 2 * (size_t)E + 2 * ((size_t)nn + 1) + na + ndr;
    const size_t nd = na + extra_d;
    if (ni > ctx->bct_cap_i || nd > ctx->bct_cap_d) {
        if (ctx->d_bct_i) (void)hipFree(ctx->d_bct_i);
        if (ctx->d_bct_f) (void)hipFree(ctx->d_bct_f);
        ctx->d_bct_i = nullptr;
        ctx->d_bct_f = nullptr;
        ctx->bct_cap_i = ctx->bct_cap_d = 0;
        RC_HIP(hipMalloc((void**)&ctx->d_bct_i, sizeof(int32_t) * std::max<size_t>(ni, 1)));
        RC_HIP(hipMalloc((void**)&ctx->d_bct_f, sizeof(double) * std::max<size_t>(nd, 1)));
        ctx->bct_cap_i = ni;
        ctx->bct_cap_d = nd;
    }
    std::vector<int32_t> hi;
    hi.reserve(ni);
    auto put = [&](const std::vector<int32_t>& v) { hi.insert(hi.end(), v.begin(), v.end()); };
    put(H.tree.start);
    put(H.tree.end);
    put(H.parent);
    put(H.leaf_of);
    put(H.tree.ord);
    put(H.adm_off);
    put(H.adm_src);
    put(H.dir_off);
    put(H.dir_src);
    RC_HIP(hipMemcpy(ctx->d_bct_i, hi.data(), sizeof(int32_t) * hi.size(), hipMemcpyHostToDevice));
    if (na) RC_HIP(hipMemcpy(ctx->d_bct_f, H.kval.data(), sizeof(double) * na, hipMemcpyHostToDevice));
    const int32_t* g = ctx->d_bct_i;
    t.start = g; g += nn;
    t.end = g; g += nn;
    t.parent = g; g += nn;
    t.leaf_of = g; g += E;
    t.ord = g; g += E;
    t.adm_off = g; g += nn + 1;
    t.adm_src = g; g += na;
    t.dir_off = g; g += nn + 1;
    t.dir_src = g;
    t.kval = ctx->d_bct_f;
    *scratch = ctx->d_bct_f + na;
    return RC_OK;
}

static rc_status bct_check(rc_ctx* ctx, double eps, double theta, double theta_t, int leaf,
                           int kind) {
    if (!std::isfinite(eps) || !(theta >= 0.0) || !std::isfinite(theta) || !(theta_t >= 0.0) ||
        !std::isfinite(theta_t) || leaf < 1 || (kind != 0 && kind != 1)) {
        ctx->err = "block-cluster-tree parameters out of range";
        return RC_BAD_ARG;
    }
    return RC_OK;
}

static rc_status bct_geometry(rc_ctx* ctx, const double* xyz, double eps) {
    rc_params p{3.0, 6.0, eps};
    rc_status s = rc_upload_positions_aos(ctx, xyz, ctx->d_pos);
    if (s == RC_OK) s = rc_launch_edge_prep(ctx, ctx->d_pos, &p);
    return s;
}

}  // namespace

extern "C" rc_status rc_bct_matvec(rc_ctx* ctx, const double* xyz, double eps, const double* psi,
                                   int32_t ncomp, int32_t kind, double theta, double theta_t,
                                   int32_t leaf_size, double* phi, int64_t* stats) {
    if (!ctx || !xyz || !psi || !phi || (ncomp != 1 && ncomp != 3)) return RC_BAD_ARG;
    rc_status s = bct_check(ctx, eps, theta, theta_t, leaf_size, kind);
    if (s != RC_OK) return s;
    const int E = ctx->E;
    BctHost H;
    bct_build_host(ctx, xyz, eps, theta, theta_t, leaf_size, kind, H);
    const int nn = (int)H.tree.start.size();
    BctDev t;
    double* scr = nullptr;
    const size_t nin = (size_t)ncomp * E, nmom = 3 * (size_t)nn, nout = (size_t)ncomp * E;
    s = bct_upload(ctx, H, nin + nmom + nout, t, &scr);
    if (s == RC_OK) s = bct_geometry(ctx, xyz, eps);
    if (s != RC_OK) return s;
    double *d_psi = scr, *d_mom = scr + nin, *d_phi = scr + nin + nmom;
    RC_HIP(hipMemcpyAsync(d_psi, psi, sizeof(double) * nin, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_bct_moments, dim3((nn + 3) / 4), dim3(256), 0, ctx->stream, ctx->ed,
                       t, nn, (const double*)d_psi, ncomp, d_mom);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bct_apply, dim3((E + 3) / 4), dim3(256), 0, ctx->stream, ctx->ed, E,
                       t, kind, eps, (const double*)d_psi, ncomp, (const double*)d_mom, d_phi);
    RC_CHECK_LAUNCH();
    RC_HIP(hipMemcpyAsync(phi, d_phi, sizeof(double) * nout, hipMemcpyDeviceToHost, ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    if (stats) {
        stats[0] = H.leaves;
        stats[1] = H.adm;
        stats[2] = H.dir;
        stats[3] = H.pairs;
    }
    return RC_OK;
}

extern "C" rc_status rc_bct_apply_b(rc_ctx* ctx, const double* xyz, double eps, const double* u,
                                    int32_t kind, double theta, double theta_t,
                                    int32_t leaf_size, double* out) {
    if (!ctx || !xyz || !u || !out) return RC_BAD_ARG;
    rc_status s = bct_check(ctx, eps, theta, theta_t, leaf_size, kind);
    if (s != RC_OK) return s;
    const int E = ctx->E, V = ctx->V;
    BctHost H;
    bct_build_host(ctx, xyz, eps, theta, theta_t, leaf_size, kind, H);
    const int nn = (int)H.tree.start.size();
    const int nc = kind == 0 ? 3 : 1;
    BctDev t;
    double* scr = nullptr;
    // u, y, ones, moments (3 per node), K1, Ky, out
    const size_t need = (size_t)V + 3 * (size_t)E + E + 3 * (size_t)nn + E + 3 * (size_t)E + V;
    s = bct_upload(ctx, H, need, t, &scr);
    if (s == RC_OK) s = bct_geometry(ctx, xyz, eps);
    if (s != RC_OK) return s;
    double* d_u = scr;
    double* d_y = d_u + V;
    double* d_one = d_y + 3 * (size_t)E;
    double* d_mom = d_one + E;
    double* d_K1 = d_mom + 3 * (size_t)nn;
    double* d_Ky = d_K1 + E;
    double* d_out = d_Ky + 3 * (size_t)E;
    RC_HIP(hipMemcpyAsync(d_u, u, sizeof(double) * V, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_bct_density, dim3((E + 255) / 256), dim3(256), 0, ctx->stream, ctx->ed,
                       E, (int)kind, (const double*)d_u, d_y, d_one);
    RC_CHECK_LAUNCH();
    const dim3 gn((nn + 3) / 4), ge((E + 3) / 4);
    hipLaunchKernelGGL(k_bct_moments, gn, dim3(256), 0, ctx->stream, ctx->ed, t, nn,
                       (const double*)d_one, 1, d_mom);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bct_apply, ge, dim3(256), 0, ctx->stream, ctx->ed, E, t, (int)kind, eps,
                       (const double*)d_one, 1, (const double*)d_mom, d_K1);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bct_moments, gn, dim3(256), 0, ctx->stream, ctx->ed, t, nn,
                       (const double*)d_y, nc, d_mom);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bct_apply, ge, dim3(256), 0, ctx->stream, ctx->ed, E, t, (int)kind, eps,
                       (const double*)d_y, nc, (const double*)d_mom, d_Ky);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bct_gather, dim3((V + 255) / 256), dim3(256), 0, ctx->stream, ctx->ed, V,
                       (int)kind, ctx->d_vOff, ctx->d_vEdge, ctx->d_vSlot, (const double*)d_y,
                       (const double*)d_K1, (const double*)d_Ky, d_out);
    RC_CHECK_LAUNCH();
    RC_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * V, hipMemcpyDeviceToHost, ctx->stream));
    RC_HIP(hipStreamSynchronize(ctx->stream));
    return RC_OK;
}

// Forward difference of the Barnes-Hut energy (the stage-2 oracle's
// finite_difference_gradient of bh_energy, tpe_stage2_oracle.py:133-141,
// 1289-1290): the tree depends on the curve, so every moved coordinate is a
// fresh tree and traversal (3V + 1 Barnes-Hut energies).
extern "C" rc_status rc_gradient_bh(rc_ctx* ctx, const double* xyz, const rc_params* p,
                                    double theta, double theta_t, int32_t leaf_size, double h,
                                    double* grad) {
    if (!ctx || !xyz || !p || !grad || !std::isfinite(h) || h == 0.0) return RC_BAD_ARG;
    double e0 = 0.0;
    rc_status s = rc_energy_bh(ctx, xyz, p, theta, theta_t, leaf_size, &e0, nullptr);
    if (s != RC_OK) return s;
    std::vector<double> pert(xyz, xyz + 3 * (size_t)ctx->V);
    for (int i = 0; i < 3 * ctx->V; ++i) {
        const double keep = pert[i];
        pert[i] = keep + h;
        double e1 = 0.0;
        s = rc_energy_bh(ctx, pert.data(), p, theta, theta_t, leaf_size, &e1, nullptr);
        pert[i] = keep;
        if (s != RC_OK) return s;
        grad[i] = (e1 - e0) / h;
    }
    return RC_OK;
}

/*
 * rc_oracle.c — CPU FP64 restatement of the reference Sobolev step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library. Only tests/,
 * the smoke check and the bench's CPU-baseline leg load it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math: the
 * energy and gradient keep the reference's IEEE op order so the reference's
 * bit-identity goldens (test/golden.json) reproduce exactly).
 */
#include "rc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Threads of the pair loops. The per-pair terms of a block of rows I are
 * computed in parallel; they are then accumulated in the reference's (I, J)
 * order — the energy by one thread, the gradient and the Gram matrices by
 * threads that each own a range of output rows and walk the pairs in order
 * applying only their rows' updates. Every output entry therefore receives
 * the same operands in the same order as the sequential loop: results are
 * bit-identical for any thread count. */
static int g_threads = 0; /* 0: the OpenMP default (OMP_NUM_THREADS) */

void rco_set_threads(int n) { g_threads = n < 0 ? 0 : n; }

int rco_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

#define RCO_IB 64 /* rows I per block of the pair loops */

/* ------------------------------------------------------------------------ */
/* JavaScript Math.pow / ** semantics as executed by the reference's engine
 * (JavaScriptCore under Bun): a non-negative int32 exponent is evaluated by
 * binary exponentiation (repeated squaring, low bit first); every other
 * exponent goes to the C library pow. The energy/gradient call sites use
 * Math.pow(x, alpha), Math.pow(x, beta), Math.pow(x, alpha-1),
 * Math.pow(x, beta+1) with integer alpha/beta in the default config
 * (tangentPointEnergy.ts:98,198-199,230-231) — the squaring form is what
 * makes the golden energies bit-exact. */
static double js_pow(double x, double y) {
    if (isnan(y)) return NAN;
    if (y >= 0.0 && y <= 2147483647.0) {
        int32_t yi = (int32_t)y;
        if ((double)yi == y) {
            double result = 1.0;
            double b = x;
            while (yi) {
                if (yi & 1) result *= b;
                b *= b;
                yi >>= 1;
            }
            return result;
        }
    }
    return pow(x, y);
}

/* addToGrad's `x || 0` (tangentPointEnergy.ts:152-156): NaN and ±0 add +0. */
static inline double js_or0(double x) { return (x != x || x == 0.0) ? 0.0 : x; }

static inline int disjoint(const int32_t* edges, int I, int J) {
    int a1 = edges[2 * I], a2 = edges[2 * I + 1];
    int b1 = edges[2 * J], b2 = edges[2 * J + 1];
    return I != J && a1 != b1 && a1 != b2 && a2 != b1 && a2 != b2;
}

void rco_default_opts(rco_step_opts* o) {
    memset(o, 0, sizeof(*o));
    o->alpha = 3.0;
    o->beta = 6.0;
    o->epsilon = 1e-10;
    o->has_energy_before = 0;
    o->projection_frozen = 1;
    o->factor_lu = 0;
    o->max_iter = 8;
    o->tol_abs = 1e-4;
    o->tol_rel = 1e-4;
    o->c1 = 1e-4;
    o->shrink = 0.5;
    o->tau0 = 1.0;
    o->tau_min = 1e-12;
    o->penalties = NULL;
    o->gradient_mode = RCO_MODE_ANALYTICAL;
    o->fd_h = 1e-6;
}

/* ------------------------------------------------------------------------ */
/* calculateEnergy — tangentPointEnergy.ts:50-111. Ordered disjoint pairs,
 * J ascending (calculateDisjointPairs :33-47), eps after every norm, /2.
 * energy_row writes the pair terms 0.25 ell_I^(1-alpha) ell_J sumK of row I
 * (term[J], disjoint J only); the caller adds them in (I, J) order. */
static void energy_row(const double* v, const int32_t* edges, int E, int I, double alpha,
                       double beta, double eps, double* term) {
    const int i1 = edges[2 * I], i2 = edges[2 * I + 1];
    const double eIx = v[3 * i2] - v[3 * i1];
    const double eIy = v[3 * i2 + 1] - v[3 * i1 + 1];
    const double eIz = v[3 * i2 + 2] - v[3 * i1 + 2];
    const double ell_I = sqrt(eIx * eIx + eIy * eIy + eIz * eIz) + eps;
    for (int J = 0; J < E; J++) {
        if (!disjoint(edges, I, J)) continue;
        const int j1 = edges[2 * J], j2 = edges[2 * J + 1];
        const double eJx = v[3 * j2] - v[3 * j1];
        const double eJy = v[3 * j2 + 1] - v[3 * j1 + 1];
        const double eJz = v[3 * j2 + 2] - v[3 * j1 + 2];
        const double ell_J = sqrt(eJx * eJx + eJy * eJy + eJz * eJz) + eps;
        const int is[2] = {i1, i2}, js[2] = {j1, j2};
        double sumK = 0.0;
        for (int a = 0; a < 2; a++) {
            for (int b = 0; b < 2; b++) {
                const int i = is[a], j = js[b];
                const double dx = v[3 * i] - v[3 * j];
                const double dy = v[3 * i + 1] - v[3 * j + 1];
                const double dz = v[3 * i + 2] - v[3 * j + 2];
                const double d_norm = sqrt(dx * dx + dy * dy + dz * dz) + eps;
                const double cx = eIy * dz - eIz * dy;
                const double cy = eIz * dx - eIx * dz;
                const double cz = eIx * dy - eIy * dx;
                const double c_norm = sqrt(cx * cx + cy * cy + cz * cz) + eps;
                sumK += js_pow(c_norm, alpha) / js_pow(d_norm, beta);
            }
        }
        term[J] = 0.25 * js_pow(ell_I, 1 - alpha) * ell_J * sumK;
    }
}

/* Pair terms of rows I0 .. I0+nI-1 into buf[r * E + J] (parallel). */
static void energy_rows(const double* v, const int32_t* edges, int E, int I0, int nI,
                        double alpha, double beta, double eps, double* buf) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(rco_threads()) schedule(dynamic, 1)
#endif
    for (int r = 0; r < nI; r++) energy_row(v, edges, E, I0 + r, alpha, beta, eps, buf + (size_t)r * E);
}

double rco_energy(int V, int E, const double* v, const int32_t* edges,
                  double alpha, double beta, double eps) {
    (void)V;
    double total = 0.0;
    double* buf = (double*)malloc(sizeof(double) * RCO_IB * (size_t)(E > 0 ? E : 1));
    for (int I0 = 0; I0 < E; I0 += RCO_IB) {
        const int nI = E - I0 < RCO_IB ? E - I0 : RCO_IB;
        energy_rows(v, edges, E, I0, nI, alpha, beta, eps, buf);
        for (int r = 0; r < nI; r++)
            for (int J = 0; J < E; J++)
                if (disjoint(edges, I0 + r, J)) total += buf[(size_t)r * E + J];
    }
    free(buf);
    return total / 2;
}

/* gradientFiniteDiff — tangentPointEnergy.ts:114-136. Forward difference of
 * the full energy: E0 once, then for each vertex v and coordinate d a copy of
 * the curve with that coordinate moved by +h (one rounding, x + h), E1 of the
 * copy, g[v][d] = (E1 - E0) / h. O(V E^2): small curves only. */
void rco_gradient_fd(int V, int E, const double* v, const int32_t* edges,
                     double alpha, double beta, double eps, double h, double* g) {
    const double e0 = rco_energy(V, E, v, edges, alpha, beta, eps);
    double* pert = (double*)malloc(sizeof(double) * 3 * (size_t)(V > 0 ? V : 1));
    for (int i = 0; i < V; i++) {
        for (int d = 0; d < 3; d++) {
            memcpy(pert, v, sizeof(double) * 3 * (size_t)V);
            pert[3 * i + d] += h;
            const double e1 = rco_energy(V, E, pert, edges, alpha, beta, eps);
            g[3 * i + d] = (e1 - e0) / h;
        }
    }
    free(pert);
}

/* step — optimizer.ts:40-68. The raw L2 step: dE at the input (analytic or
 * forward difference), next = v - stepSize * dE per coordinate (product
 * rounded, then the difference), energy at next. field (nullable) = dE. */
void rco_step(int V, int E, const double* v, const int32_t* edges, double alpha, double beta,
              double eps, int mode, double step_size, double h, double* out, double* energy,
              double* field) {
    double* g = (double*)malloc(sizeof(double) * 3 * (size_t)(V > 0 ? V : 1));
    if (mode == RCO_MODE_ANALYTICAL)
        rco_gradient(V, E, v, edges, alpha, beta, eps, g);
    else
        rco_gradient_fd(V, E, v, edges, alpha, beta, eps, h, g);
    for (int i = 0; i < 3 * V; i++) {
        const double t = step_size * g[i];
        out[i] = v[i] - t;
    }
    *energy = rco_energy(V, E, out, edges, alpha, beta, eps);
    if (field) memcpy(field, g, sizeof(double) * 3 * (size_t)V);
    free(g);
}

/* Same terms, Neumaier-compensated accumulation of the pair energies (in the
 * same (I, J) order): the sum's rounding drops to O(u) instead of O(P u), so
 * it separates the reference's left-to-right rounding from a device result
 * summed in another order at large E. */
double rco_energy_compensated(int V, int E, const double* v, const int32_t* edges,
                              double alpha, double beta, double eps) {
    (void)V;
    double s = 0.0, comp = 0.0;
    double* buf = (double*)malloc(sizeof(double) * RCO_IB * (size_t)(E > 0 ? E : 1));
    for (int I0 = 0; I0 < E; I0 += RCO_IB) {
        const int nI = E - I0 < RCO_IB ? E - I0 : RCO_IB;
        energy_rows(v, edges, E, I0, nI, alpha, beta, eps, buf);
        for (int r = 0; r < nI; r++)
            for (int J = 0; J < E; J++) {
                if (!disjoint(edges, I0 + r, J)) continue;
                const double term = buf[(size_t)r * E + J];
                const double t = s + term;
                if (fabs(s) >= fabs(term)) comp += (s - t) + term;
                else comp += (term - t) + s;
                s = t;
            }
    }
    free(buf);
    return (s + comp) / 2;
}

/* ------------------------------------------------------------------------ */
/* kernelDerivs — tangentPointEnergy.ts:165-244. */
typedef struct {
    double f, dd[3], de[3];
} kd_t;

static kd_t kernel_derivs(double ex, double ey, double ez, double dx, double dy, double dz,
                          double alpha, double beta, double eps) {
    kd_t k;
    const double rd = sqrt(dx * dx + dy * dy + dz * dz);
    double dHx = 0, dHy = 0, dHz = 0;
    if (!(rd < 1e-14)) {
        const double invRd = 1 / rd;
        dHx = invRd * dx;
        dHy = invRd * dy;
        dHz = invRd * dz;
    }
    const double d_eps = rd + eps;
    const double cx = ey * dz - ez * dy;
    const double cy = ez * dx - ex * dz;
    const double cz = ex * dy - ey * dx;
    const double rc = sqrt(cx * cx + cy * cy + cz * cz);
    const double c_eps = rc + eps;
    const double cPowA = js_pow(c_eps, alpha);
    const double dPowB = js_pow(d_eps, beta);
    k.f = cPowA / dPowB;
    double cdd[3] = {0, 0, 0}, cde[3] = {0, 0, 0};
    if (rc >= 1e-14) {
        const double invRc = 1 / rc;
        const double ccex = cy * ez - cz * ey;
        const double ccey = cz * ex - cx * ez;
        const double ccez = cx * ey - cy * ex;
        cdd[0] = invRc * ccex;
        cdd[1] = invRc * ccey;
        cdd[2] = invRc * ccez;
        const double dccx = dy * cz - dz * cy;
        const double dccy = dz * cx - dx * cz;
        const double dccz = dx * cy - dy * cx;
        cde[0] = invRc * dccx;
        cde[1] = invRc * dccy;
        cde[2] = invRc * dccz;
    }
    const double coeff_c = (alpha * js_pow(c_eps, alpha - 1)) / dPowB;
    const double coeff_d = (-beta * cPowA) / js_pow(d_eps, beta + 1);
    k.dd[0] = coeff_c * cdd[0] + coeff_d * dHx;
    k.dd[1] = coeff_c * cdd[1] + coeff_d * dHy;
    k.dd[2] = coeff_c * cdd[2] + coeff_d * dHz;
    k.de[0] = coeff_c * cde[0];
    k.de[1] = coeff_c * cde[1];
    k.de[2] = coeff_c * cde[2];
    return k;
}

static inline void add_to_grad(double* g, int idx, double x, double y, double z) {
    g[3 * idx] += js_or0(x);
    g[3 * idx + 1] += js_or0(y);
    g[3 * idx + 2] += js_or0(z);
}

/* Guarded unit tangent of edge K as the gradient forms it
 * (tangentPointEnergy.ts:181-186,261-266,293-297): e / ||e||, 0 below 1e-14;
 * *re = ||e||. */
static void edge_unit(const double* v, const int32_t* edges, int K, double* e, double* eh,
                      double* re) {
    const int k1 = edges[2 * K], k2 = edges[2 * K + 1];
    e[0] = v[3 * k2] - v[3 * k1];
    e[1] = v[3 * k2 + 1] - v[3 * k1 + 1];
    e[2] = v[3 * k2 + 2] - v[3 * k1 + 2];
    *re = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    eh[0] = eh[1] = eh[2] = 0;
    if (!(*re < 1e-14)) {
        const double inv = 1 / *re;
        eh[0] = inv * e[0];
        eh[1] = inv * e[1];
        eh[2] = inv * e[2];
    }
}

/* Per ordered pair of the gradient: sI, sJ, base, then dd[4][3], de[4][3]. */
#define GR_Q 27

/* Pair quantities of row I (gradientAnalytical's inner loop, before its
 * addToGrad calls) into q[J * GR_Q], disjoint J only. */
static void gradient_row(const double* v, const int32_t* edges, int E, int I, double alpha,
                         double beta, double eps, double* q) {
    const int i1 = edges[2 * I], i2 = edges[2 * I + 1];
    double eI[3], eIh[3], reI;
    edge_unit(v, edges, I, eI, eIh, &reI);
    const double ell_I = reI + eps;
    const double ell_I_pow = js_pow(ell_I, 1 - alpha);
    for (int J = 0; J < E; J++) {
        if (!disjoint(edges, I, J)) continue;
        const int j1 = edges[2 * J], j2 = edges[2 * J + 1];
        double eJ[3], eJh[3], reJ;
        edge_unit(v, edges, J, eJ, eJh, &reJ);
        const double ell_J = reJ + eps;
        const int ti[4] = {i1, i1, i2, i2};
        const int tj[4] = {j1, j2, j1, j2};
        double* o = q + (size_t)J * GR_Q;
        double sumF = 0;
        for (int p = 0; p < 4; p++) {
            const double dx = v[3 * ti[p]] - v[3 * tj[p]];
            const double dy = v[3 * ti[p] + 1] - v[3 * tj[p] + 1];
            const double dz = v[3 * ti[p] + 2] - v[3 * tj[p] + 2];
            const kd_t t = kernel_derivs(eI[0], eI[1], eI[2], dx, dy, dz, alpha, beta, eps);
            sumF += t.f;
            for (int c = 0; c < 3; c++) {
                o[3 + 3 * p + c] = t.dd[c];
                o[15 + 3 * p + c] = t.de[c];
            }
        }
        o[0] = 0.25 * ell_J * sumF;
        o[1] = 0.25 * ell_I_pow * sumF;
        o[2] = 0.25 * ell_I_pow * ell_J;
    }
}

/* The addToGrad calls of rows I0 .. I0+nI-1 in the reference's order,
 * applied only to vertices lo <= idx < hi. */
static void gradient_scatter(const double* v, const int32_t* edges, int E, int I0, int nI,
                             double alpha, double eps, const double* buf, int lo, int hi,
                             double* g) {
#define ADDG(idx, x, y, z) \
    do { if ((idx) >= lo && (idx) < hi) add_to_grad(g, (idx), (x), (y), (z)); } while (0)
    for (int r = 0; r < nI; r++) {
        const int I = I0 + r;
        const int i1 = edges[2 * I], i2 = edges[2 * I + 1];
        const int ihit = (i1 >= lo && i1 < hi) || (i2 >= lo && i2 < hi);
        double eI[3], eIh[3], reI;
        edge_unit(v, edges, I, eI, eIh, &reI);
        const double ell_I = reI + eps;
        const double dPowCoeff = (1 - alpha) * js_pow(ell_I, -alpha);
        const double dP1x = dPowCoeff * (-1 * eIh[0]), dP1y = dPowCoeff * (-1 * eIh[1]),
                     dP1z = dPowCoeff * (-1 * eIh[2]);
        const double dP2x = dPowCoeff * eIh[0], dP2y = dPowCoeff * eIh[1], dP2z = dPowCoeff * eIh[2];
        for (int J = 0; J < E; J++) {
            if (!disjoint(edges, I, J)) continue;
            const int j1 = edges[2 * J], j2 = edges[2 * J + 1];
            if (!ihit && !(j1 >= lo && j1 < hi) && !(j2 >= lo && j2 < hi)) continue;
            double eJ[3], eJh[3], reJ;
            edge_unit(v, edges, J, eJ, eJh, &reJ);
            const double dJ1x = -1 * eJh[0], dJ1y = -1 * eJh[1], dJ1z = -1 * eJh[2];
            const int ti[4] = {i1, i1, i2, i2};
            const int tj[4] = {j1, j2, j1, j2};
            const double* o = buf + ((size_t)r * E + J) * GR_Q;
            const double sI = o[0], sJ = o[1], base = o[2];
            ADDG(i1, sI * dP1x, sI * dP1y, sI * dP1z);
            ADDG(i2, sI * dP2x, sI * dP2y, sI * dP2z);
            ADDG(j1, sJ * dJ1x, sJ * dJ1y, sJ * dJ1z);
            ADDG(j2, sJ * eJh[0], sJ * eJh[1], sJ * eJh[2]);
            for (int p = 0; p < 4; p++) {
                const double* dd = o + 3 + 3 * p;
                const double* de = o + 15 + 3 * p;
                ADDG(ti[p], base * dd[0], base * dd[1], base * dd[2]);
                ADDG(tj[p], -base * dd[0], -base * dd[1], -base * dd[2]);
                ADDG(i1, -base * de[0], -base * de[1], -base * de[2]);
                ADDG(i2, base * de[0], base * de[1], base * de[2]);
            }
        }
    }
#undef ADDG
}

/* gradientAnalytical — tangentPointEnergy.ts:139-392 (same addToGrad order). */
void rco_gradient(int V, int E, const double* v, const int32_t* edges,
                  double alpha, double beta, double eps, double* g) {
    memset(g, 0, sizeof(double) * 3 * (size_t)V);
    const int nt = rco_threads();
    double* buf = (double*)malloc(sizeof(double) * GR_Q * RCO_IB * (size_t)(E > 0 ? E : 1));
    for (int I0 = 0; I0 < E; I0 += RCO_IB) {
        const int nI = E - I0 < RCO_IB ? E - I0 : RCO_IB;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
#endif
        for (int r = 0; r < nI; r++)
            gradient_row(v, edges, E, I0 + r, alpha, beta, eps, buf + (size_t)r * E * GR_Q);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static, 1)
#endif
        for (int t = 0; t < nt; t++)
            gradient_scatter(v, edges, E, I0, nI, alpha, eps, buf, (int)((long)V * t / nt),
                             (int)((long)V * (t + 1) / nt), g);
    }
    free(buf);
    for (int i = 0; i < V; i++) {
        g[3 * i] *= 0.5;
        g[3 * i + 1] *= 0.5;
        g[3 * i + 2] *= 0.5;
    }
}

/* ------------------------------------------------------------------------ */
/* Per-edge geometry for the metric (innerProduct.ts:62-77): ell^eps and the
 * unregularised unit tangent, zero when ||e|| < 1e-14. */
static void edge_geom_eps(const double* v, const int32_t* edges, int E, double eps,
                          double* ell, double* T) {
    for (int I = 0; I < E; I++) {
        const int i1 = edges[2 * I], i2 = edges[2 * I + 1];
        const double ex = v[3 * i2] - v[3 * i1];
        const double ey = v[3 * i2 + 1] - v[3 * i1 + 1];
        const double ez = v[3 * i2 + 2] - v[3 * i1 + 2];
        const double r = sqrt(ex * ex + ey * ey + ez * ez);
        ell[I] = r + eps;
        if (r < 1e-14) {
            T[3 * I] = T[3 * I + 1] = T[3 * I + 2] = 0;
        } else {
            const double inv = 1 / r;
            T[3 * I] = ex * inv;
            T[3 * I + 1] = ey * inv;
            T[3 * I + 2] = ez * inv;
        }
    }
}

/* Explicit symmetrisation 0.5 (x + y) over off-diagonal pairs
 * (innerProduct.ts:131-137). Row i writes the pairs (i, j > i): disjoint
 * across rows, so the rows run in parallel. */
static void symmetrize(double* B, int n) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(rco_threads()) schedule(dynamic, 16)
#endif
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            const double avg = 0.5 * (B[(size_t)i * n + j] + B[(size_t)j * n + i]);
            B[(size_t)i * n + j] = avg;
            B[(size_t)j * n + i] = avg;
        }
}

/* Pair weights of row I of the metric, disjoint J only (w[J]):
 * high order (innerProduct.ts:94-104) w = 1/4 ell_I ell_J sum 1/(d+eps)^(2 sigma+1);
 * low order (innerProduct.ts:230-251) q = 1/4 (1/4 ell_I ell_J sum k2^4/(d+eps)^(2 sigma+1)),
 * k2^4 = (||T_I x d|| + eps)^2 / (d+eps)^4 with T_I only. */
static void gram_row(const double* v, const int32_t* edges, int E, int I, const double* ell,
                     const double* T, double eps, double distExp, int low, double* w) {
    const int idxI[2] = {edges[2 * I], edges[2 * I + 1]};
    const double ellI = ell[I];
    const double* TI = T + 3 * I;
    for (int J = 0; J < E; J++) {
        if (!disjoint(edges, I, J)) continue;
        const int idxJ[2] = {edges[2 * J], edges[2 * J + 1]};
        const double ellJ = ell[J];
        double sum = 0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const int ia = idxI[a], jb = idxJ[b];
                const double dx = v[3 * ia] - v[3 * jb];
                const double dy = v[3 * ia + 1] - v[3 * jb + 1];
                const double dz = v[3 * ia + 2] - v[3 * jb + 2];
                const double dNorm = sqrt(dx * dx + dy * dy + dz * dz) + eps;
                if (!low) {
                    sum += 1 / js_pow(dNorm, distExp);
                } else {
                    const double cx = TI[1] * dz - TI[2] * dy;
                    const double cy = TI[2] * dx - TI[0] * dz;
                    const double cz = TI[0] * dy - TI[1] * dx;
                    const double crossNorm = sqrt(cx * cx + cy * cy + cz * cz) + eps;
                    const double k24 = (crossNorm * crossNorm) / (dNorm * dNorm * dNorm * dNorm);
                    sum += k24 / js_pow(dNorm, distExp);
                }
            }
        if (!low) {
            w[J] = 0.25 * ellI * ellJ * sum;
        } else {
            const double wLow = 0.25 * ellI * ellJ * sum;
            w[J] = 0.25 * wLow;
        }
    }
}

/* The 16 updates of every pair of rows I0 .. I0+nI-1 in the reference's
 * order (innerProduct.ts:111-124 signed, :259-270 unsigned), applied only to
 * matrix rows lo <= row < hi. */
static void gram_scatter(const int32_t* edges, int E, int I0, int nI, const double* ell,
                         const double* T, int low, const double* buf, int lo, int hi, int n,
                         double* B) {
#define INR(x) ((x) >= lo && (x) < hi)
    for (int r = 0; r < nI; r++) {
        const int I = I0 + r;
        const int idxI[2] = {edges[2 * I], edges[2 * I + 1]};
        const int ihit = INR(idxI[0]) || INR(idxI[1]);
        const double ellI = ell[I];
        const double* TI = T + 3 * I;
        for (int J = 0; J < E; J++) {
            if (!disjoint(edges, I, J)) continue;
            const int idxJ[2] = {edges[2 * J], edges[2 * J + 1]};
            if (!ihit && !INR(idxJ[0]) && !INR(idxJ[1])) continue;
            const double w = buf[(size_t)r * E + J];
            const double ellJ = ell[J];
            const double* TJ = T + 3 * J;
            if (!low) {
                const double dotT = TI[0] * TJ[0] + TI[1] * TJ[1] + TI[2] * TJ[2];
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++) {
                        const double sign = (a + b) % 2 == 0 ? 1 : -1;
                        const int ia = idxI[a], ib = idxI[b], ja = idxJ[a], jb = idxJ[b];
                        if (INR(ia)) B[(size_t)ia * n + ib] += (sign * w) / (ellI * ellI);
                        if (INR(ia)) B[(size_t)ia * n + jb] -= (sign * w * dotT) / (ellI * ellJ);
                        if (INR(ja)) B[(size_t)ja * n + jb] += (sign * w) / (ellJ * ellJ);
                        if (INR(ja)) B[(size_t)ja * n + ib] -= (sign * w * dotT) / (ellI * ellJ);
                    }
            } else {
                const double q = w;
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++) {
                        const int ia = idxI[a], ib = idxI[b], ja = idxJ[a], jb = idxJ[b];
                        if (INR(ia)) B[(size_t)ia * n + ib] += q;
                        if (INR(ia)) B[(size_t)ia * n + jb] -= q;
                        if (INR(ja)) B[(size_t)ja * n + ib] -= q;
                        if (INR(ja)) B[(size_t)ja * n + jb] += q;
                    }
            }
        }
    }
#undef INR
}

/* assembleBHighFlat (innerProduct.ts:38-140, low = 0) and assembleBLowFlat
 * (:177-285, low = 1): s = (beta-1)/alpha, sigma = s-1, distExp = 2 sigma+1
 * (:52-54); no 1/2 on B; symmetrised. */
static void assemble_b(int n, int E, const double* v, const int32_t* edges, double alpha,
                       double beta, double eps, int low, double* B) {
    memset(B, 0, sizeof(double) * (size_t)n * n);
    const double s = (beta - 1) / alpha;
    const double sigma = s - 1;
    const double distExp = 2 * sigma + 1;
    double* ell = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    double* T = (double*)malloc(sizeof(double) * 3 * (size_t)(E > 0 ? E : 1));
    edge_geom_eps(v, edges, E, eps, ell, T);
    const int nt = rco_threads();
    double* buf = (double*)malloc(sizeof(double) * RCO_IB * (size_t)(E > 0 ? E : 1));
    for (int I0 = 0; I0 < E; I0 += RCO_IB) {
        const int nI = E - I0 < RCO_IB ? E - I0 : RCO_IB;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
#endif
        for (int r = 0; r < nI; r++)
            gram_row(v, edges, E, I0 + r, ell, T, eps, distExp, low, buf + (size_t)r * E);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static, 1)
#endif
        for (int t = 0; t < nt; t++)
            gram_scatter(edges, E, I0, nI, ell, T, low, buf, (int)((long)n * t / nt),
                         (int)((long)n * (t + 1) / nt), n, B);
    }
    symmetrize(B, n);
    free(buf);
    free(ell);
    free(T);
}

/* A = B + B0 with every entry accumulated in long double (x87 80-bit) and
 * rounded once at the end. Not a reference function: the pair weights are
 * the reference's (FP64), only the accumulation over the ~2E contributions
 * of the diagonal and edge entries is extra-precise. The checker uses it to
 * separate the FP64 summation order's effect on the saddle solution from
 * the solver's (kappa(A_hat) ~ 1e9 at V = 8192 amplifies 1e-14 entry errors). */
void rco_assemble_a_ld(int n, int E, const double* v, const int32_t* edges, double alpha,
                       double beta, double eps, double* A) {
    long double* L = (long double*)calloc((size_t)n * n, sizeof(long double));
    const double s = (beta - 1) / alpha;
    const double sigma = s - 1;
    const double distExp = 2 * sigma + 1;
    double* ell = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    double* T = (double*)malloc(sizeof(double) * 3 * (size_t)(E > 0 ? E : 1));
    edge_geom_eps(v, edges, E, eps, ell, T);
    const int nt = rco_threads();
    double* buf = (double*)malloc(sizeof(double) * 2 * RCO_IB * (size_t)(E > 0 ? E : 1));
    for (int I0 = 0; I0 < E; I0 += RCO_IB) {
        const int nI = E - I0 < RCO_IB ? E - I0 : RCO_IB;
        double* bh = buf;
        double* bl = buf + (size_t)RCO_IB * E;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
#endif
        for (int r = 0; r < 2 * nI; r++) {
            const int low = r >= nI, rr = low ? r - nI : r;
            gram_row(v, edges, E, I0 + rr, ell, T, eps, distExp, low,
                     (low ? bl : bh) + (size_t)rr * E);
        }
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static, 1)
#endif
        for (int t = 0; t < nt; t++) {
            const int lo = (int)((long)n * t / nt), hi = (int)((long)n * (t + 1) / nt);
            for (int r = 0; r < nI; r++) {
                const int I = I0 + r;
                const int idxI[2] = {edges[2 * I], edges[2 * I + 1]};
                const double ellI = ell[I];
                const double* TI = T + 3 * I;
                for (int J = 0; J < E; J++) {
                    if (!disjoint(edges, I, J)) continue;
                    const int idxJ[2] = {edges[2 * J], edges[2 * J + 1]};
                    const double w = bh[(size_t)r * E + J], q = bl[(size_t)r * E + J];
                    const double ellJ = ell[J];
                    const double* TJ = T + 3 * J;
                    const long double dotT = (long double)TI[0] * TJ[0] +
                                             (long double)TI[1] * TJ[1] + (long double)TI[2] * TJ[2];
                    for (int a = 0; a < 2; a++)
                        for (int b = 0; b < 2; b++) {
                            const long double sign = (a + b) % 2 == 0 ? 1 : -1;
                            const int ia = idxI[a], ib = idxI[b], ja = idxJ[a], jb = idxJ[b];
                            if (ia >= lo && ia < hi) {
                                L[(size_t)ia * n + ib] += sign * w / ((long double)ellI * ellI) + q;
                                L[(size_t)ia * n + jb] -= sign * w * dotT / ((long double)ellI * ellJ) + q;
                            }
                            if (ja >= lo && ja < hi) {
                                L[(size_t)ja * n + jb] += sign * w / ((long double)ellJ * ellJ) + q;
                                L[(size_t)ja * n + ib] -= sign * w * dotT / ((long double)ellI * ellJ) + q;
                            }
                        }
                }
            }
        }
    }
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
#endif
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            A[(size_t)i * n + j] = (double)(0.5L * (L[(size_t)i * n + j] + L[(size_t)j * n + i]));
    free(buf);
    free(ell);
    free(T);
    free(L);
}

void rco_assemble_b_high(int n, int E, const double* v, const int32_t* edges,
                         double alpha, double beta, double eps, double* B) {
    assemble_b(n, E, v, edges, alpha, beta, eps, 0, B);
}

void rco_assemble_b_low(int n, int E, const double* v, const int32_t* edges,
                        double alpha, double beta, double eps, double* B0) {
    assemble_b(n, E, v, edges, alpha, beta, eps, 1, B0);
}

/* assembleAFlat — innerProduct.ts:325-350. */
void rco_assemble_a(int n, int E, const double* v, const int32_t* edges,
                    double alpha, double beta, double eps, double* A) {
    double* B0 = (double*)malloc(sizeof(double) * (size_t)n * n);
    rco_assemble_b_high(n, E, v, edges, alpha, beta, eps, A);
    rco_assemble_b_low(n, E, v, edges, alpha, beta, eps, B0);
    const long nn = (long)n * n;
#ifdef _OPENMP
#pragma omp parallel for num_threads(rco_threads()) schedule(static)
#endif
    for (long i = 0; i < nn; i++) A[i] = A[i] + B0[i];
    free(B0);
}

/* ------------------------------------------------------------------------ */
/* Constraints. Raw lengths (no eps); T = 0 when ell < 1e-14. */

/* barycenterTarget — constraints.ts:23-52. */
void rco_barycenter_target(int V, int E, const double* v, const int32_t* edges, double* x0) {
    double total = 0, ax = 0, ay = 0, az = 0;
    for (int r = 0; r < E; r++) {
        const int i1 = edges[2 * r], i2 = edges[2 * r + 1];
        const double* p1 = v + 3 * i1;
        const double* p2 = v + 3 * i2;
        const double ex = p2[0] - p1[0], ey = p2[1] - p1[1], ez = p2[2] - p1[2];
        const double ell = sqrt(ex * ex + ey * ey + ez * ez);
        total += ell;
        ax += ell * 0.5 * (p1[0] + p2[0]);
        ay += ell * 0.5 * (p1[1] + p2[1]);
        az += ell * 0.5 * (p1[2] + p2[2]);
    }
    if (total < 1e-14) {
        double m[3] = {0, 0, 0};
        for (int i = 0; i < V; i++) {
            m[0] += v[3 * i] / V;
            m[1] += v[3 * i + 1] / V;
            m[2] += v[3 * i + 2] / V;
        }
        x0[0] = m[0];
        x0[1] = m[1];
        x0[2] = m[2];
        return;
    }
    x0[0] = ax / total;
    x0[1] = ay / total;
    x0[2] = az / total;
}

/* totalLength — constraintSet.ts:64-75. */
double rco_total_length(int V, int E, const double* v, const int32_t* edges) {
    (void)V;
    double L = 0;
    for (int r = 0; r < E; r++) {
        const double* p1 = v + 3 * edges[2 * r];
        const double* p2 = v + 3 * edges[2 * r + 1];
        const double ex = p2[0] - p1[0], ey = p2[1] - p1[1], ez = p2[2] - p1[2];
        L += sqrt(ex * ex + ey * ey + ez * ez);
    }
    return L;
}

/* barycenterScale — lineSearch.ts:86-109. JS Math.max propagates NaN. */
static double js_max(double a, double b) {
    if (a != a || b != b) return NAN;
    return a > b ? a : b;
}

static double barycenter_scale(const double* v, const int32_t* edges, int E, const double* x0) {
    double L = 0, R = 0;
    for (int r = 0; r < E; r++) {
        const double* p1 = v + 3 * edges[2 * r];
        const double* p2 = v + 3 * edges[2 * r + 1];
        const double ex = p2[0] - p1[0], ey = p2[1] - p1[1], ez = p2[2] - p1[2];
        L += sqrt(ex * ex + ey * ey + ez * ez);
        const double d1x = p1[0] - x0[0], d1y = p1[1] - x0[1], d1z = p1[2] - x0[2];
        const double d2x = p2[0] - x0[0], d2y = p2[1] - x0[1], d2z = p2[2] - x0[2];
        R = js_max(R, js_max(sqrt(d1x * d1x + d1y * d1y + d1z * d1z),
                             sqrt(d2x * d2x + d2y * d2y + d2z * d2z)));
    }
    return js_max(1, L * js_max(1, R));
}

int rco_constraint_rows(int V, int E, int nblocks, const rco_block* blocks) {
    (void)V;
    int k = 0;
    for (int b = 0; b < nblocks; b++) {
        switch (blocks[b].kind) {
            case RCO_BARYCENTER: k += 3; break;
            case RCO_TOTAL_LENGTH: k += 1; break;
            case RCO_EDGE_LENGTHS: k += E; break;
            case RCO_POINT: k += 3; break;
            default: break;
        }
    }
    return k;
}

/* evaluateConstraintSet — constraintSet.ts:309-322, per-block evaluate and
 * scale: barycenterPhiAndC (constraints.ts:67-120), totalLengthBlock
 * (constraintSet.ts:119-162), edgeLengthsBlock (:206-250), pointBlock (:271-301). */
void rco_eval_constraints(int V, int E, const double* v, const int32_t* edges,
                          int nblocks, const rco_block* blocks,
                          double* phi, double* C, double* scales, int32_t* rows) {
    const int m = 3 * V;
    const int k = rco_constraint_rows(V, E, nblocks, blocks);
    if (C) memset(C, 0, sizeof(double) * (size_t)k * m);
    int row = 0;
    for (int b = 0; b < nblocks; b++) {
        const rco_block* blk = blocks + b;
        if (blk->kind == RCO_BARYCENTER) {
            double ph[3] = {0, 0, 0};
            for (int r = 0; r < E; r++) {
                const int i1 = edges[2 * r], i2 = edges[2 * r + 1];
                const double* p1 = v + 3 * i1;
                const double* p2 = v + 3 * i2;
                const double ex = p2[0] - p1[0], ey = p2[1] - p1[1], ez = p2[2] - p1[2];
                const double ell = sqrt(ex * ex + ey * ey + ez * ez);
                double T[3] = {0, 0, 0};
                if (!(ell < 1e-14)) {
                    const double inv = 1 / ell;
                    T[0] = ex * inv;
                    T[1] = ey * inv;
                    T[2] = ez * inv;
                }
                const double rv[3] = {0.5 * (p1[0] + p2[0]) - blk->x0[0],
                                      0.5 * (p1[1] + p2[1]) - blk->x0[1],
                                      0.5 * (p1[2] + p2[2]) - blk->x0[2]};
                ph[0] += ell * rv[0];
                ph[1] += ell * rv[1];
                ph[2] += ell * rv[2];
                if (C) {
                    for (int rr = 0; rr < 3; rr++)
                        for (int c = 0; c < 3; c++) {
                            const double delta = rr == c ? 1 : 0;
                            C[(size_t)(row + rr) * m + c * V + i1] += -T[c] * rv[rr] + 0.5 * ell * delta;
                            C[(size_t)(row + rr) * m + c * V + i2] += +T[c] * rv[rr] + 0.5 * ell * delta;
                        }
                }
            }
            phi[row] = ph[0];
            phi[row + 1] = ph[1];
            phi[row + 2] = ph[2];
            if (scales) scales[b] = barycenter_scale(v, edges, E, blk->x0);
            if (rows) rows[b] = 3;
            row += 3;
        } else if (blk->kind == RCO_TOTAL_LENGTH || blk->kind == RCO_EDGE_LENGTHS) {
            const int per_edge = blk->kind == RCO_EDGE_LENGTHS;
            double L = 0;
            for (int r = 0; r < E; r++) {
                const int i1 = edges[2 * r], i2 = edges[2 * r + 1];
                const double* p1 = v + 3 * i1;
                const double* p2 = v + 3 * i2;
                const double ex = p2[0] - p1[0], ey = p2[1] - p1[1], ez = p2[2] - p1[2];
                const double ell = sqrt(ex * ex + ey * ey + ez * ez);
                L += ell;
                double T[3] = {0, 0, 0};
                if (!(ell < 1e-14)) {
                    const double inv = 1 / ell;
                    T[0] = ex * inv;
                    T[1] = ey * inv;
                    T[2] = ez * inv;
                }
                const int rr = per_edge ? row + r : row;
                if (per_edge)
                    phi[rr] = (blk->ell0 && r < blk->ell0_len ? blk->ell0[r] : NAN) - ell;
                if (C)
                    for (int c = 0; c < 3; c++) {
                        C[(size_t)rr * m + c * V + i1] += T[c];
                        C[(size_t)rr * m + c * V + i2] += -T[c];
                    }
            }
            if (!per_edge) phi[row] = blk->L0 - L;
            if (scales) scales[b] = js_max(1, rco_total_length(V, E, v, edges));
            if (rows) rows[b] = per_edge ? E : 1;
            row += per_edge ? E : 1;
        } else if (blk->kind == RCO_POINT) {
            const int i = blk->vertex;
            if (i < 0 || i >= V) {
                phi[row] = phi[row + 1] = phi[row + 2] = NAN;
            } else {
                if (C)
                    for (int r = 0; r < 3; r++) C[(size_t)(row + r) * m + r * V + i] = 1;
                phi[row] = v[3 * i] - blk->target[0];
                phi[row + 1] = v[3 * i + 1] - blk->target[1];
                phi[row + 2] = v[3 * i + 2] - blk->target[2];
            }
            if (scales) {
                double R = 0;
                for (int q = 0; q < V; q++) {
                    const double dx = v[3 * q] - blk->target[0];
                    const double dy = v[3 * q + 1] - blk->target[1];
                    const double dz = v[3 * q + 2] - blk->target[2];
                    R = js_max(R, sqrt(dx * dx + dy * dy + dz * dz));
                }
                scales[b] = js_max(1, R);
            }
            if (rows) rows[b] = 3;
            row += 3;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Penalties — penalties.ts:60-241. */
typedef struct {
    double wL, wD, wF, X[3];
} pen_r;

static pen_r pen_resolve(const rco_penalties* p) {
    pen_r r = {0, 0, 0, {0, 0, 0}};
    if (!p) return r;
    r.wL = p->w_length;
    r.wD = p->w_diff;
    r.wF = p->w_field;
    if (r.wF != 0) {
        const double x = p->X[0], y = p->X[1], z = p->X[2];
        const double nn = sqrt(x * x + y * y + z * z);
        if (nn < 1e-14) r.wF = 0;
        else {
            r.X[0] = x / nn;
            r.X[1] = y / nn;
            r.X[2] = z / nn;
        }
    }
    return r;
}

int rco_penalties_active(const rco_penalties* p) {
    const pen_r r = pen_resolve(p);
    return r.wL != 0 || r.wD != 0 || r.wF != 0;
}

static void pen_geom(const double* v, const int32_t* edges, int E, double* ell, double* T) {
    for (int r = 0; r < E; r++) {
        const double* p1 = v + 3 * edges[2 * r];
        const double* p2 = v + 3 * edges[2 * r + 1];
        const double ex = p2[0] - p1[0], ey = p2[1] - p1[1], ez = p2[2] - p1[2];
        const double len = sqrt(ex * ex + ey * ey + ez * ez);
        ell[r] = len;
        if (len < 1e-14) T[3 * r] = T[3 * r + 1] = T[3 * r + 2] = 0;
        else {
            T[3 * r] = ex / len;
            T[3 * r + 1] = ey / len;
            T[3 * r + 2] = ez / len;
        }
    }
}

/* V_int: degree exactly 2, incident edges ascending (penalties.ts:116-127). */
static int interior_vertices(int n, const int32_t* edges, int E, int32_t* out /*3*n*/) {
    int* cnt = (int*)calloc((size_t)n, sizeof(int));
    int* first = (int*)malloc(sizeof(int) * (size_t)n);
    int* second = (int*)malloc(sizeof(int) * (size_t)n);
    for (int r = 0; r < E; r++)
        for (int s = 0; s < 2; s++) {
            const int q = edges[2 * r + s];
            if (cnt[q] == 0) first[q] = r;
            else if (cnt[q] == 1) second[q] = r;
            cnt[q]++;
        }
    int m = 0;
    for (int q = 0; q < n; q++)
        if (cnt[q] == 2) {
            out[3 * m] = q;
            out[3 * m + 1] = first[q];
            out[3 * m + 2] = second[q];
            m++;
        }
    free(cnt);
    free(first);
    free(second);
    return m;
}

double rco_penalty_energy(int V, int E, const double* v, const int32_t* edges,
                          const rco_penalties* p) {
    const pen_r r = pen_resolve(p);
    double* ell = (double*)malloc(sizeof(double) * (size_t)E);
    double* T = (double*)malloc(sizeof(double) * 3 * (size_t)E);
    pen_geom(v, edges, E, ell, T);
    double total = 0;
    if (r.wL != 0) {
        double acc = 0;
        for (int e = 0; e < E; e++) acc += ell[e];
        total += r.wL * acc;
    }
    if (r.wD != 0) {
        int32_t* iv = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)V);
        const int m = interior_vertices(V, edges, E, iv);
        double acc = 0;
        for (int q = 0; q < m; q++) {
            const double d = ell[iv[3 * q + 1]] - ell[iv[3 * q + 2]];
            acc += d * d;
        }
        total += r.wD * acc;
        free(iv);
    }
    if (r.wF != 0) {
        double acc = 0;
        for (int e = 0; e < E; e++) {
            const double* t = T + 3 * e;
            const double cx = t[1] * r.X[2] - t[2] * r.X[1];
            const double cy = t[2] * r.X[0] - t[0] * r.X[2];
            const double cz = t[0] * r.X[1] - t[1] * r.X[0];
            acc += ell[e] * (cx * cx + cy * cy + cz * cz);
        }
        total += r.wF * acc;
    }
    free(ell);
    free(T);
    return total;
}

void rco_penalty_gradient(int V, int E, const double* v, const int32_t* edges,
                          const rco_penalties* p, double* g) {
    const pen_r r = pen_resolve(p);
    memset(g, 0, sizeof(double) * 3 * (size_t)V);
    double* ell = (double*)malloc(sizeof(double) * (size_t)E);
    double* T = (double*)malloc(sizeof(double) * 3 * (size_t)E);
    pen_geom(v, edges, E, ell, T);
    if (r.wL != 0) {
        for (int e = 0; e < E; e++) {
            const int a = edges[2 * e], b = edges[2 * e + 1];
            const double* t = T + 3 * e;
            for (int c = 0; c < 3; c++) {
                g[3 * a + c] -= r.wL * t[c];
                g[3 * b + c] += r.wL * t[c];
            }
        }
    }
    if (r.wD != 0) {
        int32_t* iv = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)V);
        const int m = interior_vertices(V, edges, E, iv);
        for (int q = 0; q < m; q++) {
            const int rI = iv[3 * q + 1], rJ = iv[3 * q + 2];
            const double s = 2 * r.wD * (ell[rI] - ell[rJ]);
            const double* tI = T + 3 * rI;
            const double* tJ = T + 3 * rJ;
            const int aI = edges[2 * rI], bI = edges[2 * rI + 1];
            const int aJ = edges[2 * rJ], bJ = edges[2 * rJ + 1];
            for (int c = 0; c < 3; c++) g[3 * aI + c] -= s * tI[c];
            for (int c = 0; c < 3; c++) g[3 * bI + c] += s * tI[c];
            for (int c = 0; c < 3; c++) g[3 * aJ + c] += s * tJ[c];
            for (int c = 0; c < 3; c++) g[3 * bJ + c] -= s * tJ[c];
        }
        free(iv);
    }
    if (r.wF != 0) {
        for (int e = 0; e < E; e++) {
            if (ell[e] < 1e-14) continue;
            const double* t = T + 3 * e;
            const double u = t[0] * r.X[0] + t[1] * r.X[1] + t[2] * r.X[2];
            const double c = 1 + u * u;
            const double gg[3] = {c * t[0] - 2 * u * r.X[0], c * t[1] - 2 * u * r.X[1],
                                  c * t[2] - 2 * u * r.X[2]};
            const int a = edges[2 * e], b = edges[2 * e + 1];
            for (int q = 0; q < 3; q++) g[3 * a + q] -= r.wF * gg[q];
            for (int q = 0; q < 3; q++) g[3 * b + q] += r.wF * gg[q];
        }
    }
    free(ell);
    free(T);
}

/* ------------------------------------------------------------------------ */
/* Dense solvers. */

static int all_finite(const double* a, size_t n) {
    for (size_t i = 0; i < n; i++)
        if (!isfinite(a[i])) return 0;
    return 1;
}

/* ldltFactor — linsolve.ts:431-589: Bunch–Kaufman (LAPACK dsytf2, UPLO='L')
 * on the lower triangle of a row-major buffer, 1-based ipiv, negative for
 * 2x2 blocks. Returns 1 on a non-finite entry or an exactly zero column. */
int rco_ldlt_factor(double* k, int n, int32_t* ipiv) {
    if (!all_finite(k, (size_t)n * n)) return 1;
    const double ALPHA = (1 + sqrt(17.0)) / 8;
    double* w1 = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* w2 = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int col = 0;
    while (col < n) {
        const size_t colBase = (size_t)col * n;
        const double absakk = fabs(k[colBase + col]);
        int imax = col;
        double colmax = 0;
        for (int r = col + 1; r < n; r++) {
            const double a = fabs(k[(size_t)r * n + col]);
            if (a > colmax) {
                colmax = a;
                imax = r;
            }
        }
        const double mx = absakk > colmax ? absakk : colmax;
        if (!(mx > 0)) {
            free(w1);
            free(w2);
            return 1;
        }
        int kstep = 1, kp = col;
        if (absakk >= ALPHA * colmax) {
            kp = col;
        } else {
            double rowmax = 0;
            const size_t imaxBase = (size_t)imax * n;
            for (int j = col; j < imax; j++) {
                const double a = fabs(k[imaxBase + j]);
                rowmax = rowmax > a ? rowmax : a;
            }
            for (int r = imax + 1; r < n; r++) {
                const double a = fabs(k[(size_t)r * n + imax]);
                rowmax = rowmax > a ? rowmax : a;
            }
            if (absakk >= ALPHA * colmax * (colmax / rowmax)) {
                kp = col;
            } else if (fabs(k[imaxBase + imax]) >= ALPHA * rowmax) {
                kp = imax;
            } else {
                kp = imax;
                kstep = 2;
            }
        }
        const int kk = col + kstep - 1;
        if (kp != kk) {
            for (int r = kp + 1; r < n; r++) {
                const double t = k[(size_t)r * n + kk];
                k[(size_t)r * n + kk] = k[(size_t)r * n + kp];
                k[(size_t)r * n + kp] = t;
            }
            for (int j = kk + 1; j < kp; j++) {
                const double t = k[(size_t)j * n + kk];
                k[(size_t)j * n + kk] = k[(size_t)kp * n + j];
                k[(size_t)kp * n + j] = t;
            }
            const double td = k[(size_t)kk * n + kk];
            k[(size_t)kk * n + kk] = k[(size_t)kp * n + kp];
            k[(size_t)kp * n + kp] = td;
            if (kstep == 2) {
                const double t = k[(size_t)(col + 1) * n + col];
                k[(size_t)(col + 1) * n + col] = k[(size_t)kp * n + col];
                k[(size_t)kp * n + col] = t;
            }
        }
        if (kstep == 1) {
            if (col < n - 1) {
                const double d11 = 1 / k[colBase + col];
                for (int j = col + 1; j < n; j++) w1[j] = d11 * k[(size_t)j * n + col];
                for (int i = col + 1; i < n; i++) {
                    const size_t iBase = (size_t)i * n;
                    const double aik = k[iBase + col];
                    for (int j = col + 1; j <= i; j++) k[iBase + j] -= aik * w1[j];
                }
                for (int r = col + 1; r < n; r++) k[(size_t)r * n + col] *= d11;
            }
        } else {
            if (col < n - 2) {
                const double d21 = k[(size_t)(col + 1) * n + col];
                const double d11 = k[(size_t)(col + 1) * n + col + 1] / d21;
                const double d22 = k[colBase + col] / d21;
                const double t = 1 / (d11 * d22 - 1);
                const double d21t = t / d21;
                for (int j = col + 2; j < n; j++) {
                    const double ajk = k[(size_t)j * n + col];
                    const double ajk1 = k[(size_t)j * n + col + 1];
                    w1[j] = d21t * (d11 * ajk - ajk1);
                    w2[j] = d21t * (d22 * ajk1 - ajk);
                }
                for (int i = col + 2; i < n; i++) {
                    const size_t iBase = (size_t)i * n;
                    const double aik = k[iBase + col];
                    const double aik1 = k[iBase + col + 1];
                    for (int j = col + 2; j <= i; j++)
                        k[iBase + j] = k[iBase + j] - aik * w1[j] - aik1 * w2[j];
                }
                for (int j = col + 2; j < n; j++) {
                    k[(size_t)j * n + col] = w1[j];
                    k[(size_t)j * n + col + 1] = w2[j];
                }
            }
        }
        if (kstep == 1) {
            ipiv[col] = kp + 1;
        } else {
            ipiv[col] = -(kp + 1);
            ipiv[col + 1] = -(kp + 1);
        }
        col += kstep;
    }
    free(w1);
    free(w2);
    return 0;
}

/* ldltSolveFactored — linsolve.ts:603-691 (LAPACK dsytrs 'L'); b in place. */
void rco_ldlt_solve(const double* m, const int32_t* ipiv, int n, double* b) {
    int i = 0;
    while (i < n) {
        if (ipiv[i] > 0) {
            const int kp = ipiv[i] - 1;
            if (kp != i) {
                const double t = b[i];
                b[i] = b[kp];
                b[kp] = t;
            }
            for (int r = i + 1; r < n; r++) b[r] -= m[(size_t)r * n + i] * b[i];
            b[i] /= m[(size_t)i * n + i];
            i += 1;
        } else {
            const int kp = -ipiv[i] - 1;
            if (kp != i + 1) {
                const double t = b[i + 1];
                b[i + 1] = b[kp];
                b[kp] = t;
            }
            for (int r = i + 2; r < n; r++)
                b[r] = b[r] - m[(size_t)r * n + i] * b[i] - m[(size_t)r * n + i + 1] * b[i + 1];
            const double akm1k = m[(size_t)(i + 1) * n + i];
            const double akm1 = m[(size_t)i * n + i] / akm1k;
            const double ak = m[(size_t)(i + 1) * n + i + 1] / akm1k;
            const double denom = akm1 * ak - 1;
            const double bkm1 = b[i] / akm1k;
            const double bk = b[i + 1] / akm1k;
            b[i] = (ak * bkm1 - bk) / denom;
            b[i + 1] = (akm1 * bk - bkm1) / denom;
            i += 2;
        }
    }
    i = n - 1;
    while (i >= 0) {
        if (ipiv[i] > 0) {
            double s = b[i];
            for (int r = i + 1; r < n; r++) s -= m[(size_t)r * n + i] * b[r];
            b[i] = s;
            const int kp = ipiv[i] - 1;
            if (kp != i) {
                const double t = b[i];
                b[i] = b[kp];
                b[kp] = t;
            }
            i -= 1;
        } else {
            double s = b[i];
            for (int r = i + 1; r < n; r++) s -= m[(size_t)r * n + i] * b[r];
            b[i] = s;
            double s1 = b[i - 1];
            for (int r = i + 1; r < n; r++) s1 -= m[(size_t)r * n + (i - 1)] * b[r];
            b[i - 1] = s1;
            const int kp = -ipiv[i] - 1;
            if (kp != i) {
                const double t = b[i];
                b[i] = b[kp];
                b[kp] = t;
            }
            i -= 2;
        }
    }
}

/* luFactor — linsolve.ts:243-316: partial pivoting, first max wins. */
int rco_lu_factor(double* k, int n, int32_t* piv) {
    if (!all_finite(k, (size_t)n * n)) return 1;
    for (int i = 0; i < n; i++) piv[i] = i;
    for (int col = 0; col < n; col++) {
        int pivotRow = col;
        double maxAbs = fabs(k[(size_t)col * n + col]);
        for (int r = col + 1; r < n; r++) {
            const double a = fabs(k[(size_t)r * n + col]);
            if (a > maxAbs) {
                maxAbs = a;
                pivotRow = r;
            }
        }
        if (!(maxAbs > 0)) return 1;
        if (pivotRow != col) {
            for (int c = 0; c < n; c++) {
                const double t = k[(size_t)col * n + c];
                k[(size_t)col * n + c] = k[(size_t)pivotRow * n + c];
                k[(size_t)pivotRow * n + c] = t;
            }
            const int tp = piv[col];
            piv[col] = piv[pivotRow];
            piv[pivotRow] = tp;
        }
        const size_t colBase = (size_t)col * n;
        const double pivotVal = k[colBase + col];
        for (int r = col + 1; r < n; r++) {
            const size_t rBase = (size_t)r * n;
            const double mult = k[rBase + col] / pivotVal;
            k[rBase + col] = mult;
            for (int c = col + 1; c < n; c++) k[rBase + c] -= mult * k[colBase + c];
        }
    }
    return 0;
}

void rco_lu_solve(const double* m, const int32_t* piv, int n, const double* rhs, double* x) {
    double* y = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        double s = rhs[piv[i]];
        for (int j = 0; j < i; j++) s -= m[(size_t)i * n + j] * y[j];
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = y[i];
        for (int j = i + 1; j < n; j++) s -= m[(size_t)i * n + j] * x[j];
        x[i] = s / m[(size_t)i * n + i];
    }
    free(y);
}

/* structuredSaddleResidual — linsolve.ts:810-849. */
static double structured_residual(const double* a, int n, const double* C, int k,
                                  const double* z, const double* rhs) {
    const int m = 3 * n;
    double res = 0, rn = 0;
    for (int b = 0; b < 3; b++) {
        const int off = b * n;
        for (int i = 0; i < n; i++) {
            double s = 0;
            const size_t base = (size_t)i * n;
            for (int j = 0; j < n; j++) s += a[base + j] * z[off + j];
            const int row = off + i;
            for (int r = 0; r < k; r++) s += C[(size_t)r * m + row] * z[m + r];
            const double d = s - rhs[row];
            res += d * d;
            rn += rhs[row] * rhs[row];
        }
    }
    for (int r = 0; r < k; r++) {
        double s = 0;
        for (int j = 0; j < m; j++) s += C[(size_t)r * m + j] * z[j];
        const double d = s - rhs[m + r];
        res += d * d;
        rn += rhs[m + r] * rhs[m + r];
    }
    return sqrt(res) / js_max(1, sqrt(rn));
}

/* A factored frozen saddle operator (linsolve.ts:865-870). */
typedef struct {
    int n, k, size, lu;
    double* a;   /* scalar A, n x n (unfactored) */
    double* C;   /* k x 3n */
    double* K;   /* factored, size x size */
    int32_t* piv;
} frozen_t;

static void frozen_free(frozen_t* f) {
    free(f->a);
    free(f->C);
    free(f->K);
    free(f->piv);
    memset(f, 0, sizeof(*f));
}

/* Build K = [[I3 (x) A, C^T],[C, 0]] (linsolve.ts:771-790) and factor it. */
static int frozen_build(frozen_t* f, const double* A, int n, const double* C, int k, int lu) {
    const int m = 3 * n;
    const int size = m + k;
    f->n = n;
    f->k = k;
    f->size = size;
    f->lu = lu;
    f->a = (double*)malloc(sizeof(double) * (size_t)n * n);
    memcpy(f->a, A, sizeof(double) * (size_t)n * n);
    f->C = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1) * m);
    if (k > 0) memcpy(f->C, C, sizeof(double) * (size_t)k * m);
    f->K = (double*)calloc((size_t)size * size, sizeof(double));
    f->piv = (int32_t*)malloc(sizeof(int32_t) * (size_t)(size > 0 ? size : 1));
    for (int b = 0; b < 3; b++) {
        const int off = b * n;
        for (int i = 0; i < n; i++)
            memcpy(f->K + (size_t)(off + i) * size + off, A + (size_t)i * n, sizeof(double) * n);
    }
    for (int r = 0; r < k; r++)
        for (int j = 0; j < m; j++) {
            f->K[(size_t)(m + r) * size + j] = C[(size_t)r * m + j];
            f->K[(size_t)j * size + (m + r)] = C[(size_t)r * m + j];
        }
    return lu ? rco_lu_factor(f->K, size, f->piv) : rco_ldlt_factor(f->K, size, f->piv);
}

/* solveSaddleFrozen — linsolve.ts:882-905. Returns 1 on a non-finite rhs. */
static int frozen_solve(const frozen_t* f, const double* top, const double* bottom,
                        double* x, double* lambda, double* residual) {
    const int m = 3 * f->n;
    double* rhs = (double*)malloc(sizeof(double) * (size_t)(f->size > 0 ? f->size : 1));
    double* z = (double*)malloc(sizeof(double) * (size_t)(f->size > 0 ? f->size : 1));
    for (int i = 0; i < m; i++) rhs[i] = top ? top[i] : 0.0;
    for (int r = 0; r < f->k; r++) rhs[m + r] = bottom ? bottom[r] : 0.0;
    if (!all_finite(rhs, (size_t)f->size)) {
        free(rhs);
        free(z);
        return 1;
    }
    if (f->lu) {
        rco_lu_solve(f->K, f->piv, f->size, rhs, z);
    } else {
        memcpy(z, rhs, sizeof(double) * (size_t)f->size);
        rco_ldlt_solve(f->K, f->piv, f->size, z);
    }
    if (residual) *residual = structured_residual(f->a, f->n, f->C, f->k, z, rhs);
    if (x) memcpy(x, z, sizeof(double) * (size_t)m);
    if (lambda) memcpy(lambda, z + m, sizeof(double) * (size_t)f->k);
    free(rhs);
    free(z);
    return 0;
}

int rco_solve_saddle_from_a(const double* A, int V, const double* C, int k,
                            const double* rhs_top, const double* rhs_bottom,
                            int factor_lu, double* x, double* lambda, double* residual) {
    frozen_t f;
    memset(&f, 0, sizeof(f));
    int rc = frozen_build(&f, A, V, C, k, factor_lu);
    if (rc == 0) rc = frozen_solve(&f, rhs_top, rhs_bottom, x, lambda, residual);
    frozen_free(&f);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Residual of the reference saddle system K z = rhs (linsolve.ts:771-790,
 * K = [[I3 (x) A, C^T],[C, 0]], z = [x; lambda]) in double-double
 * arithmetic: every product a*x is split exactly (TwoProduct by fma), every
 * sum is an error-free TwoSum carried in a (hi, lo) pair, and the row's
 * result is rounded once. Not a reference function: it is the extra-precise
 * residual with which the checker refines a saddle solution beyond FP64
 * round-off (kappa(K) ~ 1e9 at V = 8192). A V x V row-major, C k x 3V. */
typedef struct {
    double hi, lo;
} dd_t;

static inline dd_t dd_add_d(dd_t a, double b) {
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b - bb);
    const double lo = a.lo + err;
    dd_t r;
    r.hi = s + lo;
    r.lo = lo - (r.hi - s);
    return r;
}

static inline dd_t dd_sub_prod(dd_t a, double x, double y) {
    const double p = x * y;
    const double e = fma(x, y, -p);
    a = dd_add_d(a, -p);
    return dd_add_d(a, -e);
}

void rco_saddle_residual_dd(const double* A, int V, const double* C, int k,
                            const double* r_top, const double* r_bot, const double* x,
                            const double* lambda, double* rho_top, double* rho_bot) {
    const int m = 3 * V;
#ifdef _OPENMP
#pragma omp parallel for num_threads(rco_threads()) schedule(dynamic, 16)
#endif
    for (int row = 0; row < m; row++) {
        const int b = row / V, i = row - b * V;
        dd_t s = {r_top ? r_top[row] : 0.0, 0.0};
        const double* a = A + (size_t)i * V;
        const double* xb = x + (size_t)b * V;
        for (int j = 0; j < V; j++) s = dd_sub_prod(s, a[j], xb[j]);
        for (int r = 0; r < k; r++) s = dd_sub_prod(s, C[(size_t)r * m + row], lambda[r]);
        rho_top[row] = s.hi + s.lo;
    }
    for (int r = 0; r < k; r++) {
        dd_t s = {r_bot ? r_bot[r] : 0.0, 0.0};
        for (int j = 0; j < m; j++) s = dd_sub_prod(s, C[(size_t)r * m + j], x[j]);
        rho_bot[r] = s.hi + s.lo;
    }
}

/* ------------------------------------------------------------------------ */
/* l2CurveNorm — lineSearch.ts:53-75 (values AoS). */
double rco_l2_curve_norm(int V, int E, const double* values, const double* v,
                         const int32_t* edges) {
    (void)V;
    double total = 0;
    for (int r = 0; r < E; r++) {
        const int i1 = edges[2 * r], i2 = edges[2 * r + 1];
        const double* p1 = v + 3 * i1;
        const double* p2 = v + 3 * i2;
        const double ex = p2[0] - p1[0], ey = p2[1] - p1[1], ez = p2[2] - p1[2];
        const double ell = sqrt(ex * ex + ey * ey + ez * ez);
        const double* v1 = values + 3 * i1;
        const double* v2 = values + 3 * i2;
        total += 0.5 * ell *
                 (v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2] +
                  (v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]));
    }
    return sqrt(js_max(0, total));
}

static void flatten(const double* aos, int V, double* flat) {
    for (int i = 0; i < V; i++)
        for (int c = 0; c < 3; c++) flat[c * V + i] = aos[3 * i + c];
}

static void unflatten(const double* flat, int V, double* aos) {
    for (int i = 0; i < V; i++)
        for (int c = 0; c < 3; c++) aos[3 * i + c] = flat[c * V + i];
}

/* solveConstrainedGradientSetFrozen — gradient.ts:84-115. On success the
 * frozen operator is returned in *fz (caller frees). */
static int solve_gradient_frozen(int V, int E, const double* pos, const int32_t* edges,
                                 double alpha, double beta, double eps, const double* dE,
                                 int nblocks, const rco_block* blocks, int factor_lu,
                                 double* gtilde, double* lambda, double* residual,
                                 frozen_t* fz) {
    const int m = 3 * V;
    const int k = rco_constraint_rows(V, E, nblocks, blocks);
    double* A = (double*)malloc(sizeof(double) * (size_t)V * V);
    rco_assemble_a(V, E, pos, edges, alpha, beta, eps, A);
    double* phi = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
    double* C = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1) * m);
    rco_eval_constraints(V, E, pos, edges, nblocks, blocks, phi, C, NULL, NULL);
    double* rhs = (double*)malloc(sizeof(double) * (size_t)m);
    flatten(dE, V, rhs);
    double* x = (double*)malloc(sizeof(double) * (size_t)m);
    memset(fz, 0, sizeof(*fz));
    int rc = frozen_build(fz, A, V, C, k, factor_lu);
    if (rc == 0) rc = frozen_solve(fz, rhs, NULL, x, lambda, residual);
    if (rc == 0) unflatten(x, V, gtilde);
    else frozen_free(fz);
    free(A);
    free(phi);
    free(C);
    free(rhs);
    free(x);
    return rc;
}

int rco_solve_constrained_gradient(int V, int E, const double* pos, const int32_t* edges,
                                   double alpha, double beta, double eps,
                                   const double* dE, int nblocks, const rco_block* blocks,
                                   int factor_lu, double* gtilde, double* lambda,
                                   double* residual) {
    frozen_t fz;
    const int rc = solve_gradient_frozen(V, E, pos, edges, alpha, beta, eps, dE, nblocks,
                                         blocks, factor_lu, gtilde, lambda, residual, &fz);
    if (rc == 0) frozen_free(&fz);
    return rc;
}

/* projectOntoConstraintSet — lineSearch.ts:196-295. cur (AoS) is updated in
 * place; returns ok. */
static int project(int V, int E, double* cur, const int32_t* edges, double alpha, double beta,
                   double eps, int nblocks, const rco_block* blocks, const rco_step_opts* o,
                   const frozen_t* fz, int* iterations, double* phi_norm) {
    const int m = 3 * V;
    const int k = rco_constraint_rows(V, E, nblocks, blocks);
    double* phi = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
    double* C = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1) * m);
    double* scales = (double*)malloc(sizeof(double) * (size_t)(nblocks > 0 ? nblocks : 1));
    int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nblocks > 0 ? nblocks : 1));
    double* negphi = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
    double* x = (double*)malloc(sizeof(double) * (size_t)m);
    double* zero = (double*)calloc((size_t)m, sizeof(double));
    double* A = NULL;
    int ok = 0, it;
    double fin = INFINITY;
    for (it = 0; it <= o->max_iter; it++) {
        rco_eval_constraints(V, E, cur, edges, nblocks, blocks, phi, C, scales, rows);
        double sq = 0;
        for (int r = 0; r < k; r++) sq += phi[r] * phi[r];
        fin = sqrt(sq);
        int converged = 1, row = 0;
        for (int b = 0; b < nblocks; b++) {
            double bs = 0;
            for (int r = 0; r < rows[b]; r++) bs += phi[row + r] * phi[row + r];
            row += rows[b];
            const double bn = sqrt(bs);
            if (bn > js_max(o->tol_abs, o->tol_rel * scales[b])) {
                converged = 0;
                break;
            }
        }
        if (converged) {
            ok = 1;
            break;
        }
        if (it == o->max_iter) break;
        for (int r = 0; r < k; r++) negphi[r] = -phi[r];
        int rc;
        if (fz) {
            rc = frozen_solve(fz, zero, negphi, x, NULL, NULL);
        } else {
            if (!A) A = (double*)malloc(sizeof(double) * (size_t)V * V);
            rco_assemble_a(V, E, cur, edges, alpha, beta, eps, A);
            double res;
            rc = rco_solve_saddle_from_a(A, V, C, k, zero, negphi, o->factor_lu, x, NULL, &res);
        }
        if (rc != 0) break;
        int finite = 1;
        for (int i = 0; i < m; i++)
            if (!isfinite(x[i])) finite = 0;
        if (!finite) break;
        for (int i = 0; i < V; i++)
            for (int c = 0; c < 3; c++) cur[3 * i + c] = cur[3 * i + c] + x[c * V + i];
    }
    if (it > o->max_iter) it = o->max_iter;
    *iterations = it;
    *phi_norm = fin;
    free(phi);
    free(C);
    free(scales);
    free(rows);
    free(negphi);
    free(x);
    free(zero);
    free(A);
    return ok;
}

static double objective(int V, int E, const double* v, const int32_t* edges,
                        const rco_step_opts* o) {
    const double e = rco_energy(V, E, v, edges, o->alpha, o->beta, o->epsilon);
    return rco_penalties_active(o->penalties) ? e + rco_penalty_energy(V, E, v, edges, o->penalties)
                                              : e;
}

/* lineSearchStepSet — lineSearch.ts:423-562. */
static int line_search(int V, int E, const double* pos, const int32_t* edges, const double* dE,
                       const double* gt, int nblocks, const rco_block* blocks,
                       const rco_step_opts* o, const frozen_t* fz, double* out,
                       rco_step_stats* st) {
    const double e0 = o->has_energy_before ? o->energy_before : objective(V, E, pos, edges, o);
    const double gradNorm = rco_l2_curve_norm(V, E, gt, pos, edges);
    st->energy_before = e0;
    st->trials = 0;
    memcpy(out, pos, sizeof(double) * 3 * (size_t)V);
    if (!isfinite(gradNorm) || gradNorm <= 0) {
        st->accepted = 0;
        st->reason = RCO_REASON_ZERO_OR_NONFINITE_GRADIENT_NORM;
        st->tau = 0;
        st->energy = e0;
        st->projection_iterations = 0;
        st->slope = NAN;
        st->phi_norm = NAN;
        return 0;
    }
    double* dir = (double*)malloc(sizeof(double) * 3 * (size_t)V);
    for (int i = 0; i < 3 * V; i++) dir[i] = gt[i] / gradNorm;
    double* dEf = (double*)malloc(sizeof(double) * 3 * (size_t)V);
    double* dirf = (double*)malloc(sizeof(double) * 3 * (size_t)V);
    flatten(dE, V, dEf);
    flatten(dir, V, dirf);
    double slope = 0;
    for (int i = 0; i < 3 * V; i++) slope += dEf[i] * dirf[i];
    st->slope = slope;
    st->gradient_l2_norm = gradNorm;
    int result = 0;
    if (!isfinite(slope) || slope <= 0) {
        st->accepted = 0;
        st->reason = RCO_REASON_NOT_A_DESCENT_DIRECTION;
        st->tau = 0;
        st->energy = e0;
        st->projection_iterations = 0;
        st->phi_norm = NAN;
        goto done;
    }
    double tau = o->tau0;
    int lastReason = RCO_REASON_TAU_BELOW_MIN;
    double* cand = (double*)malloc(sizeof(double) * 3 * (size_t)V);
    while (tau >= o->tau_min) {
        st->trials++;
        for (int i = 0; i < 3 * V; i++) cand[i] = pos[i] - tau * dir[i];
        int iters;
        double pn;
        const int ok = project(V, E, cand, edges, o->alpha, o->beta, o->epsilon, nblocks, blocks,
                               o, fz, &iters, &pn);
        int finite = all_finite(cand, 3 * (size_t)V);
        if (ok && finite) {
            const double e1 = objective(V, E, cand, edges, o);
            if (isfinite(e1) && e1 <= e0 - o->c1 * tau * slope) {
                st->accepted = 1;
                st->reason = RCO_REASON_NONE;
                st->tau = tau;
                st->energy = e1;
                st->projection_iterations = iters;
                st->phi_norm = pn;
                memcpy(out, cand, sizeof(double) * 3 * (size_t)V);
                free(cand);
                goto done;
            }
            lastReason = RCO_REASON_ARMIJO_FAILED;
        } else {
            lastReason = RCO_REASON_PROJECTION_FAILED;
        }
        tau *= o->shrink;
    }
    free(cand);
    st->accepted = 0;
    st->reason = lastReason;
    st->tau = 0;
    st->energy = e0;
    st->projection_iterations = -1;
    st->phi_norm = NAN;
done:
    free(dir);
    free(dEf);
    free(dirf);
    return result;
}

int rco_line_search(int V, int E, const double* pos, const int32_t* edges,
                    const double* dE, const double* gtilde,
                    int nblocks, const rco_block* blocks,
                    const rco_step_opts* o, double* pos_out, rco_step_stats* st) {
    memset(st, 0, sizeof(*st));
    frozen_t fz;
    frozen_t* fzp = NULL;
    if (o->projection_frozen) {
        const int m = 3 * V;
        const int k = rco_constraint_rows(V, E, nblocks, blocks);
        double* A = (double*)malloc(sizeof(double) * (size_t)V * V);
        rco_assemble_a(V, E, pos, edges, o->alpha, o->beta, o->epsilon, A);
        double* phi = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
        double* C = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1) * m);
        rco_eval_constraints(V, E, pos, edges, nblocks, blocks, phi, C, NULL, NULL);
        memset(&fz, 0, sizeof(fz));
        const int rc = frozen_build(&fz, A, V, C, k, o->factor_lu);
        free(A);
        free(phi);
        free(C);
        if (rc != 0) {
            frozen_free(&fz);
            return 1;
        }
        fzp = &fz;
    }
    line_search(V, E, pos, edges, dE, gtilde, nblocks, blocks, o, fzp, pos_out, st);
    if (fzp) frozen_free(fzp);
    return 0;
}

/* sobolevStepSet — optimizer.ts:231-440 (mode 'analytical'). */
int rco_sobolev_step(int V, int E, const double* pos, const int32_t* edges,
                     int nblocks, const rco_block* blocks, const rco_step_opts* o,
                     double* pos_out, double* gtilde_out, rco_step_stats* st) {
    memset(st, 0, sizeof(*st));
    st->projection_iterations = -1;
    const int k = rco_constraint_rows(V, E, nblocks, blocks);
    double* dE = (double*)malloc(sizeof(double) * 3 * (size_t)V);
    /* dE by the requested mode (optimizer.ts:278-282) */
    if (o->gradient_mode == RCO_MODE_FINITE_DIFF)
        rco_gradient_fd(V, E, pos, edges, o->alpha, o->beta, o->epsilon, o->fd_h, dE);
    else
        rco_gradient(V, E, pos, edges, o->alpha, o->beta, o->epsilon, dE);
    if (rco_penalties_active(o->penalties)) {
        double* pg = (double*)malloc(sizeof(double) * 3 * (size_t)V);
        rco_penalty_gradient(V, E, pos, edges, o->penalties, pg);
        for (int i = 0; i < 3 * V; i++) dE[i] = dE[i] + pg[i];
        free(pg);
    }
    double* gt = (double*)malloc(sizeof(double) * 3 * (size_t)V);
    double* lambda = (double*)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
    double residual = NAN;
    frozen_t fz;
    const int rc = solve_gradient_frozen(V, E, pos, edges, o->alpha, o->beta, o->epsilon, dE,
                                         nblocks, blocks, o->factor_lu, gt, lambda, &residual,
                                         &fz);
    memcpy(pos_out, pos, sizeof(double) * 3 * (size_t)V);
    if (rc != 0) {
        st->energy = o->has_energy_before ? o->energy_before : objective(V, E, pos, edges, o);
        st->accepted = 0;
        st->converged = 0;
        st->tau = 0;
        st->residual = NAN;
        st->gradient_l2_norm = NAN;
        st->projection_iterations = -1;
        st->reason = RCO_REASON_SINGULAR_SYSTEM;
        free(dE);
        free(gt);
        free(lambda);
        return 0;
    }
    if (gtilde_out) memcpy(gtilde_out, gt, sizeof(double) * 3 * (size_t)V);
    st->residual = residual;
    const double gn = rco_l2_curve_norm(V, E, gt, pos, edges);
    st->gradient_l2_norm = gn;
    if (gn < 1e-4) {
        st->energy = o->has_energy_before ? o->energy_before : objective(V, E, pos, edges, o);
        st->accepted = 0;
        st->converged = 1;
        st->tau = 0;
        st->projection_iterations = 0;
        st->reason = RCO_REASON_NONE;
    } else {
        rco_step_stats ls;
        memset(&ls, 0, sizeof(ls));
        line_search(V, E, pos, edges, dE, gt, nblocks, blocks, o,
                    o->projection_frozen ? &fz : NULL, pos_out, &ls);
        st->energy = ls.energy;
        st->energy_before = ls.energy_before;
        st->accepted = ls.accepted;
        st->converged = 0;
        st->tau = ls.tau;
        st->projection_iterations = ls.projection_iterations;
        st->reason = ls.reason;
        st->slope = ls.slope;
        st->phi_norm = ls.phi_norm;
        st->trials = ls.trials;
    }
    frozen_free(&fz);
    free(dE);
    free(gt);
    free(lambda);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Barnes-Hut energy — the reference's stage-2 oracle, bh_energy with
 * build_edge_bvh and bh_ordered_cluster_term (oracle/tpe_stage2_oracle.py:
 * 456-596). Edge BVH over midpoints and unit tangents: a node holds its edges
 * in the order its parent passed them, its mass-weighted centre and tangent,
 * the radii rx (centres) and rT (tangents); a node with more than leaf_size
 * edges splits at the median of the 6-D coordinate (midpoint, tangent) with
 * the largest spread, after a stable sort on it. Traversal per target edge I
 * from the root: a node of >= 2 edges, none sharing a vertex with I, with
 * 0 < dist, rx / dist <= theta and rT <= theta_t is replaced by one source
 * point (the centre) carrying the node's mass; a leaf otherwise sums the
 * exact ordered terms of its disjoint edges; E = 1/2 sum_I. Powers are the
 * C library pow (the Python oracle's `**`). stats (nullable): visits,
 * approximated edge references, direct pairs. */
typedef struct {
    int start, end;       /* range in the node-order array */
    int left, right;      /* children, -1 for a leaf */
    double cen[3], tan[3], mass, rx, rT;
} bh_node_t;

typedef struct {
    const double* C;      /* midpoints [E][3] */
    const double* T;      /* unit tangents (0 below 1e-14) [E][3] */
    const double* mass;   /* ell + eps */
    int* ord;             /* node-order array: each node's edges contiguous */
    int* tmp;
    double* key;
    bh_node_t* nodes;
    int nnodes, leaf;
} bh_build_t;

static void bh_stable_sort(int* a, int n, const double* key, int* tmp) {
    /* bottom-up merge sort on key[a[i]], equal keys keep their order */
    for (int w = 1; w < n; w *= 2) {
        for (int lo = 0; lo < n; lo += 2 * w) {
            int mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = (key[a[j]] < key[a[i]]) ? a[j++] : a[i++];
            while (i < mid) tmp[k++] = a[i++];
            while (j < hi) tmp[k++] = a[j++];
        }
        memcpy(a, tmp, sizeof(int) * (size_t)n);
    }
}

static int bh_build(bh_build_t* b, int start, int end) {
    const int n = end - start;
    const int id = b->nnodes++;
    bh_node_t* nd = &b->nodes[id];
    nd->start = start;
    nd->end = end;
    nd->left = nd->right = -1;
    double ms = 0.0, c[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    for (int p = start; p < end; p++) ms += b->mass[b->ord[p]];
    if (ms > 0) {
        for (int p = start; p < end; p++) {
            const int e = b->ord[p];
            for (int d = 0; d < 3; d++) {
                c[d] += b->mass[e] * b->C[3 * e + d];
                t[d] += b->mass[e] * b->T[3 * e + d];
            }
        }
        for (int d = 0; d < 3; d++) {
            c[d] /= ms;
            t[d] /= ms;
        }
    } else if (n > 0) {
        for (int p = start; p < end; p++)
            for (int d = 0; d < 3; d++) {
                c[d] += b->C[3 * b->ord[p] + d];
                t[d] += b->T[3 * b->ord[p] + d];
            }
        for (int d = 0; d < 3; d++) {
            c[d] /= n;
            t[d] /= n;
        }
    }
    double rx = 0.0, rT = 0.0;
    for (int p = start; p < end; p++) {
        const int e = b->ord[p];
        double sx = 0.0, st = 0.0;
        for (int d = 0; d < 3; d++) {
            const double u = b->C[3 * e + d] - c[d], w = b->T[3 * e + d] - t[d];
            sx += u * u;
            st += w * w;
        }
        if (sqrt(sx) > rx) rx = sqrt(sx);
        if (sqrt(st) > rT) rT = sqrt(st);
    }
    memcpy(nd->cen, c, sizeof(c));
    memcpy(nd->tan, t, sizeof(t));
    nd->mass = ms;
    nd->rx = rx;
    nd->rT = rT;
    if (n > b->leaf) {
        int dim = 0;
        double best = -1.0;
        for (int d = 0; d < 6; d++) {
            double lo = INFINITY, hi = -INFINITY;
            for (int p = start; p < end; p++) {
                const int e = b->ord[p];
                const double x = d < 3 ? b->C[3 * e + d] : b->T[3 * e + d - 3];
                if (x < lo) lo = x;
                if (x > hi) hi = x;
            }
            if (hi - lo > best) {
                best = hi - lo;
                dim = d;
            }
        }
        for (int p = start; p < end; p++) {
            const int e = b->ord[p];
            b->key[e] = dim < 3 ? b->C[3 * e + dim] : b->T[3 * e + dim - 3];
        }
        bh_stable_sort(b->ord + start, n, b->key, b->tmp);
        const int mid = n / 2;
        if (mid > 0 && n - mid > 0) {
            const int l = bh_build(b, start, start + mid);
            const int r = bh_build(b, start + mid, end);
            b->nodes[id].left = l;
            b->nodes[id].right = r;
        }
    }
    return id;
}

typedef struct {
    int V, E;
    const double* v;
    const int32_t* edges;
    double alpha, beta, eps, theta, theta_t;
    const bh_build_t* b;
    const double* C;
    long long visits, approx, direct;
} bh_trav_t;

static double bh_term_exact(const bh_trav_t* s, int I, int J) {
    const double* v = s->v;
    const int i1 = s->edges[2 * I], i2 = s->edges[2 * I + 1];
    const int j1 = s->edges[2 * J], j2 = s->edges[2 * J + 1];
    const double eI[3] = {v[3 * i2] - v[3 * i1], v[3 * i2 + 1] - v[3 * i1 + 1],
                          v[3 * i2 + 2] - v[3 * i1 + 2]};
    const double ellI = s->b->mass[I], ellJ = s->b->mass[J];
    const int is[2] = {i1, i2}, js[2] = {j1, j2};
    double sumK = 0.0;
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++) {
            const double d[3] = {v[3 * is[a]] - v[3 * js[c]], v[3 * is[a] + 1] - v[3 * js[c] + 1],
                                 v[3 * is[a] + 2] - v[3 * js[c] + 2]};
            const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + s->eps;
            const double cx = eI[1] * d[2] - eI[2] * d[1], cy = eI[2] * d[0] - eI[0] * d[2],
                         cz = eI[0] * d[1] - eI[1] * d[0];
            const double cn = sqrt(cx * cx + cy * cy + cz * cz) + s->eps;
            sumK += pow(cn, s->alpha) / pow(dn, s->beta);
        }
    return 0.25 * pow(ellI, 1.0 - s->alpha) * ellJ * sumK;
}

static double bh_term_cluster(const bh_trav_t* s, int I, const double* q, double mass) {
    const double* v = s->v;
    const int i1 = s->edges[2 * I], i2 = s->edges[2 * I + 1];
    const double eI[3] = {v[3 * i2] - v[3 * i1], v[3 * i2 + 1] - v[3 * i1 + 1],
                          v[3 * i2 + 2] - v[3 * i1 + 2]};
    const int is[2] = {i1, i2};
    double sum = 0.0;
    for (int a = 0; a < 2; a++) {
        const double d[3] = {v[3 * is[a]] - q[0], v[3 * is[a] + 1] - q[1], v[3 * is[a] + 2] - q[2]};
        const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + s->eps;
        const double cx = eI[1] * d[2] - eI[2] * d[1], cy = eI[2] * d[0] - eI[0] * d[2],
                     cz = eI[0] * d[1] - eI[1] * d[0];
        const double cn = sqrt(cx * cx + cy * cy + cz * cz) + s->eps;
        sum += pow(cn, s->alpha) / pow(dn, s->beta);
    }
    return 0.5 * pow(s->b->mass[I], 1.0 - s->alpha) * mass * sum;
}

static double bh_traverse(bh_trav_t* s, int I, int ni) {
    s->visits++;
    const bh_node_t* nd = &s->b->nodes[ni];
    const int n = nd->end - nd->start;
    if (n == 0) return 0.0;
    const double dx = s->C[3 * I] - nd->cen[0], dy = s->C[3 * I + 1] - nd->cen[1],
                 dz = s->C[3 * I + 2] - nd->cen[2];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    int admiss = n >= 2;
    if (admiss)
        for (int p = nd->start; p < nd->end; p++)
            if (!disjoint(s->edges, I, s->b->ord[p])) {
                admiss = 0;
                break;
            }
    admiss = admiss && dist > 0.0 && nd->rx / dist <= s->theta && nd->rT <= s->theta_t;
    if (admiss) {
        s->approx += n;
        return bh_term_cluster(s, I, nd->cen, nd->mass);
    }
    if (nd->left < 0) {
        double val = 0.0;
        for (int p = nd->start; p < nd->end; p++) {
            const int J = s->b->ord[p];
            if (!disjoint(s->edges, I, J)) continue;
            s->direct++;
            val += bh_term_exact(s, I, J);
        }
        return val;
    }
    return bh_traverse(s, I, nd->left) + bh_traverse(s, I, nd->right);
}

double rco_energy_bh(int V, int E, const double* v, const int32_t* edges, double alpha,
                     double beta, double eps, double theta, double theta_t, int leaf_size,
                     int64_t* stats) {
    double* C = (double*)malloc(sizeof(double) * 3 * (size_t)(E > 0 ? E : 1));
    double* T = (double*)malloc(sizeof(double) * 3 * (size_t)(E > 0 ? E : 1));
    double* mass = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    for (int K = 0; K < E; K++) {
        const int a = edges[2 * K], b = edges[2 * K + 1];
        double e[3], r2 = 0.0;
        for (int d = 0; d < 3; d++) {
            e[d] = v[3 * b + d] - v[3 * a + d];
            r2 += e[d] * e[d];
            C[3 * K + d] = 0.5 * (v[3 * a + d] + v[3 * b + d]);
        }
        const double r = sqrt(r2);
        mass[K] = r + eps;
        for (int d = 0; d < 3; d++) T[3 * K + d] = (r >= 1e-14) ? e[d] / r : 0.0;
    }
    bh_build_t b;
    b.C = C;
    b.T = T;
    b.mass = mass;
    b.ord = (int*)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
    b.tmp = (int*)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
    b.key = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    b.nodes = (bh_node_t*)malloc(sizeof(bh_node_t) * (size_t)(2 * E + 1));
    b.nnodes = 0;
    b.leaf = leaf_size;
    for (int K = 0; K < E; K++) b.ord[K] = K;
    bh_build(&b, 0, E);
    bh_trav_t s = {V, E, v, edges, alpha, beta, eps, theta, theta_t, &b, C, 0, 0, 0};
    double total = 0.0;
    for (int I = 0; I < E; I++) total += bh_traverse(&s, I, 0);
    if (stats) {
        stats[0] = s.visits;
        stats[1] = s.approx;
        stats[2] = s.direct;
    }
    free(C);
    free(T);
    free(mass);
    free(b.ord);
    free(b.tmp);
    free(b.key);
    free(b.nodes);
    return 0.5 * total;
}

/* ------------------------------------------------------------------------ */
/* Block-cluster-tree products of the two edge kernels — the reference's
 * stage-2 oracle bct_matvec_kernel, bct_kernel_node, exact_edge_kernel_pair
 * and bct_apply_B_scalar (oracle/tpe_stage2_oracle.py:598-795). Kernel
 * 'high' (kind 0): k_IJ = 2 * 1/4 ell_I ell_J sum_ij (|p_i - q_j| + eps)^-p;
 * 'low' (kind 1): 1/4 ell_I ell_J sum_ij [(|T_I x d| + eps)^2 + (|T_J x -d| +
 * eps)^2] (|d| + eps)^-(4+p), p = 7/3, disjoint pairs only. Dual recursion
 * from (root, root): a block of >= 3 edges in total, no pair sharing a vertex,
 * 0 < dist of the centres, max rx / dist <= theta, max rT <= theta_t is one
 * rank-one term kernel(centres, mean tangents) * mass_I * sum_J mass_J psi_J;
 * two leaves otherwise sum the exact pairs; else the target splits when it is
 * not a leaf and the source is a leaf or not larger, else the source. */
typedef struct {
    const double* v;
    const int32_t* edges;
    double eps, theta, theta_t;
    int kind, ncomp;
    const bh_build_t* b;
    const double* psi;
    double* phi;
    long long leaves, adm, dir, pairs;
} bct_t;

static double bct_exact_pair(const bct_t* s, int I, int J) {
    if (!disjoint(s->edges, I, J)) return 0.0;
    const double* v = s->v;
    const double p = 7.0 / 3.0;
    const int is[2] = {s->edges[2 * I], s->edges[2 * I + 1]};
    const int js[2] = {s->edges[2 * J], s->edges[2 * J + 1]};
    const double ellI = s->b->mass[I], ellJ = s->b->mass[J];
    const double* TI = s->b->T + 3 * I;
    const double* TJ = s->b->T + 3 * J;
    double sh = 0.0, sIJ = 0.0, sJI = 0.0;
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++) {
            const double d[3] = {v[3 * is[a]] - v[3 * js[c]], v[3 * is[a] + 1] - v[3 * js[c] + 1],
                                 v[3 * is[a] + 2] - v[3 * js[c] + 2]};
            const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + s->eps;
            if (s->kind == 0) {
                sh += 1.0 / pow(dn, p);
            } else {
                const double c1x = TI[1] * d[2] - TI[2] * d[1], c1y = TI[2] * d[0] - TI[0] * d[2],
                             c1z = TI[0] * d[1] - TI[1] * d[0];
                const double c2x = -(TJ[1] * d[2] - TJ[2] * d[1]),
                             c2y = -(TJ[2] * d[0] - TJ[0] * d[2]),
                             c2z = -(TJ[0] * d[1] - TJ[1] * d[0]);
                const double n1 = sqrt(c1x * c1x + c1y * c1y + c1z * c1z) + s->eps;
                const double n2 = sqrt(c2x * c2x + c2y * c2y + c2z * c2z) + s->eps;
                sIJ += pow(n1, 2.0) / pow(dn, 4.0 + p);
                sJI += pow(n2, 2.0) / pow(dn, 4.0 + p);
            }
        }
    return s->kind == 0 ? 2.0 * 0.25 * ellI * ellJ * sh : 0.25 * ellI * ellJ * (sIJ + sJI);
}

static double bct_node_kernel(const bct_t* s, const bh_node_t* t, const bh_node_t* r) {
    const double d[3] = {t->cen[0] - r->cen[0], t->cen[1] - r->cen[1], t->cen[2] - r->cen[2]};
    const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + s->eps;
    if (s->kind == 0) return 2.0 / pow(dn, 7.0 / 3.0);
    const double c1x = t->tan[1] * d[2] - t->tan[2] * d[1], c1y = t->tan[2] * d[0] - t->tan[0] * d[2],
                 c1z = t->tan[0] * d[1] - t->tan[1] * d[0];
    const double c2x = -(r->tan[1] * d[2] - r->tan[2] * d[1]),
                 c2y = -(r->tan[2] * d[0] - r->tan[0] * d[2]),
                 c2z = -(r->tan[0] * d[1] - r->tan[1] * d[0]);
    const double c1 = sqrt(c1x * c1x + c1y * c1y + c1z * c1z) + s->eps;
    const double c2 = sqrt(c2x * c2x + c2y * c2y + c2z * c2z) + s->eps;
    return (pow(c1, 2.0) + pow(c2, 2.0)) / pow(dn, 4.0 + 7.0 / 3.0);
}

static void bct_rec(bct_t* s, int ti, int si) {
    const bh_node_t* t = &s->b->nodes[ti];
    const bh_node_t* r = &s->b->nodes[si];
    const int nt = t->end - t->start, ns = r->end - r->start;
    const double dx = t->cen[0] - r->cen[0], dy = t->cen[1] - r->cen[1], dz = t->cen[2] - r->cen[2];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    int well = nt + ns >= 3;
    if (well)
        for (int p = t->start; p < t->end && well; p++)
            for (int q = r->start; q < r->end; q++)
                if (!disjoint(s->edges, s->b->ord[p], s->b->ord[q])) {
                    well = 0;
                    break;
                }
    well = well && dist > 0.0 && fmax(t->rx, r->rx) / dist <= s->theta &&
           fmax(t->rT, r->rT) <= s->theta_t;
    const int nc = s->ncomp;
    if (well) {
        s->leaves++;
        s->adm++;
        const double kval = bct_node_kernel(s, t, r);
        double tmp[3] = {0, 0, 0};
        for (int q = r->start; q < r->end; q++) {
            const int J = s->b->ord[q];
            for (int c = 0; c < nc; c++) tmp[c] += s->b->mass[J] * s->psi[nc * J + c];
        }
        for (int p = t->start; p < t->end; p++) {
            const int I = s->b->ord[p];
            for (int c = 0; c < nc; c++) s->phi[nc * I + c] += s->b->mass[I] * kval * tmp[c];
        }
        return;
    }
    if (t->left < 0 && r->left < 0) {
        s->leaves++;
        s->dir++;
        for (int p = t->start; p < t->end; p++) {
            const int I = s->b->ord[p];
            for (int q = r->start; q < r->end; q++) {
                const int J = s->b->ord[q];
                if (!disjoint(s->edges, I, J)) continue;
                const double k = bct_exact_pair(s, I, J);
                for (int c = 0; c < nc; c++) s->phi[nc * I + c] += k * s->psi[nc * J + c];
                s->pairs++;
            }
        }
        return;
    }
    const int split_t = t->left >= 0 && (r->left < 0 || nt >= ns);
    if (split_t) {
        bct_rec(s, t->left, si);
        bct_rec(s, t->right, si);
    } else {
        bct_rec(s, ti, r->left);
        bct_rec(s, ti, r->right);
    }
}

/* geometry + tree shared by the products */
static void bct_prepare(int E, const double* v, const int32_t* edges, double eps, int leaf,
                        bh_build_t* b, double** C, double** T, double** mass) {
    *C = (double*)malloc(sizeof(double) * 3 * (size_t)(E > 0 ? E : 1));
    *T = (double*)malloc(sizeof(double) * 3 * (size_t)(E > 0 ? E : 1));
    *mass = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    for (int K = 0; K < E; K++) {
        const int a = edges[2 * K], c = edges[2 * K + 1];
        double e[3], r2 = 0.0;
        for (int d = 0; d < 3; d++) {
            e[d] = v[3 * c + d] - v[3 * a + d];
            r2 += e[d] * e[d];
            (*C)[3 * K + d] = 0.5 * (v[3 * a + d] + v[3 * c + d]);
        }
        const double r = sqrt(r2);
        (*mass)[K] = r + eps;
        for (int d = 0; d < 3; d++) (*T)[3 * K + d] = (r >= 1e-14) ? e[d] / r : 0.0;
    }
    b->C = *C;
    b->T = *T;
    b->mass = *mass;
    b->ord = (int*)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
    b->tmp = (int*)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
    b->key = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    b->nodes = (bh_node_t*)malloc(sizeof(bh_node_t) * (size_t)(2 * E + 1));
    b->nnodes = 0;
    b->leaf = leaf;
    for (int K = 0; K < E; K++) b->ord[K] = K;
    bh_build(b, 0, E);
}

static void bct_release(bh_build_t* b, double* C, double* T, double* mass) {
    free(C);
    free(T);
    free(mass);
    free(b->ord);
    free(b->tmp);
    free(b->key);
    free(b->nodes);
}

static void bct_run(int E, const double* v, const int32_t* edges, double eps, const double* psi,
                    int ncomp, int kind, double theta, double theta_t, const bh_build_t* b,
                    double* phi, int64_t* stats) {
    bct_t s = {v, edges, eps, theta, theta_t, kind, ncomp, b, psi, phi, 0, 0, 0, 0};
    memset(phi, 0, sizeof(double) * (size_t)ncomp * (size_t)E);
    if (E > 0) bct_rec(&s, 0, 0);
    if (stats) {
        stats[0] = s.leaves;
        stats[1] = s.adm;
        stats[2] = s.dir;
        stats[3] = s.pairs;
    }
}

void rco_bct_matvec(int V, int E, const double* v, const int32_t* edges, double eps,
                    const double* psi, int ncomp, int kind, double theta, double theta_t,
                    int leaf_size, double* phi, int64_t* stats) {
    (void)V;
    bh_build_t b;
    double *C, *T, *mass;
    bct_prepare(E, v, edges, eps, leaf_size, &b, &C, &T, &mass);
    bct_run(E, v, edges, eps, psi, ncomp, kind, theta, theta_t, &b, phi, stats);
    bct_release(&b, C, T, mass);
}

void rco_bct_apply_b(int V, int E, const double* v, const int32_t* edges, double eps,
                     const double* u, int kind, double theta, double theta_t, int leaf_size,
                     double* out) {
    bh_build_t b;
    double *C, *T, *mass;
    bct_prepare(E, v, edges, eps, leaf_size, &b, &C, &T, &mass);
    const int nc = kind == 0 ? 3 : 1;
    double* ones = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    double* y = (double*)malloc(sizeof(double) * 3 * (size_t)(E > 0 ? E : 1));
    double* K1 = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    double* Ky = (double*)malloc(sizeof(double) * 3 * (size_t)(E > 0 ? E : 1));
    for (int I = 0; I < E; I++) {
        ones[I] = 1.0;
        const int a = edges[2 * I], c = edges[2 * I + 1];
        if (kind == 0) {
            const double du = (u[c] - u[a]) / mass[I];
            for (int d = 0; d < 3; d++) y[3 * I + d] = du * T[3 * I + d];
        } else {
            y[I] = 0.5 * (u[a] + u[c]);
        }
    }
    bct_run(E, v, edges, eps, ones, 1, kind, theta, theta_t, &b, K1, NULL);
    bct_run(E, v, edges, eps, y, nc, kind, theta, theta_t, &b, Ky, NULL);
    memset(out, 0, sizeof(double) * (size_t)V);
    for (int I = 0; I < E; I++) {
        const int a = edges[2 * I], c = edges[2 * I + 1];
        if (kind == 0) {
            double dot = 0.0;
            for (int d = 0; d < 3; d++) dot += (K1[I] * y[3 * I + d] - Ky[3 * I + d]) * T[3 * I + d];
            const double coeff = dot / mass[I];
            out[a] -= coeff;
            out[c] += coeff;
        } else {
            const double z = K1[I] * y[I] - Ky[I];
            out[a] += 0.5 * z;
            out[c] += 0.5 * z;
        }
    }
    free(ones);
    free(y);
    free(K1);
    free(Ky);
    bct_release(&b, C, T, mass);
}

/* Forward difference of the Barnes-Hut energy (the reference's stage-2
 * oracle: finite_difference_gradient of bh_energy, tpe_stage2_oracle.py:
 * 133-141, 1289-1290): the tree is rebuilt for every moved coordinate. */
void rco_gradient_bh_fd(int V, int E, const double* v, const int32_t* edges, double alpha,
                        double beta, double eps, double theta, double theta_t, int leaf_size,
                        double h, double* g) {
    const double e0 = rco_energy_bh(V, E, v, edges, alpha, beta, eps, theta, theta_t, leaf_size,
                                    NULL);
    double* pert = (double*)malloc(sizeof(double) * 3 * (size_t)(V > 0 ? V : 1));
    for (int i = 0; i < V; i++)
        for (int d = 0; d < 3; d++) {
            memcpy(pert, v, sizeof(double) * 3 * (size_t)V);
            pert[3 * i + d] += h;
            const double e1 = rco_energy_bh(V, E, pert, edges, alpha, beta, eps, theta, theta_t,
                                            leaf_size, NULL);
            g[3 * i + d] = (e1 - e0) / h;
        }
    free(pert);
}

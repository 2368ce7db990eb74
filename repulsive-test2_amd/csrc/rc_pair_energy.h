// Tangent-point energy of one unordered disjoint edge pair {I, J}: both
// ordered terms of tangentPointEnergy.ts:60-105,
//   1/4 l_I^(1-alpha) l_J sum_{i in I, j in J} (|e_I x (g_i - g_j)| + eps)^alpha
//                                               / (|g_i - g_j| + eps)^beta
// plus the same with I and J exchanged (eps after each norm, l = |e| + eps).
// e_I x (g_i - g_j) is the same for both endpoints i of I (the two
// differences differ by e_I, and e_I x e_I = 0), so the numerator of the
// I-ordered terms depends on (I, j) only and that of the J-ordered terms on
// (i, J) only: 4 cross-product norms per pair instead of 8 (the values agree
// with the per-endpoint form to rounding).
#pragma once

#include "rc_internal.h"
#include "rc_math.h"

namespace rcpair {

template <class P>
__device__ __forceinline__ double pair_energy(const P& pw, double eps, const double* PI,
                                              const double* PJ, double eIx, double eIy,
                                              double eIz, double eJx, double eJy, double eJz,
                                              double ellI, double ellJ, double lpI, double lpJ) {
    double nI[2], nJ[2], invB[2][2];
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const double dx = PI[0] - PJ[3 * jb];
        const double dy = PI[1] - PJ[3 * jb + 1];
        const double dz = PI[2] - PJ[3 * jb + 2];
        const double cx = eIy * dz - eIz * dy;
        const double cy = eIz * dx - eIx * dz;
        const double cz = eIx * dy - eIy * dx;
        nI[jb] = pw.c_alpha(rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps);
    }
#pragma unroll
    for (int ia = 0; ia < 2; ++ia) {
        const double dx = PI[3 * ia] - PJ[0];
        const double dy = PI[3 * ia + 1] - PJ[1];
        const double dz = PI[3 * ia + 2] - PJ[2];
        const double cx = eJy * dz - eJz * dy;
        const double cy = eJz * dx - eJx * dz;
        const double cz = eJx * dy - eJy * dx;
        nJ[ia] = pw.c_alpha(rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps);
    }
#pragma unroll
    for (int ia = 0; ia < 2; ++ia)
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
            const double dx = PI[3 * ia] - PJ[3 * jb];
            const double dy = PI[3 * ia + 1] - PJ[3 * jb + 1];
            const double dz = PI[3 * ia + 2] - PJ[3 * jb + 2];
            const double deps = rcm::sqrt_(dx * dx + dy * dy + dz * dz) + eps;
            invB[ia][jb] = pw.inv_beta(deps, rcm::rcp(deps));
        }
    const double sFI = nI[0] * (invB[0][0] + invB[1][0]) + nI[1] * (invB[0][1] + invB[1][1]);
    const double sFJ = nJ[0] * (invB[0][0] + invB[0][1]) + nJ[1] * (invB[1][0] + invB[1][1]);
    return 0.25 * (lpI * ellJ * sFI + lpJ * ellI * sFJ);
}

// pair_energy split for the forward-difference gradient: moving one endpoint S
// of I changes e_I (so both I-ordered numerators), the J-ordered numerator of
// endpoint S and the two distances from endpoint S; the J-ordered numerator
// and the distances of the other endpoint are reused from the unmoved pair.
struct PairParts {
    double nJ[2];
    double invB[2][2];
};

template <class P>
__device__ __forceinline__ double pair_energy_parts(const P& pw, double eps, const double* PI,
                                                    const double* PJ, double eJx, double eJy,
                                                    double eJz, double ellI, double ellJ,
                                                    double lpI, double lpJ, PairParts& q) {
    const double eIx = PI[3] - PI[0], eIy = PI[4] - PI[1], eIz = PI[5] - PI[2];
    double nI[2];
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const double dx = PI[0] - PJ[3 * jb];
        const double dy = PI[1] - PJ[3 * jb + 1];
        const double dz = PI[2] - PJ[3 * jb + 2];
        const double cx = eIy * dz - eIz * dy;
        const double cy = eIz * dx - eIx * dz;
        const double cz = eIx * dy - eIy * dx;
        nI[jb] = pw.c_alpha(rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps);
    }
#pragma unroll
    for (int ia = 0; ia < 2; ++ia) {
        const double dx = PI[3 * ia] - PJ[0];
        const double dy = PI[3 * ia + 1] - PJ[1];
        const double dz = PI[3 * ia + 2] - PJ[2];
        const double cx = eJy * dz - eJz * dy;
        const double cy = eJz * dx - eJx * dz;
        const double cz = eJx * dy - eJy * dx;
        q.nJ[ia] = pw.c_alpha(rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps);
    }
#pragma unroll
    for (int ia = 0; ia < 2; ++ia)
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
            const double dx = PI[3 * ia] - PJ[3 * jb];
            const double dy = PI[3 * ia + 1] - PJ[3 * jb + 1];
            const double dz = PI[3 * ia + 2] - PJ[3 * jb + 2];
            const double deps = rcm::sqrt_(dx * dx + dy * dy + dz * dz) + eps;
            q.invB[ia][jb] = pw.inv_beta(deps, rcm::rcp(deps));
        }
    const double sFI = nI[0] * (q.invB[0][0] + q.invB[1][0]) + nI[1] * (q.invB[0][1] + q.invB[1][1]);
    const double sFJ = q.nJ[0] * (q.invB[0][0] + q.invB[0][1]) + q.nJ[1] * (q.invB[1][0] + q.invB[1][1]);
    return 0.25 * (lpI * ellJ * sFI + lpJ * ellI * sFJ);
}

// pair_energy of I with endpoint S moved (positions Q, length ellQ, lpQ =
// ellQ^(1-alpha)), reusing the unchanged parts q of the unmoved pair.
template <int S, class P>
__device__ __forceinline__ double pair_energy_moved(const P& pw, double eps, const double* Q,
                                                    const double* PJ, double eJx, double eJy,
                                                    double eJz, double ellQ, double ellJ,
                                                    double lpQ, double lpJ, const PairParts& q) {
    const double eIx = Q[3] - Q[0], eIy = Q[4] - Q[1], eIz = Q[5] - Q[2];
    double nI[2], nJ[2], invB[2][2];
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const double dx = Q[0] - PJ[3 * jb];
        const double dy = Q[1] - PJ[3 * jb + 1];
        const double dz = Q[2] - PJ[3 * jb + 2];
        const double cx = eIy * dz - eIz * dy;
        const double cy = eIz * dx - eIx * dz;
        const double cz = eIx * dy - eIy * dx;
        nI[jb] = pw.c_alpha(rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps);
    }
    {
        const double dx = Q[3 * S] - PJ[0];
        const double dy = Q[3 * S + 1] - PJ[1];
        const double dz = Q[3 * S + 2] - PJ[2];
        const double cx = eJy * dz - eJz * dy;
        const double cy = eJz * dx - eJx * dz;
        const double cz = eJx * dy - eJy * dx;
        nJ[S] = pw.c_alpha(rcm::sqrt_(cx * cx + cy * cy + cz * cz) + eps);
        nJ[1 - S] = q.nJ[1 - S];
    }
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const double dx = Q[3 * S] - PJ[3 * jb];
        const double dy = Q[3 * S + 1] - PJ[3 * jb + 1];
        const double dz = Q[3 * S + 2] - PJ[3 * jb + 2];
        const double deps = rcm::sqrt_(dx * dx + dy * dy + dz * dz) + eps;
        invB[S][jb] = pw.inv_beta(deps, rcm::rcp(deps));
        invB[1 - S][jb] = q.invB[1 - S][jb];
    }
    const double sFI = nI[0] * (invB[0][0] + invB[1][0]) + nI[1] * (invB[0][1] + invB[1][1]);
    const double sFJ = nJ[0] * (invB[0][0] + invB[0][1]) + nJ[1] * (invB[1][0] + invB[1][1]);
    return 0.25 * (lpQ * ellJ * sFI + lpJ * ellQ * sFJ);
}

// One wave's J tile in LDS (structure of arrays, RC_TILE edges).
struct ETile {
    double P[6][RC_TILE];
    double ell[RC_TILE], lp[RC_TILE];
    int32_t i1[RC_TILE], i2[RC_TILE];
};

// Energy of the unordered tile pair (a, b), a <= b, for this wave: lane t
// owns edge I = 64a + t and walks the edges J of tile b (J > I on the
// diagonal tile) in order; disjoint pairs only. The J tile is staged in the
// wave's LDS tile T with one coalesced load per field, then read as LDS
// broadcasts. Returns the lane's sum (the caller reduces over the wave).
template <class P>
__device__ __forceinline__ double tile_pair_energy(const EdgeSoA& ed, int E, const P& pw,
                                                   double eps, int a, int b, int lane, ETile& T) {
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
    const double eIx = PI[3] - PI[0], eIy = PI[4] - PI[1], eIz = PI[5] - PI[2];
    const double ellI = ed.ell[Ic], lpI = ed.lpow[Ic];
    const int i1 = ed.i1[Ic], i2 = ed.i2[Ic];
    const bool diag = (a == b);
    const int ns = min(RC_TILE, E - b * RC_TILE);
    double sum = 0.0;
#pragma unroll 2
    for (int s = 0; s < ns; ++s) {
        const double PJ[6] = {T.P[0][s], T.P[1][s], T.P[2][s], T.P[3][s], T.P[4][s], T.P[5][s]};
        const int j1 = T.i1[s], j2 = T.i2[s];
        const bool valid = Iok && (!diag || s > lane) && i1 != j1 && i1 != j2 && i2 != j1 &&
                           i2 != j2;
        const double e = pair_energy(pw, eps, PI, PJ, eIx, eIy, eIz, PJ[3] - PJ[0],
                                     PJ[4] - PJ[1], PJ[5] - PJ[2], ellI, T.ell[s], lpI, T.lp[s]);
        sum += valid ? e : 0.0;
    }
    // the tile is rewritten by the wave's next tile pair
    __builtin_amdgcn_wave_barrier();
    return sum;
}

}  // namespace rcpair

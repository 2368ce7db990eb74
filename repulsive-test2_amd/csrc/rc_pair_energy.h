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

}  // namespace rcpair

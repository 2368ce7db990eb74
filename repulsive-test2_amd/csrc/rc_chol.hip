// Dense SPD factorisation of A_hat = A + mu 1 1^T and multi-RHS solves.
//
// The reference factors the (3V+k)^2 saddle matrix [[I3 (x) A, C^T],[C, 0]]
// with Bunch-Kaufman LDL^T (src/core/sobolev/linsolve.ts:431-589, 742-803).
// A is PSD with null space span(1) (its rows sum to zero), so A_hat is SPD and
// the saddle solution follows exactly from A_hat^{-1} applied to 3(k+1)
// right-hand sides plus a (k+3)^2 system (see rc_step.hip). Only the V x V
// Cholesky factor is formed: V^3/3 flops instead of (3V+k)^3/3.
//
// Storage: F is Vp x Vp row-major (Vp = nblk * 64, nblk even), lower
// triangle, element type T: double for the plain form, float for the mixed
// form (which factors the diagonally scaled S = D^-1/2 A_hat D^-1/2, unit
// diagonal). Padding rows/columns are the identity.
//
// FP64 form, right-looking with 128-column panels (3 launches per panel):
//   k_potrf128    one workgroup factors the 128x128 diagonal block in LDS
//                 (16-column sub-panels, MFMA updates) and inverts it
//   k_panel_trsm  L_ik = A_ik L_kk^-T as a GEMM with the stored inverse
//   k_syrk128     A_ij -= L_ik L_jk^T on 128x64 lower blocks, K = 128,
//                 v_mfma_f64_16x16x4_f64, XCD-grouped block order, on a second
//                 stream; the next panel's column goes first in the same
//                 launch (its diagonal tile is updated by k_col_diag on the
//                 critical path), so the next panel's factorisation overlaps
//                 the bulk of the update
// Mixed form, 64-column panels:
//   k_potrf_panel<float>  every block factors the 64x64 diagonal block in
//                 FP64 in LDS and inverts it; block 0 stores L_kk^-1, the
//                 others form their rows L_rk = A_rk L_kk^-T
//   k_syrk_f16x3  trailing update on the FP32 factor: each operand is split
//                 into fp16 hi + lo (x = hi + lo to ~2^-22) and
//                 hi*hi^T + hi*lo^T + lo*hi^T runs on
//                 v_mfma_f32_16x16x32_f16 with FP32 accumulation
// After either factorisation the blocks below the 128 x 128 block diagonal
// are row-scaled by the diagonal blocks' inverses (rc_sweep.hip), which the
// triangular solves read.
#include "rc_internal.h"
#include "rc_math.h"
#include "rc_handoff.h"
#include <utility>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int NB = RC_NB;

// mu = trace(A) / V^2; the diagonal of A is kept in diag (the residual reads
// A from the upper triangle of F and this diagonal once F holds the factor).
__global__ void k_trace_mu(const double* __restrict__ A, int V, int ldA, double* __restrict__ mu,
                           double* __restrict__ diag) {
    __shared__ double sh[1024];
    double s = 0.0;
#pragma unroll 8
    for (int i = threadIdx.x; i < V; i += 1024) {
        const double a = A[(size_t)i * ldA + i];
        diag[i] = a;
        s += a;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) mu[0] = sh[0] / ((double)V * (double)V);
}

// dinv[i] = 1/sqrt(A_ii + mu) (1 on padding)
__global__ void k_diag_scale(const double* __restrict__ A, int V, int ldA, int Vp,
                             const double* __restrict__ mu, double* __restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Vp) return;
    dinv[i] = (i < V) ? 1.0 / __builtin_sqrt(A[(size_t)i * ldA + i] + mu[0]) : 1.0;
}

// F32 (lower) = dinv_i (A + mu) dinv_j + shift delta_ij (mixed form; A in
// the FP64 buffer, leading dimension ldA); identity padding.
template <typename T>
__global__ void k_factor_prep(const double* __restrict__ A, int V, int ldA, int Vp,
                              const double* __restrict__ mu, const double* __restrict__ dinv,
                              double shift, T* __restrict__ F) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= Vp) return;
    double v;
    if (i < V && j < V) {
        v = (j <= i) ? A[(size_t)i * ldA + j] + mu[0] : 0.0;
        if (dinv) v = v * dinv[i] * dinv[j];
        if (i == j) v += shift;
    } else {
        v = (i == j) ? 1.0 : 0.0;
    }
    F[(size_t)i * Vp + j] = (T)v;
}

// FP64 factor storage, padding rows V <= i < Vp: identity (written before
// every factorisation, so that a breakdown of an earlier one cannot leave
// non-finite values there). Their factor is the identity again.
__global__ void k_pad_identity(double* __restrict__ F, int V, int Vp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = V + (int)blockIdx.y;
    if (j >= Vp) return;
    F[(size_t)i * Vp + j] = (i == j) ? 1.0 : 0.0;
}

// Factor the 64x64 diagonal block held in D (lower) in place and write its
// inverse into Li. Breakdown (non-positive or NaN pivot) sets *bad.
__device__ void potrf64_lds(double (*D)[NB + 1], double (*Li)[NB + 1], int* bad) {
    const int t = threadIdx.x;
    for (int j = 0; j < NB; ++j) {
        __syncthreads();
        double piv = D[j][j];
        if (!(piv > 0.0)) {
            if (t == 0) *bad = 1;
            piv = 1.0;
        }
        const double s = __builtin_sqrt(piv);
        const double inv = 1.0 / s;
        __syncthreads();
        if (t == 0) D[j][j] = s;
        if (t > j && t < NB) D[t][j] *= inv;
        __syncthreads();
        const int i = j + 1 + (t >> 2);
        if (i < NB) {
            const double lij = D[i][j];
            for (int m = j + 1 + (t & 3); m <= i; m += 4) D[i][m] -= lij * D[m][j];
        }
    }
    __syncthreads();
    if (t < NB) {
        const int j = t;
        for (int i = 0; i < j; ++i) Li[i][j] = 0.0;
        Li[j][j] = 1.0 / D[j][j];
        for (int i = j + 1; i < NB; ++i) {
            double s = 0.0;
            for (int m = j; m < i; ++m) s += D[i][m] * Li[m][j];
            Li[i][j] = -s / D[i][i];
        }
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(256) void k_potrf_panel(T* __restrict__ F, int Vp, int kb,
                                                     T* __restrict__ Linv,
                                                     int32_t* __restrict__ status, int bad_bit) {
    __shared__ double D[NB][NB + 1];
    __shared__ double Li[NB][NB + 1];
    __shared__ double R[NB][NB + 1];
    __shared__ int bad;
    const int t = threadIdx.x;
    const int k0 = kb * NB;
    if (t == 0) bad = 0;
    for (int q = t; q < NB * NB; q += 256) {
        const int i = q / NB, j = q % NB;
        D[i][j] = (j <= i) ? (double)F[(size_t)(k0 + i) * Vp + k0 + j] : 0.0;
    }
    potrf64_lds(D, Li, &bad);
    if (blockIdx.x == 0) {
        // Only L_kk^-1 is stored: the solves use it for the diagonal blocks and
        // never read F's diagonal blocks, which the other blocks of this launch
        // are still reading (so they must not be overwritten here).
        for (int q = t; q < NB * NB; q += 256) {
            const int i = q / NB, j = q % NB;
            Linv[((size_t)kb * NB + i) * NB + j] = (T)Li[i][j];
        }
        if (t == 0 && bad) atomicOr(status, bad_bit);
        return;
    }
    const int r0 = (kb + blockIdx.x) * NB;
    for (int q = t; q < NB * NB; q += 256) {
        const int i = q / NB, j = q % NB;
        R[i][j] = (double)F[(size_t)(r0 + i) * Vp + k0 + j];
    }
    __syncthreads();
    // X = R Li^T : X[i][j] = sum_m R[i][m] Li[j][m]
    const int ti = t >> 4, tj = t & 15;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int m = 0; m < NB; ++m) {
        double rv[4], lv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) rv[a] = R[4 * ti + a][m];
#pragma unroll
        for (int b = 0; b < 4; ++b) lv[b] = Li[tj + 16 * b][m];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(rv[a], lv[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            F[(size_t)(r0 + 4 * ti + a) * Vp + k0 + tj + 16 * b] = (T)acc[a][b];
}

__device__ __forceinline__ void decode_lower(int p, int& ti, int& tj) {
    int a = (int)((__builtin_sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((a + 1) * (a + 2) / 2 <= p) ++a;
    while (a * (a + 1) / 2 > p) --a;
    ti = a;
    tj = p - a * (a + 1) / 2;
}

// ---- FP64 form: 128-column panels --------------------------------------
constexpr int PB = 128;       // panel width of the FP64 factor
constexpr int KC_TRSM = 16;   // K chunk staged through LDS: panel TRSM and trailing update
// K chunk of the trailing update (2 workgroups/CU): 8 leaves its registers
// unspilled (16 spilled 14 VGPRs: 76.1 against 77.3 us per launch); a 4
// chunk takes 81 us, a 32 chunk at 1 workgroup/CU measured slower
constexpr int KC_SYRK = 8;
constexpr int LDT = PB + 4;   // LDS row length of the k-major operand tiles

// Factor the 128x128 diagonal block of panel kb and invert it, in one
// workgroup of 4 waves, block held in LDS. 16-column sub-panels p = 0..7,
// pipelined so that the pivots of sub-panel p+1 overlap the rest of p:
//   (b) wave 0 factors the 16x16 diagonal block in registers (lane (g, c) =
//       (lane/16, lane%16) keeps rows 4g..4g+3 of column c; pivots exchanged
//       by readlane and lane shuffles) and forms its inverse D_p in the same
//       sweep (right-looking substitution of L_pp X = I, applying row j-1 of
//       X at pivot j);
//   (c) all waves: L_ip = A_ip D_p^T for the sub-panel rows below (MFMA);
//   (d) wave 0: A_{p+1,p+1} -= L_{p+1,p} L_{p+1,p}^T, then (b) for p+1;
//       waves 1-3 meanwhile: the other trailing 16x16 lower tiles (MFMA) and
//       block row p of X = L^-1, X_pq = -D_p sum_{k=q}^{p-1} L_pk X_kq (MFMA;
//       each X_pq kept transposed in the unused upper 16x16 block (q, p) of
//       the LDS matrix).
// Writes L (lower) into F, L^-1 into Linv128[kb] and the two 64x64 diagonal
// blocks of L^-1 (the inverses of L's 64x64 diagonal blocks) into
// Linv64[2kb], [2kb+1].
// MFMA operands: NT products fetch both operands as P[row = lane%16][k =
// 4kk + lane/16]; NN products fetch B as Q[4kk + lane/16][lane%16]; result
// element r of a lane is D[4r + lane/16][lane%16] (this is also the B
// operand layout of k-step r, which chains two NN products in registers).
constexpr int SB = 16;
constexpr int LDA = PB + 4;

// sqrt(x) and 1/sqrt(x) for the pivots (x > 0): hardware rsqrt seed, two
// Newton steps on y = x^-1/2 (quadratic: seed error ~2^-23 -> rounding
// level), then one correction of s = x y; 9 dependent operations.
__device__ __forceinline__ void pivot_sqrt(double x, double& s, double& is) {
    const double hx = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-hx, y * y, 1.5);
    y = y * fma(-hx, y * y, 1.5);
    s = x * y;
    s = fma(fma(-s, s, x), 0.5 * y, s);
    is = y;
}

// 64-bit readlane (the value of lane `src`, wave-uniform)
__device__ __forceinline__ double readlane_f64(double v, int src) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// 64-bit lane gather through the LDS crossbar (addr = 4 x source lane)
__device__ __forceinline__ double bperm_f64(double v, int addr) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)u);
    const unsigned hi = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Every lane gets the value of lane (g, J) of its own 16-lane row g
// (DPP row_newbcast: a VALU move, no LDS round trip).
template <int J>
__device__ __forceinline__ double row_bcast(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_update_dpp(0u, (unsigned)u, 0x150 + J, 0xf, 0xf, false);
    const unsigned hi = __builtin_amdgcn_update_dpp(0u, (unsigned)(u >> 32), 0x150 + J, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <class Fn, int... J>
__device__ __forceinline__ void static_for(Fn&& fn, std::integer_sequence<int, J...>) {
    (fn(std::integral_constant<int, J>{}), ...);
}

// State of the 16 x 16 diagonal-block factorisation on one wave, lane
// (g, c) = 16 g + c: e = rows 4 g + rr of column c of the trailing block,
// mirrored from the lower triangle so that it is symmetric (a_Jc = a_cJ
// serves as the pivot row); x = the same rows of X~ = L~^-1, the inverse of
// the unit lower factor of A = L~ D L~^T; m[rr] = 1 while row 4 g + rr is
// below the pivot; lp = the previous pivot's masked column, rp = its 1/a_JJ.
struct Piv16 {
    double e[4], x[4], m[4], lp[4];
    double rp;
    int ad[4];   // bpermute addresses of lanes (R, c), R = 0..3
    bool bad;
};

// Pivot J. The chain is broadcast -> 1/a_JJ (v_rcp_f64 and two Newton
// steps) -> rank-1 update e -= a_iJ a_Jc / a_JJ on columns c > J (every row
// takes it: rows <= J are finished and never read again, so the update is a
// plain FMA). Column c keeps a_ic of its own pivot step from then on, which
// gives L after the loop (no square root in the loop). Off the chain: row
// J-1 of X~ (final now) is applied to the rows below it.
template <int J>
__device__ __forceinline__ void pivot16(Piv16& st, int g, int c) {
    constexpr int R = J >> 2, RR = J & 3;
    double aic[4];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) aic[rr] = row_bcast<J>(st.e[rr]);
    const double acj = bperm_f64(st.e[RR], st.ad[R]);
    double piv = readlane_f64(st.e[RR], R * 16 + J);
    st.bad |= !(piv > 0.0);
    piv = (piv > 0.0) ? piv : 1.0;
    double r = __builtin_amdgcn_rcp(piv);
    r = fma(r, fma(-piv, r, 1.0), r);
    r = fma(r, fma(-piv, r, 1.0), r);
    const double f = ((c > J) ? acj : 0.0) * r;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) st.e[rr] = fma(-aic[rr], f, st.e[rr]);
    if constexpr (J > 0) {
        constexpr int RP = (J - 1) / 4;
        const double t = st.rp * bperm_f64(st.x[(J - 1) & 3], st.ad[RP]);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) st.x[rr] = fma(-st.lp[rr], t, st.x[rr]);
    }
    st.m[RR] = (g == R) ? 0.0 : st.m[RR];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) st.lp[rr] = aic[rr] * st.m[rr];
    st.rp = r;
}

__global__ __launch_bounds__(512) void k_potrf128(double* __restrict__ F, int Vp, int kb,
                                                  double* __restrict__ Linv128,
                                                  const double* __restrict__ mu, int V,
                                                  int32_t* __restrict__ status) {
    __shared__ double Am[PB][LDA];
    __shared__ double Dv[PB / SB][SB][SB + 1];
    __shared__ int bad;
    // the wave index as a uniform (scalar) value, so that the per-wave task
    // control flow below compiles to scalar branches
    const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lr = lane & 15, lk = lane >> 4;
    const size_t k0 = (size_t)kb * PB;
    if (t == 0) bad = 0;
    {
        // wave w loads rows w, w + 8, ...: lane l the column pair 2l, 2l + 1
        // (one 1 KiB row per wave instruction), 16 loads in flight
        // (the first panel reads A itself: A_hat = A + mu on the V x V part)
        const int m = 2 * lane;
        const double mu0 = (kb == 0) ? mu[0] : 0.0;
        const bool sh0 = kb == 0 && m < V, sh1 = kb == 0 && m + 1 < V;
        double2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            // the lower triangle only (the upper one is masked below)
            v[q] = make_double2(0.0, 0.0);
            if (m <= w + 8 * q)
                v[q] = *reinterpret_cast<const double2*>(F + (k0 + w + 8 * q) * Vp + k0 + m);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = w + 8 * q;
            const bool si = i < V;
            const double x0 = (sh0 && si) ? v[q].x + mu0 : v[q].x;
            const double x1 = (sh1 && si) ? v[q].y + mu0 : v[q].y;
            Am[i][m] = (m <= i) ? x0 : 0.0;
            Am[i][m + 1] = (m + 1 <= i) ? x1 : 0.0;
        }
    }
    __syncthreads();
    // (b) the 16 pivots of diagonal block p on one wave (pivot16), then
    // L_pp and its inverse D_p = L_pp^-1 from the pivots (results stay in
    // registers during the pivot loop: selects, no stores, no branches).
    auto diag16 = [&](int p) __attribute__((always_inline)) {
        const int o = p * SB;
        const int g = lk, c = lr;
        Piv16 st;
        st.bad = false;
        st.rp = 0.0;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int i = 4 * g + rr;
            st.e[rr] = Am[o + (i > c ? i : c)][o + (i > c ? c : i)];
            st.x[rr] = (i == c) ? 1.0 : 0.0;
            st.m[rr] = 1.0;
            st.lp[rr] = 0.0;
            st.ad[rr] = 4 * (16 * rr + c);
        }
        static_for([&](auto jc) { pivot16<decltype(jc)::value>(st, g, c); },
                   std::make_integer_sequence<int, SB>{});
        // the pivots a_kk (e of lane (k >> 2, k), register k & 3): of this
        // lane's column c and of its rows 4 g + rr
        const int c3 = c & 3;
        const double dsel = c3 == 0 ? st.e[0] : c3 == 1 ? st.e[1] : c3 == 2 ? st.e[2] : st.e[3];
        double pc = bperm_f64(dsel, 4 * (16 * (c >> 2) + c));
        pc = (pc > 0.0) ? pc : 1.0;
        double sc, isc;
        pivot_sqrt(pc, sc, isc);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int i = 4 * g + rr;
            double pr = bperm_f64(dsel, 4 * (16 * g + i));
            pr = (pr > 0.0) ? pr : 1.0;
            double sr, isr;
            pivot_sqrt(pr, sr, isr);
            // L = L~ D^1/2: l_ic = a_ic / sqrt(a_cc); L^-1 = D^-1/2 X~
            Am[o + i][o + c] = (i > c) ? st.e[rr] * isc : ((i == c) ? sc : 0.0);
            Dv[p][i][c] = st.x[rr] * isr;
        }
        if (st.bad) bad = 1;
    };
    // (d) trailing tile number q (row by row over p < j <= i) of sub-panel p
    auto dtile = [&](int p, int q) __attribute__((always_inline)) {
        const int o = p * SB;
        int a, b;
        decode_lower(q, a, b);
        const int i = p + 1 + a, j = p + 1 + b;
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Am[16 * i + lr][o + 4 * kk + lk],
                                                       Am[16 * j + lr][o + 4 * kk + lk], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) Am[16 * i + 4 * r + lk][16 * j + lr] -= acc[r];
    };
    // Worker tasks of phase p (waves 1-7, task k for k = w - 1, w + 6, ...),
    // each a 16x16 tile C of Am updated with one 16x16x16 product (4 MFMAs):
    //  A  (k < p, q = k): X_pq = -D_p (S_pq + L_{p,p-1} X_{p-1,q}), block
    //     row p of X = L^-1, final;
    //  B  (next (7 - p) p, row i > p, q < p): S_iq += L_{i,p-1} X_{p-1,q};
    //  D  (the rest): trailing tile 1.. of sub-panel p (dtile).
    // S_iq = sum_{k=q}^{i-1} L_ik X_kq (X_qq = D_q) accumulates in the
    // place of X_iq, transposed in Am's upper block (q, i):
    // Am[16q + b][16i + a] = X_iq[a][b] (zero from the load). The tasks of a
    // phase are independent; two waves per SIMD hide each other's LDS
    // latency.
    auto xtask = [&](int p, int i, int q, bool fin) __attribute__((always_inline)) {
        const int cb = 16 * (p - 1);
        double a[4], b[4], c[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            a[kk] = Am[16 * i + lr][cb + 4 * kk + lk];
            b[kk] = (q == p - 1) ? Dv[p - 1][4 * kk + lk][lr] : Am[16 * q + lr][cb + 4 * kk + lk];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) c[r] = Am[16 * q + lr][16 * i + 4 * r + lk];
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], b[kk], acc, 0, 0, 0);
        if (fin) {
            // the sum's register r is the B operand of k-step r
            d4 xacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                xacc = __builtin_amdgcn_mfma_f64_16x16x4f64(-Dv[p][lr][4 * kk + lk], c[kk] + acc[kk],
                                                            xacc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) Am[16 * q + lr][16 * i + 4 * r + lk] = xacc[r];
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) Am[16 * q + lr][16 * i + 4 * r + lk] = c[r] + acc[r];
        }
    };
    // Stores, issued as soon as their part is final so that they overlap
    // the later phases: block row r of X = L^-1 (D_r on the diagonal, the
    // transposed upper blocks of Am left of it; the zero blocks right of it
    // are never written, the buffer starts zeroed) after phase r, for the
    // panel TRSM, the row scaling of the factor and the triangular solves.
    // The diagonal blocks of L are used only through X, and F's upper
    // triangle keeps A for the residual.
    double* X = Linv128 + k0 * PB;
    auto store_xrow = [&](int r) __attribute__((always_inline)) {
        const int per_row = 8 * (r + 1);   // double2 per row
        for (int q = t; q < 16 * per_row; q += 512) {
            const int a = q / per_row, m = 2 * (q - a * per_row), bm = m >> 4;
            const double2 xv =
                (bm == r) ? make_double2(Dv[r][a][m & 15], Dv[r][a][(m & 15) + 1])
                          : make_double2(Am[16 * bm + (m & 15)][16 * r + a],
                                         Am[16 * bm + (m & 15) + 1][16 * r + a]);
            *reinterpret_cast<double2*>(X + (size_t)(16 * r + a) * PB + m) = xv;
        }
    };
    if (w == 0) diag16(0);
    __syncthreads();
    for (int p = 0; p < PB / SB; ++p) {
        const int o = p * SB;
        if (p > 0) store_xrow(p - 1);
        // (c) rows below the diagonal block: L_ip = A_ip D_p^T
        for (int i = p + 1 + w; i < PB / SB; i += 8) {
            d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Am[16 * i + lr][o + 4 * kk + lk],
                                                           Dv[p][lr][4 * kk + lk], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) Am[16 * i + 4 * r + lk][o + lr] = acc[r];
        }
        __syncthreads();
        // wave 0: the next diagonal tile's update, then its pivots (VALU);
        // waves 1-7 meanwhile (MFMA, two waves per SIMD): the tasks above
        // (all disjoint from the next diagonal block)
        const int nt = PB / SB - 1 - p;
        if (w == 0) {
            if (nt > 0) {
                dtile(p, 0);
                // lanes exchange the updated tile through LDS
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                diag16(p + 1);
            }
        } else {
            const int nA = p, nB = nt * p, ndt = nt > 0 ? nt * (nt + 1) / 2 - 1 : 0;
            for (int k = w - 1; k < nA + nB + ndt; k += 7) {
                if (k < nA) {
                    xtask(p, p, k, true);
                } else if (k < nA + nB) {
                    const int bi = k - nA, i = p + 1 + bi / p;
                    xtask(p, i, bi - (i - p - 1) * p, false);
                } else {
                    dtile(p, k - nA - nB + 1);
                }
            }
        }
        __syncthreads();
    }
    if (t == 0 && bad) atomicOr(status, 1);
    store_xrow(PB / SB - 1);
}

// aload(x, row, col) maps each A operand value (tile-local row and column)
// on its way into LDS.
struct NoHook {
    __device__ void operator()(int) const {}
};
struct AsIs {
    __device__ double operator()(double x, int, int) const { return x; }
};

template <int MT, int KC, int NC = PB, class Hook = NoHook, class ALoad = AsIs>
__device__ __forceinline__ void gemm_nt_f64(const double* __restrict__ A, int lda,
                                            const double* __restrict__ B, int ldb,
                                            double (*As)[KC][LDT], double (*Bs)[KC][LDT],
                                            d4 (&acc)[MT / 32][NC / 32], Hook hook = Hook(),
                                            ALoad aload = ALoad()) {
    constexpr int TA = MT / 32, TB = NC / 32;
    constexpr int NA = MT * KC / 256;   // A doubles per thread per chunk
    constexpr int NBv = NC * KC / 256;  // B doubles per thread per chunk
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wr = (w >> 1) * (MT / 2), wc = (w & 1) * (NC / 2);
    const int arow = t / (KC / NA), acol = (t % (KC / NA)) * NA;
    const int brow = t / (KC / NBv), bcol = (t % (KC / NBv)) * NBv;
    const double* pa = A + (size_t)arow * lda + acol;
    const double* pb = B + (size_t)brow * ldb + bcol;
    double ra[NA], rb[NBv];
    auto gload = [&](int k) {
#pragma unroll
        for (int q = 0; q < NA; q += 2) {
            const double2 v = *reinterpret_cast<const double2*>(pa + k + q);
            ra[q] = v.x;
            ra[q + 1] = v.y;
        }
#pragma unroll
        for (int q = 0; q < NBv; q += 2) {
            const double2 v = *reinterpret_cast<const double2*>(pb + k + q);
            rb[q] = v.x;
            rb[q + 1] = v.y;
        }
    };
    auto sstore = [&](int buf, int k) {
#pragma unroll
        for (int q = 0; q < NA; ++q) As[buf][acol + q][arow] = aload(ra[q], arow, k + acol + q);
#pragma unroll
        for (int q = 0; q < NBv; ++q) Bs[buf][bcol + q][brow] = rb[q];
    };
    gload(0);
    sstore(0, 0);
    __syncthreads();
    const int lr = lane & 15, lk = lane >> 4;
    for (int ch = 0; ch < PB / KC; ++ch) {
        const int buf = ch & 1;
        if (ch + 1 < PB / KC) gload((ch + 1) * KC);
        hook(ch);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            double av[TA], bv[TB];
#pragma unroll
            for (int a = 0; a < TA; ++a) av[a] = As[buf][kk + lk][wr + 16 * a + lr];
#pragma unroll
            for (int b = 0; b < TB; ++b) bv[b] = Bs[buf][kk + lk][wc + 16 * b + lr];
#pragma unroll
            for (int a = 0; a < TA; ++a)
#pragma unroll
                for (int b = 0; b < TB; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        // keep the next chunk's LDS stores (and their wait on the global
        // loads) after this chunk's MFMAs, so the loads overlap the MFMAs
        __builtin_amdgcn_sched_barrier(0);
        if (ch + 1 < PB / KC) sstore(buf ^ 1, (ch + 1) * KC);
        __syncthreads();
    }
}

// Panel rows below the diagonal block: L_ik = A_ik L_kk^-T, 64 rows per
// workgroup, in place. wcnt (nullable): the previous panel's column update
// (k_syrk128 column blocks, other stream) must have counted `target` first.
// A_hat = A + mu on operand values of the V x V part (first panel).
struct ShiftA {
    double mu;
    int rows, cols;   // tile-local bounds of the V x V part
    __device__ double operator()(double x, int r, int c) const {
        return (r < rows && c < cols) ? x + mu : x;
    }
};

__global__ __launch_bounds__(256) void k_panel_trsm(double* __restrict__ F, int Vp, int kb,
                                                    const double* __restrict__ Linv128,
                                                    const int32_t* __restrict__ wcnt,
                                                    int32_t target, const double* __restrict__ mu,
                                                    int V, int32_t* __restrict__ status) {
    __shared__ double As[2][KC_TRSM][LDT];
    __shared__ double Bs[2][KC_TRSM][LDT];
    if (wcnt) wait_flag(wcnt, target, status);
    const size_t k0 = (size_t)kb * PB;
    const size_t r0 = k0 + PB + (size_t)blockIdx.x * 64;
    d4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
    if (kb == 0)   // the first panel reads A itself
        gemm_nt_f64<64, KC_TRSM>(F + r0 * Vp, Vp, Linv128, PB, As, Bs, acc, NoHook{},
                                 ShiftA{mu[0], V - (int)r0, V});
    else
        gemm_nt_f64<64, KC_TRSM>(F + r0 * Vp + k0, Vp, Linv128 + k0 * PB, PB, As, Bs, acc);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 64;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                F[(r0 + wr + 16 * a + 4 * r + (lane >> 4)) * Vp + k0 + wc + 16 * b + (lane & 15)] =
                    acc[a][b][r];
}

// Trailing update with panel kb: A_ij -= L_ik L_jk^T for everything right
// of column kb except the next diagonal tile (kb+1, kb+1), which k_col_diag
// updates on the critical path. One launch per panel on the second stream,
// its workgroups in three groups, dispatched in this order:
//   A  column kb+1 below its diagonal tile (2 (nt - 1) blocks of 64 x 128:
//      two per 128-row tile) and the right-hand-side rows of column kb+1
//      (nx blocks of 64 x 128): the next panel's TRSM waits for all of them
//      (col_done);
//   B  the diagonal tile (kb+2, kb+2) as two 128 x 64 blocks, which
//      k_col_diag of the next panel waits for (diag_done); then the
//      right-hand-side rows of columns kb+2 .. (nx (nt - 1) blocks of
//      64 x 128), needed only by later panels;
//   C  from the next multiple of 8: the 128 x 64 blocks of the lower triangle
//      right of column kb+1 below tile (kb+2, kb+2) (row tile ti >= 1 and
//      column half hj, 0 <= hj <= 2 ti + 1, both counted from tile kb+2).
//      Workgroup b runs on XCD b % 8, and each XCD is given a contiguous
//      range of them in row order (blocks of one row share their A panel in
//      that XCD's L2).
// A and B are small and go first, so the next panel's factorisation and
// TRSM start while C (the bulk) is still running, with nothing waiting in
// between: the next panel's TRSM is only launched after its diagonal
// factorisation, by when group A is long done. 64 x 128 blocks: 4 waves of
// 32 x 64; 128 x 64 blocks: 4 waves of 64 x 32; K = 128 (the panel) on
// v_mfma_f64_16x16x4_f64, the block's 32 C values per lane prefetched into
// registers at the start. The first panel's update reads A itself:
// A_hat = A + mu on the V x V part.
__device__ __forceinline__ void decode_rest(int q, int& ti, int& hj) {
    // q-th 128 x 64 block of the rows ti >= 1 in row order (row ti holds
    // 2 ti + 2 blocks, (ti - 1)(ti + 2) before it)
    int t = (int)((__builtin_sqrt(9.0 + 4.0 * (double)q) - 1.0) * 0.5);
    while (t > 1 && (t - 1) * (t + 2) > q) --t;
    while (t * (t + 3) <= q) ++t;
    ti = t;
    hj = q - (t - 1) * (t + 2);
}

// The same blocks of rows 1 <= ti <= nrow in 2-D order: bands of RB_ROWS
// row tiles, each band by chunks of RB_COLS column halves (then row, then
// column half), so that the blocks of a chunk, which an XCD runs at the same
// time, share RB_ROWS row tiles of the panel (A operands) and RB_COLS column
// halves (B operands) in that XCD's L2; in row order one row's blocks span
// every column half, and the B operands miss L2 row after row.
constexpr int RB_ROWS = 4, RB_COLS = 8;
__device__ __forceinline__ void decode_rest_2d(int q, int nrow, int& ti, int& hj) {
    int t0, h0;
    decode_rest(q, t0, h0);
    const int b0 = 1 + ((t0 - 1) / RB_ROWS) * RB_ROWS;
    const int b1 = min(b0 + RB_ROWS - 1, nrow);
    int r = q - (b0 - 1) * (b0 + 2);   // position in the band
    ti = t0;
    hj = h0;
    for (int c0 = 0; c0 <= 2 * b1 + 1; c0 += RB_COLS) {
        const int c1 = c0 + RB_COLS - 1;
        for (int t = b0; t <= b1; ++t) {
            const int hi = min(2 * t + 1, c1);
            if (hi < c0) continue;
            const int n = hi - c0 + 1;
            if (r < n) {
                ti = t;
                hj = c0 + r;
                return;
            }
            r -= n;
        }
    }
}

// C_rows(ri, rj) (64 x 128) -= P_i P_j^T with the panel columns k0 ..
// (WT: write-through stores, for blocks another stream's kernel waits for)
template <bool WT, int KC = KC_TRSM>
__device__ __forceinline__ void update_64x128(double* __restrict__ F, int Vp, size_t k0,
                                              size_t ri, size_t rj, double mu0, int vi, int vj,
                                              double (*As)[KC][LDT],
                                              double (*Bs)[KC][LDT]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 64;
    double* C = F + (ri + wr + (lane >> 4)) * Vp + rj + wc + (lane & 15);
    d4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
    double cv[2][4][4];
    gemm_nt_f64<64, KC>(F + ri * Vp + k0, Vp, F + rj * Vp + k0, Vp, As, Bs, acc, [&](int ch) {
        if (ch == 0) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        cv[a][b][r] = C[(size_t)(16 * a + 4 * r) * Vp + 16 * b];
        }
    });
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wr + 16 * a + 4 * r + (lane >> 4), j = wc + 16 * b + (lane & 15);
                const double c = (i < vi && j < vj) ? cv[a][b][r] + mu0 : cv[a][b][r];
                if (WT)
                    store_wt(C + (size_t)(16 * a + 4 * r) * Vp + 16 * b, c - acc[a][b][r]);
                else
                    C[(size_t)(16 * a + 4 * r) * Vp + 16 * b] = c - acc[a][b][r];
            }
}

// C(ri, rj) (128 x 64) -= P_i P_j^T; diag: the block lies on a diagonal
// tile (only j <= i is stored), jo its column offset in that tile
template <bool WT>
__device__ __forceinline__ void update_128x64(double* __restrict__ F, int Vp, size_t k0,
                                              size_t ri, size_t rj, bool diag, int jo, double mu0,
                                              int vi, int vj, double (*As)[KC_SYRK][LDT],
                                              double (*Bs)[KC_SYRK][LDT]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 64, wc = (w & 1) * 32;
    double* C = F + (ri + wr + (lane >> 4)) * Vp + rj + wc + (lane & 15);
    d4 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
    double cv[4][2][4];
    gemm_nt_f64<128, KC_SYRK, 64>(F + ri * Vp + k0, Vp, F + rj * Vp + k0, Vp, As, Bs, acc,
                                  [&](int ch) {
        if (ch == 0) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        cv[a][b][r] = C[(size_t)(16 * a + 4 * r) * Vp + 16 * b];
        }
    });
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wr + 16 * a + 4 * r + (lane >> 4);
                const int jl = wc + 16 * b + (lane & 15), j = jo + jl;
                const double c = (i < vi && jl < vj) ? cv[a][b][r] + mu0 : cv[a][b][r];
                if (!diag || j <= i) {
                    if (WT)
                        store_wt(C + (size_t)(16 * a + 4 * r) * Vp + 16 * b, c - acc[a][b][r]);
                    else
                        C[(size_t)(16 * a + 4 * r) * Vp + 16 * b] = c - acc[a][b][r];
                }
            }
}

// Workgroup counts of the trailing update of panel kb (nt = trailing
// 128-column tiles, nx right-hand-side row blocks).
struct TrailShape {
    int na, nb, rest0, nrest, grid;
    __host__ __device__ TrailShape(int nt, int nx) {
        const int nc = nt - 1;
        na = 2 * (nt - 1) + nx;
        nb = nc > 0 ? 2 + nx * nc : 0;
        rest0 = (na + nb + 7) & ~7;
        nrest = nc > 0 ? nc * (nc + 1) - 2 : 0;
        grid = rest0 + 8 * ((nrest + 7) / 8);
    }
};

__global__ __launch_bounds__(256, 2) void k_syrk128(double* __restrict__ F, int Vp, int kb,
                                                 int nt, int nx, const double* __restrict__ mu,
                                                 int V, int32_t* __restrict__ col_done,
                                                 int32_t* __restrict__ diag_done) {
    __shared__ double As[2][KC_SYRK][LDT];
    __shared__ double Bs[2][KC_SYRK][LDT];
    const TrailShape sh(nt, nx);
    const int wg = blockIdx.x;
    const size_t k0 = (size_t)kb * PB;
    const size_t c1 = k0 + PB;       // column kb+1 (the next panel)
    const size_t c2 = k0 + 2 * PB;   // column kb+2
    const double mu0 = (kb == 0) ? mu[0] : 0.0;
    if (wg < sh.na) {
        // A: column kb+1 below its diagonal tile, then its right-hand-side rows
        const int nreg = 2 * (nt - 1);
        const size_t ri = wg < nreg ? c1 + PB + (size_t)(wg >> 1) * PB + 64 * (wg & 1)
                                    : (size_t)Vp + 64 * (size_t)(wg - nreg);
        const bool vv = kb == 0 && wg < nreg;
        update_64x128<true, KC_SYRK>(F, Vp, k0, ri, c1, vv ? mu0 : 0.0, vv ? V - (int)ri : 0,
                            vv ? V - (int)c1 : 0, As, Bs);
        signal_done_wt(col_done);
        return;
    }
    if (wg < sh.na + sh.nb) {
        const int q = wg - sh.na;
        if (q < 2) {
            // B: the diagonal tile (kb+2, kb+2), column half q
            const size_t rj = c2 + 64 * (size_t)q;
            update_128x64<true>(F, Vp, k0, c2, rj, true, 64 * q, mu0, kb == 0 ? V - (int)c2 : 0,
                                kb == 0 ? V - (int)rj : 0, As, Bs);
            signal_done_wt(diag_done);
        } else {
            // B: right-hand-side rows of columns kb+2 ..
            const int r = q - 2, nc = nt - 1;
            update_64x128<false, KC_SYRK>(F, Vp, k0, (size_t)Vp + 64 * (size_t)(r / nc),
                          c2 + (size_t)(r % nc) * PB, 0.0, 0, 0, As, Bs);
        }
        return;
    }
    if (wg < sh.rest0) return;
    const int r = wg - sh.rest0, per = (sh.nrest + 7) >> 3;
    const int q = (r & 7) * per + (r >> 3);
    if (q >= sh.nrest) return;
    int ti, hj;
    decode_rest_2d(q, nt - 2, ti, hj);
    const size_t ri = c2 + (size_t)ti * PB, rj = c2 + (size_t)hj * 64;
    update_128x64<false>(F, Vp, k0, ri, rj, (hj >> 1) == ti, (hj & 1) * 64, mu0,
                  kb == 0 ? V - (int)ri : 0, kb == 0 ? V - (int)rj : 0, As, Bs);
}

// Update of the next panel's 128 x 128 diagonal tile with panel kb, on the
// factorisation's critical path: its three lower 64 x 64 blocks, one
// workgroup each (block (hi, hj) = (0,0), (1,0), (1,1)).
__global__ __launch_bounds__(256, 2) void k_col_diag(double* __restrict__ F, int Vp, int kb,
                                                     const int32_t* __restrict__ wcnt,
                                                     const double* __restrict__ mu, int V,
                                                     int32_t* __restrict__ status) {
    __shared__ double As[2][KC_TRSM][LDT];
    __shared__ double Bs[2][KC_TRSM][LDT];
    // the previous panel's trailing update must have left this tile (its
    // two blocks of it counted)
    if (wcnt) wait_flag(wcnt, 2, status);
    const int hi = blockIdx.x == 0 ? 0 : 1, hj = blockIdx.x == 2 ? 1 : 0;
    const size_t k0 = (size_t)kb * PB;
    const size_t ri = k0 + PB + 64 * hi, rj = k0 + PB + 64 * hj;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    double* C = F + (ri + wr + (lane >> 4)) * Vp + rj + wc + (lane & 15);
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
    double cv[2][2][4];
    gemm_nt_f64<64, KC_TRSM, 64>(F + ri * Vp + k0, Vp, F + rj * Vp + k0, Vp, As, Bs, acc,
                                 [&](int ch) {
        if (ch == 0) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        cv[a][b][r] = C[(size_t)(16 * a + 4 * r) * Vp + 16 * b];
        }
    });
    // first panel: A_hat = A + mu on the V x V part
    const double mu0 = (kb == 0) ? mu[0] : 0.0;
    const int vi = (kb == 0) ? V - (int)ri : 0, vj = (kb == 0) ? V - (int)rj : 0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wr + 16 * a + 4 * r + (lane >> 4);
                const int j = wc + 16 * b + (lane & 15);
                const double c = (i < vi && j < vj) ? cv[a][b][r] + mu0 : cv[a][b][r];
                if (hi != hj || j <= i)
                    C[(size_t)(16 * a + 4 * r) * Vp + 16 * b] = c - acc[a][b][r];
            }
}

// FP32 trailing update with fp16x2-split operands on MFMA (FP32 accumulate).
// v_mfma_f32_16x16x32_f16 layout: A[lane%16][8*(lane/16)+i], B[8*(lane/16)+i][lane%16],
// D[4*(lane/16)+r][lane%16].
__global__ __launch_bounds__(256) void k_syrk_f16x3(float* __restrict__ F, int Vp, int kb) {
    __shared__ _Float16 Hi[2][NB][NB + 8];
    __shared__ _Float16 Lo[2][NB][NB + 8];
    int ti, tj;
    decode_lower(blockIdx.x, ti, tj);
    const int k0 = kb * NB;
    const int ri = (kb + 1 + ti) * NB, rj = (kb + 1 + tj) * NB;
    const int t = threadIdx.x;
    for (int q = t; q < NB * NB; q += 256) {
        const int i = q / NB, m = q % NB;
        const float xi = F[(size_t)(ri + i) * Vp + k0 + m];
        const float xj = F[(size_t)(rj + i) * Vp + k0 + m];
        const _Float16 hi = (_Float16)xi, hj = (_Float16)xj;
        Hi[0][i][m] = hi;
        Lo[0][i][m] = (_Float16)(xi - (float)hi);
        Hi[1][i][m] = hj;
        Lo[1][i][m] = (_Float16)(xj - (float)hj);
    }
    __syncthreads();
    const int lane = t & 63, wave = t >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    f4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    for (int k = 0; k < NB; k += 32) {
        h8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                ah[a][i] = Hi[0][wr + 16 * a + lr][k + lk + i];
                al[a][i] = Lo[0][wr + 16 * a + lr][k + lk + i];
                bh[a][i] = Hi[1][wc + 16 * a + lr][k + lk + i];
                bl[a][i] = Lo[1][wc + 16 * a + lr][k + lk + i];
            }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
            }
    }
    const bool diag = (ti == tj);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wr + 16 * a + (lane >> 4) * 4 + r;
                const int col = wc + 16 * b + lr;
                if (!diag || col <= row) F[(size_t)(ri + row) * Vp + rj + col] -= acc[a][b][r];
            }
}

// Mixed form: 64-column panels on the FP32 factor.
rc_status factor_loop_mixed(rc_ctx* ctx, float* F, float* Linv) {
    const int Vp = ctx->nblk * NB;
    for (int kb = 0; kb < ctx->nblk; ++kb) {
        const int nrow = ctx->nblk - kb;
        // a breakdown of the FP32 factor only weakens the preconditioner (bit 8)
        hipLaunchKernelGGL(k_potrf_panel<float>, dim3(nrow), dim3(256), 0, ctx->stream, F, Vp, kb,
                           Linv, ctx->d_status, 8);
        RC_CHECK_LAUNCH();
        const int ntr = nrow - 1;
        if (ntr > 0) {
            hipLaunchKernelGGL(k_syrk_f16x3, dim3(ntr * (ntr + 1) / 2), dim3(256), 0,
                               ctx->stream, F, Vp, kb);
            RC_CHECK_LAUNCH();
        }
    }
    return RC_OK;
}

// FP64 form: 128-column panels (diagonal factor + inverse, panel TRSM as a
// GEMM with the inverse, MFMA trailing update) with a one-panel lookahead on
// two streams. Main stream, per panel kb: potrf(kb) -> panel TRSM(kb) ->
// update of the next diagonal tile (k_col_diag), the critical path. Second
// stream: the trailing update with panel kb (k_syrk128), whose first
// workgroups update column kb+1 and the diagonal tile after next, so the
// next panel's diagonal factorisation and TRSM overlap its bulk. Ordering:
// the event T (panel kb final) before the second stream's update of kb;
// device counters (d_fcnt, cleared per factorisation) for the main stream's
// waits, which a stream event would turn into a command-processor barrier of
// ~5-10 us on the critical path: panel TRSM(kb+1) polls for the column
// blocks of k_syrk128(kb), k_col_diag(kb+1) for its two diagonal-tile blocks.
// Both consumers are small, their producers were launched earlier on the
// other stream and are dispatched first, so a poll always ends; a bound on
// it sets status bit 2. (Waiting for the whole previous update with a stream
// event while it is large was measured slower on the same box: 87.6-88.6
// against 90.9-91.2 steps/s.)
// The main stream waits for the last rest-update before returning.
// A non-positive pivot sets status bit 1 (singular saddle system).
// Right-hand sides: nx blocks of 64 rows stored below the factor (rows Vp ..,
// B^T) are carried along as extra matrix rows: the panel TRSM of every panel
// (the last one included) and the trailing updates of k_syrk128 cover them,
// which leaves Y^T = (L^-1 B)^T in their place, the forward substitution of
// L L^T X = B done inside the factorisation.
rc_status factor_loop_fp64(rc_ctx* ctx, int nx) {
    const int Vp = ctx->nblk * NB;
    const int np = Vp / PB;
    if ((int)ctx->dep_ev.size() < np + 2) {
        while ((int)ctx->dep_ev.size() < np + 2) {
            // completion signals only: each kernel's own end-of-kernel release
            // makes its writes visible device-wide, so no extra fence here
            hipEvent_t e;
            RC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
            ctx->dep_ev.push_back(e);
        }
    }
    hipEvent_t* T = ctx->dep_ev.data();   // panel kb final
    hipEvent_t start = ctx->dep_ev[np];
    hipEvent_t fin = ctx->dep_ev[np + 1];
    int32_t* colcnt = ctx->d_fcnt;        // [kb]: k_syrk128(kb) column-kb+1 blocks done
    int32_t* diagcnt = ctx->d_fcnt + np;  // [kb]: k_syrk128(kb) blocks of tile kb+2 done
    const double* mu = ctx->d_scal + 1;
    RC_HIP(hipMemsetAsync(ctx->d_fcnt, 0, sizeof(int32_t) * 2 * (size_t)np, ctx->stream));
    // the second stream starts after the factor's input is prepared
    RC_HIP(hipEventRecord(start, ctx->stream));
    RC_HIP(hipStreamWaitEvent(ctx->stream2, start, 0));
    auto prof_begin = [&](hipStream_t st, double flop) -> hipEvent_t* {
        if (!ctx->kprof) return nullptr;
        if (ctx->kev_used + 2 > ctx->kev.size()) {
            // no system-scope fence: a timing marker must not write back the
            // L2 (the factor leaves it full of dirty lines)
            hipEvent_t a, b;
            if (hipEventCreateWithFlags(&a, hipEventDisableSystemFence) != hipSuccess ||
                hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess)
                return nullptr;
            ctx->kev.push_back(a);
            ctx->kev.push_back(b);
        }
        hipEvent_t* ev = &ctx->kev[ctx->kev_used];
        ctx->kev_used += 2;
        ctx->kev_flop.push_back(flop);
        (void)hipEventRecord(ev[0], st);
        return ev;
    };
    for (int kb = 0; kb < np; ++kb) {
        hipLaunchKernelGGL(k_potrf128, dim3(1), dim3(512), 0, ctx->stream, ctx->d_F, Vp, kb,
                           ctx->d_Linv128, mu, ctx->V, ctx->d_status);
        RC_CHECK_LAUNCH();
        const int nt = np - kb - 1;
        if (nt == 0 && nx == 0) break;
        // panel kb below its diagonal block (and in the right-hand-side
        // rows): needs column kb below the diagonal tile, updated by the
        // first workgroups of k_syrk128(kb-1) on the second stream
        const int prev_cols = (kb > 0) ? TrailShape(np - kb, nx).na : 0;
        hipLaunchKernelGGL(k_panel_trsm, dim3(2 * nt + nx), dim3(256), 0, ctx->stream, ctx->d_F, Vp,
                           kb, (const double*)ctx->d_Linv128,
                           (const int32_t*)(prev_cols > 0 ? colcnt + kb - 1 : nullptr), prev_cols,
                           mu, ctx->V, ctx->d_status);
        RC_CHECK_LAUNCH();
        RC_HIP(hipEventRecord(T[kb], ctx->stream));
        if (nt == 0) break;
        // next diagonal tile: after the rest-update of panel kb-1 has left it
        hipLaunchKernelGGL(k_col_diag, dim3(3), dim3(256), 0, ctx->stream, ctx->d_F, Vp, kb,
                           (const int32_t*)(kb > 0 ? diagcnt + kb - 1 : nullptr), mu, ctx->V,
                           ctx->d_status);
        RC_CHECK_LAUNCH();
        // second stream: the trailing update with panel kb (column kb+1
        // and the next diagonal tile first)
        RC_HIP(hipStreamWaitEvent(ctx->stream2, T[kb], 0));
        const TrailShape ts(nt, nx);
        if (ts.grid > 0) {
            const double mr = (double)(nt - 1) * PB;
            const double flop = (double)PB * mr * (mr + 1.0) +
                                2.0 * PB * PB * 64.0 * (2.0 * (nt - 1) + (double)nx * nt);
            hipEvent_t* ev = prof_begin(ctx->stream2, flop);
            hipLaunchKernelGGL(k_syrk128, dim3(ts.grid), dim3(256), 0, ctx->stream2, ctx->d_F, Vp,
                               kb, nt, nx, mu, ctx->V, colcnt + kb, diagcnt + kb);
            RC_CHECK_LAUNCH();
            if (ev) RC_HIP(hipEventRecord(ev[1], ctx->stream2));
        }
    }
    // the main stream continues after all second-stream work
    RC_HIP(hipEventRecord(fin, ctx->stream2));
    RC_HIP(hipStreamWaitEvent(ctx->stream, fin, 0));
    // the sweeps read the factor row-scaled by its diagonal blocks' inverses
    // (scaling a row block while the factorisation still runs, on a third
    // stream, delays the panel chain's launches by more than it saves)
    return rc_ltilde_fp64(ctx);
}

}  // namespace

// Factor A_hat = A + mu 11^T in place in ctx->d_F (FP64), which holds A (the
// Gram assembly's output, leading dimension Vp). mu = trace(A)/V^2 is kept
// in d_scal[1], the diagonal of A in d_Adiag. The factor replaces the lower
// triangle; the strict upper triangle keeps A. A non-positive pivot sets
// status bit 1 (singular). nx 64-row blocks of right-hand sides B^T below the
// factor (rows Vp .. Vp + 64 nx - 1) are overwritten with (L^-1 B)^T.
rc_status rc_factor_fp64(rc_ctx* ctx, int nx) {
    const int Vp = ctx->nblk * NB;
    RC_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(int32_t) * 4, ctx->stream));
    if (Vp > ctx->V) {
        hipLaunchKernelGGL(k_pad_identity, dim3((Vp + 255) / 256, Vp - ctx->V), dim3(256), 0,
                           ctx->stream, ctx->d_F, ctx->V, Vp);
        RC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_trace_mu, dim3(1), dim3(1024), 0, ctx->stream, (const double*)ctx->d_F,
                       ctx->V, Vp, ctx->d_scal + 1, ctx->d_Adiag);
    RC_CHECK_LAUNCH();
    // no separate pass forms A_hat: the first panel's kernels add mu where
    // they read A
    // The panel loop is launched eagerly: the host enqueues it ~3 ms ahead of
    // the GPU. (Replaying it as a captured HIP graph measured 0.25 ms slower
    // per step: hipGraphLaunch holds the host for the whole factorisation, so
    // the launches after it start late.)
    return factor_loop_fp64(ctx, nx);
}

// Factor S + delta I, S = D^-1/2 A_hat D^-1/2 (unit diagonal), into
// ctx->d_F32 (FP32 storage, FP16x3 MFMA trailing updates); D^-1/2 in
// ctx->d_dscale. kappa(A_hat) grows like V^2.4 and exceeds what an FP32
// Cholesky survives (kappa u_f32 < ~1) from a few thousand vertices on, so
// the factored matrix is shifted: delta doubles after each breakdown
// (status bit 8, read back once per try), up to 8 tries, from the last shift
// that factored. The shift only weakens the preconditioner (PCG iterations
// grow with it: 44 at 2^-23 against 109 at 2^-20 for V = 8192 in a CPU
// model of this factor); the PCG in FP64 with the unshifted A_hat fixes the
// solution.
rc_status rc_factor_mixed(rc_ctx* ctx) {
    const int Vp = ctx->nblk * NB;
    if (!ctx->d_F32) {
        RC_HIP(hipMalloc((void**)&ctx->d_F32, sizeof(float) * (size_t)Vp * Vp));
        RC_HIP(hipMalloc((void**)&ctx->d_Linv32, sizeof(float) * (size_t)ctx->nblk * NB * NB));
        RC_HIP(hipMalloc((void**)&ctx->d_dscale, sizeof(double) * (size_t)Vp));
    }
    hipLaunchKernelGGL(k_trace_mu, dim3(1), dim3(1024), 0, ctx->stream, (const double*)ctx->d_F,
                       ctx->V, Vp, ctx->d_scal + 1, ctx->d_Adiag);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_diag_scale, dim3((Vp + 255) / 256), dim3(256), 0, ctx->stream,
                       (const double*)ctx->d_F, ctx->V, Vp, Vp, ctx->d_scal + 1, ctx->d_dscale);
    RC_CHECK_LAUNCH();
    // the smallest shift that factors (nearest the unshifted S) makes the
    // best preconditioner: start from the last working one (2^-23 at first;
    // the Gram matrix changes little from step to step) and double it after
    // each breakdown
    double shift = ctx->mixed_shift_start > 0.0 ? ctx->mixed_shift_start : 0x1p-23;
    for (int attempt = 0; attempt < 8; ++attempt, shift *= 2.0) {
        RC_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(int32_t) * 4, ctx->stream));
        hipLaunchKernelGGL(k_factor_prep<float>, dim3((Vp + 255) / 256, Vp), dim3(256), 0,
                           ctx->stream, (const double*)ctx->d_F, ctx->V, Vp, Vp, ctx->d_scal + 1,
                           (const double*)ctx->d_dscale, shift, ctx->d_F32);
        RC_CHECK_LAUNCH();
        rc_status s = factor_loop_mixed(ctx, ctx->d_F32, ctx->d_Linv32);
        if (s != RC_OK) return s;
        int32_t st = 0;
        RC_HIP(hipMemcpyAsync(&st, ctx->d_status, sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        RC_HIP(hipStreamSynchronize(ctx->stream));
        ctx->mixed_shift = shift;
        ctx->mixed_shift_start = shift;
        if (!(st & 8)) break;
    }
    return rc_ltilde_f32(ctx);
}

// Dense SPD factorisation of A_hat = A + mu 1 1^T and multi-RHS solves.
//
// The reference factors the (3V+k)^2 saddle matrix [[I3 (x) A, C^T],[C, 0]]
// with Bunch-Kaufman LDL^T (src/core/sobolev/linsolve.ts:431-589, 742-803).
// A is PSD with null space span(1) (its rows sum to zero), so A_hat is SPD and
// the saddle solution follows exactly from A_hat^{-1} applied to 3(k+1)
// right-hand sides plus a (k+3)^2 system (see rc_step.hip). Only the V x V
// Cholesky factor is formed: V^3/3 flops instead of (3V+k)^3/3.
//
// Storage: F is Vp x Vp row-major (Vp = nblk * 64), lower triangle, element
// type T: double for the plain form, float for the mixed form (which factors
// the diagonally scaled S = D^-1/2 A_hat D^-1/2, unit diagonal). Padding
// rows/columns are the identity. Blocked right-looking Cholesky, 64 columns
// per panel:
//   k_potrf_panel<T>  every block factors the 64x64 diagonal block in FP64
//                     in LDS and inverts it; block 0 stores L_kk and
//                     L_kk^-1, the others form their rows L_rk = A_rk L_kk^-T
//   k_syrk_f64        trailing update A_ij -= L_ik L_jk^T, 64x64 lower tiles,
//                     v_mfma_f64_16x16x4_f64
//   k_syrk_f16x3      same update for the FP32 factor: each operand is split
//                     into fp16 hi + lo (x = hi + lo to ~2^-22) and
//                     hi*hi^T + hi*lo^T + lo*hi^T runs on
//                     v_mfma_f32_16x16x32_f16 with FP32 accumulation
// Triangular solves: one block per 64-row block, dependency flags in global
// memory (sync-free substitution, FP64 arithmetic on the stored factor); the
// diagonal solve is a GEMM with the stored L_kk^-1.
#include "rc_internal.h"
#include "rc_math.h"

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int NB = RC_NB;
constexpr int RHS_CHUNK = 32;

__global__ void k_trace_mu(const double* __restrict__ A, int V, double* __restrict__ mu) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < V; i += 1024) s += A[(size_t)i * V + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) mu[0] = sh[0] / ((double)V * (double)V);
}

// dinv[i] = 1/sqrt(A_ii + mu) (1 on padding)
__global__ void k_diag_scale(const double* __restrict__ A, int V, int Vp,
                             const double* __restrict__ mu, double* __restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Vp) return;
    dinv[i] = (i < V) ? 1.0 / __builtin_sqrt(A[(size_t)i * V + i] + mu[0]) : 1.0;
}

// F (lower) = A + mu (FP64 form) or dinv_i (A + mu) dinv_j (FP32 form); identity padding.
template <typename T>
__global__ void k_factor_prep(const double* __restrict__ A, int V, int Vp,
                              const double* __restrict__ mu, const double* __restrict__ dinv,
                              T* __restrict__ F) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= Vp) return;
    double v;
    if (i < V && j < V) {
        v = (j <= i) ? A[(size_t)i * V + j] + mu[0] : 0.0;
        if (dinv) v = v * dinv[i] * dinv[j];
    } else {
        v = (i == j) ? 1.0 : 0.0;
    }
    F[(size_t)i * Vp + j] = (T)v;
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

// FP64 trailing update on 64x64 lower tiles: C_ij -= P_i P_j^T.
// 4 waves, each a 32x32 quadrant = 2x2 MFMA 16x16 tiles, K = 64 in steps of 4.
// v_mfma_f64_16x16x4_f64 layout: A[lane%16][lane/16], B[lane/16][lane%16],
// D[4*(lane/16)+r][lane%16], r = 0..3.
__global__ __launch_bounds__(256) void k_syrk_f64(double* __restrict__ F, int Vp, int kb) {
    __shared__ double Pi[NB][NB + 1];
    __shared__ double Pj[NB][NB + 1];
    int ti, tj;
    decode_lower(blockIdx.x, ti, tj);
    const int k0 = kb * NB;
    const int ri = (kb + 1 + ti) * NB, rj = (kb + 1 + tj) * NB;
    const int t = threadIdx.x;
    for (int q = t; q < NB * NB; q += 256) {
        const int i = q / NB, m = q % NB;
        Pi[i][m] = F[(size_t)(ri + i) * Vp + k0 + m];
        Pj[i][m] = F[(size_t)(rj + i) * Vp + k0 + m];
    }
    __syncthreads();
    const int lane = t & 63, wave = t >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
    const int lr = lane & 15, lk = lane >> 4;
    for (int k = 0; k < NB; k += 4) {
        double av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = Pi[wr + 16 * a + lr][k + lk];
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = Pj[wc + 16 * b + lr][k + lk];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    const bool diag = (ti == tj);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wr + 16 * a + 4 * lk + r;
                const int col = wc + 16 * b + lr;
                if (!diag || col <= row) F[(size_t)(ri + row) * Vp + rj + col] -= acc[a][b][r];
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

__device__ __forceinline__ void wait_flag(const int32_t* flag, int32_t epoch, int32_t* status) {
    if (threadIdx.x == 0) {
        int spins = 0;
        while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1 << 24)) {
                atomicOr(status, 2);
                break;
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ double coherent_load(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void publish(int32_t* flag, int32_t epoch) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Forward substitution L Y = B on nrhs <= RHS_CHUNK columns (column-major,
// leading dimension ldb), in place. Block r handles rows [64r, 64r+64).
// skip (nullable): when *skip != 0 every block returns at once.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_fwd(const T* __restrict__ F, int Vp,
                                                  const T* __restrict__ Linv,
                                                  double* __restrict__ B, int ldb, int nrhs,
                                                  int32_t* __restrict__ flags, int32_t epoch,
                                                  int32_t* __restrict__ status,
                                                  const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double acc[NB][RHS_CHUNK + 1];
    __shared__ double Lt[NB][NB + 1];
    __shared__ double Yc[NB][RHS_CHUNK + 1];
    const int r = blockIdx.x;
    const int t = threadIdx.x;
    for (int q = t; q < NB * nrhs; q += 256) {
        const int i = q % NB, c = q / NB;
        acc[i][c] = B[(size_t)c * ldb + r * NB + i];
    }
    const int ti = t >> 2;
    const int tc = t & 3;
    for (int cb = 0; cb < r; ++cb) {
        wait_flag(flags + cb, epoch, status);
        for (int q = t; q < NB * NB; q += 256) {
            const int i = q / NB, m = q % NB;
            Lt[i][m] = (double)F[(size_t)(r * NB + i) * Vp + cb * NB + m];
        }
        for (int q = t; q < NB * nrhs; q += 256) {
            const int m = q % NB, c = q / NB;
            Yc[m][c] = coherent_load(B + (size_t)c * ldb + cb * NB + m);
        }
        __syncthreads();
        for (int c = tc; c < nrhs; c += 4) {
            double s = 0.0;
            for (int m = 0; m < NB; ++m) s = fma(Lt[ti][m], Yc[m][c], s);
            acc[ti][c] -= s;
        }
        __syncthreads();
    }
    const T* Li = Linv + (size_t)r * NB * NB;
    for (int q = t; q < NB * NB; q += 256) Lt[q / NB][q % NB] = (double)Li[q];
    __syncthreads();
    for (int c = tc; c < nrhs; c += 4) {
        double s = 0.0;
        for (int m = 0; m <= ti; ++m) s = fma(Lt[ti][m], acc[m][c], s);
        Yc[ti][c] = s;
    }
    __syncthreads();
    for (int q = t; q < NB * nrhs; q += 256) {
        const int i = q % NB, c = q / NB;
        B[(size_t)c * ldb + r * NB + i] = Yc[i][c];
    }
    publish(flags + r, epoch);
}

// Backward substitution L^T X = Y in place. Block b handles block row
// r = nblk - 1 - b so that every block only waits on lower-indexed blocks.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_bwd(const T* __restrict__ F, int Vp,
                                                  const T* __restrict__ Linv,
                                                  double* __restrict__ B, int ldb, int nrhs,
                                                  int nblk, int32_t* __restrict__ flags,
                                                  int32_t epoch, int32_t* __restrict__ status,
                                                  const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double acc[NB][RHS_CHUNK + 1];
    __shared__ double Lt[NB][NB + 1];
    __shared__ double Xc[NB][RHS_CHUNK + 1];
    const int r = nblk - 1 - (int)blockIdx.x;
    const int t = threadIdx.x;
    for (int q = t; q < NB * nrhs; q += 256) {
        const int i = q % NB, c = q / NB;
        acc[i][c] = B[(size_t)c * ldb + r * NB + i];
    }
    const int ti = t >> 2;
    const int tc = t & 3;
    for (int cb = nblk - 1; cb > r; --cb) {
        wait_flag(flags + cb, epoch, status);
        for (int q = t; q < NB * NB; q += 256) {
            const int m = q / NB, i = q % NB;
            Lt[i][m] = (double)F[(size_t)(cb * NB + m) * Vp + r * NB + i];
        }
        for (int q = t; q < NB * nrhs; q += 256) {
            const int m = q % NB, c = q / NB;
            Xc[m][c] = coherent_load(B + (size_t)c * ldb + cb * NB + m);
        }
        __syncthreads();
        for (int c = tc; c < nrhs; c += 4) {
            double s = 0.0;
            for (int m = 0; m < NB; ++m) s = fma(Lt[ti][m], Xc[m][c], s);
            acc[ti][c] -= s;
        }
        __syncthreads();
    }
    const T* Li = Linv + (size_t)r * NB * NB;
    for (int q = t; q < NB * NB; q += 256) Lt[q % NB][q / NB] = (double)Li[q];  // Lt = Linv^T
    __syncthreads();
    for (int c = tc; c < nrhs; c += 4) {
        double s = 0.0;
        for (int m = ti; m < NB; ++m) s = fma(Lt[ti][m], acc[m][c], s);
        Xc[ti][c] = s;
    }
    __syncthreads();
    for (int q = t; q < NB * nrhs; q += 256) {
        const int i = q % NB, c = q / NB;
        B[(size_t)c * ldb + r * NB + i] = Xc[i][c];
    }
    publish(flags + r, epoch);
}

template <typename T>
rc_status factor_loop(rc_ctx* ctx, T* F, T* Linv) {
    const int Vp = ctx->nblk * NB;
    for (int kb = 0; kb < ctx->nblk; ++kb) {
        const int nrow = ctx->nblk - kb;
        // a breakdown of the FP32 factor only weakens the preconditioner (bit 8);
        // in the FP64 form it means the saddle system is singular (bit 1)
        hipLaunchKernelGGL(k_potrf_panel<T>, dim3(nrow), dim3(256), 0, ctx->stream, F, Vp, kb,
                           Linv, ctx->d_status, sizeof(T) == 8 ? 1 : 8);
        RC_CHECK_LAUNCH();
        const int ntr = nrow - 1;
        if (ntr > 0) {
            if constexpr (sizeof(T) == 8)
                hipLaunchKernelGGL(k_syrk_f64, dim3(ntr * (ntr + 1) / 2), dim3(256), 0,
                                   ctx->stream, F, Vp, kb);
            else
                hipLaunchKernelGGL(k_syrk_f16x3, dim3(ntr * (ntr + 1) / 2), dim3(256), 0,
                                   ctx->stream, F, Vp, kb);
            RC_CHECK_LAUNCH();
        }
    }
    return RC_OK;
}

}  // namespace

// Factor A_hat = A + mu 11^T into ctx->d_F (FP64). mu = trace(A)/V^2 is kept
// in d_scal[1]. A non-positive pivot sets status bit 1 (singular).
rc_status rc_factor_fp64(rc_ctx* ctx) {
    const int Vp = ctx->nblk * NB;
    RC_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(int32_t) * 4, ctx->stream));
    hipLaunchKernelGGL(k_trace_mu, dim3(1), dim3(1024), 0, ctx->stream, ctx->d_A, ctx->V,
                       ctx->d_scal + 1);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_factor_prep<double>, dim3((Vp + 255) / 256, Vp), dim3(256), 0, ctx->stream,
                       ctx->d_A, ctx->V, Vp, ctx->d_scal + 1, (const double*)nullptr, ctx->d_F);
    RC_CHECK_LAUNCH();
    return factor_loop<double>(ctx, ctx->d_F, ctx->d_Linv);
}

// Factor S = D^-1/2 A_hat D^-1/2 into ctx->d_F32 (FP32 storage, FP16x3 MFMA
// trailing updates); D^-1/2 in ctx->d_dscale.
rc_status rc_factor_mixed(rc_ctx* ctx) {
    const int Vp = ctx->nblk * NB;
    if (!ctx->d_F32) {
        RC_HIP(hipMalloc((void**)&ctx->d_F32, sizeof(float) * (size_t)Vp * Vp));
        RC_HIP(hipMalloc((void**)&ctx->d_Linv32, sizeof(float) * (size_t)ctx->nblk * NB * NB));
        RC_HIP(hipMalloc((void**)&ctx->d_dscale, sizeof(double) * (size_t)Vp));
    }
    RC_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(int32_t) * 4, ctx->stream));
    hipLaunchKernelGGL(k_trace_mu, dim3(1), dim3(1024), 0, ctx->stream, ctx->d_A, ctx->V,
                       ctx->d_scal + 1);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_diag_scale, dim3((Vp + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_A,
                       ctx->V, Vp, ctx->d_scal + 1, ctx->d_dscale);
    RC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_factor_prep<float>, dim3((Vp + 255) / 256, Vp), dim3(256), 0, ctx->stream,
                       ctx->d_A, ctx->V, Vp, ctx->d_scal + 1, (const double*)ctx->d_dscale,
                       ctx->d_F32);
    RC_CHECK_LAUNCH();
    return factor_loop<float>(ctx, ctx->d_F32, ctx->d_Linv32);
}

template <typename T>
static rc_status solve_impl(rc_ctx* ctx, const T* F, const T* Linv, double* d_B, int nrhs,
                            const int32_t* skip) {
    const int Vp = ctx->nblk * NB;
    for (int c0 = 0; c0 < nrhs; c0 += RHS_CHUNK) {
        const int nc = (nrhs - c0) < RHS_CHUNK ? (nrhs - c0) : RHS_CHUNK;
        double* Bc = d_B + (size_t)c0 * Vp;
        const int32_t ef = ++ctx->epoch;
        hipLaunchKernelGGL(k_trsm_fwd<T>, dim3(ctx->nblk), dim3(256), 0, ctx->stream, F, Vp, Linv,
                           Bc, Vp, nc, ctx->d_flags, ef, ctx->d_status, skip);
        RC_CHECK_LAUNCH();
        const int32_t eb = ++ctx->epoch;
        hipLaunchKernelGGL(k_trsm_bwd<T>, dim3(ctx->nblk), dim3(256), 0, ctx->stream, F, Vp, Linv,
                           Bc, Vp, nc, ctx->nblk, ctx->d_flags, eb, ctx->d_status, skip);
        RC_CHECK_LAUNCH();
    }
    return RC_OK;
}

// Solve A_hat X = B for nrhs columns of length Vp (column-major, ld = Vp) in
// place, with the FP64 factor.
rc_status rc_solve_fp64(rc_ctx* ctx, double* d_B, int nrhs) {
    return solve_impl<double>(ctx, ctx->d_F, ctx->d_Linv, d_B, nrhs, nullptr);
}

// Solve S Y = B in place with the FP32 factor (caller applies the scaling).
rc_status rc_solve_f32_factor(rc_ctx* ctx, double* d_B, int nrhs, const int32_t* skip) {
    return solve_impl<float>(ctx, ctx->d_F32, ctx->d_Linv32, d_B, nrhs, skip);
}

int rc_trsm_max_coresident(void) {
    int dev = 0, ncu = 0, per = 0, per2 = 0, per3 = 0, per4 = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_trsm_fwd<double>, 256, 0) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, k_trsm_bwd<double>, 256, 0) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per3, k_trsm_fwd<float>, 256, 0) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per4, k_trsm_bwd<float>, 256, 0) != hipSuccess)
        return 0;
    int m = per;
    if (per2 < m) m = per2;
    if (per3 < m) m = per3;
    if (per4 < m) m = per4;
    return m * ncu;
}

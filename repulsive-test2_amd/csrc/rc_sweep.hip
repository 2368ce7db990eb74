// Triangular solves with the factor row-scaled by its diagonal blocks.
//
// Both factorisations end by replacing the blocks below the 128 x 128 block
// diagonal, L_RC (R > C), with L~_RC = X_R L_RC, X_R = L_RR^-1 (k_ltilde).
// Then L = D L~ with D = blockdiag(L_RR) and L~ unit block lower triangular,
// and with X = blockdiag(X_R)
//   L y = b    <=>  L~ y = X b                (k_bdiag, then k_usweep fwd)
//   L^T x = y  <=>  L~^T z = y,  x = X^T z     (k_usweep bwd, then k_bdiag)
// A block of a sweep over L~ needs no solve with its diagonal block: its rows
// depend only on earlier blocks, so G = RC_SWEEP_G workgroups share a block,
// 128 / G rows each. The chain of the sweep has Vp / 128 links, and the work
// after each wait is a (128 / G) x 128 x nrhs product per workgroup (against
// Vp / 64 links of a 64 x 64 x nrhs product and a diagonal-block solve on one
// workgroup when the blocks of L itself are swept).
//
// A sweep workgroup takes a ticket when it starts (take_ticket); ticket ->
// (chain position, row slice, right-hand-side chunk). It waits only for
// blocks earlier in the chain, whose workgroups hold smaller tickets and have
// started, so the sweep needs no co-residency of its grid. Results pass
// between workgroups by write-through stores and one flag per workgroup
// (publish_wt / wait_flags_wt): the flag of (block, slice) in chunk row y
// holds the epoch of the last sweep that finished that slice.
#include "rc_internal.h"
#include "rc_handoff.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int PB = 128;         // rows of a block of the row-scaled factor
constexpr int RHS_CHUNK = 32;   // right-hand-side columns per sweep workgroup

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 4 consecutive factor entries as doubles
__device__ __forceinline__ void load4(const double* p, double (&v)[4]) {
    const double2 x = *reinterpret_cast<const double2*>(p);
    const double2 y = *reinterpret_cast<const double2*>(p + 2);
    v[0] = x.x;
    v[1] = x.y;
    v[2] = y.x;
    v[3] = y.y;
}
__device__ __forceinline__ void load4(const float* p, double (&v)[4]) {
    const float4 x = *reinterpret_cast<const float4*>(p);
    v[0] = x.x;
    v[1] = x.y;
    v[2] = x.z;
    v[3] = x.w;
}

// L~_RC = X_R L_RC in place for the blocks (R, C), C < R, of the row blocks
// R >= r_lo (2 R workgroups per row block, R (R - 1) - r_lo (r_lo - 1) before
// it). Workgroup (R, C, hf) forms columns 64 hf .. 64 hf + 63 of the block:
// they are staged in LDS (as FP64) before any result is stored. Rows of X_R
// are the A operands (each lane four consecutive k, the same k of the staged
// columns as B operand); wave w forms row tile w, for which the lower
// triangular X_R has k < 16 (w + 1).
template <typename T>
__global__ __launch_bounds__(512, 4) void k_ltilde(T* __restrict__ F, int Vp,
                                                const double* __restrict__ X, int r_lo) {
    __shared__ double Ls[PB][64 + 4];
    const long q = (long)blockIdx.x + (long)r_lo * (r_lo - 1);
    int R = (int)((1.0 + __builtin_sqrt(1.0 + 4.0 * (double)q)) * 0.5);
    while (R > 1 && (long)R * (R - 1) > q) --R;
    while ((long)(R + 1) * R <= q) ++R;
    const int rem = (int)(q - (long)R * (R - 1));
    const int C = rem >> 1, hf = rem & 1;
    const size_t r0 = (size_t)R * PB, c0 = (size_t)C * PB + 64 * hf;
    const int t = threadIdx.x, lane = t & 63, lr = lane & 15, lk = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    // every load of the workgroup in flight at once: the block's 16 values
    // per thread, then this wave's rows of X_R
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int e = t + 512 * q, i = e >> 6, j = e & 63;
        v[q] = (double)F[(r0 + i) * Vp + c0 + j];
    }
    const double* xr = X + (size_t)R * PB * PB + (size_t)(16 * w + lr) * PB + 4 * lk;
    double a[8][4];
#pragma unroll
    for (int kc = 0; kc < 8; ++kc)
        if (kc <= w) load4(xr + 16 * kc, a[kc]);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int e = t + 512 * q;
        Ls[e >> 6][e & 63] = v[q];
    }
    __syncthreads();
    d4 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) {
        if (kc > w) break;
        const int kb = 16 * kc + 4 * lk;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[b] = mfma4(a[kc][j], Ls[kb + j][16 * b + lr], acc[b]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            F[(r0 + 16 * w + 4 * r + lk) * Vp + c0 + 16 * b + lr] = (T)acc[b][r];
}

// X_R = L_RR^-1 for the 128 x 128 diagonal blocks of the FP32 factor, in
// FP64, from the inverses of its 64 x 64 diagonal blocks (Linv64, row-major,
// one per block, from the panel factorisation) and the block L_10 between
// them: L_RR = [L_00 0; L_10 L_11] gives
//   X_R = [L_00^-1 0; -L_11^-1 L_10 L_00^-1  L_11^-1].
__global__ __launch_bounds__(256) void k_inv128_f32(const float* __restrict__ F, int Vp,
                                                    const float* __restrict__ Linv64,
                                                    double* __restrict__ X) {
    __shared__ double I0[64][65], I1[64][65], Cm[64][65];
    const int R = blockIdx.x, t = threadIdx.x;
    const float* i0 = Linv64 + (size_t)(2 * R) * 64 * 64;
    const float* i1 = i0 + 64 * 64;
    const size_t r1 = (size_t)(2 * R + 1) * 64, c0 = (size_t)(2 * R) * 64;
    for (int e = t; e < 64 * 64; e += 256) {
        const int i = e >> 6, j = e & 63;
        I0[i][j] = j <= i ? (double)i0[e] : 0.0;
        I1[i][j] = j <= i ? (double)i1[e] : 0.0;
        Cm[i][j] = (double)F[(r1 + i) * Vp + c0 + j];
    }
    __syncthreads();
    // T = L_10 L_00^-1 (I0[k][j] = 0 for k < j); thread: rows (t >> 6) + 4 s, column t & 63
    const int j = t & 63;
    double tv[16];
    for (int s = 0; s < 16; ++s) {
        const int i = (t >> 6) + 4 * s;
        double acc = 0.0;
        for (int k = j; k < 64; ++k) acc = fma(Cm[i][k], I0[k][j], acc);
        tv[s] = acc;
    }
    __syncthreads();
    for (int s = 0; s < 16; ++s) Cm[(t >> 6) + 4 * s][j] = tv[s];
    __syncthreads();
    double* xr = X + (size_t)R * PB * PB;
    for (int s = 0; s < 16; ++s) {
        const int i = (t >> 6) + 4 * s;
        double acc = 0.0;
        for (int k = 0; k <= i; ++k) acc = fma(I1[i][k], Cm[k][j], acc);
        xr[(size_t)(64 + i) * PB + j] = -acc;
        xr[(size_t)i * PB + j] = I0[i][j];
        xr[(size_t)i * PB + 64 + j] = 0.0;
        xr[(size_t)(64 + i) * PB + 64 + j] = I1[i][j];
    }
}

// B_R = X_R S_R (trans = 0) or X_R^T S_R (trans = 1) for the 128-row block
// R = blockIdx.x and the 32-column chunk blockIdx.y of nrhs columns
// (column-major, leading dimension ld). S = src, or din o src (rows >= V: 0)
// when din is given; the result is scaled by dout (rows >= V: 0) when given.
// src may be B: the block's columns are staged in LDS before any result is
// stored. Wave w forms row tile w; X_R is lower triangular, so k < 16 (w + 1)
// (trans = 0) or k >= 16 w (trans = 1).
__global__ __launch_bounds__(512) void k_bdiag(const double* __restrict__ X, const double* src,
                                               double* B, int ld, int nrhs, int trans,
                                               const double* __restrict__ din,
                                               const double* __restrict__ dout, int V,
                                               const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double Ss[PB][RHS_CHUNK + 1];
    const int R = blockIdx.x;
    const int nc = min(RHS_CHUNK, nrhs - (int)blockIdx.y * RHS_CHUNK);
    src += (size_t)blockIdx.y * RHS_CHUNK * ld;
    B += (size_t)blockIdx.y * RHS_CHUNK * ld;
    const size_t r0 = (size_t)R * PB;
    const int t = threadIdx.x, lane = t & 63, lr = lane & 15, lk = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    for (int e = t; e < PB * RHS_CHUNK; e += 512) {
        const int i = e & (PB - 1), c = e >> 7;
        double v = 0.0;
        if (c < nc) {
            v = src[(size_t)c * ld + r0 + i];
            if (din) v = ((int)(r0 + i) < V) ? v * din[r0 + i] : 0.0;
        }
        Ss[i][c] = v;
    }
    __syncthreads();
    const bool two = nc > 16;
    const double* xb = X + (size_t)R * PB * PB;
    d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    const int klo = trans ? w : 0, khi = trans ? PB / 16 - 1 : w;
    for (int kc = klo; kc <= khi; ++kc) {
        const int kb = 16 * kc + 4 * lk;
        double a[4];
        if (!trans) {
            load4(xb + (size_t)(16 * w + lr) * PB + kb, a);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) a[j] = xb[(size_t)(kb + j) * PB + 16 * w + lr];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc0 = mfma4(a[j], Ss[kb + j][lr], acc0);
            if (two) acc1 = mfma4(a[j], Ss[kb + j][16 + lr], acc1);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = (int)r0 + 16 * w + 4 * r + lk;
        const double s = dout ? (row < V ? dout[row] : 0.0) : 1.0;
        if (lr < nc) B[(size_t)lr * ld + row] = acc0[r] * s;
        if (two && 16 + lr < nc) B[(size_t)(16 + lr) * ld + row] = acc1[r] * s;
    }
}

// 4 consecutive doubles at byte offset off of a buffer resource, as two
// 16-byte loads with device-scope coherence (sc1: the load_wt of each
// double; published values are written once before their flag, so the
// 16-byte reads need no atomicity). Resource word 3 is the gfx9 default.
typedef int v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t wt_rsrc(const double* base, size_t doubles) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0,
                                             (int)(doubles * sizeof(double)), 0x00020000);
}
__device__ __forceinline__ void load4_wt(__amdgpu_buffer_rsrc_t r, uint32_t off, double (&v)[4]) {
    const double2 x = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 16));
    const double2 y =
        __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, 16));
    v[0] = x.x;
    v[1] = x.y;
    v[2] = y.x;
    v[3] = y.y;
}

// Sweep shape for G workgroups per 128-row block: RW rows per workgroup in
// RT tiles of 16; the 8 waves are RT row tiles x KS slices of KL k each
// (NG groups of 16 k).
template <int G>
struct Sw {
    static constexpr int RW = PB / G, RT = RW / 16, KS = 8 / RT, KL = PB / KS, NG = KL / 16;
    static_assert(RT * KS == 8 && NG >= 1, "sweep shape");
};

// One sweep over the unit block triangular L~ (forward: L~ y = b, block
// order 0 .. nb-1; backward: L~^T z = y, block order nb-1 .. 0), in place on
// nrhs columns of B (column-major, leading dimension ldb), RHS_CHUNK columns
// per grid row y (flags row y: nb G flags). The workgroup of (block, slice h)
// forms rows 128 block + RW h .. of
//   fwd: y_R = b_R - sum_{C < R} L~_RC y_C
//   bwd: z_C = y_C - sum_{R > C} L~_RC^T z_R
// taking the dependencies in chain order, each as soon as its G flags are
// up; the last one (the neighbouring block) is the chain's link. Wave
// (rt, ks) accumulates row tile rt over its k slice in MFMA registers
// (A = factor rows, each lane four consecutive k; B = the dependency's
// published values, read write-through); the slices are summed in LDS once,
// at the end.
template <typename T, int G, bool BWD>
__global__ __launch_bounds__(512) void k_usweep(const T* __restrict__ F, int Vp, double* B,
                                                int ldb, int nrhs, int nb,
                                                int32_t* __restrict__ flags, int32_t epoch,
                                                int32_t* __restrict__ status,
                                                const int32_t* __restrict__ skip,
                                                uint32_t* __restrict__ ticket, uint32_t tbase) {
    using S = Sw<G>;
    const int tk = take_ticket(ticket, tbase);
    if (skip && *skip) return;
    const int nch = (int)gridDim.y;
    if (tk < 0 || tk >= nb * G * nch) {
        if (threadIdx.x == 0) atomicOr(status, 2);
        return;
    }
    const int u = tk / nch, by = tk % nch;
    const int h = u % G, blk = BWD ? nb - 1 - u / G : u / G;
    B += (size_t)by * RHS_CHUNK * ldb;
    nrhs = min(RHS_CHUNK, nrhs - by * RHS_CHUNK);
    flags += (size_t)by * nb * G;
    __shared__ double red[8][2][16][17];
    const int t = threadIdx.x, lane = t & 63, lr = lane & 15, lk = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int rt = w % S::RT, ks = w / S::RT;
    const size_t rb = (size_t)blk * PB + S::RW * h;   // first row of this workgroup
    const bool two = nrhs > 16;
    // B operand columns, clamped into the chunk (columns >= nrhs are never stored)
    const int c0 = lr < nrhs ? lr : nrhs - 1;
    const int c1 = 16 + lr < nrhs ? 16 + lr : nrhs - 1;
    // this thread's outputs: row 16 s + oi of row tile s, column oc
    const int oi = (t >> 4) & 15, oc = (t >> 8) * 16 + (t & 15);
    const bool ocv = oc < nrhs;
    double b0[S::RT];
#pragma unroll
    for (int s = 0; s < S::RT; ++s) b0[s] = ocv ? B[(size_t)oc * ldb + rb + 16 * s + oi] : 0.0;
    d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    const int kofs = S::KL * ks + 4 * lk;   // this lane's first k in a 16-k group
    auto loadA = [&](int d, double (&v)[S::NG][4]) {
#pragma unroll
        for (int g = 0; g < S::NG; ++g) {
            const int k = kofs + 16 * g;
            if (!BWD) {
                load4(F + (rb + 16 * rt + lr) * Vp + (size_t)d * PB + k, v[g]);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    v[g][j] = (double)F[((size_t)d * PB + k + j) * Vp + rb + 16 * rt + lr];
            }
        }
    };
    // the chunk's columns as one buffer resource (byte offsets < 2^31)
    const __amdgpu_buffer_rsrc_t yr = wt_rsrc(B, (size_t)nrhs * ldb);
    auto product = [&](int d, const double (&v)[S::NG][4]) {
        const uint32_t y = (uint32_t)((size_t)d * PB + kofs) * 8u;
        double y0[S::NG][4], y1[S::NG][4];
#pragma unroll
        for (int g = 0; g < S::NG; ++g)
            load4_wt(yr, y + (uint32_t)((size_t)c0 * ldb + 16 * g) * 8u, y0[g]);
        if (two) {
#pragma unroll
            for (int g = 0; g < S::NG; ++g)
                load4_wt(yr, y + (uint32_t)((size_t)c1 * ldb + 16 * g) * 8u, y1[g]);
        }
#pragma unroll
        for (int g = 0; g < S::NG; ++g)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc0 = mfma4(-v[g][j], y0[g][j], acc0);
                if (two) acc1 = mfma4(-v[g][j], y1[g][j], acc1);
            }
    };
    const int ndep = BWD ? nb - 1 - blk : blk;
    double a[S::NG][4], an[S::NG][4];
    if (ndep > 0) loadA(BWD ? nb - 1 : 0, a);
    for (int q = 0; q < ndep; ++q) {
        const int d = BWD ? nb - 1 - q : q;
        wait_flags_wt(flags + (size_t)d * G, G, epoch, status);
        if (q + 1 < ndep) loadA(BWD ? d - 1 : d + 1, an);
        product(d, a);
        if (q + 1 < ndep) {
#pragma unroll
            for (int g = 0; g < S::NG; ++g)
#pragma unroll
                for (int j = 0; j < 4; ++j) a[g][j] = an[g][j];
        }
    }
    // sum the KS slices: wave w = rt + RT ks holds rows 4 r + lk, column lr
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        red[w][0][4 * r + lk][lr] = acc0[r];
        if (two) red[w][1][4 * r + lk][lr] = acc1[r];
    }
    __syncthreads();
    const int oct = t >> 8;
#pragma unroll
    for (int s = 0; s < S::RT; ++s) {
        double v = b0[s];
        if (oct == 0 || two) {
#pragma unroll
            for (int q = 0; q < S::KS; ++q) v += red[q * S::RT + s][oct][oi][t & 15];
        }
        if (ocv) store_wt(B + (size_t)oc * ldb + rb + 16 * s + oi, v);
    }
    publish_wt(flags + (size_t)blk * G + h, epoch);
}

// One sweep (forward or backward) over L~ on nrhs columns of B (ld = Vp),
// bwd_chunks chunks of RHS_CHUNK columns per launch.
template <typename T>
rc_status sweep(rc_ctx* ctx, const T* F, double* B, int nrhs, bool bwd, const int32_t* skip) {
    constexpr int G = RC_SWEEP_G;
    const int nb = ctx->nblk / 2, Vp = ctx->nblk * RC_NB;
    const int per = RHS_CHUNK * ctx->bwd_chunks;
    for (int c0 = 0; c0 < nrhs; c0 += per) {
        const int nc = (nrhs - c0) < per ? (nrhs - c0) : per;
        const int nch = (nc + RHS_CHUNK - 1) / RHS_CHUNK;
        const int32_t ep = ++ctx->epoch;
        const dim3 grid(nb * G, nch);
        double* Bc = B + (size_t)c0 * Vp;
        if (bwd)
            hipLaunchKernelGGL((k_usweep<T, G, true>), grid, dim3(512), 0, ctx->stream, F, Vp, Bc,
                               Vp, nc, nb, ctx->d_flags, ep, ctx->d_status, skip, ctx->d_ticket,
                               ctx->tickets);
        else
            hipLaunchKernelGGL((k_usweep<T, G, false>), grid, dim3(512), 0, ctx->stream, F, Vp, Bc,
                               Vp, nc, nb, ctx->d_flags, ep, ctx->d_status, skip, ctx->d_ticket,
                               ctx->tickets);
        RC_CHECK_LAUNCH();
        ctx->tickets += (uint32_t)(nb * G * nch);
    }
    return RC_OK;
}

rc_status bdiag(rc_ctx* ctx, const double* X, const double* src, double* B, int nrhs, int trans,
                const double* din, const double* dout, const int32_t* skip) {
    const int Vp = ctx->nblk * RC_NB;
    const dim3 grid(ctx->nblk / 2, (nrhs + RHS_CHUNK - 1) / RHS_CHUNK);
    hipLaunchKernelGGL(k_bdiag, grid, dim3(512), 0, ctx->stream, X, src, B, Vp, nrhs, trans, din,
                       dout, ctx->V, skip);
    RC_CHECK_LAUNCH();
    return RC_OK;
}

}  // namespace

// Row-scale the FP64 factor in d_F by the inverses of its 128 x 128
// diagonal blocks (d_Linv128, from the panel factorisations).
rc_status rc_ltilde_fp64(rc_ctx* ctx) {
    const int nb = ctx->nblk / 2, Vp = ctx->nblk * RC_NB;
    if (nb > 1) {
        hipLaunchKernelGGL(k_ltilde<double>, dim3(nb * (nb - 1)), dim3(512), 0, ctx->stream,
                           ctx->d_F, Vp, (const double*)ctx->d_Linv128, 1);
        RC_CHECK_LAUNCH();
    }
    return RC_OK;
}

// Inverses of the FP32 factor's 128 x 128 diagonal blocks (d_X128m, FP64),
// then the row scaling of d_F32 by them.
rc_status rc_ltilde_f32(rc_ctx* ctx) {
    const int nb = ctx->nblk / 2, Vp = ctx->nblk * RC_NB;
    if (!ctx->d_X128m)
        RC_HIP(hipMalloc((void**)&ctx->d_X128m, sizeof(double) * (size_t)nb * PB * PB));
    hipLaunchKernelGGL(k_inv128_f32, dim3(nb), dim3(256), 0, ctx->stream,
                       (const float*)ctx->d_F32, Vp, (const float*)ctx->d_Linv32, ctx->d_X128m);
    RC_CHECK_LAUNCH();
    if (nb > 1) {
        hipLaunchKernelGGL(k_ltilde<float>, dim3(nb * (nb - 1)), dim3(512), 0, ctx->stream,
                           ctx->d_F32, Vp, (const double*)ctx->d_X128m, 1);
        RC_CHECK_LAUNCH();
    }
    return RC_OK;
}

// L Y = B in place for nrhs columns of length Vp (column-major, ld = Vp)
// with the FP64 factor (the solves of the saddle refinement; the gradient
// solve's forward substitution happens inside the factorisation).
rc_status rc_solve_fp64_forward(rc_ctx* ctx, double* d_B, int nrhs) {
    rc_status s = bdiag(ctx, ctx->d_Linv128, d_B, d_B, nrhs, 0, nullptr, nullptr, nullptr);
    return s == RC_OK ? sweep<double>(ctx, ctx->d_F, d_B, nrhs, false, nullptr) : s;
}

// L^T X = Y in place for nrhs columns of length Vp (column-major, ld = Vp)
// with the FP64 factor; Y = L^-1 B comes out of the factorisation's
// right-hand-side rows.
rc_status rc_solve_fp64_backward(rc_ctx* ctx, double* d_Y, int nrhs) {
    rc_status s = sweep<double>(ctx, ctx->d_F, d_Y, nrhs, true, nullptr);
    return s == RC_OK ? bdiag(ctx, ctx->d_Linv128, d_Y, d_Y, nrhs, 1, nullptr, nullptr, nullptr)
                      : s;
}

// W = D (L L^T)^-1 D R for nrhs columns (ld = Vp) with the FP32 factor L of
// the scaled S = D A_hat D (+ shift), D = d_dscale; R may be W. Every kernel
// returns at once while *skip is set.
rc_status rc_precond_f32(rc_ctx* ctx, const double* R, double* W, int nrhs,
                         const int32_t* skip) {
    rc_status s = bdiag(ctx, ctx->d_X128m, R, W, nrhs, 0, ctx->d_dscale, nullptr, skip);
    if (s == RC_OK) s = sweep<float>(ctx, ctx->d_F32, W, nrhs, false, skip);
    if (s == RC_OK) s = sweep<float>(ctx, ctx->d_F32, W, nrhs, true, skip);
    if (s == RC_OK) s = bdiag(ctx, ctx->d_X128m, W, W, nrhs, 1, nullptr, ctx->d_dscale, skip);
    return s;
}

// Workgroups of one sweep launch the device holds at once: the smallest
// per-CU residency of the sweep kernels times the CU count.
int rc_trsm_max_coresident(void) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    int per[4] = {0, 0, 0, 0};
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[0], k_usweep<double, RC_SWEEP_G, true>, 512, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[1], k_usweep<double, RC_SWEEP_G, false>, 512, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[2], k_usweep<float, RC_SWEEP_G, true>, 512, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[3], k_usweep<float, RC_SWEEP_G, false>, 512, 0) !=
            hipSuccess)
        return 0;
    int m = per[0];
    for (int i = 1; i < 4; ++i) m = per[i] < m ? per[i] : m;
    return m * ncu;
}

// Right-hand-side chunks of one sweep launch: nb G workgroups per chunk, no
// more in one launch than the device holds at once.
int rc_sweep_chunks(int nblk) {
    const int per_chunk = (nblk / 2) * RC_SWEEP_G;
    const int cores = rc_trsm_max_coresident();
    const int c = per_chunk > 0 ? cores / per_chunk : 1;
    return c < 1 ? 1 : (c > RC_BWD_MAXCH ? RC_BWD_MAXCH : c);
}

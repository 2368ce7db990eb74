// Device-side hand-offs between workgroups of one launch (triangular
// sweeps) and between the factorisation's two streams: completion flags and
// counters in global memory, write-through tiles, and the ticket that orders
// a sweep's workgroups by start time.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

__device__ __forceinline__ void wait_flag(const int32_t* flag, int32_t epoch, int32_t* status) {
    if (threadIdx.x == 0) {
        int spins = 0;
        // poll with relaxed loads (an acquire load would invalidate the L2
        // on every spin), then one acquire fence once the flag is seen
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1 << 24)) {
                atomicOr(status, 2);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Write-through hand-off between the triangular sweeps' blocks. A block's
// result tile is stored with agent-scope atomic stores (coherent at the
// device level: never held dirty in one XCD's L2), every wave waits for its
// stores to complete before the barrier, then one thread stores the flag;
// consumers poll the flag and read such tiles only with agent-scope atomic
// loads. A release fence would write back the XCD's whole L2 instead
// (__threadfence() plus a release store: two writebacks and an invalidate),
// 3.0 against 1.3 us per hop along a 128-block chain on this device.
__device__ __forceinline__ void store_wt(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double load_wt(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void publish_wt(int32_t* flag, int32_t epoch) {
    // the signal fences keep the compiler from moving the tile's stores
    // across the wait for their completion
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wait_flag_wt(const int32_t* flag, int32_t epoch, int32_t* status) {
    if (threadIdx.x == 0) {
        int spins = 0;
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1 << 24)) {
                atomicOr(status, 2);
                break;
            }
        }
    }
    __syncthreads();
}

// wait_flag_wt for the n <= 64 flags flag[0 .. n) (thread i polls flag i)
__device__ __forceinline__ void wait_flags_wt(const int32_t* flag, int n, int32_t epoch,
                                              int32_t* status) {
    if ((int)threadIdx.x < n) {
        int spins = 0;
        while (__hip_atomic_load(flag + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
               epoch) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1 << 24)) {
                atomicOr(status, 2);
                break;
            }
        }
    }
    __syncthreads();
}

// A sweep block's position comes from a ticket counter, not from blockIdx:
// tickets are taken in the order the workgroups start, and a block waits
// only for blocks of smaller tickets, which are running or done, so the
// sweep needs no co-residency of its grid. The counter only grows; each
// launch passes the count issued before it (wraps modulo 2^32).
__device__ __forceinline__ int take_ticket(uint32_t* ctr, uint32_t base) {
    __shared__ int tk;
    if (threadIdx.x == 0) tk = (int)(atomicAdd(ctr, 1u) - base);
    __syncthreads();
    return tk;
}

// Completion counter of a producer kernel whose workgroups wrote their
// results with store_wt: every wave waits for its stores to complete, then
// one thread adds 1 (consumers poll with wait_flag for the number of
// producer workgroups; its acquire drops their stale L2 lines). A release
// here would write back the XCD's L2, which the rest of the trailing update
// keeps full of dirty lines.
__device__ __forceinline__ void signal_done_wt(int32_t* cnt) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

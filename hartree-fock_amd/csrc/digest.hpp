// digest.hpp - accumulation helpers shared by the digest kernels (k_eri.hip, k_jket.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace hfx {
namespace digest {

// Accumulation into the copy private to this XCD (no other XCD touches these lines): FP64
// atomics, or in the deterministic mode the fixed-point pairs below.
// Deterministic accumulation: v as a two-word fixed-point number, hi = round(v 2^32) and
// lo = round((v 2^32 - hi) 2^40), each added with an integer atomic.  Integer sums do not
// depend on the order of the additions, so the result is bitwise reproducible; the rounding
// error of one term is at most 2^-73 and the range is |sum| < 2^31.
__device__ __forceinline__ void fixed_add(unsigned long long* q, double v) {
    const double x = ldexp(v, 32);
    const double r = rint(x);
    const long long hi = (long long)r;
    const long long lo = (long long)rint(ldexp(x - r, 40));
    if (hi) __hip_atomic_fetch_add(q, (unsigned long long)hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lo) __hip_atomic_fetch_add(q + 1, (unsigned long long)lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Element t of an accumulator block: a double (FX false) or a fixed-point pair at [2t, 2t+1]
// (FX true).  The accumulation mode is a kernel template parameter, not a run-time flag: the
// fixed-point path's temporaries would otherwise count in every kernel's register allocation
// at the digest's peak (dsps: one wavefront per SIMD instead of two).
template <bool FX>
__device__ __forceinline__ void acc_add(double* base, long long t, double v) {
    if (v == 0.0) return;
    if constexpr (!FX)
        __hip_atomic_fetch_add(base + t, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        fixed_add(reinterpret_cast<unsigned long long*>(base) + 2 * t, v);
}
// Accumulation into the block's LDS exchange rows.
__device__ __forceinline__ void lds_add(double* p, double v) {
    if (v != 0.0) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Index of the XCD running this workgroup (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ int xcd_id() { return (int)(__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 15u); }

// Sums x over each run of consecutive lanes whose `same` flags (from run_flags) say they share
// a key; afterwards the run's first lane holds the run's total.
__device__ __forceinline__ double run_sum(double x, const bool (&same)[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double o = __shfl_down(x, 1 << k);
        if (same[k]) x += o;
    }
    return x;
}
// For the lanes of a wavefront: head = first lane of a run of equal keys; same[k] = lane + 2^k
// is in the same run.
__device__ __forceinline__ void run_flags(int key, bool& head, bool (&same)[6]) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1);
    head = (lane == 0) || (prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int run = __popcll(heads & le);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int off = 1 << k;
        const int r = __shfl_down(run, off);
        same[k] = (lane + off < 64) && (r == run);
    }
}

}  // namespace digest
}  // namespace hfx

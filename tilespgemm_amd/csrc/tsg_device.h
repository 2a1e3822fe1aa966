// Device-side wave64 primitives for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ int wave_incl_scan(int x) {
    const int l = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (l >= d) x += y;
    }
    return x;
}

__device__ __forceinline__ long long wave_incl_scan64(long long x) {
    const int l = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        long long y = __shfl_up(x, d, 64);
        if (l >= d) x += y;
    }
    return x;
}

// Inclusive prefix sum within aligned groups of G lanes (G power of two <= 64).
template <int G>
__device__ __forceinline__ int group_incl_scan(int x) {
    const int l = __lane_id() & (G - 1);
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        int y = __shfl_up(x, d, G);
        if (l >= d) x += y;
    }
    return x;
}

template <int G, class T>
__device__ __forceinline__ T group_sum(T x) {
#pragma unroll
    for (int d = G / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, G);
    return x;
}

__device__ __forceinline__ int wave_sum(int x) { return group_sum<64>(x); }

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        int y = __shfl_xor(x, d, 64);
        x = x > y ? x : y;
    }
    return x;
}

// One accumulation step of an output entry: fused multiply-add, or a rounded
// product followed by a rounded add.  The library is built with
// -ffp-contract=off so the separate form is never contracted.
template <bool FUSED>
__device__ __forceinline__ double acc_step(double acc, double a, double b) {
    if (FUSED) return __builtin_fma(a, b, acc);
    double p = a * b;
    return acc + p;
}

// Ordering point inside one wave: later LDS operations observe earlier ones
// of any lane of the same wave.
__device__ __forceinline__ void wave_lds_fence() {
    // LDS operations of one wave complete in issue order, so ordering only has
    // to be enforced against the compiler.
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Exclusive lower bound of key in sorted a[lo, hi).
__device__ __forceinline__ int lower_bound_i32(const int *a, int lo, int hi, int key) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Hash used by the row path's LDS tables (multiplicative, HASH_SCALE 107 of
// the reference's common.h:43 spreads consecutive columns).
__device__ __forceinline__ unsigned hash_col(int j, unsigned mask) {
    return ((unsigned)j * 107u) & mask;
}

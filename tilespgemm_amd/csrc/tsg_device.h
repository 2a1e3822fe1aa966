// Device-side wave64 primitives for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Inclusive prefix sum across the 64 lanes of a wave.
// DPP move of x (row_shr, row_bcast); lanes without a source, or in rows the
// row mask leaves out, read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_or_zero(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, ROW_MASK, 0xf, false);
}

// Inclusive prefix sum within aligned groups of G lanes (G = 16, 32 or 64):
// shifts within 16-lane rows, then the row broadcasts carry row totals into
// the following rows.  Every lane of a group must be active; no lane reads
// outside its own group.
template <int G>
__device__ __forceinline__ int dpp_incl_scan(int x) {
    static_assert(G == 16 || G == 32 || G == 64, "DPP scans work on 16, 32 or 64 lanes");
    x += dpp_or_zero<0x111>(x);            // row_shr:1
    x += dpp_or_zero<0x112>(x);            // row_shr:2
    x += dpp_or_zero<0x114>(x);            // row_shr:4
    x += dpp_or_zero<0x118>(x);            // row_shr:8
    if (G >= 32) x += dpp_or_zero<0x142, 0xa>(x);   // row_bcast:15 into rows 1, 3
    if (G >= 64) x += dpp_or_zero<0x143, 0xc>(x);   // row_bcast:31 into rows 2, 3
    return x;
}

__device__ __forceinline__ int wave_incl_scan(int x) { return dpp_incl_scan<64>(x); }

// Inclusive running maximum within aligned groups of G lanes (G = 16, 32 or
// 64) of values >= -1; same lane rules as dpp_incl_scan.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_or_minus1(int x) {
    return __builtin_amdgcn_update_dpp(-1, x, CTRL, ROW_MASK, 0xf, false);
}
template <int G>
__device__ __forceinline__ int dpp_incl_max(int x) {
    static_assert(G == 16 || G == 32 || G == 64, "DPP scans work on 16, 32 or 64 lanes");
    x = max(x, dpp_or_minus1<0x111>(x));
    x = max(x, dpp_or_minus1<0x112>(x));
    x = max(x, dpp_or_minus1<0x114>(x));
    x = max(x, dpp_or_minus1<0x118>(x));
    if (G >= 32) x = max(x, dpp_or_minus1<0x142, 0xa>(x));
    if (G >= 64) x = max(x, dpp_or_minus1<0x143, 0xc>(x));
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
    if constexpr (G >= 16) {
        return dpp_incl_scan<G>(x);
    } else {
        const int l = __lane_id() & (G - 1);
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
            int y = __shfl_up(x, d, G);
            if (l >= d) x += y;
        }
        return x;
    }
}

// Value of the group's first lane (a scalar read when the group is the wave).
template <int G>
__device__ __forceinline__ int group_first(int x) {
    if constexpr (G == 64) return __builtin_amdgcn_readlane(x, 0);
    else return __shfl(x, 0, G);
}

// Value of the group's last lane (a scalar read when the group is the wave).
template <int G>
__device__ __forceinline__ int group_last(int x) {
    if constexpr (G == 64) return __builtin_amdgcn_readlane(x, 63);
    else return __shfl(x, G - 1, G);
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

// Sets a bit of an LDS bitmap word unless it is already set.  While a
// bitmap is being marked its bits are only ever set, so a stale read costs
// at most a redundant atomic; repeated outputs (half the products of
// R-MAT's rows, most of a stencil's) skip the atomic, and same-word atomics
// of a wave would serialise.
__device__ __forceinline__ void lds_set_bit(unsigned *word, unsigned bit) {
    if (!(*word & bit)) atomicOr(word, bit);
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

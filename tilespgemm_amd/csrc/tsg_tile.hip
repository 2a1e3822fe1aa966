// Tile path: TileSpGEMM on 16x16 bitmask tiles for gfx950.
//
// Tile format (kept alike to the reference so the tile arrays can be checked
// against it; main_spgemm-col.cu:4968-5360): per tile a 16-entry uchar row
// pointer (rows past the matrix edge hold the tile nnz), 16 ushort row masks
// with bit (15 - c) for in-tile column c, packed uchar (r << 4) | c indices and
// FP64 values, tiles sorted by tile column inside a tile row.
//
// Stages (reference numbering in brackets):
//   convert  CSR -> tile-CSR [host code main_spgemm-col.cu:4968-5360]: one
//            16-lane group per tile row walks its 16 CSR rows and extracts
//            tiles in column order by a group-wide min reduction.
//   step1    C tile structure [step1+step2, spgemm_nsparse_kernel.h hash on
//            tile patterns]: one wave per tile row marks the tile columns
//            reachable through B in an LDS bitmap (flat expansion over the
//            tile pattern), so the tile list comes out sorted.
//   step2    C row masks and counts [step3, main_spgemm-col.cu:997-1325]:
//            one wave per C tile row, four tiles at a time, lane r of a tile
//            owns row r; matched (A_IK, B_KJ) pairs come from a ballot-driven
//            intersection; maskC[r] = OR of B_KJ row masks selected by the set
//            bits of A_IK row r.  Also produces every (tile, row) segment's
//            position inside its CSR row, so step3 writes CSR directly, and
//            each C tile's count of matched pairs.
//   pairs    the matched pairs themselves, stored per C tile in ascending K,
//            and the numeric phase's work units: each C tile's 64-output
//            parts and its remainder, binned by size (<= 16, <= 32, more).
//   step3    FP64 numeric [step4, main_spgemm-col.cu:2094-2397,3386-3754]:
//            one lane per output of a work unit keeps the output's sum in a
//            register and takes its products -- the set bits of
//            maskA[r] & cmaskB[c] of every pair, pairs in ascending K -- with
//            one FMA each, in A-row order (no atomics, no LDS accumulator),
//            then writes the entry straight into C's CSR.
#include "tsg_internal.h"
#include "tsg_device.h"
#include "tsg_expand.h"

#include <algorithm>
#include <climits>

namespace {

template <int G>
__device__ __forceinline__ int group_min(int x) {
#pragma unroll
    for (int d = G / 2; d > 0; d >>= 1) {
        int y = __shfl_xor(x, d, G);
        x = x < y ? x : y;
    }
    return x;
}

__device__ __forceinline__ int wave_min(int x) { return group_min<64>(x); }

// ---------------------------------------------------------------- convert
// One 16-lane group per tile row; lane r handles CSR row 16*I + r.  A
// block's 16 tile rows hold one contiguous range of ja: when it has at most
// CV_STAGE entries it is staged in LDS first (coalesced loads), and the rows
// are walked there instead of with a dependent global load per entry (the
// 27-point stencil's 6,912 entries per block fit).
constexpr int CV_STAGE = 8192;

__global__ void __launch_bounds__(256) k_tile_convert(
    int m, int blkm, const int *__restrict__ ia, const int *__restrict__ ja,
    const int *__restrict__ tile_ptr, int *__restrict__ tile_col, int *__restrict__ tile_off,
    uint16_t *__restrict__ mask, uint8_t *__restrict__ rowptr, uint8_t *__restrict__ idx,
    int *__restrict__ perm) {
    __shared__ int s_col[CV_STAGE];
    const int I0 = blockIdx.x * (blockDim.x >> 4);
    const int e0 = ia[I0 * 16];
    const int e1 = ia[min(m, (I0 + (int)(blockDim.x >> 4)) * 16)];
    const bool staged = e1 - e0 <= CV_STAGE;
    if (staged)
        for (int i = threadIdx.x; i < e1 - e0; i += blockDim.x) s_col[i] = ja[e0 + i];
    __syncthreads();
    auto col_at = [&](int p) { return staged ? s_col[p - e0] : ja[p]; };
    const int I = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int r = threadIdx.x & 15;
    if (I >= blkm) return;
    const int row = I * 16 + r;
    int cur = 0, end = 0;
    if (row < m) { cur = ia[row]; end = ia[row + 1]; }
    int t = tile_ptr[I];
    int off = ia[I * 16];
    while (true) {
        const int Jr = cur < end ? (col_at(cur) >> 4) : INT_MAX;
        const int J = group_min<16>(Jr);
        if (J == INT_MAX) break;
        const int c0 = cur;
        unsigned mk = 0;
        while (cur < end) {
            const int col = col_at(cur);
            if ((col >> 4) != J) break;
            mk |= 1u << (15 - (col & 15));
            cur++;
        }
        const int cnt = cur - c0;
        const int incl = group_incl_scan<16>(cnt);
        const int tot = __shfl(incl, 15, 16);
        const int rp = incl - cnt;
        rowptr[(size_t)t * 16 + r] = (uint8_t)rp;
        mask[(size_t)t * 16 + r] = (uint16_t)mk;
        if (r == 0) {
            tile_col[t] = J;
            tile_off[t] = off;
        }
        for (int i = 0; i < cnt; i++) {
            const int q = off + rp + i;
            idx[q] = (uint8_t)((r << 4) | (col_at(c0 + i) & 15));
            perm[c0 + i] = q;
        }
        off += tot;
        t++;
    }
    if (r == 0 && I == blkm - 1) tile_off[t] = off;
}

// Tile count of each tile row through an LDS bitmap over its tile columns
// (all entries of a tile row are contiguous in CSR).  One wave per tile row;
// the bitmap window covers 65536 tile columns per pass.
constexpr int TC_WAVES = 4;
constexpr int TC_WORDS = 2048;
constexpr int TC_U = 4;

// Row pointers are clamped to [0, nnz] and columns to [0, ncols): every
// access stays in bounds (and the tile-column loop bounded) whatever the
// entries hold.  A bit already set is not set again (power-law rows repeat a
// few tile columns, whose words would serialise the LDS atomics); TC_U
// entries per lane and pass have their loads issued together.
template <int NT>
struct TileCols {
    int p1, cmax;
    const int *ja;
    __device__ void load(int p, int (&J)[TC_U]) const {
#pragma unroll
        for (int q = 0; q < TC_U; q++) {
            const int pq = p + q * NT;
            J[q] = pq < p1 ? min(max(ja[pq], 0), cmax) >> 4 : -1;
        }
    }
    // tile-column range of the entries [p0 + t, p1) taken by thread t
    __device__ void range(int p0, int t, int &lo, int &hi) const {
        lo = INT_MAX;
        hi = -1;
        for (int p = p0 + t; p < p1; p += NT * TC_U) {
            int J[TC_U];
            load(p, J);
#pragma unroll
            for (int q = 0; q < TC_U; q++)
                if (J[q] >= 0) {
                    lo = min(lo, J[q]);
                    hi = max(hi, J[q]);
                }
        }
    }
    __device__ void mark(int p0, int t, int w0, int w1, unsigned *bits) const {
        for (int p = p0 + t; p < p1; p += NT * TC_U) {
            int J[TC_U];
            load(p, J);
#pragma unroll
            for (int q = 0; q < TC_U; q++)
                if (J[q] >= w0 && J[q] < w1) {
                    const unsigned bit = 1u << ((J[q] - w0) & 31);
                    unsigned *word = &bits[(J[q] - w0) >> 5];
                    if (!(*word & bit)) atomicOr(word, bit);
                }
        }
    }
};

__global__ void __launch_bounds__(64 * TC_WAVES) k_tile_count(int m, int ncols, int64_t nnz,
                                                              int blkm,
                                                              const int *__restrict__ ia,
                                                              const int *__restrict__ ja,
                                                              int *__restrict__ tcount) {
    __shared__ unsigned s_bits[TC_WAVES][TC_WORDS];
    unsigned *bits = s_bits[threadIdx.x >> 6];
    const int I = blockIdx.x * TC_WAVES + (threadIdx.x >> 6);
    const int l = __lane_id();
    if (I >= blkm) return;
    const int p0 = (int)min((int64_t)max(ia[I * 16], 0), nnz);
    const int p1 = (int)min((int64_t)max(ia[min(m, I * 16 + 16)], p0), nnz);
    const TileCols<64> tc{p1, max(ncols - 1, 0), ja};
    int lo, hi;
    tc.range(p0, l, lo, hi);
    lo = wave_min(lo);
    hi = wave_max(hi);
    int count = 0;
    for (int w0 = lo; w0 <= hi; w0 += TC_WORDS * 32) {
        const int w1 = min(hi + 1, w0 + TC_WORDS * 32);
        const int nwords = (w1 - w0 + 31) >> 5;
        for (int i = l; i < nwords; i += 64) bits[i] = 0u;
        wave_lds_fence();
        tc.mark(p0, l, w0, w1, bits);
        wave_lds_fence();
        for (int i = l; i < nwords; i += 64) count += __popc(bits[i]);
        wave_lds_fence();
    }
    count = wave_sum(count);
    if (l == 0) tcount[I] = count;
}

// AUTO's density sample: tile rows 0, step, 2*step, ... (the grid has one
// block per sampled row, so every block's row is < blkm), one 256-thread
// block each sharing a bitmap of 262144 tile columns per pass -- a
// power-law matrix puts most of its entries in its first tile rows, and one
// wave per row left the sample waiting on those.  The tile and entry counts
// are added to sums[0] and sums[1].  Indices are clamped as above, so the
// sample can run beside the input's validation instead of after it.
constexpr int TS_WORDS = 8192;

__global__ void __launch_bounds__(256) k_tile_sample(int m, int ncols, int64_t nnz, int step,
                                                     const int *__restrict__ ia,
                                                     const int *__restrict__ ja,
                                                     unsigned long long *__restrict__ sums) {
    __shared__ unsigned bits[TS_WORDS];
    __shared__ int s_lo[4], s_hi[4];
    __shared__ int s_cnt[4];
    const int I = blockIdx.x * step;
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int p0 = (int)min((int64_t)max(ia[I * 16], 0), nnz);
    const int p1 = (int)min((int64_t)max(ia[min(m, I * 16 + 16)], p0), nnz);
    const TileCols<256> tc{p1, max(ncols - 1, 0), ja};
    int lo, hi;
    tc.range(p0, t, lo, hi);
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (l == 0) {
        s_lo[w] = lo;
        s_hi[w] = hi;
    }
    __syncthreads();
    lo = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
    hi = max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3]));
    int count = 0;
    // lo and hi are the block's: every thread takes every pass and barrier
    for (int w0 = lo; w0 <= hi; w0 += TS_WORDS * 32) {
        const int w1 = min(hi + 1, w0 + TS_WORDS * 32);
        const int nwords = (w1 - w0 + 31) >> 5;
        for (int i = t; i < nwords; i += 256) bits[i] = 0u;
        __syncthreads();
        tc.mark(p0, t, w0, w1, bits);
        __syncthreads();
        for (int i = t; i < nwords; i += 256) count += __popc(bits[i]);
        __syncthreads();
    }
    count = wave_sum(count);
    if (l == 0) s_cnt[w] = count;
    __syncthreads();
    if (t == 0) {
        atomicAdd(&sums[0], (unsigned long long)(s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3]));
        atomicAdd(&sums[1], (unsigned long long)(p1 - p0));
    }
}

__global__ void k_perm_scatter(int64_t nnz, const int *__restrict__ perm,
                               const double *__restrict__ v, double *__restrict__ out) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) out[perm[p]] = v[p];
}

// The same values into two layouts (C = A*A: A's row-major payload and B's
// column-major one), read once.
__global__ void k_perm_scatter2(int64_t nnz, const int *__restrict__ perm,
                                const int *__restrict__ perm2, const double *__restrict__ v,
                                double *__restrict__ out, double *__restrict__ out2) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) {
        const double x = v[p];
        out[perm[p]] = x;
        out2[perm2[p]] = x;
    }
}

// ------------------------------------------------------------------ step1
constexpr int S1_WAVES = 4;
constexpr int S1_WORDS = 2048;           // 65536 tile columns per pass

template <bool FILL>
__global__ void __launch_bounds__(64 * S1_WAVES) k_step1(
    int blkm, const int *__restrict__ tptrA, const int *__restrict__ tcolA,
    const int *__restrict__ tptrB, const int *__restrict__ tcolB, int *__restrict__ cnt,
    const int *__restrict__ tptrC, int *__restrict__ tcolC, int *__restrict__ trowC) {
    __shared__ unsigned s_bits[S1_WAVES][S1_WORDS];
    __shared__ int s_mark[S1_WAVES][64];
    unsigned *bits = s_bits[threadIdx.x >> 6];
    const int I = blockIdx.x * S1_WAVES + (threadIdx.x >> 6);
    const int l = __lane_id();
    s_mark[threadIdx.x >> 6][l] = -1;
    if (I >= blkm) return;
    const int pa = tptrA[I], pe = tptrA[I + 1];
    int lo = INT_MAX, hi = -1;
    for (int p = pa + l; p < pe; p += 64) {
        const int K = tcolA[p];
        const int b0 = tptrB[K], b1 = tptrB[K + 1];
        if (b0 < b1) {
            lo = min(lo, tcolB[b0]);
            hi = max(hi, tcolB[b1 - 1]);
        }
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    int count = 0;
    const int base = FILL ? tptrC[I] : 0;
    for (int w0 = lo; w0 <= hi; w0 += S1_WORDS * 32) {
        const int w1 = min(hi + 1, w0 + S1_WORDS * 32);
        const int nwords = (w1 - w0 + 31) >> 5;
        for (int i = l; i < nwords; i += 64) bits[i] = 0u;
        wave_lds_fence();
        for_products<64, 4, false, false>(pa, pe, 0, 64, tcolA, nullptr, tptrB, tcolB, nullptr,
                                          BlockPtrs(), 0, 0, s_mark[threadIdx.x >> 6],
                                          [&](bool act, int J, double, double, bool) {
                                           if (act && J >= w0 && J < w1) {
                                               const int c = J - w0;
                                               lds_set_bit(&bits[c >> 5], 1u << (c & 31));
                                           }
                                       });
        wave_lds_fence();
        for (int i0 = 0; i0 < nwords; i0 += 64) {
            const int i = i0 + l;
            unsigned word = i < nwords ? bits[i] : 0u;
            const int pc = __popc(word);
            const int incl = wave_incl_scan(pc);
            if (FILL) {
                int pos = base + count + incl - pc;
                while (word) {
                    const int b = __ffs(word) - 1;
                    word &= word - 1;
                    tcolC[pos] = w0 + (i << 5) + b;
                    trowC[pos] = I;
                    pos++;
                }
            }
            count += __shfl(incl, 63, 64);
        }
        wave_lds_fence();
    }
    if (!FILL && l == 0) cnt[I] = count;
}

// ------------------------------------------------- pair intersection helper
// Calls f(pa, pb) for the matched pairs of C tile (I, J) in ascending K:
// pa = position of A_IK in A's tile-CSR, pb = position of B_KJ in B's
// tile-CSR.  Executed by the 16 lanes of group g (lanes 16g .. 16g+15).
// The tile columns K of one A tile row as an LDS bitmap over the window
// [kmin, kmin + 32 nw) with the count of set bits before every word, built by
// a whole wave: the position in A's tile-CSR of tile K is then pa0 + pre[w] +
// popc(word below K) -- no search.  ok is false (and the bitmap unused) when
// the window exceeds RT_WORDS words.
constexpr int RT_WORDS = 512;            // 16384 tile columns

struct RowTiles {
    unsigned *words;
    uint16_t *pre;
    int kmin = 0, nw = 0;
    bool ok = false;

    __device__ __forceinline__ void build(const int *__restrict__ tcolA, int pa0, int pa1) {
        const int l = __lane_id();
        ok = false;
        nw = 0;
        if (pa1 <= pa0) return;
        kmin = tcolA[pa0];
        nw = ((tcolA[pa1 - 1] - kmin) >> 5) + 1;
        if (nw > RT_WORDS) return;
        ok = true;
        for (int w = l; w < nw; w += 64) words[w] = 0u;
        wave_lds_fence();
        for (int p = pa0 + l; p < pa1; p += 64) {
            const int d = tcolA[p] - kmin;
            atomicOr(&words[d >> 5], 1u << (d & 31));
        }
        wave_lds_fence();
        int running = 0;
        for (int base = 0; base < nw; base += 64) {
            const int w = base + l;
            const int c = w < nw ? __popc(words[w]) : 0;
            const int incl = wave_incl_scan(c);
            if (w < nw) pre[w] = (uint16_t)(running + incl - c);
            running += __builtin_amdgcn_readlane(incl, 63);
        }
        wave_lds_fence();
    }
    // position of tile column K in the row (pa0-relative), or -1
    __device__ __forceinline__ int find(int K) const {
        const int d = K - kmin;
        if (d < 0 || d >= nw * 32) return -1;
        const unsigned word = words[d >> 5];
        const unsigned bit = 1u << (d & 31);
        if (!(word & bit)) return -1;
        return (int)pre[d >> 5] + __popc(word & (bit - 1u));
    }
};

// With rt built for A's tile row (ok), a B column list of at most 8 times
// the row's length is walked in order and each entry looked up in the
// bitmap; otherwise the shorter list is walked with a binary search in the
// other.  Either way matches come in ascending K.
template <class F>
__device__ __forceinline__ void for_pairs(bool valid, int pa0, int pa1, int cb0, int cb1,
                                          const int *__restrict__ tcolA,
                                          const int *__restrict__ csc_row,
                                          const int *__restrict__ csc_tile, F &&f,
                                          const RowTiles *rt = nullptr) {
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
    const int lenA = valid ? pa1 - pa0 : 0, lenB = valid ? cb1 - cb0 : 0;
    const bool bm = rt && rt->ok && lenB <= 8 * lenA;
    const bool iterA = !bm && lenA <= lenB;
    const int len = iterA ? lenA : lenB;
    for (int i0 = 0; i0 < len; i0 += 16) {
        const int i = i0 + r;
        bool match = false;
        int pa = 0, pb = 0;
        if (i < len) {
            if (bm) {
                const int K = csc_row[cb0 + i];
                const int pos = rt->find(K);
                if (pos >= 0) { match = true; pa = pa0 + pos; pb = csc_tile[cb0 + i]; }
            } else if (iterA) {
                const int K = tcolA[pa0 + i];
                const int pos = lower_bound_i32(csc_row, cb0, cb1, K);
                if (pos < cb1 && csc_row[pos] == K) { match = true; pa = pa0 + i; pb = csc_tile[pos]; }
            } else {
                const int K = csc_row[cb0 + i];
                const int pos = lower_bound_i32(tcolA, pa0, pa1, K);
                if (pos < pa1 && tcolA[pos] == K) { match = true; pa = pos; pb = csc_tile[cb0 + i]; }
            }
        }
        unsigned bits = (unsigned)((__ballot(match) >> (g * 16)) & 0xffffull);
        while (bits) {
            const int j = __ffs(bits) - 1;
            bits &= bits - 1;
            const int spa = __shfl(pa, g * 16 + j, 64);
            const int spb = __shfl(pb, g * 16 + j, 64);
            f(spa, spb);
        }
    }
}

// ------------------------------------------------------------------ step2
constexpr int S2_WAVES = 4;

// With pslot (slots of C tile row I from pslot[I], each of its C tiles
// taking as many as the row has A tiles -- a bound on its matched pairs),
// the matched pairs are stored as they are found, at tslot[t], and npair[t]
// counts them; without, npair[t] counts them for k_fill_pairs.
__global__ void __launch_bounds__(64 * S2_WAVES) k_step2(
    int m, int blkm, const int *__restrict__ tptrA, const int *__restrict__ tcolA,
    const uint16_t *__restrict__ maskA, const int *__restrict__ csc_ptr,
    const int *__restrict__ csc_row, const int *__restrict__ csc_tile,
    const uint16_t *__restrict__ maskB, const int *__restrict__ tptrC,
    const int *__restrict__ tcolC, uint16_t *__restrict__ maskC, uint8_t *__restrict__ rowptrC,
    int64_t *__restrict__ tnnz, int *__restrict__ rowoffC, int64_t *__restrict__ rownnz,
    int64_t *__restrict__ npair, const int64_t *__restrict__ pslot,
    int64_t *__restrict__ tslot, int2 *__restrict__ pairs) {
    __shared__ unsigned s_words[S2_WAVES][RT_WORDS];
    __shared__ uint16_t s_pre[S2_WAVES][RT_WORDS];
    const int I = blockIdx.x * S2_WAVES + (threadIdx.x >> 6);
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
    if (I >= blkm) return;
    const int pa0 = tptrA[I], pa1 = tptrA[I + 1];
    const int t0 = tptrC[I], t1 = tptrC[I + 1];
    RowTiles rt{s_words[threadIdx.x >> 6], s_pre[threadIdx.x >> 6]};
    rt.build(tcolA, pa0, pa1);
    const int64_t slot0 = pslot ? pslot[I] : 0;
    int running = 0;
    for (int base = t0; base < t1; base += 4) {
        const int t = base + g;
        const bool valid = t < t1;
        const int J = valid ? tcolC[t] : 0;
        const int cb0 = valid ? csc_ptr[J] : 0, cb1 = valid ? csc_ptr[J + 1] : 0;
        unsigned mc = 0;
        int np = 0;
        const int64_t slot = slot0 + (int64_t)(t - t0) * (pa1 - pa0);
        for_pairs(valid, pa0, pa1, cb0, cb1, tcolA, csc_row, csc_tile, [&](int pa, int pb) {
            if (pslot && r == 0) pairs[slot + np] = make_int2(pa, pb);
            np++;
            // B_KJ's 16 row masks in registers (the group's lanes read the
            // same 32 bytes), then the rows selected by A_IK's row r
            const unsigned ma = maskA[(size_t)pa * 16 + r];
            const uint4 *m4 = reinterpret_cast<const uint4 *>(maskB + (size_t)pb * 16);
            const uint4 lo = m4[0], hi = m4[1];
            const unsigned w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
            for (int k = 0; k < 16; k++)
                if (ma & (0x8000u >> k)) mc |= (w[k >> 1] >> ((k & 1) * 16)) & 0xffffu;
        }, &rt);
        const int pc = __popc(mc);
        const int incl = group_incl_scan<16>(pc);
        if (valid) {
            maskC[(size_t)t * 16 + r] = (uint16_t)mc;
            rowptrC[(size_t)t * 16 + r] = (uint8_t)(incl - pc);
            if (r == 15) tnnz[t] = incl;
            if (r == 0) npair[t] = np;
            if (pslot && r == 1) tslot[t] = slot;
        }
        int pre = 0, tot = 0;
#pragma unroll
        for (int gg = 0; gg < 4; gg++) {
            const int v = __shfl(pc, gg * 16 + r, 64);
            if (gg < g) pre += v;
            tot += v;
        }
        if (valid) rowoffC[(size_t)t * 16 + r] = running + pre;
        running += tot;
    }
    const int row = I * 16 + r;
    if (g == 0 && row < m) rownnz[row] = running;
}

// Pair slots of C tile row I: its C tiles times its A tiles.
__global__ void __launch_bounds__(256) k_pair_slots(int blkm, const int *__restrict__ tptrA,
                                                    const int *__restrict__ cntC,
                                                    int64_t *__restrict__ slots) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I < blkm) slots[I] = (int64_t)cntC[I] * (tptrA[I + 1] - tptrA[I]);
}

// ------------------------------------------------------------- pairs, bins
// Matched (A_IK, B_KJ) pairs of every C tile, in ascending K, at the tile's
// offset tpair[t] (the exclusive scan of the counts step2 wrote): one wave
// per C tile row, four tiles at a time, the intersection of step2 again.
constexpr int FP_WAVES = 4;

__global__ void __launch_bounds__(64 * FP_WAVES) k_fill_pairs(
    int blkm, const int *__restrict__ tptrA, const int *__restrict__ tcolA,
    const int *__restrict__ csc_ptr, const int *__restrict__ csc_row,
    const int *__restrict__ csc_tile, const int *__restrict__ tptrC,
    const int *__restrict__ tcolC, const int64_t *__restrict__ tpair, int2 *__restrict__ pairs) {
    __shared__ unsigned s_words[FP_WAVES][RT_WORDS];
    __shared__ uint16_t s_pre[FP_WAVES][RT_WORDS];
    const int I = blockIdx.x * FP_WAVES + (threadIdx.x >> 6);
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
    if (I >= blkm) return;
    const int pa0 = tptrA[I], pa1 = tptrA[I + 1];
    const int t0 = tptrC[I], t1 = tptrC[I + 1];
    RowTiles rt{s_words[threadIdx.x >> 6], s_pre[threadIdx.x >> 6]};
    rt.build(tcolA, pa0, pa1);
    for (int base = t0; base < t1; base += 4) {
        const int t = base + g;
        const bool valid = t < t1;
        const int J = valid ? tcolC[t] : 0;
        const int cb0 = valid ? csc_ptr[J] : 0, cb1 = valid ? csc_ptr[J + 1] : 0;
        int64_t pos = valid ? tpair[t] : 0;
        for_pairs(valid, pa0, pa1, cb0, cb1, tcolA, csc_row, csc_tile, [&](int pa, int pb) {
            if (r == 0) pairs[pos] = make_int2(pa, pb);
            pos++;
        }, &rt);
    }
}

// Work units of the numeric phase.  A C tile of n outputs gives n / 64
// units of 64 outputs and, when n % 64 is not 0, one unit of the remaining
// outputs; every lane owns one output.  Units go to bins by their size:
// bin 0 <= 16 outputs (four units per wave), bin 1 <= 32 (two per wave),
// bin 2 more (one per wave).  A unit is t * 4 + part, its outputs the tile's
// 64 * part onwards (int64).  cnt[0..2] counts units, then the scatter
// places them at their bin's start (bins back to back) plus a slot from
// fill.  (The 27-point stencil's larger C tiles hold 74 outputs; their
// 10-output remainders, a wave each, had cost as much as the 64-output
// parts: 1.44 ms of one-unit waves before, with half of them remainders.
// A wave per 64-output part beats one wave with two outputs per lane, 1.43
// vs 1.78 ms, and staging pair values in LDS as well as the metadata, 1.77
// ms: the kernel is bound by its VALU instructions per product and by the
// busiest lane of each wave, not by load latency.)
constexpr int TBINS = 3;

__device__ __forceinline__ int rem_bin(int64_t nnz) {
    const int rem = (int)(nnz & 63);
    return rem == 0 ? -1 : rem <= 16 ? 0 : rem <= 32 ? 1 : 2;
}

// units of a tile with nnz outputs in bin q
__device__ __forceinline__ int tile_units_in(int64_t nnz, int q) {
    return (q == 2 ? (int)(nnz >> 6) : 0) + (rem_bin(nnz) == q ? 1 : 0);
}

// Units per bin: summed over each wave first (one LDS atomic per wave and
// bin; a thread's own atomics on the three counters had serialised).
__global__ void __launch_bounds__(256) k_tbin_count(int n, const int64_t *__restrict__ off,
                                                    int *__restrict__ cnt) {
    __shared__ int s_cnt[TBINS];
    if (threadIdx.x < TBINS) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nz = t < n ? off[t + 1] - off[t] : 0;
#pragma unroll
    for (int q = 0; q < TBINS; q++) {
        const int u = wave_sum(tile_units_in(nz, q));
        if (__lane_id() == 0 && u) atomicAdd(&s_cnt[q], u);
    }
    __syncthreads();
    if (threadIdx.x < TBINS && s_cnt[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], s_cnt[threadIdx.x]);
}

// A unit's record carries its tile's pairs (first one, count), so the
// numeric kernel's first loads do not wait on each other.
// Slots come from a wave scan of the unit counts per bin, one LDS atomic
// per wave and bin, and one global atomic per block and bin (a global
// atomic per tile on the counters serialised the scatter).  Bin 2 takes a
// tile's parts 0, 1, ... (the remainder last, when it is there); bins 0 and
// 1 take only a remainder, part n / 64.
// The tile's pairs: from tslot[t] with npair[t] of them when step2 stored
// them in slots, else from the scanned counts tpair.
__global__ void __launch_bounds__(256) k_tbin_scatter(int n, const int64_t *__restrict__ off,
                                                      const int64_t *__restrict__ tpair,
                                                      const int64_t *__restrict__ tslot,
                                                      const int *__restrict__ cnt,
                                                      int *__restrict__ fill,
                                                      int4 *__restrict__ units) {
    __shared__ int s_cnt[TBINS], s_base[TBINS];
    if (threadIdx.x < TBINS) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int64_t nz = t < n ? off[t + 1] - off[t] : 0;
    int pos[TBINS] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < TBINS; q++) {
        const int v = tile_units_in(nz, q);
        if (!__ballot(v != 0)) continue;
        const int incl = wave_incl_scan(v);
        const int wtot = __builtin_amdgcn_readlane(incl, 63);
        int wbase = 0;
        if (lane == 0) wbase = atomicAdd(&s_cnt[q], wtot);
        wbase = __builtin_amdgcn_readlane(wbase, 0);
        pos[q] = wbase + incl - v;
    }
    __syncthreads();
    if (threadIdx.x < TBINS && s_cnt[threadIdx.x]) {
        int start = 0;
        for (int i = 0; i < (int)threadIdx.x; i++) start += cnt[i];
        s_base[threadIdx.x] = start + atomicAdd(&fill[threadIdx.x], s_cnt[threadIdx.x]);
    }
    __syncthreads();
    const int full = (int)(nz >> 6);
    const int rb = rem_bin(nz);
    const int u2 = tile_units_in(nz, 2);
    if (u2 == 0 && rb < 0) return;
    const int64_t p0 = tslot ? tslot[t] : tpair[t];
    const int np = (int)(tslot ? tpair[t] : tpair[t + 1] - p0);
    auto rec = [&](int part) {
        return make_int4(t, part | (np << 2), (int)(uint32_t)(uint64_t)p0,
                         (int)(uint32_t)((uint64_t)p0 >> 32));
    };
    for (int i = 0; i < u2; i++) units[s_base[2] + pos[2] + i] = rec(i);
    if (rb == 0 || rb == 1) units[s_base[rb] + pos[rb]] = rec(full);
}

// Tile-CSC index of a square tile pattern that is symmetric (tile (K, J)
// present exactly when (J, K) is, as in stencil and mesh matrices): column
// J's tiles are then row J's tile columns K, in the same ascending order, and
// each one's index is that of tile (K, J), found by a search in row K.  16
// lanes per tile row.  A tile whose partner is missing sets *asym (the
// general transpose then replaces the index).  csc_ptr is tile_ptr.
__global__ void __launch_bounds__(256) k_sym_csc(int blk, const int *__restrict__ tile_ptr,
                                                 const int *__restrict__ tile_col,
                                                 int *__restrict__ csc_row,
                                                 int *__restrict__ csc_tile,
                                                 unsigned *__restrict__ asym) {
    const int J = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    bool missing = false;
    if (J < blk) {
        for (int q = tile_ptr[J] + l; q < tile_ptr[J + 1]; q += 16) {
            const int K = tile_col[q];
            const int lo = tile_ptr[K], hi = tile_ptr[K + 1];
            const int pos = lower_bound_i32(tile_col, lo, hi, J);
            const bool found = pos < hi && tile_col[pos] == J;
            csc_row[q] = K;
            csc_tile[q] = found ? pos : 0;
            missing |= !found;
        }
    }
    // one plain store per wave that found a tile without its partner (an
    // asymmetric pattern has millions: atomics on one word would queue)
    const unsigned long long bad = __ballot(missing);
    if (bad && __lane_id() == __ffsll(bad) - 1) *asym = 1u;
}

// Column layout of 16x16 tiles (B only), the transposed sub-tiles of
// main_spgemm-col.cu:5255-5290, as tile records: mask[c] has bit 15 - k set
// when the tile holds entry (k, c) -- the transpose of the row masks, in the
// same bit order; start[c] counts the tile's entries in columns before c;
// and cpos[q] is the position of row-major payload entry q in the tile's
// column-major payload (columns ascending, rows ascending in a column).
// 16 lanes per tile, lane c builds column c (a group shares its tile, so
// whole groups leave together).
__global__ void __launch_bounds__(256) k_col_layout(int numblk, const int *__restrict__ tile_off,
                                                    const uint16_t *__restrict__ mask,
                                                    const uint8_t *__restrict__ rowptr,
                                                    TileRec *__restrict__ rec,
                                                    int *__restrict__ cpos) {
    const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int c = threadIdx.x & 15;
    if (t >= numblk) return;
    const uint4 *m4 = reinterpret_cast<const uint4 *>(mask + t * 16);
    const uint4 lo = m4[0], hi = m4[1];
    const unsigned w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    unsigned cm = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const unsigned row = (w[k >> 1] >> ((k & 1) * 16)) & 0xffffu;
        cm |= ((row >> (15 - c)) & 1u) << (15 - k);
    }
    const int cnt = __popc(cm);
    const int cstart = group_incl_scan<16>(cnt) - cnt;
    const int off = tile_off[t];
    rec[t].mask[c] = (uint16_t)cm;
    rec[t].start[c] = (uint8_t)cstart;
    if (c == 0) rec[t].off = off;
    const uint4 rp = *reinterpret_cast<const uint4 *>(rowptr + t * 16);
    int j = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const unsigned row = (w[k >> 1] >> ((k & 1) * 16)) & 0xffffu;
        if ((row >> (15 - c)) & 1u) {
            const int rw = (int)((k < 8 ? (k < 4 ? rp.x : rp.y) : (k < 12 ? rp.z : rp.w)) >>
                                 ((k & 3) * 8)) & 255;
            cpos[off + rw + __popc(row >> (16 - c))] = off + cstart + j;
            j++;
        }
    }
}

// Row layout of 16x16 tiles as tile records (A): row masks, row starts and
// payload offset, four lanes per tile writing 16 bytes each.
__global__ void __launch_bounds__(256) k_row_records(int numblk, const int *__restrict__ tile_off,
                                                     const uint16_t *__restrict__ mask,
                                                     const uint8_t *__restrict__ rowptr,
                                                     TileRec *__restrict__ rec) {
    const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    const int j = threadIdx.x & 3;
    if (t >= numblk) return;
    uint4 v;
    if (j < 2)
        v = reinterpret_cast<const uint4 *>(mask + t * 16)[j];
    else if (j == 2)
        v = *reinterpret_cast<const uint4 *>(rowptr + t * 16);
    else
        v = make_uint4((unsigned)tile_off[t], 0u, 0u, 0u);
    reinterpret_cast<uint4 *>(rec + t)[j] = v;
}

// permc[i] = cpos[perm[i]]: CSR position -> column-major payload position.
__global__ void k_compose_perm(int64_t nnz, const int *__restrict__ perm,
                               const int *__restrict__ cpos, int *__restrict__ permc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) permc[i] = cpos[perm[i]];
}

// ------------------------------------------------------------------ step3
// FP64 numeric phase.  Every output's accumulator stays in a register of
// the lane that owns the output and takes the output's products in A-row
// order: matched pairs in ascending K (stored by k_fill_pairs), then the
// columns k of A tile row r in ascending order.  The products of output
// (r, c) of a pair are the k set in maskA[r] & cmaskB[c] (row r of A_IK
// and column c of B_KJ both hold k).  A's values are row-major in its tile
// and B's column-major, so both operands of a product sit at a base fixed
// for the pair plus the rank of k in a mask fixed for the pair: A row r's
// start + popc of maskA[r] above bit 15-k, B column c's start + popc of
// cmaskB[c] above it.  A chunk of the tile's pairs has its fixed-size
// metadata (payload offsets, A row starts and masks, B column starts and
// masks: 104 bytes a pair) staged in LDS by one lane per pair with coalesced
// loads, so a product costs a few bit operations and two value loads; no
// atomics.
struct TileArgs {
    const int *trowC, *tcolC;
    const uint16_t *maskC;
    const uint8_t *rowptrC;
    const int *rowoffC;
    const int2 *pairs;
    const TileRec *recA;         // row layout
    const double *valA;
    const TileRec *recB;         // column layout
    const double *valB;          // column-major inside each tile
    const int64_t *ic;
    int *jc;
    double *c;
};

__device__ __forceinline__ unsigned byte_at(const uint4 &v, int i) {
    const unsigned w = i < 8 ? (i < 4 ? v.x : v.y) : (i < 12 ? v.z : v.w);
    return (w >> ((i & 3) * 8)) & 255u;
}

// A matched pair's two tile records, as staged in LDS.
struct PairMeta {
    TileRec a, b;
};

// Stage pair pr's metadata (called by one lane per pair): two 64-byte
// records.
__device__ __forceinline__ void stage_pair(const TileArgs &T, int2 pr, PairMeta &M) {
    const uint4 *ra = reinterpret_cast<const uint4 *>(T.recA + pr.x);
    const uint4 *rb = reinterpret_cast<const uint4 *>(T.recB + pr.y);
    const uint4 a0 = ra[0], a1 = ra[1], a2 = ra[2], a3 = ra[3];
    const uint4 b0 = rb[0], b1 = rb[1], b2 = rb[2], b3 = rb[3];
    uint4 *d = reinterpret_cast<uint4 *>(&M);
    d[0] = a0;
    d[1] = a1;
    d[2] = a2;
    d[3] = a3;
    d[4] = b0;
    d[5] = b1;
    d[6] = b2;
    d[7] = b3;
}

// A lane's products over the staged pairs, in order: pair by pair (the
// pairs with products for this lane are the set bits of live, taken from the
// lowest), and inside a pair the k set in maskA[r] & cmaskB[c] from the
// smallest.  next() yields the A and B value positions of the next product
// (false when none is left), so a lane's loop runs as many times as it has
// products, not as many as the busiest lane has in every pair.
struct ProductWalk {
    const PairMeta *meta;
    int r, c;
    unsigned live;
    unsigned hits = 0, ma = 0, cm = 0;
    int abase = 0, bbase = 0;
    __device__ __forceinline__ bool next(int &ai, int &bi) {
        if (hits == 0) {
            if (live == 0) return false;
            const PairMeta &M = meta[__builtin_ctz(live)];
            live &= live - 1;
            ma = M.a.mask[r];
            cm = M.b.mask[c];
            hits = ma & cm;
            abase = M.a.off + M.a.start[r];
            bbase = M.b.off + M.b.start[c];
        }
        const int bit = 31 - __clz(hits);          // highest bit: smallest k
        hits ^= 1u << bit;
        ai = abase + __popc(ma >> (bit + 1));
        bi = bbase + __popc(cm >> (bit + 1));
        return true;
    }
};

// Accumulates a lane's products over the staged pairs meta[0..n) into acc,
// in order, STEP_U at a time: the positions of the next STEP_U products come
// from the walk (LDS and bit operations only), their 2 STEP_U value loads
// are issued together, then the FMAs run in order (positions past the
// lane's last product repeat the first and are not accumulated).
constexpr int STEP_U = 4;

template <bool FUSED>
__device__ __forceinline__ double walk_products(const TileArgs &T, const PairMeta *meta, int n,
                                                int r, int c, double acc) {
    unsigned live = 0;
    for (int j = 0; j < n; j++)
        if (meta[j].a.mask[r] & meta[j].b.mask[c]) live |= 1u << j;
    ProductWalk w{meta, r, c, live};
    while (true) {
        int a[STEP_U], b[STEP_U];
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < STEP_U; u++) {
            a[u] = u ? a[0] : 0;
            b[u] = u ? b[0] : 0;
            cnt += w.next(a[u], b[u]) ? 1 : 0;
        }
        if (cnt == 0) break;
        double x[STEP_U], y[STEP_U];
#pragma unroll
        for (int u = 0; u < STEP_U; u++) {
            x[u] = T.valA[a[u]];
            y[u] = T.valB[b[u]];
        }
#pragma unroll
        for (int u = 0; u < STEP_U; u++)
            if (u < cnt) acc = acc_step<FUSED>(acc, x[u], y[u]);
        if (cnt < STEP_U) break;
    }
    return acc;
}

// Work units of <= 64 / TPW outputs, TPW at a time per wave (G = 64 / TPW
// lanes each), UPW rounds per wave over consecutive units; a unit is a tile
// and its 64-output part, and lane lq owns the part's lq-th output in
// row-major order.  Pairs are staged CP = 32 / TPW at a time per unit (3328
// bytes of LDS a wave).  The loads that need only the unit (the tile's row
// starts and masks, its tile row and column, the first chunk of pairs) are
// issued together, and the output's position in C right after the lane's
// row is known, so they overlap the staging.
constexpr int UPW = 1;

template <bool FUSED, int TPW>
__global__ void __launch_bounds__(256) k_step3_out(int nunits, const int4 *__restrict__ units,
                                                   TileArgs T) {
    constexpr int G = 64 / TPW;
    constexpr int CP = 32 / TPW;
    static_assert(CP <= 32, "a lane's live pairs are a 32-bit mask");
    __shared__ PairMeta s_meta[4][TPW][CP];
    const int lane = __lane_id();
    const int lq = lane % G, g = lane / G;
    PairMeta *meta = s_meta[threadIdx.x >> 6][g];
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    for (int round = 0; round < UPW; round++) {
        const int64_t slot = (wave * UPW + round) * TPW + g;
        if (slot >= nunits) break;
        const int4 u = units[slot];
        const int t = u.x;
        const int q = (u.y & 3) * 64 + lq;       // the tile's output this lane owns
        const int np = (int)((unsigned)u.y >> 2);
        const int64_t p0 = (int64_t)(((uint64_t)(uint32_t)u.w << 32) | (uint32_t)u.z);
        const uint4 rp = *reinterpret_cast<const uint4 *>(T.rowptrC + (size_t)t * 16);
        const uint4 mk0 = *reinterpret_cast<const uint4 *>(T.maskC + (size_t)t * 16);
        const uint4 mk1 = *reinterpret_cast<const uint4 *>(T.maskC + (size_t)t * 16 + 8);
        const int I = T.trowC[t], J = T.tcolC[t];
        int2 pr0 = make_int2(0, 0);
        if (lq < min(CP, np)) pr0 = T.pairs[p0 + lq];
        // row of output q: the last row whose first output is <= q (rows hold
        // their tile-local start in rowptrC), then its (q - start)-th column;
        // lanes past the unit's outputs stage pairs but own no output
        int r = 0, start = 0;
#pragma unroll
        for (int i = 1; i < 16; i++) {
            const int st = (int)byte_at(rp, i);
            if (st <= q) {
                r = i;
                start = st;
            }
        }
        // rows after r with the same start are empty, so r holds q exactly
        // when q is below r's end
        const unsigned mw = r < 8 ? (r < 4 ? (r < 2 ? mk0.x : mk0.y) : (r < 6 ? mk0.z : mk0.w))
                                  : (r < 12 ? (r < 10 ? mk1.x : mk1.y) : (r < 14 ? mk1.z : mk1.w));
        const unsigned mrow = (mw >> ((r & 1) * 16)) & 0xffffu;
        const bool owner = q - start < __popc(mrow);
        int64_t o = 0;
        if (owner) o = T.ic[(int64_t)I * 16 + r] + T.rowoffC[(size_t)t * 16 + r] + (q - start);
        int c = 0;
        if (owner) {
            unsigned mm = mrow;
            for (int j = 0; j < q - start; j++) mm &= ~(1u << (31 - __clz(mm)));
            c = 15 - (31 - __clz(mm));
        }
        double acc = 0.0;
        for (int pc = 0; pc < np; pc += CP) {
            const int n = min(CP, np - pc);
            wave_lds_fence();
            if (lq < n) stage_pair(T, pc == 0 ? pr0 : T.pairs[p0 + pc + lq], meta[lq]);
            wave_lds_fence();
            if (owner) acc = walk_products<FUSED>(T, meta, n, r, c, acc);
        }
        if (owner) {
            T.jc[o] = J * 16 + c;
            T.c[o] = acc;
        }
    }
}

// tile counts per tile row, scanned into tile_ptr
void tile_count_scan(const CsrIn &X, int blkm, int *tile_ptr, Arena &ws, hipStream_t s) {
    int *tcount = ws.alloc_n<int>((size_t)blkm + 1);
    int64_t *tmp = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)blkm + 1));
    if (blkm > 0) {
        hipLaunchKernelGGL(k_tile_count, dim3(grid_for(blkm, TC_WAVES)), dim3(64 * TC_WAVES), 0,
                           s, X.rows, X.cols, X.nnz, blkm, X.ptr, X.idx, tcount);
        TSG_LAUNCH_CHECK();
    }
    exclusive_scan_i32(tcount, tile_ptr, (size_t)blkm, tmp, s);
}

// early: the tile count already taken (tile_count_early), else counted here
// with a host round trip of its own
void convert(TileMat &T, const CsrIn &X, bool fill_values, Arena &ws, hipStream_t s,
             const EarlyTileCount *early) {
    T.rows = X.rows;
    T.cols = X.cols;
    T.blkm = (X.rows + 15) / 16;
    T.blkn = (X.cols + 15) / 16;
    T.nnz = X.nnz;
    int nt = 0;
    if (early && early->nt >= 0) {
        T.tile_ptr = early->tile_ptr;
        nt = early->nt;
    } else {
        T.tile_ptr = ws.alloc_n<int>((size_t)T.blkm + 1);
        tile_count_scan(X, T.blkm, T.tile_ptr, ws, s);
        TSG_HIP_TRY(
            hipMemcpyAsync(&nt, T.tile_ptr + T.blkm, sizeof(int), hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipStreamSynchronize(s));
    }
    T.numblk = nt;
    T.tile_col = ws.alloc_n<int>((size_t)nt + 1);
    T.tile_off = ws.alloc_n<int>((size_t)nt + 1);
    T.mask = ws.alloc_n<uint16_t>((size_t)nt * 16 + 16);
    T.rowptr = ws.alloc_n<uint8_t>((size_t)nt * 16 + 16);
    T.idx = ws.alloc_n<uint8_t>((size_t)X.nnz + 1);
    T.val = ws.alloc_n<double>((size_t)X.nnz + 1);
    T.perm = ws.alloc_n<int>((size_t)X.nnz + 1);
    if (nt == 0) {
        TSG_HIP_TRY(hipMemsetAsync(T.tile_off, 0, sizeof(int), s));
    } else {
        hipLaunchKernelGGL(k_tile_convert, dim3(grid_for((size_t)T.blkm * 16, 256)), dim3(256),
                           0, s, X.rows, T.blkm, X.ptr, X.idx, T.tile_ptr, T.tile_col,
                           T.tile_off, T.mask, T.rowptr, T.idx, T.perm);
        TSG_LAUNCH_CHECK();
    }
    if (fill_values && X.val && X.nnz > 0) {
        hipLaunchKernelGGL(k_perm_scatter, dim3(grid_for((size_t)X.nnz, 256)), dim3(256), 0, s,
                           X.nnz, T.perm, X.val, T.val);
        TSG_LAUNCH_CHECK();
    }
}

}  // namespace

void tile_count_early(const CsrIn &X, EarlyTileCount &e, Arena &ws, hipStream_t s) {
    const int blkm = (X.rows + 15) / 16;
    e.tile_ptr = ws.alloc_n<int>((size_t)blkm + 1);
    tile_count_scan(X, blkm, e.tile_ptr, ws, s);
    TSG_HIP_TRY(
        hipMemcpyAsync(&e.nt, e.tile_ptr + blkm, sizeof(int), hipMemcpyDeviceToHost, s));
}

void check_tile_count(int64_t n_c_tiles) {
    TSG_CHECK(n_c_tiles >= 0 && n_c_tiles <= INT_MAX, TSG_ERR_OVERFLOW,
              "the tile path indexes C tiles with int32: " + std::to_string(n_c_tiles) +
                  " C tiles exceed 2^31-1 (use the row path)");
}

void tile_symbolic(TilePlan &p, const CsrIn &A, const CsrIn &B, bool b_is_a, int64_t *d_ic,
                   int64_t *nnz_c, Arena &ws, hipStream_t s, StageTimer &tm, int cus,
                   const EarlyTileCount *earlyA, const EarlyTileCount *earlyB) {
    (void)cus;
    p = TilePlan();
    p.b_is_a = b_is_a;
    const int m = A.rows;
    tm.begin("tile.convert");
    convert(p.A, A, false, ws, s, earlyA);
    if (b_is_a) {
        // B shares A's structure (its column-major values are its own, below)
        p.B = p.A;
    } else {
        convert(p.B, B, false, ws, s, earlyB);
    }
    // tile-CSC index of B: transpose of its tile pattern.  For C = A*A the
    // pattern is first taken as symmetric (its transpose is itself, with the
    // partner tiles' indices), which the C tile count's round trip confirms;
    // otherwise, or when B is not A, a sort of the tiles by column.
    TileMat &TB = p.B;
    TB.csc_row = ws.alloc_n<int>((size_t)TB.numblk + 1);
    TB.csc_tile = ws.alloc_n<int>((size_t)TB.numblk + 1);
    const bool try_sym = b_is_a && TB.blkm == TB.blkn && TB.numblk > 0;
    unsigned *d_asym = nullptr;
    auto transpose = [&] {
        TB.csc_ptr = ws.alloc_n<int>((size_t)TB.blkn + 1);
        csr_transpose(TB.blkm, TB.blkn, TB.tile_ptr, TB.tile_col, nullptr, TB.numblk,
                      TB.csc_ptr, TB.csc_row, nullptr, ws, s, TB.csc_tile);
    };
    if (try_sym) {
        d_asym = ws.alloc_n<unsigned>(1);
        TSG_HIP_TRY(hipMemsetAsync(d_asym, 0, sizeof(unsigned), s));
        TB.csc_ptr = TB.tile_ptr;
        hipLaunchKernelGGL(k_sym_csc, dim3(grid_for((size_t)TB.blkm * 16, 256)), dim3(256), 0, s,
                           TB.blkm, TB.tile_ptr, TB.tile_col, TB.csc_row, TB.csc_tile, d_asym);
        TSG_LAUNCH_CHECK();
    } else {
        transpose();
    }
    // B's column layout: column masks and starts, and every entry's position
    // in the column-major payload the numeric phase reads
    TB.rec = ws.alloc_n<TileRec>((size_t)TB.numblk + 1);
    TB.permc = ws.alloc_n<int>((size_t)TB.nnz + 1);
    TB.valc = ws.alloc_n<double>((size_t)TB.nnz + 1);
    if (TB.numblk > 0) {
        int *cpos = ws.alloc_n<int>((size_t)TB.nnz + 1);
        hipLaunchKernelGGL(k_col_layout, dim3(grid_for((size_t)TB.numblk * 16, 256)), dim3(256),
                           0, s, TB.numblk, TB.tile_off, TB.mask, TB.rowptr, TB.rec, cpos);
        TSG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_compose_perm, dim3(grid_for((size_t)TB.nnz, 256)), dim3(256), 0, s,
                           TB.nnz, TB.perm, cpos, TB.permc);
        TSG_LAUNCH_CHECK();
    }
    // A's row layout records
    p.A.rec = ws.alloc_n<TileRec>((size_t)p.A.numblk + 1);
    if (p.A.numblk > 0) {
        hipLaunchKernelGGL(k_row_records, dim3(grid_for((size_t)p.A.numblk * 4, 256)), dim3(256),
                           0, s, p.A.numblk, p.A.tile_off, p.A.mask, p.A.rowptr, p.A.rec);
        TSG_LAUNCH_CHECK();
    }
    tm.end();

    const int blkm = p.A.blkm;
    tm.begin("tile.step1");
    int *cnt = ws.alloc_n<int>((size_t)blkm + 1);
    int64_t *tmp = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)blkm + 1));
    p.tptrC = ws.alloc_n<int>((size_t)blkm + 1);
    if (blkm > 0) {
        hipLaunchKernelGGL(k_step1<false>, dim3(grid_for(blkm, S1_WAVES)), dim3(64 * S1_WAVES), 0,
                           s, blkm, p.A.tile_ptr, p.A.tile_col, TB.tile_ptr, TB.tile_col, cnt,
                           nullptr, nullptr, nullptr);
        TSG_LAUNCH_CHECK();
    }
    // C tiles are counted in int64: the tile path indexes them with int32,
    // so a product with 2^31 or more C tiles is refused before anything is
    // sized from the count
    int64_t *d_nC = ws.alloc_n<int64_t>(1);
    exclusive_scan_i32(cnt, p.tptrC, (size_t)blkm, tmp, s, d_nC);
    // pair slots per C tile row (its C tiles times its A tiles), so step2
    // can store the matched pairs as it finds them
    int64_t *pslot = ws.alloc_n<int64_t>((size_t)blkm + 1);
    if (blkm > 0) {
        hipLaunchKernelGGL(k_pair_slots, dim3(grid_for((size_t)blkm, 256)), dim3(256), 0, s, blkm,
                           p.A.tile_ptr, cnt, pslot);
        TSG_LAUNCH_CHECK();
    }
    exclusive_scan_i64(pslot, pslot, (size_t)blkm, tmp, s);
    int64_t nC64 = 0, nslots = 0;
    unsigned h_asym = 0;
    TSG_HIP_TRY(hipMemcpyAsync(&nC64, d_nC, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    TSG_HIP_TRY(hipMemcpyAsync(&nslots, pslot + blkm, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (d_asym)
        TSG_HIP_TRY(hipMemcpyAsync(&h_asym, d_asym, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    TSG_HIP_TRY(hipStreamSynchronize(s));
    if (h_asym) transpose();
    if (b_is_a) {
        p.A.csc_ptr = TB.csc_ptr;
        p.A.csc_row = TB.csc_row;
        p.A.csc_tile = TB.csc_tile;
    }
    check_tile_count(nC64);
    // slots when they stay within a few times the C tiles (a tile row of
    // many A tiles whose C tiles match few of them -- hypersparse inputs --
    // would make them too many); else a second pass stores the pairs
    const bool direct_pairs = nslots <= 32 * nC64 + (int64_t(1) << 20);
    const int nC = (int)nC64;
    p.numblkC = nC;
    p.tcolC = ws.alloc_n<int>((size_t)nC + 1);
    p.trowC = ws.alloc_n<int>((size_t)nC + 1);
    if (nC > 0) {
        hipLaunchKernelGGL(k_step1<true>, dim3(grid_for(blkm, S1_WAVES)), dim3(64 * S1_WAVES), 0,
                           s, blkm, p.A.tile_ptr, p.A.tile_col, TB.tile_ptr, TB.tile_col, nullptr,
                           p.tptrC, p.tcolC, p.trowC);
        TSG_LAUNCH_CHECK();
    }
    tm.end();

    tm.begin("tile.step2");
    p.maskC = ws.alloc_n<uint16_t>((size_t)nC * 16 + 16);
    p.rowptrC = ws.alloc_n<uint8_t>((size_t)nC * 16 + 16);
    p.offC = ws.alloc_n<int64_t>((size_t)nC + 1);
    p.rowoffC = ws.alloc_n<int>((size_t)nC * 16 + 16);
    p.tpair = ws.alloc_n<int64_t>((size_t)nC + 1);
    p.tslot = direct_pairs ? ws.alloc_n<int64_t>((size_t)nC + 1) : nullptr;
    if (direct_pairs) {
        p.npairs = nslots;
        p.pairs = ws.alloc_n<int2>((size_t)nslots + 1);
    }
    p.rownnz = ws.alloc_n<int64_t>((size_t)m + 1);
    TSG_HIP_TRY(hipMemsetAsync(p.rownnz, 0, sizeof(int64_t) * ((size_t)m + 1), s));
    if (blkm > 0) {
        hipLaunchKernelGGL(k_step2, dim3(grid_for(blkm, S2_WAVES)), dim3(64 * S2_WAVES), 0, s, m,
                           blkm, p.A.tile_ptr, p.A.tile_col, p.A.mask, TB.csc_ptr, TB.csc_row,
                           TB.csc_tile, TB.mask, p.tptrC, p.tcolC, p.maskC, p.rowptrC, p.offC,
                           p.rowoffC, p.rownnz, p.tpair, direct_pairs ? pslot : nullptr, p.tslot,
                           p.pairs);
        TSG_LAUNCH_CHECK();
    }
    int64_t *tmpC = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)nC + 1));
    exclusive_scan_i64(p.offC, p.offC, (size_t)nC, tmpC, s);
    if (!direct_pairs) exclusive_scan_i64(p.tpair, p.tpair, (size_t)nC, tmpC, s);
    int64_t *tmpm = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)m + 1));
    exclusive_scan_i64(p.rownnz, d_ic, (size_t)m, tmpm, s);
    // C tiles by output count for the numeric kernels
    int *bcnt = ws.alloc_n<int>(2 * TBINS);
    TSG_HIP_TRY(hipMemsetAsync(bcnt, 0, 2 * TBINS * sizeof(int), s));
    if (nC > 0) {
        hipLaunchKernelGGL(k_tbin_count, dim3(grid_for((size_t)nC, 256)), dim3(256), 0, s, nC,
                           p.offC, bcnt);
        TSG_LAUNCH_CHECK();
    }
    tm.end();
    // nnz(C), the matched-pair total and the bin counts in one round trip
    int64_t h64[2] = {0, 0};
    int hb[TBINS] = {0, 0, 0};
    TSG_HIP_TRY(hipMemcpyAsync(&h64[0], d_ic + m, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (!direct_pairs)
        TSG_HIP_TRY(hipMemcpyAsync(&h64[1], p.tpair + nC, sizeof(int64_t), hipMemcpyDeviceToHost,
                                   s));
    TSG_HIP_TRY(hipMemcpyAsync(hb, bcnt, sizeof(hb), hipMemcpyDeviceToHost, s));
    TSG_HIP_TRY(hipStreamSynchronize(s));
    // the numeric phase's work units, sized by the counts just read
    int nunits = 0;
    for (int b = 0; b < TBINS; b++) nunits += hb[b];
    p.tbin_units = ws.alloc_n<int4>((size_t)nunits + 1);
    tm.begin("tile.pairs");
    if (nunits > 0) {
        hipLaunchKernelGGL(k_tbin_scatter, dim3(grid_for((size_t)nC, 256)), dim3(256), 0, s, nC,
                           p.offC, p.tpair, p.tslot, bcnt, bcnt + TBINS, p.tbin_units);
        TSG_LAUNCH_CHECK();
    }
    // the matched pairs themselves, unless step2 stored them (the numeric
    // phase reads them instead of intersecting tile lists again)
    if (!direct_pairs) {
        p.npairs = h64[1];
        p.pairs = ws.alloc_n<int2>((size_t)p.npairs + 1);
    }
    if (!direct_pairs && blkm > 0 && p.npairs > 0) {
        hipLaunchKernelGGL(k_fill_pairs, dim3(grid_for(blkm, FP_WAVES)), dim3(64 * FP_WAVES), 0,
                           s, blkm, p.A.tile_ptr, p.A.tile_col, TB.csc_ptr, TB.csc_row,
                           TB.csc_tile, p.tptrC, p.tcolC, p.tpair, p.pairs);
        TSG_LAUNCH_CHECK();
    }
    tm.end();
    int start = 0;
    for (int b = 0; b < TBINS; b++) {
        p.tbin_start[b] = start;
        start += hb[b];
    }
    p.tbin_start[TBINS] = start;
    p.nnzC = h64[0];
    *nnz_c = h64[0];
    p.valid = true;
}

template <bool FUSED>
static void tile_numeric_t(TilePlan &p, const double *valB, const int64_t *d_ic, int *d_jc,
                           double *d_c, hipStream_t s) {
    const TileMat &TA = p.A, &TB = p.B;
    const TileArgs T{p.trowC, p.tcolC, p.maskC, p.rowptrC, p.rowoffC, p.pairs, TA.rec,
                     TA.val,  TB.rec,  valB,    d_ic,      d_jc,      d_c};
    // four waves per block; units per wave by bin
    const int tpw[TBINS] = {4, 2, 1};
    for (int b = 0; b < TBINS; b++) {
        const int n = p.tbin_start[b + 1] - p.tbin_start[b];
        if (n == 0) continue;
        const int4 *units = p.tbin_units + p.tbin_start[b];
        const dim3 grid(grid_for((size_t)n, 4 * tpw[b] * UPW));
        if (b == 0)
            hipLaunchKernelGGL((k_step3_out<FUSED, 4>), grid, dim3(256), 0, s, n, units, T);
        else if (b == 1)
            hipLaunchKernelGGL((k_step3_out<FUSED, 2>), grid, dim3(256), 0, s, n, units, T);
        else
            hipLaunchKernelGGL((k_step3_out<FUSED, 1>), grid, dim3(256), 0, s, n, units, T);
        TSG_LAUNCH_CHECK();
    }
}

void tile_numeric(TilePlan &p, const CsrIn &A, const CsrIn &B, const int64_t *d_ic, int *d_jc,
                  double *d_c, int fused, Arena &ws, hipStream_t s, StageTimer &tm, int cus) {
    (void)ws;
    (void)cus;
    TSG_CHECK(p.valid && p.A.rows == A.rows && p.A.cols == A.cols && p.B.cols == B.cols,
              TSG_ERR_STATE, "tile numeric phase without a matching symbolic phase");
    tm.begin("tile.values");
    // B's values go into its column-major payload (also when B is A: A's
    // payload is row-major); when B is A with the same values, one pass
    // writes both
    const bool same_vals = p.b_is_a && A.val == B.val && A.nnz == B.nnz;
    if (same_vals && A.nnz > 0) {
        hipLaunchKernelGGL(k_perm_scatter2, dim3(grid_for((size_t)A.nnz, 256)), dim3(256), 0, s,
                           A.nnz, p.A.perm, p.B.permc, A.val, p.A.val, p.B.valc);
        TSG_LAUNCH_CHECK();
    } else if (A.nnz > 0) {
        hipLaunchKernelGGL(k_perm_scatter, dim3(grid_for((size_t)A.nnz, 256)), dim3(256), 0, s,
                           A.nnz, p.A.perm, A.val, p.A.val);
        TSG_LAUNCH_CHECK();
    }
    if (!same_vals && B.nnz > 0) {
        hipLaunchKernelGGL(k_perm_scatter, dim3(grid_for((size_t)B.nnz, 256)), dim3(256), 0, s,
                           B.nnz, p.B.permc, B.val, p.B.valc);
        TSG_LAUNCH_CHECK();
    }
    const double *valB = p.B.valc;
    tm.end();
    tm.begin("tile.step3");
    if (fused)
        tile_numeric_t<true>(p, valB, d_ic, d_jc, d_c, s);
    else
        tile_numeric_t<false>(p, valB, d_ic, d_jc, d_c, s);
    tm.end();
}

// Entries per non-empty 16x16 tile of X, used to choose between the tile
// path and the row path.  A deterministic sample of <= 1024 tile rows
// decides unless it lands near the threshold, where every tile row is
// counted.  tile_sample_launch enqueues the sample's count (sums = {tiles,
// entries}, zeroed by the caller) so that it can share a host round trip
// with other work; tile_density takes its result, or counts itself.
static int tile_sample_step(int blkm) { return blkm > 1024 ? (blkm + 1023) / 1024 : 1; }

static void tile_count_launch(const CsrIn &X, int step, unsigned long long *sums,
                              hipStream_t s) {
    const int blkm = (X.rows + 15) / 16;
    // one block per sampled tile row: block b takes tile row b*step < blkm
    const int nrows = (blkm + step - 1) / step;
    if (nrows == 0) return;
    hipLaunchKernelGGL(k_tile_sample, dim3((unsigned)nrows), dim3(256), 0, s, X.rows, X.cols,
                       X.nnz, step, X.ptr, X.idx, sums);
    TSG_LAUNCH_CHECK();
}

void tile_sample_launch(const CsrIn &X, unsigned long long *sums, hipStream_t s) {
    const int blkm = (X.rows + 15) / 16;
    if (blkm == 0 || X.nnz == 0) return;
    tile_count_launch(X, tile_sample_step(blkm), sums, s);
}

double tile_density(const CsrIn &X, double threshold, Arena &ws, hipStream_t s,
                    const unsigned long long *sample) {
    const int blkm = (X.rows + 15) / 16;
    if (blkm == 0 || X.nnz == 0) return 0.0;
    unsigned long long *sums = ws.alloc_n<unsigned long long>(2);
    auto ratio = [](const unsigned long long *h) {
        return h[0] > 0 ? (double)h[1] / (double)h[0] : 0.0;
    };
    auto count = [&](int step) {
        TSG_HIP_TRY(hipMemsetAsync(sums, 0, 2 * sizeof(unsigned long long), s));
        tile_count_launch(X, step, sums, s);
        unsigned long long h[2] = {0, 0};
        TSG_HIP_TRY(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipStreamSynchronize(s));
        return ratio(h);
    };
    const int step = tile_sample_step(blkm);
    const double est = sample ? ratio(sample) : count(step);
    if (step == 1 || est < 0.75 * threshold || est > 1.33 * threshold) return est;
    return count(1);
}

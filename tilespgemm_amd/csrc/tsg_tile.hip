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
//            position inside its CSR row, so step3 writes CSR directly.
//   step3    FP64 numeric [step4, main_spgemm-col.cu:2094-2397,3386-3754]:
//            lane r of a tile walks A_IK row r and the selected B_KJ rows in
//            ascending K, accumulating a padded 16x17 LDS tile with one FMA
//            per product (no atomics; products of every output in A-row
//            order), then writes its row's entries straight into C's CSR.
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
// One 16-lane group per tile row; lane r handles CSR row 16*I + r.
template <bool FILL>
__global__ void __launch_bounds__(256) k_tile_convert(
    int m, int blkm, const int *__restrict__ ia, const int *__restrict__ ja,
    int *__restrict__ tcount, const int *__restrict__ tile_ptr, int *__restrict__ tile_col,
    int *__restrict__ tile_off, uint16_t *__restrict__ mask, uint8_t *__restrict__ rowptr,
    uint8_t *__restrict__ idx, int *__restrict__ perm) {
    const int I = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int r = threadIdx.x & 15;
    if (I >= blkm) return;
    const int row = I * 16 + r;
    int cur = 0, end = 0;
    if (row < m) { cur = ia[row]; end = ia[row + 1]; }
    int t = FILL ? tile_ptr[I] : 0;
    int off = FILL ? ia[I * 16] : 0;
    int ntiles = 0;
    while (true) {
        const int Jr = cur < end ? (ja[cur] >> 4) : INT_MAX;
        const int J = group_min<16>(Jr);
        if (J == INT_MAX) break;
        const int c0 = cur;
        unsigned mk = 0;
        while (cur < end) {
            const int col = ja[cur];
            if ((col >> 4) != J) break;
            mk |= 1u << (15 - (col & 15));
            cur++;
        }
        if (FILL) {
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
                idx[q] = (uint8_t)((r << 4) | (ja[c0 + i] & 15));
                perm[c0 + i] = q;
            }
            off += tot;
            t++;
        }
        ntiles++;
    }
    if (!FILL && r == 0) tcount[I] = ntiles;
    if (FILL && r == 0 && I == blkm - 1) tile_off[t] = off;
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
                                               atomicOr(&bits[c >> 5], 1u << (c & 31));
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
template <class F>
__device__ __forceinline__ void for_pairs(bool valid, int pa0, int pa1, int cb0, int cb1,
                                          const int *__restrict__ tcolA,
                                          const int *__restrict__ csc_row,
                                          const int *__restrict__ csc_tile, F &&f) {
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
    const int lenA = valid ? pa1 - pa0 : 0, lenB = valid ? cb1 - cb0 : 0;
    const bool iterA = lenA <= lenB;
    const int len = iterA ? lenA : lenB;
    for (int i0 = 0; i0 < len; i0 += 16) {
        const int i = i0 + r;
        bool match = false;
        int pa = 0, pb = 0;
        if (i < len) {
            if (iterA) {
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

__global__ void __launch_bounds__(64 * S2_WAVES) k_step2(
    int m, int blkm, const int *__restrict__ tptrA, const int *__restrict__ tcolA,
    const uint16_t *__restrict__ maskA, const int *__restrict__ csc_ptr,
    const int *__restrict__ csc_row, const int *__restrict__ csc_tile,
    const uint16_t *__restrict__ maskB, const int *__restrict__ tptrC,
    const int *__restrict__ tcolC, uint16_t *__restrict__ maskC, uint8_t *__restrict__ rowptrC,
    int64_t *__restrict__ tnnz, int *__restrict__ rowoffC, int64_t *__restrict__ rownnz) {
    const int I = blockIdx.x * S2_WAVES + (threadIdx.x >> 6);
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
    if (I >= blkm) return;
    const int pa0 = tptrA[I], pa1 = tptrA[I + 1];
    const int t0 = tptrC[I], t1 = tptrC[I + 1];
    int running = 0;
    for (int base = t0; base < t1; base += 4) {
        const int t = base + g;
        const bool valid = t < t1;
        const int J = valid ? tcolC[t] : 0;
        const int cb0 = valid ? csc_ptr[J] : 0, cb1 = valid ? csc_ptr[J + 1] : 0;
        unsigned mc = 0;
        for_pairs(valid, pa0, pa1, cb0, cb1, tcolA, csc_row, csc_tile, [&](int pa, int pb) {
            unsigned ma = maskA[(size_t)pa * 16 + r];
            const uint16_t *mb = maskB + (size_t)pb * 16;
            while (ma) {
                const int bit = 31 - __clz(ma);   // highest set bit = lowest column
                ma &= ~(1u << bit);
                mc |= mb[15 - bit];
            }
        });
        const int pc = __popc(mc);
        const int incl = group_incl_scan<16>(pc);
        if (valid) {
            maskC[(size_t)t * 16 + r] = (uint16_t)mc;
            rowptrC[(size_t)t * 16 + r] = (uint8_t)(incl - pc);
            if (r == 15) tnnz[t] = incl;
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

// ------------------------------------------------------------------ step3
constexpr int S3_WAVES = 4;
constexpr int S3_TILES = S3_WAVES * 4;
constexpr int ACC_LD = 17;               // padded row stride (doubles)

template <bool FUSED>
__global__ void __launch_bounds__(64 * S3_WAVES) k_step3(
    int numblkC, const int *__restrict__ trowC, const int *__restrict__ tcolC,
    const int *__restrict__ tptrA, const int *__restrict__ tcolA,
    const int *__restrict__ offA, const uint8_t *__restrict__ rowptrA,
    const uint8_t *__restrict__ idxA, const double *__restrict__ valA,
    const int *__restrict__ csc_ptr, const int *__restrict__ csc_row,
    const int *__restrict__ csc_tile, const int *__restrict__ offB,
    const uint8_t *__restrict__ rowptrB, const uint8_t *__restrict__ idxB,
    const double *__restrict__ valB, const uint16_t *__restrict__ maskC,
    const int *__restrict__ rowoffC, const int64_t *__restrict__ ic, int *__restrict__ jc,
    double *__restrict__ c) {
    __shared__ double s_acc[S3_TILES * 16 * ACC_LD];
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
    const int slot = (threadIdx.x >> 6) * 4 + g;
    double *acc = s_acc + (slot * 16 + r) * ACC_LD;
    const int t = blockIdx.x * S3_TILES + slot;
    const bool valid = t < numblkC;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0;
    int I = 0, J = 0, cb0 = 0, cb1 = 0, pa0 = 0, pa1 = 0;
    if (valid) {
        I = trowC[t];
        J = tcolC[t];
        cb0 = csc_ptr[J];
        cb1 = csc_ptr[J + 1];
        pa0 = tptrA[I];
        pa1 = tptrA[I + 1];
    }
    for_pairs(valid, pa0, pa1, cb0, cb1, tcolA, csc_row, csc_tile, [&](int pa, int pb) {
        const int abase = offA[pa];
        const int alo = abase + rowptrA[(size_t)pa * 16 + r];
        const int ahi = r == 15 ? offA[pa + 1] : abase + rowptrA[(size_t)pa * 16 + r + 1];
        const int bbase = offB[pb];
        const int bend = offB[pb + 1];
        const uint8_t *rpb = rowptrB + (size_t)pb * 16;
        for (int q = alo; q < ahi; q++) {
            const int cc = idxA[q] & 15;
            const double av = valA[q];
            const int blo = bbase + rpb[cc];
            const int bhi = cc == 15 ? bend : bbase + rpb[cc + 1];
            for (int u = blo; u < bhi; u++) {
                const int col = idxB[u] & 15;
                acc[col] = acc_step<FUSED>(acc[col], av, valB[u]);
            }
        }
    });
    if (!valid) return;
    const unsigned mk = maskC[(size_t)t * 16 + r];
    if (!mk) return;
    const int64_t o = ic[I * 16 + r] + rowoffC[(size_t)t * 16 + r];
    int i = 0;
#pragma unroll
    for (int cc = 0; cc < 16; cc++) {
        if (mk & (1u << (15 - cc))) {
            jc[o + i] = J * 16 + cc;
            c[o + i] = acc[cc];
            i++;
        }
    }
}

void convert(TileMat &T, const CsrIn &X, bool fill_values, Arena &ws, hipStream_t s) {
    T.rows = X.rows;
    T.cols = X.cols;
    T.blkm = (X.rows + 15) / 16;
    T.blkn = (X.cols + 15) / 16;
    T.nnz = X.nnz;
    int *tcount = ws.alloc_n<int>((size_t)T.blkm + 1);
    int64_t *tmp = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)T.blkm + 1));
    T.tile_ptr = ws.alloc_n<int>((size_t)T.blkm + 1);
    if (T.blkm > 0) {
        hipLaunchKernelGGL(k_tile_count, dim3(grid_for(T.blkm, TC_WAVES)), dim3(64 * TC_WAVES), 0,
                           s, X.rows, X.cols, X.nnz, T.blkm, X.ptr, X.idx, tcount);
        TSG_LAUNCH_CHECK();
    }
    exclusive_scan_i32(tcount, T.tile_ptr, (size_t)T.blkm, tmp, s);
    int nt = 0;
    TSG_HIP_TRY(hipMemcpyAsync(&nt, T.tile_ptr + T.blkm, sizeof(int), hipMemcpyDeviceToHost, s));
    TSG_HIP_TRY(hipStreamSynchronize(s));
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
        hipLaunchKernelGGL(k_tile_convert<true>, dim3(grid_for((size_t)T.blkm * 16, 256)),
                           dim3(256), 0, s, X.rows, T.blkm, X.ptr, X.idx, nullptr, T.tile_ptr,
                           T.tile_col, T.tile_off, T.mask, T.rowptr, T.idx, T.perm);
        TSG_LAUNCH_CHECK();
    }
    if (fill_values && X.val && X.nnz > 0) {
        hipLaunchKernelGGL(k_perm_scatter, dim3(grid_for((size_t)X.nnz, 256)), dim3(256), 0, s,
                           X.nnz, T.perm, X.val, T.val);
        TSG_LAUNCH_CHECK();
    }
}

}  // namespace

void check_tile_count(int64_t n_c_tiles) {
    TSG_CHECK(n_c_tiles >= 0 && n_c_tiles <= INT_MAX, TSG_ERR_OVERFLOW,
              "the tile path indexes C tiles with int32: " + std::to_string(n_c_tiles) +
                  " C tiles exceed 2^31-1 (use the row path)");
}

void tile_symbolic(TilePlan &p, const CsrIn &A, const CsrIn &B, bool b_is_a, int64_t *d_ic,
                   int64_t *nnz_c, Arena &ws, hipStream_t s, StageTimer &tm, int cus) {
    (void)cus;
    p = TilePlan();
    p.b_is_a = b_is_a;
    const int m = A.rows;
    tm.begin("tile.convert");
    convert(p.A, A, false, ws, s);
    if (b_is_a) {
        // B shares A's structure; its values get a buffer of their own, used
        // when a numeric phase passes B values other than A's
        p.B = p.A;
        p.B.val = ws.alloc_n<double>((size_t)B.nnz + 1);
    } else {
        convert(p.B, B, false, ws, s);
    }
    // tile-CSC index of B: transpose of its tile pattern
    TileMat &TB = p.B;
    TB.csc_ptr = ws.alloc_n<int>((size_t)TB.blkn + 1);
    TB.csc_row = ws.alloc_n<int>((size_t)TB.numblk + 1);
    TB.csc_tile = ws.alloc_n<int>((size_t)TB.numblk + 1);
    csr_transpose(TB.blkm, TB.blkn, TB.tile_ptr, TB.tile_col, nullptr, TB.numblk, TB.csc_ptr,
                  TB.csc_row, nullptr, ws, s, TB.csc_tile);
    if (b_is_a) {
        p.A.csc_ptr = TB.csc_ptr;
        p.A.csc_row = TB.csc_row;
        p.A.csc_tile = TB.csc_tile;
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
    int64_t nC64 = 0;
    TSG_HIP_TRY(hipMemcpyAsync(&nC64, d_nC, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    TSG_HIP_TRY(hipStreamSynchronize(s));
    check_tile_count(nC64);
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
    p.rownnz = ws.alloc_n<int64_t>((size_t)m + 1);
    TSG_HIP_TRY(hipMemsetAsync(p.rownnz, 0, sizeof(int64_t) * ((size_t)m + 1), s));
    if (blkm > 0) {
        hipLaunchKernelGGL(k_step2, dim3(grid_for(blkm, S2_WAVES)), dim3(64 * S2_WAVES), 0, s, m,
                           blkm, p.A.tile_ptr, p.A.tile_col, p.A.mask, TB.csc_ptr, TB.csc_row,
                           TB.csc_tile, TB.mask, p.tptrC, p.tcolC, p.maskC, p.rowptrC, p.offC,
                           p.rowoffC, p.rownnz);
        TSG_LAUNCH_CHECK();
    }
    int64_t *tmpC = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)nC + 1));
    exclusive_scan_i64(p.offC, p.offC, (size_t)nC, tmpC, s);
    int64_t *tmpm = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)m + 1));
    exclusive_scan_i64(p.rownnz, d_ic, (size_t)m, tmpm, s);
    tm.end();
    int64_t tot = 0;
    TSG_HIP_TRY(hipMemcpyAsync(&tot, d_ic + m, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    TSG_HIP_TRY(hipStreamSynchronize(s));
    p.nnzC = tot;
    *nnz_c = tot;
    p.valid = true;
}

template <bool FUSED>
static void tile_numeric_t(TilePlan &p, const double *valB, const int64_t *d_ic, int *d_jc,
                           double *d_c, hipStream_t s) {
    const TileMat &TA = p.A, &TB = p.B;
    if (p.numblkC > 0) {
        hipLaunchKernelGGL(k_step3<FUSED>, dim3(grid_for(p.numblkC, S3_TILES)),
                           dim3(64 * S3_WAVES), 0, s, p.numblkC, p.trowC, p.tcolC, TA.tile_ptr,
                           TA.tile_col, TA.tile_off, TA.rowptr, TA.idx, TA.val, TB.csc_ptr,
                           TB.csc_row, TB.csc_tile, TB.tile_off, TB.rowptr, TB.idx, valB,
                           p.maskC, p.rowoffC, d_ic, d_jc, d_c);
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
    if (A.nnz > 0) {
        hipLaunchKernelGGL(k_perm_scatter, dim3(grid_for((size_t)A.nnz, 256)), dim3(256), 0, s,
                           A.nnz, p.A.perm, A.val, p.A.val);
        TSG_LAUNCH_CHECK();
    }
    // B's values: A's tile payload when B is A (same structure, same value
    // array), else scattered into B's own payload (with A's structure, by
    // A's permutation)
    p.b_val_shared = p.b_is_a && B.val == A.val;
    if (!p.b_val_shared && B.nnz > 0) {
        hipLaunchKernelGGL(k_perm_scatter, dim3(grid_for((size_t)B.nnz, 256)), dim3(256), 0, s,
                           B.nnz, p.B.perm, B.val, p.B.val);
        TSG_LAUNCH_CHECK();
    }
    const double *valB = p.b_val_shared ? p.A.val : p.B.val;
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

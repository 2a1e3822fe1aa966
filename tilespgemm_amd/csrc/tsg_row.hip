// Row path: row-binned SpGEMM for matrices whose 16x16 tiles are too sparse
// for the tile path (hypersparse / power-law inputs such as R-MAT).
//
// Lineage: the reference bins rows by intermediate products and runs per-bin
// hash kernels (nsparse, spgemm_nsparse_kernel.h:221-311 binning, :331-687
// symbolic hash, :696-1119 structure fill; its value-carrying variant is
// commented out at :805-907).  This path is built for gfx950 instead:
//
//   * Every output entry is accumulated from +0.0 over its products in A-row
//     order (thmkl_dcsrmultcsr.c:61-72), so results are bitwise equal to the
//     CPU reference.  Products are expanded flat, a wave at a time in A-row
//     order ("t-major"); lanes of one batch that hit the same accumulator slot
//     are serialized in lane order by tag rounds (LDS atomicMin on a small
//     bucket table), so the parallel sum equals the sequential one.
//   * Small rows (<= SMALL_PROD products): one lane group per row, LDS hash
//     table, output sorted by an in-LDS bitonic sort.
//   * Big rows: the symbolic phase cuts each row's column span into bands of
//     65536 columns, marks the row's columns of a band in an LDS bitmap and
//     records the count of every 1024-column sub-block (and, when memory
//     allows, the band's bitmap in sparse form: a mask of non-empty words per
//     sub-block, and those words); the numeric phase splits the row into
//     items of <= ITEM_CAP outputs and <= 16384 columns, each accumulated by
//     one wave into a rank-ordered LDS array (rank from the stored bitmap)
//     and written in column order.  A B block pointer table gives each B row's
//     sub-range of a band or item without a search; bands and items run in
//     order of their first column, so those in flight share a band of B.
#include "tsg_internal.h"

#include <cmath>
#include "tsg_device.h"
#include "tsg_expand.h"

#include <algorithm>
#include <map>
#include <mutex>

namespace {

constexpr int SMALL_PROD = 1024;       // rows up to this many products: small path
constexpr int NBINS = 5;               // 0: empty, 1..3: small, 4: big
constexpr int BIN_BIG = 4;
constexpr int bin_limit[NBINS] = {0, 64, 256, SMALL_PROD, 0x7fffffff};

constexpr int SUB_SHIFT = 10;          // sub-block = 1024 columns
constexpr int BAND_SHIFT = 6;          // symbolic band = 64 sub-blocks (65536 columns)
constexpr int BAND_T = 128;            // threads per band unit
constexpr int BAND_WORDS = (1 << (BAND_SHIFT + SUB_SHIFT)) / 32;
constexpr int BAND_PAD = 4;            // band word lists: words rounded up to this
constexpr int ITEM_CAP = 1024;         // outputs per item
constexpr int ITEM_SLOTS = 2048;       // hash slots per item (load <= 1/2)
constexpr int ITEM_WBITS_SHIFT = 14;   // item width <= 16384 columns
constexpr int ITEM_WORDS = (1 << ITEM_WBITS_SHIFT) / 32;
// items are whole sub-blocks, and one sub-block can hold 1 << SUB_SHIFT outputs
static_assert(ITEM_CAP >= (1 << SUB_SHIFT), "an item must hold a full sub-block");
constexpr int ITEM_TAGS = 256;
constexpr int ITEM_WAVES = 1;          // waves per block (one item per wave)
constexpr int RANK_U = 8;
// Product batches in flight per lane in the small-row kernels: their LDS
// tables already cap residency, so registers for more loads in flight are
// free up to 8 (48 VGPRs; 16 would cut residency by registers).  2 -> 8:
// ER batch rows of <= 1024 products 62.0 -> 57.2 ms, 27-point 64^3 3.05 ->
// 2.93 ms.
constexpr int SMALL_U = 8;
// Rows of the 64-lane bin whose column span fits 1024 bitmap words and whose
// outputs number at most 512 (meshes and stencils) run in the rank kernel's
// 1024-word variant: 10 KB of LDS per row instead of 25 KB.
constexpr int NARROW_WORDS = 1024;
constexpr int NARROW_OUT = NARROW_WORDS / 2;
constexpr int BIG_SYM_U = 4;
constexpr int ITEM_U = 4;

struct BinStarts { int s[NBINS + 1]; };

// Row bins by intermediate products: (0, l1] -> 16-lane groups with 128 hash
// slots, (l1, l2] -> 32 lanes and 512 slots, (l2, l3] -> 64 lanes and 2048
// slots, above l3 -> the big-row path.  Each small kernel is exact for any
// row up to bin_limit[b] products (hash load <= 1/2), so any limits with
// l1 <= 64, l2 <= 256, l3 <= 1024 (checked on the host) are valid; they are
// tuning parameters (tsg_set_row_bins).
struct BinLimits { int l1, l2, l3; };

__device__ __forceinline__ int bin_of(int64_t prod, BinLimits lim) {
    if (prod == 0) return 0;
    if (prod <= lim.l1) return 1;
    if (prod <= lim.l2) return 2;
    if (prod <= lim.l3) return 3;
    return BIN_BIG;
}

// ------------------------------------------------------------------ binning
__global__ void k_bin_count(int m, const int64_t *__restrict__ rowprod, BinLimits lim,
                            int *__restrict__ cnt) {
    __shared__ int s_cnt[NBINS];
    if (threadIdx.x < NBINS) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) atomicAdd(&s_cnt[bin_of(rowprod[i], lim)], 1);
    __syncthreads();
    if (threadIdx.x < NBINS && s_cnt[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], s_cnt[threadIdx.x]);
}

// Rows go to their bin's list in any order: per wave, one ballot per bin
// gives each lane its slot among the wave's rows of that bin; per block, one
// LDS atomic per wave and bin; per grid, one global atomic per block and bin.
__global__ void __launch_bounds__(256) k_bin_scatter(int m, const int64_t *__restrict__ rowprod,
                                                     BinLimits lim, BinStarts st,
                                                     int *__restrict__ fill,
                                                     int *__restrict__ bin_rows,
                                                     int64_t *__restrict__ rownnz) {
    __shared__ int s_cnt[NBINS], s_base[NBINS];
    if (threadIdx.x < NBINS) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int b = i < m ? bin_of(rowprod[i], lim) : -1;
    if (b == 0 || b == BIN_BIG) rownnz[i] = 0;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    int pos = 0;
    for (int q = 1; q < NBINS; q++) {
        const uint64_t mask = __ballot(b == q);
        if (!mask) continue;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int wbase = 0;
        if (lane == leader) wbase = atomicAdd(&s_cnt[q], __popcll(mask));
        wbase = __shfl(wbase, leader);
        if (b == q) pos = wbase + __popcll(mask & below);
    }
    __syncthreads();
    if (threadIdx.x > 0 && threadIdx.x < NBINS && s_cnt[threadIdx.x])
        s_base[threadIdx.x] = atomicAdd(&fill[threadIdx.x], s_cnt[threadIdx.x]);
    __syncthreads();
    if (b > 0) bin_rows[st.s[b] + s_base[b] + pos] = i;
}

// Rows of the <= 1024-product bin split by the class the symbolic phase gave
// them (see k_small_symbolic): class-1 rows to the front of out, class-0
// rows to its back, class-2 rows to the front of out2, class-3 rows to its
// back (any order inside each part, as rows are independent); count[c]
// (zero on entry) ends as class c's size.
__global__ void __launch_bounds__(256) k_split_bin3(int n, const int *__restrict__ rows,
                                                    const uint8_t *__restrict__ cls_of,
                                                    int *__restrict__ out, int *__restrict__ out2,
                                                    int *__restrict__ count) {
    __shared__ int s_cnt[4], s_base[4];
    if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int row = i < n ? rows[i] : -1;
    const int cls = i < n ? (int)cls_of[row] : -1;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    int pos = 0;
    for (int q = 0; q < 4; q++) {
        const uint64_t mask = __ballot(cls == q);
        if (!mask) continue;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int wbase = 0;
        if (lane == leader) wbase = atomicAdd(&s_cnt[q], __popcll(mask));
        wbase = __shfl(wbase, leader);
        if (cls == q) pos = wbase + __popcll(mask & below);
    }
    __syncthreads();
    if (threadIdx.x < 4 && s_cnt[threadIdx.x])
        s_base[threadIdx.x] = atomicAdd(&count[threadIdx.x], s_cnt[threadIdx.x]);
    __syncthreads();
    if (cls == 1) out[s_base[1] + pos] = row;
    else if (cls == 0) out[n - 1 - (s_base[0] + pos)] = row;
    else if (cls == 2) out2[s_base[2] + pos] = row;
    else if (cls == 3) out2[n - 1 - (s_base[3] + pos)] = row;
}

// Column span [lo, hi] of a row's products, from the first and last column
// of each B row its A entries name (lo > hi for a row without products); G
// lanes of a group.
template <int G>
__device__ __forceinline__ void row_span(int pa, int pe, const int *__restrict__ ja,
                                         const int *__restrict__ ib,
                                         const int *__restrict__ jb, int &lo, int &hi) {
    const int l = __lane_id() & (G - 1);
    lo = 0x7fffffff;
    hi = -1;
    for (int p = pa + l; p < pe; p += G) {
        const int t = ja[p];
        const int b0 = ib[t], b1 = ib[t + 1];
        if (b0 < b1) {
            lo = min(lo, jb[b0]);
            hi = max(hi, jb[b1 - 1]);
        }
    }
#pragma unroll
    for (int d = G / 2; d > 0; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, G));
        hi = max(hi, __shfl_xor(hi, d, G));
    }
}

// ------------------------------------------------------ small rows: symbolic
// Output count of each small row (rownnz), G lanes per row: a bitmap over
// the row's column span when it fits SLOTS words, else an LDS hash of SLOTS
// keys.  With narrow (the <= 1024-product bin), each row's numeric class: 1
// when its span fits NARROW_WORDS words and its outputs NARROW_OUT, 3 when
// its span fits SLOTS words and its outputs NARROW_OUT, 2 when its span fits
// SLOTS words, 0 otherwise.
template <int G, int SLOTS>
__global__ void __launch_bounds__(256) k_small_symbolic(int nrows, const int *__restrict__ rows,
                                                        const int *__restrict__ ia,
                                                        const int *__restrict__ ja,
                                                        const int *__restrict__ ib,
                                                        const int *__restrict__ jb,
                                                        int64_t *__restrict__ rownnz,
                                                        int2 *__restrict__ span,
                                                        uint8_t *__restrict__ narrow) {
    constexpr int RPB = 256 / G;
    __shared__ __align__(16) int s_keys[RPB * SLOTS];
    __shared__ int s_mark[256];
    const int grp = threadIdx.x / G;
    const int l = threadIdx.x & (G - 1);
    int *keys = s_keys + grp * SLOTS;
    const int r = blockIdx.x * RPB + grp;
    s_mark[threadIdx.x] = -1;
    wave_lds_fence();
    if (r >= nrows) return;
    const int row = rows[r];
    int lo, hi;
    row_span<G>(ia[row], ia[row + 1], ja, ib, jb, lo, hi);
    if (l == 0) span[row] = make_int2(lo, hi);
    const int nw = hi >= lo ? ((hi - lo) >> 5) + 1 : 0;
    if (nw <= SLOTS) {
        // the row's columns fit a bitmap over its span in the table's space:
        // mark them and count the bits (no hashing)
        unsigned *bmp = reinterpret_cast<unsigned *>(keys);
        uint4 *bmp4 = reinterpret_cast<uint4 *>(keys);
        const int nw4 = (nw + 3) >> 2;
        for (int i = l; i < nw4; i += G) bmp4[i] = make_uint4(0u, 0u, 0u, 0u);
        wave_lds_fence();
        for_products<G, SMALL_U, false, false>(ia[row], ia[row + 1], 0, G, ja, nullptr, ib, jb,
                                               nullptr, BlockPtrs(), 0, 0, s_mark + grp * G,
                                               [&](bool act, int col, double, double, bool) {
                                                   if (act) {
                                                       const int cc = col - lo;
                                                       lds_set_bit(&bmp[cc >> 5], 1u << (cc & 31));
                                                   }
                                               });
        wave_lds_fence();
        int cnt = 0;
        for (int i = l; i < nw4; i += G) {
            const uint4 wv = bmp4[i];
            cnt += __popc(wv.x) + __popc(wv.y) + __popc(wv.z) + __popc(wv.w);
        }
        cnt = group_sum<G>(cnt);
        if (l == 0) {
            rownnz[row] = cnt;
            if (narrow)
                narrow[row] = cnt <= NARROW_OUT ? (nw <= NARROW_WORDS ? 1 : 3) : 2;
        }
        return;
    }
    for (int i = l; i < SLOTS; i += G) keys[i] = -1;
    wave_lds_fence();
    int created_cnt = 0;
    for_products<G, SMALL_U, false, false>(ia[row], ia[row + 1], 0, G, ja, nullptr, ib, jb,
                                           nullptr, BlockPtrs(), 0, 0, s_mark + grp * G,
                                           [&](bool act, int col, double, double, bool) {
                                      if (act) {
                                          bool created;
                                          hash_insert<SLOTS>(keys, col, created);
                                          created_cnt += created;
                                      }
                                  });
    created_cnt = group_sum<G>(created_cnt);
    if (l == 0) {
        rownnz[row] = created_cnt;
        if (narrow) narrow[row] = 0;
    }
}

// ------------------------------------------------------- small rows: numeric
// RPB rows per block of RPB*G threads; LDS per row: SLOTS keys and values,
// then the tags of the ordered accumulation, reused as the 256-entry radix
// histogram once accumulation is over (rows of <= 64 products never need
// the radix sort: at most 64 outputs go to the bitonic sort).
template <int G, int SLOTS>
constexpr int small_tag_words() {
    return SLOTS > 128 ? (4 * G > 256 ? 4 * G : 256) : 4 * G;
}
template <int G, int SLOTS>
constexpr size_t small_numeric_row_bytes() {
    return (size_t)SLOTS * 12 + (size_t)small_tag_words<G, SLOTS>() * 4;
}
template <int G, int SLOTS, bool FUSED, int RPB>
__global__ void __launch_bounds__(RPB * G) k_small_numeric(int nrows, const int *__restrict__ rows,
                                                       const int *__restrict__ ia,
                                                       const int *__restrict__ ja,
                                                       const double *__restrict__ a,
                                                       const int *__restrict__ ib,
                                                       const int *__restrict__ jb,
                                                       const double *__restrict__ b,
                                                       const int64_t *__restrict__ ic,
                                                       int *__restrict__ jc,
                                                       double *__restrict__ c, int b_strict,
                                                       const int2 *__restrict__ span) {
    constexpr int NT = 4 * G;
    constexpr int TW = small_tag_words<G, SLOTS>();
    extern __shared__ __align__(16) unsigned char smem[];
    double *s_vals = reinterpret_cast<double *>(smem);
    int *s_keys = reinterpret_cast<int *>(s_vals + RPB * SLOTS);
    unsigned *s_tags = reinterpret_cast<unsigned *>(s_keys + RPB * SLOTS);
    const int grp = threadIdx.x / G;
    const int l = threadIdx.x & (G - 1);
    int *keys = s_keys + grp * SLOTS;
    double *vals = s_vals + grp * SLOTS;
    unsigned *tags = s_tags + grp * TW;
    unsigned *hist = tags;      // after the accumulation
    const int r = blockIdx.x * RPB + grp;
    __shared__ int s_mark[RPB * G];
    for (int i = l; i < NT; i += G) tags[i] = 0xffffffffu;
    s_mark[threadIdx.x] = -1;
    wave_lds_fence();
    if (r >= nrows) return;
    const int row = rows[r];
    {
        // the row's column span, from the symbolic phase
        const int2 sp = span[row];
        const int lo = sp.x, hi = sp.y;
        const int nw = hi >= lo ? ((hi - lo) >> 5) + 1 : 0;
        if (nw <= SLOTS) {
            // Rank path: the row's span fits a bitmap in the table's space.
            // Mark the columns, prefix-count the words, then accumulate every
            // product at its column's rank (outputs come out sorted: no hash,
            // compaction or sort).  Key space: SLOTS bitmap words; value space:
            // the words' prefix counts (first half, as ints) and the outputs
            // (second half, <= SLOTS/2 of them).
            // The bitmap loops take four words per lane (16-byte LDS accesses;
            // SLOTS is a multiple of 4, so the rounded-up word count fits).
            unsigned *bmp = reinterpret_cast<unsigned *>(keys);
            uint4 *bmp4 = reinterpret_cast<uint4 *>(keys);
            int *pfx = reinterpret_cast<int *>(vals);
            int4 *pfx4 = reinterpret_cast<int4 *>(vals);
            double *acc = vals + SLOTS / 2;
            const int64_t o = ic[row];
            const int nnz = (int)(ic[row + 1] - o);
            const int nw4 = (nw + 3) >> 2;
            for (int i = l; i < nw4; i += G) bmp4[i] = make_uint4(0u, 0u, 0u, 0u);
            wave_lds_fence();
            for_products<G, SMALL_U, false, false>(
                ia[row], ia[row + 1], 0, G, ja, nullptr, ib, jb, nullptr, BlockPtrs(), 0, 0,
                s_mark + grp * G, [&](bool act, int col, double, double, bool) {
                    if (act) {
                        const int cc = col - lo;
                        lds_set_bit(&bmp[cc >> 5], 1u << (cc & 31));
                    }
                });
            wave_lds_fence();
            int carry = 0;
            for (int q0 = 0; q0 < nw4; q0 += G) {
                const int q = q0 + l;
                const uint4 wv = q < nw4 ? bmp4[q] : make_uint4(0u, 0u, 0u, 0u);
                const int p0 = __popc(wv.x), p1 = __popc(wv.y), p2 = __popc(wv.z);
                const int pc = p0 + p1 + p2 + __popc(wv.w);
                const int incl = group_incl_scan<G>(pc);
                if (q < nw4) {
                    const int b0 = carry + incl - pc;
                    pfx4[q] = make_int4(b0, b0 + p0, b0 + p0 + p1, b0 + p0 + p1 + p2);
                }
                carry += group_last<G>(incl);
            }
            for (int i = l; i < nnz; i += G) acc[i] = 0.0;
            wave_lds_fence();
            for_products<G, SMALL_U, false, true>(
                ia[row], ia[row + 1], 0, G, ja, a, ib, jb, b, BlockPtrs(), 0, 0,
                s_mark + grp * G, [&](bool act, int col, double av, double bv, bool single) {
                    int rank = 0;
                    if (act) {
                        const int cc = col - lo;
                        const int w = cc >> 5;
                        rank = pfx[w] + __popc(bmp[w] & ((1u << (cc & 31)) - 1u));
                    }
                    ordered_accumulate<FUSED, NT>(act && rank < nnz, rank, av, bv, acc, tags,
                                                  (unsigned)l, single && b_strict);
                });
            wave_lds_fence();
            for (int i = l; i < nnz; i += G) c[o + i] = acc[i];
            for (int q = l; q < nw4; q += G) {
                const uint4 wv = bmp4[q];
                const int4 pv = pfx4[q];
                const unsigned ws[4] = {wv.x, wv.y, wv.z, wv.w};
                const int ps[4] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    unsigned word = ws[j];
                    int pos = ps[j];
                    const int cb = lo + ((4 * q + j) << 5);
                    while (word) {
                        const int bit = __ffs(word) - 1;
                        word &= word - 1;
                        if (pos < nnz) jc[o + pos] = cb + bit;
                        pos++;
                    }
                }
            }
            return;
        }
    }
    for (int i = l; i < SLOTS; i += G) { keys[i] = -1; vals[i] = 0.0; }
    wave_lds_fence();
    for_products<G, SMALL_U, false, true>(ia[row], ia[row + 1], 0, G, ja, a, ib, jb, b,
                                          BlockPtrs(), 0, 0, s_mark + grp * G,
                                          [&](bool act, int col, double av, double bv, bool single) {
                                     int slot = 0;
                                     if (act) {
                                         bool created;
                                         slot = hash_insert<SLOTS>(keys, col, created);
                                     }
                                     wave_lds_fence();
                                     ordered_accumulate<FUSED, NT>(act, slot, av, bv, vals, tags,
                                                                   (unsigned)l, single && b_strict);
                                 });
    // compact the occupied slots to the front, in slot order: chunks of G
    // slots are visited in ascending order and every entry moves to a lower
    // or equal position, so no entry is overwritten before it is read
    const int gb = __lane_id() & ~(G - 1);
    const unsigned long long gm = G == 64 ? ~0ull : ((1ull << G) - 1);
    int nocc = 0;
    for (int c0 = 0; c0 < SLOTS; c0 += G) {
        const int k = keys[c0 + l];
        const double v = vals[c0 + l];
        const bool occ = k != -1;
        const unsigned long long bal = (__ballot(occ) >> gb) & gm;
        const int before = __popcll(bal & ((1ull << l) - 1));
        wave_lds_fence();
        if (occ) {
            keys[nocc + before] = k;
            vals[nocc + before] = v;
        }
        nocc += __popcll(bal);
        wave_lds_fence();
    }
    const int64_t o = ic[row];
    const int nnz = (int)(ic[row + 1] - o);
    if (nocc <= 64) {
        int n2 = 1;
        while (n2 < nocc) n2 <<= 1;
        for (int i = nocc + l; i < n2; i += G) keys[i] = 0x7fffffff;
        wave_lds_fence();
        bitonic_sort_group<G>(keys, vals, n2);
        for (int i = l; i < nnz; i += G) {
            jc[o + i] = keys[i];
            c[o + i] = vals[i];
        }
        return;
    }
    // radix sort on keys relative to the row's smallest column; the back
    // half of the table (occupancy <= SLOTS/2) is the ping-pong buffer
    if constexpr (SLOTS > 128) {
        int kmin = 0x7fffffff, kmax = 0;
        for (int i = l; i < nocc; i += G) {
            kmin = min(kmin, keys[i]);
            kmax = max(kmax, keys[i]);
        }
#pragma unroll
        for (int d = G / 2; d > 0; d >>= 1) {
            kmin = min(kmin, __shfl_xor(kmin, d, G));
            kmax = max(kmax, __shfl_xor(kmax, d, G));
        }
        const int span = kmax - kmin;
        const int nw = (span >> 5) + 1;              // bitmap words over [kmin, kmax]
        if (nw <= SLOTS) {
            // rank by bitmap: the free back halves of the table hold one bit per
            // column of the row's span (the value half, SLOTS words) and the
            // outputs before every pair of words (the key half); an entry's
            // position is its pair's prefix plus the bits below it
            unsigned *bmp = reinterpret_cast<unsigned *>(vals + SLOTS / 2);
            int *pfx = keys + SLOTS / 2;
            for (int i = l; i < nw; i += G) bmp[i] = 0u;
            wave_lds_fence();
            for (int i = l; i < nocc; i += G) {
                const int kk = keys[i] - kmin;
                atomicOr(&bmp[kk >> 5], 1u << (kk & 31));
            }
            wave_lds_fence();
            const int npair = (nw + 1) >> 1;
            int carry = 0;
            for (int p0 = 0; p0 < npair; p0 += G) {
                const int p = p0 + l;
                int pc = 0;
                if (p < npair) {
                    pc = __popc(bmp[2 * p]);
                    if (2 * p + 1 < nw) pc += __popc(bmp[2 * p + 1]);
                }
                const int incl = group_incl_scan<G>(pc);
                if (p < npair) pfx[p] = carry + incl - pc;
                carry += group_last<G>(incl);
            }
            wave_lds_fence();
            for (int i = l; i < nocc; i += G) {
                const int kk = keys[i] - kmin;
                const int w = kk >> 5;
                int rank = pfx[w >> 1] + __popc(bmp[w] & ((1u << (kk & 31)) - 1u));
                if (w & 1) rank += __popc(bmp[w - 1]);
                if (rank < nnz) {
                    jc[o + rank] = keys[i];
                    c[o + rank] = vals[i];
                }
            }
            return;
        }
        const int bits = span > 0 ? 32 - __clz(span) : 1;
        const bool in_dst = radix_sort_group<G>(keys, vals, keys + SLOTS / 2, vals + SLOTS / 2,
                                                nocc, kmin, bits, hist);
        const int *sk = in_dst ? keys + SLOTS / 2 : keys;
        const double *sv = in_dst ? vals + SLOTS / 2 : vals;
        for (int i = l; i < nnz; i += G) {
            jc[o + i] = sk[i];
            c[o + i] = sv[i];
        }
    }
}

// Rows of the <= 1024-product bin whose column span fits RK_WORDS bitmap
// words and whose outputs number at most RK_OUT: the rank path alone, 64
// lanes per row.  LDS: the span bitmap, one output count per 4 words, the
// row's values and the accumulation tags -- 19.7 KB for (2048, 1024)
// instead of the hash-capable kernel's 25 KB (8 rows per CU resident instead
// of 6), 15.6 KB for (2048, 512) and 10.5 KB for the narrow rows' (1024,
// 512).  Mark, prefix, accumulate at each column's rank, write.
template <bool FUSED, int RK_WORDS, int RK_OUT>
__global__ void __launch_bounds__(64) k_small_numeric_rank(
    int nrows, const int *__restrict__ rows, const int *__restrict__ ia,
    const int *__restrict__ ja, const double *__restrict__ a, const int *__restrict__ ib,
    const int *__restrict__ jb, const double *__restrict__ b, const int64_t *__restrict__ ic,
    int *__restrict__ jc, double *__restrict__ c, int b_strict, const int2 *__restrict__ span) {
    constexpr int G = 64, NT = 4 * G;
    __shared__ __align__(16) unsigned bmp[RK_WORDS];
    __shared__ int pfx[RK_WORDS / 4];
    __shared__ double acc[RK_OUT];
    __shared__ unsigned tags[NT];
    __shared__ int s_mark[G];
    const int l = threadIdx.x;
    for (int i = l; i < NT; i += G) tags[i] = 0xffffffffu;
    s_mark[l] = -1;
    wave_lds_fence();
    if ((int)blockIdx.x >= nrows) return;
    const int row = rows[blockIdx.x];
    // the row's column span, from the symbolic phase, which put the row here
    // because the span fits RK_WORDS
    const int2 sp = span[row];
    const int lo = sp.x, hi = sp.y;
    const int nw = hi >= lo ? min(((hi - lo) >> 5) + 1, RK_WORDS) : 0;
    const int nw4 = (nw + 3) >> 2;
    uint4 *bmp4 = reinterpret_cast<uint4 *>(bmp);
    const int64_t o = ic[row];
    const int64_t nrow = ic[row + 1] - o;
    const int nnz = (int)(nrow < RK_OUT ? nrow : RK_OUT);
    // columns outside [lo, lo + 32*nw) cannot occur (B rows are ascending);
    // the unsigned test keeps every LDS index in range regardless
    const unsigned span_cols = (unsigned)nw << 5;
    for (int i = l; i < nw4; i += G) bmp4[i] = make_uint4(0u, 0u, 0u, 0u);
    wave_lds_fence();
    for_products<G, SMALL_U, false, false>(ia[row], ia[row + 1], 0, G, ja, nullptr, ib, jb,
                                           nullptr, BlockPtrs(), 0, 0, s_mark,
                                           [&](bool act, int col, double, double, bool) {
                                               const int cc = col - lo;
                                               if (act && (unsigned)cc < span_cols)
                                                   lds_set_bit(&bmp[cc >> 5], 1u << (cc & 31));
                                           });
    wave_lds_fence();
    int carry = 0;
    for (int q0 = 0; q0 < nw4; q0 += G) {
        const int q = q0 + l;
        const uint4 wv = q < nw4 ? bmp4[q] : make_uint4(0u, 0u, 0u, 0u);
        const int pc = __popc(wv.x) + __popc(wv.y) + __popc(wv.z) + __popc(wv.w);
        const int incl = group_incl_scan<G>(pc);
        if (q < nw4) pfx[q] = carry + incl - pc;
        carry += group_last<G>(incl);
    }
    for (int i = l; i < nnz; i += G) acc[i] = 0.0;
    wave_lds_fence();
    for_products<G, SMALL_U, false, true>(
        ia[row], ia[row + 1], 0, G, ja, a, ib, jb, b, BlockPtrs(), 0, 0, s_mark,
        [&](bool act, int col, double av, double bv, bool single) {
            int rank = 0;
            const int cc = col - lo;
            const int w = cc >> 5;
            const bool in = act && (unsigned)cc < span_cols;
            if (in) {
                const uint4 wv = bmp4[w >> 2];
                const int j = w & 3;
                rank = pfx[w >> 2] + (j > 0 ? __popc(wv.x) : 0) + (j > 1 ? __popc(wv.y) : 0) +
                       (j > 2 ? __popc(wv.z) : 0);
                const unsigned cur = j == 0 ? wv.x : j == 1 ? wv.y : j == 2 ? wv.z : wv.w;
                rank += __popc(cur & ((1u << (cc & 31)) - 1u));
            }
            ordered_accumulate<FUSED, NT>(in && rank < nnz, rank, av, bv, acc, tags,
                                          (unsigned)l, single && b_strict);
        });
    wave_lds_fence();
    for (int i = l; i < nnz; i += G) c[o + i] = acc[i];
    for (int q = l; q < nw4; q += G) {
        const uint4 wv = bmp4[q];
        const unsigned ws[4] = {wv.x, wv.y, wv.z, wv.w};
        int pos = pfx[q];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            unsigned word = ws[j];
            const int cb = lo + ((4 * q + j) << 5);
            while (word) {
                const int bit = __ffs(word) - 1;
                word &= word - 1;
                if (pos < nnz) jc[o + pos] = cb + bit;
                pos++;
            }
        }
    }
}

// -------------------------------------------------------- big rows: span
// Column span of each big row, in 1024-column sub-blocks: from the first
// sub-block of the first B column to the last sub-block of the last B column
// over the row's A entries, and a bound on the words of the row's band word
// lists: min(products, span words) plus each band's padding.  One wave per
// big row.
__global__ void __launch_bounds__(256) k_big_span(int nbig, const int *__restrict__ rows,
                                                  const int *__restrict__ ia,
                                                  const int *__restrict__ ja,
                                                  const int *__restrict__ ib,
                                                  const int *__restrict__ jb,
                                                  const int64_t *__restrict__ rowprod,
                                                  int *__restrict__ sb0,
                                                  int64_t *__restrict__ nsb,
                                                  int64_t *__restrict__ nent,
                                                  int64_t *__restrict__ nprod,
                                                  int64_t *__restrict__ lbound) {
    const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int l = threadIdx.x & 63;
    if (r >= nbig) return;
    const int row = rows[r];
    int lo = 0x7fffffff, hi = -1;
    for (int p = ia[row] + l; p < ia[row + 1]; p += 64) {
        const int t = ja[p];
        const int b0 = ib[t], b1 = ib[t + 1];
        if (b0 < b1) {
            lo = min(lo, jb[b0] >> SUB_SHIFT);
            hi = max(hi, jb[b1 - 1] >> SUB_SHIFT);
        }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
    }
    if (l == 0) {
        const int64_t ns = hi >= lo ? (int64_t)(hi - lo + 1) : 0;
        sb0[r] = hi >= lo ? lo : 0;
        nsb[r] = ns;
        nent[r] = ia[row + 1] - ia[row];
        nprod[r] = rowprod[row];
        // a multiple of BAND_PAD, so every band's list starts 16-byte aligned
        lbound[r] = ((min(rowprod[row], ns << (SUB_SHIFT - 5)) + BAND_PAD - 1) & -BAND_PAD) +
                    BAND_PAD * ((ns + (1 << BAND_SHIFT) - 1) >> BAND_SHIFT);
    }
}

// -------------------------------------------------------- big rows: symbolic
// A big row's span is cut into bands of 64 sub-blocks; a band unit {r, u}
// covers span sub-blocks [64u, 64u + 64) of big row r.  Units are run in
// order of their first column (like the numeric items), so the units in
// flight read B columns of one narrow band, which the L2 caches keep.
__global__ void __launch_bounds__(256) k_band_count(int nbig, const int64_t *__restrict__ nsb,
                                                    int64_t *__restrict__ cnt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nbig) cnt[r] = (nsb[r] + (1 << BAND_SHIFT) - 1) >> BAND_SHIFT;
}

__global__ void __launch_bounds__(256) k_band_fill(int nbig, const int *__restrict__ span_sb0,
                                                   const int64_t *__restrict__ off,
                                                   int2 *__restrict__ units,
                                                   int *__restrict__ keys) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nbig) return;
    const int64_t o = off[r], n = off[r + 1] - o;
    for (int u = 0; u < n; u++) {
        units[o + u] = make_int2(r, u);
        keys[o + u] = span_sb0[r] + (u << BAND_SHIFT);
    }
}

__global__ void __launch_bounds__(256) k_gather_units(int64_t n, const int *__restrict__ order,
                                                      const int2 *__restrict__ units,
                                                      int2 *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = units[order[j]];
}

// One block per band unit: LDS bitmap over the band, the row's products in
// the band (B rows restricted through the block pointers) marked by the 2
// waves taking every other batch of 64 products; then the band's sub-block
// output counts, its share of the row's output count (rownnz of big rows is
// zero on entry) and, when words is given, the band's bitmap in sparse form:
// per sub-block a mask of its non-empty words (subocc), and the non-empty
// words themselves in column order (the band word list), in the row's region
// [row_off[r], row_off[r + 1]) at a position taken from the row's cursor
// (row_cur[r], starting at row_off[r]) and recorded in
// band_base[band_off[r] + u].  On R-MAT's big rows a fifth of the span's
// words are non-empty, so the lists and masks take a quarter of the bitmap's
// bytes, and no step expands bits.  A band takes its word count rounded up to
// BAND_PAD, so its list starts 16-byte aligned and goes out in whole 16-byte
// stores, staged in the bitmap's LDS once the words are in registers (one
// cursor for all bands had them queue on a single address; the cursor's
// result is first needed after the staging).  A band whose list would pass
// its row's region writes none and sets *overflow, which the host checks.
// 8 waves per SIMD, as with no word lists (the kernel waits on its gathers;
// its arguments alone would take the scalar registers past that).
__global__ void __launch_bounds__(BAND_T) __attribute__((amdgpu_waves_per_eu(8))) k_big_symbolic(
    int64_t nunits, const int2 *__restrict__ units, const int *__restrict__ rows, int k,
    const int *__restrict__ ia, const int *__restrict__ ja, const int *__restrict__ ib,
    const int *__restrict__ jb, BlockPtrs bp, const int *__restrict__ span_sb0,
    const int64_t *__restrict__ sb_off, uint16_t *__restrict__ subcnt,
    int64_t *__restrict__ rownnz, const int64_t *__restrict__ band_off,
    int64_t *__restrict__ band_base, unsigned long long *__restrict__ row_cur,
    const int64_t *__restrict__ row_off, unsigned *__restrict__ overflow,
    unsigned *__restrict__ subocc, unsigned *__restrict__ words_out) {
    constexpr int NW = BAND_T / 64;
    constexpr int TW = BAND_WORDS / BAND_T;        // words per thread, half a sub-block
    static_assert(TW == 16 && 2 * TW == (1 << (SUB_SHIFT - 5)), "a sub-block is two threads' words");
    __shared__ __align__(16) unsigned bits[BAND_WORDS];
    __shared__ int s_mark[NW][64];
    __shared__ int s_wsum[NW];
    __shared__ long long s_base;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    s_mark[wave][lane] = -1;
    for (int64_t ui = blockIdx.x; ui < nunits; ui += gridDim.x) {
        const int2 un = units[ui];
        const int r = un.x, u = un.y;
        const int row = rows[r];
        const int64_t so = sb_off[r] + ((int64_t)u << BAND_SHIFT);
        const int nsb = (int)min((int64_t)1 << BAND_SHIFT, sb_off[r + 1] - so);
        const int sb0 = span_sb0[r] + (u << BAND_SHIFT);
        const int c0 = sb0 << SUB_SHIFT;
        const int c1 = (int)min((int64_t)k, (int64_t)(sb0 + nsb) << SUB_SHIFT);
        const int words = (c1 - c0 + 31) >> 5;
        for (int i = threadIdx.x; i < words; i += BAND_T) bits[i] = 0u;
        __syncthreads();
        for_products<64, BIG_SYM_U, true, false>(
            ia[row], ia[row + 1], 0, 64, ja, nullptr, ib, jb, nullptr, bp, c0, c1, s_mark[wave],
            [&](bool act, int col, double, double, bool) {
                if (act) {
                    const int cc = col - c0;
                    lds_set_bit(&bits[cc >> 5], 1u << (cc & 31));
                }
            },
            wave, NW);
        __syncthreads();
        // thread t takes words [TW t, TW t + TW) (words past the band's end
        // are stale); threads 2s and 2s + 1 hold sub-block s
        unsigned wv[TW];
        int tot = 0;
        unsigned nzm = 0;
        {
            const uint4 *src = reinterpret_cast<const uint4 *>(bits) + threadIdx.x * (TW / 4);
#pragma unroll
            for (int q = 0; q < TW / 4; q++) {
                const uint4 v = src[q];
                wv[4 * q] = v.x;
                wv[4 * q + 1] = v.y;
                wv[4 * q + 2] = v.z;
                wv[4 * q + 3] = v.w;
            }
#pragma unroll
            for (int j = 0; j < TW; j++) {
                if ((int)threadIdx.x * TW + j >= words) wv[j] = 0u;
                tot += __popc(wv[j]);
                nzm |= (wv[j] != 0u ? 1u : 0u) << j;
            }
        }
        const int sbc = tot + __shfl_xor(tot, 1, 64);
        const unsigned occ = nzm | (__shfl_xor(nzm, 1, 64) << TW);
        if (!(threadIdx.x & 1) && (int)(threadIdx.x >> 1) < nsb) {
            subcnt[so + (threadIdx.x >> 1)] = (uint16_t)sbc;
            if (words_out) subocc[so + (threadIdx.x >> 1)] = occ;
        }
        // block-wide scans: outputs (for the row's count) and non-empty words
        const int nz = __popc(nzm);
        const int packed_incl = wave_incl_scan(tot | (nz << 16));
        if (lane == 63) s_wsum[wave] = packed_incl;
        __syncthreads();
        int before = (packed_incl >> 16) - nz, total = 0, total_nz = 0;
#pragma unroll
        for (int i = 0; i < NW; i++) {
            const int wsum = s_wsum[i];
            if (i < wave) before += wsum >> 16;
            total += wsum & 0xffff;
            total_nz += wsum >> 16;
        }
        const int padded = (total_nz + BAND_PAD - 1) & -BAND_PAD;
        // thread 0 takes the band's place in its row's region; the result is
        // first needed after the list is staged, so the atomic's round trip
        // overlaps the staging
        unsigned long long ticket = 0;
        if (threadIdx.x == 0) {
            if (total)
                atomicAdd(reinterpret_cast<unsigned long long *>(rownnz + row),
                          (unsigned long long)total);
            if (words_out && padded)
                ticket = atomicAdd(row_cur + r, (unsigned long long)padded);
        }
        if (words_out && total_nz > 0) {
            // the band's <= 2048 non-empty words fit the bitmap's LDS
            int pos = before;
#pragma unroll
            for (int j = 0; j < TW; j++) {
                if (wv[j]) bits[pos] = wv[j];
                pos += wv[j] ? 1 : 0;
            }
            if (threadIdx.x == 0) {
                long long base = (long long)ticket;
                if (base + padded > row_off[r + 1]) {
                    atomicOr(overflow, 1u);
                    base = -1;
                }
                band_base[band_off[r] + u] = base;
                s_base = base;
            }
            __syncthreads();
            const long long base = s_base;
            if (base >= 0) {
                // whole 16-byte stores, the last into the band's padding.
                // Written once, read once by the numeric phase much later:
                // streaming stores keep the L2 for B's rows
                typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
                u32x4 *dst = reinterpret_cast<u32x4 *>(words_out + base);
                const u32x4 *sv = reinterpret_cast<const u32x4 *>(bits);
                for (int i = threadIdx.x; i < padded / 4; i += BAND_T)
                    __builtin_nontemporal_store(sv[i], dst + i);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------ B block pointers
// Row-major table: bptr[t*(nblk+1) + b] = first position of B row t with
// column >= b << bs; one wave per B row, a lane per block.
__global__ void __launch_bounds__(256) k_bptr_rows(int n, const int *__restrict__ ib,
                                                   const int *__restrict__ jb, int nblk, int bs,
                                                   int *__restrict__ bptr) {
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int l = threadIdx.x & 63;
    if (t >= n) return;
    const int lo = ib[t], hi = ib[t + 1];
    for (int b = l; b <= nblk; b += 64) {
        const int pos = b == nblk ? hi : lower_bound_i32(jb, lo, hi, b << bs);
        bptr[(size_t)t * (nblk + 1) + b] = pos;
    }
}

// Block-major table: bptr[b*n + t] = first position of B row t with column
// >= b << bs.  A block of 4 waves takes 64 B rows; for each range of 64 blocks
// the waves fill an LDS tile [block][row], a lane per block searching inside
// its row, then write it out as 64 consecutive rows per block (256-byte runs).
// The 64 rows' columns are one contiguous range of jb: when it holds at most
// BPT_STAGE entries (R-MAT 2^20's B averages ~940 per 64 rows) it is staged
// in LDS once and every search runs there; otherwise the searches read jb.
// The searches wait on their loads, so residency mattered with global
// searches (a 128-block tile took 6.2 ms instead of 4.5 ms for R-MAT 2^20's
// 16,385 x 2^20 table, a 32-block one 7.1 ms), and still does with staged
// ones: blocks of 128 or 256 rows (512- or 1024-byte runs) took 6.6 and
// 8.4 ms instead of 4.5.
constexpr int BPT_ROWS = 64, BPT_BLK = 64, BPT_STAGE = 2048;

__device__ __forceinline__ int lds_lower_bound(const int *a, int lo, int hi, int key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_bptr(int n, const int *__restrict__ ib,
                                              const int *__restrict__ jb, int nblk, int bs,
                                              int *__restrict__ bptr) {
    __shared__ int tile[BPT_BLK][BPT_ROWS + 1];
    __shared__ int s_col[BPT_STAGE];
    __shared__ int s_rp[BPT_ROWS + 1];
    const int t0 = blockIdx.x * BPT_ROWS;
    const int nr = min(BPT_ROWS, n - t0);
    for (int i = threadIdx.x; i <= nr; i += blockDim.x) s_rp[i] = ib[t0 + i];
    __syncthreads();
    const int e0 = s_rp[0];
    const bool staged = s_rp[nr] - e0 <= BPT_STAGE;
    if (staged)
        for (int i = threadIdx.x; i < s_rp[nr] - e0; i += blockDim.x) s_col[i] = jb[e0 + i];
    __syncthreads();
    for (int b0 = 0; b0 <= nblk; b0 += BPT_BLK) {
        // (row, block) pairs, a thread each: a row's blocks in consecutive lanes
        for (int i = threadIdx.x; i < nr * BPT_BLK; i += blockDim.x) {
            const int r = i / BPT_BLK, bb = i % BPT_BLK;
            const int b = b0 + bb;
            if (b > nblk) continue;
            const int lo = s_rp[r], hi = s_rp[r + 1];
            int pos = hi;
            if (b < nblk)
                pos = staged ? e0 + lds_lower_bound(s_col, lo - e0, hi - e0, b << bs)
                             : lower_bound_i32(jb, lo, hi, b << bs);
            tile[bb][r] = pos;
        }
        __syncthreads();
        // a block's rows in consecutive lanes: runs of 4 * BPT_ROWS bytes
        for (int i = threadIdx.x; i < BPT_BLK * BPT_ROWS; i += blockDim.x) {
            const int bb = i / BPT_ROWS, r = i % BPT_ROWS;
            const int b = b0 + bb;
            if (b <= nblk && r < nr) bptr[(size_t)b * n + t0 + r] = tile[bb][r];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- item plan

// Split of each big row into items: the row's span is cut into windows of
// WSB sub-blocks (at most ITEM_WBITS columns); inside a window, runs of
// sub-blocks with at most ITEM_CAP outputs, leading empty sub-blocks skipped.
// One wave per big row, one lane per window (64 windows per pass): a lane
// loads its window's counts at once and splits them in registers, and wave
// scans place the items and their outputs.  FILL=false counts items per
// row; FILL=true (item_cnt = exclusive item offsets) writes each item's
// record -- {row, c0, c1, big-row index}, its first output position -- and
// output count, and, in rank mode (band_base given), where its first
// non-empty word is in the band word lists of the symbolic phase (from the
// sub-blocks' masks of non-empty words); the record's fourth field then
// holds the item's first sub-block's index in subcnt and subocc (< 2^31 in
// that mode) instead of the big-row index, so the numeric phase finds the
// item's masks with no other lookup.  Windows lie inside bands, and a
// band's windows inside one pass of 64 windows, so the item's place in its
// band's list is its count of non-empty words before it less that of the
// band's first window.
template <bool FILL, int WSB>
__global__ void __launch_bounds__(256) k_items(int nbig, const int *__restrict__ rows, int k,
                                               const int *__restrict__ span_sb0,
                                               const int64_t *__restrict__ sb_off,
                                               const uint16_t *__restrict__ subcnt,
                                               const int64_t *__restrict__ ic,
                                               int64_t *__restrict__ item_cnt,
                                               ItemRec *__restrict__ items,
                                               int *__restrict__ item_n,
                                               const unsigned *__restrict__ subocc,
                                               const int64_t *__restrict__ band_off,
                                               const int64_t *__restrict__ band_base) {
    static_assert(WSB <= (1 << (ITEM_WBITS_SHIFT - SUB_SHIFT)), "item wider than its LDS");
    constexpr int WPB = (1 << BAND_SHIFT) / WSB;   // windows per band
    static_assert(WPB * WSB == (1 << BAND_SHIFT) && 64 % WPB == 0, "windows tile the bands");
    const bool words = band_base != nullptr;   // rank mode: word-list positions
    const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int l = threadIdx.x & 63;
    if (r >= nbig) return;
    const int row = rows[r];
    const uint16_t *sc = subcnt + sb_off[r];
    const int nsb = (int)(sb_off[r + 1] - sb_off[r]);
    const int first = span_sb0[r];
    const int nwin = (nsb + WSB - 1) / WSB;
    int64_t wbase = FILL ? item_cnt[r] : 0;
    int64_t obase = FILL ? ic[row] : 0;
    for (int w0 = 0; w0 < nwin; w0 += 64) {
        const int win = w0 + l;
        const int j0 = win * WSB;
        const int len = win < nwin ? min(WSB, nsb - j0) : 0;
        int cs[WSB], os[WSB];
        int nzt = 0;
#pragma unroll
        for (int q = 0; q < WSB; q++) {
            cs[q] = q < len ? (int)sc[j0 + q] : 0;
            os[q] = FILL && words && q < len ? __popc(subocc[sb_off[r] + j0 + q]) : 0;
            nzt += os[q];
        }
        int nit = 0, tot = 0;
        {
            int start = -1, cnt = 0;
#pragma unroll
            for (int q = 0; q < WSB; q++) {
                if (start >= 0 && cnt + cs[q] > ITEM_CAP) {
                    nit++;
                    start = -1;
                    cnt = 0;
                }
                if (cs[q] != 0 || start >= 0) {
                    if (start < 0) start = q;
                    cnt += cs[q];
                }
                tot += cs[q];
            }
            if (start >= 0 && cnt > 0) nit++;
        }
        const int it_incl = wave_incl_scan(nit);
        const int out_incl = wave_incl_scan(tot);
        // non-empty words of this pass before this window, and before the
        // first window of this lane's band
        const int nz_excl = FILL ? wave_incl_scan(nzt) - nzt : 0;
        const int band_nz = FILL ? __shfl(nz_excl, (l / WPB) * WPB, 64) : 0;
        if (FILL && nit > 0) {
            int64_t w = wbase + it_incl - nit;
            int64_t out = obase + out_incl - tot;
            const int cbase = first + j0;
            // this window's first non-empty word in the band word lists
            const int64_t wb =
                words ? band_base[band_off[r] + win / WPB] + (nz_excl - band_nz) : 0;
            int start = -1, cnt = 0, nzrun = 0, nzs = 0;
#pragma unroll
            for (int q = 0; q < WSB; q++) {
                if (start >= 0 && cnt + cs[q] > ITEM_CAP) {
                    items[w] = ItemRec{make_int4(row, (cbase + start) << SUB_SHIFT,
                                                 min(k, (cbase + q) << SUB_SHIFT),
                                                 words ? (int)(sb_off[r] + j0 + start) : r),
                                       out, wb + nzs};
                    item_n[w] = cnt;
                    w++;
                    out += cnt;
                    start = -1;
                    cnt = 0;
                }
                if (cs[q] != 0 || start >= 0) {
                    if (start < 0) {
                        start = q;
                        nzs = nzrun;
                    }
                    cnt += cs[q];
                }
                nzrun += os[q];
            }
            if (start >= 0 && cnt > 0) {
                items[w] = ItemRec{make_int4(row, (cbase + start) << SUB_SHIFT,
                                             min(k, (cbase + len) << SUB_SHIFT),
                                             words ? (int)(sb_off[r] + j0 + start) : r),
                                   out, wb + nzs};
                item_n[w] = cnt;
            }
        }
        wbase += group_last<64>(it_incl);
        obase += group_last<64>(out_incl);
    }
    if (!FILL && l == 0) item_cnt[r] = wbase;
}

// Items in order of their first column: keys, then the gather of the items
// and their output offsets into the sorted order.  The numeric phase takes
// items in this order, so the items in flight at any time gather B entries
// of one narrow column band, which the L2 caches keep, instead of bands
// spread over all of B.
// Items sort by size class -- at most SMALL_ITEM outputs, at most MID_ITEM,
// more -- then by first column: key = class << class_bit | first sub-block
// (class_bit < 0: no classes).  The numeric phase runs each class in a
// kernel whose value array holds just that class's items, so smaller items
// take less LDS and more of them are resident.  block_small[b] and
// block_mid[b] count block b's items of the first two classes.
constexpr int SMALL_ITEM = 256;
constexpr int MID_ITEM = 512;
// The small-item kernel is bound by how many items are resident: with its
// 5.3 KB of LDS a wave, registers set the limit, so it runs 4 product
// batches in flight per lane at 6 waves per SIMD (<= 80 VGPRs) instead of 8
// at 4.  Big numeric on R-MAT 2^20 A*A^T: 149.7 ms at (8, 4); 142.5 at
// (6, 5); 136.4 at (4, 6); 139.7 at (3, 7) and 138.9 at (2, 8), both with
// a few registers spilled.
constexpr int SMALL_RANK_U = 4;
constexpr int SMALL_MINW = 6;
// the middle class: 7.3 KB of LDS a wave allows 21 waves per CU, so 5 per
// SIMD, with 6 batches in flight (<= 96 VGPRs).  (The last class stays at 8
// batches: its LDS caps it at 14 waves per CU, and 10 or 12 batches in the
// registers that leaves gave 133.3 and 135.1 ms against 134.2.)  R-MAT 2^20 A*A^T: items
// of > 256 outputs 96.8 ms in one class, 77.3 + 16.9 ms in two.  (Cutting
// items at 512 outputs where whole sub-blocks allow, so that fewer land in
// the last class: 148 ms in all instead of 134, more (A entry, item) pairs.)
constexpr int MID_RANK_U = 6;
constexpr int MID_MINW = 5;

__global__ void __launch_bounds__(256) k_item_keys(int64_t n, const ItemRec *__restrict__ items,
                                                   const int *__restrict__ item_n, int class_bit,
                                                   int *__restrict__ keys,
                                                   int *__restrict__ block_small,
                                                   int *__restrict__ block_mid) {
    __shared__ int s_cnt[2][4];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int cls = 0;
    if (i < n && class_bit >= 0) cls = item_n[i] <= SMALL_ITEM ? 0 : item_n[i] <= MID_ITEM ? 1 : 2;
    if (i < n) keys[i] = (cls << max(class_bit, 0)) | (items[i].d.y >> SUB_SHIFT);
    const unsigned long long b0 = __ballot(i < n && class_bit >= 0 && cls == 0);
    const unsigned long long b1 = __ballot(i < n && class_bit >= 0 && cls == 1);
    if (__lane_id() == 0) {
        s_cnt[0][threadIdx.x >> 6] = __popcll(b0);
        s_cnt[1][threadIdx.x >> 6] = __popcll(b1);
    }
    __syncthreads();
    if (threadIdx.x < 2)
        (threadIdx.x ? block_mid : block_small)[blockIdx.x] =
            s_cnt[threadIdx.x][0] + s_cnt[threadIdx.x][1] + s_cnt[threadIdx.x][2] +
            s_cnt[threadIdx.x][3];
}

__global__ void __launch_bounds__(256) k_gather_items(int64_t n, const int *__restrict__ order,
                                                      const ItemRec *__restrict__ items,
                                                      ItemRec *__restrict__ items_out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) items_out[j] = items[order[j]];
}

// --------------------------------------------------------- big rows: numeric
// One wave per item (grid-stride over items).  LDS per wave: a key hash with
// a 16-bit id per slot, values compacted by insertion id (only the item's
// <= ITEM_CAP outputs are stored), the item's column bitmap (emission order)
// and the tag table of the ordered accumulation.
struct ItemSmem {
    double vals[ITEM_CAP];
    int keys[ITEM_SLOTS];
    unsigned short ids[ITEM_SLOTS];
    unsigned bits[ITEM_WORDS];
    unsigned tags[ITEM_TAGS];
    int mark[64];                 // segment marks of the product expansion
    int nid;
};

template <bool FUSED>
__global__ void __launch_bounds__(64 * ITEM_WAVES) k_big_numeric(
    int64_t nitems, const ItemRec *__restrict__ items,
    const int *__restrict__ ia, const int *__restrict__ ja, const double *__restrict__ a,
    const int *__restrict__ ib, const int *__restrict__ jb, const double *__restrict__ b,
    BlockPtrs bp, int *__restrict__ jc,
    double *__restrict__ c, int b_strict) {
    extern __shared__ __align__(16) unsigned char smem[];
    ItemSmem &S = reinterpret_cast<ItemSmem *>(smem)[threadIdx.x >> 6];
    const int l = __lane_id();
    for (int i = l; i < ITEM_TAGS; i += 64) S.tags[i] = 0xffffffffu;
    S.mark[l] = -1;
    const int64_t wave_id = (int64_t)blockIdx.x * ITEM_WAVES + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * ITEM_WAVES;
    for (int64_t it = wave_id; it < nitems; it += nwaves) {
        const int4 item = items[it].d;
        const int row = item.x, c0 = item.y, c1 = item.z;
        const int nwords = (c1 - c0 + 31) >> 5;
        for (int i = l; i < ITEM_SLOTS; i += 64) S.keys[i] = -1;
        for (int i = l; i < nwords; i += 64) S.bits[i] = 0u;
        if (l == 0) S.nid = 0;
        wave_lds_fence();
        for_products<64, ITEM_U, true, true>(
            ia[row], ia[row + 1], 0, 64, ja, a, ib, jb, b, bp, c0, c1, S.mark,
            [&](bool act, int col, double av, double bv, bool single) {
                int slot = -1;
                if (act) {
                    bool created;
                    slot = hash_insert<ITEM_SLOTS>(S.keys, col, created);
                    if (created) {
                        const int id = atomicAdd(&S.nid, 1);
                        S.ids[slot] = (unsigned short)id;
                        S.vals[id] = 0.0;
                        const int cc = col - c0;
                        atomicOr(&S.bits[cc >> 5], 1u << (cc & 31));
                    }
                }
                wave_lds_fence();
                const int id = slot >= 0 ? (int)S.ids[slot] : -1;
                ordered_accumulate<FUSED, ITEM_TAGS>(act, id, av, bv, S.vals, S.tags,
                                                     (unsigned)l, single && b_strict);
            });
        wave_lds_fence();
        // emit in column order: bitmap words -> ranks -> hash lookups
        int64_t o = items[it].off;
        for (int w0 = 0; w0 < nwords; w0 += 64) {
            const int w = w0 + l;
            unsigned word = w < nwords ? S.bits[w] : 0u;
            const int pc = __popc(word);
            const int incl = wave_incl_scan(pc);
            int64_t pos = o + incl - pc;
            while (word) {
                const int bit = __ffs(word) - 1;
                word &= word - 1;
                const int col = c0 + (w << 5) + bit;
                const int slot = hash_find<ITEM_SLOTS>(S.keys, col);
                jc[pos] = col;
                c[pos] = slot >= 0 ? S.vals[S.ids[slot]] : 0.0;
                pos++;
            }
            o += __shfl(incl, 63, 64);
        }
        wave_lds_fence();
    }
}

// Rank-indexed variant: the symbolic phase stored every big row's column
// bitmap (in sparse form: per sub-block a mask of its non-empty words, and
// those words), so an item's outputs are known before its products are
// visited.  The rank of column j inside the item is prefix[w] + popc(bits[w]
// below j) -- one LDS read of a {bits, prefix} pair -- and values accumulate
// in a dense rank-ordered array that is written out coalesced.
// WORDS: bitmap words of the widest item (256 or 512: items of 8192 or
// 16384 columns).
template <int WORDS, int CAP = ITEM_CAP>
struct RankSmem {
    uint2 words[WORDS];           // {bitmap word, outputs before it}
    double vals[CAP];
    unsigned tags[ITEM_TAGS];     // bucketed by rank
    int mark[64];                 // segment marks of the product expansion
};

// Rank table of one item, in two parts so that other loads can be issued
// between them: item_words loads the item's bitmap words (<= ITEM_WORDS,
// every load issued before the first is used); item_rank_build clears bits
// at or past c1, stores {word, outputs before it} pairs in S.words, writes
// the item's columns to jc at offset o (expanded into S.vals, then stored
// coalesced; the stores drain while the products are processed), zeroes the
// item's first cnt values and returns cnt, the item's output count.
// item_words rebuilds the words from the band word list: lane l takes words
// l, l + 64, ..., and word w of sub-block s (lanes 0..15 of occ hold the
// item's sub-block masks, of pre their exclusive popcount prefix) is there
// when its bit of occ[s] is set, at list + pre[s] + popc(occ[s] below it).
template <int WCH>
__device__ __forceinline__ void item_words(const unsigned *__restrict__ list, unsigned occ,
                                           int pre, int c0, int c1, unsigned (&wv)[WCH]) {
    static_assert(2 * WCH <= 16, "an item's sub-block masks sit in lanes 0..15");
    const int l = __lane_id();
    const int nwords = (c1 - c0 + 31) >> 5;
    const unsigned below = (1u << (l & 31)) - 1u;
#pragma unroll
    for (int j = 0; j < WCH; j++) {
        const int w = j * 64 + l;
        // word w is in sub-block 2j (lanes 0..31) or 2j + 1 (lanes 32..63)
        const unsigned o0 = __builtin_amdgcn_readlane(occ, 2 * j);
        const unsigned o1 = __builtin_amdgcn_readlane(occ, 2 * j + 1);
        const int p0 = __builtin_amdgcn_readlane(pre, 2 * j);
        const int p1 = __builtin_amdgcn_readlane(pre, 2 * j + 1);
        const unsigned om = l < 32 ? o0 : o1;
        const int pm = l < 32 ? p0 : p1;
        const bool present = w < nwords && ((om >> (l & 31)) & 1u);
        // read once: a streaming load keeps the L2 for B's rows.  A 32-bit
        // offset from the wave's scalar base keeps one address register
        // per load
        const unsigned off = (unsigned)pm + (unsigned)__popc(om & below);
        wv[j] = present ? __builtin_nontemporal_load(list + off) : 0u;
    }
}

template <int WORDS, int CAP, int WCH = WORDS / 64>
__device__ __forceinline__ int item_rank_build(RankSmem<WORDS, CAP> &S, unsigned (&wv)[WCH], int c0,
                                               int c1, int64_t o, int *__restrict__ jc) {
    const int l = __lane_id();
    const int nwords = (c1 - c0 + 31) >> 5;
    // the 64-word chunks are scanned independently
    int incl[WCH], tot[WCH];
#pragma unroll
    for (int j = 0; j < WCH; j++) {
        const int w = j * 64 + l;
        const int lastbits = (c1 - c0) - (w << 5);
        if (w < nwords && lastbits < 32) wv[j] &= (1u << lastbits) - 1u;
        incl[j] = wave_incl_scan(__popc(wv[j]));
        tot[j] = group_last<64>(incl[j]);
    }
    int cnt = 0;
    int *cols = reinterpret_cast<int *>(S.vals);
#pragma unroll
    for (int j = 0; j < WCH; j++) {
        const int w = j * 64 + l;
        const int base = cnt + incl[j] - __popc(wv[j]);
        if (w < nwords) S.words[w] = make_uint2(wv[j], (unsigned)base);
        unsigned word = wv[j];
        int pos = base;
        while (word) {
            const int bit = __ffs(word) - 1;
            word &= word - 1;
            cols[pos++] = c0 + (w << 5) + bit;
        }
        cnt += tot[j];
    }
    wave_lds_fence();
    for (int i = l; i < cnt; i += 64) __builtin_nontemporal_store(cols[i], jc + o + i);
    wave_lds_fence();
    for (int i = l; i < cnt; i += 64) S.vals[i] = 0.0;
    wave_lds_fence();
    return cnt;
}

// Writes an item's cnt values to C at offset o, straight from LDS.
template <int WORDS, int CAP>
__device__ __forceinline__ void item_store_values(RankSmem<WORDS, CAP> &S, int cnt, int64_t o,
                                                  double *__restrict__ c) {
    const int l = __lane_id();
    wave_lds_fence();
    for (int i = l; i < cnt; i += 64) __builtin_nontemporal_store(S.vals[i], c + o + i);
    wave_lds_fence();
}

// Accumulates one batch of products into the item's rank-ordered values:
// a batch from one A entry (distinct columns when B's rows are strictly
// ascending) adds directly, any other batch by tag rounds in lane order.
template <bool FUSED, int WORDS, int CAP>
__device__ __forceinline__ void item_accumulate(RankSmem<WORDS, CAP> &S, int c0, bool act, int col,
                                                double av, double bv, bool single, int b_strict) {
    int rank = -1;
    if (act) {
        const int cc = col - c0;
        const uint2 wd = S.words[cc >> 5];
        rank = (int)wd.y + __popc(wd.x & ((1u << (cc & 31)) - 1u));
    }
    if (single && b_strict) {
        if (rank >= 0) S.vals[rank] = acc_step<FUSED>(S.vals[rank], av, bv);
        wave_lds_fence();
        return;
    }
    ordered_accumulate<FUSED, ITEM_TAGS>(rank >= 0, rank, av, bv, S.vals, S.tags,
                                         (unsigned)__lane_id(), false);
}

template <bool FUSED, int WORDS, int CAP = ITEM_CAP, int U = RANK_U, int MINW = 1>
__global__ void __launch_bounds__(64 * ITEM_WAVES) __attribute__((amdgpu_waves_per_eu(MINW)))
k_big_numeric_rank(
    int64_t nitems, const ItemRec *__restrict__ items, const unsigned *__restrict__ wlist,
    const unsigned *__restrict__ subocc, const int *__restrict__ ia, const int *__restrict__ ja,
    const double *__restrict__ a, const int *__restrict__ jb, const double *__restrict__ b,
    const int4 *__restrict__ bpk, BlockPtrs bp, int *__restrict__ jc, double *__restrict__ c,
    int b_strict) {
    extern __shared__ __align__(16) unsigned char smem[];
    RankSmem<WORDS, CAP> &S = reinterpret_cast<RankSmem<WORDS, CAP> *>(smem)[threadIdx.x >> 6];
    const int l = __lane_id();
    for (int i = l; i < ITEM_TAGS; i += 64) S.tags[i] = 0xffffffffu;
    S.mark[l] = -1;
    const int64_t wave_id = (int64_t)blockIdx.x * ITEM_WAVES + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * ITEM_WAVES;
    // the next item's descriptor is loaded when an item starts, and its
    // sub-block masks, word-list position and A row bounds while the item's
    // outputs are stored, so an item's first product loads need no chain of
    // dependent loads.  Those and the item's output offset go into one
    // register, a lane each (lanes 0..15 the masks, 16 and 17 the position,
    // 18 and 19 the row bounds, 20 and 21 the offset: one load instruction,
    // and they reach scalar registers at the item's start -- the small-item
    // kernel has no vector registers to spare).
    int4 item = make_int4(0, 0, 0, 0);
    unsigned nx = 0u;
    auto locate = [&](int64_t at) {
        // item.w: the item's first sub-block in subocc
        const int nsb = (item.z - item.y + (1 << SUB_SHIFT) - 1) >> SUB_SHIFT;
        const unsigned *src = nullptr;
        if (l < min(nsb, 2 * (WORDS / 64)))
            src = subocc + ((int64_t)item.w + l);
        else if (l == 16 || l == 17)
            src = reinterpret_cast<const unsigned *>(&items[at].wpos) + (l - 16);
        else if (l == 18 || l == 19)
            src = reinterpret_cast<const unsigned *>(ia + item.x) + (l - 18);
        else if (l == 20 || l == 21)
            src = reinterpret_cast<const unsigned *>(&items[at].off) + (l - 20);
        nx = src ? *src : 0u;
    };
    // (streaming loads of the item descriptors and A entries, read once per
    // item, were measured slower: 158 -> 166 ms on R-MAT 2^20 A*A^T)
    if (wave_id < nitems) {
        item = items[wave_id].d;
        locate(wave_id);
    }
    for (int64_t it = wave_id; it < nitems; it += nwaves) {
        const int c0 = __builtin_amdgcn_readfirstlane(item.y);
        const int c1 = __builtin_amdgcn_readfirstlane(item.z);
        const unsigned iocc = l < 16 ? nx : 0u;
        const int ipa = __builtin_amdgcn_readlane((int)nx, 18);
        const int ipe = __builtin_amdgcn_readlane((int)nx, 19);
        const int64_t iwpos =
            (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)nx, 17) << 32) |
                      (uint32_t)__builtin_amdgcn_readlane((int)nx, 16));
        const unsigned *ilist = wlist + iwpos;
        const int64_t o =
            (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)nx, 21) << 32) |
                      (uint32_t)__builtin_amdgcn_readlane((int)nx, 20));
        const int64_t next = it + nwaves;
        if (next < nitems) item = items[next].d;
        // the product pipeline of for_products, with the rank table built
        // after the first groups' A entries and B ranges are issued, so the
        // word loads and that chain of dependent loads overlap
        auto load_t = [&](int g, int &t, double &av) {
            const int p = g + l;
            t = -1;
            av = 0.0;
            if (g < ipe && p < ipe) {
                t = ja[p];
                av = a[p];
            }
        };
        auto load_range = [&](int t, int &lo, int &hi) {
            lo = 0;
            hi = 0;
            if (t >= 0) {
                lo = brange_lo(bp, jb, t, c0);
                hi = brange_lo(bp, jb, t, c1);
            }
        };
        int t0, t1;
        double a0, a1;
        load_t(ipa, t0, a0);
        load_t(ipa + 64, t1, a1);
        const int pc = __popc(iocc);
        const int pre = dpp_incl_scan<16>(pc) - pc;
        unsigned wv[WORDS / 64];
        item_words(ilist, iocc, pre, c0, c1, wv);
        int lo, hi;
        load_range(t0, lo, hi);
        const int cnt = item_rank_build<WORDS, CAP>(S, wv, c0, c1, o, jc);
        auto f = [&](bool act, int col, double av, double bv, bool single) {
            item_accumulate<FUSED, WORDS, CAP>(S, c0, act, col, av, bv, single, b_strict);
        };
        for (int g = ipa; g < ipe; g += 64) {
            const double ag = a0;
            int lo_n, hi_n, t2;
            double a2;
            load_range(t1, lo_n, hi_n);
            load_t(g + 128, t2, a2);
            expand_group<64, U, true, true>(lo, hi - lo, ag, jb, b, S.mark, f, 0, 1, bpk);
            lo = lo_n;
            hi = hi_n;
            t1 = t2;
            a0 = a1;
            a1 = a2;
        }
        if (next < nitems) locate(next);
        item_store_values(S, cnt, o, c);
    }
}

// B's entries as 16-byte {column, 0, value lo, value hi} records for the
// big-row numeric kernel's gathers.
__global__ void __launch_bounds__(256) k_pack_b(int64_t nnz, const int *__restrict__ jb,
                                                const double *__restrict__ b,
                                                int4 *__restrict__ bpk) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    const double v = b[q];
    bpk[q] = make_int4(jb[q], 0, __double2loint(v), __double2hiint(v));
}

// The plan's block pointers in the layout they were built in; without a
// table, B's row pointers b_ptr (of the call at hand) as a one-block table.
static BlockPtrs block_ptrs(const RowPlan &p, const int *b_ptr) {
    BlockPtrs bp;
    bp.p = p.bptr;
    bp.bs = p.blk_shift;
    if (p.bptr_rowptr) {
        bp.p = b_ptr;
        bp.tstride = 1;
        bp.bstride = 1;
    } else if (p.column_order) {
        bp.tstride = 1;
        bp.bstride = p.n;
    } else {
        bp.tstride = p.nblk + 1;
        bp.bstride = 1;
    }
    return bp;
}

// One block per item (the hardware dispatcher hands the next item to
// whichever CU frees a slot first, so no CU idles behind a long tail); the
// item kernels keep a grid-stride loop for item counts beyond the grid limit.
static unsigned item_grid(int64_t nitems) {
    return (unsigned)std::min<int64_t>((nitems + ITEM_WAVES - 1) / ITEM_WAVES, int64_t(1) << 30);
}

// Raises a kernel's dynamic LDS limit once per process (later calls, e.g.
// inside a captured numeric phase, touch no runtime state).
template <class K>
void set_smem(K kernel, size_t bytes) {
    static std::mutex mu;
    static std::map<const void *, size_t> done;
    const void *key = reinterpret_cast<const void *>(kernel);
    std::lock_guard<std::mutex> lk(mu);
    auto it = done.find(key);
    if (it != done.end() && it->second >= bytes) return;
    TSG_HIP_TRY(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done[key] = bytes;
}

}  // namespace

// =================================================================== host
void row_symbolic(RowPlan &p, const CsrIn &A, const CsrIn &B, int64_t *d_ic, int64_t *nnz_c,
                  Arena &ws, hipStream_t s, StageTimer &tm, int /*cus*/, const RowTuning &tune) {
    const int m = A.rows, n = A.cols, k = B.cols;
    TSG_CHECK(tune.bin1 >= 0 && tune.bin1 <= bin_limit[1] && tune.bin1 <= tune.bin2 &&
                  tune.bin2 <= bin_limit[2] && tune.bin2 <= tune.bin3 &&
                  tune.bin3 <= bin_limit[3],
              TSG_ERR_INVALID, "row bins out of range");
    const BinLimits lim{tune.bin1, tune.bin2, tune.bin3};
    p = RowPlan();
    p.m = m;
    p.n = n;
    p.k = k;
    p.rowprod = ws.alloc_n<int64_t>((size_t)m + 1);
    p.rownnz = ws.alloc_n<int64_t>((size_t)m + 1);
    p.bin_rows = ws.alloc_n<int>((size_t)m + 1);
    int *d_cnt = ws.alloc_n<int>(2 * NBINS + 4);
    int64_t *scan_tmp = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)m + 1));

    tm.begin("row.products+bins");
    csr_row_products(m, A.ptr, A.idx, A.nnz, B.ptr, p.rowprod, ws, s);
    TSG_HIP_TRY(hipMemsetAsync(d_cnt, 0, sizeof(int) * (2 * NBINS + 4), s));
    if (m > 0) {
        hipLaunchKernelGGL(k_bin_count, dim3(grid_for(m, 256)), dim3(256), 0, s, m, p.rowprod, lim,
                           d_cnt);
        TSG_LAUNCH_CHECK();
    }
    int h_cnt[NBINS];
    TSG_HIP_TRY(hipMemcpyAsync(h_cnt, d_cnt, sizeof(h_cnt), hipMemcpyDeviceToHost, s));
    TSG_HIP_TRY(hipStreamSynchronize(s));
    BinStarts st;
    st.s[0] = 0;
    for (int b = 0; b < NBINS; b++) st.s[b + 1] = st.s[b] + (b == 0 ? 0 : h_cnt[b]);
    for (int b = 0; b <= NBINS; b++) p.bin_start_h[b] = st.s[b];
    p.nbins = NBINS;
    if (m > 0) {
        hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(m, 256)), dim3(256), 0, s, m, p.rowprod, lim,
                           st, d_cnt + NBINS, p.bin_rows, p.rownnz);
        TSG_LAUNCH_CHECK();
    }
    tm.end();

    // small rows
    tm.begin("row.small_symbolic");
    {
        p.span = ws.alloc_n<int2>((size_t)m);
        const int *rows1 = p.bin_rows + st.s[1];
        int n1 = st.s[2] - st.s[1];
        if (n1 > 0) {
            hipLaunchKernelGGL((k_small_symbolic<16, 128>), dim3(grid_for(n1, 16)), dim3(256), 0,
                               s, n1, rows1, A.ptr, A.idx, B.ptr, B.idx, p.rownnz, p.span, nullptr);
            TSG_LAUNCH_CHECK();
        }
        const int *rows2 = p.bin_rows + st.s[2];
        int n2 = st.s[3] - st.s[2];
        if (n2 > 0) {
            hipLaunchKernelGGL((k_small_symbolic<32, 512>), dim3(grid_for(n2, 8)), dim3(256), 0,
                               s, n2, rows2, A.ptr, A.idx, B.ptr, B.idx, p.rownnz, p.span, nullptr);
            TSG_LAUNCH_CHECK();
        }
        const int *rows3 = p.bin_rows + st.s[3];
        int n3 = st.s[4] - st.s[3];
        if (n3 > 0) {
            p.narrow3 = ws.alloc_n<uint8_t>((size_t)m);
            hipLaunchKernelGGL((k_small_symbolic<64, 2048>), dim3(grid_for(n3, 4)), dim3(256), 0,
                               s, n3, rows3, A.ptr, A.idx, B.ptr, B.idx, p.rownnz, p.span, p.narrow3);
            TSG_LAUNCH_CHECK();
            // each numeric variant launches over its own rows; the class
            // sizes come back with the phase's last host round trip
            p.rows3 = ws.alloc_n<int>((size_t)n3);
            p.rows3r = ws.alloc_n<int>((size_t)n3);
            hipLaunchKernelGGL(k_split_bin3, dim3(grid_for(n3, 256)), dim3(256), 0, s, n3, rows3,
                               p.narrow3, p.rows3, p.rows3r, d_cnt + 2 * NBINS);
            TSG_LAUNCH_CHECK();
            TSG_HIP_TRY(hipMemcpyAsync(p.n3cls, d_cnt + 2 * NBINS, 4 * sizeof(int),
                                       hipMemcpyDeviceToHost, s));
        }
    }
    tm.end();

    // big rows
    p.nbig = st.s[5] - st.s[4];
    p.sub_shift = SUB_SHIFT;
    int64_t *band_off = nullptr, *band_base = nullptr, *row_loff = nullptr;
    unsigned long long *row_cur = nullptr;
    unsigned *d_overflow = nullptr;
    if (p.nbig > 0) {
        const int *rowsb = p.bin_rows + st.s[4];
        tm.begin("row.big_symbolic");
        // column span of every big row, in sub-blocks, span offsets and the
        // band units over the spans
        p.span_sb0 = ws.alloc_n<int>((size_t)p.nbig + 1);
        p.sb_off = ws.alloc_n<int64_t>((size_t)p.nbig + 1);
        band_off = ws.alloc_n<int64_t>((size_t)p.nbig + 1);
        int64_t *tmp_sb = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)p.nbig + 1));
        int64_t *a_ent = ws.alloc_n<int64_t>((size_t)p.nbig + 1);
        int64_t *a_prod = ws.alloc_n<int64_t>((size_t)p.nbig + 1);
        row_loff = ws.alloc_n<int64_t>((size_t)p.nbig + 1);
        hipLaunchKernelGGL(k_big_span, dim3(grid_for((size_t)p.nbig * 64, 256)), dim3(256), 0, s,
                           p.nbig, rowsb, A.ptr, A.idx, B.ptr, B.idx, p.rowprod, p.span_sb0,
                           p.sb_off, a_ent, a_prod, row_loff);
        TSG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_band_count, dim3(grid_for((size_t)p.nbig, 256)), dim3(256), 0, s,
                           p.nbig, p.sb_off, band_off);
        TSG_LAUNCH_CHECK();
        exclusive_scan_i64(p.sb_off, p.sb_off, (size_t)p.nbig, tmp_sb, s);
        exclusive_scan_i64(band_off, band_off, (size_t)p.nbig, tmp_sb, s);
        exclusive_scan_i64(a_ent, a_ent, (size_t)p.nbig, tmp_sb, s);
        exclusive_scan_i64(a_prod, a_prod, (size_t)p.nbig, tmp_sb, s);
        exclusive_scan_i64(row_loff, row_loff, (size_t)p.nbig, tmp_sb, s);
        int64_t tot2[5] = {0, 0, 0, 0, 0};
        TSG_HIP_TRY(hipMemcpyAsync(&tot2[0], p.sb_off + p.nbig, sizeof(int64_t),
                                   hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipMemcpyAsync(&tot2[1], band_off + p.nbig, sizeof(int64_t),
                                   hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipMemcpyAsync(&tot2[2], a_ent + p.nbig, sizeof(int64_t),
                                   hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipMemcpyAsync(&tot2[3], a_prod + p.nbig, sizeof(int64_t),
                                   hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipMemcpyAsync(&tot2[4], row_loff + p.nbig, sizeof(int64_t),
                                   hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipStreamSynchronize(s));
        const int64_t nsb_total = tot2[0], nunits = tot2[1];
        p.nsb_total = nsb_total;
        p.subcnt = ws.alloc_n<uint16_t>((size_t)nsb_total + 1);
        // keep every big row's bitmap in sparse form (masks of non-empty
        // words per sub-block, and band word lists: each row's region sized
        // by the bound min(products, span words) plus its bands' padding) for
        // the rank-indexed numeric phase when they fit a budget of device
        // memory; else the numeric phase uses per-item hash tables
        p.words_cap = tot2[4];
        const size_t cols_bytes =
            ((size_t)p.words_cap + (size_t)nsb_total) * sizeof(unsigned);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { (void)hipGetLastError(); free_b = 0; }
        const size_t budget = tune.bitmap_budget >= 0
                                  ? (size_t)tune.bitmap_budget
                                  : std::min<size_t>(size_t(48) << 30, free_b / 10 * 4);
        p.use_rank = cols_bytes <= budget && nsb_total < (int64_t(1) << 31);
        p.bandwords = p.use_rank ? ws.alloc_n<unsigned>((size_t)p.words_cap + 1) : nullptr;
        p.subocc = p.use_rank ? ws.alloc_n<unsigned>((size_t)nsb_total + 1) : nullptr;
        if (p.use_rank) {
            band_base = ws.alloc_n<int64_t>((size_t)nunits + 1);
            row_cur = ws.alloc_n<unsigned long long>((size_t)p.nbig + 1);
            d_overflow = ws.alloc_n<unsigned>(1);
            TSG_HIP_TRY(hipMemcpyAsync(row_cur, row_loff, sizeof(int64_t) * (size_t)p.nbig,
                                       hipMemcpyDeviceToDevice, s));
            TSG_HIP_TRY(hipMemsetAsync(d_overflow, 0, sizeof(unsigned), s));
        }
        // B as {column, value} records for the rank kernel's gathers
        p.bpk = p.use_rank ? ws.alloc_n<int4>((size_t)B.nnz + 1) : nullptr;
        // Spans this wide (at least item_sort_min windows of 16 sub-blocks)
        // make enough bands and items for them to run in column order, which
        // a radix sort arranges; the block pointers are then laid out block
        // major and the items are 8 sub-blocks wide.
        p.column_order = nsb_total / 16 >= tune.item_sort_min;
        tm.end();

        // B block pointers at the sub-block width (item and band boundaries
        // then need no search) while the table stays <= 6 GiB, coarser beyond.
        // No table at all when building it would cost more than the searches
        // it saves -- more than 8 max(1, log2(L / 16)) entries per (A entry,
        // window of 16 sub-blocks) pair of the big rows, with the pairs
        // estimated from their A entries and average span and L, the mean
        // length of the B rows they reference, from their products per A
        // entry (measured on batches: an entry costs ~5 ps to build, a search
        // ~26 ps in B rows of ~30 entries, which span one or two cache lines,
        // and several times that in rows of hundreds, where each step misses)
        // -- or when the budget is 0: B's row pointers then act as a one-block
        // table (block width 2^30 >= k) and every range comes from a search
        // inside its B row.
        tm.begin("row.bptr");
        const size_t bptr_cap = tune.bptr_budget >= 0 ? (size_t)tune.bptr_budget
                                                      : (size_t(6) << 30);
        const double fine_entries =
            (double)n * (double)((((size_t)k + (1u << SUB_SHIFT) - 1) >> SUB_SHIFT) + 1);
        const double pairs = (double)tot2[2] * ((double)nsb_total / (16.0 * p.nbig) + 1.0);
        const double brow = tot2[2] > 0 ? (double)tot2[3] / (double)tot2[2] : 1.0;
        p.bptr_rowptr = k <= (1 << 30) &&
                        (tune.bptr_budget == 0 ||
                         (tune.bptr_budget < 0 &&
                          fine_entries > 8.0 * pairs * std::max(1.0, std::log2(brow / 16.0))));
        int bs = SUB_SHIFT;
        while (bs < 30) {
            size_t nblk = ((size_t)k + (1u << bs) - 1) >> bs;
            if ((size_t)n * (nblk + 1) * sizeof(int) <= bptr_cap) break;
            bs++;
        }
        if (p.bptr_rowptr) bs = 30;
        p.blk_shift = bs;
        p.nblk = (int)(((size_t)k + (1u << bs) - 1) >> bs);
        p.bptr = p.bptr_rowptr ? nullptr : ws.alloc_n<int>((size_t)n * (p.nblk + 1));
        if (p.bptr_rowptr) {
            // nothing to build
        } else if (n > 0 && p.column_order) {
            hipLaunchKernelGGL(k_bptr, dim3(grid_for((size_t)n, BPT_ROWS)), dim3(256), 0, s, n,
                               B.ptr, B.idx, p.nblk, bs, p.bptr);
            TSG_LAUNCH_CHECK();
        } else if (n > 0) {
            hipLaunchKernelGGL(k_bptr_rows, dim3(grid_for((size_t)n * 64, 256)), dim3(256), 0,
                               s, n, B.ptr, B.idx, p.nblk, bs, p.bptr);
            TSG_LAUNCH_CHECK();
        }
        tm.end();

        tm.begin("row.big_symbolic");
        int2 *units = ws.alloc_n<int2>((size_t)nunits + 1);
        int *ukeys = ws.alloc_n<int>((size_t)nunits + 1);
        hipLaunchKernelGGL(k_band_fill, dim3(grid_for((size_t)p.nbig, 256)), dim3(256), 0, s,
                           p.nbig, p.span_sb0, band_off, units, ukeys);
        TSG_LAUNCH_CHECK();
        if (p.column_order && nunits > 1 && nunits < (int64_t(1) << 31)) {
            int *order = ws.alloc_n<int>((size_t)nunits);
            int end_bit = 1;
            while (end_bit < 31 && (int64_t(1) << end_bit) <= ((int64_t)k >> SUB_SHIFT)) end_bit++;
            sort_order_i32(ukeys, nunits, end_bit, order, ws, s);
            int2 *sorted = ws.alloc_n<int2>((size_t)nunits + 1);
            hipLaunchKernelGGL(k_gather_units, dim3(grid_for((size_t)nunits, 256)), dim3(256), 0,
                               s, nunits, order, units, sorted);
            TSG_LAUNCH_CHECK();
            units = sorted;
        }
        hipLaunchKernelGGL(k_big_symbolic,
                           dim3((unsigned)std::min<int64_t>(nunits, int64_t(1) << 30)),
                           dim3(BAND_T), 0, s, nunits, units, rowsb, k, A.ptr, A.idx, B.ptr,
                           B.idx, block_ptrs(p, B.ptr), p.span_sb0, p.sb_off, p.subcnt, p.rownnz,
                           band_off, band_base, row_cur, row_loff, d_overflow, p.subocc,
                           p.bandwords);
        TSG_LAUNCH_CHECK();
        tm.end();
    }

    tm.begin("row.scan");
    exclusive_scan_i64(p.rownnz, d_ic, (size_t)m, scan_tmp, s);
    tm.end();

    int64_t h_tot[2] = {0, 0};
    if (p.nbig > 0) {
        tm.begin("row.items");
        const int *rowsb = p.bin_rows + st.s[4];
        p.item_cnt = ws.alloc_n<int64_t>((size_t)p.nbig + 1);
        int64_t *tmp2 = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)p.nbig + 1));
        // Narrow (8-sub-block) windows when the spans are wide enough that
        // the items will be run in column order: the items in flight then
        // read a band of B half as wide, which the L2 holds better; wide
        // windows otherwise (fewer items, fewer (A entry, item) pairs).
        const bool narrow = p.column_order;
        p.narrow_items = narrow;
        auto items_kernel = [&](auto kern, int64_t *cnt, ItemRec *it, int *itn) {
            hipLaunchKernelGGL(kern, dim3(grid_for((size_t)p.nbig * 64, 256)), dim3(256), 0, s,
                               p.nbig, rowsb, k, p.span_sb0, p.sb_off, p.subcnt, d_ic, cnt, it,
                               itn, p.subocc, band_off, band_base);
            TSG_LAUNCH_CHECK();
        };
        if (narrow)
            items_kernel(k_items<false, 8>, p.item_cnt, nullptr, nullptr);
        else
            items_kernel(k_items<false, 16>, p.item_cnt, nullptr, nullptr);
        exclusive_scan_i64(p.item_cnt, p.item_cnt, (size_t)p.nbig, tmp2, s);
        unsigned h_overflow = 0;
        TSG_HIP_TRY(hipMemcpyAsync(&h_tot[1], p.item_cnt + p.nbig, sizeof(int64_t),
                                   hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipMemcpyAsync(&h_tot[0], d_ic + m, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        if (d_overflow)
            TSG_HIP_TRY(hipMemcpyAsync(&h_overflow, d_overflow, sizeof(h_overflow),
                                       hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipStreamSynchronize(s));
        // the lists are sized by a bound on the outputs, so this holds unless
        // the symbolic phase itself went wrong
        TSG_CHECK(h_overflow == 0, TSG_ERR_OVERFLOW, "big-row word lists exceed their bound");
        p.nitems = h_tot[1];
        p.items = ws.alloc_n<ItemRec>((size_t)p.nitems + 1);
        int *item_n = ws.alloc_n<int>((size_t)p.nitems + 1);
        if (narrow)
            items_kernel(k_items<true, 8>, p.item_cnt, p.items, item_n);
        else
            items_kernel(k_items<true, 16>, p.item_cnt, p.items, item_n);
        p.nitems_small = 0;
        p.nitems_mid = 0;
        if (p.column_order && p.nitems > 1 && p.nitems < (int64_t(1) << 31)) {
            const size_t ni = (size_t)p.nitems;
            int *keys = ws.alloc_n<int>(ni);
            int *order = ws.alloc_n<int>(ni);
            const unsigned kblocks = grid_for(ni, 256);
            int *bsmall = ws.alloc_n<int>((size_t)kblocks + 1);
            int *bmid = ws.alloc_n<int>((size_t)kblocks + 1);
            int64_t *tmpk = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)kblocks + 1));
            int64_t *d_small = ws.alloc_n<int64_t>(2);
            int end_bit = 1;
            while (end_bit < 29 && (int64_t(1) << end_bit) <= ((int64_t)k >> SUB_SHIFT)) end_bit++;
            // the size class splits the rank path's items only
            hipLaunchKernelGGL(k_item_keys, dim3(kblocks), dim3(256), 0, s, p.nitems,
                               p.items, item_n, p.use_rank ? end_bit : -1, keys, bsmall, bmid);
            TSG_LAUNCH_CHECK();
            exclusive_scan_i32(bsmall, bsmall, (size_t)kblocks, tmpk, s, d_small);
            exclusive_scan_i32(bmid, bmid, (size_t)kblocks, tmpk, s, d_small + 1);
            sort_order_i32(keys, p.nitems, p.use_rank ? end_bit + 2 : end_bit, order, ws, s);
            ItemRec *items_s = ws.alloc_n<ItemRec>(ni + 1);
            hipLaunchKernelGGL(k_gather_items, dim3(grid_for(ni, 256)), dim3(256), 0, s,
                               p.nitems, order, p.items, items_s);
            TSG_LAUNCH_CHECK();
            p.items = items_s;
            if (p.use_rank) {
                int64_t h_cls[2] = {0, 0};
                TSG_HIP_TRY(hipMemcpyAsync(h_cls, d_small, sizeof(h_cls), hipMemcpyDeviceToHost,
                                           s));
                TSG_HIP_TRY(hipStreamSynchronize(s));
                p.nitems_small = h_cls[0];
                p.nitems_mid = h_cls[0] + h_cls[1];
            }
        }
        tm.end();
    } else {
        TSG_HIP_TRY(hipMemcpyAsync(&h_tot[0], d_ic + m, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipStreamSynchronize(s));
    }
    *nnz_c = h_tot[0];
    p.valid = true;
}

template <bool FUSED>
static void row_numeric_t(RowPlan &p, const CsrIn &A, const CsrIn &B, const int64_t *d_ic,
                          int *d_jc, double *d_c, hipStream_t s, StageTimer &tm) {
    const int *bins = p.bin_rows;
    const int *st = p.bin_start_h;
    tm.begin("row.small_numeric");
    {
        // rows of <= 64 and <= 256 products: 16 and 2 rows per block of the
        // hash-capable kernel; rows of <= 1024 products by the class the
        // symbolic phase gave them: the rank-only kernel's 1024- and 2048-word
        // variants, the hash-capable kernel for wider spans (1 row per block)
        auto launch = [&](auto kern, int rpb, int g, size_t per_row, const int *rows, int nr) {
            if (nr <= 0) return;
            const size_t sm = per_row * rpb;
            set_smem(kern, sm);
            hipLaunchKernelGGL(kern, dim3(grid_for(nr, rpb)), dim3(rpb * g), sm, s, nr, rows,
                               A.ptr, A.idx, A.val, B.ptr, B.idx, B.val, d_ic, d_jc, d_c,
                               (int)p.b_strict, p.span);
            TSG_LAUNCH_CHECK();
        };
        launch(k_small_numeric<16, 128, FUSED, 16>, 16, 16, small_numeric_row_bytes<16, 128>(),
               bins + st[1], st[2] - st[1]);
        launch(k_small_numeric<32, 512, FUSED, 2>, 2, 32, small_numeric_row_bytes<32, 512>(),
               bins + st[2], st[3] - st[2]);
        const int n3 = st[4] - st[3];
        if (n3 > 0) {
            // classes from the symbolic phase: [1 | ... | 0] in rows3,
            // [2 | ... | 3] in rows3r
            const int n_hash = p.n3cls[0], n_narrow = p.n3cls[1], n_rank = p.n3cls[2],
                      n_rank512 = p.n3cls[3];
            auto launch_rank = [&](auto kern, const int *rows, int nr) {
                if (nr <= 0) return;
                hipLaunchKernelGGL(kern, dim3((unsigned)nr), dim3(64), 0, s, nr, rows, A.ptr,
                                   A.idx, A.val, B.ptr, B.idx, B.val, d_ic, d_jc, d_c,
                                   (int)p.b_strict, p.span);
                TSG_LAUNCH_CHECK();
            };
            launch_rank(k_small_numeric_rank<FUSED, NARROW_WORDS, NARROW_OUT>, p.rows3, n_narrow);
            launch_rank(k_small_numeric_rank<FUSED, 2048, 1024>, p.rows3r, n_rank);
            launch_rank(k_small_numeric_rank<FUSED, 2048, NARROW_OUT>, p.rows3r + (n3 - n_rank512),
                        n_rank512);
            launch(k_small_numeric<64, 2048, FUSED, 1>, 1, 64, small_numeric_row_bytes<64, 2048>(),
                   p.rows3 + (n3 - n_hash), n_hash);
        }
    }
    tm.end();
    if (p.nitems > 0 && p.use_rank) {
        tm.begin("row.big_numeric");
        if (B.nnz > 0) {
            hipLaunchKernelGGL(k_pack_b, dim3(grid_for((size_t)B.nnz, 256)), dim3(256), 0, s,
                               B.nnz, B.idx, B.val, p.bpk);
            TSG_LAUNCH_CHECK();
        }
        auto launch_rank = [&](auto kern, size_t sm, int64_t i0, int64_t i1) {
            if (i1 <= i0) return;
            set_smem(kern, sm);
            const unsigned grid = item_grid(i1 - i0);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * ITEM_WAVES), sm, s, i1 - i0,
                               p.items + i0, p.bandwords, p.subocc, A.ptr, A.idx, A.val, B.idx, B.val, p.bpk, block_ptrs(p, B.ptr),
                               d_jc, d_c, (int)p.b_strict);
            TSG_LAUNCH_CHECK();
        };
        // the size classes (sorted in this order) in kernels with value
        // arrays of their size
        const int64_t ns = p.nitems_small, nm = std::max(p.nitems_mid, ns);
        if (p.narrow_items) {
            constexpr int W = ITEM_WORDS / 2;
            launch_rank(k_big_numeric_rank<FUSED, W, SMALL_ITEM, SMALL_RANK_U, SMALL_MINW>,
                        sizeof(RankSmem<W, SMALL_ITEM>) * ITEM_WAVES, 0, ns);
            launch_rank(k_big_numeric_rank<FUSED, W, MID_ITEM, MID_RANK_U, MID_MINW>,
                        sizeof(RankSmem<W, MID_ITEM>) * ITEM_WAVES, ns, nm);
            launch_rank(k_big_numeric_rank<FUSED, W>, sizeof(RankSmem<W>) * ITEM_WAVES, nm,
                        p.nitems);
        } else {
            constexpr int W = ITEM_WORDS;
            launch_rank(k_big_numeric_rank<FUSED, W, SMALL_ITEM, SMALL_RANK_U, SMALL_MINW>,
                        sizeof(RankSmem<W, SMALL_ITEM>) * ITEM_WAVES, 0, ns);
            launch_rank(k_big_numeric_rank<FUSED, W, MID_ITEM, MID_RANK_U, MID_MINW>,
                        sizeof(RankSmem<W, MID_ITEM>) * ITEM_WAVES, ns, nm);
            launch_rank(k_big_numeric_rank<FUSED, W>, sizeof(RankSmem<W>) * ITEM_WAVES, nm,
                        p.nitems);
        }
        tm.end();
    } else if (p.nitems > 0) {
        tm.begin("row.big_numeric");
        constexpr size_t sm = sizeof(ItemSmem) * ITEM_WAVES;
        auto kern = k_big_numeric<FUSED>;
        set_smem(kern, sm);
        const unsigned grid = item_grid(p.nitems);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * ITEM_WAVES), sm, s, p.nitems, p.items,
                           A.ptr, A.idx, A.val, B.ptr, B.idx, B.val, block_ptrs(p, B.ptr),
                           d_jc, d_c, (int)p.b_strict);
        TSG_LAUNCH_CHECK();
        tm.end();
    }
}

void row_numeric(RowPlan &p, const CsrIn &A, const CsrIn &B, const int64_t *d_ic, int *d_jc,
                 double *d_c, int fused, Arena & /*ws*/, hipStream_t s, StageTimer &tm,
                 int /*cus*/) {
    TSG_CHECK(p.valid && p.m == A.rows && p.n == A.cols && p.k == B.cols, TSG_ERR_STATE,
              "row numeric phase without a matching symbolic phase");
    if (fused)
        row_numeric_t<true>(p, A, B, d_ic, d_jc, d_c, s, tm);
    else
        row_numeric_t<false>(p, A, B, d_ic, d_jc, d_c, s, tm);
}

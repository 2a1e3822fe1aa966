// Flat product expansion, LDS hash tables, ordered accumulation and
// group-wide sorts shared by the row path and the tile path (gfx950, wave64).
#pragma once

#include "tsg_device.h"

// ------------------------------------------------- flat product expansion
// B block pointers: entry (row t, block b) = first position of B row t with
// column >= b << bs, at p[t*tstride + b*bstride].  Row-major (tstride =
// blocks+1, bstride = 1) keeps a row's entries together; block-major
// (tstride = 1, bstride = rows of B) keeps a block's entries for all rows
// together, which suits work that runs in column order: it looks up a few
// blocks at a time, whose entries stay in L2.
struct BlockPtrs {
    const int *p = nullptr;
    int64_t tstride = 0, bstride = 0;
    int bs = 0;
};

// Range of B row t restricted to columns >= c, through the block pointers
// (a binary search inside the block when c is not a block boundary).
__device__ __forceinline__ int brange_lo(const BlockPtrs &bp, const int *__restrict__ jb, int t,
                                         int c) {
    const int b = c >> bp.bs;
    const int64_t e = (int64_t)t * bp.tstride + (int64_t)b * bp.bstride;
    int lo = bp.p[e];
    if (c & ((1 << bp.bs) - 1)) {
        int hi = bp.p[e + bp.bstride];
        lo = lower_bound_i32(jb, lo, hi, c);
    }
    return lo;
}

// Products of one group of G A entries: lane l's A entry contributes the B
// positions [lo, lo + cnt) with value a.  Batches of G products are handed to
// f in product order (see for_products); bofs/bstep select every bstep-th
// batch starting at bofs.  s_mark: the group's G ints of LDS, all -1 on entry
// and on return.
// PACKED: B's entries are also given as 16-byte {column, -, value} records
// (bpk), so a product's column and value come from one load and, for a short
// range, one cache line.
template <int G, int U, bool VALUES, bool PACKED = false, class F>
__device__ __forceinline__ void expand_group(int lo, int cnt, double a,
                                             const int *__restrict__ jb,
                                             const double *__restrict__ bval, int *s_mark,
                                             F &&f, int bofs = 0, int bstep = 1,
                                             const int4 *__restrict__ bpk = nullptr) {
    static_assert(G >= 16, "segment lookup scans 16, 32 or 64 lanes");
    const int l = __lane_id() & (G - 1);
    const int gbase = __lane_id() & ~(G - 1);
    const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1);
    const int incl = group_incl_scan<G>(cnt);
    const int total = group_last<G>(incl);
    const int start = incl - cnt;
    // position in B of product q of source lane s: off_s + q
    const int off = lo - start;
    for (int base = bofs * G; base < total; base += G * U * bstep) {
        int e[U], sl[U];
        bool act[U];
        double as[U];
        // batches of this chunk holding products: the rest are skipped
        // (uniform over the group; with G == 64, over the wave)
        const int nb = (total - base + G * bstep - 1) / (G * bstep);
#pragma unroll
        for (int u = 0; u < U; u++) {
            act[u] = false;
            as[u] = 0.0;
            e[u] = 0;
            sl[u] = 0;
            if (u < nb) {
                const int qb = base + u * G * bstep;
                const int q = qb + l;
                act[u] = q < total;
                wave_lds_fence();
                if (cnt > 0 && start < qb + G && incl > qb)
                    s_mark[start > qb ? start - qb : 0] = l;
                wave_lds_fence();
                const int v = s_mark[l];
                s_mark[l] = -1;
                const int s = dpp_incl_max<G>(v);
                as[u] = VALUES ? __shfl(a, s, G) : 0.0;
                e[u] = __shfl(off, s, G) + q;
                sl[u] = s;
            }
        }
        int col[U];
        double bv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (PACKED) {
                const int4 rec = act[u] ? bpk[e[u]] : make_int4(0, 0, 0, 0);
                col[u] = rec.x;
                bv[u] = __hiloint2double(rec.w, rec.z);
            } else {
                col[u] = act[u] ? jb[e[u]] : 0;
                bv[u] = (VALUES && act[u]) ? bval[e[u]] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (base + u * G * bstep < total) {
                const int s0 = group_first<G>(sl[u]);
                const unsigned long long diff = __ballot(act[u] && sl[u] != s0);
                const bool single = ((diff >> gbase) & gmask) == 0;
                f(act[u], col[u], as[u], bv[u], single);
            }
        }
    }
}

// Visits the products of the A entries [pa, pe) of one row in A-row order,
// G lanes at a time, optionally restricted to B columns [c0, c1).  The
// A entries are taken in groups of G starting at pa + gofs, stride gstride.
// f(active, col, a, b, single) is called once per batch by every lane of the
// group; batches are issued in product order and lane order inside a batch is
// product order.  U batches are gathered before the first is handed to f, so
// each lane keeps U independent loads in flight.  single is true when all
// active lanes of the batch come from the same A entry (one B row: distinct
// columns whenever B's rows are duplicate-free).
// Batch striding (bofs, bstep): only batches bofs, bofs + bstep, ... of
// every group are visited, so several waves can share one group's products
// (used where the order of products does not matter).
// s_mark: the group's G ints of LDS, all -1 on entry and on return.  A
// batch finds the A entry of each of its products there: every non-empty
// segment (the products of one A entry) that overlaps the batch's window of
// G products writes its lane at its first position in the window, and a
// running maximum over the window gives each position its segment.
template <int G, int U, bool RANGED, bool VALUES, class F>
__device__ __forceinline__ void for_products(int pa, int pe, int gofs, int gstride,
                                             const int *__restrict__ ja,
                                             const double *__restrict__ aval,
                                             const int *__restrict__ ib,
                                             const int *__restrict__ jb,
                                             const double *__restrict__ bval,
                                             BlockPtrs bp, int c0, int c1, int *s_mark,
                                             F &&f, int bofs = 0,
                                             int bstep = 1) {
    static_assert(G >= 16, "segment lookup scans 16, 32 or 64 lanes");
    const int l = __lane_id() & (G - 1);
    // Software pipeline over groups of A entries: the A entry (t, a) of group
    // g+2 and the B range of group g+1 are loaded while group g's products
    // are processed.
    auto load_t = [&](int g, int &t, double &a) {
        const int p = g + l;
        t = -1;
        a = 0.0;
        if (g < pe && p < pe) {
            t = ja[p];
            if (VALUES) a = aval[p];
        }
    };
    auto load_range = [&](int t, int &lo, int &hi) {
        lo = 0;
        hi = 0;
        if (t >= 0) {
            if (RANGED) {
                lo = brange_lo(bp, jb, t, c0);
                hi = brange_lo(bp, jb, t, c1);
            } else {
                lo = ib[t];
                hi = ib[t + 1];
            }
        }
    };
    int g = pa + gofs;
    int t0, t1;
    double a0, a1;
    load_t(g, t0, a0);
    load_t(g + gstride, t1, a1);
    int lo, hi;
    load_range(t0, lo, hi);
    for (; g < pe; g += gstride) {
        const double a = a0;
        int lo_n, hi_n, t2;
        double a2;
        load_range(t1, lo_n, hi_n);                 // next group's B ranges
        load_t(g + 2 * gstride, t2, a2);            // group after next
        expand_group<G, U, VALUES>(lo, hi - lo, a, jb, bval, s_mark, f, bofs, bstep);
        lo = lo_n;
        hi = hi_n;
        t1 = t2;
        a0 = a1;
        a1 = a2;
    }
}

// Insert col into an LDS linear-probing table; returns the slot and whether
// this call created the key.  At most SLOTS probes: a full table (which the
// symbolic phase's sizing rules out) yields -1 instead of spinning.
template <int SLOTS>
__device__ __forceinline__ int hash_insert(int *keys, int col, bool &created) {
    unsigned h = hash_col(col, SLOTS - 1);
    created = false;
    for (int probe = 0; probe < SLOTS; probe++) {
        int k = keys[h];
        if (k == col) return (int)h;
        if (k == -1) {
            int old = atomicCAS(&keys[h], -1, col);
            if (old == -1) { created = true; return (int)h; }
            if (old == col) return (int)h;
        }
        h = (h + 1) & (SLOTS - 1);
    }
    return -1;
}

template <int SLOTS>
__device__ __forceinline__ int hash_find(const int *keys, int col) {
    unsigned h = hash_col(col, SLOTS - 1);
    for (int probe = 0; probe < SLOTS; probe++) {
        if (keys[h] == col) return (int)h;
        h = (h + 1) & (SLOTS - 1);
    }
    return -1;
}

// Ordered read-modify-write of vals[slot] by the active lanes of a batch:
// lanes sharing a tag bucket go one per round, lowest lane (earliest product)
// first, so every slot sees its products in product order.
// With single (all active lanes from one duplicate-free B row, hence distinct
// slots) the update is done in one step.
template <bool FUSED, int NTAGS>
__device__ __forceinline__ void ordered_accumulate(bool act, int slot, double a, double b,
                                                   double *vals, unsigned *tags,
                                                   unsigned me, bool single) {
    if (single) {
        if (act && slot >= 0) vals[slot] = acc_step<FUSED>(vals[slot], a, b);
        wave_lds_fence();
        return;
    }
    bool pending = act && slot >= 0;
    const unsigned bucket = (unsigned)slot & (NTAGS - 1);
    while (__ballot(pending)) {
        if (pending) atomicMin(&tags[bucket], me);
        wave_lds_fence();
        bool win = pending && tags[bucket] == me;
        wave_lds_fence();
        if (win) {
            vals[slot] = acc_step<FUSED>(vals[slot], a, b);
            tags[bucket] = 0xffffffffu;
            pending = false;
        }
        wave_lds_fence();
    }
}

// Bitonic sort of n (power of two) keys with values, by the G lanes of a
// group (G <= 64, so wave-synchronous) or by a whole block (GROUP_IS_BLOCK).
template <int G>
__device__ __forceinline__ void bitonic_sort_group(int *keys, double *vals, int n) {
    const int l = __lane_id() & (G - 1);
    for (int kk = 2; kk <= n; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = l; i < n; i += G) {
                int ixj = i ^ j;
                if (ixj > i) {
                    int ki = keys[i], kx = keys[ixj];
                    bool asc = (i & kk) == 0;
                    if ((ki > kx) == asc) {
                        keys[i] = kx;
                        keys[ixj] = ki;
                        double t = vals[i];
                        vals[i] = vals[ixj];
                        vals[ixj] = t;
                    }
                }
            }
            wave_lds_fence();
        }
    }
}

// Stable LSD radix sort of n keys (with values) by the G lanes of a group:
// 8-bit digits of (key - kmin) over `bits` bits; src/dst ping-pong buffers of
// n entries each; hist is a 256-entry LDS array owned by the group.  Each
// chunk of G entries is split by digit with ballots (lanes with equal digits
// keep their order), so the sort is stable.  Returns true when the sorted
// data ends in dst.
template <int G>
__device__ __forceinline__ bool radix_sort_group(int *skeys, double *svals, int *dkeys,
                                                 double *dvals, int n, int kmin, int bits,
                                                 unsigned *hist) {
    const int l = __lane_id() & (G - 1);
    const int gb = __lane_id() & ~(G - 1);
    const unsigned long long gm = G == 64 ? ~0ull : ((1ull << G) - 1);
    const unsigned long long below = (1ull << l) - 1;
    bool in_dst = false;
    for (int shift = 0; shift < bits; shift += 8) {
        int *ik = in_dst ? dkeys : skeys;
        double *iv = in_dst ? dvals : svals;
        int *ok = in_dst ? skeys : dkeys;
        double *ov = in_dst ? svals : dvals;
        for (int i = l; i < 256; i += G) hist[i] = 0u;
        wave_lds_fence();
        // histogram: one add per digit group of a chunk
        for (int c0 = 0; c0 < n; c0 += G) {
            const bool act = c0 + l < n;
            const unsigned d = act ? ((unsigned)(ik[c0 + l] - kmin) >> shift) & 255u : 0u;
            unsigned long long m = (__ballot(act) >> gb) & gm;
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const unsigned long long bb = (__ballot((d >> b) & 1u) >> gb) & gm;
                m &= ((d >> b) & 1u) ? bb : ~bb;
            }
            if (act && (m & below) == 0) hist[d] += (unsigned)__popcll(m);
            wave_lds_fence();
        }
        // exclusive scan of the 256 counts: 256/G consecutive bins per lane,
        // then one group scan of the lanes' sums
        {
            constexpr int PER = 256 / G;
            int cnt[PER];
            int sum = 0;
#pragma unroll
            for (int q = 0; q < PER; q++) {
                cnt[q] = (int)hist[l * PER + q];
                sum += cnt[q];
            }
            int run = group_incl_scan<G>(sum) - sum;
#pragma unroll
            for (int q = 0; q < PER; q++) {
                hist[l * PER + q] = (unsigned)run;
                run += cnt[q];
            }
        }
        wave_lds_fence();
        // stable scatter
        for (int c0 = 0; c0 < n; c0 += G) {
            const bool act = c0 + l < n;
            const int key = act ? ik[c0 + l] : 0;
            const double val = act ? iv[c0 + l] : 0.0;
            const unsigned d = act ? ((unsigned)(key - kmin) >> shift) & 255u : 0u;
            unsigned long long m = (__ballot(act) >> gb) & gm;
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const unsigned long long bb = (__ballot((d >> b) & 1u) >> gb) & gm;
                m &= ((d >> b) & 1u) ? bb : ~bb;
            }
            const int rank = __popcll(m & below);
            const unsigned base = act ? hist[d] : 0u;
            wave_lds_fence();
            if (act) {
                ok[base + rank] = key;
                ov[base + rank] = val;
                if ((m >> l) == 1ull) hist[d] = base + (unsigned)__popcll(m);   // last of its digit
            }
            wave_lds_fence();
        }
        in_dst = !in_dst;
    }
    return in_dst;
}

__device__ __forceinline__ void bitonic_sort_keys_group64(int *keys, int n) {
    const int l = __lane_id();
    for (int kk = 2; kk <= n; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = l; i < n; i += 64) {
                int ixj = i ^ j;
                if (ixj > i) {
                    int ki = keys[i], kx = keys[ixj];
                    bool asc = (i & kk) == 0;
                    if ((ki > kx) == asc) { keys[i] = kx; keys[ixj] = ki; }
                }
            }
            wave_lds_fence();
        }
    }
}


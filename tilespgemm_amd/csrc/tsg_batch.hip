// Batched products: C_i = A_i * B_i for a batch of independent products,
// laid out as one block-diagonal product diag(A_i) * diag(B_i) = diag(C_i).
// The structure of the concatenated operands is gathered once in the
// symbolic phase; the numeric phase only gathers values, runs the single
// product's numeric kernels and makes C's columns local to each product, so
// it issues kernels only and can be captured in a HIP graph.
#include "tsg_internal.h"
#include "tsg_device.h"

#include <algorithm>
#include <cstring>

namespace {

struct BDesc {
    const int *ia, *ja;
    const double *a;
    const int *ib, *jb;
    const double *b;
};

// Offsets per product (count+1 entries each): rows of A_i / C_i, rows of B_i
// (= columns of A_i), columns of B_i / C_i, entries of A_i, entries of B_i.
enum { OFF_ROW = 0, OFF_BROW = 1, OFF_BCOL = 2, OFF_A = 3, OFF_B = 4, OFF_N = 5 };

// Largest p in [0, count) with off[p] <= x.
__device__ __forceinline__ int find_seg(const int64_t *__restrict__ off, int count, int64_t x) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Row pointers of the concatenated A (B = false) or B (B = true), plus the
// closing entry.  A block takes CAT_PER * 256 consecutive rows, thread t
// those at t, t + 256, ...; the product of the block's first row is searched
// once and each thread steps to the next product at a boundary (products
// without rows are stepped over).
constexpr int CAT_PER = 8;
template <bool B>
__global__ void __launch_bounds__(256) k_cat_rows(int count, const BDesc *__restrict__ desc,
                                                  const int64_t *__restrict__ off, int64_t rows,
                                                  int *__restrict__ out_ptr) {
    const int64_t r0 = (int64_t)blockIdx.x * (256 * CAT_PER);
    if (r0 > rows) return;
    const int64_t *roff = off + (B ? OFF_BROW : OFF_ROW) * (int64_t)(count + 1);
    const int64_t *eoff = off + (B ? OFF_B : OFF_A) * (int64_t)(count + 1);
    __shared__ int s_p;
    if (threadIdx.x == 0) s_p = r0 < rows ? find_seg(roff, count, r0) : 0;
    __syncthreads();
    int p = s_p;
#pragma unroll
    for (int j = 0; j < CAT_PER; j++) {
        const int64_t r = r0 + j * 256 + threadIdx.x;
        if (r > rows) break;
        if (r == rows) {
            out_ptr[r] = (int)eoff[count];
            break;
        }
        while (r >= roff[p + 1]) p++;   // stops below count: r < roff[count]
        const int *ptr = B ? desc[p].ib : desc[p].ia;
        out_ptr[r] = (int)(eoff[p] + ptr[r - roff[p]]);
    }
}

// Entries of the concatenated A (B = false) or B (B = true).  A's columns
// index rows of the concatenated B, so they move by the product's B-row
// offset; B's columns move by the product's column offset unless LOCAL (each
// product's C then comes out with its own columns).
// A block takes CAT_PER * 256 consecutive entries, thread t those at t,
// t + 256, ... (coalesced); the product of the block's first entry is
// searched once, and each thread steps to the next product at a boundary
// (products are contiguous and far longer than a block on average).
template <bool B, bool VALUES, bool LOCAL = false>
__global__ void __launch_bounds__(256) k_cat_entries(int count, const BDesc *__restrict__ desc,
                                                     const int64_t *__restrict__ off,
                                                     int64_t nnz, int *__restrict__ out_idx,
                                                     double *__restrict__ out_val) {
    const int64_t q0 = (int64_t)blockIdx.x * (256 * CAT_PER);
    if (q0 >= nnz) return;
    const int64_t *eoff = off + (B ? OFF_B : OFF_A) * (int64_t)(count + 1);
    const int64_t *coff = off + (B ? OFF_BCOL : OFF_BROW) * (int64_t)(count + 1);
    __shared__ int s_p;
    if (threadIdx.x == 0) s_p = find_seg(eoff, count, q0);
    __syncthreads();
    int p = s_p;
#pragma unroll
    for (int j = 0; j < CAT_PER; j++) {
        const int64_t q = q0 + j * 256 + threadIdx.x;
        if (q >= nnz) break;
        while (q >= eoff[p + 1]) p++;   // stops below count: q < eoff[count]
        const int64_t local = q - eoff[p];
        if (VALUES) {
            out_val[q] = B ? desc[p].b[local] : desc[p].a[local];
        } else {
            const int *idx = B ? desc[p].jb : desc[p].ja;
            out_idx[q] = (B && LOCAL) ? idx[local] : (int)(idx[local] + coff[p]);
        }
    }
}

static unsigned cat_grid(int64_t nnz) { return grid_for((size_t)nnz, 256 * CAT_PER); }

// Makes C's columns local to each product: 8 lanes per row of C.
__global__ void __launch_bounds__(256) k_local_cols(int count, const int64_t *__restrict__ off,
                                                    int64_t rows, const int64_t *__restrict__ ic,
                                                    int *__restrict__ jc) {
    const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const int l = threadIdx.x & 7;
    if (r >= rows) return;
    const int p = find_seg(off + OFF_ROW * (int64_t)(count + 1), count, r);
    const int sub = (int)off[OFF_BCOL * (int64_t)(count + 1) + p];
    if (sub == 0) return;
    for (int64_t e = ic[r] + l; e < ic[r + 1]; e += 8) jc[e] -= sub;
}

__global__ void k_gather_row_ptrs(int count, const int64_t *__restrict__ off,
                                  const int64_t *__restrict__ ic, int64_t *__restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= count) out[p] = ic[off[OFF_ROW * (int64_t)(count + 1) + p]];
}

// Per-product row pointers starting at 0: product i's m_i+1 entries go to
// out[row_off[i] + i ...].
__global__ void k_rebase(int count, const int64_t *__restrict__ off, int64_t rows,
                         const int64_t *__restrict__ ic, int64_t *__restrict__ out) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= rows + count) return;
    const int64_t *roff = off + OFF_ROW * (int64_t)(count + 1);
    int lo = 0, hi = count - 1;                 // largest i with roff[i] + i <= o
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (roff[mid] + mid <= o) lo = mid; else hi = mid - 1;
    }
    const int64_t r = roff[lo] + (o - (roff[lo] + lo));
    out[o] = ic[r] - ic[roff[lo]];
}

bool same_product(const tsg_product_t &x, const tsg_product_t &y) {
    return x.m == y.m && x.n == y.n && x.k == y.k && x.ia == y.ia && x.ja == y.ja &&
           x.a == y.a && x.nnz_a == y.nnz_a && x.ib == y.ib && x.jb == y.jb && x.b == y.b &&
           x.nnz_b == y.nnz_b;
}

}  // namespace

void batch_symbolic(tsg_handle_t h, int count, const tsg_product_t *prods, int64_t *d_ic,
                    int64_t *nnz_c_total, int64_t *nnz_c) {
    TSG_CHECK(count > 0 && prods && d_ic && nnz_c_total, TSG_ERR_INVALID, "bad batch arguments");
    BatchState &S = h->batch;
    S.valid = false;
    std::vector<int64_t> off((size_t)OFF_N * (count + 1), 0);
    auto O = [&](int which, int p) -> int64_t & { return off[(size_t)which * (count + 1) + p]; };
    bool b_is_a = true;
    for (int p = 0; p < count; p++) {
        const tsg_product_t &x = prods[p];
        TSG_CHECK(x.m >= 0 && x.n >= 0 && x.k >= 0 && x.nnz_a >= 0 && x.nnz_b >= 0,
                  TSG_ERR_INVALID, "batch product " + std::to_string(p) + ": negative size");
        TSG_CHECK(x.ia && x.ib && (x.nnz_a == 0 || (x.ja && x.a)) &&
                      (x.nnz_b == 0 || (x.jb && x.b)),
                  TSG_ERR_INVALID, "batch product " + std::to_string(p) + ": null pointer");
        O(OFF_ROW, p + 1) = O(OFF_ROW, p) + x.m;
        O(OFF_BROW, p + 1) = O(OFF_BROW, p) + x.n;
        O(OFF_BCOL, p + 1) = O(OFF_BCOL, p) + x.k;
        O(OFF_A, p + 1) = O(OFF_A, p) + x.nnz_a;
        O(OFF_B, p + 1) = O(OFF_B, p) + x.nnz_b;
        b_is_a = b_is_a && x.ib == x.ia && x.jb == x.ja && x.b == x.a && x.m == x.n &&
                 x.n == x.k && x.nnz_a == x.nnz_b;
    }
    const int64_t lim = 0x7fffffffLL;
    TSG_CHECK(O(OFF_ROW, count) < lim && O(OFF_BROW, count) < lim && O(OFF_BCOL, count) < lim &&
                  O(OFF_A, count) < lim && O(OFF_B, count) < lim,
              TSG_ERR_OVERFLOW, "batch exceeds int32 rows, columns or entries");
    int max_k = 0;
    for (int p = 0; p < count; p++) max_k = std::max(max_k, prods[p].k);
    S.count = count;
    S.b_is_a = b_is_a;
    S.rows = (int)O(OFF_ROW, count);
    S.cols_a = (int)O(OFF_BROW, count);
    S.cols_b = (int)O(OFF_BCOL, count);
    S.nnz_a = O(OFF_A, count);
    S.nnz_b = O(OFF_B, count);
    S.host.assign(prods, prods + count);

    std::vector<BDesc> desc(count);
    for (int p = 0; p < count; p++)
        desc[p] = BDesc{prods[p].ia, prods[p].ja, prods[p].a, prods[p].ib, prods[p].jb, prods[p].b};

    hipStream_t s = h->stream;
    h->bws.reset();
    S.desc = h->bws.alloc(sizeof(BDesc) * count);
    S.row_off = h->bws.alloc_n<int64_t>(off.size());
    S.ia = h->bws.alloc_n<int>((size_t)S.rows + 1);
    S.ja = h->bws.alloc_n<int>((size_t)S.nnz_a + 1);
    S.a = h->bws.alloc_n<double>((size_t)S.nnz_a + 1);
    // B keeps its own arrays except when every product is A_i * A_i: then its
    // row pointers and values are A's, and so are its columns unless they are
    // kept local (below)
    if (b_is_a) {
        S.ib = S.ia;
        S.b = S.a;
    } else {
        S.ib = h->bws.alloc_n<int>((size_t)S.cols_a + 1);
        S.b = h->bws.alloc_n<double>((size_t)S.nnz_b + 1);
    }
    int64_t *d_pc = h->bws.alloc_n<int64_t>((size_t)count + 1);
    TSG_HIP_TRY(hipMemcpyAsync(S.desc, desc.data(), sizeof(BDesc) * count, hipMemcpyHostToDevice, s));
    TSG_HIP_TRY(hipMemcpyAsync(S.row_off, off.data(), sizeof(int64_t) * off.size(),
                               hipMemcpyHostToDevice, s));
    const BDesc *dd = static_cast<const BDesc *>(S.desc);
    hipLaunchKernelGGL(k_cat_rows<false>, dim3(cat_grid((int64_t)S.rows + 1)), dim3(256), 0, s,
                       count, dd, S.row_off, (int64_t)S.rows, S.ia);
    TSG_LAUNCH_CHECK();
    if (S.nnz_a > 0) {
        hipLaunchKernelGGL((k_cat_entries<false, false>), dim3(cat_grid(S.nnz_a)),
                           dim3(256), 0, s, count, dd, S.row_off, S.nnz_a, S.ja, nullptr);
        TSG_LAUNCH_CHECK();
    }
    if (!b_is_a) {
        hipLaunchKernelGGL(k_cat_rows<true>, dim3(cat_grid((int64_t)S.cols_a + 1)), dim3(256),
                           0, s, count, dd, S.row_off, (int64_t)S.cols_a, S.ib);
        TSG_LAUNCH_CHECK();
    }
    // Method of the whole batch, decided before B's columns are laid out: on
    // the row path every row of A_i reaches only rows of B_i, so B keeps each
    // product's own columns and C comes out with them (and the column range,
    // which sizes the row path's structures, is max k_i, not sum k_i); the
    // tile path keeps the block-diagonal columns (its B tile columns would
    // otherwise merge across products) and C's columns are made local after.
    // (AUTO with tile-dense A keeps AUTO: the symbolic phase still sends
    // unsorted input to the row path, whose result is then made local too)
    int method = h->method;
    if (method == TSG_METHOD_AUTO) {
        const CsrIn Acat{S.rows, S.cols_a, S.ia, S.ja, nullptr, S.nnz_a};
        if (S.nnz_a == 0 || tile_density(Acat, h->tile_density, h->bws, s) < h->tile_density)
            method = TSG_METHOD_ROW;
    }
    S.local_cols = method == TSG_METHOD_ROW;
    if (S.local_cols) S.cols_b = max_k;
    if (b_is_a && !S.local_cols)
        S.jb = S.ja;
    else
        S.jb = h->bws.alloc_n<int>((size_t)S.nnz_b + 1);
    if (S.nnz_b > 0 && S.jb != S.ja) {
        if (S.local_cols)
            hipLaunchKernelGGL((k_cat_entries<true, false, true>),
                               dim3(cat_grid(S.nnz_b)), dim3(256), 0, s, count, dd,
                               S.row_off, S.nnz_b, S.jb, nullptr);
        else
            hipLaunchKernelGGL((k_cat_entries<true, false, false>),
                               dim3(cat_grid(S.nnz_b)), dim3(256), 0, s, count, dd,
                               S.row_off, S.nnz_b, S.jb, nullptr);
        TSG_LAUNCH_CHECK();
    }
    int64_t tot = 0;
    const int saved_method = h->method;
    h->method = method;
    int rc = tsg_spgemm_symbolic(h, S.rows, S.cols_a, S.cols_b, S.ia, S.ja, S.nnz_a, S.ib, S.jb,
                                 S.nnz_b, d_ic, &tot);
    h->method = saved_method;
    if (rc != TSG_OK) throw TsgError{rc, h->err};
    *nnz_c_total = tot;
    if (nnz_c) {
        hipLaunchKernelGGL(k_gather_row_ptrs, dim3(grid_for((size_t)count + 1, 256)), dim3(256), 0,
                           s, count, S.row_off, d_ic, d_pc);
        TSG_LAUNCH_CHECK();
        std::vector<int64_t> pc((size_t)count + 1);
        TSG_HIP_TRY(hipMemcpyAsync(pc.data(), d_pc, sizeof(int64_t) * (count + 1),
                                   hipMemcpyDeviceToHost, s));
        TSG_HIP_TRY(hipStreamSynchronize(s));
        for (int p = 0; p < count; p++) nnz_c[p] = pc[p + 1] - pc[p];
    }
    S.valid = true;
}

void batch_numeric(tsg_handle_t h, int count, const tsg_product_t *prods, const int64_t *d_ic,
                   int *d_jc, double *d_c) {
    BatchState &S = h->batch;
    TSG_CHECK(S.valid && count == S.count && prods, TSG_ERR_STATE,
              "batch numeric phase without a matching batch symbolic phase");
    for (int p = 0; p < count; p++)
        TSG_CHECK(same_product(prods[p], S.host[p]), TSG_ERR_STATE,
                  "batch product " + std::to_string(p) + " differs from its symbolic phase");
    hipStream_t s = h->stream;
    const BDesc *dd = static_cast<const BDesc *>(S.desc);
    if (S.nnz_a > 0) {
        hipLaunchKernelGGL((k_cat_entries<false, true>), dim3(cat_grid(S.nnz_a)),
                           dim3(256), 0, s, count, dd, S.row_off, S.nnz_a, nullptr, S.a);
        TSG_LAUNCH_CHECK();
    }
    if (!S.b_is_a && S.nnz_b > 0) {
        hipLaunchKernelGGL((k_cat_entries<true, true>), dim3(cat_grid(S.nnz_b)),
                           dim3(256), 0, s, count, dd, S.row_off, S.nnz_b, nullptr, S.b);
        TSG_LAUNCH_CHECK();
    }
    int rc = tsg_spgemm_numeric(h, S.rows, S.cols_a, S.cols_b, S.ia, S.ja, S.a, S.nnz_a, S.ib,
                                S.jb, S.b, S.nnz_b, d_ic, d_jc, d_c);
    if (rc != TSG_OK) throw TsgError{rc, h->err};
    if (S.rows > 0 && !S.local_cols) {
        hipLaunchKernelGGL(k_local_cols, dim3(grid_for((size_t)S.rows * 8, 256)), dim3(256), 0, s,
                           count, S.row_off, (int64_t)S.rows, d_ic, d_jc);
        TSG_LAUNCH_CHECK();
    }
}

void batch_rebase(tsg_handle_t h, const int64_t *d_ic, int64_t *d_out) {
    BatchState &S = h->batch;
    TSG_CHECK(S.valid, TSG_ERR_STATE, "no batch");
    const int64_t total = (int64_t)S.rows + S.count;
    hipLaunchKernelGGL(k_rebase, dim3(grid_for((size_t)total, 256)), dim3(256), 0, h->stream,
                       S.count, S.row_off, (int64_t)S.rows, d_ic, d_out);
    TSG_LAUNCH_CHECK();
}

// C ABI of libtsg.so: handles, the device two-phase SpGEMM, the
// reference-compatible host entry thmkl_dcsrmultcsr, plan inspection.
#include "tsg_internal.h"

#include <algorithm>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>


// Scratch of tsg_spgemm_noalloc: every workspace array of its symbolic and
// numeric phases lives here instead of in hipMalloc'd memory.  Each device
// the library runs on holds its own copy until the process ends.  Host code
// never uses this name as an address: tsg_create looks up each device's copy
// with hipGetSymbolAddress (scratch_setup) and only that pointer is passed on.
static __device__ __align__(256) unsigned char scratch[TSG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
// per device: the address of its copy of `scratch`, and a lock that
// serialises tsg_spgemm_noalloc calls on that device
void *g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_scratch_mu[MAX_DEVICES];

// Looks up the current device's copy of `scratch` (once per device).
void scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu[device]);
    if (g_scratch_ptr[device]) return;
    void *ptr = nullptr;
    TSG_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    TSG_CHECK(ptr != nullptr, TSG_ERR_HIP, "hipGetSymbolAddress returned no address for scratch");
    g_scratch_ptr[device] = ptr;
}

std::mutex g_err_mu;
std::string g_err;

void set_global_error(const std::string &e) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = e;
}

// Makes the handle's device current for the duration of a call.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) TSG_HIP_TRY(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

template <class F>
int guarded(tsg_handle_t h, F &&f) {
    try {
        f();
        if (h) h->err.clear();
        return TSG_OK;
    } catch (const TsgError &e) {
        if (h) h->err = e.msg;
        set_global_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        if (h) h->err = "host allocation failed";
        set_global_error("host allocation failed");
        return TSG_ERR_ALLOC;
    } catch (...) {
        if (h) h->err = "unknown error";
        set_global_error("unknown error");
        return TSG_ERR_INVALID;
    }
}

void check_csr_args(int m, int n, const int *ia, const int *ja, int64_t nnz, const char *what) {
    TSG_CHECK(m >= 0 && n >= 0 && nnz >= 0, TSG_ERR_INVALID,
              std::string(what) + ": negative size");
    TSG_CHECK(ia != nullptr, TSG_ERR_INVALID, std::string(what) + ": null row pointer");
    TSG_CHECK(nnz == 0 || ja != nullptr, TSG_ERR_INVALID, std::string(what) + ": null columns");
    TSG_CHECK(nnz < (int64_t(1) << 31), TSG_ERR_OVERFLOW,
              std::string(what) + ": inputs are limited to 2^31-1 entries (int32 indices)");
}

int choose_method(tsg_handle_t h, const CsrIn &A, bool sorted,
                  const unsigned long long *sample, Arena &ws) {
    if (h->method == TSG_METHOD_TILE)
        TSG_CHECK(sorted, TSG_ERR_INVALID,
                  "the tile path needs column-sorted, duplicate-free rows");
    if (h->method != TSG_METHOD_AUTO) return h->method;
    if (!sorted) return TSG_METHOD_ROW;
    if (A.nnz == 0 || A.rows == 0) return TSG_METHOD_ROW;
    // Tile path when A's 16x16 tiles hold enough entries to amortize the
    // per-tile metadata (~60 bytes per tile), row path otherwise.
    const double per_tile = tile_density(A, h->tile_density, ws, h->stream, sample);
    return per_tile >= h->tile_density ? TSG_METHOD_TILE : TSG_METHOD_ROW;
}

}  // namespace

int device_cu_count(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 256;
    return prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
}

namespace {

// Symbolic phase of C = A*B on the handle's stream and settings, into the
// given plans: validation (with AUTO's density sample), the method choice,
// then the tile or the row path.  The plan's arrays come from ws; aux holds
// temporaries of B's row sort.  Arguments are checked by the caller.
void symbolic_into(tsg_handle_t h, RowPlan &row, TilePlan &tile, int &method_used, Arena &ws,
                   Arena &aux, StageTimer &tm, const RowTuning &tune, int m, int n, int k,
                   const int *d_ia, const int *d_ja, int64_t nnz_a, const int *d_ib,
                   const int *d_jb, int64_t nnz_b, int64_t *d_ic, int64_t *nnz_c) {
    CsrIn A{m, n, d_ia, d_ja, nullptr, nnz_a};
    CsrIn B{n, k, d_ib, d_jb, nullptr, nnz_b};
    // with AUTO, A's tile-density sample shares the validation's round trip;
    // with the tile path forced, the tile counts of A and B do
    unsigned long long sample[2] = {0, 0};
    const bool want_sample = h->method == TSG_METHOD_AUTO;
    const bool b_is_a = d_ia == d_ib && d_ja == d_jb && m == n && n == k;
    EarlyTileCount earlyA, earlyB;
    std::function<void()> early;
    if (h->method == TSG_METHOD_TILE)
        early = [&] {
            tile_count_early(A, earlyA, ws, h->stream);
            if (!b_is_a) tile_count_early(B, earlyB, ws, h->stream);
        };
    const int vf =
        validate_csr_pair(A, B, ws, h->stream, want_sample ? sample : nullptr, early);
    TSG_CHECK((vf & 0x303) == 0, TSG_ERR_INVALID,
              std::string("invalid CSR input: ") + ((vf & 0x3) ? "A" : "B") +
                  ((vf & 0x101) ? " has bad row pointers" : " has a column out of range"));
    int method = choose_method(h, A, (vf & 0x404) == 0, want_sample ? sample : nullptr, ws);
    method_used = method;
    if (method == TSG_METHOD_TILE) {
        try {
            tile_symbolic(tile, A, B, b_is_a, d_ic, nnz_c, ws, h->stream, tm, h->cus, &earlyA,
                          &earlyB);
        } catch (const TsgError &e) {
            // AUTO's choice falls back to the row path when the tile path's
            // int32 tile indices would overflow or its structures do not fit
            // in memory; a forced tile path reports the error
            if ((e.code != TSG_ERR_OVERFLOW && e.code != TSG_ERR_ALLOC) ||
                h->method != TSG_METHOD_AUTO)
                throw;
            TSG_HIP_TRY(hipStreamSynchronize(h->stream));
            tile.valid = false;
            method = TSG_METHOD_ROW;
            method_used = method;
        }
    }
    if (method == TSG_METHOD_ROW) {
        // The row path reads every B row in column order (spans from a row's
        // first and last column, block pointers by search), so B rows that
        // are not strictly ascending are replaced by a stably sorted copy:
        // equal columns keep their storage order, so every output still
        // accumulates its products in the reference's order.
        const bool b_validated = B.ptr != A.ptr || B.idx != A.idx || B.cols != A.cols;
        bool b_strict = b_validated ? (vf & 0x400) == 0 : (vf & 0x4) == 0;
        CsrIn Bs = B;
        int *b_idx = nullptr, *b_perm = nullptr;
        double *b_val = nullptr;
        if (!b_strict && nnz_b > 0) {
            b_idx = ws.alloc_n<int>((size_t)nnz_b);
            b_perm = ws.alloc_n<int>((size_t)nnz_b);
            b_val = ws.alloc_n<double>((size_t)nnz_b);
            csr_sort_rows(n, k, d_ib, d_jb, nnz_b, b_idx, b_perm, ws, aux, h->stream);
            Bs.idx = b_idx;
            // duplicates remain after sorting: strict only if none
            b_strict = (validate_csr_pair(Bs, Bs, aux, h->stream) & 0x4) == 0;
        }
        row_symbolic(row, A, Bs, d_ic, nnz_c, ws, h->stream, tm, h->cus, tune);
        row.b_strict = b_strict;
        row.b_canon = b_idx != nullptr;
        row.b_idx = b_idx;
        row.b_perm = b_perm;
        row.b_val = b_val;
    }
}

// Numeric phase from plans made by symbolic_into (same structure; values
// may differ).  Arguments are checked by the caller.
void numeric_from(tsg_handle_t h, int method_used, RowPlan &row, TilePlan &tile, Arena &ws,
                  StageTimer &tm, int m, int n, int k, const int *d_ia, const int *d_ja,
                  const double *d_a, int64_t nnz_a, const int *d_ib, const int *d_jb,
                  const double *d_b, int64_t nnz_b, const int64_t *d_ic, int *d_jc,
                  double *d_c) {
    CsrIn A{m, n, d_ia, d_ja, d_a, nnz_a};
    CsrIn B{n, k, d_ib, d_jb, d_b, nnz_b};
    const int fused = h->fp_mode == TSG_FP_FUSED;
    if (method_used == TSG_METHOD_TILE) {
        TSG_CHECK(tile.valid, TSG_ERR_STATE, "numeric phase without a symbolic phase");
        tile_numeric(tile, A, B, d_ic, d_jc, d_c, fused, ws, h->stream, tm, h->cus);
    } else {
        TSG_CHECK(row.valid, TSG_ERR_STATE, "numeric phase without a symbolic phase");
        CsrIn Bs = B;
        if (row.b_canon) {
            gather_values(nnz_b, row.b_perm, d_b, row.b_val, h->stream);
            Bs.idx = row.b_idx;
            Bs.val = row.b_val;
        }
        row_numeric(row, A, Bs, d_ic, d_jc, d_c, fused, ws, h->stream, tm, h->cus);
    }
}

}  // namespace

extern "C" {

int tsg_create(tsg_handle_t *handle, int device) {
    if (!handle) return TSG_ERR_INVALID;
    *handle = nullptr;
    return guarded(nullptr, [&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
            (void)hipGetLastError();
            throw TsgError{TSG_ERR_NODEVICE, "no HIP device available"};
        }
        TSG_CHECK(device >= 0 && device < ndev, TSG_ERR_INVALID, "device index out of range");
        std::unique_ptr<tsg_handle_s> h(new tsg_handle_s());
        h->device = device;
        DeviceGuard dg(device);
        h->cus = device_cu_count(device);
        // A blocking stream: it orders with the null stream, which is where
        // callers that never create streams (and torch's default stream) work.
        TSG_HIP_TRY(hipStreamCreate(&h->own_stream));
        h->stream = h->own_stream;
        h->tm.stream = h->stream;
        scratch_setup(device);
        *handle = h.release();
    });
}

int tsg_destroy(tsg_handle_t h) {
    if (!h) return TSG_OK;
    {
        DeviceGuard dg(h->device);
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        h->tm.clear();
        h->ws.release();
        h->bws.release();
        h->aux.release();
        if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    }
    delete h;
    return TSG_OK;
}

int tsg_set_stream(tsg_handle_t h, void *stream) {
    if (!h) return TSG_ERR_INVALID;
    h->stream = stream ? static_cast<hipStream_t>(stream) : h->own_stream;
    h->tm.stream = h->stream;
    return TSG_OK;
}

int tsg_set_method(tsg_handle_t h, int method) {
    if (!h || method < TSG_METHOD_AUTO || method > TSG_METHOD_ROW) return TSG_ERR_INVALID;
    h->method = method;
    return TSG_OK;
}

int tsg_set_tile_density(tsg_handle_t h, double entries_per_tile) {
    if (!h || !(entries_per_tile > 0.0)) return TSG_ERR_INVALID;
    h->tile_density = entries_per_tile;
    return TSG_OK;
}

int tsg_set_row_bins(tsg_handle_t h, int bin1, int bin2, int bin3) {
    if (!h) return TSG_ERR_INVALID;
    if (bin1 < 0 || bin1 > 64 || bin2 < bin1 || bin2 > 256 || bin3 < bin2 || bin3 > 1024)
        return TSG_ERR_INVALID;
    h->row_tune.bin1 = bin1;
    h->row_tune.bin2 = bin2;
    h->row_tune.bin3 = bin3;
    return TSG_OK;
}

int tsg_set_item_order(tsg_handle_t h, int64_t min_items) {
    if (!h || min_items < 0) return TSG_ERR_INVALID;
    h->row_tune.item_sort_min = min_items;
    return TSG_OK;
}

int tsg_set_bitmap_budget(tsg_handle_t h, int64_t bytes) {
    if (!h) return TSG_ERR_INVALID;
    h->row_tune.bitmap_budget = bytes < 0 ? -1 : bytes;
    return TSG_OK;
}

int tsg_set_block_pointer_budget(tsg_handle_t h, int64_t bytes) {
    if (!h) return TSG_ERR_INVALID;
    h->row_tune.bptr_budget = bytes < 0 ? -1 : bytes;
    return TSG_OK;
}

int tsg_get_item_order(tsg_handle_t h, int64_t *min_items) {
    if (!h || !min_items) return TSG_ERR_INVALID;
    *min_items = h->row_tune.item_sort_min;
    return TSG_OK;
}

int tsg_get_tuning(tsg_handle_t h, double *tile_density, int *bin1, int *bin2, int *bin3) {
    if (!h) return TSG_ERR_INVALID;
    if (tile_density) *tile_density = h->tile_density;
    if (bin1) *bin1 = h->row_tune.bin1;
    if (bin2) *bin2 = h->row_tune.bin2;
    if (bin3) *bin3 = h->row_tune.bin3;
    return TSG_OK;
}

int tsg_set_fp_mode(tsg_handle_t h, int fp_mode) {
    if (!h || (fp_mode != TSG_FP_FUSED && fp_mode != TSG_FP_SEPARATE)) return TSG_ERR_INVALID;
    h->fp_mode = fp_mode;
    return TSG_OK;
}

int tsg_get_method_used(tsg_handle_t h) { return h ? h->method_used : TSG_ERR_INVALID; }

size_t tsg_workspace_bytes(tsg_handle_t h) { return h ? h->ws.held() : 0; }

int tsg_enable_timing(tsg_handle_t h, int enable) {
    if (!h) return TSG_ERR_INVALID;
    h->tm.enabled = enable != 0;
    return TSG_OK;
}

int tsg_stage_count(tsg_handle_t h) {
    if (!h) return 0;
    if (h->tm.ms.size() != h->tm.names.size()) {
        if (guarded(h, [&] { DeviceGuard dg(h->device); h->tm.finalize(); }) != TSG_OK) return 0;
    }
    return (int)h->tm.names.size();
}

const char *tsg_stage_name(tsg_handle_t h, int i) {
    if (!h || i < 0 || i >= (int)h->tm.names.size()) return "";
    return h->tm.names[i].c_str();
}

double tsg_stage_ms(tsg_handle_t h, int i) {
    if (!h || i < 0) return -1.0;
    if (h->tm.ms.size() != h->tm.names.size()) tsg_stage_count(h);
    if (i >= (int)h->tm.ms.size()) return -1.0;
    return h->tm.ms[i];
}

const char *tsg_last_error(tsg_handle_t h) {
    if (h) return h->err.c_str();
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(g_err_mu);
    copy = g_err;
    return copy.c_str();
}

int tsg_spgemm_symbolic(tsg_handle_t h, int m, int n, int k, const int *d_ia, const int *d_ja,
                        int64_t nnz_a, const int *d_ib, const int *d_jb, int64_t nnz_b,
                        int64_t *d_ic, int64_t *nnz_c) {
    if (!h) return TSG_ERR_INVALID;
    return guarded(h, [&] {
        check_csr_args(m, n, d_ia, d_ja, nnz_a, "A");
        check_csr_args(n, k, d_ib, d_jb, nnz_b, "B");
        TSG_CHECK(d_ic && nnz_c, TSG_ERR_INVALID, "null output");
        DeviceGuard dg(h->device);
        h->ws.reset();
        h->aux.reset();
        h->tm.clear();
        h->tm.stream = h->stream;
        h->row.valid = false;
        h->tile.valid = false;
        symbolic_into(h, h->row, h->tile, h->method_used, h->ws, h->aux, h->tm, h->row_tune, m,
                      n, k, d_ia, d_ja, nnz_a, d_ib, d_jb, nnz_b, d_ic, nnz_c);
    });
}

int tsg_spgemm_numeric(tsg_handle_t h, int m, int n, int k, const int *d_ia, const int *d_ja,
                       const double *d_a, int64_t nnz_a, const int *d_ib, const int *d_jb,
                       const double *d_b, int64_t nnz_b, const int64_t *d_ic, int *d_jc,
                       double *d_c) {
    if (!h) return TSG_ERR_INVALID;
    return guarded(h, [&] {
        check_csr_args(m, n, d_ia, d_ja, nnz_a, "A");
        check_csr_args(n, k, d_ib, d_jb, nnz_b, "B");
        TSG_CHECK(nnz_a == 0 || d_a, TSG_ERR_INVALID, "null A values");
        TSG_CHECK(nnz_b == 0 || d_b, TSG_ERR_INVALID, "null B values");
        DeviceGuard dg(h->device);
        numeric_from(h, h->method_used, h->row, h->tile, h->ws, h->tm, m, n, k, d_ia, d_ja, d_a,
                     nnz_a, d_ib, d_jb, d_b, nnz_b, d_ic, d_jc, d_c);
    });
}

int tsg_spgemm_noalloc(tsg_handle_t h, int m, int n, int k, const int *d_ia, const int *d_ja,
                       const double *d_a, int64_t nnz_a, const int *d_ib, const int *d_jb,
                       const double *d_b, int64_t nnz_b, int64_t *d_ic, int *d_jc, double *d_c,
                       int64_t nzmax, int64_t *nnz_c) {
    if (!h) return TSG_ERR_INVALID;
    return guarded(h, [&] {
        check_csr_args(m, n, d_ia, d_ja, nnz_a, "A");
        check_csr_args(n, k, d_ib, d_jb, nnz_b, "B");
        TSG_CHECK(nnz_a == 0 || d_a, TSG_ERR_INVALID, "null A values");
        TSG_CHECK(nnz_b == 0 || d_b, TSG_ERR_INVALID, "null B values");
        TSG_CHECK(d_ic && nnz_c && nzmax >= 0 && (nzmax == 0 || (d_jc && d_c)), TSG_ERR_INVALID,
                  "null output or negative nzmax");
        const size_t cap = TSG_NOALLOC_SCRATCH_BYTES;
        // the row arrays and per-entry product counts every product needs
        // (the rest is checked as it is carved out of the scratch)
        const double least = 28.0 * ((double)m + 1) + 8.0 * ((double)nnz_a + 1);
        TSG_CHECK(least <= (double)cap, TSG_ERR_INVALID,
                  "tsg_spgemm_noalloc: m = " + std::to_string(m) + " and nnz(A) = " +
                      std::to_string(nnz_a) + " need more than its " + std::to_string(cap) +
                      "-byte scratch");
        TSG_CHECK(h->device >= 0 && h->device < MAX_DEVICES && g_scratch_ptr[h->device],
                  TSG_ERR_STATE, "tsg_spgemm_noalloc: no scratch set up on this device");
        DeviceGuard dg(h->device);
        std::lock_guard<std::mutex> lk(g_scratch_mu[h->device]);
        // on every way out (errors included) the work queued on the stream
        // finishes before the lock is released, so the next call finds the
        // scratch idle
        struct Drain {
            hipStream_t s;
            ~Drain() { (void)hipStreamSynchronize(s); }
        } drain{h->stream};
        Arena ws;
        ws.attach(g_scratch_ptr[h->device], cap, "tsg_spgemm_noalloc scratch");
        RowPlan row;
        TilePlan tile;
        int method_used = TSG_METHOD_AUTO;
        StageTimer tm;   // disabled: records no events
        tm.stream = h->stream;
        RowTuning tune = h->row_tune;
        // the big rows' bitmaps and the block pointers must fit beside the rest
        if (tune.bitmap_budget < 0 || (size_t)tune.bitmap_budget > cap / 4)
            tune.bitmap_budget = (int64_t)(cap / 4);
        if (tune.bptr_budget < 0 || (size_t)tune.bptr_budget > cap / 8)
            tune.bptr_budget = (int64_t)(cap / 8);
        int64_t nz = 0;
        symbolic_into(h, row, tile, method_used, ws, ws, tm, tune, m, n, k, d_ia, d_ja, nnz_a,
                      d_ib, d_jb, nnz_b, d_ic, &nz);
        *nnz_c = nz;
        if (nz > nzmax) {
            throw TsgError{TSG_ERR_OVERFLOW, "tsg_spgemm_noalloc: nnz(C) = " +
                                                 std::to_string(nz) + " exceeds nzmax = " +
                                                 std::to_string(nzmax)};
        }
        numeric_from(h, method_used, row, tile, ws, tm, m, n, k, d_ia, d_ja, d_a, nnz_a, d_ib,
                     d_jb, d_b, nnz_b, d_ic, d_jc, d_c);
        TSG_HIP_TRY(hipStreamSynchronize(h->stream));
    });
}

int tsg_csr_transpose(tsg_handle_t h, int m, int n, const int *d_ia, const int *d_ja,
                      const double *d_a, int64_t nnz, int *d_tp, int *d_ti, double *d_tv) {
    if (!h) return TSG_ERR_INVALID;
    return guarded(h, [&] {
        check_csr_args(m, n, d_ia, d_ja, nnz, "A");
        TSG_CHECK(d_tp && (nnz == 0 || d_ti), TSG_ERR_INVALID, "null output");
        DeviceGuard dg(h->device);
        h->aux.reset();
        csr_transpose(m, n, d_ia, d_ja, d_a, nnz, d_tp, d_ti, d_a ? d_tv : nullptr, h->aux,
                      h->stream);
        TSG_HIP_TRY(hipStreamSynchronize(h->stream));
    });
}

int tsg_batch_symbolic(tsg_handle_t h, int count, const tsg_product_t *prods, int64_t *d_ic,
                       int64_t *nnz_c_total, int64_t *nnz_c) {
    if (!h) return TSG_ERR_INVALID;
    return guarded(h, [&] {
        DeviceGuard dg(h->device);
        batch_symbolic(h, count, prods, d_ic, nnz_c_total, nnz_c);
    });
}

int tsg_batch_numeric(tsg_handle_t h, int count, const tsg_product_t *prods, const int64_t *d_ic,
                      int *d_jc, double *d_c) {
    if (!h) return TSG_ERR_INVALID;
    return guarded(h, [&] {
        DeviceGuard dg(h->device);
        batch_numeric(h, count, prods, d_ic, d_jc, d_c);
    });
}

int tsg_intermediate_products(tsg_handle_t h, int m, int n, const int *d_ia, const int *d_ja,
                              int64_t nnz_a, const int *d_ib, uint64_t *products) {
    if (!h || !products) return TSG_ERR_INVALID;
    return guarded(h, [&] {
        check_csr_args(m, n, d_ia, d_ja, nnz_a, "A");
        TSG_CHECK(d_ib != nullptr, TSG_ERR_INVALID, "B: null row pointer");
        DeviceGuard dg(h->device);
        h->aux.reset();
        CsrIn A{m, n, d_ia, d_ja, nullptr, nnz_a};
        const int vf = validate_csr_pair(A, A, h->aux, h->stream);
        TSG_CHECK((vf & 0x3) == 0, TSG_ERR_INVALID,
                  std::string("invalid CSR input: A") +
                      ((vf & 0x1) ? " has bad row pointers" : " has a column out of range"));
        int64_t *rp = h->aux.alloc_n<int64_t>((size_t)m + 1);
        int64_t *sc = h->aux.alloc_n<int64_t>(scan_tmp_elems((size_t)m + 1));
        csr_row_products(m, d_ia, d_ja, nnz_a, d_ib, rp, h->aux, h->stream);
        exclusive_scan_i64(rp, rp, (size_t)m, sc, h->stream);
        int64_t tot = 0;
        TSG_HIP_TRY(hipMemcpyAsync(&tot, rp + m, sizeof(int64_t), hipMemcpyDeviceToHost,
                                   h->stream));
        TSG_HIP_TRY(hipStreamSynchronize(h->stream));
        *products = (uint64_t)tot;
    });
}

// ------------------------------------------------------- plan inspection
static bool plan_array(tsg_handle_t h, int which, const void **ptr, size_t *bytes) {
    const TilePlan &p = h->tile;
    if (!p.valid) return false;
    const TileMat &A = p.A, &B = p.B;
    const size_t nC = (size_t)p.numblkC;
    switch (which) {
    case TSG_A_TILE_PTR: *ptr = A.tile_ptr; *bytes = 4 * ((size_t)A.blkm + 1); return true;
    case TSG_A_TILE_COL: *ptr = A.tile_col; *bytes = 4 * (size_t)A.numblk; return true;
    case TSG_A_TILE_OFF: *ptr = A.tile_off; *bytes = 4 * ((size_t)A.numblk + 1); return true;
    case TSG_A_MASK: *ptr = A.mask; *bytes = 2 * 16 * (size_t)A.numblk; return true;
    case TSG_A_ROWPTR: *ptr = A.rowptr; *bytes = 16 * (size_t)A.numblk; return true;
    case TSG_A_IDX: *ptr = A.idx; *bytes = (size_t)A.nnz; return true;
    case TSG_A_VAL: *ptr = A.val; *bytes = 8 * (size_t)A.nnz; return true;
    case TSG_B_TILE_PTR: *ptr = B.tile_ptr; *bytes = 4 * ((size_t)B.blkm + 1); return true;
    case TSG_B_TILE_COL: *ptr = B.tile_col; *bytes = 4 * (size_t)B.numblk; return true;
    case TSG_B_TILE_OFF: *ptr = B.tile_off; *bytes = 4 * ((size_t)B.numblk + 1); return true;
    case TSG_B_MASK: *ptr = B.mask; *bytes = 2 * 16 * (size_t)B.numblk; return true;
    case TSG_B_ROWPTR: *ptr = B.rowptr; *bytes = 16 * (size_t)B.numblk; return true;
    case TSG_B_IDX: *ptr = B.idx; *bytes = (size_t)B.nnz; return true;
    case TSG_B_VAL: *ptr = B.valc; *bytes = 8 * (size_t)B.nnz; return true;
    case TSG_B_CSC_PTR: *ptr = B.csc_ptr; *bytes = 4 * ((size_t)B.blkn + 1); return true;
    case TSG_B_CSC_ROW: *ptr = B.csc_row; *bytes = 4 * (size_t)B.numblk; return true;
    case TSG_B_CSC_TILE: *ptr = B.csc_tile; *bytes = 4 * (size_t)B.numblk; return true;
    case TSG_C_TILE_PTR: *ptr = p.tptrC; *bytes = 4 * ((size_t)A.blkm + 1); return true;
    case TSG_C_TILE_COL: *ptr = p.tcolC; *bytes = 4 * nC; return true;
    case TSG_C_MASK: *ptr = p.maskC; *bytes = 2 * 16 * nC; return true;
    case TSG_C_ROWPTR: *ptr = p.rowptrC; *bytes = 16 * nC; return true;
    case TSG_C_TILE_OFF: *ptr = p.offC; *bytes = 8 * (nC + 1); return true;
    default: return false;
    }
}

size_t tsg_plan_array_bytes(tsg_handle_t h, int which) {
    if (!h) return 0;
    const void *p;
    size_t b;
    return plan_array(h, which, &p, &b) ? b : 0;
}

int tsg_plan_array_copy(tsg_handle_t h, int which, void *dst, size_t bytes) {
    if (!h || !dst) return TSG_ERR_INVALID;
    return guarded(h, [&] {
        const void *p;
        size_t b;
        TSG_CHECK(plan_array(h, which, &p, &b), TSG_ERR_STATE, "no such plan array");
        TSG_CHECK(bytes >= b, TSG_ERR_INVALID, "destination too small");
        DeviceGuard dg(h->device);
        TSG_HIP_TRY(hipStreamSynchronize(h->stream));
        if (b) TSG_HIP_TRY(hipMemcpy(dst, p, b, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"

// =================================================================
// Host-pointer entries.  One process-wide context per device holds the
// device copies and the plan between the request=1 and request=0 calls.
namespace {

// Host-side copy split over a few threads: a pageable destination fills at
// the speed of several cores (page first-touch included) instead of one.
class CopyPool {
public:
    explicit CopyPool(int nthreads) : n_(std::max(1, nthreads)) {
        for (int i = 1; i < n_; i++) th_.emplace_back([this, i] { work(i); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void copy(void *dst, const void *src, size_t bytes) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char *>(dst);
            src_ = static_cast<const char *>(src);
            bytes_ = bytes;
            pending_ = n_ - 1;
            gen_++;
        }
        cv_.notify_all();
        stripe(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
    }

private:
    void stripe(int i) {
        // 4 KiB-aligned stripes, one per thread
        const size_t per = ((bytes_ + n_ - 1) / n_ + 4095) & ~size_t(4095);
        const size_t b0 = std::min(bytes_, per * i), b1 = std::min(bytes_, b0 + per);
        if (b1 > b0) std::memcpy(dst_ + b0, src_ + b0, b1 - b0);
    }
    void work(int i) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            stripe(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    const int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0;
    int pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Device -> pageable host copy through two pinned staging buffers: chunk
// i+1 is DMA'd into one while the pool copies chunk i out of the other.
// Waits for work queued on s before the first chunk; returns when dst is
// complete.
constexpr size_t STAGE_BYTES = size_t(64) << 20;

struct Stager {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::unique_ptr<CopyPool> pool;

    void init() {
        if (pool) return;
        for (int i = 0; i < 2; i++) {
            TSG_HIP_TRY(hipHostMalloc(&buf[i], STAGE_BYTES, hipHostMallocDefault));
            TSG_HIP_TRY(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
        }
        const unsigned hc = std::thread::hardware_concurrency();
        pool.reset(new CopyPool((int)std::min(8u, std::max(1u, hc))));
    }
    void d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
        if (bytes == 0) return;
        init();
        const size_t nch = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
        auto issue = [&](size_t i) {
            const size_t off = i * STAGE_BYTES, len = std::min(STAGE_BYTES, bytes - off);
            TSG_HIP_TRY(hipMemcpyAsync(buf[i & 1], static_cast<const char *>(src) + off, len,
                                       hipMemcpyDeviceToHost, s));
            TSG_HIP_TRY(hipEventRecord(ev[i & 1], s));
        };
        issue(0);
        for (size_t i = 0; i < nch; i++) {
            // chunk i+1 goes to the buffer whose chunk (i-1) was copied out
            // in the previous iteration
            if (i + 1 < nch) issue(i + 1);
            TSG_HIP_TRY(hipEventSynchronize(ev[i & 1]));
            const size_t off = i * STAGE_BYTES, len = std::min(STAGE_BYTES, bytes - off);
            pool->copy(static_cast<char *>(dst) + off, buf[i & 1], len);
        }
    }
};

struct HostCtx {
    std::mutex mu;
    Stager stage;   // C's way back to the caller's arrays
    tsg_handle_t h = nullptr;
    Arena buf;     // device copies of the inputs, ic, and C when kept whole
    Arena sbuf;    // per-slab row pointers and C of the slabbed numeric phase
    // identity of the cached symbolic phase: sizes, pointers and a checksum
    // of the index arrays' contents (a caller may rewrite them in place)
    bool have = false;
    bool plan_whole = false;   // the handle holds the whole product's plan
    int m = -1, n = -1, k = -1;
    const void *ia = nullptr, *ja = nullptr, *ib = nullptr, *jb = nullptr;
    uint64_t sum = 0;
    int64_t nnz_a = -1, nnz_b = -1, nnz_c = -1;
    int *d_ia = nullptr, *d_ja = nullptr, *d_ib = nullptr, *d_jb = nullptr;
    bool shared_idx = false;   // B's row pointers and columns are A's (same host arrays)
    // B's values always have a buffer of their own: a caller may pass B with
    // A's index arrays and different values
    double *d_a = nullptr, *d_b = nullptr;
    int64_t *d_ic = nullptr;
    int *d_jc = nullptr;       // whole C (allocated by the first numeric call)
    double *d_c = nullptr;
    // C entries kept on the device at once (0: automatic) and the slabs the
    // last numeric call ran
    int64_t slab_cap = 0;
    int last_slabs = 0;
};

HostCtx &host_ctx() {
    static HostCtx ctx;
    return ctx;
}

int env_device() {
    const char *e = std::getenv("TSG_DEVICE");
    return e ? std::atoi(e) : 0;
}

// 64-bit checksum of an int32 array: four independent multiply-rotate lanes
// over 8-byte words (several GB/s on one core), then the tail.
uint64_t checksum_i32(const int *p, int64_t n, uint64_t seed) {
    const uint64_t K = 0x9E3779B97F4A7C15ull;
    auto mix = [K](uint64_t h, uint64_t v) {
        h ^= v * K;
        return ((h << 29) | (h >> 35)) * 0xBF58476D1CE4E5B9ull;
    };
    uint64_t h[4] = {seed, seed ^ 1, seed ^ 2, seed ^ 3};
    const int64_t words = n / 2;
    int64_t w = 0;
    for (; w + 4 <= words; w += 4) {
        uint64_t v[4];
        std::memcpy(v, p + 2 * w, sizeof(v));
        for (int j = 0; j < 4; j++) h[j] = mix(h[j], v[j]);
    }
    uint64_t r = mix(mix(h[0], h[1]), mix(h[2], h[3]));
    for (int64_t i = 2 * w; i < n; i++) r = mix(r, (uint32_t)p[i]);
    return mix(r, (uint64_t)n);
}

uint64_t index_checksum(int m, int n, const int *ia, const int *ja, const int *ib,
                        const int *jb) {
    const int64_t nnz_a = ia[m] - ia[0], nnz_b = ib[n] - ib[0];
    uint64_t s = checksum_i32(ia, (int64_t)m + 1, 1);
    if (nnz_a) s = checksum_i32(ja, nnz_a, s);
    if (ib != ia || jb != ja) {
        s = checksum_i32(ib, (int64_t)n + 1, s);
        if (nnz_b) s = checksum_i32(jb, nnz_b, s);
    }
    return s;
}

// Uploads the inputs, runs the symbolic phase and remembers its identity.
void host_symbolic(HostCtx &x, int m, int n, int k, const int *ia, const int *ja,
                   const int *ib, const int *jb) {
    if (!x.h) {
        int rc = tsg_create(&x.h, env_device());
        if (rc != TSG_OK) throw TsgError{rc, tsg_last_error(nullptr)};
    }
    TSG_CHECK(m >= 0 && n >= 0 && k >= 0 && ia && ib, TSG_ERR_INVALID, "bad arguments");
    DeviceGuard dg(x.h->device);
    const int64_t nnz_a = ia[m] - ia[0], nnz_b = ib[n] - ib[0];
    TSG_CHECK(ia[0] == 0 && ib[0] == 0, TSG_ERR_INVALID, "row pointers must start at 0");
    x.have = false;
    x.plan_whole = false;
    x.buf.reset();
    x.d_jc = nullptr;
    x.d_c = nullptr;
    x.d_ia = x.buf.alloc_n<int>((size_t)m + 1);
    x.d_ja = x.buf.alloc_n<int>((size_t)nnz_a + 1);
    x.d_a = x.buf.alloc_n<double>((size_t)nnz_a + 1);
    const bool same = ia == ib && ja == jb && m == n;
    x.shared_idx = same;
    if (same) {
        x.d_ib = x.d_ia;
        x.d_jb = x.d_ja;
    } else {
        x.d_ib = x.buf.alloc_n<int>((size_t)n + 1);
        x.d_jb = x.buf.alloc_n<int>((size_t)nnz_b + 1);
    }
    x.d_b = x.buf.alloc_n<double>((size_t)nnz_b + 1);
    x.d_ic = x.buf.alloc_n<int64_t>((size_t)m + 1);
    hipStream_t s = x.h->stream;
    TSG_HIP_TRY(hipMemcpyAsync(x.d_ia, ia, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice, s));
    if (nnz_a)
        TSG_HIP_TRY(hipMemcpyAsync(x.d_ja, ja, sizeof(int) * nnz_a, hipMemcpyHostToDevice, s));
    if (!same) {
        TSG_HIP_TRY(hipMemcpyAsync(x.d_ib, ib, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice, s));
        if (nnz_b)
            TSG_HIP_TRY(hipMemcpyAsync(x.d_jb, jb, sizeof(int) * nnz_b, hipMemcpyHostToDevice, s));
    }
    int64_t nnz_c = 0;
    int rc = tsg_spgemm_symbolic(x.h, m, n, k, x.d_ia, x.d_ja, nnz_a, x.d_ib, x.d_jb, nnz_b,
                                 x.d_ic, &nnz_c);
    if (rc != TSG_OK) throw TsgError{rc, tsg_last_error(x.h)};
    x.have = true;
    x.plan_whole = true;
    x.m = m; x.n = n; x.k = k;
    x.ia = ia; x.ja = ja; x.ib = ib; x.jb = jb;
    x.sum = index_checksum(m, n, ia, ja, ib, jb);
    x.nnz_a = nnz_a; x.nnz_b = nnz_b; x.nnz_c = nnz_c;
}

// The cached plan serves a request=0 call only for the same sizes, the same
// arrays and the same contents of those arrays: a caller that rewrote ia, ja,
// ib or jb in place between the two calls gets a fresh symbolic phase.
bool host_matches(const HostCtx &x, int m, int n, int k, const int *ia, const int *ja,
                  const int *ib, const int *jb) {
    return x.have && x.m == m && x.n == n && x.k == k && x.ia == ia && x.ja == ja &&
           x.ib == ib && x.jb == jb && x.nnz_a == (int64_t)(ia[m] - ia[0]) &&
           x.nnz_b == (int64_t)(ib[n] - ib[0]) &&
           x.sum == index_checksum(m, n, ia, ja, ib, jb);
}

// C entries that may stay on the device at once: the caller's cap, else all
// of C when an earlier call holds it whole already or its 12 bytes per entry
// fit in 90% of the free device memory, else 35% of the free memory plus
// what the whole product's plan holds (the slabs' plans reuse that memory).
int64_t host_slab_cap(const HostCtx &x) {
    if (x.slab_cap > 0) return x.slab_cap;
    if (x.d_jc) return x.nnz_c;   // C already held whole by an earlier call
    size_t fr = 0, tot = 0;
    TSG_HIP_TRY(hipMemGetInfo(&fr, &tot));
    if (12.0 * (double)x.nnz_c + double(1 << 20) <= 0.9 * (double)fr) return x.nnz_c;
    return std::max<int64_t>(1, (int64_t)(0.35 * (double)(fr + x.h->ws.held()) / 12.0));
}

// Numeric phase from host buffers into host buffers.  When C exceeds the
// slab cap, A's rows run in slabs: consecutive rows whose entries (from the
// caller's ic) fit the cap, a row larger than the cap alone; each slab is
// its own product A[r0:r1]·B (A's row pointers rebased, its column and
// value arrays offset), symbolic then numeric, and its columns and values
// are copied to jc and c at ic[r0].  Every row's entries and sums depend on
// that row alone, so the slabs give the whole product's C bit for bit.
template <class IC>
void host_numeric(HostCtx &x, const int *ia, const IC *ic, const double *a, const double *b,
                  int *jc, double *c) {
    DeviceGuard dg(x.h->device);
    hipStream_t s = x.h->stream;
    if (x.nnz_a)
        TSG_HIP_TRY(hipMemcpyAsync(x.d_a, a, sizeof(double) * x.nnz_a, hipMemcpyHostToDevice, s));
    // B's values are A's device copy only when B is A (same index arrays and
    // the same value array); otherwise they go to B's own buffer
    const double *d_b = x.d_a;
    if (!(x.shared_idx && b == a)) {
        if (x.nnz_b)
            TSG_HIP_TRY(hipMemcpyAsync(x.d_b, b, sizeof(double) * x.nnz_b, hipMemcpyHostToDevice,
                                       s));
        d_b = x.d_b;
    }
    const int64_t cap = host_slab_cap(x);
    if (x.nnz_c <= cap) {
        if (!x.plan_whole) {
            int64_t nz = 0;
            int rc = tsg_spgemm_symbolic(x.h, x.m, x.n, x.k, x.d_ia, x.d_ja, x.nnz_a, x.d_ib,
                                         x.d_jb, x.nnz_b, x.d_ic, &nz);
            if (rc != TSG_OK) throw TsgError{rc, tsg_last_error(x.h)};
            TSG_CHECK(nz == x.nnz_c, TSG_ERR_STATE, "symbolic phase disagrees with ic");
            x.plan_whole = true;
        }
        if (!x.d_jc) {
            x.d_jc = x.buf.alloc_n<int>((size_t)x.nnz_c + 1);
            x.d_c = x.buf.alloc_n<double>((size_t)x.nnz_c + 1);
        }
        int rc = tsg_spgemm_numeric(x.h, x.m, x.n, x.k, x.d_ia, x.d_ja, x.d_a, x.nnz_a, x.d_ib,
                                    x.d_jb, d_b, x.nnz_b, x.d_ic, x.d_jc, x.d_c);
        if (rc != TSG_OK) throw TsgError{rc, tsg_last_error(x.h)};
        x.stage.d2h(jc, x.d_jc, sizeof(int) * x.nnz_c, s);
        x.stage.d2h(c, x.d_c, sizeof(double) * x.nnz_c, s);
        TSG_HIP_TRY(hipStreamSynchronize(s));
        x.last_slabs = 1;
        return;
    }
    // the whole product's plan and C give way to the slabs'
    TSG_HIP_TRY(hipStreamSynchronize(s));
    x.h->row.valid = false;
    x.h->tile.valid = false;
    x.h->ws.release();
    x.plan_whole = false;
    // the copies land at the caller's ic: it must ascend to ic[m] = nnz(C)
    for (int r = 0; r < x.m; r++)
        TSG_CHECK(ic[r + 1] >= ic[r], TSG_ERR_STATE, "ic is not non-decreasing");
    std::vector<int> row0{0};
    int max_rows = 0;
    int64_t max_nnz = 0;
    for (int r = 0; r < x.m;) {
        int r1 = r + 1;
        while (r1 < x.m && (int64_t)ic[r1 + 1] - (int64_t)ic[r] <= cap) r1++;
        row0.push_back(r1);
        max_rows = std::max(max_rows, r1 - r);
        max_nnz = std::max<int64_t>(max_nnz, (int64_t)ic[r1] - (int64_t)ic[r]);
        r = r1;
    }
    x.sbuf.reset();
    int *d_ia_s = x.sbuf.alloc_n<int>((size_t)max_rows + 1);
    int64_t *d_ic_s = x.sbuf.alloc_n<int64_t>((size_t)max_rows + 1);
    int *d_jc_s = x.sbuf.alloc_n<int>((size_t)max_nnz + 1);
    double *d_c_s = x.sbuf.alloc_n<double>((size_t)max_nnz + 1);
    std::vector<int> ia_s((size_t)max_rows + 1);
    const int nslab = (int)row0.size() - 1;
    for (int i = 0; i < nslab; i++) {
        const int r0 = row0[i], r1 = row0[i + 1], ms = r1 - r0;
        const int64_t e0 = ia[r0], ea = (int64_t)ia[r1] - ia[r0];
        for (int j = 0; j <= ms; j++) ia_s[j] = ia[r0 + j] - ia[r0];
        TSG_HIP_TRY(hipMemcpyAsync(d_ia_s, ia_s.data(), sizeof(int) * ((size_t)ms + 1),
                                   hipMemcpyHostToDevice, s));
        int64_t nz = 0;
        int rc = tsg_spgemm_symbolic(x.h, ms, x.n, x.k, d_ia_s, x.d_ja + e0, ea, x.d_ib, x.d_jb,
                                     x.nnz_b, d_ic_s, &nz);
        if (rc != TSG_OK) throw TsgError{rc, tsg_last_error(x.h)};
        const int64_t o = ic[r0];
        TSG_CHECK(nz == (int64_t)ic[r1] - o, TSG_ERR_STATE,
                  "a slab's symbolic phase disagrees with ic");
        rc = tsg_spgemm_numeric(x.h, ms, x.n, x.k, d_ia_s, x.d_ja + e0, x.d_a + e0, ea, x.d_ib,
                                x.d_jb, d_b, x.nnz_b, d_ic_s, d_jc_s, d_c_s);
        if (rc != TSG_OK) throw TsgError{rc, tsg_last_error(x.h)};
        x.stage.d2h(jc + o, d_jc_s, sizeof(int) * nz, s);
        x.stage.d2h(c + o, d_c_s, sizeof(double) * nz, s);
        // ia_s is rewritten for the next slab
        TSG_HIP_TRY(hipStreamSynchronize(s));
    }
    x.last_slabs = nslab;
}

template <class IC>
int host_entry(int request, int m, int n, int k, const double *a, const int *ja, const int *ia,
               const double *b, const int *jb, const int *ib, double *c, int *jc, IC *ic) {
    HostCtx &x = host_ctx();
    std::lock_guard<std::mutex> lk(x.mu);
    return guarded(nullptr, [&] {
        TSG_CHECK(request == 0 || request == 1, TSG_ERR_INVALID, "request must be 0 or 1");
        TSG_CHECK(ic != nullptr, TSG_ERR_INVALID, "null ic");
        if (request == 1) {
            host_symbolic(x, m, n, k, ia, ja, ib, jb);
            if (sizeof(IC) == 4)
                TSG_CHECK(x.nnz_c <= 0x7fffffffLL, TSG_ERR_OVERFLOW,
                          "nnz(C) exceeds int32; use tsg_dcsrmultcsr_i64");
            std::vector<int64_t> tmp((size_t)m + 1);
            DeviceGuard dg(x.h->device);
            TSG_HIP_TRY(hipMemcpy(tmp.data(), x.d_ic, sizeof(int64_t) * ((size_t)m + 1),
                                  hipMemcpyDeviceToHost));
            for (int i = 0; i <= m; i++) ic[i] = (IC)tmp[i];
            return;
        }
        if (!host_matches(x, m, n, k, ia, ja, ib, jb)) host_symbolic(x, m, n, k, ia, ja, ib, jb);
        TSG_CHECK((int64_t)ic[m] == x.nnz_c && ic[0] == 0, TSG_ERR_STATE,
                  "ic does not match the symbolic phase of these inputs");
        TSG_CHECK(x.nnz_c == 0 || (jc && c), TSG_ERR_INVALID, "null jc or c");
        host_numeric(x, ia, ic, a, b, jc, c);
    });
}

}  // namespace

extern "C" {

void thmkl_dcsrmultcsr(const char *trans, const int *request, const int *sort, const int *m,
                       const int *n, const int *k, double *a, int *ja, int *ia, double *b,
                       int *jb, int *ib, double *c, int *jc, int *ic, const int *nzmax,
                       int *info) {
    (void)trans;
    (void)sort;
    (void)nzmax;
    if (!request || !m || !n || !k || !info) {
        if (info) *info = TSG_ERR_INVALID;
        return;
    }
    int rc = host_entry<int>(*request, *m, *n, *k, a, ja, ia, b, jb, ib, c, jc, ic);
    *info = rc != TSG_OK ? rc : (*request == 1 ? 1 : 0);
}

int tsg_dcsrmultcsr_i64(int request, int m, int n, int k, const double *a, const int *ja,
                        const int *ia, const double *b, const int *jb, const int *ib, double *c,
                        int *jc, int64_t *ic) {
    return host_entry<int64_t>(request, m, n, k, a, ja, ia, b, jb, ib, c, jc, ic);
}

int tsg_set_host_slab_entries(int64_t entries) {
    if (entries < 0) return TSG_ERR_INVALID;
    HostCtx &x = host_ctx();
    std::lock_guard<std::mutex> lk(x.mu);
    x.slab_cap = entries;
    return TSG_OK;
}

int64_t tsg_get_host_slab_entries(void) {
    HostCtx &x = host_ctx();
    std::lock_guard<std::mutex> lk(x.mu);
    return x.slab_cap;
}

int tsg_host_last_slabs(void) {
    HostCtx &x = host_ctx();
    std::lock_guard<std::mutex> lk(x.mu);
    return x.last_slabs;
}

}  // extern "C"

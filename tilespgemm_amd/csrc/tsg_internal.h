// Internal declarations shared by the tsg HIP sources (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <functional>
#include <string>
#include <vector>

#include "../../include/tsg.h"

#define TSG_WAVE 64
#define TSG_TILE 16

// ------------------------------------------------------------ error plumbing
struct TsgError {
    int code;
    std::string msg;
};

#define TSG_HIP_TRY(expr)                                                      \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess)                                                  \
            throw TsgError{TSG_ERR_HIP, std::string(#expr) + ": " +            \
                                            hipGetErrorString(_e)};            \
    } while (0)

#define TSG_CHECK(cond, code, text)                                            \
    do {                                                                       \
        if (!(cond)) throw TsgError{(code), (text)};                           \
    } while (0)

// Launch check: surfaces bad launch configurations immediately.
#define TSG_LAUNCH_CHECK() TSG_HIP_TRY(hipGetLastError())

// -------------------------------------------------------------- workspace
// Chunked bump arena over hipMalloc'd blocks.  reset() keeps the blocks, so
// repeated calls of the same size reach a steady state with no allocation.
// An arena attached to an external block (attach) bumps inside that block
// only: it never calls hipMalloc or hipFree, and a request past the block's
// end throws TSG_ERR_ALLOC instead of growing.
class Arena {
public:
    Arena() = default;
    Arena(const Arena &) = delete;
    Arena &operator=(const Arena &) = delete;
    ~Arena();
    void reset();
    void *alloc(size_t bytes);
    template <class T> T *alloc_n(size_t n) {
        return static_cast<T *>(alloc(n * sizeof(T) + 1));
    }
    size_t held() const { return held_; }
    void release();
    // Serve every allocation from [base, base + bytes) (256-byte aligned);
    // what is held is dropped first.
    void attach(void *base, size_t bytes, const char *what);
    size_t used() const;

private:
    struct Block { char *base; size_t size; size_t used; };
    std::vector<Block> blocks_;
    size_t cur_ = 0;
    size_t held_ = 0;
    bool external_ = false;
    const char *what_ = "";
};

// -------------------------------------------------------------- scans
// Exclusive scan of n int64 counts in place into out[0..n] (out[n] = total).
// in and out may alias.  tmp must hold scan_tmp_elems(n) int64.
size_t scan_tmp_elems(size_t n);
void exclusive_scan_i64(const int64_t *in, int64_t *out, size_t n, int64_t *tmp,
                        hipStream_t s);
// The int32 scan sums in int64; with total set, the exact int64 grand total
// is written there too (the caller checks it against INT32_MAX).
void exclusive_scan_i32(const int32_t *in, int32_t *out, size_t n, int64_t *tmp,
                        hipStream_t s, int64_t *total = nullptr);

// -------------------------------------------------------------- helpers
static inline unsigned grid_for(size_t work, unsigned per_block, unsigned cap = 1u << 30) {
    size_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

int device_cu_count(int device);

// Input description shared by both paths.
struct CsrIn {
    int rows, cols;
    const int *ptr;
    const int *idx;
    const double *val;   // may be null in the symbolic phase
    int64_t nnz;
};

// Stage timer: records (name, start, stop) HIP events on a stream.
struct StageTimer {
    bool enabled = false;
    hipStream_t stream = nullptr;
    std::vector<std::string> names;
    std::vector<hipEvent_t> ev;          // 2 per stage
    std::vector<double> ms;
    void clear();
    void begin(const char *name);
    void end();
    void finalize();                     // synchronizes, fills ms
    ~StageTimer();
};

// ------------------------------------------------------------ row path
// Tuning parameters of the row path (tsg_set_row_bins): rows with at most
// bin1 / bin2 / bin3 intermediate products go to the 16-, 32- and 64-lane
// hash kernels, the rest to the big-row path.
struct RowTuning {
    int bin1 = 64, bin2 = 256, bin3 = 1024;
    int64_t item_sort_min = int64_t(1) << 18;  // big-row items column-ordered from this many
    int64_t bitmap_budget = -1;  // bytes of big-row bitmaps (sparse form) kept; < 0: min(48 GiB, 40% of free)
    int64_t bptr_budget = -1;    // bytes of the B block-pointer table; < 0: 6 GiB
};

// One big-row item: d = {row, first column, end column, first sub-block in
// subocc (rank mode) or big-row index}, the position of its first output in
// C, and of its first word in the band word lists (rank mode).  One record,
// so that gathering the items into column order reads one line per item.
struct ItemRec {
    int4 d;
    int64_t off;
    int64_t wpos;
};
static_assert(sizeof(ItemRec) == 32, "item records are 32 bytes");

struct RowPlan {
    bool valid = false;
    bool b_strict = false;           // B rows strictly ascending (no duplicates)
    int m = 0, n = 0, k = 0;
    // per-row products and output counts
    int64_t *rowprod = nullptr;      // [m]
    int64_t *rownnz = nullptr;       // [m+1] (scratch for the scan)
    // bins: small rows (by products) and big rows
    int *bin_rows = nullptr;         // [m] rows grouped by bin
    int *bin_start = nullptr;        // [nbins+1] device
    int2 *span = nullptr;            // [m] column span (lo, hi) of each small row
    uint8_t *narrow3 = nullptr;      // [m] class of each row of the 1024-product bin:
                                     // 1 narrow, 2/3 span <= 2048 words with > / <= 512
                                     // outputs, 0 wider
    int *rows3 = nullptr;            // that bin's class-1 rows first, class-0 rows last
    int *rows3r = nullptr;           // its class-2 rows first, class-3 rows last
    int n3cls[4] = {0, 0, 0, 0};     // (host) rows per class
    int bin_start_h[16] = {0};       // host copy (after symbolic)
    int nbins = 0;
    // big-row structures
    int nbig = 0;
    int sub_shift = 0;               // log2 sub-block width (columns)
    int *span_sb0 = nullptr;         // [nbig] first sub-block of the row's span
    int64_t *sb_off = nullptr;       // [nbig+1] span offsets (sub-blocks)
    int64_t nsb_total = 0;
    uint16_t *subcnt = nullptr;      // [nsb_total] outputs per span sub-block
    // B block-pointer table for big-row items
    int blk_shift = 0;               // log2 block width (columns)
    int nblk = 0;                    // blocks per B row (+1 stored)
    int *bptr = nullptr;             // [(nblk+1) * n]: block-major with column_order
    bool bptr_rowptr = false;        // no table: B's row pointers serve as one block
    bool column_order = false;       // big-row bands and items run in column order
    // items
    int64_t nitems = 0;
    int64_t nitems_small = 0;        // items of <= 256 outputs, first in items (column order)
    int64_t nitems_mid = 0;          // end of the items of <= 512 outputs, next
    int64_t *item_cnt = nullptr;     // [nbig+1] items per big row -> offsets
    ItemRec *items = nullptr;        // by c0 (column order) or by row
    bool narrow_items = false;       // items of <= 8192 columns (else <= 16384)
    // big rows' column bitmaps kept from the symbolic phase (rank mode), in
    // sparse form: per sub-block a mask of its non-empty words, and per
    // 65536-column band those words in column order (a band word list),
    // bands placed in their row's region as they finished
    bool use_rank = false;
    int64_t words_cap = 0;           // words of bandwords (a bound)
    unsigned *bandwords = nullptr;   // [words_cap]
    unsigned *subocc = nullptr;      // [nsb_total] non-empty-word masks

    int4 *bpk = nullptr;             // [nnz(B)] {column, 0, value} records (rank mode)

    // B with rows sorted, when the input's were not (set after row_symbolic):
    // the plan ran on (B.ptr, b_idx); numeric phases gather values by b_perm
    bool b_canon = false;
    int *b_idx = nullptr;
    int *b_perm = nullptr;
    double *b_val = nullptr;
};

void row_symbolic(RowPlan &p, const CsrIn &A, const CsrIn &B, int64_t *d_ic,
                  int64_t *nnz_c, Arena &ws, hipStream_t s, StageTimer &tm, int cus,
                  const RowTuning &tune = RowTuning());
void row_numeric(RowPlan &p, const CsrIn &A, const CsrIn &B, const int64_t *d_ic,
                 int *d_jc, double *d_c, int fused, Arena &ws, hipStream_t s,
                 StageTimer &tm, int cus);

// ------------------------------------------------------------ tile path
// One tile's numeric-phase metadata in 64 bytes (one aligned half line), so
// a matched pair's two tiles cost two reads.  A tiles: row masks (bit 15-c)
// and row starts of the row-major payload; B tiles: column masks (bit 15-k)
// and column starts of the column-major payload.
struct TileRec {
    uint16_t mask[16];
    uint8_t start[16];
    int off;                     // payload offset of the tile
    int pad[3];
};
static_assert(sizeof(TileRec) == 64, "tile records are 64 bytes");

struct TileMat {
    int rows = 0, cols = 0, blkm = 0, blkn = 0, numblk = 0;
    int *tile_ptr = nullptr;     // [blkm+1]
    int *tile_col = nullptr;     // [numblk]
    int *tile_off = nullptr;     // [numblk+1]
    uint16_t *mask = nullptr;    // [16*numblk]
    uint8_t *rowptr = nullptr;   // [16*numblk]
    uint8_t *idx = nullptr;      // [nnz]
    double *val = nullptr;       // [nnz]
    int64_t nnz = 0;
    // tile-CSC index (B only)
    int *csc_ptr = nullptr;      // [blkn+1]
    int *csc_row = nullptr;      // [numblk]
    int *csc_tile = nullptr;     // [numblk]
    TileRec *rec = nullptr;      // [numblk] A: row layout; B: column layout
    // entry -> payload position permutation (values refreshed per numeric)
    int *perm = nullptr;         // [nnz] CSR position -> tile payload position
    int *permc = nullptr;        // [nnz] CSR position -> column-major payload position (B only)
    double *valc = nullptr;      // [nnz] values column-major inside each tile (B only)
};

struct TilePlan {
    bool valid = false;
    bool b_is_a = false;         // B has A's structure (same index arrays)
    TileMat A, B;
    int numblkC = 0;
    int *tptrC = nullptr;        // [blkm+1]
    int *tcolC = nullptr;        // [numblkC]
    int *trowC = nullptr;        // [numblkC]
    uint16_t *maskC = nullptr;   // [16*numblkC]
    uint8_t *rowptrC = nullptr;  // [16*numblkC]
    int64_t *offC = nullptr;     // [numblkC+1] payload offsets (tile order)
    int *rowoffC = nullptr;      // [numblkC*16] start of (tile,row) inside CSR row
    int64_t *rownnz = nullptr;   // [m+1] CSR row counts
    int64_t nnzC = 0;
    // matched (A tile, B tile) pairs of every C tile in ascending K
    int64_t *tpair = nullptr;    // [numblkC+1] offsets into pairs, or counts with tslot
    int64_t *tslot = nullptr;    // [numblkC] first pair of each C tile (pairs in slots)
    int2 *pairs = nullptr;       // [npairs] {A tile position, B tile position}
    int64_t npairs = 0;
    // numeric work units (tile * 4 + part) by C tile output count: bins 0..3
    // (<= 16, <= 32, <= 64, 64-output parts of larger tiles) back to back,
    // tiles without outputs left out
    int4 *tbin_units = nullptr;      // [units] {tile, part | pairs << 2, first pair (lo, hi)}
    int tbin_start[5] = {0, 0, 0, 0, 0};
};

// Throws TSG_ERR_OVERFLOW unless 0 <= n_c_tiles <= INT32_MAX.
void check_tile_count(int64_t n_c_tiles);
// A matrix's tile count per tile row, scanned, enqueued ahead of the tile
// path (beside the input's validation, whose host round trip then brings
// the total back into nt): the count clamps every index, so it is safe on
// unchecked input.  tile_ptr and nt are used only once the input is valid.
struct EarlyTileCount {
    int *tile_ptr = nullptr;     // [blkm+1] device
    int nt = -1;                 // host, after the round trip
};
void tile_count_early(const CsrIn &X, EarlyTileCount &e, Arena &ws, hipStream_t s);
void tile_symbolic(TilePlan &p, const CsrIn &A, const CsrIn &B, bool b_is_a,
                   int64_t *d_ic, int64_t *nnz_c, Arena &ws, hipStream_t s,
                   StageTimer &tm, int cus, const EarlyTileCount *earlyA = nullptr,
                   const EarlyTileCount *earlyB = nullptr);
void tile_numeric(TilePlan &p, const CsrIn &A, const CsrIn &B, const int64_t *d_ic,
                  int *d_jc, double *d_c, int fused, Arena &ws, hipStream_t s,
                  StageTimer &tm, int cus);

// ------------------------------------------------------------ batches
// A batch laid out as one block-diagonal product (tsg_batch.hip).
struct BatchState {
    bool valid = false;
    int count = 0;
    bool b_is_a = false;
    bool local_cols = false;                   // B (and so C) keeps each product's own columns
    int rows = 0, cols_a = 0, cols_b = 0;     // sum m, sum n; columns of C (see local_cols)
    int64_t nnz_a = 0, nnz_b = 0;
    void *desc = nullptr;                      // device descriptors
    int64_t *row_off = nullptr;                // [count+1] C/A row offsets
    int *ia = nullptr, *ja = nullptr, *ib = nullptr, *jb = nullptr;
    double *a = nullptr, *b = nullptr;
    std::vector<tsg_product_t> host;           // structure identity check
};

// ------------------------------------------------------------ handle
struct tsg_handle_s {
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int method = TSG_METHOD_AUTO;
    int fp_mode = TSG_FP_FUSED;
    int method_used = TSG_METHOD_AUTO;
    double tile_density = 32.0;  // AUTO: tile path from this many entries per tile
    RowTuning row_tune;
    Arena ws;       // plan of the last symbolic phase
    Arena bws;      // batch layout (outlives the plan's reset)
    Arena aux;      // scratch of utility calls (transpose, product counts)
    RowPlan row;
    TilePlan tile;
    BatchState batch;
    StageTimer tm;
    std::string err;
};

void batch_symbolic(tsg_handle_t h, int count, const tsg_product_t *prods, int64_t *d_ic,
                    int64_t *nnz_c_total, int64_t *nnz_c);
void batch_numeric(tsg_handle_t h, int count, const tsg_product_t *prods, const int64_t *d_ic,
                   int *d_jc, double *d_c);
// Row pointers of every product of the last batch rebased to start at 0,
// m_i + 1 entries each, back to back (rows + count entries).
void batch_rebase(tsg_handle_t h, const int64_t *d_ic, int64_t *d_out);

// CSR utilities
// Validation of A and B on the device, one host round trip (flags: A in
// bits 0-7, B in bits 8-15; 1 bad row pointers, 2 column out of range, 4 not
// strictly ascending).  With density_sample, A's tile-density sample
// (tile_sample_launch) runs beside the checks and its {tiles, entries} come
// back in the same round trip; before_sync enqueues more work (and read-
// backs) ahead of that round trip.
int validate_csr_pair(const CsrIn &A, const CsrIn &B, Arena &ws, hipStream_t s,
                      unsigned long long *density_sample = nullptr,
                      const std::function<void()> &before_sync = {});
// entries per non-empty 16x16 tile of X (tsg_tile.hip; a sample decides when
// clear of the threshold)
void tile_sample_launch(const CsrIn &X, unsigned long long *sums, hipStream_t s);
double tile_density(const CsrIn &X, double threshold, Arena &ws, hipStream_t s,
                    const unsigned long long *sample = nullptr);
void csr_transpose(int m, int n, const int *ia, const int *ja, const double *a,
                   int64_t nnz, int *tp, int *ti, double *tv, Arena &ws,
                   hipStream_t s, int *tpos = nullptr);
// Row-sorted copy of X's columns (stable: equal columns keep their order) and
// the original position of every sorted entry; temporaries from tmp.
void csr_sort_rows(int m, int n, const int *ia, const int *ja, int64_t nnz, int *sorted_ja,
                   int *perm, Arena &keep, Arena &tmp, hipStream_t s);
void gather_values(int64_t n, const int *perm, const double *src, double *dst, hipStream_t s);
void sort_order_i32(const int *keys, int64_t n, int end_bit, int *order, Arena &ws,
                    hipStream_t s);
void csr_row_products(int m, const int *ia, const int *ja, int64_t nnz, const int *ib,
                      int64_t *rowprod, Arena &ws, hipStream_t s);

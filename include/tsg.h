/*
 * tsg -- MI355X (gfx950) tiled FP64 SpGEMM, C = A * B with A, B, C in CSR.
 *
 * C ABI of libtsg.so.  No HIP or torch types appear in any signature: device
 * buffers are plain pointers, streams are passed as void *.
 *
 * Arithmetic: every output entry c_ij is accumulated from +0.0 over its
 * products a_ik * b_kj in ascending position of k within row i of A (A-row
 * order, as the reference's thmkl_dcsrmultcsr.c:61-72), one fused multiply-add
 * per product by default (TSG_FP_FUSED) or a rounded product followed by a
 * rounded add (TSG_FP_SEPARATE).  Given the same mode, results are bitwise
 * identical to the CPU reference on column-sorted, duplicate-free inputs.
 * Structural zeros (products that cancel to 0.0) are kept.
 *
 * Index types: inputs use int32 row pointers and columns (the reference's
 * MAT_PTR_TYPE, common.h:15-17); the device API returns int64 row pointers
 * for C so outputs past 2^31 entries are expressible.  The reference-
 * compatible host entry thmkl_dcsrmultcsr keeps int32 and reports overflow
 * through *info.
 */
#ifndef TSG_H
#define TSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    TSG_OK = 0,
    TSG_ERR_INVALID = -1,   /* bad argument (negative size, null pointer) */
    TSG_ERR_HIP = -2,       /* a HIP runtime call failed                 */
    TSG_ERR_ALLOC = -3,     /* device or host allocation failed          */
    TSG_ERR_OVERFLOW = -4,  /* nnz(C) does not fit the requested index   */
    TSG_ERR_STATE = -5,     /* numeric phase without a matching symbolic */
    TSG_ERR_NODEVICE = -6   /* no usable HIP device                      */
};

/* ------------------------------------------------------------- methods */
enum {
    TSG_METHOD_AUTO = 0,    /* tile path for tile-dense A, row path else */
    TSG_METHOD_TILE = 1,    /* 16x16 bitmask tiles (TileSpGEMM)          */
    TSG_METHOD_ROW = 2      /* row-binned hash / bitmap accumulation     */
};

enum { TSG_FP_FUSED = 0, TSG_FP_SEPARATE = 1 };

typedef struct tsg_handle_s *tsg_handle_t;

/* ====================================================================
 * Reference-compatible host entry.
 *
 * Replaces thmkl_dcsrmultcsr (reference thmkl_dcsrmultcsr.c:7-10), same
 * argument list and meaning:
 *   *request == 1  fill ic[0..m] (0-based, ic[m] = nnz(C)); *info = 1
 *   *request == 0  with ic from the previous call, fill jc[ic[i]..ic[i+1])
 *                  ascending and c; *info = 0
 * m = rows of A and C, n = columns of A = rows of B, k = columns of B and C.
 * trans, sort and nzmax are accepted and ignored, as in the reference.
 * Host pointers in, host pointers out; the device work runs on device 0 (or
 * $TSG_DEVICE).  The symbolic plan of the request=1 call is cached and reused
 * by the request=0 call when (m, n, k, ia, ja, ib, jb) match and a checksum
 * of the index arrays' contents (row pointers and columns of A and B) is
 * unchanged, so arrays rewritten in place between the two calls are seen;
 * otherwise the symbolic phase is recomputed.
 * The request=0 call keeps C whole on the device when it fits, else it
 * streams C in row slabs (tsg_set_host_slab_entries): each slab of
 * consecutive rows of A is a product of its own (symbolic, numeric, copy of
 * its jc and c to the caller's arrays at ic[first row]); the result is the
 * same bit for bit.
 * Errors (the reference has none): *info < 0 with a TSG_ERR_* code;
 * TSG_ERR_OVERFLOW when nnz(C) > INT32_MAX (use tsg_dcsrmultcsr_i64).
 * ==================================================================== */
void thmkl_dcsrmultcsr(const char *trans, const int *request, const int *sort,
                       const int *m, const int *n, const int *k,
                       double *a, int *ja, int *ia,
                       double *b, int *jb, int *ib,
                       double *c, int *jc, int *ic,
                       const int *nzmax, int *info);

/* Same two-phase contract with an int64 ic (no overflow at nnz(C) >= 2^31).
 * Returns TSG_OK or a TSG_ERR_* code. */
int tsg_dcsrmultcsr_i64(int request, int m, int n, int k,
                        const double *a, const int *ja, const int *ia,
                        const double *b, const int *jb, const int *ib,
                        double *c, int *jc, int64_t *ic);

/* Row-slab streaming of the two host entries above (no counterpart in the
 * reference, whose C lives in host memory).  entries caps the C entries the
 * request=0 call holds on the device at once (12 bytes each); a product with
 * more runs in slabs of consecutive rows whose entries fit, a row with more
 * entries than the cap in a slab of its own.  0 (the default) chooses: C
 * whole when it fits in 90% of the free device memory, else slabs of 35% of
 * it.  Returns TSG_ERR_INVALID for entries < 0. */
int tsg_set_host_slab_entries(int64_t entries);
int64_t tsg_get_host_slab_entries(void);
/* Slabs run by the last request=0 call of the host entries (1: C whole). */
int tsg_host_last_slabs(void);

/* ====================================================================
 * Device API.  All matrix pointers are device pointers on the handle's
 * device; calls are asynchronous on the handle's stream except where noted.
 * ==================================================================== */
int tsg_create(tsg_handle_t *handle, int device);
int tsg_destroy(tsg_handle_t handle);
/* stream: a hipStream_t, or NULL for the handle's own stream */
int tsg_set_stream(tsg_handle_t handle, void *stream);
int tsg_set_method(tsg_handle_t handle, int method);
int tsg_set_fp_mode(tsg_handle_t handle, int fp_mode);
/* Method the last symbolic phase used (TSG_METHOD_TILE or TSG_METHOD_ROW). */
int tsg_get_method_used(tsg_handle_t handle);

/* Tuning parameters (no counterpart in the reference, where the matching
 * thresholds are compile-time: main_spgemm-col.cu:26-31,
 * spgemm_nsparse_kernel.h:4-14).  They change speed, never results.
 * tile_density: TSG_METHOD_AUTO takes the tile path when A holds at least
 * this many entries per non-empty 16x16 tile (default 32: measured on
 * MI355X, 5-point and 27-point stencils, at 16-19 entries per tile, run
 * faster on the row path and banded or block-dense matrices, at 90 and
 * more, 1.5-4x faster on the tile path).
 * Row bins: rows with at most bin1 (<= 64), bin2 (<= 256) and bin3 (<= 1024)
 * intermediate products go to the 16-, 32- and 64-lane LDS-hash kernels,
 * longer rows to the big-row path (defaults 64, 256, 1024; bin1 <= bin2 <=
 * bin3).  tilespgemm_amd.autotune searches them for a workload. */
int tsg_set_tile_density(tsg_handle_t handle, double entries_per_tile);
int tsg_set_row_bins(tsg_handle_t handle, int bin1, int bin2, int bin3);
int tsg_get_tuning(tsg_handle_t handle, double *tile_density, int *bin1, int *bin2,
                   int *bin3);
/* Big-row work order (no counterpart in the reference): when the column
 * spans of the row path's big rows add up to at least min_items windows of
 * 16384 columns, the symbolic band units and the numeric items are run in
 * order of their first column (a radix sort in the symbolic phase), so the
 * units in flight together read B entries of one narrow column band and
 * find them in L2; the B block pointers are then laid out block-major and
 * the items cut to 8192 columns.  Below that the sort costs more than it
 * saves.  Default 2^18 (R-MAT n=2^20 A*A^T: 24M windows; n=2^18 with 4
 * draws per row: 523K, 6.48 -> 6.10 ms in column order; n=2^16: 108K,
 * even; n=2^14: 1.68 -> 1.89 ms).  0 always sorts.  Changes speed only. */
int tsg_set_item_order(tsg_handle_t handle, int64_t min_items);
/* Memory for the big rows' column bitmaps (no counterpart in the
 * reference): the symbolic phase keeps every big row's bitmap in sparse form
 * -- per 1024-column sub-block a mask of its non-empty 32-bit words, and
 * those words -- for the rank-indexed numeric kernel when the space reserved
 * for it (4 bytes per sub-block of the rows' spans plus, per row, 4 bytes
 * times min(products, span words)) is at most this many bytes; beyond that
 * the numeric phase builds each item's columns in an LDS hash table instead
 * (same results, slower).  R-MAT n=2^20 A*A^T: a fifth of its spans' words
 * are non-empty, so the sparse form takes about a quarter of the 47 GB the
 * dense bitmaps would.  Negative (the default): min(48 GiB, 40% of free
 * device memory). */
int tsg_set_bitmap_budget(tsg_handle_t handle, int64_t bytes);
/* Memory for the B block-pointer table of the big-row path (no counterpart
 * in the reference): one position per B row and 1024-column block while the
 * table takes at most this many bytes, else blocks of 2048, 4096, ...
 * columns (up to 2^30) with a binary search inside a block.  0: no table,
 * every range is searched inside its B row.  Negative (the default): 6 GiB,
 * and no table when building it would cost more than 8 entries per lookup
 * it serves (a few big rows against many B rows, as in batches of small
 * products).  Changes speed only. */
int tsg_set_block_pointer_budget(tsg_handle_t handle, int64_t bytes);
int tsg_get_item_order(tsg_handle_t handle, int64_t *min_items);
/* Device workspace currently held by the handle, bytes. */
size_t tsg_workspace_bytes(tsg_handle_t handle);

/* Symbolic phase (thmkl request=1, device side).  Builds the plan (tile
 * structures or row bins) kept in the handle, writes C's row pointers d_ic
 * (m+1 int64) and returns nnz(C) in *nnz_c.  Synchronizes the stream once
 * (nnz(C) is needed on the host to size jc and c). */
int tsg_spgemm_symbolic(tsg_handle_t handle, int m, int n, int k,
                        const int *d_ia, const int *d_ja, int64_t nnz_a,
                        const int *d_ib, const int *d_jb, int64_t nnz_b,
                        int64_t *d_ic, int64_t *nnz_c);

/* Numeric phase (thmkl request=0, device side).  Must follow a symbolic
 * phase on the same handle with the same structure (values may differ, so
 * one symbolic plan serves any number of numeric calls).  Fills d_jc/d_c
 * (nnz(C) entries, columns ascending per row).  Asynchronous. */
int tsg_spgemm_numeric(tsg_handle_t handle, int m, int n, int k,
                       const int *d_ia, const int *d_ja, const double *d_a,
                       int64_t nnz_a,
                       const int *d_ib, const int *d_jb, const double *d_b,
                       int64_t nnz_b,
                       const int64_t *d_ic, int *d_jc, double *d_c);

/* Symbolic and numeric phase in one call that allocates nothing (no
 * counterpart in the reference; for callers that cannot call hipMalloc:
 * serving loops and harnesses that forbid allocation).  Every workspace
 * array comes from a fixed device scratch of TSG_NOALLOC_SCRATCH_BYTES per
 * device, set up by tsg_create; inputs and outputs are the caller's device
 * buffers: d_ic (m+1 int64) always, d_jc and d_c with room for nzmax
 * entries.  Writes d_ic and *nnz_c; when nnz(C) <= nzmax also d_jc and d_c,
 * the same bits as tsg_spgemm_symbolic + tsg_spgemm_numeric on the handle's
 * settings (method, FP mode, tuning), else returns TSG_ERR_OVERFLOW.  A
 * product whose workspace exceeds the scratch returns TSG_ERR_INVALID
 * (m and nnz(A) too large at a glance) or TSG_ERR_ALLOC (the scratch ran
 * out), with the sizes in tsg_last_error.  The handle's own plan is left as
 * it was.  Calls on one device share the scratch: they are serialised and
 * each returns after its work on the handle's stream has finished (the
 * call synchronises that stream), so calls never overlap on the device. */
#define TSG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tsg_spgemm_noalloc(tsg_handle_t handle, int m, int n, int k,
                       const int *d_ia, const int *d_ja, const double *d_a, int64_t nnz_a,
                       const int *d_ib, const int *d_jb, const double *d_b, int64_t nnz_b,
                       int64_t *d_ic, int *d_jc, double *d_c, int64_t nzmax, int64_t *nnz_c);

/* ====================================================================
 * Batched products (no counterpart in the reference, whose driver runs one
 * product per process).  The batch is laid out as one block-diagonal
 * product diag(A_i) * diag(B_i) = diag(C_i), so every stage is one launch
 * over all products.  Device pointers; int32 inputs.
 * ==================================================================== */
typedef struct {
    int m, n, k;                 /* A_i is m x n, B_i is n x k            */
    const int *ia, *ja;          /* A_i CSR (row pointers start at 0)     */
    const double *a;
    int64_t nnz_a;
    const int *ib, *jb;          /* B_i CSR                               */
    const double *b;
    int64_t nnz_b;
} tsg_product_t;

/* Symbolic phase of a batch.  d_ic (sum of m_i, plus 1, int64) receives the
 * row pointers of all C_i back to back: product i owns rows
 * [sum_{j<i} m_j, sum_{j<=i} m_j) and entries [d_ic[first row], d_ic[last
 * row + 1]).  *nnz_c_total = nnz of all C_i; nnz_c (host, count entries, may
 * be NULL) receives each nnz(C_i).  Synchronizes the stream. */
int tsg_batch_symbolic(tsg_handle_t handle, int count, const tsg_product_t *prods,
                       int64_t *d_ic, int64_t *nnz_c_total, int64_t *nnz_c);

/* Numeric phase of a batch (same products, same structure, values may
 * differ).  Column indices in d_jc are local to each C_i.  Issues kernels
 * only (no allocation, no synchronization), so it can be captured in a HIP
 * graph and replayed. */
int tsg_batch_numeric(tsg_handle_t handle, int count, const tsg_product_t *prods,
                      const int64_t *d_ic, int *d_jc, double *d_c);

/* ====================================================================
 * Serving (no counterpart in the reference).  A request loop with
 * device-resident operands and continuous batching: `workers` threads, each
 * with its own handle and stream, take in FIFO order every waiting request
 * that fits the memory budget into one batched product, so requests that
 * arrive while others run are admitted as memory frees up and their work
 * overlaps on the device.  Admission charges a request
 * 16*(m+2) + 12*min(products, m*k) bytes (products = intermediate products,
 * computed on the device by an estimator thread after submission) from
 * admission until release; when no batch is running the head request is
 * admitted even if it exceeds the budget, so progress never depends on the
 * caller releasing results first.  max_batch = 1 and workers = 1 give plain
 * per-call processing in arrival order.
 * Outputs are device buffers owned by the server until released; calls on
 * one server are thread-safe.
 * The estimator and the workers use workers + 1 streams; give the process at
 * least that many HIP hardware queues beyond the caller's own busy streams
 * (GPU_MAX_HW_QUEUES, read when the runtime starts; HIP's default is 4).
 * Streams that share a hardware queue run in turn, so a worker's host round
 * trips would wait behind another worker's kernels.
 * ==================================================================== */
typedef struct tsg_server_s *tsg_server_t;

typedef struct {
    int workers;                 /* worker threads (handles, streams)       */
    int max_batch;               /* requests per batch                      */
    size_t mem_budget;           /* bytes charged to admitted, unreleased
                                    requests; 0 = 64 GiB                    */
    int method;                  /* TSG_METHOD_*                            */
    double tile_density;         /* tsg_set_tile_density, 0 = default       */
    int bin1, bin2, bin3;        /* tsg_set_row_bins; bin3 = 0: 64/256/512,
                                    not the library's 64/256/1024 (mid-size
                                    requests: rows of 513-1024 products run
                                    faster on the big-row path)            */
} tsg_server_config_t;

typedef struct {
    int status;                  /* TSG_OK or TSG_ERR_*                     */
    int m;                       /* rows of C                               */
    int64_t nnz;                 /* nnz(C)                                  */
    const int64_t *ic;           /* device, m+1 entries, ic[0] = 0          */
    const int *jc;               /* device, nnz entries                     */
    const double *c;             /* device, nnz entries                     */
    double queue_ms;             /* submission -> batch start               */
    double total_ms;             /* submission -> result ready              */
    int batch_size;              /* requests in the batch that served it    */
    int64_t products;            /* intermediate products of A * B (as
                                    tsg_intermediate_products), the basis
                                    of the admission estimate              */
} tsg_result_t;

int tsg_server_create(tsg_server_t *server, int device, const tsg_server_config_t *config);
/* Enqueues C = A * B for the device operands of *req (kept alive by the
 * caller until the result is released).  Returns a ticket. */
int tsg_server_submit(tsg_server_t server, const tsg_product_t *req, uint64_t *ticket);
/* Blocks until the ticket's result is ready. */
int tsg_server_wait(tsg_server_t server, uint64_t ticket, tsg_result_t *result);
/* Returns the result's device buffers to the server. */
int tsg_server_release(tsg_server_t server, uint64_t ticket);
/* Drains the queue, stops the workers, frees everything. */
int tsg_server_destroy(tsg_server_t server);

/* Device CSR transpose (reference utils.h:158-195): row j of the output
 * holds column j of the input, entries by ascending source row.  d_tp has
 * n+1 entries.  Used for C = A * A^T, whose B the reference builds on the
 * host (main_spgemm-col.cu:4914-4937). */
int tsg_csr_transpose(tsg_handle_t handle, int m, int n,
                      const int *d_ia, const int *d_ja, const double *d_a,
                      int64_t nnz, int *d_tp, int *d_ti, double *d_tv);

/* Sum over A's entries a_ik of |row k of B| -- the intermediate-product
 * count (reference main_spgemm-col.cu:4950-4955); flops = 2 * this.  A is
 * m x n (n = rows of B, d_ib has n+1 entries) and is validated first
 * (TSG_ERR_INVALID for bad row pointers or columns outside [0, n)).
 * Synchronizes the stream. */
int tsg_intermediate_products(tsg_handle_t handle, int m, int n, const int *d_ia,
                              const int *d_ja, int64_t nnz_a, const int *d_ib,
                              uint64_t *products);

/* ====================================================================
 * Plan inspection (tile path).  After a symbolic phase with
 * TSG_METHOD_TILE, copies an internal tile array to host memory so the tile
 * format and the tile-level steps can be checked against the reference's
 * CPU restatement.  tsg_plan_array_bytes returns the size (0 if absent).
 * Synchronizes the stream.
 * ==================================================================== */
enum {
    /* A in tile-CSR (main_spgemm-col.cu:4976-5240) */
    TSG_A_TILE_PTR = 0,   /* int32[blkm+1]                                 */
    TSG_A_TILE_COL,       /* int32[numblkA]                                */
    TSG_A_TILE_OFF,       /* int32[numblkA+1]  payload offsets (nnzb_A)     */
    TSG_A_MASK,           /* uint16[16*numblkA] row masks, bit 15-c        */
    TSG_A_ROWPTR,         /* uint8[16*numblkA]  in-tile row starts          */
    TSG_A_IDX,            /* uint8[nnzA]        (r<<4)|c                    */
    TSG_A_VAL,            /* double[nnzA]       (after a numeric phase)     */
    /* B: tile-CSR payload + tile-CSC index (main_spgemm-col.cu:5003-5360) */
    TSG_B_TILE_PTR,
    TSG_B_TILE_COL,
    TSG_B_TILE_OFF,
    TSG_B_MASK,
    TSG_B_ROWPTR,
    TSG_B_IDX,
    TSG_B_VAL,            /* double[nnzB] column-major inside each tile:
                             columns ascending, rows ascending in a column
                             (the reference's transposed B sub-tiles,
                             main_spgemm-col.cu:5255-5290); after a numeric
                             phase                                        */
    TSG_B_CSC_PTR,        /* int32[blknB+1]                                */
    TSG_B_CSC_ROW,        /* int32[numblkB]  tile row K per tile column     */
    TSG_B_CSC_TILE,       /* int32[numblkB]  tile-CSR position of (K, J)   */
    /* C tile structure (steps 1-2, main_spgemm-col.cu:997-1325) */
    TSG_C_TILE_PTR,       /* int32[blkm+1]                                 */
    TSG_C_TILE_COL,       /* int32[numblkC]                                */
    TSG_C_MASK,           /* uint16[16*numblkC]                            */
    TSG_C_ROWPTR,         /* uint8[16*numblkC]                             */
    TSG_C_TILE_OFF,       /* int64[numblkC+1]  payload offsets (nnzb_C)    */
    TSG_PLAN_ARRAY_COUNT
};
size_t tsg_plan_array_bytes(tsg_handle_t handle, int which);
int tsg_plan_array_copy(tsg_handle_t handle, int which, void *host_dst,
                        size_t bytes);

/* Per-stage device time of the last symbolic + numeric phases, measured with
 * HIP events on the handle's stream when timing is enabled (off by default:
 * events add a few microseconds).  Stage names are returned through
 * tsg_stage_name; times in milliseconds. */
int tsg_enable_timing(tsg_handle_t handle, int enable);
int tsg_stage_count(tsg_handle_t handle);
const char *tsg_stage_name(tsg_handle_t handle, int stage);
double tsg_stage_ms(tsg_handle_t handle, int stage);

/* Human-readable text of the last error on this handle (or globally for
 * NULL). */
const char *tsg_last_error(tsg_handle_t handle);

/* ====================================================================
 * Matrix Market input (host only; replaces the reference loader
 * mmio_allinone, mmio_highlevel.h:593-759, with the same semantics).
 * Reads a coordinate file into a freshly malloc'ed 0-based CSR: entries of a
 * row in file order, no sorting, duplicates kept; symmetric and hermitian
 * files mirrored off the diagonal (same value); pattern entries 1.0, integer
 * entries their value, complex entries their real part.  *symmetric reports
 * the mirroring.  Release the arrays with tsg_free.  Returns TSG_OK, a
 * TSG_MTX_ERR_* code (the reference's -1/-2/-4 less 100, plus bad entries),
 * TSG_ERR_OVERFLOW (more than 2^31-1 entries) or TSG_ERR_ALLOC.
 * ==================================================================== */
enum {
    TSG_MTX_ERR_OPEN = -101,     /* file cannot be opened                  */
    TSG_MTX_ERR_BANNER = -102,   /* not "%%MatrixMarket matrix coordinate" */
    TSG_MTX_ERR_SIZE = -104,     /* bad size line (or non-square symmetric) */
    TSG_MTX_ERR_ENTRIES = -105   /* short file or an index out of range    */
};
int tsg_read_mtx(const char *path, int *m, int *n, int64_t *nnz, int *symmetric, int **ia,
                 int **ja, double **a);
void tsg_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* TSG_H */

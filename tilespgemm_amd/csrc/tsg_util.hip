// Utilities: workspace arena, device-wide scans, stage timer, CSR transpose,
// intermediate-product counts.
#include "tsg_internal.h"
#include "tsg_device.h"

#include <algorithm>
#include <cstring>
#include <hipcub/hipcub.hpp>

// ================================================================= arena
Arena::~Arena() { release(); }

void Arena::release() {
    if (!external_)
        for (auto &b : blocks_) (void)hipFree(b.base);
    blocks_.clear();
    cur_ = 0;
    held_ = 0;
    external_ = false;
}

void Arena::attach(void *base, size_t bytes, const char *what) {
    release();
    external_ = true;
    what_ = what;
    blocks_.push_back(Block{static_cast<char *>(base), bytes & ~size_t(255), 0});
}

size_t Arena::used() const {
    size_t u = 0;
    for (const auto &b : blocks_) u += b.used;
    return u;
}

void Arena::reset() {
    for (auto &b : blocks_) b.used = 0;
    cur_ = 0;
}

void *Arena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    while (cur_ < blocks_.size()) {
        Block &b = blocks_[cur_];
        if (b.size - b.used >= bytes) {
            void *p = b.base + b.used;
            b.used += bytes;
            return p;
        }
        cur_++;
    }
    if (external_)
        throw TsgError{TSG_ERR_ALLOC, std::string(what_) + " of " +
                                          std::to_string(blocks_.empty() ? 0 : blocks_[0].size) +
                                          " bytes exhausted: " + std::to_string(used()) +
                                          " bytes in use, " + std::to_string(bytes) +
                                          " more needed"};
    size_t sz = std::max(bytes, size_t(64) << 20);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw TsgError{TSG_ERR_ALLOC, "workspace hipMalloc of " + std::to_string(sz) +
                                          " bytes failed: " + hipGetErrorString(e)};
    }
    blocks_.push_back(Block{static_cast<char *>(p), sz, bytes});
    cur_ = blocks_.size() - 1;
    held_ += sz;
    return p;
}

// ================================================================= scans
// Three-phase exclusive scan: per-block totals, recursive scan of the totals,
// per-block rescan with the block's offset.  1024 elements per 256-thread
// block (4 per thread).
namespace {
constexpr int SCAN_T = 256;
constexpr int SCAN_PER = 4;
constexpr int SCAN_BLK = SCAN_T * SCAN_PER;

template <class T>
__device__ __forceinline__ long long block_excl_scan(long long x, long long *total) {
    __shared__ long long wsum[SCAN_T / 64];
    long long inc = wave_incl_scan64(x);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 63) wsum[w] = inc;
    __syncthreads();
    long long off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SCAN_T / 64; i++) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + inc - x;
}

template <class T>
__global__ void __launch_bounds__(SCAN_T) k_scan_reduce(const T *in, size_t n, long long *sums) {
    size_t base = (size_t)blockIdx.x * SCAN_BLK + threadIdx.x * SCAN_PER;
    long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; i++)
        if (base + i < n) s += (long long)in[base + i];
    long long tot;
    (void)block_excl_scan<T>(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class T>
__global__ void __launch_bounds__(SCAN_T) k_scan_apply(const T *in, T *out, size_t n,
                                                        const long long *offs, int has_offs,
                                                        long long *total) {
    size_t base = (size_t)blockIdx.x * SCAN_BLK + threadIdx.x * SCAN_PER;
    T v[SCAN_PER];
    long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; i++) {
        v[i] = base + i < n ? in[base + i] : T(0);
        s += (long long)v[i];
    }
    long long tot;
    long long ex = block_excl_scan<T>(s, &tot);
    long long run = ex + (has_offs ? offs[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < SCAN_PER; i++) {
        if (base + i < n) out[base + i] = (T)run;
        run += (long long)v[i];
    }
    // out[n] = grand total, written by the thread that owns element n-1; the
    // exact 64-bit total also goes to *total when asked for (T may be narrower)
    if (n > 0 && base <= n - 1 && n - 1 < base + SCAN_PER) {
        out[n] = (T)run;
        if (total) *total = run;
    }
    if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        out[0] = T(0);
        if (total) *total = 0;
    }
}

size_t level_elems(size_t n) { return (n + SCAN_BLK - 1) / SCAN_BLK; }

template <class T>
void scan_rec(const T *in, T *out, size_t n, long long *tmp, hipStream_t s,
              long long *total = nullptr) {
    size_t nb = level_elems(n);
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(SCAN_T), 0, s, in, out, n,
                           (const long long *)nullptr, 0, total);
        TSG_LAUNCH_CHECK();
        return;
    }
    long long *sums = tmp;
    long long *rest = tmp + nb + 1;
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(SCAN_T), 0, s, in, n, sums);
    TSG_LAUNCH_CHECK();
    scan_rec<long long>(sums, sums, nb, rest, s);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(SCAN_T), 0, s, in, out, n,
                       (const long long *)sums, 1, total);
    TSG_LAUNCH_CHECK();
}
}  // namespace

size_t scan_tmp_elems(size_t n) {
    size_t total = 0;
    size_t nb = level_elems(n);
    while (nb > 1) {
        total += nb + 1;
        nb = level_elems(nb);
    }
    return total + 8;
}

// The apply pass reads in[] before any block writes out[] only within its own
// block range, so in == out is safe: every block reads its elements before
// writing them, and out[n] lies outside every block's input range.
void exclusive_scan_i64(const int64_t *in, int64_t *out, size_t n, int64_t *tmp, hipStream_t s) {
    scan_rec<long long>((const long long *)in, (long long *)out, n, (long long *)tmp, s);
}

void exclusive_scan_i32(const int32_t *in, int32_t *out, size_t n, int64_t *tmp, hipStream_t s,
                        int64_t *total) {
    scan_rec<int>(in, out, n, (long long *)tmp, s, (long long *)total);
}

// ================================================================= timer
void StageTimer::clear() {
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.clear();
    names.clear();
    ms.clear();
}

void StageTimer::begin(const char *name) {
    if (!enabled) return;
    hipEvent_t a, b;
    TSG_HIP_TRY(hipEventCreate(&a));
    TSG_HIP_TRY(hipEventCreate(&b));
    TSG_HIP_TRY(hipEventRecord(a, stream));
    names.emplace_back(name);
    ev.push_back(a);
    ev.push_back(b);
}

void StageTimer::end() {
    if (!enabled || ev.empty()) return;
    TSG_HIP_TRY(hipEventRecord(ev.back(), stream));
}

void StageTimer::finalize() {
    if (!enabled) return;
    ms.assign(names.size(), 0.0);
    for (size_t i = 0; i < names.size(); i++) {
        TSG_HIP_TRY(hipEventSynchronize(ev[2 * i + 1]));
        float t = 0.f;
        TSG_HIP_TRY(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
        ms[i] = t;
    }
}

StageTimer::~StageTimer() { clear(); }

// ======================================================= validation
// flags: bit 0 bad row pointers, bit 1 column out of range, bit 2 a row not
// strictly ascending (unsorted or duplicate columns).  Reads stay inside
// [0, nnz) whatever the row pointers hold.
// The kernels below run between kernels that stream gigabytes, so their
// operands are cold in every cache: each keeps its chain of dependent loads
// to two or three (no searches, no per-row loops), which is what sets their
// duration.

// Threads check row pointers t < m and entries t < nnz (grid-stride).  Rows
// are sorted and duplicate-free iff every descent ja[p-1] >= ja[p] (p > 0)
// sits at the first entry of a non-empty row, i.e. iff the number of all
// descents equals the number of descents at row starts: out[1] counts the
// former, out[2] the latter (summed per block, one atomic per block and
// counter).  Every access stays inside ia[0..m] and ja[0..nnz) whatever the
// contents.
__global__ void __launch_bounds__(256) k_validate(int m, int n, const int *__restrict__ ia,
                                                  const int *__restrict__ ja, int64_t nnz,
                                                  int *__restrict__ out) {
    __shared__ int s_cnt[2][4];
    int f = 0;
    int at_start = 0, descent = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t lim = nnz > m ? nnz : m;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < lim; t += stride) {
        if (t < m) {
            const int64_t lo = ia[t], hi = ia[t + 1];
            if (lo > hi || lo < 0 || hi > nnz) f |= 1;
            if (t == 0 && lo != 0) f |= 1;
            if (t == m - 1 && hi != nnz) f |= 1;
            if (lo > 0 && lo < hi && hi <= nnz) at_start += ja[lo - 1] >= ja[lo];
        }
        if (t < nnz) {
            const int c = ja[t];
            if (c < 0 || c >= n) f |= 2;
            if (t > 0) descent += ja[t - 1] >= c;
        }
    }
    descent = wave_sum(descent);
    at_start = wave_sum(at_start);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_cnt[0][w] = descent;
        s_cnt[1][w] = at_start;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int d = s_cnt[0][0] + s_cnt[0][1] + s_cnt[0][2] + s_cnt[0][3];
        const int st = s_cnt[1][0] + s_cnt[1][1] + s_cnt[1][2] + s_cnt[1][3];
        if (d) atomicAdd(&out[1], d);
        if (st) atomicAdd(&out[2], st);
    }
    if (f) atomicOr(&out[0], f);
}

int validate_csr_pair(const CsrIn &A, const CsrIn &B, Arena &ws, hipStream_t s,
                      unsigned long long *density_sample,
                      const std::function<void()> &before_sync) {
    // six flag/count ints, then the two sums of the density sample
    unsigned long long *buf = ws.alloc_n<unsigned long long>(5);
    int *d = reinterpret_cast<int *>(buf);
    TSG_HIP_TRY(hipMemsetAsync(buf, 0, 5 * sizeof(unsigned long long), s));
    if (density_sample) tile_sample_launch(A, buf + 3, s);
    // a fixed grid (grid-stride inside): few per-block atomics on the counters
    auto threads = [](const CsrIn &X) {
        return std::min(std::max((size_t)X.rows, (size_t)X.nnz), (size_t)256 * 2048);
    };
    if (A.rows > 0) {
        hipLaunchKernelGGL(k_validate, dim3(grid_for(threads(A), 256)), dim3(256), 0, s, A.rows,
                           A.cols, A.ptr, A.idx, A.nnz, d);
        TSG_LAUNCH_CHECK();
    }
    if (B.rows > 0 && (B.ptr != A.ptr || B.idx != A.idx || B.cols != A.cols)) {
        hipLaunchKernelGGL(k_validate, dim3(grid_for(threads(B), 256)), dim3(256), 0, s, B.rows,
                           B.cols, B.ptr, B.idx, B.nnz, d + 3);
        TSG_LAUNCH_CHECK();
    }
    if (before_sync) before_sync();
    unsigned long long hb[5] = {0, 0, 0, 0, 0};
    TSG_HIP_TRY(hipMemcpyAsync(hb, buf, sizeof(hb), hipMemcpyDeviceToHost, s));
    TSG_HIP_TRY(hipStreamSynchronize(s));
    int h[6];
    std::memcpy(h, hb, sizeof(h));
    if (density_sample) {
        density_sample[0] = hb[3];
        density_sample[1] = hb[4];
    }
    int fa = h[0], fb = h[3];
    if (h[1] != h[2]) fa |= 4;
    if (h[4] != h[5]) fb |= 4;
    if (A.rows == 0 && A.nnz != 0) fa |= 1;
    if (B.rows == 0 && B.nnz != 0) fb |= 1;
    return fa | (fb << 8);
}

// ======================================================= row products
// rowprod[i] = sum over A's entries a_ik in row i of |row k of B|
// (main_spgemm-col.cu:4950-4955), as P[ia[i+1]] - P[ia[i]] with P the
// exclusive scan of the per-entry counts |row ja[p] of B|.  A must be a valid
// CSR with columns indexing rows of B.
__global__ void __launch_bounds__(256) k_entry_products(int64_t nnz, const int *__restrict__ ja,
                                                        const int *__restrict__ ib,
                                                        int64_t *__restrict__ ep) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int k = ja[p];
    ep[p] = ib[k + 1] - ib[k];
}

__global__ void __launch_bounds__(256) k_row_diff(int m, const int *__restrict__ ia,
                                                  const int64_t *__restrict__ P,
                                                  int64_t *__restrict__ rowprod) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) rowprod[i] = P[ia[i + 1]] - P[ia[i]];
}

__global__ void k_zero_i64(int64_t n, int64_t *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = 0;
}

void csr_row_products(int m, const int *ia, const int *ja, int64_t nnz, const int *ib,
                      int64_t *rowprod, Arena &ws, hipStream_t s) {
    if (m <= 0) return;
    if (nnz <= 0) {
        hipLaunchKernelGGL(k_zero_i64, dim3(grid_for((size_t)m, 256)), dim3(256), 0, s,
                           (int64_t)m, rowprod);
        TSG_LAUNCH_CHECK();
        return;
    }
    int64_t *P = ws.alloc_n<int64_t>((size_t)nnz + 1);
    int64_t *tmp = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)nnz + 1));
    hipLaunchKernelGGL(k_entry_products, dim3(grid_for((size_t)nnz, 256)), dim3(256), 0, s, nnz,
                       ja, ib, P);
    TSG_LAUNCH_CHECK();
    exclusive_scan_i64(P, P, (size_t)nnz, tmp, s);
    hipLaunchKernelGGL(k_row_diff, dim3(grid_for((size_t)m, 256)), dim3(256), 0, s, m, ia, P,
                       rowprod);
    TSG_LAUNCH_CHECK();
}

// ============================================================ transpose
// Stable CSR transpose: a radix sort of (column, position) pairs keeps equal
// columns in position order, i.e. ascending source row (utils.h:158-195).
__global__ void k_col_count(int64_t nnz, const int *__restrict__ ja, int *__restrict__ cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) atomicAdd(&cnt[ja[i]], 1);
}

__global__ void k_row_index(int m, const int *__restrict__ ia, int *__restrict__ rowidx) {
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const int l = threadIdx.x & 7;
    if (row >= m) return;
    for (int p = ia[row] + l; p < ia[row + 1]; p += 8) rowidx[p] = row;
}

__global__ void k_iota(int64_t n, int *__restrict__ v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (int)i;
}

__global__ void k_gather_t(int64_t nnz, const int *__restrict__ pos, const int *__restrict__ rowidx,
                           const double *__restrict__ a, int *__restrict__ ti,
                           double *__restrict__ tv, int *__restrict__ tpos) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    int p = pos[q];
    ti[q] = rowidx[p];
    if (a) tv[q] = a[p];
    if (tpos) tpos[q] = p;
}

void csr_transpose(int m, int n, const int *ia, const int *ja, const double *a, int64_t nnz,
                   int *tp, int *ti, double *tv, Arena &ws, hipStream_t s, int *tpos) {
    int *cnt = ws.alloc_n<int>((size_t)n + 1);
    int64_t *tmp = ws.alloc_n<int64_t>(scan_tmp_elems((size_t)n));
    TSG_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(int) * ((size_t)n + 1), s));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_col_count, dim3(grid_for((size_t)nnz, 256)), dim3(256), 0, s, nnz,
                           ja, cnt);
        TSG_LAUNCH_CHECK();
    }
    exclusive_scan_i32(cnt, tp, (size_t)n, tmp, s);
    if (nnz <= 0) return;
    int *rowidx = ws.alloc_n<int>((size_t)nnz);
    int *pos_in = ws.alloc_n<int>((size_t)nnz);
    int *pos_out = ws.alloc_n<int>((size_t)nnz);
    int *keys_out = ws.alloc_n<int>((size_t)nnz);
    hipLaunchKernelGGL(k_row_index, dim3(grid_for((size_t)m * 8, 256)), dim3(256), 0, s, m, ia,
                       rowidx);
    TSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_iota, dim3(grid_for((size_t)nnz, 256)), dim3(256), 0, s, nnz, pos_in);
    TSG_LAUNCH_CHECK();
    int end_bit = 1;
    while (end_bit < 31 && (1ll << end_bit) < (long long)n) end_bit++;
    size_t tbytes = 0;
    TSG_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, ja, keys_out, pos_in, pos_out,
                                                   (int)nnz, 0, end_bit, s));
    void *tstore = ws.alloc(tbytes);
    TSG_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tstore, tbytes, ja, keys_out, pos_in, pos_out,
                                                   (int)nnz, 0, end_bit, s));
    hipLaunchKernelGGL(k_gather_t, dim3(grid_for((size_t)nnz, 256)), dim3(256), 0, s, nnz,
                       pos_out, rowidx, a, ti, tv, tpos);
    TSG_LAUNCH_CHECK();
}

// ====================================================== stable key order
// order[j] = index of the j-th smallest of keys[0..n) (each in [0, 2^end_bit)),
// equal keys in index order.
void sort_order_i32(const int *keys, int64_t n, int end_bit, int *order, Arena &ws,
                    hipStream_t s) {
    if (n <= 0) return;
    TSG_CHECK(n < ((int64_t)1 << 31), TSG_ERR_OVERFLOW, "sort_order_i32: more than 2^31 keys");
    int *iota = ws.alloc_n<int>((size_t)n);
    int *keys_out = ws.alloc_n<int>((size_t)n);
    hipLaunchKernelGGL(k_iota, dim3(grid_for((size_t)n, 256)), dim3(256), 0, s, n, iota);
    TSG_LAUNCH_CHECK();
    size_t tbytes = 0;
    TSG_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, keys, keys_out, iota, order,
                                                   (int)n, 0, end_bit, s));
    void *tstore = ws.alloc(tbytes);
    TSG_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tstore, tbytes, keys, keys_out, iota, order,
                                                   (int)n, 0, end_bit, s));
}

// ================================================== stable row sort of B
// Row-sorted copy of a CSR matrix whose rows may be unsorted or hold
// duplicates: two stable transposes (each a stable radix sort of columns by
// position) give every row ascending columns with equal columns in their
// original order; perm[q] is the original position of sorted entry q.  The
// row pointers are unchanged.
__global__ void k_compose_pos(int64_t n, const int *__restrict__ outer,
                              const int *__restrict__ inner, int *__restrict__ perm) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) perm[q] = inner[outer[q]];
}

__global__ void k_gather_vals(int64_t n, const int *__restrict__ perm,
                              const double *__restrict__ src, double *__restrict__ dst) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = src[perm[q]];
}

void csr_sort_rows(int m, int n, const int *ia, const int *ja, int64_t nnz, int *sorted_ja,
                   int *perm, Arena &keep, Arena &tmp, hipStream_t s) {
    if (nnz <= 0) return;
    int *tp = tmp.alloc_n<int>((size_t)n + 1);
    int *ti = tmp.alloc_n<int>((size_t)nnz);
    int *pos1 = tmp.alloc_n<int>((size_t)nnz);
    int *tp2 = tmp.alloc_n<int>((size_t)m + 1);
    int *pos2 = tmp.alloc_n<int>((size_t)nnz);
    csr_transpose(m, n, ia, ja, nullptr, nnz, tp, ti, nullptr, tmp, s, pos1);
    csr_transpose(n, m, tp, ti, nullptr, nnz, tp2, sorted_ja, nullptr, tmp, s, pos2);
    hipLaunchKernelGGL(k_compose_pos, dim3(grid_for((size_t)nnz, 256)), dim3(256), 0, s, nnz,
                       pos2, pos1, perm);
    TSG_LAUNCH_CHECK();
    (void)keep;
}

void gather_values(int64_t n, const int *perm, const double *src, double *dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for((size_t)n, 256)), dim3(256), 0, s, n, perm,
                       src, dst);
    TSG_LAUNCH_CHECK();
}

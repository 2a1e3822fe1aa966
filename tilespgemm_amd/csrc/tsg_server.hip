// Request server with continuous batching (include/tsg.h, tsg_server_*).
//
// Requests carry device-resident operands.  Submission only enqueues; an
// estimator thread (own handle and stream) computes each request's output
// upper bound from its intermediate products (the admission estimate) and
// moves it, in arrival order, to the admission queue.  Each worker thread owns
// a tsg handle (its own stream) and repeatedly takes, from the front of that
// queue, every request that fits the memory budget and runs them as one
// batched product (block-diagonal layout, tsg_batch.hip).  With several
// workers, the next batch is formed and launched while earlier ones are still
// on the device.  Output buffers come from a best-fit device cache and are
// returned when the caller releases the result.
#include "tsg_internal.h"
#include "tsg_device.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Best-fit cache of device blocks (hipMalloc on a miss, reuse after put).
class DevicePool {
public:
    ~DevicePool() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto &kv : size_) (void)hipFree(kv.first);
        size_.clear();
        free_.clear();
    }
    void *get(size_t bytes) {
        bytes = std::max<size_t>(256, (bytes + 255) & ~size_t(255));
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(bytes);
            if (it != free_.end() && it->first <= 2 * bytes + (size_t(1) << 20)) {
                void *p = it->second;
                free_.erase(it);
                return p;
            }
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            // drop the cache and retry once
            clear();
            e = hipMalloc(&p, bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw TsgError{TSG_ERR_ALLOC, "server output allocation failed"};
            }
        }
        std::lock_guard<std::mutex> lk(mu_);
        size_[p] = bytes;
        return p;
    }
    void put(void *p) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(size_[p], p);
    }
    void clear() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto &kv : free_) {
            (void)hipFree(kv.second);
            size_.erase(kv.second);
        }
        free_.clear();
    }

private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::unordered_map<void *, size_t> size_;
};

// ------------------------------------------------------ admission estimates
// The estimates of every request taken from the submission queue come from
// one host round trip: each request's A (row pointers; columns in [0, n))
// and, unless B is A itself, its B (columns in [0, k)) are checked, and A's
// intermediate products summed (main_spgemm-col.cu:4950-4955), over the
// requests' arrays laid end to end.  One list entry per matrix checked;
// ib is set on A's entries only, whose products are counted.
struct EstDesc {
    const int *ptr, *idx;
    const int *ib;
    int rows, cols;
    int64_t nnz;
    int req;
};

constexpr int EST_PER = 8;

// Largest p in [0, count) with off[p] <= x.
__device__ __forceinline__ int est_find(const int64_t *__restrict__ off, int count, int64_t x) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Row pointers: ptr[0] = 0, ptr[rows] = nnz, non-decreasing, inside
// [0, nnz].  Bit 0 of flags[req] marks a failure.  A block takes
// EST_PER * 256 consecutive rows of the list; the entry of its first row is
// searched once, each thread steps to the next entry at a boundary.  Reads
// stay inside ptr[0..rows].
__global__ void __launch_bounds__(256) k_est_rows(int count, const EstDesc *__restrict__ desc,
                                                  const int64_t *__restrict__ roff, int64_t rows,
                                                  int *__restrict__ flags) {
    const int64_t r0 = (int64_t)blockIdx.x * (256 * EST_PER);
    if (r0 >= rows) return;
    __shared__ int s_p;
    if (threadIdx.x == 0) s_p = est_find(roff, count, r0);
    __syncthreads();
    int p = s_p;
#pragma unroll
    for (int j = 0; j < EST_PER; j++) {
        const int64_t r = r0 + j * 256 + threadIdx.x;
        if (r >= rows) break;
        while (r >= roff[p + 1]) p++;   // stops below count: r < roff[count]
        const EstDesc &d = desc[p];
        const int64_t i = r - roff[p];
        const int64_t lo = d.ptr[i], hi = d.ptr[i + 1];
        const bool bad = lo > hi || lo < 0 || hi > d.nnz || (i == 0 && lo != 0) ||
                         (i == d.rows - 1 && hi != d.nnz);
        if (bad) atomicOr(&flags[d.req], 1);
    }
}

// Columns in [0, cols) (bit 1 of flags[req]); with PRODUCTS, the length of
// B's row at each valid column added to prods[req].  Laid out as
// k_est_rows; a lane's sums are flushed when its entry changes, and at the
// end a wave whose lanes all hold the same entry adds one sum.
template <bool PRODUCTS>
__global__ void __launch_bounds__(256) k_est_entries(int count, const EstDesc *__restrict__ desc,
                                                     const int64_t *__restrict__ eoff, int64_t nnz,
                                                     int *__restrict__ flags,
                                                     unsigned long long *__restrict__ prods) {
    const int64_t q0 = (int64_t)blockIdx.x * (256 * EST_PER);
    if (q0 >= nnz) return;
    __shared__ int s_p;
    if (threadIdx.x == 0) s_p = est_find(eoff, count, q0);
    __syncthreads();
    int p = s_p;
    int cur = -1;
    unsigned long long acc = 0;
#pragma unroll
    for (int j = 0; j < EST_PER; j++) {
        const int64_t q = q0 + j * 256 + threadIdx.x;
        if (q >= nnz) break;
        while (q >= eoff[p + 1]) p++;   // stops below count: q < eoff[count]
        const EstDesc &d = desc[p];
        const int c = d.idx[q - eoff[p]];
        if (c < 0 || c >= d.cols) {
            atomicOr(&flags[d.req], 2);
            continue;
        }
        if (PRODUCTS) {
            if (p != cur) {
                if (cur >= 0 && acc) atomicAdd(&prods[desc[cur].req], acc);
                cur = p;
                acc = 0;
            }
            acc += (unsigned long long)((int64_t)d.ib[c + 1] - (int64_t)d.ib[c]);
        }
    }
    if (PRODUCTS) {
        const int c0 = __shfl(cur, 0, 64);
        if (__ballot(cur != c0) == 0) {
            const unsigned long long sum = group_sum<64>(acc);
            if (lane_id() == 0 && c0 >= 0 && sum) atomicAdd(&prods[desc[c0].req], sum);
        } else if (cur >= 0 && acc) {
            atomicAdd(&prods[desc[cur].req], acc);
        }
    }
}

struct BatchOut {
    void *ic = nullptr, *jc = nullptr, *c = nullptr;
    int refs = 0;
};

struct Req {
    uint64_t ticket = 0;
    tsg_product_t prod{};
    size_t est = 0;
    int64_t products = 0;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    tsg_result_t res{};
    std::shared_ptr<BatchOut> out;
};

}  // namespace

struct tsg_server_s {
    int device = 0;
    tsg_server_config_t cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_est;
    std::deque<Req *> pending;         // submitted, not yet estimated
    std::deque<Req *> queue;           // estimated, waiting for admission
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    bool est_done = false;             // estimator has drained and exited
    size_t in_flight = 0;              // estimated bytes of admitted, unreleased requests
    int running = 0;                   // batches on the device now
    std::vector<std::thread> workers;
    std::thread estimator;
    tsg_handle_t est = nullptr;        // owned by the estimator thread
    DevicePool pool;

    void estimate(std::deque<Req *> &take, std::vector<int> &status);
    void run_batch(tsg_handle_t h, std::vector<Req *> &batch);
    void worker_main();
    void estimator_main();
};

void tsg_server_s::estimator_main() {
    const bool dev_ok = hipSetDevice(device) == hipSuccess;
    for (;;) {
        std::deque<Req *> take;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_est.wait(lk, [&] { return stop || !pending.empty(); });
            if (pending.empty()) {
                est_done = true;
                cv_work.notify_all();
                return;
            }
            take.swap(pending);
        }
        std::vector<int> status(take.size(), TSG_OK);
        if (dev_ok)
            estimate(take, status);
        else
            std::fill(status.begin(), status.end(), (int)TSG_ERR_NODEVICE);
        std::lock_guard<std::mutex> lk(mu);
        bool failed = false;
        for (size_t i = 0; i < take.size(); i++) {
            Req *r = take[i];
            if (status[i] == TSG_OK) {
                queue.push_back(r);
                continue;
            }
            r->done = true;
            r->t_start = r->t_done = Clock::now();
            r->res.status = status[i];
            r->res.m = r->prod.m;
            failed = true;
        }
        cv_work.notify_all();
        if (failed) cv_done.notify_all();
    }
}

void tsg_server_s::estimate(std::deque<Req *> &take, std::vector<int> &status) {
    const int n = (int)take.size();
    std::vector<EstDesc> la, lb;          // A of every request, B where it is not A
    std::vector<int64_t> ra(1, 0), ea(1, 0), rb(1, 0), eb(1, 0);
    for (int i = 0; i < n; i++) {
        const tsg_product_t &x = take[i]->prod;
        int rc = TSG_OK;
        if (x.m < 0 || x.n < 0 || x.k < 0 || x.nnz_a < 0 || x.nnz_b < 0)
            rc = TSG_ERR_INVALID;
        else if (!x.ia || !x.ib || (x.nnz_a > 0 && !x.ja) || (x.nnz_b > 0 && !x.jb))
            rc = TSG_ERR_INVALID;
        else if (x.nnz_a >= (int64_t(1) << 31) || x.nnz_b >= (int64_t(1) << 31))
            rc = TSG_ERR_OVERFLOW;
        else if ((x.m == 0 && x.nnz_a != 0) || (x.n == 0 && x.nnz_b != 0))
            rc = TSG_ERR_INVALID;
        status[i] = rc;
        if (rc != TSG_OK) continue;
        la.push_back(EstDesc{x.ia, x.ja, x.ib, x.m, x.n, x.nnz_a, i});
        ra.push_back(ra.back() + x.m);
        ea.push_back(ea.back() + x.nnz_a);
        const bool b_is_a = x.ib == x.ia && x.jb == x.ja && x.n == x.m && x.k == x.n &&
                            x.nnz_b == x.nnz_a;
        if (!b_is_a) {
            lb.push_back(EstDesc{x.ib, x.jb, nullptr, x.n, x.k, x.nnz_b, i});
            rb.push_back(rb.back() + x.n);
            eb.push_back(eb.back() + x.nnz_b);
        }
    }
    std::vector<unsigned long long> prods(n, 0);
    if (!la.empty()) {
        try {
            hipStream_t st = est->stream;
            Arena &ws = est->aux;
            ws.reset();
            const int na = (int)la.size(), nb = (int)lb.size();
            EstDesc *d_la = ws.alloc_n<EstDesc>(na + nb + 1);
            EstDesc *d_lb = d_la + na;
            int64_t *d_off = ws.alloc_n<int64_t>(2 * ((size_t)na + 1) + 2 * ((size_t)nb + 1));
            int64_t *d_ra = d_off, *d_ea = d_ra + na + 1, *d_rb = d_ea + na + 1,
                    *d_eb = d_rb + nb + 1;
            // flags (int) then sums (u64) in one block, cleared and read back together
            unsigned long long *d_sum = ws.alloc_n<unsigned long long>(2 * (size_t)n);
            int *d_flag = reinterpret_cast<int *>(d_sum + n);
            std::vector<EstDesc> dl(la);
            dl.insert(dl.end(), lb.begin(), lb.end());
            std::vector<int64_t> offs(ra);
            offs.insert(offs.end(), ea.begin(), ea.end());
            offs.insert(offs.end(), rb.begin(), rb.end());
            offs.insert(offs.end(), eb.begin(), eb.end());
            TSG_HIP_TRY(hipMemcpyAsync(d_la, dl.data(), sizeof(EstDesc) * dl.size(),
                                       hipMemcpyHostToDevice, st));
            TSG_HIP_TRY(hipMemcpyAsync(d_off, offs.data(), sizeof(int64_t) * offs.size(),
                                       hipMemcpyHostToDevice, st));
            TSG_HIP_TRY(hipMemsetAsync(d_sum, 0, 2 * sizeof(unsigned long long) * n, st));
            auto grid = [](int64_t w) { return dim3(grid_for((size_t)w, 256 * EST_PER)); };
            if (ra[na] > 0)
                hipLaunchKernelGGL(k_est_rows, grid(ra[na]), dim3(256), 0, st, na, d_la, d_ra,
                                   ra[na], d_flag);
            if (ea[na] > 0)
                hipLaunchKernelGGL(k_est_entries<true>, grid(ea[na]), dim3(256), 0, st, na, d_la,
                                   d_ea, ea[na], d_flag, d_sum);
            if (nb > 0 && rb[nb] > 0)
                hipLaunchKernelGGL(k_est_rows, grid(rb[nb]), dim3(256), 0, st, nb, d_lb, d_rb,
                                   rb[nb], d_flag);
            if (nb > 0 && eb[nb] > 0)
                hipLaunchKernelGGL((k_est_entries<false>), grid(eb[nb]), dim3(256), 0, st, nb,
                                   d_lb, d_eb, eb[nb], d_flag, nullptr);
            TSG_LAUNCH_CHECK();
            std::vector<unsigned long long> back(2 * (size_t)n);
            TSG_HIP_TRY(hipMemcpyAsync(back.data(), d_sum, sizeof(unsigned long long) * back.size(),
                                       hipMemcpyDeviceToHost, st));
            TSG_HIP_TRY(hipStreamSynchronize(st));
            const int *flags = reinterpret_cast<const int *>(back.data() + n);
            for (int i = 0; i < n; i++) {
                prods[i] = back[i];
                if (status[i] == TSG_OK && flags[i]) status[i] = TSG_ERR_INVALID;
            }
        } catch (const TsgError &e) {
            for (int i = 0; i < n; i++)
                if (status[i] == TSG_OK) status[i] = e.code;
        }
    }
    for (int i = 0; i < n; i++) {
        Req *r = take[i];
        const uint64_t cap = std::min<uint64_t>(prods[i], (uint64_t)r->prod.m * (uint64_t)r->prod.k);
        r->est = status[i] == TSG_OK ? 16 * ((size_t)r->prod.m + 2) + 12 * (size_t)cap : 0;
        r->products = status[i] == TSG_OK ? (int64_t)prods[i] : 0;
    }
}

void tsg_server_s::run_batch(tsg_handle_t h, std::vector<Req *> &batch) {
    const int n = (int)batch.size();
    std::vector<tsg_product_t> prods(n);
    int rows = 0;
    for (int i = 0; i < n; i++) {
        prods[i] = batch[i]->prod;
        rows += batch[i]->prod.m;
    }
    auto t0 = Clock::now();
    for (auto *r : batch) r->t_start = t0;
    int status = TSG_OK;
    auto out = std::make_shared<BatchOut>();
    std::vector<int64_t> per(n, 0);
    void *ic_glob = nullptr;
    try {
        ic_glob = pool.get(sizeof(int64_t) * ((size_t)rows + 1));
        int64_t tot = 0;
        int rc = tsg_batch_symbolic(h, n, prods.data(), (int64_t *)ic_glob, &tot, per.data());
        if (rc != TSG_OK) throw TsgError{rc, tsg_last_error(h)};
        out->ic = pool.get(sizeof(int64_t) * ((size_t)rows + n));
        out->jc = pool.get(sizeof(int) * ((size_t)tot + 1));
        out->c = pool.get(sizeof(double) * ((size_t)tot + 1));
        rc = tsg_batch_numeric(h, n, prods.data(), (const int64_t *)ic_glob, (int *)out->jc,
                               (double *)out->c);
        if (rc != TSG_OK) throw TsgError{rc, tsg_last_error(h)};
        batch_rebase(h, (const int64_t *)ic_glob, (int64_t *)out->ic);
        TSG_HIP_TRY(hipStreamSynchronize(h->stream));
    } catch (const TsgError &e) {
        status = e.code;
    }
    pool.put(ic_glob);
    auto t1 = Clock::now();
    std::lock_guard<std::mutex> lk(mu);
    running--;
    cv_work.notify_all();
    int64_t e0 = 0, r0 = 0;
    out->refs = n;
    for (int i = 0; i < n; i++) {
        Req *r = batch[i];
        r->t_done = t1;
        r->done = true;
        r->out = out;
        r->res.status = status;
        r->res.m = r->prod.m;
        r->res.nnz = status == TSG_OK ? per[i] : 0;
        r->res.ic = status == TSG_OK ? (const int64_t *)out->ic + r0 + i : nullptr;
        r->res.jc = status == TSG_OK ? (const int *)out->jc + e0 : nullptr;
        r->res.c = status == TSG_OK ? (const double *)out->c + e0 : nullptr;
        r->res.queue_ms = ms_between(r->t_submit, r->t_start);
        r->res.total_ms = ms_between(r->t_submit, t1);
        r->res.batch_size = n;
        r->res.products = r->products;
        e0 += per[i];
        r0 += r->prod.m;
    }
    cv_done.notify_all();
}

void tsg_server_s::worker_main() {
    if (hipSetDevice(device) != hipSuccess) return;
    tsg_handle_t h = nullptr;
    if (tsg_create(&h, device) != TSG_OK) return;
    tsg_set_method(h, cfg.method);
    if (cfg.tile_density > 0.0) tsg_set_tile_density(h, cfg.tile_density);
    if (cfg.bin3 > 0) tsg_set_row_bins(h, cfg.bin1, cfg.bin2, cfg.bin3);
    for (;;) {
        std::vector<Req *> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (est_done && queue.empty()) {
                    lk.unlock();
                    tsg_destroy(h);
                    return;
                }
                // admission: FIFO, every waiting request that fits the budget
                // (estimates of admitted requests count until released).  When
                // no batch is running, the head request is admitted even if it
                // does not fit, so a caller holding results it has not released
                // while waiting for a later ticket still makes progress; at
                // shutdown the queue drains regardless of the budget
                while (!queue.empty() && (int)batch.size() < cfg.max_batch) {
                    Req *r = queue.front();
                    const bool fits = stop || in_flight + r->est <= cfg.mem_budget;
                    const bool alone = running == 0 && batch.empty();
                    if (!fits && !alone) break;
                    queue.pop_front();
                    in_flight += r->est;
                    batch.push_back(r);
                }
                if (!batch.empty()) {
                    running++;
                    break;
                }
                cv_work.wait(lk);
            }
        }
        run_batch(h, batch);
    }
}

extern "C" {

int tsg_server_create(tsg_server_t *server, int device, const tsg_server_config_t *config) {
    if (!server || !config || config->workers < 1 || config->max_batch < 1) return TSG_ERR_INVALID;
    *server = nullptr;
    auto s = std::make_unique<tsg_server_s>();
    s->device = device;
    s->cfg = *config;
    if (s->cfg.mem_budget == 0) s->cfg.mem_budget = size_t(64) << 30;
    // a server's requests are many mid-size products, whose rows of 513 to
    // 1024 products run faster on the big-row path than in the hash-capable
    // small-row kernel (R-MAT 2^18 A*A^T with 4 draws per row 6.04 -> 5.43
    // ms; the 512-product ER batch 67.3 -> 62.7 ms); the library default of
    // 1024 suits single products of a million rows and more (R-MAT 2^20:
    // 183 -> 187 ms with 512)
    if (s->cfg.bin3 <= 0) {
        s->cfg.bin1 = 64;
        s->cfg.bin2 = 256;
        s->cfg.bin3 = 512;
    }
    int rc = tsg_create(&s->est, device);
    if (rc != TSG_OK) return rc;
    // the estimator's handle checks the tuning the workers will apply
    if ((s->cfg.tile_density > 0.0 &&
         tsg_set_tile_density(s->est, s->cfg.tile_density) != TSG_OK) ||
        (s->cfg.bin3 > 0 &&
         tsg_set_row_bins(s->est, s->cfg.bin1, s->cfg.bin2, s->cfg.bin3) != TSG_OK) ||
        s->cfg.tile_density < 0.0) {
        tsg_destroy(s->est);
        return TSG_ERR_INVALID;
    }
    s->estimator = std::thread([p = s.get()] { p->estimator_main(); });
    for (int i = 0; i < s->cfg.workers; i++)
        s->workers.emplace_back([p = s.get()] { p->worker_main(); });
    *server = s.release();
    return TSG_OK;
}

int tsg_server_submit(tsg_server_t s, const tsg_product_t *req, uint64_t *ticket) {
    if (!s || !req || !ticket) return TSG_ERR_INVALID;
    if (req->m < 0 || req->n < 0 || req->k < 0 || req->nnz_a < 0 || req->nnz_b < 0)
        return TSG_ERR_INVALID;
    if (!req->ia || !req->ib || (req->nnz_a > 0 && (!req->ja || !req->a)) ||
        (req->nnz_b > 0 && (!req->jb || !req->b)))
        return TSG_ERR_INVALID;
    auto r = std::make_unique<Req>();
    r->prod = *req;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return TSG_ERR_STATE;
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->pending.push_back(r.get());
    s->reqs.emplace(r->ticket, std::move(r));
    s->cv_est.notify_one();
    return TSG_OK;
}

int tsg_server_wait(tsg_server_t s, uint64_t ticket, tsg_result_t *result) {
    if (!s || !result) return TSG_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return TSG_ERR_INVALID;
    Req *r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    *result = r->res;
    return r->res.status;
}

int tsg_server_release(tsg_server_t s, uint64_t ticket) {
    if (!s) return TSG_ERR_INVALID;
    std::shared_ptr<BatchOut> out;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end() || !it->second->done) return TSG_ERR_INVALID;
        Req *r = it->second.get();
        s->in_flight -= r->est;
        out = r->out;
        if (out && --out->refs == 0) {
            s->pool.put(out->ic);
            s->pool.put(out->jc);
            s->pool.put(out->c);
        }
        s->reqs.erase(it);
        s->cv_work.notify_all();
    }
    return TSG_OK;
}

int tsg_server_destroy(tsg_server_t s) {
    if (!s) return TSG_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_est.notify_all();
        s->cv_work.notify_all();
    }
    s->estimator.join();
    for (auto &t : s->workers) t.join();
    tsg_destroy(s->est);
    delete s;
    return TSG_OK;
}

}  // extern "C"

// tsg_spgemm: command-line SpGEMM on one device, the counterpart of the
// reference driver `./test -d <dev> A.mtx` (main_spgemm-col.cu:4824-4960,
// 5518-5560), written against the C-ABI of libtsg only.
//
//   tsg_spgemm -d <device> [--aat] [--method auto|tile|row] [--repeat N]
//              [--csv FILE] [--keep-values] A.mtx
//
// Loads A with the reference loader's semantics (tsg_read_mtx), replaces its
// values by i % 10 as the reference does (main_spgemm-col.cu:4895-4896;
// --keep-values keeps the file's), and computes C = A*A, or C = A*A^T with
// --aat (a compile-time switch in the reference).  A square symmetric A is
// refused for A*A^T and a non-square A for A*A, with the reference's
// messages.  Each of N runs (default 20, as REPEAT_NUM) times the symbolic
// and numeric phases on device-resident operands; the minimum is reported,
// GFLOP/s = 2 * nnzCub / time (main_spgemm-col.cu:4787), and one line
//   filename,rowA,colA,nnzA,nnzCub,nnzC,compression_rate,time_ms,gflops
// is appended to results_tile.csv (or FILE), as main_spgemm-col.cu:5554-5559.
#include "../../include/tsg.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

int usage() {
    std::fprintf(stderr,
                 "usage: tsg_spgemm -d <device> [--aat] [--method auto|tile|row] [--repeat N]\n"
                 "                  [--csv FILE] [--keep-values] A.mtx\n");
    return 2;
}

#define HIP_OK(x)                                                                            \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                          \
            return 1;                                                                        \
        }                                                                                    \
    } while (0)

#define TSG_OK_OR(x, h)                                                                      \
    do {                                                                                     \
        int r_ = (x);                                                                        \
        if (r_ != TSG_OK) {                                                                  \
            std::fprintf(stderr, "tsg error %d: %s\n", r_, tsg_last_error(h));               \
            return 1;                                                                        \
        }                                                                                    \
    } while (0)

}  // namespace

int main(int argc, char **argv) {
    int device = -1, repeat = 20, method = TSG_METHOD_AUTO;
    bool aat = false, keep_values = false;
    std::string csv = "results_tile.csv";
    const char *file = nullptr;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "-d" && i + 1 < argc) device = std::atoi(argv[++i]);
        else if (a == "--aat") aat = true;
        else if (a == "--keep-values") keep_values = true;
        else if (a == "--repeat" && i + 1 < argc) repeat = std::atoi(argv[++i]);
        else if (a == "--csv" && i + 1 < argc) csv = argv[++i];
        else if (a == "--method" && i + 1 < argc) {
            const std::string m = argv[++i];
            if (m == "auto") method = TSG_METHOD_AUTO;
            else if (m == "tile") method = TSG_METHOD_TILE;
            else if (m == "row") method = TSG_METHOD_ROW;
            else return usage();
        } else if (!a.empty() && a[0] != '-' && !file) file = argv[i];
        else return usage();
    }
    if (device < 0 || !file || repeat < 1) return usage();
    HIP_OK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, device));
    std::printf("Device [ %d ] %s, %d CUs\n", device, prop.name, prop.multiProcessorCount);
    std::printf("MAT: -------------- %s --------------\n", file);

    int m = 0, n = 0, sym = 0;
    int64_t nnz = 0;
    int *ia = nullptr, *ja = nullptr;
    double *va = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = tsg_read_mtx(file, &m, &n, &nnz, &sym, &ia, &ja, &va);
    const double load_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != TSG_OK) {
        std::fprintf(stderr, "cannot read %s (error %d)\n", file, rc);
        return 1;
    }
    std::printf("input matrix A: ( %d, %d ) nnz = %lld\n loadfile time    = %4.5f sec\n", m, n,
                (long long)nnz, load_s);
    if (!aat && m != n) {
        std::printf("matrix squaring must have rowA == colA. Exit.\n");
        return 0;
    }
    if (aat && m == n && sym) {
        std::printf("matrix AAT does not do symmetric matrix. Exit.\n");
        return 0;
    }
    if (!keep_values)
        for (int64_t i = 0; i < nnz; i++) va[i] = (double)(i % 10);

    tsg_handle_t h = nullptr;
    if (tsg_create(&h, device) != TSG_OK) {
        std::fprintf(stderr, "tsg_create failed: %s\n", tsg_last_error(nullptr));
        return 1;
    }
    TSG_OK_OR(tsg_set_method(h, method), h);
    int *d_ia, *d_ja, *d_ib = nullptr, *d_jb = nullptr;
    double *d_a, *d_b = nullptr;
    HIP_OK(hipMalloc(&d_ia, sizeof(int) * ((size_t)m + 1)));
    HIP_OK(hipMalloc(&d_ja, sizeof(int) * ((size_t)nnz + 1)));
    HIP_OK(hipMalloc(&d_a, sizeof(double) * ((size_t)nnz + 1)));
    HIP_OK(hipMemcpy(d_ia, ia, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_ja, ja, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_a, va, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice));
    const int rows_b = aat ? n : m, cols_b = aat ? m : n;
    if (aat) {
        HIP_OK(hipMalloc(&d_ib, sizeof(int) * ((size_t)n + 1)));
        HIP_OK(hipMalloc(&d_jb, sizeof(int) * ((size_t)nnz + 1)));
        HIP_OK(hipMalloc(&d_b, sizeof(double) * ((size_t)nnz + 1)));
        TSG_OK_OR(tsg_csr_transpose(h, m, n, d_ia, d_ja, d_a, nnz, d_ib, d_jb, d_b), h);
    } else {
        d_ib = d_ia;
        d_jb = d_ja;
        d_b = d_a;
    }
    uint64_t nnz_cub = 0;
    TSG_OK_OR(tsg_intermediate_products(h, m, rows_b, d_ia, d_ja, nnz, d_ib, &nnz_cub), h);
    std::printf("SpGEMM nnzCub = %llu\n", (unsigned long long)nnz_cub);

    int64_t *d_ic = nullptr;
    HIP_OK(hipMalloc(&d_ic, sizeof(int64_t) * ((size_t)m + 1)));
    int64_t nnz_c = 0;
    TSG_OK_OR(tsg_spgemm_symbolic(h, m, n, cols_b, d_ia, d_ja, nnz, d_ib, d_jb, nnz, d_ic,
                                  &nnz_c), h);
    int *d_jc = nullptr;
    double *d_c = nullptr;
    HIP_OK(hipMalloc(&d_jc, sizeof(int) * ((size_t)nnz_c + 1)));
    HIP_OK(hipMalloc(&d_c, sizeof(double) * ((size_t)nnz_c + 1)));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    hipStream_t stream = nullptr;
    HIP_OK(hipStreamCreate(&stream));
    TSG_OK_OR(tsg_set_stream(h, stream), h);
    double best = 0.0;
    for (int r = 0; r < repeat; r++) {
        HIP_OK(hipEventRecord(e0, stream));
        int64_t nz = 0;
        TSG_OK_OR(tsg_spgemm_symbolic(h, m, n, cols_b, d_ia, d_ja, nnz, d_ib, d_jb, nnz, d_ic,
                                      &nz), h);
        TSG_OK_OR(tsg_spgemm_numeric(h, m, n, cols_b, d_ia, d_ja, d_a, nnz, d_ib, d_jb, d_b,
                                     nnz, d_ic, d_jc, d_c), h);
        HIP_OK(hipEventRecord(e1, stream));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("run %2d takes %f ms\n", r, ms);
        if (r == 0 || ms < best) best = ms;
    }
    const int used = tsg_get_method_used(h);
    const double compression = nnz_c > 0 ? (double)nnz_cub / (double)nnz_c : 0.0;
    const double gflops = best > 0 ? 2.0 * (double)nnz_cub / (best * 1e6) : 0.0;
    std::printf("method = %s, nnzC = %lld\n", used == TSG_METHOD_TILE ? "tile" : "row",
                (long long)nnz_c);
    std::printf("spgemm time = %4.3f ms, gflops = %4.2f\n", best, gflops);
    FILE *fout = std::fopen(csv.c_str(), "a");
    if (!fout) {
        std::fprintf(stderr, "Writing results fails.\n");
    } else {
        std::fprintf(fout, "%s,%i,%i,%lld,%llu,%lld,%f,%f,%f\n", file, m, n, (long long)nnz,
                     (unsigned long long)nnz_cub, (long long)nnz_c, compression, best, gflops);
        std::fclose(fout);
    }
    hipFree(d_c);
    hipFree(d_jc);
    hipFree(d_ic);
    if (aat) {
        hipFree(d_b);
        hipFree(d_jb);
        hipFree(d_ib);
    }
    hipFree(d_a);
    hipFree(d_ja);
    hipFree(d_ia);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    tsg_destroy(h);
    hipStreamDestroy(stream);
    tsg_free(ia);
    tsg_free(ja);
    tsg_free(va);
    return 0;
}

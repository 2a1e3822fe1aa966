// Matrix Market input (include/tsg.h, tsg_read_mtx): host code.
//
// Semantics follow the reference loader mmio_allinone (mmio_highlevel.h:593-759):
// the banner's field (real, integer, complex, pattern) and symmetry decide how
// entries are read; indices are 1-based in the file and 0-based in the CSR;
// a symmetric or hermitian file is expanded by mirroring every off-diagonal
// entry with the same value; entries of a row keep file order (a mirrored
// entry lands where its turn in the file puts it), with no sorting and no
// merging of duplicates; a pattern entry is 1.0, an integer entry its value,
// a complex entry its real part.  Unlike the reference, a malformed file (an
// index outside the declared size, a short entry list, a bad size line) is
// an error instead of undefined behaviour.
#include "tsg_internal.h"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

enum Field { F_REAL, F_INTEGER, F_COMPLEX, F_PATTERN };

struct Banner {
    Field field = F_REAL;
    bool symmetric = false;    // symmetric or hermitian: mirrored
};

std::string lower(std::string s) {
    for (auto &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

// "%%MatrixMarket matrix coordinate <field> <symmetry>"
int parse_banner(const char *line, Banner &b) {
    char tag[64] = {0}, obj[64] = {0}, fmt[64] = {0}, field[64] = {0}, sym[64] = {0};
    if (std::sscanf(line, "%63s %63s %63s %63s %63s", tag, obj, fmt, field, sym) != 5) return -2;
    if (std::strcmp(tag, "%%MatrixMarket") != 0) return -2;
    if (lower(obj) != "matrix" || lower(fmt) != "coordinate") return -2;
    const std::string f = lower(field), s = lower(sym);
    if (f == "real" || f == "double") b.field = F_REAL;
    else if (f == "integer") b.field = F_INTEGER;
    else if (f == "complex") b.field = F_COMPLEX;
    else if (f == "pattern") b.field = F_PATTERN;
    else return -2;
    if (s == "symmetric" || s == "hermitian") b.symmetric = true;
    else if (s == "general" || s == "skew-symmetric") b.symmetric = false;
    else return -2;
    return 0;
}

// Next line that is not a comment (%) or blank; false at end of file.
bool next_data_line(FILE *f, std::vector<char> &buf) {
    for (;;) {
        buf[0] = 0;
        if (!std::fgets(buf.data(), (int)buf.size(), f)) return false;
        const char *p = buf.data();
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '%' || *p == '\n' || *p == '\r' || *p == 0) continue;
        return true;
    }
}

}  // namespace

extern "C" {

int tsg_read_mtx(const char *path, int *m, int *n, int64_t *nnz, int *symmetric, int **ia,
                 int **ja, double **a) {
    if (!path || !m || !n || !nnz || !symmetric || !ia || !ja || !a) return TSG_ERR_INVALID;
    *ia = nullptr;
    *ja = nullptr;
    *a = nullptr;
    FILE *f = std::fopen(path, "r");
    if (!f) return TSG_MTX_ERR_OPEN;
    std::vector<char> buf(1 << 16);
    int rc = 0;
    Banner b;
    long long rows = 0, cols = 0, entries = 0;
    if (!std::fgets(buf.data(), (int)buf.size(), f) || parse_banner(buf.data(), b) != 0) {
        rc = TSG_MTX_ERR_BANNER;
    } else if (!next_data_line(f, buf) ||
               std::sscanf(buf.data(), "%lld %lld %lld", &rows, &cols, &entries) != 3 ||
               rows < 0 || cols < 0 || entries < 0 || rows > 0x7fffffffLL ||
               cols > 0x7fffffffLL) {
        rc = TSG_MTX_ERR_SIZE;
    }
    std::vector<int> ri, ci;
    std::vector<double> vi;
    if (rc == 0) {
        ri.resize((size_t)entries);
        ci.resize((size_t)entries);
        vi.resize((size_t)entries);
        for (long long e = 0; e < entries && rc == 0; e++) {
            if (!next_data_line(f, buf)) {
                rc = TSG_MTX_ERR_ENTRIES;
                break;
            }
            long long r = 0, c = 0;
            double v = 1.0, im = 0.0;
            long long iv = 0;
            int got = 0;
            switch (b.field) {
            case F_REAL:
                got = std::sscanf(buf.data(), "%lld %lld %lg", &r, &c, &v) == 3;
                break;
            case F_INTEGER:
                got = std::sscanf(buf.data(), "%lld %lld %lld", &r, &c, &iv) == 3;
                v = (double)iv;
                break;
            case F_COMPLEX:
                got = std::sscanf(buf.data(), "%lld %lld %lg %lg", &r, &c, &v, &im) == 4;
                break;
            case F_PATTERN:
                got = std::sscanf(buf.data(), "%lld %lld", &r, &c) == 2;
                v = 1.0;
                break;
            }
            if (!got || r < 1 || r > rows || c < 1 || c > cols) {
                rc = TSG_MTX_ERR_ENTRIES;
                break;
            }
            ri[(size_t)e] = (int)(r - 1);
            ci[(size_t)e] = (int)(c - 1);
            vi[(size_t)e] = v;
        }
    }
    std::fclose(f);
    if (rc != 0) return rc;
    if (b.symmetric && rows != cols) return TSG_MTX_ERR_SIZE;

    // counting sort by row, file order kept; mirrored entries interleaved
    std::vector<long long> cnt((size_t)rows + 1, 0);
    for (long long e = 0; e < entries; e++) {
        cnt[(size_t)ri[(size_t)e]]++;
        if (b.symmetric && ri[(size_t)e] != ci[(size_t)e]) cnt[(size_t)ci[(size_t)e]]++;
    }
    long long total = 0;
    for (long long i = 0; i <= rows; i++) {
        const long long c = cnt[(size_t)i];
        cnt[(size_t)i] = total;
        total += c;
    }
    if (total > 0x7fffffffLL) return TSG_ERR_OVERFLOW;
    int *pia = static_cast<int *>(std::malloc(sizeof(int) * ((size_t)rows + 1)));
    int *pja = static_cast<int *>(std::malloc(sizeof(int) * ((size_t)total + 1)));
    double *pa = static_cast<double *>(std::malloc(sizeof(double) * ((size_t)total + 1)));
    if (!pia || !pja || !pa) {
        std::free(pia);
        std::free(pja);
        std::free(pa);
        return TSG_ERR_ALLOC;
    }
    for (long long i = 0; i <= rows; i++) pia[i] = (int)cnt[(size_t)i];
    std::vector<long long> fill(cnt.begin(), cnt.end() - 1);
    for (long long e = 0; e < entries; e++) {
        const int r = ri[(size_t)e], c = ci[(size_t)e];
        const double v = vi[(size_t)e];
        long long p = fill[(size_t)r]++;
        pja[p] = c;
        pa[p] = v;
        if (b.symmetric && r != c) {
            p = fill[(size_t)c]++;
            pja[p] = r;
            pa[p] = v;
        }
    }
    *m = (int)rows;
    *n = (int)cols;
    *nnz = total;
    *symmetric = b.symmetric ? 1 : 0;
    *ia = pia;
    *ja = pja;
    *a = pa;
    return TSG_OK;
}

void tsg_free(void *p) { std::free(p); }

}  // extern "C"

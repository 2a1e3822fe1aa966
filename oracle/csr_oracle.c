/*
 * ORACLE -- test infrastructure only.  Nothing in the product links this
 * file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and only as the checker.
 *
 * CPU restatement of the reference's CSR SpGEMM
 * thmkl_dcsrmultcsr (thmkl_dcsrmultcsr.c:7-95):
 *   symbolic (request=1, :12-51): per row of A, mark every column reached
 *       through B in a dense length-k flag array; the row count is the number
 *       of marked columns (structural zeros included); ic is the exclusive scan.
 *   numeric (request=0, :52-91): per row, a dense length-k value array starts
 *       at +0.0 and accumulates value[j] += a_ik * b_kj in A-row order, then
 *       B-row order (:61-72); flagged columns are emitted ascending (:76-84).
 *
 * Deviation (complexity only, results identical): the dense arrays are
 * allocated once per thread and reset through the list of touched columns
 * instead of a full memset per row, and the emission pass sorts the touched
 * list instead of scanning all k columns.  Row pointers of C are int64 so
 * outputs past 2^31 entries are expressible.
 *
 * FP mode: fused != 0 accumulates with fma(a, b, acc) -- what gcc emits for
 * the reference's `+=` under its -O3 -march=native build; fused == 0 rounds
 * the product first.  This file is compiled with -ffp-contract=off so the
 * non-fused form is not contracted behind our back.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int cmp_int(const void *x, const void *y)
{
    int a = *(const int *)x, b = *(const int *)y;
    return (a > b) - (a < b);
}

/* Per-thread scratch: flag[k], val[k], touched[k]. */
typedef struct { unsigned char *flag; double *val; int *touched; } scratch_t;

static int scratch_init(scratch_t *s, int k, int with_val)
{
    s->flag = (unsigned char *)calloc((size_t)k > 0 ? (size_t)k : 1, 1);
    s->touched = (int *)malloc(sizeof(int) * ((size_t)k > 0 ? (size_t)k : 1));
    s->val = with_val ? (double *)calloc((size_t)k > 0 ? (size_t)k : 1, sizeof(double)) : NULL;
    return s->flag && s->touched && (!with_val || s->val);
}

static void scratch_free(scratch_t *s)
{
    free(s->flag); free(s->val); free(s->touched);
}

/* Mark the columns of row i of C; returns the number of distinct columns and
 * leaves them (unsorted) in s->touched. */
static int64_t mark_row(int64_t i, const int64_t *ia, const int *ja,
                        const int64_t *ib, const int *jb, scratch_t *s)
{
    int64_t cnt = 0;
    for (int64_t p = ia[i]; p < ia[i + 1]; p++) {
        int kk = ja[p];
        for (int64_t q = ib[kk]; q < ib[kk + 1]; q++) {
            int j = jb[q];
            if (!s->flag[j]) { s->flag[j] = 1; s->touched[cnt++] = j; }
        }
    }
    return cnt;
}

/* thmkl_dcsrmultcsr.c:12-51.  ic has m+1 entries; returns nnz(C) or -1 on
 * allocation failure. */
int64_t orc_csr_symbolic(int m, int k,
                         const int64_t *ia, const int *ja,
                         const int64_t *ib, const int *jb, int64_t *ic)
{
    int failed = 0;
    #pragma omp parallel
    {
        scratch_t s;
        int ok = scratch_init(&s, k, 0);
        if (!ok) {
            #pragma omp atomic write
            failed = 1;
        }
        #pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < m; i++) {
            if (!ok) continue;
            int64_t cnt = mark_row(i, ia, ja, ib, jb, &s);
            for (int64_t t = 0; t < cnt; t++) s.flag[s.touched[t]] = 0;
            ic[i] = cnt;
        }
        scratch_free(&s);
    }
    if (failed) return -1;
    /* exclusive scan, as thmkl_dcsrmultcsr.c:38-47 */
    int64_t run = 0;
    for (int i = 0; i < m; i++) { int64_t c = ic[i]; ic[i] = run; run += c; }
    ic[m] = run;
    return run;
}

/* thmkl_dcsrmultcsr.c:52-91.  ic must come from orc_csr_symbolic. */
int orc_csr_numeric(int m, int k,
                    const int64_t *ia, const int *ja, const double *a,
                    const int64_t *ib, const int *jb, const double *b,
                    const int64_t *ic, int *jc, double *c, int fused)
{
    int failed = 0;
    #pragma omp parallel
    {
        scratch_t s;
        int ok = scratch_init(&s, k, 1);
        if (!ok) {
            #pragma omp atomic write
            failed = 1;
        }
        #pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < m; i++) {
            if (!ok) continue;
            int64_t cnt = 0;
            for (int64_t p = ia[i]; p < ia[i + 1]; p++) {
                int kk = ja[p];
                double av = a[p];
                for (int64_t q = ib[kk]; q < ib[kk + 1]; q++) {
                    int j = jb[q];
                    if (!s.flag[j]) { s.flag[j] = 1; s.touched[cnt++] = j; }
                    if (fused) s.val[j] = fma(av, b[q], s.val[j]);
                    else       s.val[j] = s.val[j] + av * b[q];
                }
            }
            qsort(s.touched, (size_t)cnt, sizeof(int), cmp_int);
            int64_t base = ic[i];
            for (int64_t t = 0; t < cnt; t++) {
                int j = s.touched[t];
                jc[base + t] = j;
                c[base + t] = s.val[j];
                s.flag[j] = 0;
                s.val[j] = 0.0;
            }
        }
        scratch_free(&s);
    }
    return failed ? -1 : 0;
}

/* Sum over A's entries a_ik of |row k of B| (main_spgemm-col.cu:4950-4955). */
uint64_t orc_intermediate_products(int64_t nnza, const int *ja, const int64_t *ib)
{
    uint64_t s = 0;
    for (int64_t p = 0; p < nnza; p++) s += (uint64_t)(ib[ja[p] + 1] - ib[ja[p]]);
    return s;
}

/* For each C entry, the sum of |a_ik * b_kj| over its products: the scale
 * against which an accumulation-order difference is measured. */
int orc_csr_absdot(int m, int k,
                   const int64_t *ia, const int *ja, const double *a,
                   const int64_t *ib, const int *jb, const double *b,
                   const int64_t *ic, const int *jc, double *out)
{
    int failed = 0;
    #pragma omp parallel
    {
        double *acc = (double *)calloc((size_t)k > 0 ? (size_t)k : 1, sizeof(double));
        if (!acc) {
            #pragma omp atomic write
            failed = 1;
        }
        #pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < m; i++) {
            if (!acc) continue;
            for (int64_t p = ia[i]; p < ia[i + 1]; p++)
                for (int64_t q = ib[ja[p]]; q < ib[ja[p] + 1]; q++)
                    acc[jb[q]] += fabs(a[p] * b[q]);
            for (int64_t t = ic[i]; t < ic[i + 1]; t++) {
                out[t] = acc[jc[t]];
                acc[jc[t]] = 0.0;
            }
        }
        free(acc);
    }
    return failed ? -1 : 0;
}

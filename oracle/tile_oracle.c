/*
 * ORACLE -- test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this file, and only as the checker.
 *
 * CPU restatement of the reference's 16x16 bitmask-tile format and of its
 * tile-level SpGEMM steps:
 *   - CSR -> tile conversion: main_spgemm-col.cu:4968-5360 (blocknnz :44-68,
 *     blkmessage :70-120, tile-column sort :4987-4990, per-tile nnz
 *     :5055-5085, payload/masks :5148-5240).
 *   - tile-level symbolic (C tile structure): CPU step1/step2,
 *     main_spgemm1.cu:104 and :571.
 *   - per-tile masks / nnz: CPU step3 main_spgemm1.cu:917 (the GPU kernel
 *     main_spgemm-col.cu:997-1325 additionally keeps the row masks and the
 *     uchar in-tile row pointers, which are produced here as well).
 *   - FP64 numeric per tile: CPU step4 main_spgemm1.cu:2588 (dense 16x16
 *     accumulator, matched tile pairs in ascending K, A rows, A-row entries,
 *     B-row entries).
 *   - tile -> CSR: absent from the reference; defined here as the unique CSR
 *     with the same entries, columns ascending.
 *
 * Tile format (per matrix with m rows, n columns; blkm = ceil(m/16)):
 *   tile_ptr[blkm+1]   tiles per tile-row (exclusive scan)
 *   tile_col[numblk]   tile column, ascending within a tile-row
 *   tile_off[numblk+1] payload offset of each tile (exclusive scan of nnz)
 *   mask[16*numblk]    row masks, bit (15 - c) for in-tile column c
 *   rowptr[16*numblk]  uchar start of in-tile row r; rows past the matrix
 *                      edge hold the tile nnz (main_spgemm-col.cu:5225-5228)
 *   idx[nnz]           packed (r << 4) | c   (main_spgemm-col.cu:5214)
 *   val[nnz]           values, row-major inside a tile, CSR order in a row
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TS 16

static int cmp_int(const void *x, const void *y)
{
    int a = *(const int *)x, b = *(const int *)y;
    return (a > b) - (a < b);
}

/* blocknnz + blkmessage (main_spgemm-col.cu:44-120): tiles per tile-row.
 * Returns numblk. */
int orc_tile_count(int m, int n, const int64_t *ia, const int *ja, int *tile_ptr)
{
    int blkm = (m + TS - 1) / TS, blkn = (n + TS - 1) / TS;
    char *flag = (char *)calloc(blkn > 0 ? blkn : 1, 1);
    int *list = (int *)malloc(sizeof(int) * (blkn > 0 ? blkn : 1));
    int total = 0;
    for (int I = 0; I < blkm; I++) {
        int cnt = 0;
        int r1 = (I + 1) * TS < m ? (I + 1) * TS : m;
        for (int64_t p = ia[I * TS]; p < ia[r1]; p++) {
            int J = ja[p] / TS;
            if (!flag[J]) { flag[J] = 1; list[cnt++] = J; }
        }
        for (int t = 0; t < cnt; t++) flag[list[t]] = 0;
        tile_ptr[I] = total;
        total += cnt;
    }
    tile_ptr[blkm] = total;
    free(flag); free(list);
    return total;
}

/* Payload of every tile (main_spgemm-col.cu:4976-5240). */
int orc_tile_fill(int m, int n, const int64_t *ia, const int *ja, const double *a,
                  const int *tile_ptr, int *tile_col, int64_t *tile_off,
                  uint16_t *mask, uint8_t *rowptr, uint8_t *idx, double *val)
{
    int blkm = (m + TS - 1) / TS, blkn = (n + TS - 1) / TS;
    int *pos = (int *)malloc(sizeof(int) * (blkn > 0 ? blkn : 1));   /* J -> local tile */
    int64_t off = 0;
    for (int I = 0; I < blkm; I++) {
        int r0 = I * TS, r1 = (I + 1) * TS < m ? (I + 1) * TS : m;
        int t0 = tile_ptr[I], nt = tile_ptr[I + 1] - t0;
        /* distinct tile columns, ascending (quick_sort_key at :4987-4990) */
        int cnt = 0;
        for (int64_t p = ia[r0]; p < ia[r1]; p++) {
            int J = ja[p] / TS, seen = 0;
            for (int t = 0; t < cnt; t++) if (tile_col[t0 + t] == J) { seen = 1; break; }
            if (!seen) tile_col[t0 + cnt++] = J;
        }
        if (cnt != nt) { free(pos); return -1; }
        qsort(tile_col + t0, (size_t)nt, sizeof(int), cmp_int);
        for (int t = 0; t < nt; t++) pos[tile_col[t0 + t]] = t;
        /* per-tile, per-row counts */
        int *rc = (int *)calloc((size_t)nt * TS + 1, sizeof(int));
        for (int r = r0; r < r1; r++)
            for (int64_t p = ia[r]; p < ia[r + 1]; p++)
                rc[pos[ja[p] / TS] * TS + (r - r0)]++;
        for (int t = 0; t < nt; t++) {
            int tile = t0 + t, run = 0;
            tile_off[tile] = off;
            for (int r = 0; r < TS; r++) {
                rowptr[(int64_t)tile * TS + r] = (uint8_t)run;  /* padded rows get nnz */
                mask[(int64_t)tile * TS + r] = 0;
                run += rc[t * TS + r];
            }
            off += run;
        }
        /* scatter entries: row-major inside a tile, CSR order inside a row */
        int *fillp = (int *)calloc((size_t)nt * TS + 1, sizeof(int));
        for (int r = r0; r < r1; r++) {
            for (int64_t p = ia[r]; p < ia[r + 1]; p++) {
                int t = pos[ja[p] / TS], tile = t0 + t, rr = r - r0, c = ja[p] % TS;
                int64_t q = tile_off[tile] + rowptr[(int64_t)tile * TS + rr] + fillp[t * TS + rr]++;
                idx[q] = (uint8_t)((rr << 4) | c);
                val[q] = a[p];
                mask[(int64_t)tile * TS + rr] |= (uint16_t)(1u << (TS - 1 - c));
            }
        }
        free(rc); free(fillp);
    }
    tile_off[tile_ptr[blkm]] = off;
    free(pos);
    return 0;
}

/* Tile-CSC index of a tile-CSR matrix: for every tile column J the tile rows K
 * (ascending) and the tile-CSR position of tile (K, J).  The reference builds
 * B's tile-CSC by transposing B (main_spgemm-col.cu:5003-5040); for a
 * duplicate-free column-sorted B both give the same tiles in the same order. */
int orc_tile_csc(int blkm, int blkn, const int *tile_ptr, const int *tile_col,
                 int *col_ptr, int *col_row, int *col_tile)
{
    memset(col_ptr, 0, sizeof(int) * (blkn + 1));
    int nt = tile_ptr[blkm];
    for (int t = 0; t < nt; t++) col_ptr[tile_col[t] + 1]++;
    for (int J = 0; J < blkn; J++) col_ptr[J + 1] += col_ptr[J];
    int *cur = (int *)malloc(sizeof(int) * (blkn > 0 ? blkn : 1));
    memcpy(cur, col_ptr, sizeof(int) * blkn);
    for (int K = 0; K < blkm; K++)
        for (int t = tile_ptr[K]; t < tile_ptr[K + 1]; t++) {
            int q = cur[tile_col[t]]++;
            col_row[q] = K;
            col_tile[q] = t;
        }
    free(cur);
    return 0;
}

/* Does tile-row I of A share a K with tile-column J of B?  Sorted merge as in
 * main_spgemm1.cu:116-150. */
static int tiles_meet(const int *a, int la, const int *b, int lb)
{
    int i = 0, j = 0;
    while (i < la && j < lb) {
        if (a[i] == b[j]) return 1;
        if (a[i] < b[j]) i++; else j++;
    }
    return 0;
}

/* CPU step1 (main_spgemm1.cu:104-160): count C tiles per tile-row by testing
 * every (I, J) -- O(blkm * blkn) merges.  Returns numblkC. */
int orc_tile_step1_intersect(int blkmA, const int *tptrA, const int *tcolA,
                             int blknB, const int *cptrB, const int *crowB,
                             int *tptrC)
{
    int total = 0;
    for (int I = 0; I < blkmA; I++) {
        int cnt = 0;
        for (int J = 0; J < blknB; J++)
            cnt += tiles_meet(tcolA + tptrA[I], tptrA[I + 1] - tptrA[I],
                              crowB + cptrB[J], cptrB[J + 1] - cptrB[J]);
        tptrC[I] = total;
        total += cnt;
    }
    tptrC[blkmA] = total;
    return total;
}

/* The same tile set by the Gustavson route (union over K in tile-row I of A
 * of B's tile-row K), as the GPU computes it; fills the sorted column list
 * when tcolC != NULL (CPU step2, main_spgemm1.cu:571).  Returns numblkC. */
int orc_tile_step1(int blkmA, const int *tptrA, const int *tcolA,
                   int blknB, const int *tptrB, const int *tcolB,
                   int *tptrC, int *tcolC)
{
    char *flag = (char *)calloc(blknB > 0 ? blknB : 1, 1);
    int *list = (int *)malloc(sizeof(int) * (blknB > 0 ? blknB : 1));
    int total = 0;
    for (int I = 0; I < blkmA; I++) {
        int cnt = 0;
        for (int p = tptrA[I]; p < tptrA[I + 1]; p++) {
            int K = tcolA[p];
            for (int q = tptrB[K]; q < tptrB[K + 1]; q++) {
                int J = tcolB[q];
                if (!flag[J]) { flag[J] = 1; list[cnt++] = J; }
            }
        }
        qsort(list, (size_t)cnt, sizeof(int), cmp_int);
        for (int t = 0; t < cnt; t++) {
            flag[list[t]] = 0;
            if (tcolC) tcolC[total + t] = list[t];
        }
        tptrC[I] = total;
        total += cnt;
    }
    tptrC[blkmA] = total;
    free(flag); free(list);
    return total;
}

/* Matched pairs of C tile (I, J): positions pa (in A's tile-CSR) and pb (in
 * B's tile-CSR) with equal K, ascending K. */
static int matched_pairs(int I, int J, const int *tptrA, const int *tcolA,
                         const int *cptrB, const int *crowB, const int *ctileB,
                         int *pa, int *pb)
{
    int i = tptrA[I], ie = tptrA[I + 1], j = cptrB[J], je = cptrB[J + 1], n = 0;
    while (i < ie && j < je) {
        if (tcolA[i] == crowB[j]) { pa[n] = i; pb[n] = ctileB[j]; n++; i++; j++; }
        else if (tcolA[i] < crowB[j]) i++;
        else j++;
    }
    return n;
}

/* Row r of tile t: [lo, hi) into its payload. */
static void tile_row(const uint8_t *rowptr, const int64_t *off, int t, int r,
                     int64_t *lo, int64_t *hi)
{
    int64_t base = off[t];
    *lo = base + rowptr[(int64_t)t * TS + r];
    *hi = r == TS - 1 ? off[t + 1] : base + rowptr[(int64_t)t * TS + r + 1];
}

/* Per-C-tile row masks, uchar row pointers and payload offsets
 * (CPU step3 main_spgemm1.cu:917; GPU step3 main_spgemm-col.cu:1167-1262).
 * maskC row r = OR over matched K and A_IK row r entries c of B_KJ row mask c.
 * Returns nnz(C). */
int64_t orc_tile_step2(int blkmA, const int *tptrA, const int *tcolA,
                       const int64_t *offA, const uint8_t *rowptrA, const uint8_t *idxA,
                       int blknB, const int *cptrB, const int *crowB, const int *ctileB,
                       const uint16_t *maskB,
                       const int *tptrC, const int *tcolC,
                       uint16_t *maskC, uint8_t *rowptrC, int64_t *offC)
{
    (void)rowptrA; (void)blknB;
    int maxp = 1;
    for (int I = 0; I < blkmA; I++)
        if (tptrA[I + 1] - tptrA[I] > maxp) maxp = tptrA[I + 1] - tptrA[I];
    int *pa = (int *)malloc(sizeof(int) * maxp), *pb = (int *)malloc(sizeof(int) * maxp);
    int64_t off = 0;
    for (int I = 0; I < blkmA; I++) {
        for (int t = tptrC[I]; t < tptrC[I + 1]; t++) {
            uint32_t mk[TS] = {0};
            int np = matched_pairs(I, tcolC[t], tptrA, tcolA, cptrB, crowB, ctileB, pa, pb);
            for (int s = 0; s < np; s++)
                for (int64_t q = offA[pa[s]]; q < offA[pa[s] + 1]; q++)
                    mk[idxA[q] >> 4] |= maskB[(int64_t)pb[s] * TS + (idxA[q] & 15)];
            int run = 0;
            offC[t] = off;
            for (int r = 0; r < TS; r++) {
                maskC[(int64_t)t * TS + r] = (uint16_t)mk[r];
                rowptrC[(int64_t)t * TS + r] = (uint8_t)run;
                run += __builtin_popcount(mk[r]);
            }
            off += run;
        }
    }
    offC[tptrC[blkmA]] = off;
    free(pa); free(pb);
    return off;
}

/* FP64 numeric per C tile (CPU step4 main_spgemm1.cu:2588-2690): a dense
 * 16x16 accumulator visited in ascending K, A-tile rows, A-row entries, B-row
 * entries; emitted row-major as in-tile column + value. */
int orc_tile_step3(int blkmA, const int *tptrA, const int *tcolA,
                   const int64_t *offA, const uint8_t *rowptrA, const uint8_t *idxA,
                   const double *valA,
                   const int *cptrB, const int *crowB, const int *ctileB,
                   const int64_t *offB, const uint8_t *rowptrB, const uint8_t *idxB,
                   const double *valB,
                   const int *tptrC, const int *tcolC, const int64_t *offC,
                   const uint16_t *maskC, uint8_t *colC, double *valC, int fused)
{
    int maxp = 1;
    for (int I = 0; I < blkmA; I++)
        if (tptrA[I + 1] - tptrA[I] > maxp) maxp = tptrA[I + 1] - tptrA[I];
    int *pa = (int *)malloc(sizeof(int) * maxp), *pb = (int *)malloc(sizeof(int) * maxp);
    for (int I = 0; I < blkmA; I++) {
        for (int t = tptrC[I]; t < tptrC[I + 1]; t++) {
            double acc[TS * TS];
            memset(acc, 0, sizeof(acc));
            int np = matched_pairs(I, tcolC[t], tptrA, tcolA, cptrB, crowB, ctileB, pa, pb);
            for (int s = 0; s < np; s++) {
                for (int r = 0; r < TS; r++) {
                    int64_t alo, ahi;
                    tile_row(rowptrA, offA, pa[s], r, &alo, &ahi);
                    for (int64_t q = alo; q < ahi; q++) {
                        int c = idxA[q] & 15;
                        int64_t blo, bhi;
                        tile_row(rowptrB, offB, pb[s], c, &blo, &bhi);
                        for (int64_t u = blo; u < bhi; u++) {
                            int cc = idxB[u] & 15;
                            double *d = &acc[r * TS + cc];
                            if (fused) *d = fma(valA[q], valB[u], *d);
                            else       *d = *d + valA[q] * valB[u];
                        }
                    }
                }
            }
            int64_t q = offC[t];
            for (int r = 0; r < TS; r++) {
                uint32_t mk = maskC[(int64_t)t * TS + r];
                for (int c = 0; c < TS; c++)
                    if (mk & (1u << (TS - 1 - c))) { colC[q] = (uint8_t)c; valC[q] = acc[r * TS + c]; q++; }
            }
        }
    }
    free(pa); free(pb);
    return 0;
}

/* C tiles -> CSR with int64 row pointers, columns ascending.  ic[m] = nnz. */
int orc_tile_to_csr(int m, const int *tptrC, const int *tcolC, const int64_t *offC,
                    const uint16_t *maskC, const uint8_t *rowptrC,
                    const uint8_t *colC, const double *valC,
                    int64_t *ic, int *jc, double *c)
{
    int blkm = (m + TS - 1) / TS;
    ic[0] = 0;
    for (int i = 0; i < m; i++) {
        int I = i / TS, r = i % TS;
        int64_t cnt = 0;
        for (int t = tptrC[I]; t < tptrC[I + 1]; t++)
            cnt += __builtin_popcount(maskC[(int64_t)t * TS + r]);
        ic[i + 1] = ic[i] + cnt;
    }
    if (jc == NULL) return 0;
    for (int I = 0; I < blkm; I++) {
        for (int r = 0; r < TS && I * TS + r < m; r++) {
            int64_t w = ic[I * TS + r];
            for (int t = tptrC[I]; t < tptrC[I + 1]; t++) {
                int64_t lo, hi;
                tile_row(rowptrC, offC, t, r, &lo, &hi);
                for (int64_t q = lo; q < hi; q++) {
                    jc[w] = tcolC[t] * TS + colC[q];
                    c[w] = valC[q];
                    w++;
                }
            }
        }
    }
    return 0;
}

"""Deterministic synthetic CSR generators for the SpGEMM workloads.

Every generator returns a canonical CSR triple ``(indptr, indices, data)``
with int32 ``indptr``/``indices`` (int64 ``indptr`` when nnz >= 2**31),
float64 ``data``, rows sorted by column and free of duplicates -- the form
the reference's drivers feed to ``thmkl_dcsrmultcsr``
(thmkl_dcsrmultcsr.c:7-10) and to the tile conversion
(main_spgemm-col.cu:4968-5360).

Workloads:
  * ``laplacian2d(k)``   5-point Laplacian on a k x k grid, 4 on the
                          diagonal, -1 off it (n = k*k).
  * ``stencil27(m, seed)`` 27-point stencil on an m^3 grid, values N(0,1).
  * ``rmat(scale, draws, seed)`` R-MAT with (a,b,c,d)=(.57,.19,.19,.05),
                          one uniform draw per level, duplicates merged,
                          values U[-1,1].
  * ``erdos_renyi(n, deg, seed)`` ER matrix with mean row degree ``deg``.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "laplacian2d", "stencil27", "rmat", "erdos_renyi", "er_batch", "served_mix", "from_coo",
    "transpose", "intermediate_products", "ref_values",
]


def from_coo(m: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray,
             dedupe: bool = True):
    """Sort COO triplets by (row, col) into CSR.  With ``dedupe`` the first
    value of a repeated (row, col) is kept (generators only)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    key = rows * (int(cols.max()) + 1 if cols.size else 1) + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    rows, cols, vals = rows[order], cols[order], vals[order]
    if dedupe and key.size:
        keep = np.ones(key.size, dtype=bool)
        keep[1:] = key[1:] != key[:-1]
        rows, cols, vals = rows[keep], cols[keep], vals[keep]
    counts = np.bincount(rows, minlength=m)
    indptr = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    ptr_t = np.int64 if indptr[-1] >= 2**31 else np.int32
    return indptr.astype(ptr_t), cols.astype(np.int32), vals


def laplacian2d(k: int):
    """5-point Laplacian, natural ordering of a k x k grid."""
    n = k * k
    idx = np.arange(n, dtype=np.int64)
    x, y = idx % k, idx // k
    rows, cols, vals = [idx], [idx], [np.full(n, 4.0)]
    for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        ok = (x + dx >= 0) & (x + dx < k) & (y + dy >= 0) & (y + dy < k)
        rows.append(idx[ok])
        cols.append(idx[ok] + dx + dy * k)
        vals.append(np.full(int(ok.sum()), -1.0))
    return from_coo(n, np.concatenate(rows), np.concatenate(cols),
                    np.concatenate(vals))


def stencil27(m: int, seed: int = 0):
    """27-point stencil on an m^3 grid, values drawn N(0,1) in CSR order."""
    n = m ** 3
    idx = np.arange(n, dtype=np.int64)
    x, y, z = idx % m, (idx // m) % m, idx // (m * m)
    rows, cols = [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                ok = ((x + dx >= 0) & (x + dx < m) & (y + dy >= 0) &
                      (y + dy < m) & (z + dz >= 0) & (z + dz < m))
                rows.append(idx[ok])
                cols.append(idx[ok] + dx + dy * m + dz * m * m)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    indptr, indices, _ = from_coo(n, rows, cols, np.zeros(rows.size))
    data = np.random.default_rng(seed).standard_normal(indices.size)
    return indptr, indices, data


def rmat(scale: int, draws: int, seed: int = 1,
         a: float = 0.57, b: float = 0.19, c: float = 0.19):
    """R-MAT: each of ``draws`` edges descends ``scale`` levels; at each level
    one uniform u picks the quadrant (u<a: top-left, a<=u<a+b: top-right,
    a+b<=u<a+b+c: bottom-left, else bottom-right).  Duplicates are merged and
    the surviving entries get U[-1,1] values in CSR order."""
    rng = np.random.default_rng(seed)
    n = 1 << scale
    rows = np.zeros(draws, dtype=np.int64)
    cols = np.zeros(draws, dtype=np.int64)
    for _ in range(scale):
        u = rng.random(draws)
        rbit = u >= a + b
        cbit = ((u >= a) & (u < a + b)) | (u >= a + b + c)
        rows = (rows << 1) | rbit
        cols = (cols << 1) | cbit
        del u, rbit, cbit
    key = np.unique(rows * n + cols)
    rows, cols = key // n, key % n
    del key
    counts = np.bincount(rows, minlength=n)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    data = rng.uniform(-1.0, 1.0, cols.size)
    return indptr.astype(np.int32), cols.astype(np.int32), data


def erdos_renyi(n: int, deg: float, seed: int = 0, ncols: int | None = None):
    """Erdos-Renyi: each row draws Poisson(deg) distinct columns uniformly."""
    rng = np.random.default_rng(seed)
    ncols = n if ncols is None else ncols
    per_row = np.minimum(rng.poisson(deg, n), ncols)
    rows = np.repeat(np.arange(n, dtype=np.int64), per_row)
    cols = rng.integers(0, ncols, rows.size)
    indptr, indices, _ = from_coo(n, rows, cols, np.zeros(rows.size))
    data = rng.uniform(-1.0, 1.0, indices.size)
    return indptr, indices, data


def er_batch(count: int, seed: int = 0, nmin: int = 1000, nmax: int = 65536,
             dmin: float = 4.0, dmax: float = 32.0):
    """``count`` independent ER matrices: n log-uniform in [nmin, nmax], mean
    row degree U[dmin, dmax], values U[-1,1]; one seeded stream for the whole
    batch.  Returns a list of (n, (indptr, indices, data))."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(np.int64)
    degs = rng.uniform(dmin, dmax, count)
    out = []
    for n, d in zip(ns, degs):
        n = int(n)
        per_row = np.minimum(rng.poisson(d, n), n)
        rows = np.repeat(np.arange(n, dtype=np.int64), per_row)
        cols = rng.integers(0, n, rows.size)
        key = np.unique(rows * n + cols)
        counts = np.bincount(key // n, minlength=n)
        indptr = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(counts, out=indptr[1:])
        out.append((n, (indptr.astype(np.int32), (key % n).astype(np.int32),
                        rng.uniform(-1.0, 1.0, key.size))))
    return out


def served_mix(count: int, seed: int = 11, edge_factor: int = 4):
    """The served request mix: 70% ER (n uniform in [1000, 16000], mean
    degree 8) C=A*A, 25% 2D 5-point Laplacian (grid side log-uniform in
    [64, 512]) C=A*A, 5% R-MAT scale 18 (edge_factor * 2^18 draws) C=A*A^T.
    Operands come from a pool of distinct matrices (64 ER, up to 16
    Laplacians, 2 R-MAT) drawn per request.  Returns (kind per request: 0 ER,
    1 Laplacian, 2 R-MAT; index into that kind's pool; ER pool; Laplacian pool
    of (side, CSR); R-MAT pool)."""
    rng = np.random.default_rng(seed)
    er = [erdos_renyi(int(n), 8.0, seed=1000 + i)
          for i, n in enumerate(rng.integers(1000, 16001, 64))]
    sides = sorted(set(int(x) for x in np.exp(rng.uniform(np.log(64), np.log(512), 16))))
    lap = [(s, laplacian2d(s)) for s in sides]
    rm = [rmat(18, edge_factor << 18, seed=s) for s in (1, 2)]
    u = rng.random(count)
    kinds = np.where(u < 0.70, 0, np.where(u < 0.95, 1, 2))
    pick = np.array([rng.integers(0, (len(er), len(lap), len(rm))[k]) for k in kinds])
    return kinds, pick, er, lap, rm


def transpose(m: int, n: int, indptr, indices, data):
    """CSR transpose by counting sort: column j of the input becomes row j,
    its entries ordered by ascending source row (utils.h:158-195)."""
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices)
    rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(indptr))
    order = np.argsort(indices, kind="stable")
    counts = np.bincount(indices, minlength=n)
    tptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=tptr[1:])
    ptr_t = np.int64 if tptr[-1] >= 2**31 else np.int32
    return tptr.astype(ptr_t), rows[order].astype(np.int32), \
        np.asarray(data)[order]


def intermediate_products(a_indptr, a_indices, b_indptr) -> int:
    """nnzCub = sum over A's entries a_ik of |row k of B|
    (main_spgemm-col.cu:4950-4955); flops = 2 * nnzCub."""
    blen = np.diff(np.asarray(b_indptr, dtype=np.int64))
    return int(blen[np.asarray(a_indices)].sum())


def ref_values(nnz: int) -> np.ndarray:
    """The reference drivers overwrite A's values with i % 10
    (main_spgemm-col.cu:4895-4896), making FP64 products exact."""
    return (np.arange(nnz) % 10).astype(np.float64)

"""The benchmark product at full size: R-MAT n=2^20, 16M draws, seed 1,
C = A*A^T (9.1e9 outputs, 109 GB).  The oracle cannot form all of C on the
CPU in test time, so the whole product is checked through properties that do
not depend on size, and a sample of rows bit for bit:

* nnz(C) equals the structural count of A*A^T (9,099,789,340, the count of
  scipy's boolean product of the same matrix);
* every column index lies in [0, n) and every row's columns ascend strictly
  (every descent of the flat column array sits at a row boundary);
* for every 997th row and the 16 rows with the most products (up to 9.7M
  products each, all on the big-row path), row pointers, columns and values
  equal the oracle's C = A[rows] * A^T bit for bit.

Runs after the other GPU tests (the file name sorts after theirs): it holds
~170 GB of device memory while it runs and releases it at the end."""
from __future__ import annotations

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import gen
from tests.util import assert_csr_equal

pytestmark = pytest.mark.gpu

NNZ_C = 9_099_789_340


def _sub_rows(A, rows):
    ia, ja, a = A
    lens = (ia[rows + 1] - ia[rows]).astype(np.int64)
    ptr = np.zeros(rows.size + 1, dtype=np.int64)
    np.cumsum(lens, out=ptr[1:])
    idx = np.concatenate([np.arange(ia[r], ia[r + 1]) for r in rows])
    return ptr, ja[idx], a[idx]


def test_rmat20_aat_full_size():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    n = 1 << 20
    A = gen.rmat(20, 16_000_000, seed=1)
    print(f"[fullsize] R-MAT 2^20: nnz(A)={A[1].size}", flush=True)
    h = tsg.Handle(0)
    try:
        _check_full_product(h, n, A, torch)
    finally:
        h.close()
        torch.cuda.empty_cache()


def _check_full_product(h, n, A, torch):
    dA = tsg.DeviceCsr.from_host(n, n, *A)
    dB = h.transpose(dA)
    ic, jc, c = h.spgemm(dA, dB)
    torch.cuda.synchronize()
    assert h.method_used == tsg.METHOD_ROW
    ic_h = ic.cpu().numpy()
    assert int(ic_h[-1]) == NNZ_C
    assert np.all(np.diff(ic_h) >= 0)

    # column range and within-row strict ascent, in chunks of the flat array
    nnz = int(ic_h[-1])
    chunk = 1 << 30
    descents = 0
    for p0 in range(0, nnz, chunk):
        p1 = min(nnz, p0 + chunk + 1)
        x = jc[p0:p1]
        assert int(x.min()) >= 0 and int(x.max()) < n
        descents += int((x[1:] <= x[:-1]).sum())
        del x
    # descents at row boundaries (positions where a later row starts, empty
    # rows collapsed): last column of one row >= first column of the next
    at = torch.unique(ic[1:-1])
    at = at[(at > 0) & (at < nnz)]
    boundary = int((jc[at - 1] >= jc[at]).sum())
    assert descents == boundary, (descents, boundary)
    print(f"[fullsize] nnz(C)={nnz}, columns in range and ascending", flush=True)

    # sampled rows, bit for bit
    B = (dB.indptr.cpu().numpy(), dB.indices.cpu().numpy(), dB.data.cpu().numpy())
    prods = prods_per_row(A, B[0])
    heavy = np.argsort(prods)[-16:]
    rows = np.unique(np.concatenate([np.arange(0, n, 997), heavy])).astype(np.int64)
    sub = _sub_rows(A, rows)
    want = O.csr_spgemm(rows.size, n, sub, B)
    starts, stops = ic_h[rows], ic_h[rows + 1]
    pos = np.concatenate([np.arange(s, e) for s, e in zip(starts, stops)])
    pos_d = torch.from_numpy(pos).to(jc.device)
    got_jc = jc[pos_d].cpu().numpy()
    got_c = c[pos_d].cpu().numpy()
    got_ic = np.zeros(rows.size + 1, dtype=np.int64)
    np.cumsum(stops - starts, out=got_ic[1:])
    print(f"[fullsize] {rows.size} sampled rows ({int(prods[rows].sum())} products), "
          f"{pos.size} outputs", flush=True)
    assert_csr_equal((got_ic, got_jc, got_c), want)


def prods_per_row(A, ib):
    """Intermediate products of every row of A against B (row pointers ib)."""
    blen = np.diff(ib.astype(np.int64))
    ia = A[0].astype(np.int64)
    per_entry = blen[A[1]]
    cs = np.zeros(per_entry.size + 1, dtype=np.int64)
    np.cumsum(per_entry, out=cs[1:])
    return cs[ia[1:]] - cs[ia[:-1]]

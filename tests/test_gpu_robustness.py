"""GPU robustness: operand aliasing, index overflow of the tile path, and the
host entry's cached plan against inputs rewritten in place.  Results are
checked against the CPU oracle bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import gen
from tests.util import assert_csr_equal

pytestmark = pytest.mark.gpu

METHODS = [("tile", tsg.METHOD_TILE), ("row", tsg.METHOD_ROW)]


@pytest.fixture(scope="module")
def handle():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    h = tsg.Handle(0)
    yield h
    h.close()


def _shared_index_cases():
    S = gen.stencil27(10, seed=4)                        # tile path by AUTO too
    E = gen.erdos_renyi(3000, 9.0, seed=12)
    return [("st27_10", 1000, S), ("er3000", 3000, E)]


@pytest.mark.parametrize("case", _shared_index_cases(), ids=lambda c: c[0])
@pytest.mark.parametrize("mname,method", METHODS)
def test_shared_indices_distinct_values(handle, case, mname, method):
    """B passes A's own row-pointer and column arrays with values of its own
    (b = 2a + 1): C = A*B, not A*A, on both paths through the device API,
    also after a symbolic phase that saw the same pointers."""
    import torch
    name, n, A = case
    bv = 2.0 * A[2] + 1.0
    want = O.csr_spgemm(n, n, A, (A[0], A[1], bv))
    dA = tsg.DeviceCsr.from_host(n, n, *A)
    dB = tsg.DeviceCsr(n, n, dA.indptr, dA.indices,
                       torch.from_numpy(bv).to(dA.indptr.device))
    handle.set_method(method)
    handle.set_fp_mode(tsg.FP_FUSED)
    ic, jc, c = handle.spgemm(dA, dB)
    print(f"[{name} {mname}] nnzC={int(want[0][-1])} used={handle.method_used}", flush=True)
    assert_csr_equal((ic.cpu().numpy(), jc.cpu().numpy(), c.cpu().numpy()), want)
    # the same plan, then B = A again: A's values for both operands
    ic2, nnz = handle.symbolic(dA, dA)
    jc2 = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda:0")
    c2 = torch.empty(max(nnz, 1), dtype=torch.float64, device="cuda:0")
    handle.numeric(dA, dB, ic2, jc2, c2)
    assert_csr_equal((ic2.cpu().numpy(), jc2[:nnz].cpu().numpy(), c2[:nnz].cpu().numpy()), want)
    handle.numeric(dA, dA, ic2, jc2, c2)
    assert_csr_equal((ic2.cpu().numpy(), jc2[:nnz].cpu().numpy(), c2[:nnz].cpu().numpy()),
                     O.csr_spgemm(n, n, A, A))
    handle.set_method(tsg.METHOD_AUTO)


@pytest.mark.parametrize("case", _shared_index_cases(), ids=lambda c: c[0])
def test_host_entry_shared_indices_distinct_values(case):
    """The reference entry with ib == ia, jb == ja (the same host arrays) and
    b != a: B's values are uploaded on their own."""
    name, n, A = case
    ia = np.ascontiguousarray(A[0], dtype=np.int32)
    ja = np.ascontiguousarray(A[1], dtype=np.int32)
    a = np.ascontiguousarray(A[2], dtype=np.float64)
    b = 2.0 * a + 1.0
    want = O.csr_spgemm(n, n, (ia, ja, a), (ia, ja, b))
    for entry in (tsg.thmkl_dcsrmultcsr, tsg.dcsrmultcsr_i64):
        ic = entry(1, n, n, n, (ia, ja, a), (ia, ja, b))
        jc, c = entry(0, n, n, n, (ia, ja, a), (ia, ja, b), ic=ic)
        assert_csr_equal((ic, jc, c), want)
        # B = A (same arrays, same values) right after
        jc, c = entry(0, n, n, n, (ia, ja, a), (ia, ja, a), ic=ic)
        assert_csr_equal((ic, jc, c), O.csr_spgemm(n, n, (ia, ja, a), (ia, ja, a)))


def test_host_entry_indices_rewritten_in_place():
    """request=1, then the caller rewrites ja in place (same sizes, same
    nnz(C), different columns), then request=0: the cached plan must not be
    used for the new contents.  A has one entry per row, so C's row i is row
    ja[i] of B, and B's rows all hold 5 entries: ic is the same for any ja."""
    rng = np.random.default_rng(5)
    n, k = 500, 800
    ia = np.arange(n + 1, dtype=np.int32)
    ja = rng.permutation(n).astype(np.int32)
    a = rng.uniform(-1, 1, n)
    rows = np.repeat(np.arange(n), 5)
    cols = np.concatenate([np.sort(rng.choice(k, 5, replace=False)) for _ in range(n)])
    B = gen.from_coo(n, rows, cols, rng.uniform(-1, 1, rows.size))
    ic = tsg.thmkl_dcsrmultcsr(1, n, n, k, (ia, ja, a), B)
    ja[:] = rng.permutation(n).astype(np.int32)          # same array, new contents
    want = O.csr_spgemm(n, k, (ia, ja, a), B)
    np.testing.assert_array_equal(ic, want[0])
    jc, c = tsg.thmkl_dcsrmultcsr(0, n, n, k, (ia, ja, a), B, ic=ic)
    assert_csr_equal((ic, jc, c), want)


def _outer_product_operands():
    """A (2^20 x 1): one entry in every 16th row, so each of its 2^16 tile
    rows holds one tile; B (1 x 2^20): one entry in every 16th column, 2^16
    tile columns.  C = A*B has 2^32 tiles (and 2^32 entries)."""
    m = k = 1 << 20
    rows = np.arange(0, m, 16)
    ia = np.zeros(m + 1, dtype=np.int32)
    ia[rows + 1] = 1
    ia = np.cumsum(ia).astype(np.int32)
    A = (ia, np.zeros(rows.size, dtype=np.int32), np.ones(rows.size))
    B = (np.array([0, rows.size], dtype=np.int32), rows.astype(np.int32), np.ones(rows.size))
    return m, 1, k, A, B


def test_tile_path_refuses_c_tile_overflow(handle):
    """The tile path indexes C tiles with int32: 2^32 C tiles are refused
    with TSG_ERR_OVERFLOW before anything is sized from the count; AUTO
    pushed onto the tile path (density threshold below A's) falls back to the
    row path, whose int64 row pointers hold nnz(C) = 2^32."""
    m, n, k, A, B = _outer_product_operands()
    dA = tsg.DeviceCsr.from_host(m, n, *A)
    dB = tsg.DeviceCsr.from_host(n, k, *B)
    handle.set_method(tsg.METHOD_TILE)
    with pytest.raises(tsg.TsgError) as e:
        handle.symbolic(dA, dB)
    assert e.value.code == -4, e.value
    print(f"[overflow] tile path: {e.value}", flush=True)
    handle.set_method(tsg.METHOD_AUTO)
    handle.set_tuning({"tile_density": 0.5})
    try:
        ic, nnz = handle.symbolic(dA, dB)
    finally:
        handle.set_tuning({"tile_density": 32.0})
    assert handle.method_used == tsg.METHOD_ROW
    assert nnz == 1 << 32
    got = ic.cpu().numpy()
    want = np.zeros(m + 1, dtype=np.int64)
    want[1:] = np.cumsum(np.where(np.arange(m) % 16 == 0, 1 << 16, 0))
    np.testing.assert_array_equal(got, want)

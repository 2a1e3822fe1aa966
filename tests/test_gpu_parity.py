"""GPU parity: the HIP paths (through the C ABI) against the reference's
fixtures and the CPU oracle.  Structure bit-exact; values bitwise in both FP
modes (every output is accumulated in A-row order, as the reference's CPU
SpGEMM does).
"""
from __future__ import annotations

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import gen
from tests.util import assert_csr_equal, golden_names, load_golden

pytestmark = pytest.mark.gpu

METHODS = [("tile", tsg.METHOD_TILE), ("row", tsg.METHOD_ROW)]


@pytest.fixture(scope="module")
def handle():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    h = tsg.Handle(0)
    yield h
    h.close()


def to_dev(m, n, X):
    return tsg.DeviceCsr.from_host(m, n, X[0], X[1], X[2])


def run(handle, m, n, k, A, B, method, fp=tsg.FP_FUSED, same=False):
    dA = to_dev(m, n, A)
    dB = dA if same else to_dev(n, k, B)
    handle.set_method(method)
    handle.set_fp_mode(fp)
    ic, jc, c = handle.spgemm(dA, dB)
    return ic.cpu().numpy(), jc.cpu().numpy(), c.cpu().numpy()


def is_square_aa(g):
    return (g["m"] == g["n"] == g["k"] and np.array_equal(g["A"][0], g["B"][0])
            and np.array_equal(g["A"][1], g["B"][1]) and np.array_equal(g["A"][2], g["B"][2]))


@pytest.mark.parametrize("name", golden_names())
@pytest.mark.parametrize("mname,method", METHODS)
def test_fixture_parity(handle, name, mname, method):
    g = load_golden(name)
    same = is_square_aa(g)
    got = run(handle, g["m"], g["n"], g["k"], g["A"], g["B"], method, tsg.FP_FUSED, same)
    print(f"[{name} {mname}] nnzC={got[0][-1]} used={handle.method_used}", flush=True)
    assert handle.method_used == method
    assert_csr_equal(got, (g["ic"], g["jc"], g["c_fma"]))
    got = run(handle, g["m"], g["n"], g["k"], g["A"], g["B"], method, tsg.FP_SEPARATE, same)
    assert_csr_equal(got, (g["ic"], g["jc"], g["c_nofma"]))


def _cases():
    L64 = gen.laplacian2d(64)
    S12 = gen.stencil27(12, seed=3)
    E = gen.erdos_renyi(4000, 12.0, seed=21)
    # rows of ~576 products spanning <= 20000 columns: those with more than
    # 512 outputs run in the rank-only kernel, the others in the narrow one
    E24 = gen.erdos_renyi(20000, 24.0, seed=22)
    # rows of ~400 products and outputs spanning up to 50000 columns (more
    # than 1024 bitmap words): the rank-only kernel's 512-output variant
    E20 = gen.erdos_renyi(50000, 20.0, seed=23)
    R12 = gen.rmat(12, 40_000, seed=5)
    R12T = gen.transpose(4096, 4096, *R12)
    R14 = gen.rmat(14, 250_000, seed=6)
    R14T = gen.transpose(16384, 16384, *R14)
    return [
        ("lap64_AA", 4096, 4096, 4096, L64, L64, True),
        ("st27_12_AA", 1728, 1728, 1728, S12, S12, True),
        ("er4000_AA", 4000, 4000, 4000, E, E, True),
        ("er20000_d24_AA", 20000, 20000, 20000, E24, E24, True),
        ("er50000_d20_AA", 50000, 50000, 50000, E20, E20, True),
        ("rmat12_AAT", 4096, 4096, 4096, R12, R12T, False),
        ("rmat14_AAT", 16384, 16384, 16384, R14, R14T, False),
    ]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
@pytest.mark.parametrize("mname,method", METHODS)
def test_generated_parity(handle, case, mname, method):
    name, m, n, k, A, B, same = case
    want = O.csr_spgemm(m, k, A, B, fused=True)
    print(f"[{name} {mname}] nnzC={want[0][-1]}", flush=True)
    got = run(handle, m, n, k, A, B, method, tsg.FP_FUSED, same)
    assert_csr_equal(got, want)


def _wide_case():
    """Rows with > 1024 products and > 1024 outputs spread over k > 65536
    columns: big-row symbolic, item splitting by count and by width."""
    rng = np.random.default_rng(17)
    m, n, k = 40, 3000, 300_000
    A = gen.erdos_renyi(m, 400.0, seed=1, ncols=n)
    rows = np.repeat(np.arange(n), 30)
    cols = rng.integers(0, k, rows.size)
    cols[::7] = rng.integers(0, 2000, cols[::7].size)      # a dense column band
    B = gen.from_coo(n, rows, cols, rng.uniform(-1, 1, rows.size))
    return m, n, k, A, B


def test_row_path_big_rows_and_items(handle):
    m, n, k, A, B = _wide_case()
    want = O.csr_spgemm(m, k, A, B, fused=True)
    print(f"[wide] nnzC={want[0][-1]} max row={np.diff(want[0]).max()}", flush=True)
    got = run(handle, m, n, k, A, B, tsg.METHOD_ROW)
    assert_csr_equal(got, want)


def test_big_row_items_in_column_order(handle):
    """Items run in first-column order (forced for any item count) give the
    oracle's bits: wide rows with many items, and R-MAT A*A^T with
    duplicate-free B (single-entry batches) and ordered accumulation."""
    assert handle.item_order() == 1 << 18
    handle.set_item_order(0)
    try:
        m, n, k, A, B = _wide_case()
        got = run(handle, m, n, k, A, B, tsg.METHOD_ROW)
        assert_csr_equal(got, O.csr_spgemm(m, k, A, B, fused=True))
        R = gen.rmat(12, 60_000, seed=3)
        RT = gen.transpose(4096, 4096, *R)
        got = run(handle, 4096, 4096, 4096, R, RT, tsg.METHOD_ROW)
        assert_csr_equal(got, O.csr_spgemm(4096, 4096, R, RT))
    finally:
        handle.set_item_order(1 << 18)
    with pytest.raises(tsg.TsgError):
        handle.set_item_order(-1)


def _size_class_case():
    """Big rows whose items (8192-column windows, <= 1024 outputs) fall in
    all three size classes of the rank kernels: rows of A with 10, 60 or 200
    entries against B rows of 128 entries spread over 131072 columns give
    ~80, ~470 and (cut by count) ~900 outputs per item."""
    rng = np.random.default_rng(23)
    n, k = 4096, 1 << 17
    degs = np.repeat([10, 60, 200], 300)
    rng.shuffle(degs)
    m = degs.size
    arows = np.repeat(np.arange(m), degs)
    acols = np.concatenate([rng.choice(n, d, replace=False) for d in degs])
    A = gen.from_coo(m, arows, acols, rng.uniform(-1, 1, arows.size))
    brows = np.repeat(np.arange(n), 128)
    bcols = np.concatenate([rng.choice(k, 128, replace=False) for _ in range(n)])
    B = gen.from_coo(n, brows, bcols, rng.uniform(-1, 1, brows.size))
    return m, n, k, A, B


@pytest.mark.parametrize("fp", [tsg.FP_FUSED, tsg.FP_SEPARATE])
def test_big_row_item_size_classes(handle, fp):
    """Items in column order split into the three size classes, each run by
    its own rank kernel (both FP modes): the oracle's bits."""
    fused = fp == tsg.FP_FUSED
    m, n, k, A, B = _size_class_case()
    handle.set_item_order(0)
    try:
        got = run(handle, m, n, k, A, B, tsg.METHOD_ROW, fp=fp)
    finally:
        handle.set_item_order(1 << 18)
        handle.set_fp_mode(tsg.FP_FUSED)
    want = O.csr_spgemm(m, k, A, B, fused=fused)
    print(f"[size classes fp{fp}] nnzC={want[0][-1]}", flush=True)
    assert_csr_equal(got, want)


def _tight_bound_case():
    """Big rows whose stored bitmaps fill their reserved space exactly: row 0
    is dense over 201,600 columns (three full 65536-column bands, every
    bitmap word non-empty, and a partial fourth band; k ends 16 columns past
    it), row 1 has 1,100 products on 1,100 columns 181 apart (one per word:
    as many non-empty words as products), row 2 both at once, then small
    rows."""
    rng = np.random.default_rng(31)
    k = 3 * 65536 + 5000 + 8
    n = 1300
    seg = 1008
    brows, bcols = [], []
    for i in range(200):                       # B rows 0..199: dense segments
        brows.append(np.full(seg, i))
        bcols.append(np.arange(i * seg, (i + 1) * seg))
    sparse_cols = (np.arange(1100) * 181) % k
    for j in range(1100):                      # B rows 200..1299: one entry each
        brows.append(np.array([200 + j]))
        bcols.append(np.array([sparse_cols[j]]))
    brows = np.concatenate(brows)
    bcols = np.concatenate(bcols)
    B = gen.from_coo(n, brows, bcols, rng.uniform(-1, 1, brows.size))
    arows = [np.zeros(200, int), np.ones(1100, int), np.full(1300, 2)]
    acols = [np.arange(200), np.arange(200, 1300), np.arange(1300)]
    for r in range(3, 51):
        d = int(rng.integers(1, 6))
        arows.append(np.full(d, r))
        acols.append(rng.choice(n, d, replace=False))
    arows = np.concatenate(arows)
    acols = np.concatenate(acols)
    m = 51
    A = gen.from_coo(m, arows, acols, rng.uniform(-1, 1, arows.size))
    return m, n, k, A, B


@pytest.mark.parametrize("order", [0, 1 << 18])
def test_big_rows_bitmaps_fill_their_bound(handle, order):
    """Stored big-row bitmaps (non-empty words per band, masks per
    sub-block) that fill the space reserved for them exactly, in rows' order
    (16384-column items) and column order (8192-column items in three size
    classes): the oracle's bits."""
    m, n, k, A, B = _tight_bound_case()
    handle.set_item_order(order)
    try:
        got = run(handle, m, n, k, A, B, tsg.METHOD_ROW)
    finally:
        handle.set_item_order(1 << 18)
    want = O.csr_spgemm(m, k, A, B, fused=True)
    print(f"[tight bound order={order}] nnzC={want[0][-1]}", flush=True)
    assert want[0][1] == 201600 and want[0][2] - want[0][1] == 1100
    assert_csr_equal(got, want)


@pytest.mark.parametrize("fp", [tsg.FP_FUSED, tsg.FP_SEPARATE])
def test_big_rows_hash_fallback(handle, fp):
    """No memory for span bitmaps: the big-row numeric phase keeps each
    item's columns in an LDS hash table and emits them in column order;
    results stay the oracle's bits (wide rows, R-MAT A*A^T, every row big,
    B rows unsorted with repeated columns)."""
    fused = fp == tsg.FP_FUSED
    handle.set_bitmap_budget(0)
    try:
        m, n, k, A, B = _wide_case()
        got = run(handle, m, n, k, A, B, tsg.METHOD_ROW, fp=fp)
        assert_csr_equal(got, O.csr_spgemm(m, k, A, B, fused=fused))
        R = gen.rmat(12, 60_000, seed=3)
        RT = gen.transpose(4096, 4096, *R)
        handle.set_tuning({"row_bins": (0, 0, 0)})
        got = run(handle, 4096, 4096, 4096, R, RT, tsg.METHOD_ROW, fp=fp)
        assert_csr_equal(got, O.csr_spgemm(4096, 4096, R, RT, fused=fused))
        U = _scramble_rows(RT, seed=1, dup_every=3)
        got = run(handle, 4096, 4096, 4096, R, U, tsg.METHOD_ROW, fp=fp)
        assert_csr_equal(got, O.csr_spgemm(4096, 4096, R, U, fused=fused))
    finally:
        handle.set_bitmap_budget(-1)
        handle.set_tuning({"row_bins": (64, 256, 1024)})


@pytest.mark.parametrize("cap", [0, 4096 * 40])
def test_coarse_block_pointers(handle, cap):
    """A block-pointer table capped below one entry per 1024 columns: blocks
    widen, or with cap 0 no table is built and B's row pointers serve as one
    block; every band and item range comes from a binary search inside a
    block; same bits as the oracle, also with the hash fallback and with the
    work sorted by column."""
    handle.set_block_pointer_budget(cap)
    try:
        m, n, k, A, B = _wide_case()
        want = O.csr_spgemm(m, k, A, B, fused=True)
        assert_csr_equal(run(handle, m, n, k, A, B, tsg.METHOD_ROW), want)
        handle.set_item_order(0)
        assert_csr_equal(run(handle, m, n, k, A, B, tsg.METHOD_ROW), want)
        handle.set_bitmap_budget(0)
        assert_csr_equal(run(handle, m, n, k, A, B, tsg.METHOD_ROW), want)
        R = gen.rmat(12, 60_000, seed=3)
        RT = gen.transpose(4096, 4096, *R)
        got = run(handle, 4096, 4096, 4096, R, RT, tsg.METHOD_ROW)
        assert_csr_equal(got, O.csr_spgemm(4096, 4096, R, RT))
    finally:
        handle.set_block_pointer_budget(-1)
        handle.set_bitmap_budget(-1)
        handle.set_item_order(1 << 18)


@pytest.mark.parametrize("n,deg,bins", [(40000, 12, (64, 256, 1024)),
                                          (80000, 20, (64, 256, 1024)),
                                          (40000, 12, (16, 256, 256))])
def test_small_rows_wide_spans(handle, n, deg, bins):
    """Small rows (<= 1024 products) whose columns spread over more than the
    LDS bitmap covers (16K columns for the 32-lane kernel, 64K for the
    64-lane one): hash accumulation, then a radix sort of > 64 outputs."""
    A = gen.erdos_renyi(n, deg, seed=5)
    handle.set_tuning({"row_bins": bins})
    try:
        got = run(handle, n, n, n, A, A, tsg.METHOD_ROW, same=True)
        assert_csr_equal(got, O.csr_spgemm(n, n, A, A))
        got = run(handle, n, n, n, A, A, tsg.METHOD_ROW, fp=tsg.FP_SEPARATE, same=True)
        assert_csr_equal(got, O.csr_spgemm(n, n, A, A, fused=False))
    finally:
        handle.set_tuning({"row_bins": (64, 256, 1024)})


def test_row_path_more_than_2p20_columns(handle):
    """k > 2^20: the big-row symbolic bitmap runs in several column passes."""
    rng = np.random.default_rng(4)
    m, n, k = 24, 500, 1_500_000
    A = gen.erdos_renyi(m, 300.0, seed=2, ncols=n)
    rows = np.repeat(np.arange(n), 20)
    B = gen.from_coo(n, rows, rng.integers(0, k, rows.size), rng.uniform(-1, 1, rows.size))
    want = O.csr_spgemm(m, k, A, B, fused=True)
    got = run(handle, m, n, k, A, B, tsg.METHOD_ROW)
    assert_csr_equal(got, want)


def test_tile_plan_arrays_match_tile_oracle(handle):
    """The tile path's intermediate arrays equal the restated reference tile
    format and tile steps (main_spgemm-col.cu:4968-5360, :997-1325)."""
    A = gen.stencil27(9, seed=2)
    R = gen.rmat(10, 6000, seed=8)
    RT = gen.transpose(1024, 1024, *R)
    for (m, n, k, XA, XB, same) in ((729, 729, 729, A, A, True),
                                    (1024, 1024, 1024, R, RT, False)):
        run(handle, m, n, k, XA, XB, tsg.METHOD_TILE, same=same)
        tA = O.tile_convert(m, n, XA)
        tB = O.tile_convert(n, k, XB)
        csc = O.tile_csc(tB)
        tptr, tcol = O.tile_step1(tA, tB)
        maskC, rowptrC, offC = O.tile_step2(tA, tB, csc, tptr, tcol)
        P = handle.plan_array
        np.testing.assert_array_equal(P("A_TILE_PTR"), tA["tile_ptr"])
        np.testing.assert_array_equal(P("A_TILE_COL"), tA["tile_col"])
        np.testing.assert_array_equal(P("A_TILE_OFF"), tA["tile_off"])
        np.testing.assert_array_equal(P("A_MASK"), tA["mask"])
        np.testing.assert_array_equal(P("A_ROWPTR"), tA["rowptr"])
        np.testing.assert_array_equal(P("A_IDX"), tA["idx"])
        np.testing.assert_array_equal(P("A_VAL").view(np.int64), tA["val"].view(np.int64))
        np.testing.assert_array_equal(P("B_TILE_PTR"), tB["tile_ptr"])
        np.testing.assert_array_equal(P("B_MASK"), tB["mask"])
        # B's values column-major inside each tile (columns, then rows)
        tiles = np.repeat(np.arange(tB["tile_col"].size), np.diff(tB["tile_off"]))
        order = np.lexsort((tB["idx"] >> 4, tB["idx"] & 15, tiles))
        np.testing.assert_array_equal(P("B_VAL").view(np.int64),
                                      tB["val"][order].view(np.int64))
        np.testing.assert_array_equal(P("B_CSC_PTR"), csc["col_ptr"])
        np.testing.assert_array_equal(P("B_CSC_ROW"), csc["col_row"])
        np.testing.assert_array_equal(P("B_CSC_TILE"), csc["col_tile"])
        np.testing.assert_array_equal(P("C_TILE_PTR"), tptr)
        np.testing.assert_array_equal(P("C_TILE_COL"), tcol)
        np.testing.assert_array_equal(P("C_MASK"), maskC)
        np.testing.assert_array_equal(P("C_ROWPTR"), rowptrC)
        np.testing.assert_array_equal(P("C_TILE_OFF"), offC)


@pytest.mark.parametrize("mname,method", METHODS)
def test_symbolic_reuse_with_new_values(handle, mname, method):
    """One symbolic plan serves numeric calls with different values."""
    import torch
    S = gen.stencil27(10, seed=1)
    n = 1000
    dA = to_dev(n, n, S)
    handle.set_method(method)
    handle.set_fp_mode(tsg.FP_FUSED)
    ic, nnz = handle.symbolic(dA, dA)
    for seed in (1, 2):
        vals = np.random.default_rng(seed).standard_normal(S[1].size)
        dA.data.copy_(torch.from_numpy(vals))
        jc = torch.empty(nnz, dtype=torch.int32, device="cuda:0")
        c = torch.empty(nnz, dtype=torch.float64, device="cuda:0")
        handle.numeric(dA, dA, ic, jc, c)
        want = O.csr_spgemm(n, n, (S[0], S[1], vals), (S[0], S[1], vals))
        assert_csr_equal((ic.cpu().numpy(), jc.cpu().numpy(), c.cpu().numpy()), want)


@pytest.mark.parametrize("mname,method", METHODS)
def test_empty_and_degenerate(handle, mname, method):
    # all-zero matrix
    Z = (np.zeros(33, np.int32), np.zeros(0, np.int32), np.zeros(0))
    got = run(handle, 32, 32, 32, Z, Z, method)
    assert got[0].tolist() == [0] * 33 and got[1].size == 0
    # 1 x n times n x 1
    g = load_golden("row_times_col")
    got = run(handle, 1, 40, 1, g["A"], g["B"], method)
    assert_csr_equal(got, (g["ic"], g["jc"], g["c_fma"]))


def test_unsorted_input_goes_to_row_path(handle):
    rng = np.random.default_rng(3)
    A = gen.erdos_renyi(300, 7.0, seed=9)
    ip, ix, d = (x.copy() for x in A)
    for i in range(300):                      # reverse every row
        ix[ip[i]:ip[i + 1]] = ix[ip[i]:ip[i + 1]][::-1].copy()
        d[ip[i]:ip[i + 1]] = d[ip[i]:ip[i + 1]][::-1].copy()
    U = (ip, ix, d)
    want = O.ref_thmkl(300, 300, 300, U, A, "fma") if O.ref_available() else \
        O.csr_spgemm(300, 300, U, A)
    got = run(handle, 300, 300, 300, U, A, tsg.METHOD_AUTO)
    assert handle.method_used == tsg.METHOD_ROW
    assert_csr_equal(got, want)
    with pytest.raises(tsg.TsgError):
        run(handle, 300, 300, 300, U, A, tsg.METHOD_TILE)
    _ = rng


@pytest.mark.parametrize("mname,method", [("auto", tsg.METHOD_AUTO), ("tile", tsg.METHOD_TILE)])
def test_invalid_input_is_rejected(handle, mname, method):
    """Malformed A is refused.  AUTO's tile-density sample, and with the tile
    path forced the tile counts of A and B, run beside the checks and must
    stay in bounds whatever the arrays hold."""
    A = gen.erdos_renyi(50, 4.0, seed=1)
    bad = (A[0], A[1].copy(), A[2])
    bad[1][3] = 50                            # column out of range
    with pytest.raises(tsg.TsgError):
        run(handle, 50, 50, 50, bad, A, method)
    badp = (A[0].copy(), A[1], A[2])
    badp[0][10] = badp[0][11] + 5             # non-monotone row pointers
    with pytest.raises(tsg.TsgError):
        run(handle, 50, 50, 50, badp, A, method)
    # far out of range
    for row, v in [(5, 10**9), (7, -5)]:
        badq = (A[0].copy(), A[1], A[2])
        badq[0][row] = v
        with pytest.raises(tsg.TsgError):
            run(handle, 50, 50, 50, badq, A, method)
    badc = (A[0], A[1].copy(), A[2])
    badc[1][0], badc[1][1] = -1, 2**31 - 1
    with pytest.raises(tsg.TsgError):
        run(handle, 50, 50, 50, badc, A, method)
    # the handle still works afterwards
    got = run(handle, 50, 50, 50, A, A, method)
    handle.set_method(tsg.METHOD_AUTO)
    assert_csr_equal(got, O.csr_spgemm(50, 50, A, A, fused=True))


def test_transpose_matches_host(handle):
    R = gen.rmat(12, 30_000, seed=2)
    dR = to_dev(4096, 4096, R)
    T = handle.transpose(dR)
    want = gen.transpose(4096, 4096, *R)
    np.testing.assert_array_equal(T.indptr.cpu().numpy(), want[0])
    np.testing.assert_array_equal(T.indices.cpu().numpy(), want[1])
    np.testing.assert_array_equal(T.data.cpu().numpy(), want[2])


def test_intermediate_products(handle):
    S = gen.stencil27(11, seed=0)
    dS = to_dev(1331, 1331, S)
    assert handle.intermediate_products(dS, dS) == (9 * 11 - 10) ** 3
    # heavy-tailed rows (one row spans many threads' entry slices) and empty rows
    R = gen.rmat(14, 200_000, seed=4)
    RT = gen.transpose(1 << 14, 1 << 14, *R)
    dR, dRT = to_dev(1 << 14, 1 << 14, R), to_dev(1 << 14, 1 << 14, RT)
    assert np.diff(R[0]).max() > 1000 and (np.diff(R[0]) == 0).any()
    assert handle.intermediate_products(dR, dRT) == gen.intermediate_products(R[0], R[1], RT[0])
    # A with a column outside B's rows is rejected before any product is counted
    bad = (R[0], R[1].copy(), R[2])
    bad[1][5] = 1 << 14
    with pytest.raises(tsg.TsgError):
        handle.intermediate_products(to_dev(1 << 14, 1 << 14, bad), dRT)


@pytest.mark.parametrize("where", ["slice_edge", "mid_long_row", "duplicate", "row_start"])
def test_sortedness_detection_on_long_rows(handle, where):
    """The tile path needs sorted, duplicate-free rows; the entry-parallel
    check must see a descent anywhere inside a long row (including across
    the 8-entry slices its threads walk) and must not flag the descent from
    one row's last column to the next row's first."""
    R = gen.rmat(12, 60_000, seed=6)
    ip, ix, d = (x.copy() for x in R)
    lens = np.diff(ip)
    r = int(np.argmax(lens))
    lo = int(ip[r])
    if where == "row_start":
        # rows untouched: every row start is a descent but the input is valid
        run(handle, 4096, 4096, 4096, (ip, ix, d), R, tsg.METHOD_TILE)
        assert handle.method_used == tsg.METHOD_TILE
        return
    if where == "slice_edge":
        q = lo + ((8 - lo % 8) % 8) + 8          # first entry of a thread slice inside row r
    elif where == "mid_long_row":
        q = lo + lens[r] // 2 + 3
    else:
        q = lo + lens[r] // 3
    if where == "duplicate":
        ix[q] = ix[q - 1]
        U = (ip, ix, d)
    else:
        ix[q - 1], ix[q] = ix[q], ix[q - 1]
        d[q - 1], d[q] = d[q], d[q - 1]
        U = (ip, ix, d)
    with pytest.raises(tsg.TsgError):
        run(handle, 4096, 4096, 4096, U, R, tsg.METHOD_TILE)
    got = run(handle, 4096, 4096, 4096, U, R, tsg.METHOD_AUTO)
    assert handle.method_used == tsg.METHOD_ROW
    want = O.ref_thmkl(4096, 4096, 4096, U, R, "fma") if O.ref_available() else \
        O.csr_spgemm(4096, 4096, U, R)
    assert_csr_equal(got, want)


@pytest.mark.parametrize("name", ["lap16_AA_refvals", "rmat9_AAT", "rect_203x157x91"])
def test_reference_host_entry(name):
    """thmkl_dcsrmultcsr: same signature and two-phase contract as the
    reference (thmkl_dcsrmultcsr.c:7-95), host pointers in and out."""
    g = load_golden(name)
    m, n, k = g["m"], g["n"], g["k"]
    A = g["A"]
    B = A if is_square_aa(g) else g["B"]
    ic = tsg.thmkl_dcsrmultcsr(1, m, n, k, A, B)
    np.testing.assert_array_equal(ic, g["ic"])
    jc, c = tsg.thmkl_dcsrmultcsr(0, m, n, k, A, B, ic=ic)
    assert_csr_equal((ic, jc, c), (g["ic"], g["jc"], g["c_fma"]))
    ic64 = tsg.dcsrmultcsr_i64(1, m, n, k, A, B)
    jc64, c64 = tsg.dcsrmultcsr_i64(0, m, n, k, A, B, ic=ic64)
    assert_csr_equal((ic64, jc64, c64), (g["ic"], g["jc"], g["c_fma"]))


@pytest.mark.parametrize("case", ["rmat_aat", "banded_tile", "rect_203x157x91"])
def test_host_entry_row_slabs(case):
    """Row-slab streaming of the host entries: with C capped below nnz(C),
    request=0 runs slabs of consecutive rows of A as products of their own
    and copies each slab's columns and values out at ic[first row]; C equals
    the whole-C call and the oracle bit for bit.  Caps below the longest row
    put that row in a slab of its own; a cap of one entry runs every row
    alone.  R-MAT rows reach the big-row path, the banded matrix the tile
    path (AUTO), the rectangular golden product k != n."""
    if case == "rmat_aat":
        m = n = k = 2048
        A = gen.rmat(11, 20_000, seed=4)
        B = gen.transpose(2048, 2048, *A)
    elif case == "banded_tile":
        m = n = k = 3000
        A = _banded(3000, 16, seed=2)
        B = A
    else:
        g = load_golden(case)
        m, n, k = g["m"], g["n"], g["k"]
        A, B = g["A"], g["B"]
    want = O.csr_spgemm(m, k, A, B)
    nnz = int(want[0][-1])
    longest = int(np.diff(want[0]).max())
    try:
        tsg.set_host_slab_entries(0)
        ic = tsg.dcsrmultcsr_i64(1, m, n, k, A, B)
        np.testing.assert_array_equal(ic, want[0])
        whole = tsg.dcsrmultcsr_i64(0, m, n, k, A, B, ic=ic)
        assert tsg.host_last_slabs() == 1
        assert_csr_equal((ic,) + whole, want)
        caps = [max(1, nnz // 5), max(1, longest // 2)]
        if m <= 256:
            caps.append(1)
        for cap in caps:
            tsg.set_host_slab_entries(cap)
            assert tsg.host_slab_entries() == cap
            jc, c = tsg.dcsrmultcsr_i64(0, m, n, k, A, B, ic=ic)
            slabs = tsg.host_last_slabs()
            assert slabs > 1, (cap, slabs)
            if cap == 1:
                assert slabs == m
            np.testing.assert_array_equal(jc, whole[0])
            np.testing.assert_array_equal(c.view(np.int64), whole[1].view(np.int64))
        # the int32 reference entry streams the same way
        tsg.set_host_slab_entries(max(1, nnz // 3))
        ic32 = tsg.thmkl_dcsrmultcsr(1, m, n, k, A, B)
        jc32, c32 = tsg.thmkl_dcsrmultcsr(0, m, n, k, A, B, ic=ic32)
        assert tsg.host_last_slabs() > 1
        assert_csr_equal((ic32, jc32, c32), want)
        # slabs place C by the caller's ic, which must ascend
        if nnz > 1 and m > 1:
            bad = ic.copy()
            r = int(np.argmax(np.diff(ic) > 0))
            bad[r + 1] = ic[-1]
            assert np.any(np.diff(bad) < 0)
            with pytest.raises(tsg.TsgError):
                tsg.dcsrmultcsr_i64(0, m, n, k, A, B, ic=bad)
        # back to C whole after slabs: the whole product's plan is rebuilt
        tsg.set_host_slab_entries(0)
        jc, c = tsg.dcsrmultcsr_i64(0, m, n, k, A, B, ic=ic)
        assert tsg.host_last_slabs() == 1
        assert_csr_equal((ic, jc, c), want)
    finally:
        tsg.set_host_slab_entries(0)


def test_auto_near_threshold_counts_every_tile_row(handle):
    """AUTO's density sample (every 2nd of R-MAT 2^15's 2048 tile rows here)
    decides only when clear of the threshold; near it every tile row is
    counted.  Thresholds a hair below and above A's exact entries per
    non-empty 16x16 tile (numpy) must pick the tile and the row path."""
    n = 1 << 15
    R = gen.rmat(15, 1 << 19, seed=7)
    ia, ja = np.asarray(R[0]), np.asarray(R[1])
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ia))
    ntiles = np.unique((rows >> 4) * (n >> 4) + (ja >> 4)).size
    d = ja.size / ntiles
    print(f"[near threshold] {ja.size} entries in {ntiles} tiles: {d:.6f} per tile", flush=True)
    want = O.csr_spgemm(n, n, R, R)
    try:
        for thr, method in ((d * (1 - 1e-9), tsg.METHOD_TILE), (d * (1 + 1e-9), tsg.METHOD_ROW)):
            handle.set_tuning({"tile_density": thr})
            got = run(handle, n, n, n, R, R, tsg.METHOD_AUTO, same=True)
            assert handle.method_used == method, (thr, d)
            assert_csr_equal(got, want)
    finally:
        handle.set_tuning({"tile_density": 32.0})


def test_host_entry_large_c_matches_device_api(handle):
    """C of 165M entries (columns 661 MB, values 1.3 GB) back through the host
    entry's pinned staging buffers, many chunks each, whole and in slabs:
    equal bit for bit to the device API's C for the same product (R-MAT
    n=2^16 A*A^T)."""
    m = 1 << 16
    A = gen.rmat(16, 16 << 16, seed=1)
    B = gen.transpose(m, m, *A)
    handle.set_method(tsg.METHOD_AUTO)
    handle.set_fp_mode(tsg.FP_FUSED)
    dA, dB = to_dev(m, m, A), to_dev(m, m, B)
    ic_d, jc_d, c_d = handle.spgemm(dA, dB)
    want = (ic_d.cpu().numpy(), jc_d.cpu().numpy(), c_d.cpu().numpy())
    del ic_d, jc_d, c_d, dA, dB
    nnz = int(want[0][-1])
    print(f"[host large] nnz(C)={nnz}", flush=True)
    try:
        tsg.set_host_slab_entries(0)
        ic = tsg.dcsrmultcsr_i64(1, m, m, m, A, B)
        np.testing.assert_array_equal(ic, want[0])
        for cap in (0, nnz // 3):
            tsg.set_host_slab_entries(cap)
            jc, c = tsg.dcsrmultcsr_i64(0, m, m, m, A, B, ic=ic)
            slabs = tsg.host_last_slabs()
            assert (slabs == 1) if cap == 0 else (slabs >= 3), (cap, slabs)
            np.testing.assert_array_equal(jc, want[1])
            np.testing.assert_array_equal(c.view(np.int64), want[2].view(np.int64))
            del jc, c
    finally:
        tsg.set_host_slab_entries(0)


def _banded(n, bw, seed=0):
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), 2 * bw + 1)
    cols = rows + np.tile(np.arange(-bw, bw + 1), n)
    ok = (cols >= 0) & (cols < n)
    return gen.from_coo(n, rows[ok], cols[ok], rng.standard_normal(int(ok.sum())))


def test_auto_method_choice(handle):
    """AUTO takes the tile path for tile-dense A (a band of half-width 16:
    ~100 entries per 16x16 tile) and the row path for a 5-point Laplacian
    (~18 per tile) and hypersparse A (R-MAT), with identical results either
    way."""
    handle.set_tuning({"tile_density": 32.0})
    D = _banded(3000, 16, seed=2)
    got = run(handle, 3000, 3000, 3000, D, D, tsg.METHOD_AUTO, same=True)
    assert handle.method_used == tsg.METHOD_TILE
    assert_csr_equal(got, O.csr_spgemm(3000, 3000, D, D))
    L = gen.laplacian2d(48)
    n = 48 * 48
    got = run(handle, n, n, n, L, L, tsg.METHOD_AUTO, same=True)
    assert handle.method_used == tsg.METHOD_ROW
    assert_csr_equal(got, O.csr_spgemm(n, n, L, L))
    R = gen.rmat(11, 20_000, seed=4)
    RT = gen.transpose(2048, 2048, *R)
    got = run(handle, 2048, 2048, 2048, R, RT, tsg.METHOD_AUTO)
    assert handle.method_used == tsg.METHOD_ROW
    assert_csr_equal(got, O.csr_spgemm(2048, 2048, R, RT))


def _scramble_rows(X, seed, dup_every=0):
    """Reverse every row's entries (and, with dup_every, repeat every k-th
    entry right after itself with a new value): rows no longer ascending."""
    rng = np.random.default_rng(seed)
    ip, ix, d = (np.asarray(x) for x in X)
    rows, cols, vals = [], [], []
    newp = [0]
    for i in range(ip.size - 1):
        c = ix[ip[i]:ip[i + 1]][::-1]
        v = d[ip[i]:ip[i + 1]][::-1]
        rc, rv = [], []
        for j, (cc, vv) in enumerate(zip(c, v)):
            rc.append(cc)
            rv.append(vv)
            if dup_every and j % dup_every == 0:
                rc.append(cc)
                rv.append(rng.uniform(-1, 1))
        cols += rc
        vals += rv
        newp.append(newp[-1] + len(rc))
    return (np.array(newp, dtype=np.int32), np.array(cols, dtype=np.int32),
            np.array(vals, dtype=np.float64))


@pytest.mark.parametrize("dup_every", [0, 3])
@pytest.mark.parametrize("bins", [(64, 256, 1024), (0, 0, 0), (16, 64, 256)])
def test_unsorted_and_duplicate_b_rows(handle, dup_every, bins):
    """B rows out of order (and with repeated columns): the row path sorts B
    stably, so every output accumulates in the reference's order; small rows
    (hash and rank paths) and big rows (every row big with bins 0,0,0)."""
    R = gen.rmat(12, 60_000, seed=8)
    RT = gen.transpose(4096, 4096, *R)
    U = _scramble_rows(RT, seed=dup_every, dup_every=dup_every)
    handle.set_tuning({"row_bins": bins})
    try:
        got = run(handle, 4096, 4096, 4096, R, U, tsg.METHOD_AUTO)
        assert handle.method_used == tsg.METHOD_ROW
        want = O.ref_thmkl(4096, 4096, 4096, R, U, "fma") if O.ref_available() else \
            O.csr_spgemm(4096, 4096, R, U)
        assert_csr_equal(got, want)
        # a second numeric phase with new values reuses the sorted copy's map
        dA, dU = to_dev(4096, 4096, R), to_dev(4096, 4096, U)
        ic, nnz = handle.symbolic(dA, dU)
        import torch
        jc = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda:0")
        c = torch.empty(max(nnz, 1), dtype=torch.float64, device="cuda:0")
        U2 = (U[0], U[1], U[2] * 2.0 + 1.0)
        dU.data.copy_(torch.from_numpy(U2[2]))
        handle.numeric(dA, dU, ic, jc, c)
        want2 = O.ref_thmkl(4096, 4096, 4096, R, U2, "fma") if O.ref_available() else \
            O.csr_spgemm(4096, 4096, R, U2)
        assert_csr_equal((ic.cpu().numpy(), jc[:nnz].cpu().numpy(), c[:nnz].cpu().numpy()),
                         want2)
    finally:
        handle.set_tuning({"row_bins": (64, 256, 1024)})

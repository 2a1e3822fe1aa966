"""Pins the CPU oracle (oracle/) before anything is checked against it.

* golden fixtures: outputs of the reference's own thmkl_dcsrmultcsr compiled
  from its sources (tests/golden/make_golden.py) -- bitwise, both FP builds;
* the live reference build (oracle/_ref) on fresh seeded inputs;
* closed forms (5-point Laplacian and 27-point stencil squares);
* brute-force dense products on small integer matrices;
* the tile-format restatement against independent numpy formulations of the
  reference's tile construction (main_spgemm-col.cu:4968-5360);
* sensitivity: plausible mistakes (reversed sum order, transposed operand,
  dropped structural zero) must fail these pins.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import py_oracle as O
from tilespgemm_amd import gen
from tests.util import assert_csr_equal, bits, golden_names, load_golden, max_rel_error


# ----------------------------------------------------------------- golden
@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture(name):
    g = load_golden(name)
    want_fma = (g["ic"], g["jc"], g["c_fma"])
    want_sep = (g["ic"], g["jc"], g["c_nofma"])
    assert_csr_equal(O.csr_spgemm(g["m"], g["k"], g["A"], g["B"], fused=True), want_fma)
    assert_csr_equal(O.csr_spgemm(g["m"], g["k"], g["A"], g["B"], fused=False), want_sep)


@pytest.mark.parametrize("name", golden_names())
def test_tile_oracle_chain_matches_reference_fixture(name):
    g = load_golden(name)
    got = O.tile_spgemm(g["m"], g["n"], g["k"], g["A"], g["B"], fused=True)
    assert_csr_equal(got, (g["ic"], g["jc"], g["c_fma"]))
    got = O.tile_spgemm(g["m"], g["n"], g["k"], g["A"], g["B"], fused=False)
    assert_csr_equal(got, (g["ic"], g["jc"], g["c_nofma"]))


@pytest.mark.skipif(not (O.ref_available("fma") and O.ref_available("nofma")),
                    reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_matches_live_reference(seed):
    rng = np.random.default_rng(seed)
    m, n, k = (int(x) for x in rng.integers(20, 400, 3))
    A = gen.erdos_renyi(m, float(rng.uniform(1, 9)), seed=seed, ncols=n)
    B = gen.erdos_renyi(n, float(rng.uniform(1, 9)), seed=seed + 100, ncols=k)
    for variant, fused in (("fma", True), ("nofma", False)):
        want = O.ref_thmkl(m, n, k, A, B, variant)
        assert_csr_equal(O.csr_spgemm(m, k, A, B, fused=fused), want)


def test_fp_modes_are_distinguished():
    """Random values make FMA and separate rounding differ somewhere, so the
    fixtures pin the accumulation mode, not just the structure."""
    g = load_golden("st27_6_AA")
    assert np.any(bits(g["c_fma"]) != bits(g["c_nofma"]))


# ------------------------------------------------------------- closed forms
def _lap_a2_nnz(k):
    x, y = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    tot = 0
    for dx in range(-2, 3):
        for dy in range(-2, 3):
            if abs(dx) + abs(dy) <= 2:
                tot += int(((x + dx >= 0) & (x + dx < k) & (y + dy >= 0) & (y + dy < k)).sum())
    return tot


@pytest.mark.parametrize("k", [4, 16, 37, 64])
def test_laplacian_square_closed_form(k):
    A = gen.laplacian2d(k)
    n = k * k
    ic, jc, c = O.csr_spgemm(n, n, A, A)
    assert int(ic[-1]) == _lap_a2_nnz(k)
    # values: (A^2)_ii = 16 + deg(i); neighbours -8; straight distance 2: 1;
    # diagonal distance 2: 2  (all exact in FP64)
    deg = np.diff(A[0]) - 1
    rows = np.repeat(np.arange(n), np.diff(ic))
    dx = (jc % k) - (rows % k)
    dy = (jc // k) - (rows // k)
    want = np.where((dx == 0) & (dy == 0), 16.0 + deg[rows], 0.0)
    want = np.where(np.abs(dx) + np.abs(dy) == 1, -8.0, want)
    want = np.where(((np.abs(dx) == 2) & (dy == 0)) | ((np.abs(dy) == 2) & (dx == 0)), 1.0, want)
    want = np.where((np.abs(dx) == 1) & (np.abs(dy) == 1), 2.0, want)
    np.testing.assert_array_equal(c, want)


def test_laplacian_256_nnz_matches_survey_count():
    A = gen.laplacian2d(256)
    ic = O.csr_symbolic(65536, 65536, A[:2], A[:2])
    assert int(ic[-1]) == 846_852 == _lap_a2_nnz(256)


@pytest.mark.parametrize("m", [3, 5, 8, 12])
def test_stencil27_square_closed_form(m):
    A = gen.stencil27(m, seed=1)
    n = m ** 3
    ic = O.csr_symbolic(n, n, A[:2], A[:2])
    assert int(ic[-1]) == (5 * m - 6) ** 3
    assert O.intermediate_products(A[1], A[0]) == (9 * m - 10) ** 3


# ------------------------------------------------------------- brute force
@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_dense_brute_force(seed):
    rng = np.random.default_rng(seed)
    m, n, k = (int(x) for x in rng.integers(1, 48, 3))
    Ad = (rng.random((m, n)) < rng.uniform(0.05, 0.5)) * rng.integers(-4, 5, (m, n))
    Bd = (rng.random((n, k)) < rng.uniform(0.05, 0.5)) * rng.integers(-4, 5, (n, k))
    # explicit structure: keep entries whose value drew 0 as stored zeros
    pa = rng.random((m, n)) < 0.3
    pb = rng.random((n, k)) < 0.3
    ra, ca = np.nonzero(pa | (Ad != 0))
    rb, cb = np.nonzero(pb | (Bd != 0))
    A = gen.from_coo(m, ra, ca, Ad[ra, ca].astype(float))
    B = gen.from_coo(n, rb, cb, Bd[rb, cb].astype(float))
    ic, jc, c = O.csr_spgemm(m, k, A, B)
    Pa = np.zeros((m, n), dtype=np.int64)
    Pa[ra, ca] = 1
    Pb = np.zeros((n, k), dtype=np.int64)
    Pb[rb, cb] = 1
    struct = (Pa @ Pb) > 0                      # structural union, zeros kept
    dense = Ad.astype(float) @ Bd.astype(float)  # exact: small integers
    got_struct = np.zeros((m, k), dtype=bool)
    rows = np.repeat(np.arange(m), np.diff(ic))
    got_struct[rows, jc] = True
    np.testing.assert_array_equal(got_struct, struct)
    np.testing.assert_array_equal(c, dense[rows, jc])
    for i in range(m):                          # ascending columns per row
        assert np.all(np.diff(jc[ic[i]:ic[i + 1]]) > 0)


@pytest.mark.parametrize("seed", range(4))
def test_absdot_matches_dense_brute_force(seed):
    """csr_absdot (the scale of the relative error bound: sum over an
    output's products of |a_ik * b_kj|) equals |A| @ |B| at C's structure,
    exactly, on small integer matrices with negative values and stored
    zeros; a sum of signed products or one operand's sign kept would not."""
    rng = np.random.default_rng(100 + seed)
    m, n, k = (int(x) for x in rng.integers(2, 40, 3))
    Ad = (rng.random((m, n)) < 0.35) * rng.integers(-5, 6, (m, n))
    Bd = (rng.random((n, k)) < 0.35) * rng.integers(-5, 6, (n, k))
    pa, pb = rng.random((m, n)) < 0.2, rng.random((n, k)) < 0.2
    ra, ca = np.nonzero(pa | (Ad != 0))
    rb, cb = np.nonzero(pb | (Bd != 0))
    A = gen.from_coo(m, ra, ca, Ad[ra, ca].astype(float))
    B = gen.from_coo(n, rb, cb, Bd[rb, cb].astype(float))
    ic, jc, c = O.csr_spgemm(m, k, A, B)
    got = O.csr_absdot(m, k, A, B, ic, jc)
    rows = np.repeat(np.arange(m), np.diff(ic))
    want = (np.abs(Ad) @ np.abs(Bd)).astype(float)[rows, jc]
    np.testing.assert_array_equal(got, want)
    assert np.all(got >= np.abs(c))
    signed = (Ad @ Bd).astype(float)[rows, jc]
    half = (Ad @ np.abs(Bd)).astype(float)[rows, jc]
    if np.any(signed != want):
        assert not np.array_equal(got, signed)
    if np.any(half != want):
        assert not np.array_equal(got, half)


def test_absdot_is_positive_where_products_cancel():
    """Outputs whose products cancel to exactly 0.0 keep a positive scale, so
    a relative error there is measured against the products' magnitude."""
    z = load_golden("cancel_32")
    got = O.csr_absdot(z["m"], z["k"], z["A"], z["B"], z["ic"], z["jc"])
    zero = z["c_fma"] == 0.0
    assert zero.any() and np.all(got[zero] > 0)


def test_max_rel_error_measure():
    """The reported error is max |got - want| / scale; exact results give 0,
    one ulp moved at a unit-scale output gives 2^-52, and an output of scale
    0 must match exactly (else the error is infinite)."""
    want = np.array([1.0, -3.0, 0.0, 0.5])
    scale = np.array([1.0, 4.0, 2.0, 0.0])
    assert max_rel_error(want.copy(), want, scale) == 0.0
    got = want.copy()
    got[0] = np.nextafter(1.0, 2.0)
    assert max_rel_error(got, want, scale) == 2.0 ** -52
    got = want.copy()
    got[3] = 0.5000001
    assert max_rel_error(got, want, scale) == np.inf


# ----------------------------------------------------- sensitivity (CPU)
def _py_row_sum(A, B, i, j, order=1):
    """Sum of row i, column j products in A-row order (order=1) or reversed."""
    ia, ja, a = A
    ib, jb, b = B
    terms = []
    for p in range(ia[i], ia[i + 1]):
        t = ja[p]
        for q in range(ib[t], ib[t + 1]):
            if jb[q] == j:
                terms.append((a[p], b[q]))
    acc = 0.0
    for x, y in terms[::order]:
        acc = float(np.fma(x, y, acc)) if hasattr(np, "fma") else acc + x * y
    return acc, len(terms)


def test_fixture_detects_reversed_sum_order():
    g = load_golden("st27_6_AA")
    ic, jc, c = g["ic"], g["jc"], g["c_nofma"]
    A, B = g["A"], g["B"]
    rows = np.repeat(np.arange(g["m"]), np.diff(ic))
    caught = 0
    for q in range(0, jc.size, 97):
        ia, ja, a = A
        ib, jb, b = B
        terms = [(a[p], b[r]) for p in range(ia[rows[q]], ia[rows[q] + 1])
                 for r in range(ib[ja[p]], ib[ja[p] + 1]) if jb[r] == jc[q]]
        fwd = 0.0
        for x, y in terms:
            fwd = fwd + x * y
        rev = 0.0
        for x, y in terms[::-1]:
            rev = rev + x * y
        assert fwd == c[q]                   # the restatement's order is the fixture's
        caught += rev != c[q]
    assert caught > 0                         # a reversed order would be detected


def test_fixture_detects_transposed_operand_and_dropped_zero():
    g = load_golden("rmat9_AAT")
    # B^T in place of B changes the product
    BT = gen.transpose(g["n"], g["k"], *g["B"])
    ic, jc, c = O.csr_spgemm(g["m"], g["k"], g["A"], BT)
    assert not (np.array_equal(ic, g["ic"]) and np.array_equal(jc, g["jc"]) and
                np.array_equal(bits(c), bits(g["c_fma"])))
    # dropping exact-zero sums (scipy's behaviour) loses a structural entry
    z = load_golden("cancel_32")
    assert np.any(z["c_fma"] == 0.0)
    SA = sp.csr_matrix((z["A"][2], z["A"][1], z["A"][0]), shape=(32, 32))
    S = (SA @ SA).tocsr()
    S.eliminate_zeros()
    assert S.nnz < int(z["ic"][-1])


# ------------------------------------------------------------- tile format
def _np_tiles(m, n, A):
    """Independent numpy construction of the reference's tile arrays
    (main_spgemm-col.cu:4976-5240)."""
    ia, ja, a = (np.asarray(x) for x in A)
    rows = np.repeat(np.arange(m), np.diff(ia.astype(np.int64)))
    I, J = rows >> 4, ja >> 4
    key = I.astype(np.int64) * ((n + 15) // 16 + 1) + J
    order = np.lexsort((np.arange(ja.size), rows, key))   # tile, then row, then CSR order
    tk, first = np.unique(key[order], return_index=True)
    blkm = (m + 15) // 16
    tI = tk // ((n + 15) // 16 + 1)
    tile_ptr = np.zeros(blkm + 1, np.int32)
    np.add.at(tile_ptr, tI + 1, 1)
    tile_ptr = np.cumsum(tile_ptr).astype(np.int32)
    tile_col = (tk % ((n + 15) // 16 + 1)).astype(np.int32)
    tile_of = np.searchsorted(tk, key)
    nt = tk.size
    cnt = np.bincount(tile_of, minlength=nt)
    tile_off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    r, c = rows & 15, ja & 15
    mask = np.zeros((nt, 16), np.uint16)
    np.bitwise_or.at(mask, (tile_of, r), (1 << (15 - c)).astype(np.uint16))
    rc = np.zeros((nt, 16), np.int64)
    np.add.at(rc, (tile_of, r), 1)
    rowptr = (np.cumsum(rc, axis=1) - rc).astype(np.uint8)
    idx = ((r << 4) | c).astype(np.uint8)[order]
    val = a[order]
    return dict(tile_ptr=tile_ptr, tile_col=tile_col, tile_off=tile_off,
                mask=mask.ravel(), rowptr=rowptr.ravel(), idx=idx, val=val)


@pytest.mark.parametrize("name", golden_names())
def test_tile_conversion_matches_numpy_construction(name):
    g = load_golden(name)
    for (mm, nn, X) in ((g["m"], g["n"], g["A"]), (g["n"], g["k"], g["B"])):
        t = O.tile_convert(mm, nn, X)
        w = _np_tiles(mm, nn, X)
        for key in ("tile_ptr", "tile_col", "tile_off", "mask", "rowptr", "idx"):
            np.testing.assert_array_equal(t[key], w[key], err_msg=key)
        np.testing.assert_array_equal(bits(t["val"]), bits(w["val"]))
        # invariants: popcount(mask row) = entries of that tile row
        pc = np.array([bin(int(v)).count("1") for v in t["mask"]], dtype=np.int64)
        assert pc.sum() == X[1].size


def _np_tile_csc_by_transpose(n, k, B):
    """B's tile-CSC the way the reference builds it: transpose B, tile the
    transpose, transpose every tile back (main_spgemm-col.cu:5003-5360).
    Returns, per tile column J in order, the list of (K, mask rows)."""
    BT = gen.transpose(n, k, *B)
    tT = _np_tiles(k, n, BT)
    out = []
    for J in range(len(tT["tile_ptr"]) - 1):
        for t in range(tT["tile_ptr"][J], tT["tile_ptr"][J + 1]):
            K = int(tT["tile_col"][t])
            mt = tT["mask"][16 * t:16 * t + 16]          # rows of B^T tile = cols of B tile
            m_rows = np.zeros(16, np.uint16)
            for cc in range(16):
                for rr in range(16):
                    if mt[cc] & (1 << (15 - rr)):
                        m_rows[rr] |= 1 << (15 - cc)
            out.append((J, K, m_rows))
    return out


@pytest.mark.parametrize("name", ["rmat9_AAT", "rect_203x157x91", "holes_70", "denseblock_48"])
def test_tile_csc_matches_reference_transpose_route(name):
    g = load_golden(name)
    tB = O.tile_convert(g["n"], g["k"], g["B"])
    csc = O.tile_csc(tB)
    ref = _np_tile_csc_by_transpose(g["n"], g["k"], g["B"])
    got = []
    for J in range(tB["blkn"]):
        for q in range(csc["col_ptr"][J], csc["col_ptr"][J + 1]):
            t = csc["col_tile"][q]
            got.append((J, int(csc["col_row"][q]), tB["mask"][16 * t:16 * t + 16]))
    assert len(got) == len(ref)
    for (J1, K1, m1), (J2, K2, m2) in zip(got, ref):
        assert (J1, K1) == (J2, K2)
        np.testing.assert_array_equal(m1, m2)


@pytest.mark.parametrize("name", golden_names())
def test_tile_step1_formulations_agree(name):
    """Gustavson union (the GPU's formulation) == the reference CPU step1's
    all-pairs intersection test (main_spgemm1.cu:104-160)."""
    g = load_golden(name)
    tA = O.tile_convert(g["m"], g["n"], g["A"])
    tB = O.tile_convert(g["n"], g["k"], g["B"])
    tptr, tcol = O.tile_step1(tA, tB)
    tptr2 = O.tile_step1_intersect(tA, O.tile_csc(tB), tB["blkn"])
    np.testing.assert_array_equal(tptr, tptr2)
    # and equals the boolean product of the tile patterns
    PA = sp.csr_matrix((np.ones(tA["numblk"]), tA["tile_col"], tA["tile_ptr"]),
                       shape=(tA["blkm"], tA["blkn"]))
    PB = sp.csr_matrix((np.ones(tB["numblk"]), tB["tile_col"], tB["tile_ptr"]),
                       shape=(tB["blkm"], tB["blkn"]))
    P = (PA @ PB).tocsr()
    P.sort_indices()
    np.testing.assert_array_equal(tptr, P.indptr)
    np.testing.assert_array_equal(tcol, P.indices)


@pytest.mark.parametrize("name", golden_names())
def test_tile_step2_masks_match_structure(name):
    """maskC rows = pattern of C inside each C tile; rowptr = in-tile scan."""
    g = load_golden(name)
    tA = O.tile_convert(g["m"], g["n"], g["A"])
    tB = O.tile_convert(g["n"], g["k"], g["B"])
    csc = O.tile_csc(tB)
    tptr, tcol = O.tile_step1(tA, tB)
    maskC, rowptrC, offC = O.tile_step2(tA, tB, csc, tptr, tcol)
    ic, jc = g["ic"], g["jc"]
    rows = np.repeat(np.arange(g["m"]), np.diff(ic))
    want = {}
    for r, c in zip(rows, jc):
        want.setdefault((r >> 4, c >> 4), np.zeros(16, np.uint16))[r & 15] |= 1 << (15 - (c & 15))
    nnz = 0
    for I in range(tA["blkm"]):
        for t in range(tptr[I], tptr[I + 1]):
            w = want.get((I, int(tcol[t])), np.zeros(16, np.uint16))
            np.testing.assert_array_equal(maskC[16 * t:16 * t + 16], w)
            pc = np.array([bin(int(v)).count("1") for v in w])
            np.testing.assert_array_equal(rowptrC[16 * t:16 * t + 16], np.cumsum(pc) - pc)
            assert offC[t] == nnz
            nnz += pc.sum()
    assert offC[-1] == nnz == ic[-1]

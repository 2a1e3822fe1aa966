"""GPU parity of the batched form: many independent products laid out as
one block-diagonal product; every C_i must equal the oracle's C_i bitwise,
alone and in the mix, and a HIP-graph capture of the numeric phase must give
the same result on every replay."""
from __future__ import annotations

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import gen
from tests.util import assert_csr_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def handle():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    h = tsg.Handle(0)
    yield h
    h.close()


def dev(m, n, X):
    return tsg.DeviceCsr.from_host(m, n, *X)


def mixed_products(seed=0, count=24):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = i % 4
        if kind == 0:
            n = int(rng.integers(50, 3000))
            A = gen.erdos_renyi(n, float(rng.uniform(2, 16)), seed=seed * 100 + i)
            out.append((n, n, n, A, A, True))
        elif kind == 1:
            g = int(rng.integers(4, 40))
            A = gen.laplacian2d(g)
            out.append((g * g, g * g, g * g, A, A, True))
        elif kind == 2:
            s = int(rng.integers(6, 11))
            A = gen.rmat(s, 8 << s, seed=seed * 100 + i)
            AT = gen.transpose(1 << s, 1 << s, *A)
            out.append((1 << s, 1 << s, 1 << s, A, AT, False))
        else:
            m, n, k = (int(x) for x in rng.integers(1, 700, 3))
            A = gen.erdos_renyi(m, 4.0, seed=seed * 100 + i, ncols=n)
            B = gen.erdos_renyi(n, 5.0, seed=seed * 100 + i + 50, ncols=k)
            out.append((m, n, k, A, B, False))
    return out


def to_host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("method", [tsg.METHOD_AUTO, tsg.METHOD_ROW, tsg.METHOD_TILE])
def test_mixed_batch_matches_oracle(handle, method):
    prods = mixed_products(seed=1)
    pairs = []
    for m, n, k, A, B, same in prods:
        dA = dev(m, n, A)
        pairs.append((dA, dA if same else dev(n, k, B)))
    handle.set_method(method)
    res = handle.batch_spgemm(pairs)
    for (m, n, k, A, B, same), (ic, jc, c) in zip(prods, res):
        want = O.csr_spgemm(m, k, A, B, fused=True)
        assert_csr_equal((to_host(ic), to_host(jc), to_host(c)), want)


@pytest.mark.parametrize("method", [tsg.METHOD_ROW, tsg.METHOD_AUTO])
def test_batch_with_empty_products(handle, method):
    """Products whose A, B or both hold no entries between ordinary ones:
    the concatenation steps over their empty entry ranges (a block's first
    product is searched once, then each thread moves past empty products) and
    every C_i equals the oracle's."""
    E = gen.erdos_renyi(300, 6.0, seed=9)
    Z = (np.zeros(301, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float64))
    L = gen.laplacian2d(12)
    prods = [(300, 300, 300, E, E), (300, 300, 300, Z, E), (144, 144, 144, L, L),
             (300, 300, 300, E, Z), (300, 300, 300, Z, Z), (300, 300, 300, E, E)]
    pairs = [(dev(m, n, A), dev(n, k, B)) for m, n, k, A, B in prods]
    handle.set_method(method)
    res = handle.batch_spgemm(pairs)
    for (m, n, k, A, B), (ic, jc, c) in zip(prods, res):
        assert_csr_equal((to_host(ic), to_host(jc), to_host(c)), O.csr_spgemm(m, k, A, B))


def test_all_square_batch_uses_one_operand(handle):
    """All products A_i * A_i: B is the concatenated A itself."""
    prods = [p for p in mixed_products(seed=2, count=16) if p[5]]
    pairs = [(dev(m, n, A),) * 2 for m, n, k, A, B, same in prods]
    handle.set_method(tsg.METHOD_AUTO)
    res = handle.batch_spgemm(pairs)
    for (m, n, k, A, B, same), (ic, jc, c) in zip(prods, res):
        assert_csr_equal((to_host(ic), to_host(jc), to_host(c)), O.csr_spgemm(m, k, A, A))


def test_batch_result_equals_single_call(handle):
    """Each product gives the same result alone and inside the batch."""
    prods = mixed_products(seed=3, count=12)
    pairs = []
    for m, n, k, A, B, same in prods:
        dA = dev(m, n, A)
        pairs.append((dA, dA if same else dev(n, k, B)))
    handle.set_method(tsg.METHOD_AUTO)
    res = handle.batch_spgemm(pairs)
    for (dA, dB), (ic, jc, c) in zip(pairs, res):
        ic1, jc1, c1 = handle.spgemm(dA, dB)
        assert_csr_equal((to_host(ic), to_host(jc), to_host(c)),
                         (to_host(ic1), to_host(jc1), to_host(c1)))


@pytest.mark.parametrize("method,big", [(tsg.METHOD_ROW, False), (tsg.METHOD_TILE, False),
                                        (tsg.METHOD_ROW, True)])
def test_graph_captured_numeric_replays(handle, method, big):
    """The batched numeric phase captured once in a HIP graph and replayed
    with new values: every replay equals the oracle.  big: every row on the
    big-row path, bands and items in column order (block-major block
    pointers, 8192-column items)."""
    import torch
    prods = [p for p in mixed_products(seed=4, count=12) if p[5]]
    pairs = [(dev(m, n, A),) * 2 for m, n, k, A, B, same in prods]
    s = torch.cuda.Stream()
    handle.set_stream(s)
    if big:
        handle.set_tuning({"row_bins": (0, 0, 0)})
        handle.set_item_order(0)
    try:
        handle.set_method(method)
        with torch.cuda.stream(s):
            arr, ic, per, roff, tot = handle.batch_symbolic(pairs)
            jc = torch.empty(tot, dtype=torch.int32, device="cuda:0")
            c = torch.empty(tot, dtype=torch.float64, device="cuda:0")
            handle.batch_numeric(arr, ic, jc, c)          # warm-up outside capture
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            handle.batch_numeric(arr, ic, jc, c)
        for rep in range(3):
            vals = []
            for (m, n, k, A, B, same), (dA, _) in zip(prods, pairs):
                v = np.random.default_rng(10 * rep + m).uniform(-1, 1, A[1].size)
                dA.data.copy_(torch.from_numpy(v))
                vals.append(v)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for i, ((m, n, k, A, B, same), v) in enumerate(zip(prods, vals)):
                r0, r1 = int(roff[i]), int(roff[i + 1])
                icp = to_host(ic[r0:r1 + 1])
                e0, e1 = int(icp[0]), int(icp[-1])
                got = (icp - e0, to_host(jc[e0:e1]), to_host(c[e0:e1]))
                X = (A[0], A[1], v)
                assert_csr_equal(got, O.csr_spgemm(m, k, X, X))
    finally:
        handle.set_stream(None)
        handle.set_tuning({"row_bins": (64, 256, 1024)})
        handle.set_item_order(1 << 18)


def test_auto_batch_tile_dense_with_unsorted_rows(handle):
    """A tile-dense batch (27-point stencils) lays out global columns for the
    tile path; one product with reversed rows sends the whole batch to the
    row path, whose block-diagonal columns are then made local.  Both batches
    equal the oracle."""
    S = [gen.stencil27(m, seed=m) for m in (5, 6, 7)]
    dims = [m ** 3 for m in (5, 6, 7)]
    handle.set_method(tsg.METHOD_AUTO)
    pairs = [(dev(n, n, X),) * 2 for n, X in zip(dims, S)]
    res = handle.batch_spgemm(pairs)
    assert handle.method_used == tsg.METHOD_TILE
    for n, X, (ic, jc, c) in zip(dims, S, res):
        assert_csr_equal((to_host(ic), to_host(jc), to_host(c)), O.csr_spgemm(n, n, X, X))
    ip, ix, d = (x.copy() for x in S[1])
    for i in range(dims[1]):
        ix[ip[i]:ip[i + 1]] = ix[ip[i]:ip[i + 1]][::-1].copy()
        d[ip[i]:ip[i + 1]] = d[ip[i]:ip[i + 1]][::-1].copy()
    U = (ip, ix, d)
    pairs[1] = (dev(dims[1], dims[1], U), dev(dims[1], dims[1], S[1]))
    res = handle.batch_spgemm(pairs)
    assert handle.method_used == tsg.METHOD_ROW
    want = [O.csr_spgemm(dims[0], dims[0], S[0], S[0]),
            O.ref_thmkl(dims[1], dims[1], dims[1], U, S[1], "fma") if O.ref_available()
            else O.csr_spgemm(dims[1], dims[1], U, S[1]),
            O.csr_spgemm(dims[2], dims[2], S[2], S[2])]
    for w, (ic, jc, c) in zip(want, res):
        assert_csr_equal((to_host(ic), to_host(jc), to_host(c)), w)

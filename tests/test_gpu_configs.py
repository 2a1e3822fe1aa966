"""GPU parity at the benchmark configurations, at their full sizes, against
the CPU oracle (the restated thmkl_dcsrmultcsr, pinned to the reference's
own CPU build).  Structure is compared exactly and values bit for bit; each
test also reports the max relative error, |got - want| / sum |a_ik b_kj|,
which the 1e-12 bound is stated against (0 when bitwise equal).

* 2D 5-point Laplacian 256x256, C = A*A: whole C, tile and row paths and
  AUTO, both FP modes;
* 27-point stencil 64^3, C = A*A: whole C, tile and row paths;
  nnz(C) = (5*64-6)^3 = 30,959,144 and (9*64-10)^3 intermediate products;
* 512 independent ER products (gen.er_batch(512, seed=7)) as one batch: the
  nnz of every C_i against the oracle's symbolic phase, every 16th C_i bit
  for bit (32 products), and the whole batch's structure sum;
* the served mix (gen.served_mix: 2000 requests, 70% ER, 25% Laplacian
  64^2-512^2, 5% R-MAT 2^18 A*A^T) through the continuous-batching server:
  every result equals its pair's alone-call bit for bit, the ER and
  Laplacian alone-calls equal the oracle's whole C, and the R-MAT ones the
  oracle's C on every 997th row plus the 16 rows with the most products.
"""
from __future__ import annotations

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import gen
from tests.util import assert_csr_equal, max_rel_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def handle():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    h = tsg.Handle(0)
    yield h
    h.close()


def _host(t):
    return t.cpu().numpy()


def _report(tag, A, B, m, k, got, want):
    scale = O.csr_absdot(m, k, A, B, want[0], want[1])
    err = max_rel_error(got[2], want[2], scale)
    print(f"[{tag}] nnz(C)={int(want[0][-1])} max rel err {err:.3g}", flush=True)
    assert err <= 1e-12
    return err


@pytest.mark.parametrize("mname,method", [("tile", tsg.METHOD_TILE), ("row", tsg.METHOD_ROW),
                                          ("auto", tsg.METHOD_AUTO)])
@pytest.mark.parametrize("fp", [tsg.FP_FUSED, tsg.FP_SEPARATE])
def test_laplacian256_aa(handle, mname, method, fp):
    A = gen.laplacian2d(256)
    n = 65536
    want = O.csr_spgemm(n, n, A, A, fused=fp == tsg.FP_FUSED)
    assert int(want[0][-1]) == 846_852
    dA = tsg.DeviceCsr.from_host(n, n, *A)
    handle.set_method(method)
    handle.set_fp_mode(fp)
    ic, jc, c = handle.spgemm(dA, dA)
    got = (_host(ic), _host(jc), _host(c))
    handle.set_fp_mode(tsg.FP_FUSED)
    handle.set_method(tsg.METHOD_AUTO)
    _report(f"lap256 {mname} fp{fp}", A, A, n, n, got, want)
    assert_csr_equal(got, want)


@pytest.mark.parametrize("mname,method", [("tile", tsg.METHOD_TILE), ("row", tsg.METHOD_ROW)])
def test_stencil27_64_aa(handle, mname, method):
    A = gen.stencil27(64, seed=0)
    n = 64 ** 3
    dA = tsg.DeviceCsr.from_host(n, n, *A)
    assert handle.intermediate_products(dA, dA) == (9 * 64 - 10) ** 3
    want = O.csr_spgemm(n, n, A, A)
    assert int(want[0][-1]) == (5 * 64 - 6) ** 3 == 30_959_144
    handle.set_method(method)
    ic, jc, c = handle.spgemm(dA, dA)
    got = (_host(ic), _host(jc), _host(c))
    handle.set_method(tsg.METHOD_AUTO)
    _report(f"st27_64 {mname}", A, A, n, n, got, want)
    assert_csr_equal(got, want)


def test_er_batch_512(handle):
    """The batched configuration exactly as benchmarked: every product's
    nnz, and every 16th product bit for bit."""
    import torch
    mats = gen.er_batch(512, seed=7)
    pairs = []
    for n, X in mats:
        d = tsg.DeviceCsr.from_host(n, n, *X)
        pairs.append((d, d))
    handle.set_method(tsg.METHOD_AUTO)
    out = handle.batch_spgemm(pairs)
    torch.cuda.synchronize()
    per = np.array([int(x[1].numel()) for x in out], dtype=np.int64)
    want_per = np.array([int(O.csr_symbolic(n, n, X, X)[-1]) for n, X in mats], dtype=np.int64)
    np.testing.assert_array_equal(per, want_per)
    print(f"[er512] nnz(C) total {int(per.sum())}, per-product counts equal", flush=True)
    worst = 0.0
    for i in range(0, 512, 16):
        n, X = mats[i]
        want = O.csr_spgemm(n, n, X, X)
        ic, jc, c = out[i]
        got = (_host(ic), _host(jc), _host(c))
        assert_csr_equal(got, want)
        worst = max(worst, _report(f"er512 #{i}", X, X, n, n, got, want))
    print(f"[er512] 32 sampled products bitwise, max rel err {worst:.3g}", flush=True)


def _sub_rows(A, rows):
    ia, ja, a = (np.asarray(x) for x in A)
    lens = (ia[rows + 1] - ia[rows]).astype(np.int64)
    ptr = np.zeros(rows.size + 1, dtype=np.int64)
    np.cumsum(lens, out=ptr[1:])
    idx = np.concatenate([np.arange(ia[r], ia[r + 1]) for r in rows])
    return ptr, ja[idx], a[idx]


def _row_products(A, B):
    """Intermediate products of every row of A against B."""
    per_entry = np.diff(np.asarray(B[0], np.int64))[np.asarray(A[1])]
    cs = np.zeros(per_entry.size + 1, dtype=np.int64)
    np.cumsum(per_entry, out=cs[1:])
    ia = np.asarray(A[0], np.int64)
    return cs[ia[1:]] - cs[ia[:-1]]


def _check_rows_against_oracle(tag, A, B, n, k, alone, rows):
    ic, jc, c = alone
    ic_h = _host(ic)
    sub = _sub_rows(A, rows)
    want = O.csr_spgemm(rows.size, k, sub, B)
    pos = np.concatenate([np.arange(ic_h[r], ic_h[r + 1]) for r in rows])
    import torch
    pos_d = torch.from_numpy(pos).to(jc.device)
    got_ic = np.zeros(rows.size + 1, dtype=np.int64)
    np.cumsum(ic_h[rows + 1] - ic_h[rows], out=got_ic[1:])
    got = (got_ic, _host(jc[pos_d]), _host(c[pos_d]))
    assert_csr_equal(got, want)
    _report(tag, sub, B, rows.size, k, got, want)


def test_served_mix_results_equal_alone_and_oracle():
    """The served configuration as benchmarked (2000 requests, R-MAT 2^18
    with 4*2^18 draws, three workers, batches up to 64, 32 GiB budget, all
    submitted at once): each result, fetched in submission order and
    released, equals its operand pair's alone-call."""
    import torch
    kinds, pick, er, lap, rm = gen.served_mix(2000, seed=11, edge_factor=4)
    h = tsg.Handle(0)
    d_er = [tsg.DeviceCsr.from_host(X[0].size - 1, X[0].size - 1, *X) for X in er]
    d_lap = [tsg.DeviceCsr.from_host(s * s, s * s, *X) for s, X in lap]
    d_rm = []
    for X in rm:
        dA = tsg.DeviceCsr.from_host(1 << 18, 1 << 18, *X)
        d_rm.append((dA, h.transpose(dA)))
    pool = {}
    for kd, p in zip(kinds, pick):
        pool[(int(kd), int(p))] = None
    ops = {}
    for kd, p in pool:
        ops[(kd, p)] = ((d_er[p], d_er[p]) if kd == 0 else
                        (d_lap[p], d_lap[p]) if kd == 1 else d_rm[p])
    print(f"[served] 2000 requests over {len(pool)} operand pairs "
          f"(R-MAT requests {int((kinds == 2).sum())})", flush=True)
    alone = {key: h.spgemm(*ops[key]) for key in pool}
    torch.cuda.synchronize()
    srv = tsg.Server(0, workers=3, max_batch=64, mem_budget=32 << 30)
    try:
        tickets = [srv.submit(*ops[(int(kd), int(p))]) for kd, p in zip(kinds, pick)]
        batches = []
        for t, kd, p in zip(tickets, kinds, pick):
            ic, jc, c, res = srv.fetch_device(t)
            a_ic, a_jc, a_c = alone[(int(kd), int(p))]
            assert torch.equal(ic, a_ic)
            assert torch.equal(jc, a_jc)
            assert torch.equal(c.view(torch.int64), a_c.view(torch.int64))
            batches.append(res.batch_size)
    finally:
        srv.close()
    print(f"[served] every result equals its alone-call; mean batch "
          f"{np.mean(batches):.1f}, max {max(batches)}", flush=True)
    # the alone-calls against the oracle
    for (kd, p), res in alone.items():
        if kd == 0:
            X = er[p]
            n = X[0].size - 1
            got = tuple(_host(x) for x in res)
            want = O.csr_spgemm(n, n, X, X)
            assert_csr_equal(got, want)
        elif kd == 1:
            s, X = lap[p]
            got = tuple(_host(x) for x in res)
            want = O.csr_spgemm(s * s, s * s, X, X)
            assert_csr_equal(got, want)
        else:
            n = 1 << 18
            A = rm[p]
            B = gen.transpose(n, n, *A)
            prods = _row_products(A, B)
            heavy = np.argsort(prods)[-16:]
            rows = np.unique(np.concatenate([np.arange(0, n, 997), heavy])).astype(np.int64)
            _check_rows_against_oracle(f"served rmat18 #{p}", A, B, n, n, res, rows)
    h.close()

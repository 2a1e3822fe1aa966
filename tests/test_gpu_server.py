"""GPU tests of the continuous-batching server: every request's result is
bitwise the same served in a mix as computed alone, for batched (several
workers) and per-call (one worker, batch 1) configurations, including a
request larger than the memory budget."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import gen
from tests.util import assert_csr_equal

pytestmark = pytest.mark.gpu


def request_mix(seed=0, count=40):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        u = rng.random()
        if u < 0.7:
            n = int(rng.integers(1000, 16000))
            A = gen.erdos_renyi(n, 8.0, seed=seed * 1000 + i)
            out.append((n, n, n, A, A, True))
        elif u < 0.95:
            g = int(rng.integers(8, 128))
            A = gen.laplacian2d(g)
            out.append((g * g, g * g, g * g, A, A, True))
        else:
            A = gen.rmat(12, 4 << 12, seed=seed * 1000 + i)
            AT = gen.transpose(4096, 4096, *A)
            out.append((4096, 4096, 4096, A, AT, False))
    return out


def devs(mix):
    out = []
    for m, n, k, A, B, same in mix:
        dA = tsg.DeviceCsr.from_host(m, n, *A)
        out.append((dA, dA if same else tsg.DeviceCsr.from_host(n, k, *B)))
    return out


@pytest.mark.parametrize("workers,max_batch", [(2, 64), (1, 1), (3, 8)])
def test_served_results_equal_alone(workers, max_batch):
    import torch
    assert torch.cuda.is_available()
    mix = request_mix(seed=workers * 10 + max_batch)
    ops = devs(mix)
    srv = tsg.Server(0, workers=workers, max_batch=max_batch, mem_budget=1 << 30)
    try:
        tickets = [srv.submit(A, B) for A, B in ops]
        got = [srv.fetch(t) for t in tickets]
    finally:
        srv.close()
    h = tsg.Handle(0)
    sizes = []
    for (m, n, k, A, B, same), (dA, dB), (ic, jc, c, res) in zip(mix, ops, got):
        ic1, jc1, c1 = h.spgemm(dA, dB)
        alone = (ic1.cpu().numpy(), jc1.cpu().numpy(), c1.cpu().numpy())
        assert_csr_equal((ic, jc, c), alone)
        # the estimator's products, summed for many requests at once
        assert res.products == h.intermediate_products(dA, dB)
        sizes.append(res.batch_size)
    # the last request against the oracle too (structure and bits)
    m, n, k, A, B, same = mix[-1]
    assert_csr_equal(got[-1][:3], O.csr_spgemm(m, k, A, B))
    if max_batch == 1:
        assert max(sizes) == 1
    h.close()


def test_concurrent_submitters_and_budget():
    """Two client threads submit while workers run; a budget smaller than one
    request still serves it (alone)."""
    import torch
    assert torch.cuda.is_available()
    mix = request_mix(seed=5, count=24)
    ops = devs(mix)
    srv = tsg.Server(0, workers=2, max_batch=16, mem_budget=4 << 20)
    tickets = [None] * len(ops)

    def client(idx):
        for i in idx:
            tickets[i] = srv.submit(*ops[i])

    th = [threading.Thread(target=client, args=(range(j, len(ops), 2),)) for j in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    try:
        got = [srv.fetch(t) for t in tickets]
    finally:
        srv.close()
    for (m, n, k, A, B, same), g in zip(mix, got):
        assert_csr_equal(g[:3], O.csr_spgemm(m, k, A, B))


def _malformed(kind):
    """(A, B) host operands, 2000 x 2000, malformed as `kind` says."""
    A = gen.erdos_renyi(2000, 6.0, seed=3)
    ia, ja, a = (x.copy() for x in A)
    B = gen.transpose(2000, 2000, *A)
    ib, jb, b = (x.copy() for x in B)
    if kind == "a_col_high":
        ja[7] = 5000
    elif kind == "a_col_negative":
        ja[ja.size - 3] = -1
    elif kind == "a_ptr_descent":
        ia[5] = ia[7] + 1
    elif kind == "a_ptr_end":
        ia[-1] -= 1
    elif kind == "b_col_high":
        jb[11] = 2000
    elif kind == "b_ptr_start":
        ib[0] = 1
    return (ia, ja, a), (ib, jb, b), kind.startswith("a_")


@pytest.mark.parametrize("kind", ["a_col_high", "a_col_negative", "a_ptr_descent", "a_ptr_end",
                                  "b_col_high", "b_ptr_start"])
def test_malformed_request_fails_alone(kind):
    """A request with a malformed operand is rejected by the estimator with
    an error of its own; the requests around it (submitted with it, so
    estimated in the same group) are served.  A-side faults are submitted as
    A*A (B is A), B-side faults as A*B with B malformed."""
    import torch
    assert torch.cuda.is_available()
    A = gen.erdos_renyi(2000, 6.0, seed=3)
    badA, badB, a_side = _malformed(kind)
    dA = tsg.DeviceCsr.from_host(2000, 2000, *A)
    if a_side:
        dbad = tsg.DeviceCsr.from_host(2000, 2000, *badA)
        pair = (dbad, dbad)
    else:
        pair = (dA, tsg.DeviceCsr.from_host(2000, 2000, *badB))
    srv = tsg.Server(0, workers=2, max_batch=8)
    try:
        t = [srv.submit(dA, dA), srv.submit(*pair), srv.submit(dA, dA)]
        with pytest.raises(tsg.TsgError):
            srv.wait(t[1])
        srv.release(t[1])
        want = O.csr_spgemm(2000, 2000, A, A)
        for x in (t[0], t[2]):
            assert_csr_equal(srv.fetch(x)[:3], want)
    finally:
        srv.close()

"""tsg_spgemm_noalloc, the allocation-free symbolic + numeric call: through
the C ABI it must give the bits of the allocating device entry
(tsg_spgemm_symbolic + tsg_spgemm_numeric) on the same handle settings, on
consecutive calls that share the device scratch (the second with a smaller
input, so stale contents of the first would show), and refuse what does not
fit its scratch or the caller's nzmax with a clear error, after which it
still works."""
from __future__ import annotations

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import gen
from tests.util import assert_csr_equal

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20   # TSG_NOALLOC_SCRATCH_BYTES


@pytest.fixture(scope="module")
def handle():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    h = tsg.Handle(0)
    yield h
    h.close()


def _dev(m, n, X):
    return tsg.DeviceCsr.from_host(m, n, *X)


def _noalloc(handle, dA, dB, nzmax):
    import torch
    ic = torch.empty(dA.rows + 1, dtype=torch.int64, device="cuda:0")
    jc = torch.full((max(nzmax, 1),), -7, dtype=torch.int32, device="cuda:0")
    c = torch.full((max(nzmax, 1),), float("nan"), dtype=torch.float64, device="cuda:0")
    nnz = handle.spgemm_noalloc(dA, dB, ic, jc[:nzmax], c[:nzmax])
    return ic.cpu().numpy(), jc[:nnz].cpu().numpy(), c[:nnz].cpu().numpy()


def _allocating(handle, dA, dB):
    ic, jc, c = handle.spgemm(dA, dB)
    return ic.cpu().numpy(), jc.cpu().numpy(), c.cpu().numpy()


def _pairs(method):
    """(name, m, n, k, A, B, same): a large then a smaller product of each
    kind -- R-MAT A*A^T with rows on the big-row path, a 27-point stencil
    (tile path under AUTO's default density only when forced).  The forced
    tile path keeps every C tile's structure and matched tile pairs, so its
    large R-MAT case is scale 12 (about 17 MiB of the scratch; scale 13 with
    100K draws would need about 70 MiB and is refused)."""
    sc = 12 if method == tsg.METHOD_TILE else 13
    R = gen.rmat(sc, 40_000 if sc == 12 else 100_000, seed=3)
    RT = gen.transpose(1 << sc, 1 << sc, *R)
    r = gen.rmat(10, 6_000, seed=4)
    rT = gen.transpose(1 << 10, 1 << 10, *r)
    S = gen.stencil27(14, seed=1)
    s = gen.stencil27(6, seed=2)
    return [(f"rmat{sc}_aat", 1 << sc, 1 << sc, 1 << sc, R, RT, False),
            ("rmat10_aat", 1 << 10, 1 << 10, 1 << 10, r, rT, False),
            ("st27_14", 14 ** 3, 14 ** 3, 14 ** 3, S, S, True),
            ("st27_6", 6 ** 3, 6 ** 3, 6 ** 3, s, s, True)]


@pytest.mark.parametrize("mname,method", [("auto", tsg.METHOD_AUTO), ("row", tsg.METHOD_ROW),
                                          ("tile", tsg.METHOD_TILE)])
def test_noalloc_matches_allocating_entry(handle, mname, method):
    handle.set_method(method)
    try:
        for name, m, n, k, A, B, same in _pairs(method):
            dA = _dev(m, n, A)
            dB = dA if same else _dev(n, k, B)
            want = _allocating(handle, dA, dB)
            nnz = int(want[0][-1])
            got = _noalloc(handle, dA, dB, nnz)
            print(f"[noalloc {mname} {name}] nnz(C)={nnz} used={handle.method_used}", flush=True)
            assert_csr_equal(got, want)
            assert_csr_equal(got, O.csr_spgemm(m, k, A, B))
    finally:
        handle.set_method(tsg.METHOD_AUTO)


def test_noalloc_leaves_the_handle_plan(handle):
    """A plan made by tsg_spgemm_symbolic still serves tsg_spgemm_numeric
    after a no-alloc call on the same handle."""
    import torch
    S = gen.stencil27(9, seed=5)
    n = 729
    dS = _dev(n, n, S)
    ic, nnz = handle.symbolic(dS, dS)
    E = gen.erdos_renyi(2000, 7.0, seed=2)
    dE = _dev(2000, 2000, E)
    _noalloc(handle, dE, dE, 200_000)
    jc = torch.empty(nnz, dtype=torch.int32, device="cuda:0")
    c = torch.empty(nnz, dtype=torch.float64, device="cuda:0")
    handle.numeric(dS, dS, ic, jc, c)
    assert_csr_equal((ic.cpu().numpy(), jc.cpu().numpy(), c.cpu().numpy()),
                     O.csr_spgemm(n, n, S, S))


def test_noalloc_refusals_then_recovers(handle):
    import torch
    E = gen.erdos_renyi(3000, 6.0, seed=8)
    dE = _dev(3000, 3000, E)
    want = O.csr_spgemm(3000, 3000, E, E)
    nnz = int(want[0][-1])
    # nzmax below nnz(C): TSG_ERR_OVERFLOW, with ic and nnz(C) written
    ic = torch.empty(3001, dtype=torch.int64, device="cuda:0")
    jc = torch.empty(nnz - 1, dtype=torch.int32, device="cuda:0")
    c = torch.empty(nnz - 1, dtype=torch.float64, device="cuda:0")
    with pytest.raises(tsg.TsgError) as e:
        handle.spgemm_noalloc(dE, dE, ic, jc, c)
    assert e.value.code == -4, e.value
    np.testing.assert_array_equal(ic.cpu().numpy(), want[0])
    # rows far beyond the scratch: refused before any work
    big = 3_000_000
    Z = (np.zeros(big + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    dZ = _dev(big, 3000, Z)
    with pytest.raises(tsg.TsgError) as e:
        _noalloc(handle, dZ, dE, 1)
    assert e.value.code == -1 and "scratch" in str(e.value), e.value
    # a workspace that outgrows the scratch part-way (B's 3M entries in
    # unsorted rows need a sorted copy and its transposes): TSG_ERR_ALLOC
    rng = np.random.default_rng(1)
    nb, kb = 60_000, 100_000
    rows = np.repeat(np.arange(nb), 50)
    cols = rng.integers(0, kb, rows.size)
    ip = np.arange(0, rows.size + 1, 50, dtype=np.int32)
    U = (ip, cols.astype(np.int32), rng.uniform(-1, 1, rows.size))   # rows unsorted
    # sorted columns, permutation and values (16 B per entry) and the first
    # transpose's rows and positions (8 B per entry)
    assert 24 * rows.size > SCRATCH
    A = gen.erdos_renyi(16, 3.0, seed=2, ncols=nb)
    with pytest.raises(tsg.TsgError) as e:
        _noalloc(handle, _dev(16, nb, A), _dev(nb, kb, U), 10_000)
    assert e.value.code == -3 and "scratch" in str(e.value), e.value
    # and the scratch serves the next call as before
    assert_csr_equal(_noalloc(handle, dE, dE, nnz), want)

"""Autotuner: the search on synthetic cost functions (CPU) and the tuning
parameters' invariance of results on the GPU (every setting gives the
oracle's bits; only speed may change)."""
from __future__ import annotations

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import autotune as AT
from tilespgemm_amd import gen
from tests.util import assert_csr_equal


def test_descent_finds_separable_minimum():
    target = {"bin1": 32, "bin2": 128, "bin3": 512, "tile_density": 64.0}

    def cost(t):
        b1, b2, b3 = t["row_bins"]
        return (1.0 + abs(np.log2(b1 / target["bin1"])) + abs(np.log2(b2 / target["bin2"]))
                + abs(np.log2(b3 / target["bin3"]))
                + abs(np.log2(t["tile_density"] / target["tile_density"])))

    best, ms, trials = AT.coordinate_descent(cost)
    assert best == {"tile_density": 64.0, "row_bins": (32, 128, 512), "block_pointer_budget": -1}
    assert ms == pytest.approx(1.0)
    # every measured setting is valid and measured once
    keys = [repr(sorted(t.items())) for t, _ in trials]
    assert len(keys) == len(set(keys))
    assert all(AT.valid(t) for t, _ in trials)


def test_descent_keeps_default_without_gain():
    best, ms, trials = AT.coordinate_descent(lambda t: 5.0)
    assert best == {"tile_density": 32.0, "row_bins": (64, 256, 1024), "block_pointer_budget": -1}
    assert ms == 5.0


def test_descent_respects_bin_order():
    # cheapest at bin3 = 256 only if bin2 <= 256: always valid; bin2 = 256 with
    # bin3 = 256 allowed, bin1 > bin2 never proposed
    seen = []
    AT.coordinate_descent(lambda t: seen.append(t) or 1.0)
    for t in seen:
        b1, b2, b3 = t["row_bins"]
        assert b1 <= b2 <= b3


def test_tune_takes_a_measure_and_applies_the_best():
    """tune() with a measure of its own (the numeric-phase-only objective a
    captured graph needs): the cheapest setting is applied to the handle and
    reported beside the default's time."""
    class H:
        applied = None

        def set_tuning(self, t):
            H.applied = dict(t)

    def cost(t):
        return 10.0 - (t["row_bins"][2] == 256) * 3.0

    res = AT.tune(H(), None, measure=cost)
    assert res["row_bins"] == [64, 256, 256] and res["block_pointer_budget"] == -1
    assert res["ms"] == 7.0 and res["default_ms"] == 10.0
    assert tuple(H.applied["row_bins"]) == (64, 256, 256)


def test_signature_is_stable():
    class M:
        def __init__(self, rows, nnz):
            self.rows, self.nnz = rows, nnz
    pairs = [(M(1000, 8000), None), (M(60000, 1_900_000), None)]
    assert AT.signature(pairs) == AT.signature(list(pairs))
    assert AT.signature(pairs).startswith("n2_rows10-16")


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", [
    {"tile_density": 1e9, "row_bins": (0, 0, 0)},        # every row to the big-row path
    {"tile_density": 0.5, "row_bins": (16, 64, 256)},
    {"tile_density": 6.0, "row_bins": (64, 64, 64)},
    {"tile_density": 3.0, "row_bins": (32, 256, 1024)}])
def test_tuning_never_changes_results(tuning):
    import torch
    assert torch.cuda.is_available()
    h = tsg.Handle(0)
    h.set_tuning(tuning)
    t = h.tuning()
    assert t["row_bins"] == tuple(tuning["row_bins"])
    cases = [(gen.erdos_renyi(3000, 9.0, seed=2), None), (gen.stencil27(9, seed=3), None),
             (gen.rmat(11, 30000, seed=5), "aat")]
    for A, kind in cases:
        m = A[0].size - 1
        B = gen.transpose(m, m, *A) if kind == "aat" else A
        dA = tsg.DeviceCsr.from_host(m, m, *A)
        dB = tsg.DeviceCsr.from_host(m, m, *B)
        ic, jc, c = h.spgemm(dA, dB)
        got = (ic.cpu().numpy(), jc.cpu().numpy(), c.cpu().numpy())
        assert_csr_equal(got, O.csr_spgemm(m, m, A, B))
    # a batch under the same tuning
    mats = gen.er_batch(12, seed=9, nmax=4000)
    pairs = [(tsg.DeviceCsr.from_host(n, n, *X),) * 2 for n, X in mats]
    outs = h.batch_spgemm(pairs)
    for (n, X), (ic, jc, c) in zip(mats, outs):
        got = (ic.cpu().numpy(), jc.cpu().numpy(), c.cpu().numpy())
        assert_csr_equal(got, O.csr_spgemm(n, n, X, X))
    h.close()


@pytest.mark.gpu
def test_invalid_tuning_rejected():
    h = tsg.Handle(0)
    for bins in [(65, 100, 200), (64, 300, 400), (64, 256, 2000), (100, 50, 60)]:
        with pytest.raises(tsg.TsgError):
            h.set_tuning({"row_bins": bins})
    with pytest.raises(tsg.TsgError):
        h.set_tuning({"tile_density": 0.0})
    assert h.tuning() == {"tile_density": 32.0, "row_bins": (64, 256, 1024)}
    h.close()

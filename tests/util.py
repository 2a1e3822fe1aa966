"""Shared helpers for the tests (fixture loading, comparisons)."""
from __future__ import annotations

import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    A = (z["a_ptr"], z["a_idx"], z["a_val"])
    B = (z["b_ptr"], z["b_idx"], z["b_val"])
    return dict(m=int(z["m"]), n=int(z["n"]), k=int(z["k"]), A=A, B=B, ic=z["ic"], jc=z["jc"],
                c_fma=z["c_fma"], c_nofma=z["c_nofma"])


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.int64)


def assert_csr_equal(got, want, values="bitwise", rtol=1e-12):
    """got/want = (ic, jc, c).  Structure must match exactly; values bitwise
    or within rtol relative to max(|want|, tiny)."""
    gic, gjc, gc = (np.asarray(x) for x in got)
    wic, wjc, wc = (np.asarray(x) for x in want)
    assert gic.shape == wic.shape, (gic.shape, wic.shape)
    np.testing.assert_array_equal(gic.astype(np.int64), wic.astype(np.int64))
    np.testing.assert_array_equal(gjc.astype(np.int64), wjc.astype(np.int64))
    if values == "bitwise":
        bad = np.nonzero(bits(gc) != bits(wc))[0]
        assert bad.size == 0, f"{bad.size} values differ, first at {bad[:5]}: " \
                              f"{gc[bad[:5]]} vs {wc[bad[:5]]}"
    else:
        err = np.abs(gc - wc) / np.maximum(np.abs(wc), 1e-300)
        assert err.max(initial=0.0) <= rtol, f"max rel err {err.max()}"


def max_rel_error(got_c, want_c, scale):
    """max over outputs of |got - want| / scale, where scale is the output's
    sum of |a_ik * b_kj| over its products (oracle.csr_absdot): the bound a
    sequential FP64 sum meets independently of summation order.  An output of
    scale 0 must match exactly (its error counts as infinite otherwise)."""
    g = np.asarray(got_c, dtype=np.float64)
    w = np.asarray(want_c, dtype=np.float64)
    s = np.asarray(scale, dtype=np.float64)
    d = np.abs(g - w)
    if d.size == 0:
        return 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(s > 0, d / np.where(s > 0, s, 1.0), np.where(d > 0, np.inf, 0.0))
    return float(r.max())

"""Writes the golden fixtures in this directory.

Each fixture holds seeded input matrices (A, B) and the outputs of the
reference's own CPU SpGEMM, thmkl_dcsrmultcsr.c compiled from its sources by
oracle/oracle.mk (oracle/_ref/libthmkl_ref_{fma,nofma}.so):
  ic, jc          structure (identical for both builds)
  c_fma           values from the -march=x86-64-v3 build (gcc fuses the
                  accumulate into an FMA, as the reference's -march=native
                  build does)
  c_nofma         values from the -ffp-contract=off build

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import py_oracle as O  # noqa: E402
from tilespgemm_amd import gen  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def csr_from_dense(M):
    M = np.asarray(M)
    m, n = M.shape
    rows, cols = np.nonzero(M)
    return gen.from_coo(m, rows, cols, M[rows, cols]) if rows.size else (
        np.zeros(m + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))


def with_explicit_zeros(A, every=7):
    ip, ix, d = A
    d = d.copy()
    d[::every] = 0.0
    return ip, ix, d


def cases():
    """(name, m, n, k, A, B) -- B is A for A*A, A^T for A*A^T."""
    out = []
    L = gen.laplacian2d(16)
    out.append(("lap16_AA", 256, 256, 256, L, L))
    Lr = (L[0], L[1], gen.ref_values(L[1].size))           # reference's i%10 values
    out.append(("lap16_AA_refvals", 256, 256, 256, Lr, Lr))
    S = gen.stencil27(6, seed=11)
    out.append(("st27_6_AA", 216, 216, 216, S, S))
    E = gen.erdos_renyi(300, 6.0, seed=5)
    out.append(("er300_AA", 300, 300, 300, E, E))
    R = gen.rmat(9, 3000, seed=3)
    RT = gen.transpose(512, 512, *R)
    out.append(("rmat9_AAT", 512, 512, 512, R, RT))
    # rectangular, m and k not multiples of 16
    A = gen.erdos_renyi(203, 5.0, seed=7, ncols=157)
    B = gen.erdos_renyi(157, 4.0, seed=8, ncols=91)
    out.append(("rect_203x157x91", 203, 157, 91, A, B))
    # empty rows and empty tile rows
    rng = np.random.default_rng(9)
    D = (rng.random((70, 70)) < 0.05) * rng.uniform(-1, 1, (70, 70))
    D[5:40, :] = 0.0
    D[:, 20:30] = 0.0
    Dz = csr_from_dense(D)
    out.append(("holes_70", 70, 70, 70, Dz, Dz))
    # a dense 16x16 block inside a sparse matrix, with explicit zeros
    D2 = (rng.random((48, 48)) < 0.04) * rng.uniform(-1, 1, (48, 48))
    D2[16:32, 16:32] = rng.uniform(-1, 1, (16, 16))
    Dd = csr_from_dense(D2)
    Dd = with_explicit_zeros(Dd, every=5)
    out.append(("denseblock_48", 48, 48, 48, Dd, Dd))
    # cancellation: integer values whose products cancel to exact 0.0
    P = np.zeros((32, 32))
    P[0, 1], P[0, 2] = 1.0, -1.0
    P[1, 5], P[2, 5] = 3.0, 3.0
    P[3, 3] = 2.0
    Pc = csr_from_dense(P)
    out.append(("cancel_32", 32, 32, 32, Pc, Pc))
    # one row / one column
    Z1 = gen.erdos_renyi(1, 3.0, seed=2, ncols=40)
    Z2 = gen.erdos_renyi(40, 2.0, seed=4, ncols=1)
    out.append(("row_times_col", 1, 40, 1, Z1, Z2))
    return out


def main():
    if not (O.ref_available("fma") and O.ref_available("nofma")):
        O.build()
    for name, m, n, k, A, B in cases():
        ic, jc, c_fma = O.ref_thmkl(m, n, k, A, B, "fma")
        ic2, jc2, c_nofma = O.ref_thmkl(m, n, k, A, B, "nofma")
        assert np.array_equal(ic, ic2) and np.array_equal(jc, jc2)
        np.savez_compressed(
            os.path.join(OUT, f"{name}.npz"), m=m, n=n, k=k,
            a_ptr=np.asarray(A[0], np.int32), a_idx=np.asarray(A[1], np.int32),
            a_val=np.asarray(A[2], np.float64),
            b_ptr=np.asarray(B[0], np.int32), b_idx=np.asarray(B[1], np.int32),
            b_val=np.asarray(B[2], np.float64),
            ic=ic.astype(np.int32), jc=jc.astype(np.int32), c_fma=c_fma, c_nofma=c_nofma)
        print(f"{name}: m={m} n={n} k={k} nnzA={A[1].size} nnzC={int(ic[-1])}")


if __name__ == "__main__":
    main()

"""ctypes front-end of the CPU oracle -- test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg
import this module, and only as the checker or the timed CPU baseline.

* ``liboracle.so``  -- this repository's C restatement of the reference
  (csr_oracle.c: thmkl_dcsrmultcsr.c:7-95; tile_oracle.c: the tile format and
  tile steps of main_spgemm-col.cu / main_spgemm1.cu).
* ``_ref/libthmkl_ref_{fma,nofma}.so`` -- the reference's own
  thmkl_dcsrmultcsr.c compiled from its sources by oracle/oracle.mk.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TS = 16

_lib = None
_ref = {}


def build():
    """Compile liboracle.so (and _ref/ when the reference sources exist)."""
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "oracle.mk")],
                   check=True, cwd=os.path.dirname(HERE))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = C.CDLL(path)
        _lib.orc_csr_symbolic.restype = C.c_int64
        _lib.orc_csr_numeric.restype = C.c_int
        _lib.orc_intermediate_products.restype = C.c_uint64
        _lib.orc_tile_count.restype = C.c_int
        _lib.orc_tile_step1.restype = C.c_int
        _lib.orc_tile_step1_intersect.restype = C.c_int
        _lib.orc_tile_step2.restype = C.c_int64
    return _lib


def ref_available(variant: str = "fma") -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libthmkl_ref_{variant}.so"))


def ref_lib(variant: str = "fma"):
    if variant not in _ref:
        _ref[variant] = C.CDLL(os.path.join(HERE, "_ref", f"libthmkl_ref_{variant}.so"))
    return _ref[variant]


def _i64(x):
    return np.ascontiguousarray(x, dtype=np.int64)


def _i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


# ----------------------------------------------------------------- CSR oracle
def csr_symbolic(m, k, A, B):
    ia, ja = _i64(A[0]), _i32(A[1])
    ib, jb = _i64(B[0]), _i32(B[1])
    ic = np.zeros(m + 1, dtype=np.int64)
    nnz = lib().orc_csr_symbolic(C.c_int(m), C.c_int(k), _p(ia), _p(ja), _p(ib), _p(jb), _p(ic))
    if nnz < 0:
        raise MemoryError("oracle symbolic failed")
    return ic


def csr_spgemm(m, k, A, B, fused: bool = True):
    """C = A*B by the restated thmkl_dcsrmultcsr; returns (ic, jc, c)."""
    ia, ja, a = _i64(A[0]), _i32(A[1]), _f64(A[2])
    ib, jb, b = _i64(B[0]), _i32(B[1]), _f64(B[2])
    ic = csr_symbolic(m, k, (ia, ja), (ib, jb))
    nnz = int(ic[-1])
    jc = np.empty(nnz, dtype=np.int32)
    c = np.empty(nnz, dtype=np.float64)
    rc = lib().orc_csr_numeric(C.c_int(m), C.c_int(k), _p(ia), _p(ja), _p(a),
                               _p(ib), _p(jb), _p(b), _p(ic), _p(jc), _p(c),
                               C.c_int(1 if fused else 0))
    if rc != 0:
        raise MemoryError("oracle numeric failed")
    return ic, jc, c


def csr_absdot(m, k, A, B, ic, jc):
    """Per output entry, sum of |a_ik b_kj| over its products."""
    ia, ja, a = _i64(A[0]), _i32(A[1]), _f64(A[2])
    ib, jb, b = _i64(B[0]), _i32(B[1]), _f64(B[2])
    ic, jc = _i64(ic), _i32(jc)
    out = np.empty(jc.size, dtype=np.float64)
    lib().orc_csr_absdot(C.c_int(m), C.c_int(k), _p(ia), _p(ja), _p(a),
                         _p(ib), _p(jb), _p(b), _p(ic), _p(jc), _p(out))
    return out


def intermediate_products(A_indices, B_indptr) -> int:
    ja, ib = _i32(A_indices), _i64(B_indptr)
    return int(lib().orc_intermediate_products(C.c_int64(ja.size), _p(ja), _p(ib)))


@contextlib.contextmanager
def _quiet_stdout():
    """The reference prints "nnzc:%d" in its symbolic phase
    (thmkl_dcsrmultcsr.c:48); keep it off our stdout."""
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    try:
        os.dup2(devnull, 1)
        yield
    finally:
        try:
            C.CDLL(None).fflush(None)
        except Exception:
            pass
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)


def ref_thmkl(m, n, k, A, B, variant: str = "fma"):
    """Call the reference's own thmkl_dcsrmultcsr (both phases, as its main()
    does at thmkl_dcsrmultcsr.c:190-210).  int32 throughout, as the reference."""
    fn = ref_lib(variant).thmkl_dcsrmultcsr
    ia, ja, a = _i32(A[0]), _i32(A[1]), _f64(A[2]).copy()
    ib, jb, b = _i32(B[0]), _i32(B[1]), _f64(B[2]).copy()
    ic = np.zeros(m + 1, dtype=np.int32)
    req, srt, nzmax, info = C.c_int(1), C.c_int(1), C.c_int(0), C.c_int(-99)
    cm, cn, ck = C.c_int(m), C.c_int(n), C.c_int(k)
    dummy_j = np.zeros(1, dtype=np.int32)
    dummy_c = np.zeros(1, dtype=np.float64)
    with _quiet_stdout():
        fn(b"n", C.byref(req), C.byref(srt), C.byref(cm), C.byref(cn), C.byref(ck),
           _p(a), _p(ja), _p(ia), _p(b), _p(jb), _p(ib), _p(dummy_c), _p(dummy_j),
           _p(ic), C.byref(nzmax), C.byref(info))
    assert info.value == 1
    nnz = int(ic[-1])
    jc = np.zeros(max(nnz, 1), dtype=np.int32)
    c = np.zeros(max(nnz, 1), dtype=np.float64)
    req = C.c_int(0)
    fn(b"n", C.byref(req), C.byref(srt), C.byref(cm), C.byref(cn), C.byref(ck),
       _p(a), _p(ja), _p(ia), _p(b), _p(jb), _p(ib), _p(c), _p(jc),
       _p(ic), C.byref(nzmax), C.byref(info))
    assert info.value == 0
    return ic, jc[:nnz], c[:nnz]


# ---------------------------------------------------------------- tile oracle
def tile_convert(m, n, A):
    """CSR -> 16x16 bitmask tiles (main_spgemm-col.cu:4968-5240)."""
    ia, ja, a = _i64(A[0]), _i32(A[1]), _f64(A[2])
    blkm = (m + TS - 1) // TS
    tile_ptr = np.zeros(blkm + 1, dtype=np.int32)
    nt = lib().orc_tile_count(C.c_int(m), C.c_int(n), _p(ia), _p(ja), _p(tile_ptr))
    nnz = ja.size
    t = dict(
        m=m, n=n, blkm=blkm, blkn=(n + TS - 1) // TS, numblk=nt,
        tile_ptr=tile_ptr,
        tile_col=np.zeros(nt, dtype=np.int32),
        tile_off=np.zeros(nt + 1, dtype=np.int64),
        mask=np.zeros(nt * TS, dtype=np.uint16),
        rowptr=np.zeros(nt * TS, dtype=np.uint8),
        idx=np.zeros(nnz, dtype=np.uint8),
        val=np.zeros(nnz, dtype=np.float64),
    )
    rc = lib().orc_tile_fill(C.c_int(m), C.c_int(n), _p(ia), _p(ja), _p(a),
                             _p(t["tile_ptr"]), _p(t["tile_col"]), _p(t["tile_off"]),
                             _p(t["mask"]), _p(t["rowptr"]), _p(t["idx"]), _p(t["val"]))
    if rc != 0:
        raise RuntimeError("tile_fill failed")
    return t


def tile_csc(t):
    blkm, blkn = t["blkm"], t["blkn"]
    nt = t["numblk"]
    out = dict(col_ptr=np.zeros(blkn + 1, dtype=np.int32),
               col_row=np.zeros(nt, dtype=np.int32),
               col_tile=np.zeros(nt, dtype=np.int32))
    lib().orc_tile_csc(C.c_int(blkm), C.c_int(blkn), _p(t["tile_ptr"]), _p(t["tile_col"]),
                       _p(out["col_ptr"]), _p(out["col_row"]), _p(out["col_tile"]))
    return out


def tile_step1(tA, tB, fill: bool = True):
    blkm = tA["blkm"]
    tptrC = np.zeros(blkm + 1, dtype=np.int32)
    n = lib().orc_tile_step1(C.c_int(blkm), _p(tA["tile_ptr"]), _p(tA["tile_col"]),
                             C.c_int(tB["blkn"]), _p(tB["tile_ptr"]), _p(tB["tile_col"]),
                             _p(tptrC), None)
    tcolC = np.zeros(max(n, 1), dtype=np.int32)
    if fill:
        lib().orc_tile_step1(C.c_int(blkm), _p(tA["tile_ptr"]), _p(tA["tile_col"]),
                             C.c_int(tB["blkn"]), _p(tB["tile_ptr"]), _p(tB["tile_col"]),
                             _p(tptrC), _p(tcolC))
    return tptrC, tcolC[:n]


def tile_step1_intersect(tA, cscB, blknB):
    blkm = tA["blkm"]
    tptrC = np.zeros(blkm + 1, dtype=np.int32)
    lib().orc_tile_step1_intersect(C.c_int(blkm), _p(tA["tile_ptr"]), _p(tA["tile_col"]),
                                   C.c_int(blknB), _p(cscB["col_ptr"]), _p(cscB["col_row"]),
                                   _p(tptrC))
    return tptrC


def tile_step2(tA, tB, cscB, tptrC, tcolC):
    nC = int(tptrC[-1])
    maskC = np.zeros(max(nC, 1) * TS, dtype=np.uint16)
    rowptrC = np.zeros(max(nC, 1) * TS, dtype=np.uint8)
    offC = np.zeros(nC + 1, dtype=np.int64)
    tcolC = _i32(tcolC) if nC else np.zeros(1, dtype=np.int32)
    lib().orc_tile_step2(C.c_int(tA["blkm"]), _p(tA["tile_ptr"]), _p(tA["tile_col"]),
                         _p(tA["tile_off"]), _p(tA["rowptr"]), _p(tA["idx"]),
                         C.c_int(tB["blkn"]), _p(cscB["col_ptr"]), _p(cscB["col_row"]),
                         _p(cscB["col_tile"]), _p(tB["mask"]),
                         _p(tptrC), _p(tcolC), _p(maskC), _p(rowptrC), _p(offC))
    return maskC[:nC * TS], rowptrC[:nC * TS], offC


def tile_step3(tA, tB, cscB, tptrC, tcolC, offC, maskC, fused: bool = True):
    nnz = int(offC[-1])
    colC = np.zeros(max(nnz, 1), dtype=np.uint8)
    valC = np.zeros(max(nnz, 1), dtype=np.float64)
    nC = int(tptrC[-1])
    tcolC = _i32(tcolC) if nC else np.zeros(1, dtype=np.int32)
    maskC = maskC if nC else np.zeros(TS, dtype=np.uint16)
    lib().orc_tile_step3(C.c_int(tA["blkm"]), _p(tA["tile_ptr"]), _p(tA["tile_col"]),
                         _p(tA["tile_off"]), _p(tA["rowptr"]), _p(tA["idx"]), _p(tA["val"]),
                         _p(cscB["col_ptr"]), _p(cscB["col_row"]), _p(cscB["col_tile"]),
                         _p(tB["tile_off"]), _p(tB["rowptr"]), _p(tB["idx"]), _p(tB["val"]),
                         _p(tptrC), _p(tcolC), _p(offC), _p(maskC), _p(colC), _p(valC),
                         C.c_int(1 if fused else 0))
    return colC[:nnz], valC[:nnz]


def tile_to_csr(m, tptrC, tcolC, offC, maskC, rowptrC, colC, valC):
    ic = np.zeros(m + 1, dtype=np.int64)
    nnz = int(offC[-1])
    jc = np.zeros(max(nnz, 1), dtype=np.int32)
    c = np.zeros(max(nnz, 1), dtype=np.float64)
    nC = int(tptrC[-1])
    tcolC = _i32(tcolC) if nC else np.zeros(1, dtype=np.int32)
    maskC = maskC if nC else np.zeros(TS, dtype=np.uint16)
    rowptrC = rowptrC if nC else np.zeros(TS, dtype=np.uint8)
    colC = colC if nnz else np.zeros(1, dtype=np.uint8)
    valC = valC if nnz else np.zeros(1, dtype=np.float64)
    lib().orc_tile_to_csr(C.c_int(m), _p(tptrC), _p(tcolC), _p(offC), _p(maskC),
                          _p(rowptrC), _p(colC), _p(valC), _p(ic), _p(jc), _p(c))
    return ic, jc[:nnz], c[:nnz]


def tile_spgemm(m, n, k, A, B, fused: bool = True):
    """Whole tile-level chain: convert, step1, step2, step3, tile->CSR."""
    tA = tile_convert(m, n, A)
    tB = tile_convert(n, k, B)
    cscB = tile_csc(tB)
    tptrC, tcolC = tile_step1(tA, tB)
    maskC, rowptrC, offC = tile_step2(tA, tB, cscB, tptrC, tcolC)
    colC, valC = tile_step3(tA, tB, cscB, tptrC, tcolC, offC, maskC, fused)
    return tile_to_csr(m, tptrC, tcolC, offC, maskC, rowptrC, colC, valC)


def mtx_read(text: str):
    """Matrix Market coordinate text -> (m, n, indptr, indices, data,
    symmetric), restating the reference loader mmio_allinone
    (mmio_highlevel.h:593-759): 1-based to 0-based; symmetric/hermitian
    entries off the diagonal mirrored with the same value, the mirror placed
    right after the entry in its row's file order (:700-725); rows keep file
    order, no sort, duplicates kept; pattern -> 1.0, integer -> value,
    complex -> real part (:648-672)."""
    lines = [ln for ln in text.splitlines()]
    banner = lines[0].split()
    field, symmetry = banner[3].lower(), banner[4].lower()
    body = [ln for ln in lines[1:] if ln.strip() and not ln.lstrip().startswith("%")]
    m, n, nz = (int(x) for x in body[0].split())
    mirror = symmetry in ("symmetric", "hermitian")
    rows = [[] for _ in range(m)]
    for ln in body[1:1 + nz]:
        tok = ln.split()
        r, c = int(tok[0]) - 1, int(tok[1]) - 1
        v = 1.0 if field == "pattern" else float(tok[2])
        rows[r].append((c, v))
        if mirror and r != c:
            rows[c].append((r, v))
    indptr = np.zeros(m + 1, dtype=np.int32)
    indptr[1:] = np.cumsum([len(x) for x in rows])
    indices = np.array([c for x in rows for c, _ in x], dtype=np.int32)
    data = np.array([v for x in rows for _, v in x], dtype=np.float64)
    return m, n, indptr, indices, data, mirror

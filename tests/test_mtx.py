"""Matrix Market input (tsg_read_mtx) and the command-line driver.

CPU tests pin the loader to hand-computed CSR arrays for each case the
reference loader handles (mmio_highlevel.h:593-759): field types, the
interleaved placement of mirrored entries of symmetric/hermitian files,
file order kept inside rows, duplicates kept, 1-based indices; the error
codes; and agreement with the oracle's restatement on a generated file.
The GPU test runs the tsg_spgemm binary end to end and checks its CSV line
against the oracle.
"""
from __future__ import annotations

import csv
import os
import subprocess

import numpy as np
import pytest

import tilespgemm_amd as tsg
from oracle import py_oracle as O
from tilespgemm_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tilespgemm_amd", "tsg_spgemm")


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def write_mtx(path, m, n, rows, cols, vals, field="real", symmetry="general"):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {symmetry}\n% generated\n")
        f.write(f"{m} {n} {len(rows)}\n")
        for r, c, v in zip(rows, cols, vals):
            if field == "pattern":
                f.write(f"{r + 1} {c + 1}\n")
            elif field == "integer":
                f.write(f"{r + 1} {c + 1} {int(v)}\n")
            else:
                f.write(f"{r + 1} {c + 1} {float(v)!r}\n")


def test_general_keeps_file_order_and_duplicates(tmp_path):
    p = write(tmp_path, "g.mtx", """%%MatrixMarket matrix coordinate real general
% comment line
3 4 5
2 4 1.5
1 3 -2
2 1 0.25
2 4 7
3 2 1e-3
""")
    m, n, ip, ix, dv, sym = tsg.read_mtx(p)
    assert (m, n, sym) == (3, 4, False)
    assert ip.tolist() == [0, 1, 4, 5]
    assert ix.tolist() == [2, 3, 0, 3, 1]
    assert dv.tolist() == [-2.0, 1.5, 0.25, 7.0, 1e-3]


def test_symmetric_mirrors_in_file_order(tmp_path):
    # entries (1,1,a) (3,1,b) (2,1,c) (3,3,d): row 0 gets a, then b's mirror,
    # then c's mirror; row 2 gets b before d
    p = write(tmp_path, "s.mtx", """%%MatrixMarket matrix coordinate real symmetric
3 3 4
1 1 1
3 1 2
2 1 3
3 3 4
""")
    m, n, ip, ix, dv, sym = tsg.read_mtx(p)
    assert (m, n, sym) == (3, 3, True)
    assert ip.tolist() == [0, 3, 4, 6]
    assert ix.tolist() == [0, 2, 1, 0, 0, 2]
    assert dv.tolist() == [1.0, 2.0, 3.0, 3.0, 2.0, 4.0]


@pytest.mark.parametrize("field,line,value", [
    ("pattern", "2 1", 1.0), ("integer", "2 1 -7", -7.0), ("complex", "2 1 2.5 -9", 2.5),
    ("real", "2 1 3.25", 3.25)])
def test_field_types(tmp_path, field, line, value):
    p = write(tmp_path, "f.mtx", f"%%MatrixMarket matrix coordinate {field} general\n2 2 1\n"
                                 f"{line}\n")
    m, n, ip, ix, dv, _ = tsg.read_mtx(p)
    assert ip.tolist() == [0, 0, 1] and ix.tolist() == [0] and dv.tolist() == [value]


def test_hermitian_mirrors_and_skew_does_not(tmp_path):
    h = write(tmp_path, "h.mtx",
              "%%MatrixMarket matrix coordinate complex hermitian\n2 2 1\n2 1 1.5 2\n")
    _, _, ip, ix, dv, sym = tsg.read_mtx(h)
    assert sym and ip.tolist() == [0, 1, 2] and ix.tolist() == [1, 0]
    assert dv.tolist() == [1.5, 1.5]
    k = write(tmp_path, "k.mtx",
              "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1.5\n")
    _, _, ip, ix, dv, sym = tsg.read_mtx(k)
    assert not sym and ip.tolist() == [0, 0, 1] and ix.tolist() == [0]


@pytest.mark.parametrize("text,code", [
    (None, -101),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -102),
    ("%%MatrixMarket matrix coordinate real general\n2 two 1\n1 1 1\n", -104),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", -105),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", -105),
    ("%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 1 1\n", -104)])
def test_errors(tmp_path, text, code):
    p = str(tmp_path / "missing.mtx") if text is None else write(tmp_path, "e.mtx", text)
    with pytest.raises(tsg.TsgError) as e:
        tsg.read_mtx(p)
    assert e.value.code == code


def test_oracle_restatement_agrees(tmp_path):
    """A larger generated file: symmetric, comments, duplicates, unsorted."""
    rng = np.random.default_rng(4)
    n = 300
    rows = rng.integers(0, n, 2000)
    cols = rng.integers(0, n, 2000)
    lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
    vals = rng.standard_normal(2000)
    p = str(tmp_path / "r.mtx")
    write_mtx(p, n, n, hi, lo, vals, symmetry="symmetric")
    got = tsg.read_mtx(p)
    want = O.mtx_read(open(p).read())
    assert got[:2] == want[:2] and got[5] == want[5]
    for g, w in zip(got[2:5], want[2:5]):
        np.testing.assert_array_equal(g, w)


def test_oracle_restatement_on_hand_case():
    """The restatement itself, on the symmetric hand case above."""
    m, n, ip, ix, dv, sym = O.mtx_read("""%%MatrixMarket matrix coordinate real symmetric
3 3 4
1 1 1
3 1 2
2 1 3
3 3 4
""")
    assert ip.tolist() == [0, 3, 4, 6] and ix.tolist() == [0, 2, 1, 0, 0, 2]
    assert dv.tolist() == [1.0, 2.0, 3.0, 3.0, 2.0, 4.0] and sym


def test_cli_is_built():
    assert os.access(CLI, os.X_OK), "tsg_spgemm missing: run the build"
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("aat", [False, True])
def test_cli_end_to_end(tmp_path, aat):
    """tsg_spgemm on a generated file: CSV line fields against the oracle
    (values i % 10 as the reference driver sets them)."""
    if aat:
        A = gen.rmat(10, 5000, seed=3)
        m = n = 1 << 10
        ip, ix = A[0], A[1]
        A = (ip, ix, np.zeros(ix.size))
        # a general file (the driver refuses a symmetric one for A*A^T)
    else:
        A = gen.stencil27(7, seed=1)
        m = n = 343
    rows = np.repeat(np.arange(m), np.diff(A[0]))
    p = str(tmp_path / ("aat.mtx" if aat else "aa.mtx"))
    write_mtx(p, m, n, rows, A[1], np.ones(A[1].size))
    out = str(tmp_path / "results.csv")
    cmd = [CLI, "-d", "0", "--repeat", "3", "--csv", out] + (["--aat"] if aat else []) + [p]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = next(csv.reader(open(out)))
    assert line[0] == p and int(line[1]) == m and int(line[2]) == n
    assert int(line[3]) == A[1].size
    vals = gen.ref_values(A[1].size)
    Av = (A[0], A[1], vals)
    B = gen.transpose(m, n, *Av) if aat else Av
    want = O.csr_spgemm(m, m if aat else n, Av, B)
    assert int(line[4]) == gen.intermediate_products(A[0], A[1], B[0])
    assert int(line[5]) == want[1].size
    assert float(line[6]) == pytest.approx(int(line[4]) / int(line[5]), rel=1e-6)
    assert float(line[7]) > 0 and float(line[8]) > 0


@pytest.mark.gpu
def test_cli_refusals(tmp_path):
    """The reference driver's refusals: A*A of a non-square matrix, A*A^T of
    a square symmetric one (both exit 0 with a message, no CSV line)."""
    rect = str(tmp_path / "rect.mtx")
    write_mtx(rect, 3, 4, [0, 2], [3, 1], [1.0, 2.0])
    sym = str(tmp_path / "sym.mtx")
    write_mtx(sym, 3, 3, [1, 2], [0, 2], [1.0, 2.0], symmetry="symmetric")
    out = str(tmp_path / "r.csv")
    r = subprocess.run([CLI, "-d", "0", "--csv", out, rect], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and "rowA == colA" in r.stdout
    r = subprocess.run([CLI, "-d", "0", "--aat", "--csv", out, sym], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0 and "symmetric" in r.stdout
    assert not os.path.exists(out)

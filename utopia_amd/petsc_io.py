"""PETSc binary matrix / vector files (the on-disk format of utopia's read/write
with the PETSc backend: MatLoad / VecLoad, backend/petsc/utopia_petsc_Matrix.cpp:181-184,
backend/petsc/utopia_petsc_Vector.cpp:356-436).

Layout (big-endian): Mat: int32 classid 1211216, rows, cols, nnz, int32 row
lengths[rows], int32 cols[nnz], float64 values[nnz]. Vec: int32 classid
1211214, n, float64 values[n].
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import Csr

MAT_FILE_CLASSID = 1211216
VEC_FILE_CLASSID = 1211214


def read_mat(path) -> Csr:
    raw = Path(path).read_bytes()
    head = np.frombuffer(raw, ">i4", 4)
    if int(head[0]) != MAT_FILE_CLASSID:
        raise ValueError(f"{path}: not a PETSc binary matrix (classid {int(head[0])})")
    m, n, nnz = (int(v) for v in head[1:])
    off = 16
    lens = np.frombuffer(raw, ">i4", m, off).astype(np.int64)
    off += 4 * m
    col = np.frombuffer(raw, ">i4", nnz, off).astype(np.int32)
    off += 4 * nnz
    val = np.frombuffer(raw, ">f8", nnz, off).astype(np.float64)
    rowptr = np.zeros(m + 1, np.int64)
    np.cumsum(lens, out=rowptr[1:])
    if int(rowptr[-1]) != nnz:
        raise ValueError(f"{path}: row lengths sum to {int(rowptr[-1])}, header says {nnz}")
    # PETSc stores each row with sorted columns; keep that invariant explicit.
    for i in range(m):
        a, b = rowptr[i], rowptr[i + 1]
        seg = col[a:b]
        if b - a > 1 and np.any(seg[1:] <= seg[:-1]):
            order = np.argsort(seg, kind="stable")
            col[a:b] = seg[order]
            val[a:b] = val[a:b][order]
    return Csr(rowptr.astype(np.int32), col, val, (m, n))


def read_vec(path) -> np.ndarray:
    raw = Path(path).read_bytes()
    head = np.frombuffer(raw, ">i4", 2)
    if int(head[0]) != VEC_FILE_CLASSID:
        raise ValueError(f"{path}: not a PETSc binary vector (classid {int(head[0])})")
    n = int(head[1])
    return np.frombuffer(raw, ">f8", n, 8).astype(np.float64)


def write_mat(path, A: Csr):
    m, n = A.shape
    lens = np.diff(A.rowptr.astype(np.int64)).astype(">i4")
    with open(path, "wb") as f:
        f.write(np.array([MAT_FILE_CLASSID, m, n, A.nnz], ">i4").tobytes())
        f.write(lens.tobytes())
        f.write(np.asarray(A.col, ">i4").tobytes())
        f.write(np.asarray(A.val, ">f8").tobytes())


def write_vec(path, x: np.ndarray):
    with open(path, "wb") as f:
        f.write(np.array([VEC_FILE_CLASSID, x.size], ">i4").tobytes())
        f.write(np.asarray(x, ">f8").tobytes())


# ---------------------------------------------------------------- MatrixMarket
# Coordinate MatrixMarket files (the text format utopia reads through the Trilinos
# backend, tests/trilinos/utopia_TrilinosTest.cpp:983-988, and that
# data/matrixmarket/gre_343_343_crg.mm holds). Header
# "%%MatrixMarket matrix coordinate <real|integer|pattern> <general|symmetric|skew-symmetric>",
# '%' comment lines, a "rows cols entries" line, then 1-based "i j [value]" lines.
# symmetric / skew-symmetric files store one triangle; the other is mirrored
# (negated for skew). Entries repeated at one position are summed; rows come out
# with sorted columns.

def read_matrix_market(path) -> Csr:
    with open(path, "r") as f:
        header = f.readline().split()
        if len(header) < 5 or header[0].lower() != "%%matrixmarket" or header[1].lower() != "matrix":
            raise ValueError(f"{path}: not a MatrixMarket matrix file")
        fmt, field, sym = (h.lower() for h in header[2:5])
        if fmt != "coordinate":
            raise ValueError(f"{path}: only coordinate MatrixMarket files are supported (got {fmt})")
        if field not in ("real", "integer", "pattern") or sym not in ("general", "symmetric", "skew-symmetric"):
            raise ValueError(f"{path}: unsupported field/symmetry {field}/{sym}")
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        m, n, nz = (int(t) for t in line.split()[:3])
        data = np.loadtxt(f, ndmin=2) if nz > 0 else np.zeros((0, 3))
    if data.shape[0] != nz:
        raise ValueError(f"{path}: header says {nz} entries, found {data.shape[0]}")
    i = data[:, 0].astype(np.int64) - 1
    j = data[:, 1].astype(np.int64) - 1
    v = np.ones(nz) if field == "pattern" else data[:, 2].astype(np.float64)
    if sym != "general":
        off = i != j
        i, j, v = (np.concatenate([i, j[off]]), np.concatenate([j, i[off]]),
                   np.concatenate([v, (-1.0 if sym == "skew-symmetric" else 1.0) * v[off]]))
    if nz and (i.min() < 0 or j.min() < 0 or i.max() >= m or j.max() >= n):
        raise ValueError(f"{path}: index out of range")
    # sum repeated positions in file order, then sort each row by column
    key = i * n + j
    order = np.argsort(key, kind="stable")
    key, v = key[order], v[order]
    uniq, start = np.unique(key, return_index=True)
    vals = np.add.reduceat(v, start) if v.size else v
    rows = (uniq // n).astype(np.int64)
    cols = (uniq % n).astype(np.int32)
    rowptr = np.zeros(m + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    np.cumsum(rowptr, out=rowptr)
    return Csr(rowptr.astype(np.int32), cols, np.asarray(vals, np.float64), (m, n))


def write_matrix_market(path, A: Csr):
    m, n = A.shape
    rows = np.repeat(np.arange(m), np.diff(A.rowptr.astype(np.int64)))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{m} {n} {A.nnz}\n")
        for r, c, v in zip(rows, A.col, A.val):
            f.write(f"{r + 1} {c + 1} {float(v)!r}\n")

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

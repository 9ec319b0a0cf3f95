"""The reference's binary file formats (oracle; test infrastructure only).

Restates, field by field:
  * MPX_matrix::fprint_binary / load_MPX_matrix_binary (MPX_matrix.cpp:219-247,
    594-612): size_t number of indices, int64 number of row indices, the
    indices, the CSC matrix;
  * MPXIndex::fprint_binary / load_MPXIndex_binary (MPXIndex.cpp:8-24, 98-120):
    bool inward, bool physical, size_t size, then (non-physical only) size
    states of nrules int16 charges (State::fprint_binary, states.cpp:59-61);
    a physical index takes its charges from the basis on loading;
  * SparseMatrix::fprint_binary / load_SparseMatrix_binary
    (sparse_interface.cpp:727-835): int64 nzmax, nz (-1), rows, cols; int64
    p[cols+1]; int64 i[nzmax]; complex128 x[nzmax]; the reader stops after p
    when nzmax == 0 or p[cols] == 0;
  * Basis::fprint_binary / load_Basis_binary (states.cpp:119-189): size_t
    nrules, int16 rules, size_t d, int16 charges[d][nrules], double
    energies[d];
  * UnitCell::store / load_UnitCell_binary (UnitCell.cpp:70-145): the basis,
    size_t n, n MPX_matrix records, n times (size_t length, double values).
Little-endian, 8-byte size_t, 1-byte bool (the reference's x86-64 build).

Parity: the layouts are pinned by a literal byte string assembled in
tests/test_oracle_formats.py from these rules for a small tensor, by round
trips, and on the GPU by byte-for-byte equality with the library's writer.
No file written by the reference exists in its repository to compare with
(parity of the layout with the reference's own files is unpinned beyond the
source reading above).
"""
from __future__ import annotations

import io
import struct

import numpy as np
import scipy.sparse as sp

from .mpx import MPX
from .qn import MPXIndex


def _u64(b, v):
    b.write(struct.pack("<Q", v))


def _i64(b, v):
    b.write(struct.pack("<q", v))


def write_mpx(buf, A: MPX, nrules: int):
    _u64(buf, len(A.indices))
    _i64(buf, A.numrows)
    for ix in A.indices:
        buf.write(struct.pack("<??", bool(ix.inward), bool(ix.physical)))
        _u64(buf, ix.size)
        if not ix.physical:
            ch = np.asarray(ix.charges, dtype=np.int16).reshape(ix.size, nrules)
            buf.write(ch.astype("<i2").tobytes())
    m = A.matrix.tocsc()
    m.eliminate_zeros()
    m.sort_indices()
    nnz = int(m.indptr[-1])
    _i64(buf, nnz)
    _i64(buf, -1)
    _i64(buf, m.shape[0])
    _i64(buf, m.shape[1])
    buf.write(np.asarray(m.indptr, dtype="<i8").tobytes())
    if nnz:
        buf.write(np.asarray(m.indices, dtype="<i8").tobytes())
        buf.write(np.asarray(m.data, dtype="<c16").tobytes())


def read_mpx(buf, rules, phys_charges) -> MPX:
    nr = len(rules)
    nidx, = struct.unpack("<Q", buf.read(8))
    numrows, = struct.unpack("<q", buf.read(8))
    idx = []
    for _ in range(nidx):
        inward, physical = struct.unpack("<??", buf.read(2))
        size, = struct.unpack("<Q", buf.read(8))
        if physical:
            ch = np.asarray(phys_charges, dtype=np.int64).reshape(-1, nr)
            if ch.shape[0] != size:
                raise ValueError("physical index size differs from the basis")
        else:
            ch = np.frombuffer(buf.read(2 * nr * size), dtype="<i2").astype(np.int64).reshape(size, nr)
        idx.append(MPXIndex(inward, ch, physical))
    nzmax, nz, rows, cols = struct.unpack("<qqqq", buf.read(32))
    if nz != -1:
        raise ValueError("not in CSC form")
    p = np.frombuffer(buf.read(8 * (cols + 1)), dtype="<i8").astype(np.int64)
    if nzmax and p[-1]:
        i = np.frombuffer(buf.read(8 * nzmax), dtype="<i8").astype(np.int64)[:p[-1]]
        x = np.frombuffer(buf.read(16 * nzmax), dtype="<c16")[:p[-1]]
    else:
        i = np.zeros(0, np.int64)
        x = np.zeros(0, np.complex128)
    A = MPX(rules, idx, numrows, sp.csc_matrix((x, i, p), shape=(rows, cols)))
    return A


def write_basis(buf, rules, charges, energies):
    r = np.asarray(list(rules), dtype="<i2")
    _u64(buf, len(r))
    buf.write(r.tobytes())
    ch = np.asarray(charges, dtype=np.int64).reshape(-1, len(r))
    _u64(buf, ch.shape[0])
    buf.write(ch.astype("<i2").tobytes())
    buf.write(np.asarray(energies, dtype="<f8").tobytes())


def read_basis(buf):
    nr, = struct.unpack("<Q", buf.read(8))
    rules = np.frombuffer(buf.read(2 * nr), dtype="<i2").astype(np.int64)
    d, = struct.unpack("<Q", buf.read(8))
    ch = np.frombuffer(buf.read(2 * nr * d), dtype="<i2").astype(np.int64).reshape(d, nr)
    en = np.frombuffer(buf.read(8 * d), dtype="<f8").copy()
    return rules, ch, en


def write_unitcell(buf, rules, charges, energies, matrices, lambdas):
    if len(matrices) != len(lambdas):
        raise ValueError("Invalid Unit Cell")
    write_basis(buf, rules, charges, energies)
    _u64(buf, len(matrices))
    for m in matrices:
        write_mpx(buf, m, len(list(rules)))
    for l in lambdas:
        _u64(buf, len(l))
        buf.write(np.asarray(l, dtype="<f8").tobytes())


def read_unitcell(buf, charge_rules):
    rules, ch, en = read_basis(buf)
    n, = struct.unpack("<Q", buf.read(8))
    mats = [read_mpx(buf, charge_rules, ch) for _ in range(n)]
    lams = []
    for _ in range(n):
        m, = struct.unpack("<Q", buf.read(8))
        lams.append(np.frombuffer(buf.read(8 * m), dtype="<f8").copy())
    return rules, ch, en, mats, lams


def to_bytes(fn, *args):
    b = io.BytesIO()
    fn(b, *args)
    return b.getvalue()

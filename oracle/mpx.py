"""MPX_matrix restatement on scipy CSC (oracle; test infrastructure only).

Restates the reference's general tensor class and its index choreography:
  * storage: one complex128 CSC whose rows are the mixed radix of the first
    ``numrows`` indices and whose columns are the mixed radix of the rest,
    first index fastest (MPX_matrix.hpp:61-66; sparse_interface.cpp:85-91);
  * zero-flux rule ``IngoingRowState(r) == OutgoingColState(c)``
    (MPX_matrix.cpp:275-302, 380-413);
  * generalised ``reshape`` (sparse_interface.cpp:1005-1102) and the plain
    row-count reshape (sparse_interface.cpp:966-1002);
  * ``reorder`` (MPX_matrix.cpp:630-638), ``ShiftNumRowIndices`` (72-85),
    ``RemoveDummyIndices`` (415-457), ``MoveDummyIndices`` (87-105);
  * ``contract`` / ``contract_to_sparse`` (MPX_matrix.cpp:640-791): kept A
    indices (flipped if conjA) then kept B indices (flipped if conjB),
    numrows = #kept A; contracted indices in pair order;
  * ``GetAllBlockColumns`` (MPX_matrix.cpp:331-355);
  * ``SVD`` with the reference truncation rule (MPX_matrix.cpp:157-192,
    sparse_interface.cpp:1248-1406).

The SpGEMM itself is scipy's (a Gustavson/SMMP product, the same algorithm
family as sparse_interface.cpp:1892-2170, single-threaded); the timed CPU
baseline runs the reference-order matvec through it on one core.

Parity pins (tests/test_oracle_mpx.py): every reshape/contract result is
checked against numpy.einsum on the dense tensors built from the mixed-radix
convention, with random index permutations and pair orders; the zero-flux
invariant is checked on every contraction output; SVD against
numpy.linalg.svd of the dense matrix (singular values, reconstruction,
orthonormality) and the truncation rule against a brute-force restatement
on hand-made spectra (degenerate extension, max_trunc tail, D cap).
Structural zeros kept by the reference SpGEMM and the tie order of exactly
equal singular values (std::sort) are parity unpinned.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .qn import ChargeRules, MPXIndex, match

EPS = np.finfo(np.float64).eps          # SPARSETOL (sparse_interface.hpp:15)


def _prod(xs):
    p = 1
    for x in xs:
        p *= int(x)
    return p


def index_decompose(flat, dims):
    """sparse_interface.cpp:85-91 (first index fastest); vectorised."""
    flat = np.asarray(flat, dtype=np.int64)
    out = []
    for d in dims[:-1]:
        out.append(flat % d)
        flat = flat // d
    out.append(flat)
    return out


def index_compose(parts, dims, n=None):
    if not parts:
        return np.zeros(0 if n is None else n, dtype=np.int64)
    flat = np.zeros_like(np.asarray(parts[0], dtype=np.int64))
    mult = 1
    for p, d in zip(parts, dims):
        flat = flat + np.asarray(p, dtype=np.int64) * mult
        mult *= int(d)
    return flat


class MPX:
    """Restatement of ``ajaj::MPX_matrix`` (MPX_matrix.hpp:55-160)."""

    def __init__(self, rules: ChargeRules, indices, numrows, matrix):
        self.rules = rules
        self.indices = list(indices)
        self.numrows = int(numrows)
        m = sp.csc_matrix(matrix, dtype=np.complex128)
        m.sum_duplicates()
        m.sort_indices()
        self.matrix = m
        r = _prod(i.size for i in self.indices[:self.numrows])
        c = _prod(i.size for i in self.indices[self.numrows:])
        if m.shape != (r, c):
            raise ValueError(f"matrix shape {m.shape} does not match indices ({r},{c})")

    # -- shape helpers ------------------------------------------------------
    def dims(self):
        return [i.size for i in self.indices]

    @property
    def rows(self):
        return self.matrix.shape[0]

    @property
    def cols(self):
        return self.matrix.shape[1]

    def copy(self):
        return MPX(self.rules, list(self.indices), self.numrows, self.matrix.copy())

    # -- QN states: MPX_matrix.cpp:362-413 -------------------------------------
    def _row_states(self, rows):
        acc = np.zeros((len(rows), len(self.rules)), dtype=np.int64)
        parts = index_decompose(rows, self.dims()[:self.numrows]) if self.numrows else []
        for k, ix in enumerate(self.indices[:self.numrows]):
            ch = ix.charges[parts[k]]
            acc = self.rules.add(acc, ch) if ix.inward else self.rules.sub(acc, ch)
        return acc

    def _col_states(self, cols):
        acc = np.zeros((len(cols), len(self.rules)), dtype=np.int64)
        cdims = self.dims()[self.numrows:]
        parts = index_decompose(cols, cdims)
        for k, ix in enumerate(self.indices[self.numrows:]):
            ch = ix.charges[parts[k]]
            acc = self.rules.add(acc, ch) if ix.outgoing else self.rules.sub(acc, ch)
        return acc

    def ingoing_row_state(self, r):
        return self._row_states(np.array([r]))[0]

    def outgoing_col_state(self, c):
        return self._col_states(np.array([c]))[0]

    def is_consistent(self):
        """MPX_matrix.cpp:275-302."""
        m = self.matrix.tocoo()
        if m.nnz == 0:
            return True
        rs = self._row_states(m.row.astype(np.int64))
        cs = self._col_states(m.col.astype(np.int64))
        return bool(np.array_equal(rs, cs))

    def get_all_block_columns(self):
        """MPX_matrix.cpp:331-355: groups of columns with equal outgoing
        state, in order of first appearance."""
        cs = self._col_states(np.arange(self.cols, dtype=np.int64))
        groups = []
        lookup = {}
        for c in range(self.cols):
            key = tuple(int(x) for x in cs[c])
            if key in lookup:
                groups[lookup[key]][1].append(c)
            else:
                lookup[key] = len(groups)
                groups.append((key, [c]))
        return groups

    # -- dense view for tests ------------------------------------------------
    def to_dense_tensor(self):
        """Dense N-index tensor T[i0, i1, ...] (numpy C-order axes =
        index order) from the mixed-radix convention."""
        m = self.matrix.toarray()
        dims = self.dims()
        flat = m.reshape(-1, order="F")          # flat = row + rows*col
        return flat.reshape(dims[::-1]).transpose(list(range(len(dims)))[::-1])

    @staticmethod
    def from_dense_tensor(rules, indices, numrows, T, drop_zeros=True):
        dims = [i.size for i in indices]
        r = _prod(dims[:numrows])
        flat = np.asarray(T, dtype=np.complex128).transpose(list(range(len(dims)))[::-1]).reshape(-1)
        m = flat.reshape((r, -1), order="F")
        s = sp.csc_matrix(m)
        if drop_zeros:
            s.eliminate_zeros()
        return MPX(rules, indices, numrows, s)

    # -- index choreography --------------------------------------------------
    def shift_num_row_indices(self, numrows):
        """MPX_matrix.cpp:72-85 + reshape(old,newrows) (966-1002)."""
        if numrows > len(self.indices) or numrows < 1:
            raise ValueError("Incorrect MPX reshape params for change rows")
        newrows = _prod(i.size for i in self.indices[:numrows])
        self.matrix = reshape_rows(self.matrix, newrows)
        self.numrows = numrows
        return self

    def remove_dummy_indices(self, removal):
        """MPX_matrix.cpp:415-457."""
        for i in removal:
            if self.indices[i].size != 1:
                raise ValueError("Not a dummy index!")
        nrem_rows = sum(1 for i in removal if i < self.numrows)
        if nrem_rows == self.numrows:
            raise ValueError("All row indices would be removed")
        keep = [n for n in range(len(self.indices)) if n not in set(removal)]
        self.indices = [self.indices[n] for n in keep]
        self.numrows -= nrem_rows
        return self

    def move_dummy_indices(self, order):
        """MPX_matrix.cpp:87-105."""
        if len(order) != len(self.indices):
            raise ValueError("Incorrect MPX reshape params")
        nondummy = []
        new = []
        newnum = self.numrows
        for pos, n in enumerate(order):
            if self.indices[n].size > 1:
                nondummy.append(n)
            if n == self.numrows:
                newnum = pos
            new.append(self.indices[n])
        if any(b <= a for a, b in zip(nondummy, nondummy[1:])):
            raise ValueError("Non dummy order shouldn't change!")
        self.indices = new
        self.numrows = newnum
        return self

    def rescale(self, factor):
        self.matrix = self.matrix * factor
        return self


def reshape_rows(m, newrows):
    """reshape(old,newrows): same flat order (sparse_interface.cpp:966-1002)."""
    rows, cols = m.shape
    if rows * cols % newrows:
        raise ValueError("Incorrect reshape params")
    if newrows == rows:
        return m.copy()
    coo = m.tocoo()
    flat = coo.row.astype(np.int64) + rows * coo.col.astype(np.int64)
    newcols = rows * cols // newrows
    out = sp.csc_matrix((coo.data, (flat % newrows, flat // newrows)), shape=(newrows, newcols))
    out.sort_indices()
    return out


def reshape_general(m, old_numrows, new_numrows, old_dims, new_order, conjugate):
    """Generalised reshape (sparse_interface.cpp:1005-1102): permute the
    mixed-radix index to ``new_order`` and split rows/cols at
    ``new_numrows``; optionally conjugate values."""
    old_dims = [int(d) for d in old_dims]
    if _prod(old_dims) != m.shape[0] * m.shape[1]:
        raise ValueError("Incorrect old params")
    if len(old_dims) != len(new_order):
        raise ValueError("Incorrect size params")
    coo = m.tocoo()
    rparts = index_decompose(coo.row.astype(np.int64), old_dims[:old_numrows]) if old_numrows else []
    cparts = index_decompose(coo.col.astype(np.int64), old_dims[old_numrows:])
    parts = list(rparts) + list(cparts)
    new_dims = [old_dims[o] for o in new_order]
    new_parts = [parts[o] for o in new_order]
    nr = _prod(new_dims[:new_numrows])
    nc = _prod(new_dims[new_numrows:])
    r = index_compose(new_parts[:new_numrows], new_dims[:new_numrows], coo.nnz)
    c = index_compose(new_parts[new_numrows:], new_dims[new_numrows:], coo.nnz)
    data = np.conj(coo.data) if conjugate else coo.data
    out = sp.csc_matrix((data, (r, c)), shape=(nr, nc))
    out.sort_indices()
    return out


def reorder(A: MPX, conjA: bool, new_order, numrows) -> MPX:
    """MPX_matrix.cpp:630-638."""
    if len(new_order) != len(A.indices):
        raise ValueError("Incorrect MPX reshape params")
    idx = [A.indices[o].flip() if conjA else A.indices[o] for o in new_order]
    return MPX(A.rules, idx, numrows,
               reshape_general(A.matrix, A.numrows, numrows, A.dims(), new_order, conjA))


def _check_pairs(A, conjA, B, conjB, pairs):
    for (a, b) in pairs:
        if not match(A.indices[a], conjA, B.indices[b], conjB):
            raise ValueError(f"Indices don't match for pair ({a},{b})")


def contract_to_sparse(A: MPX, conjA: bool, B: MPX, conjB: bool, pairs):
    """MPX_matrix.cpp:640-742.  All four branches (ready/not ready A/B)
    compute the same product; this restatement uses the general one:
    A reshaped to rows=(kept A in order), cols=(contracted, pair order);
    B reshaped to rows=(contracted, pair order), cols=(kept B in order)."""
    if A.rules != B.rules:
        raise ValueError("Physical Indices don't match!")
    _check_pairs(A, conjA, B, conjB, pairs)
    na, nb = len(A.indices), len(B.indices)
    ca = [p[0] for p in pairs]
    cb = [p[1] for p in pairs]
    if len(set(ca)) != len(ca) or len(set(cb)) != len(cb):
        raise ValueError("Illegal contraction, repeated indices!")
    orderA = [i for i in range(na) if i not in ca] + ca
    orderB = cb + [i for i in range(nb) if i not in cb]
    Ar = reshape_general(A.matrix, A.numrows, na - len(pairs), A.dims(), orderA, conjA)
    Br = reshape_general(B.matrix, B.numrows, len(pairs), B.dims(), orderB, conjB)
    out = (Ar @ Br).tocsc()
    out.sort_indices()
    return out


def contract(A: MPX, conjA: bool, B: MPX, conjB: bool, pairs) -> MPX:
    """MPX_matrix.cpp:744-791."""
    ca = [p[0] for p in pairs]
    cb = [p[1] for p in pairs]
    idx = [(A.indices[a].flip() if conjA else A.indices[a])
           for a in range(len(A.indices)) if a not in ca]
    nrows = len(idx)
    idx += [(B.indices[b].flip() if conjB else B.indices[b])
            for b in range(len(B.indices)) if b not in cb]
    return MPX(A.rules, idx, nrows, contract_to_sparse(A, conjA, B, conjB, pairs))


def reshape_to_vector(A: MPX):
    """MPX_matrix.cpp:793-795."""
    return reshape_rows(A.matrix, A.rows * A.cols)


# ---------------------------------------------------------------------------
# SVD with the reference truncation rule
# ---------------------------------------------------------------------------

def truncation_rule(values, D, max_trunc):
    """Length selection of SparseMatrix::SVD (sparse_interface.cpp:1343-1387).

    ``values`` must already be sorted descending.  Returns
    (length, kept_weight, discarded_weight)."""
    n = len(values)
    length = D if (D > 0 and D < n) else n
    trunc_error = 0.0
    trunc_length = n
    while trunc_error < max_trunc and trunc_length > 0:
        trunc_error += values[trunc_length - 1] ** 2
        trunc_length -= 1
    trunc_length += 1
    if trunc_length < length:
        length = trunc_length
    while length < n and (values[length - 1] - values[length]) / values[length - 1] < 1.0e-3:
        length += 1
    kept = float(np.sum(np.asarray(values[:length]) ** 2)) if length else 0.0
    disc = float(np.sum(np.asarray(values[length:]) ** 2))
    return length, kept, disc


def sparse_svd(m, blocks, D, max_trunc):
    """SparseMatrix::SVD (sparse_interface.cpp:1248-1406).

    ``blocks``: list of column lists (GetAllBlockColumns order).  Returns
    dict(values, U (rows x k csc), Vd (k x cols csc), kept, discarded,
    order) where ``order`` lists (block number, local index) of each kept
    value.  Tie order among exactly equal singular values is by
    (block, local index) here (stable sort); the reference's std::sort
    leaves it unspecified (parity unpinned)."""
    m = sp.csc_matrix(m)
    rows, cols = m.shape
    unsorted = []   # (value, block, local, ucol(dict row->val), vrow(dict col->val))
    for bno, bcols in enumerate(blocks):
        sub = m[:, bcols]
        weight = float(np.sum(np.abs(sub.data) ** 2))
        if weight <= EPS * EPS:
            continue
        nonempty = [c for c in bcols if m.indptr[c] != m.indptr[c + 1]]
        subc = m[:, nonempty]
        rowl = np.unique(subc.tocoo().row)
        dense = subc[rowl, :].toarray()
        u, s, vh = np.linalg.svd(dense, full_matrices=False)
        uth = EPS / float(u.shape[0])
        vth = EPS / float(vh.shape[1])
        for v in range(len(s)):
            ucol = {int(rowl[i]): u[i, v] for i in range(u.shape[0]) if abs(u[i, v]) > uth}
            vrow = {int(nonempty[j]): vh[v, j] for j in range(vh.shape[1]) if abs(vh[v, j]) > vth}
            unsorted.append((float(s[v]), bno, v, ucol, vrow))
    if not unsorted:
        raise ValueError("Error: No singular values. Empty blocks?")
    order = sorted(range(len(unsorted)), key=lambda i: (-unsorted[i][0], unsorted[i][1], unsorted[i][2]))
    vals = [unsorted[i][0] for i in order]
    length, kept, disc = truncation_rule(vals, D, max_trunc)
    keep = order[:length]
    ur, uc, ux, vr, vc, vx = [], [], [], [], [], []
    for k, i in enumerate(keep):
        for r, x in unsorted[i][3].items():
            ur.append(r); uc.append(k); ux.append(x)
        for c, x in unsorted[i][4].items():
            vr.append(k); vc.append(c); vx.append(x)
    U = sp.csc_matrix((ux, (ur, uc)), shape=(rows, length))
    Vd = sp.csc_matrix((vx, (vr, vc)), shape=(length, cols))
    U.sort_indices(); Vd.sort_indices()
    return dict(values=np.array(vals[:length]), U=U, Vd=Vd, kept=kept, discarded=disc,
                order=[(unsorted[i][1], unsorted[i][2]) for i in keep])


def mpx_svd(A: MPX, D=0, min_s=0.0):
    """MPX_matrix::SVD (MPX_matrix.cpp:157-192): rows become U's rows,
    the new bond index charges are the ingoing row state of each U column's
    first stored row; U gets (row indices, new out index), Vd gets (new in
    index, column indices)."""
    blocks = [c for (_, c) in A.get_all_block_columns()]
    r = sparse_svd(A.matrix, blocks, D, min_s)
    U = r["U"]
    newq = []
    for c in range(U.shape[1]):
        if U.indptr[c] == U.indptr[c + 1]:
            raise ValueError("Empty column! Malformed SVD")
        newq.append(A.ingoing_row_state(int(U.indices[U.indptr[c]])))
    newq = np.array(newq, dtype=np.int64).reshape(-1, len(A.rules))
    left = A.indices[:A.numrows] + [MPXIndex(False, newq)]
    right = [MPXIndex(True, newq)] + A.indices[A.numrows:]
    Um = MPX(A.rules, left, A.numrows, U)
    Vm = MPX(A.rules, right, 1, r["Vd"])
    return dict(U=Um, values=r["values"], Vd=Vm, truncation=r["discarded"],
                kept=r["kept"], order=r["order"])


def diagonal_mpx(rules, lindex: MPXIndex, rindex: MPXIndex, values, inverse=False):
    """Diagonal constructor MPX_matrix(spectrum, Lindex, Rindex, values)
    (MPX_matrix.cpp:44-54): indices (in Lindex, out Rindex), 1 row index."""
    v = np.asarray(values, dtype=np.complex128)
    if inverse:
        v = 1.0 / v
    n = len(v)
    m = sp.csc_matrix((v, (np.arange(n), np.arange(n))), shape=(n, n))
    return MPX(rules, [lindex.with_direction(True), rindex.with_direction(False)], 1, m)

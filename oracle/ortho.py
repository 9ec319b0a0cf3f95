"""Orthogonalisation of a two-vertex unit cell, its transfer-matrix
eigensolves and its energy per vertex (oracle; test infrastructure only).

Restates:
  * UnitCell(MPSDecomposition, PreviousLambda) (UnitCell.hpp:42-48): the cell
    [U, reorder(diag(values) . Vd . diag(previous)^-1, 102)] with lambdas
    [previous, values];
  * TransferMatrixComponents (measurement.cpp:405-449): the left boundary
    space is (bra inward index as in, ket inward index as out), the right
    one (bra outward index as out, ket outward index as in); the solve runs
    on the entries (r, c) of the reshaped vector whose two charges agree
    (target = identity);
  * LeftComponentsMultiply / RightComponentsMultiply (582-629):
    v -> sum_s A_s^H v A_s over the cell, and v -> sum_s conj(A_s) v A_s^T
    backwards over the cell;
  * LeftED / RightED (451-580): ARPACK znaupd, nev = 1 (NUMEVALS, line 13),
    which = "LM", ncv = 20, tol = 0 (arpack_interface.cpp:28-39); the
    leading vector is completed to a Hermitian matrix by adding the
    conjugates at the transposed positions (entries summed on finalise,
    512-529 and 553-562; sparse_interface.cpp:430-444);
  * SqrtDR (MPX_matrix.cpp:829-913): block Hermitian eigensolve, eigenvalues
    with |w| <= 100 eps max|w| dropped with their vectors, mixed signs
    rejected, returns sqrt|w| and R = V^H (rows on a deduced index);
  * OrthogonaliseInversionSymmetric (measurement.cpp:131-178): the left
    fixed point only, reused on the right boundary, SVD of X.diag(l0).Y;
  * Orthogonalise (measurement.cpp:180-246): X = diag(sqrt w_L) R_L,
    X^-1 = R_L^H diag(1/sqrt w_L), Y = R_R^H diag(sqrt w_R), SVD of X.Y,
    singular values rescaled to unit square sum, A0' = U^H X A0 / sqrt(l_L),
    A1' = A1 X^-1 U, lambdas [values, previous second lambda];
  * iTwoVertexEnergy (measurement.cpp:378-397) with OneVertexMeasurement
    (317-324) and TwoVertexMeasurement (310-315): 0.5 (on-vertex terms of
    both vertices + the bond inside the cell + the bond to the next cell),
    every term closed with Lambda0^2 on the right.

The right fixed point enters transposed: RightComponentsMultiply maps
v -> sum_s conj(A_s) v A_s^T, whose fixed point M_R is the transpose of the
textbook sum_s A_s R A_s^H = l R, and Orthogonalise factors Y Y^H = M_R.  For
real cells (every iDMRG cell of a real Hamiltonian) M_R = R and the new lambdas
are the Schmidt values; for complex cells they are not, in the reference and
in this restatement alike.

Values the reference leaves open, fixed here (and in the library) as:
  * ARPACK's eigenvector phase is arbitrary, so the completed matrix
    v + v^H carries a factor 2 cos(phase).  The vector is rotated so that its
    trace is real and positive before the completion.  Everything compared
    in the tests (the leading eigenvalues, the new lambdas, the entropy,
    energies, the canonical conditions) is independent of that factor.
  * For lengths <= 1000 the reference takes the accumulated transfer matrix
    (measurement.cpp:458-466) without the Hermitian completion; both routes
    have the same fixed point, and every length takes the completion here.
  * The ARPACK start vector is the identity on the boundary space (the fixed
    point of a canonical cell) instead of ARPACK's random vector.
  * Entries below machine epsilon are kept (the reference drops them).

Parity pins (tests/test_oracle_ortho.py): the matvecs against an explicit
dense transfer matrix; the leading eigenvalue against numpy.linalg.eig of it;
the new lambdas of real cells against the Schmidt values from the textbook
formula eig(sqrt(l) r sqrt(l)) of the dense fixed points; the canonical
conditions of the output cell; the energy against two-site expectation values
of the full bond Hamiltonians on a four-vertex reduced density matrix, and the
Heisenberg energy against 1/4 - ln 2 to the accuracy of the bond dimension;
SqrtDR's sign rule on hand-made diagonal matrices.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .mpx import MPX, contract, contract_to_sparse, diagonal_mpx, mpx_svd, reorder
from .qn import MPXIndex
from .dmrg import extract_mpo_block

EPS = np.finfo(float).eps
IMAGTOL = 100.0 * EPS          # measurement.cpp:12
NCV = 20                       # arpack_interface.cpp:39 with nev = 1


class OrthogonaliseError(RuntimeError):
    """Orthogonalise returned an empty cell (complex leading eigenvalue or
    SqrtDR failure; measurement.cpp:187-226)."""


def unit_cell_from_decomposition(U: MPX, values, Vd: MPX, previous):
    """UnitCell.hpp:42-48."""
    rules = U.rules
    lam = diagonal_mpx(rules, Vd.indices[0], Vd.indices[0], values)
    inv_prev = diagonal_mpx(rules, Vd.indices[2], Vd.indices[2], previous, inverse=True)
    A1 = reorder(contract(contract(lam, False, Vd, False, [(1, 0)]), False, inv_prev, False, [(2, 0)]),
                 False, [1, 0, 2], 2)
    return [U.copy(), A1], [np.asarray(previous, float).copy(), np.asarray(values, float).copy()]


def _dense(A: MPX):
    return A.to_dense_tensor()          # [sigma, a_in, a_out]


def boundary_charges(mats, side):
    return mats[0].indices[1].charges if side == 0 else mats[-1].indices[2].charges


def allowed_entries(q):
    """(r, c) with equal charges, loop order c outer, r inner
    (measurement.cpp:431-442); flat position r + n c."""
    q = np.asarray(q)
    n = q.shape[0]
    same = np.all(q[:, None, :] == q[None, :, :], axis=2)      # same[r, c]
    rr, cc = np.nonzero(same.T)                                  # rows of .T = c
    return cc + n * rr, (cc, rr)                                 # (flat, (r, c))


def left_apply(dense_cell, v):
    """LeftComponentsMultiply (measurement.cpp:582-602) on the full n x n
    matrix v (rows bra, columns ket)."""
    for A in dense_cell:
        v = np.einsum("sab,ac,scd->bd", A.conj(), v, A)
    return v


def right_apply(dense_cell, v):
    """RightComponentsMultiply (measurement.cpp:603-629)."""
    for A in reversed(dense_cell):
        v = np.einsum("sbr,rc,sdc->bd", A.conj(), v, A)
    return v


def mixed_left_apply(bra_cell, ket_cell, v):
    """LeftComponentsMultiply with Bra != Ket (measurement.cpp:582-602):
    v -> sum_s B_s^H v A_s over the cells, v rows on the bra index."""
    for B, A in zip(bra_cell, ket_cell):
        v = np.einsum("sab,ac,scd->bd", B.conj(), v, A)
    return v


def overlap(bra_mats, ket_mats, ncv=NCV):
    """Overlap(bra, ket) (measurement.cpp:282-288): the largest-magnitude
    eigenvalue of TransferMatrixComponents(bra, ket, 0, State(rules)), i.e.
    of the left map on the entries (r, c) of (bra inward) x (ket inward)
    with equal charges (no Hermitian completion)."""
    B = [_dense(m) for m in bra_mats]
    A = [_dense(m) for m in ket_mats]
    qb, qk = bra_mats[0].indices[1].charges, ket_mats[0].indices[1].charges
    nb, nk = qb.shape[0], qk.shape[0]
    if nb * nk != bra_mats[-1].indices[2].size * ket_mats[-1].indices[2].size:
        raise ValueError("Index dimensions mismatch between ends of translation unit!")
    same = np.all(qb[:, None, :] == qk[None, :, :], axis=2)         # same[r, c]
    cc, rr = np.nonzero(same.T)                                      # c outer, r inner
    m = len(rr)

    def mv(x):
        V = np.zeros((nb, nk), complex)
        V[rr, cc] = x
        return mixed_left_apply(B, A, V)[rr, cc]

    if m <= ncv + 1:
        T = np.stack([mv(e) for e in np.eye(m, dtype=complex)], axis=1)
        w = np.linalg.eigvals(T)
        return complex(w[int(np.argmax(np.abs(w)))])
    op = spla.LinearOperator((m, m), matvec=mv, dtype=np.complex128)
    w = spla.eigs(op, k=1, which="LM", ncv=min(ncv, m - 1), tol=0.0, return_eigenvectors=False)
    return complex(w[0])


def transfer_leading(mats, side, ncv=NCV, tol=0.0):
    """LeftED / RightED with NUMEVALS = 1, LARGESTMAGNITUDE
    (measurement.cpp:451-580).  Returns (eigenvalue, completed Hermitian
    matrix M with rows bra, columns ket, allowed flat positions)."""
    dense_cell = [_dense(A) for A in mats]
    q = boundary_charges(mats, side)
    n = q.shape[0]
    if n != (mats[-1].indices[2].size if side == 0 else mats[0].indices[1].size):
        raise ValueError("Index dimensions mismatch between ends of translation unit!")
    flat, (r, c) = allowed_entries(q)
    m = len(flat)
    app = left_apply if side == 0 else right_apply

    def mv(x):
        V = np.zeros((n, n), complex)
        V[r, c] = x
        return app(dense_cell, V)[r, c]

    start = (r == c).astype(complex)
    if m <= ncv + 1:
        T = np.stack([mv(e) for e in np.eye(m, dtype=complex)], axis=1)
        w, vecs = np.linalg.eig(T)
        k = int(np.argmax(np.abs(w)))
        lam, x = complex(w[k]), vecs[:, k]
    else:
        op = spla.LinearOperator((m, m), matvec=mv, dtype=np.complex128)
        w, vecs = spla.eigs(op, k=1, which="LM", ncv=min(ncv, m - 1), tol=tol, maxiter=50 * m, v0=start)
        lam, x = complex(w[0]), vecs[:, 0]
    x = x / np.linalg.norm(x)
    tr = np.sum(x[r == c])
    if abs(tr) > 0:
        x = x * (abs(tr) / tr)
    V = np.zeros((n, n), complex)
    V[r, c] = x
    return lam, V + V.conj().T, flat


def sqrt_dr(M: MPX):
    """SqrtDR (MPX_matrix.cpp:829-913): returns (roots, R) with R = V^H on
    the rows (deduced index, inward) x columns (M's first index, outward);
    None on the mixed-sign failure."""
    dense = M.matrix.toarray()
    q = M.indices[0].charges
    vals, vecs_cols, ded = [], [], []
    for state, cols in M.get_all_block_columns():
        blk = dense[np.ix_(cols, cols)]
        w, v = np.linalg.eigh(0.5 * (blk + blk.conj().T))
        for j in range(len(w)):
            col = np.zeros(dense.shape[0], complex)
            col[cols] = v[:, j]
            vals.append(float(w[j]))
            vecs_cols.append(col)
            ded.append(q[cols[int(np.argmax(np.abs(v[:, j])))]])
    vals = np.asarray(vals)
    tolerance = np.max(np.abs(vals)) * 100.0 * EPS
    kept = np.nonzero(np.abs(vals) > tolerance)[0]
    neg = int(np.sum(vals[kept] < 0.0))
    if neg != 0 and neg != len(kept):
        return None
    roots = np.sqrt(np.abs(vals[kept]))
    V = np.stack([vecs_cols[k] for k in kept], axis=1)
    R = sp.csc_matrix(V.conj().T)
    idx = [MPXIndex(True, np.asarray([ded[k] for k in kept]).reshape(len(kept), -1)),
           MPXIndex(False, q)]
    return roots, MPX(M.rules, idx, 1, R)


def orthogonalise(mats, lambdas, ncv=NCV):
    """Orthogonalise (measurement.cpp:180-246) of a two-vertex cell.
    Returns (new matrices, new lambdas, info dict).  ``ncv`` is the Krylov
    dimension of the fixed-point eigensolves (the reference's 20; a larger
    one reaches the same fixed point in fewer restarts)."""
    A0, A1 = mats
    rules = A0.rules
    lamL, ML, _ = transfer_leading(mats, 0, ncv)
    if abs(lamL.imag) >= IMAGTOL * abs(lamL.real):
        raise OrthogonaliseError("Eigenvalue has non negligible imaginary part")
    q0 = A0.indices[1].charges
    VL = MPX(rules, [MPXIndex(True, q0), MPXIndex(False, q0)], 1, sp.csc_matrix(ML))
    dl = sqrt_dr(VL)
    if dl is None:
        raise OrthogonaliseError("SqrtDR failure")
    rootsL, RL = dl
    X = contract(diagonal_mpx(rules, RL.indices[0], RL.indices[0], rootsL), False, RL, False, [(1, 0)])
    Xinv = contract(reorder(RL, True, [1, 0], 1), False,
                    diagonal_mpx(rules, RL.indices[0], RL.indices[0], rootsL, inverse=True), False, [(1, 0)])
    lamR, MR, _ = transfer_leading(mats, 1, ncv)
    if abs(lamR.imag) >= IMAGTOL * abs(lamR.real):
        raise OrthogonaliseError("Eigenvalue has non negligible imaginary part")
    q1 = A1.indices[2].charges
    VR = MPX(rules, [MPXIndex(True, q1), MPXIndex(False, q1)], 1, sp.csc_matrix(MR))
    dr = sqrt_dr(VR)
    if dr is None:
        raise OrthogonaliseError("VRDecomp failure")
    rootsR, RR = dr
    LY = contract(reorder(RR, True, [1, 0], 1), False,
                  diagonal_mpx(rules, RR.indices[0], RR.indices[0], rootsR), False, [(1, 0)])
    XLY = contract(X, False, LY, False, [(1, 0)])
    return _finish(mats, lambdas, X, Xinv, XLY, lamL, dict(left_value=lamL, right_value=lamR,
                                                           kept_left=len(rootsL), kept_right=len(rootsR)))


def _finish(mats, lambdas, X, Xinv, XLY, lamL, info):
    """The common end of Orthogonalise and OrthogonaliseInversionSymmetric
    (measurement.cpp:166-177, 229-245): SVD of X.(lambda).Y, values rescaled
    to unit square sum, A0' = U^H X A0 / sqrt(l), A1' = A1 X^-1 U."""
    A0, A1 = mats
    rules = A0.rules
    svd = mpx_svd(XLY)
    vals = np.asarray(svd["values"], float)
    vals = vals * np.sqrt(1.0 / float(np.sum(vals * vals)))
    s = 1.0 / np.sqrt(lamL.real)
    sX = MPX(rules, X.indices, 1, X.matrix * s)
    U = svd["U"]
    A0n = reorder(contract(contract(U, True, sX, False, [(0, 0)]), False, A0, False, [(1, 1)]), False, [1, 0, 2], 2)
    A1n = contract(A1, False, contract(Xinv, False, U, False, [(1, 0)]), False, [(2, 0)])
    nz = vals[vals > 0]
    info = dict(info, entropy=float(-np.sum(nz * nz * np.log(nz * nz))))
    return [A0n, A1n], [vals, np.asarray(lambdas[1], float).copy()], info


def orthogonalise_inversion_symmetric(mats, lambdas):
    """OrthogonaliseInversionSymmetric (measurement.cpp:131-178): only the
    left fixed point is solved; under inversion symmetry it serves as the
    right one on the outward index of the last matrix, and the boundary
    lambda enters between the factors: SVD of X . diag(lambda0) . Y."""
    A0, A1 = mats
    rules = A0.rules
    lamL, ML, _ = transfer_leading(mats, 0)
    if abs(lamL.imag) >= IMAGTOL * abs(lamL.real):
        raise OrthogonaliseError("Eigenvalue has non negligible imaginary part")
    q0 = A0.indices[1].charges
    dl = sqrt_dr(MPX(rules, [MPXIndex(True, q0), MPXIndex(False, q0)], 1, sp.csc_matrix(ML)))
    if dl is None:
        raise OrthogonaliseError("SqrtDR failure")
    rootsL, RL = dl
    X = contract(diagonal_mpx(rules, RL.indices[0], RL.indices[0], rootsL), False, RL, False, [(1, 0)])
    Xinv = contract(reorder(RL, True, [1, 0], 1), False,
                    diagonal_mpx(rules, RL.indices[0], RL.indices[0], rootsL, inverse=True), False, [(1, 0)])
    q1 = A1.indices[2].charges
    dr = sqrt_dr(MPX(rules, [MPXIndex(True, q1), MPXIndex(False, q1)], 1, sp.csc_matrix(ML)))
    if dr is None:
        raise OrthogonaliseError("SqrtDR failure")
    rootsR, RR = dr
    Y = contract(reorder(RR, True, [1, 0], 1), False,
                 diagonal_mpx(rules, RR.indices[0], RR.indices[0], rootsR), False, [(1, 0)])
    L0 = diagonal_mpx(rules, A1.indices[2], A1.indices[2], lambdas[0])
    XLY = contract(contract(X, False, L0, False, [(1, 0)]), False, Y, False, [(1, 0)])
    return _finish(mats, lambdas, X, Xinv, XLY, lamL, dict(left_value=lamL, right_value=lamL,
                                                           kept_left=len(rootsL), kept_right=len(rootsR)))


# ---------------------------------------------------------------------------
# energy per vertex
# ---------------------------------------------------------------------------

def _trace(m):
    return complex(m.diagonal().sum())


def _lambda_sq(A0, lam0):
    L = diagonal_mpx(A0.rules, A0.indices[1], A0.indices[1], lam0)
    return contract(L, False, L, False, [(1, 0)])


def one_vertex_measurement(W, mats, lambdas):
    """OneVertexMeasurement (measurement.cpp:317-324): W on the second vertex."""
    A0, A1 = mats
    L2 = _lambda_sq(A0, lambdas[0])
    t = contract(A0, True, A0, False, [(0, 0), (1, 1)])
    t = contract(t, False, A1, False, [(1, 1)])
    t = contract(t, False, W, False, [(1, 2)])
    t = contract(A1, True, t, False, [(0, 2), (1, 0)])
    return _trace(contract_to_sparse(t, False, L2, False, [(0, 1), (1, 0)]))


def two_vertex_measurement(W1, W2, A1, A2, L2):
    """TwoVertexMeasurement (measurement.cpp:310-315): W1 on A1, W2 on A2."""
    t = contract(A1, True, A1, False, [(1, 1)])
    t = contract(t, False, W1, False, [(0, 0), (2, 2)])
    t = contract(t, False, A2, True, [(0, 1)])
    t = contract(t, False, W2, False, [(2, 1), (3, 0)])
    t = contract(t, False, A2, False, [(0, 1), (3, 0)])
    return _trace(contract_to_sparse(t, False, L2, False, [(1, 1), (3, 0)]))


def mpo_pieces(W):
    """H1, ColX, RowX as iDMRG_DRV.cpp:50-55 / measurement.cpp:371-375."""
    w = W.indices[1].size
    H1 = extract_mpo_block(W, (w - 1, w - 1), (0, 0))
    ColX = extract_mpo_block(W, (1, w - 2), (0, 0))
    RowX = extract_mpo_block(W, (w - 1, w - 1), (1, W.indices[3].size - 2))
    return H1, ColX, RowX


def i_two_vertex_energy(W, mats, lambdas):
    """iTwoVertexEnergy (measurement.cpp:378-397).  Returns the complex
    energy per vertex and its four terms."""
    H1, ColX, RowX = mpo_pieces(W)
    A0, A1 = mats
    L2 = _lambda_sq(A0, lambdas[0])
    local = 0.0 + 0.0j
    if H1.matrix.count_nonzero():
        t = contract(H1, False, A0, False, [(2, 0)])
        t = contract(A0, True, t, False, [(0, 0), (1, 3)])
        t = contract(t, False, A1, False, [(3, 1)])
        t = contract(A1, True, t, False, [(0, 3), (1, 0)])
        local += _trace(contract_to_sparse(t, False, L2, False, [(0, 1), (3, 0)]))
        local += one_vertex_measurement(H1, mats, lambdas)
    bond1 = two_vertex_measurement(RowX, ColX, A0, A1, L2)
    t = contract(A0, True, A0, False, [(0, 0), (1, 1)])
    t = contract(t, False, A1, False, [(1, 1)])
    t = contract(t, False, RowX, False, [(1, 2)])
    t = contract(A1, True, t, False, [(0, 2), (1, 0)])
    t = contract(A0, False, t, False, [(1, 1)])
    t = contract(t, False, ColX, False, [(0, 2), (4, 1)])
    t = contract(A0, True, t, False, [(0, 3), (1, 1)])
    t = contract(t, False, A1, False, [(1, 1)])
    t = contract(A1, True, t, False, [(0, 3), (1, 0)])
    bond2 = _trace(contract_to_sparse(t, False, L2, False, [(0, 1), (3, 0)]))
    return 0.5 * (local + bond1 + bond2), dict(local=local, bond1=bond1, bond2=bond2)

"""iTEBD restatement (oracle; test infrastructure only).

Follows TEBD_routines.cpp:
  * ``MakeOddBondHamiltonian`` / ``MakeEvenBondHamiltonian`` (826-867):
    odd bond = row w-1 of W (all columns) times column 0 (all rows), even
    bond = row w-1 / columns 1..w-2 times rows 1..w-2 / column 0, each
    contracted on (3,1) and reordered 032415 with two row indices;
  * ``MakeBondEvolutionOperator`` (869-880): exp(-i tau H_bond) from the
    blockwise Hermitian eigendecomposition (Exponentiate,
    sparse_interface.cpp:1707-1714), keeping the first four indices;
  * ``TrotterDecomposition`` order 2 (41-51): (odd tau/2, even tau, odd tau);
  * ``iTEBD::apply_and_decompose`` (233-266) and the unit-cell swap
    (UnitCell.hpp:49).

Parity pins (tests/test_oracle_tebd.py): exp(-i tau H) against
scipy.linalg.expm of the dense bond Hamiltonian; unitarity; an exact
two-site evolution without truncation (norm conservation, energy
conservation of the single bond); the gauge-invariant two-site tensor.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .qn import MPXIndex
from .mpx import MPX, contract, reorder, mpx_svd, diagonal_mpx
from .dmrg import extract_mpo_block


def odd_bond_hamiltonian(W: MPX) -> MPX:
    w = W.indices[1].size
    left = extract_mpo_block(W, (w - 1, w - 1), (0, W.indices[3].size - 1))
    right = extract_mpo_block(W, (0, w - 1), (0, 0))
    return reorder(contract(left, False, right, False, [(3, 1)]), False, [0, 3, 2, 4, 1, 5], 2)


def even_bond_hamiltonian(W: MPX) -> MPX:
    w = W.indices[1].size
    left = extract_mpo_block(W, (w - 1, w - 1), (1, W.indices[3].size - 2))
    right = extract_mpo_block(W, (1, w - 2), (0, 0))
    return reorder(contract(left, False, right, False, [(3, 1)]), False, [0, 3, 2, 4, 1, 5], 2)


def bond_eigensystem(H: MPX):
    """Per charge block of the bond Hamiltonian: (columns, eigenvalues,
    eigenvectors) of its Hermitian part (sparse_interface.cpp:1414-1441)."""
    m = H.matrix.toarray()
    out = []
    for _, cols in H.get_all_block_columns():
        blk = m[np.ix_(cols, cols)]
        e, v = np.linalg.eigh(0.5 * (blk + blk.conj().T))
        out.append((cols, e, v))
    return out


def bond_evolution_operator(H: MPX, tau: float, eig=None) -> MPX:
    """exp(-i tau H) per charge block of the (d^2 x d^2) bond Hamiltonian,
    indices = the first four of H (Exponentiate, sparse_interface.cpp:
    1707-1714); ``eig`` reuses a bond_eigensystem of H."""
    eig = bond_eigensystem(H) if eig is None else eig
    n = H.matrix.shape[0]
    U = np.zeros((n, n), np.complex128)
    for cols, e, v in eig:
        U[np.ix_(cols, cols)] = (v * np.exp(-1j * tau * e)) @ v.conj().T
    s = sp.csc_matrix(U)
    s.eliminate_zeros()
    return MPX(H.rules, H.indices[:4], 2, s)


def apply_and_decompose(G: MPX, A0: MPX, A1: MPX, lambda0, D=0, minS=0.0):
    """Returns (new A0, new A1, new lambda0, truncation) after the unit-cell
    swap; the other bond weight becomes the old lambda0."""
    C = reorder(contract(G, False, contract(A0, False, A1, False, [(2, 1)]), False, [(2, 0), (3, 2)]),
                False, [0, 2, 1, 3], 2)
    lam = diagonal_mpx(C.rules, C.indices[3], C.indices[3], lambda0)
    CL = contract(C, False, lam, False, [(3, 0)]).shift_num_row_indices(2)
    d = mpx_svd(CL, D, minS)
    U = d["U"]
    A1p = reorder(contract(U, True, C, False, [(0, 0), (1, 1)]), False, [1, 0, 2], 2)
    A1p.rescale(1.0 / np.sqrt(np.sum(np.asarray(lambda0) ** 2)))
    return A1p, U, np.asarray(d["values"]), d["truncation"]


def product_state_cell(rules, phys, level=0):
    """Two-site unit cell of a product state in basis level `level`
    (MakeProductStateUnitCell, UnitCell.cpp:190-210): A[s, a_l, a_r] with
    1x1 bonds; the left bond carries charge 0, the right bond the level's
    charge, repeated with accumulated charge on the second site."""
    phys = np.asarray(phys, dtype=np.int64).reshape(-1, len(rules))
    q = phys[level]
    z = rules.zero(1)
    q1 = q.reshape(1, -1)
    q2 = rules.add(q, q).reshape(1, -1)
    d = phys.shape[0]

    def site(ql, qr):
        idx = [MPXIndex(True, phys, True), MPXIndex(True, ql), MPXIndex(False, qr)]
        T = np.zeros((d, 1, 1), complex)
        T[level, 0, 0] = 1.0
        return MPX.from_dense_tensor(rules, idx, 2, T)
    return site(z, q1), site(q1, q2), np.array([1.0]), np.array([1.0])


def trotter_operators(W: MPX, tau: float, order: int):
    """TrotterDecomposition (TEBD_routines.cpp:14-85): OrderedOperatorPtrs
    of order 1 (odd tau, even tau), order 2 (odd tau/2, even tau, odd tau)
    and order 4 (the 11 pieces of Suzuki's product)."""
    Hodd, Heven = odd_bond_hamiltonian(W), even_bond_hamiltonian(W)
    if order == 1:
        return [bond_evolution_operator(Hodd, tau), bond_evolution_operator(Heven, tau)]
    if order == 2:
        eo = bond_eigensystem(Hodd)
        return [bond_evolution_operator(Hodd, 0.5 * tau, eo), bond_evolution_operator(Heven, tau),
                bond_evolution_operator(Hodd, tau, eo)]
    if order == 4:
        from .ftebd import trotter_operators as pieces
        ops, seq = pieces(W, tau, 4)
        return [ops[k] for k in seq]
    raise ValueError("Only 1st, 2nd and 4th order decompositions are supported")


def evolve_measured(ops, order, cell, steps, D, minS, interval, measured=(), initial=None):
    """iTEBD::evolve (TEBD_routines.cpp:120-231) with measurement steps every
    `interval` steps (m_current_time_step counted from 1 in this call; the
    time step tau enters through ``ops``).  At a measurement step the cell is
    canonised with OrthogonaliseInversionSymmetric (measurement.cpp:131-178)
    and do_measurements (99-118) records {step, truncation of the last gate,
    entropy of the new lambda0, |overlap|, overlap with the initial cell,
    OneVertexMeasurement of each measured operator}.  Order 2 follows the
    reference's measuring branch (odd(tau/2), swap, canonise, odd(tau/2)) and
    uses the odd full step [2] between steps elsewhere (see evolve)."""
    from .ortho import orthogonalise_inversion_symmetric, overlap, one_vertex_measurement
    A0, A1, l0, l1 = cell
    init = [A0, A1] if initial is None else initial
    tr = 0.0
    records = []

    def gate(G):
        nonlocal A0, A1, l0, l1, tr
        nA0, nA1, nl0, tr = apply_and_decompose(G, A0, A1, l0, D, minS)
        A0, A1, l0, l1 = nA0, nA1, nl0, l0

    def canonise():
        nonlocal A0, A1, l0, l1
        (A0, A1), (l0, l1), _ = orthogonalise_inversion_symmetric([A0, A1], [l0, l1])

    def measure(step):
        ov = overlap(init, [A0, A1])
        v = np.asarray(l0, float)
        nz = v[v > 0]
        rec = dict(step=step, truncation=tr, entropy=float(-np.sum(nz * nz * np.log(nz * nz))),
                   abs_overlap=abs(ov), overlap=ov,
                   values=[one_vertex_measurement(op, [A0, A1], [l0, l1]) for op in measured])
        records.append(rec)

    def swap():
        nonlocal A0, A1, l0, l1
        A0, A1, l0, l1 = A1, A0, l1, l0

    due = lambda n: interval > 0 and n % interval == 0
    if order == 1:
        for n in range(1, steps + 1):
            gate(ops[0])
            gate(ops[1])
            if due(n):
                canonise()
                measure(n)
    elif order == 2:
        if steps > 0:
            gate(ops[0])
            for n in range(1, steps):
                gate(ops[1])
                if due(n):
                    gate(ops[0])
                    swap()
                    canonise()
                    measure(n)
                    gate(ops[0])
                else:
                    gate(ops[2])
            gate(ops[1])
            gate(ops[0])
            swap()
            if due(steps):
                canonise()
                measure(steps)
    elif order == 4:
        gate(ops[0])
        for n in range(1, steps + 1):
            for i in range(1, 10):
                gate(ops[i])
                if i % 2:
                    canonise()
            if due(n):
                gate(ops[10])
                swap()
                canonise()
                measure(n)
                gate(ops[0])
            else:
                gate(ops[2])
    else:
        raise ValueError("order 1, 2 or 4")
    return [A0, A1, l0, l1], tr, records


def evolve(ops, order, cell, steps, D=0, minS=0.0):
    """iTEBD::evolve (TEBD_routines.cpp:120-224) for orders 1, 2 and 4 with no
    measurement step inside the run.  ``cell`` = [A0, A1, lambda0, lambda1];
    returns the new cell and the truncation of the last gate.

    Second order: odd(tau/2), then per step even(tau) and odd(tau), closing
    with even(tau), odd(tau/2) and the cell swap.  The reference's
    no-measurement branch applies OrderedOperatorPtrs[1] (even) where the
    odd full-step operator [2] belongs (TEBD_routines.cpp:163); its
    measuring branch applies odd(tau/2) twice around the canonisation, which
    is the same odd(tau) up to truncation.  This restatement uses [2]."""
    A0, A1, l0, l1 = cell
    tr = 0.0

    def gate(G):
        nonlocal A0, A1, l0, l1, tr
        nA0, nA1, nl0, tr = apply_and_decompose(G, A0, A1, l0, D, minS)
        A0, A1, l0, l1 = nA0, nA1, nl0, l0

    if order == 1:
        for _ in range(steps):
            gate(ops[0])
            gate(ops[1])
    elif order == 2:
        if steps > 0:
            gate(ops[0])
            for _ in range(steps - 1):
                gate(ops[1])
                gate(ops[2])
            gate(ops[1])
            gate(ops[0])
            A0, A1, l0, l1 = A1, A0, l1, l0         # m_unit.swap(0,1)
    elif order == 4:
        # TEBD_routines.cpp:194-224 without measurement steps: the special
        # start, pieces 1..9 re-canonised after every even-bond piece, and
        # the odd full step that joins this step's last piece to the next
        # step's first
        from .ortho import orthogonalise_inversion_symmetric
        gate(ops[0])
        for _ in range(steps):
            for i in range(1, 10):
                gate(ops[i])
                if i % 2:
                    (A0, A1), (l0, l1), _ = orthogonalise_inversion_symmetric([A0, A1], [l0, l1])
            gate(ops[2])
    else:
        raise ValueError("order 1, 2 or 4")
    return [A0, A1, l0, l1], tr

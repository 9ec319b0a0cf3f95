"""Finite-chain TEBD restatement (oracle; test infrastructure only).

Follows TEBD_routines.cpp with the disk-resident MPS replaced by lists of
site tensors (site v = 1..N; a "left" tensor has the shape [s, a_l, a_r]
with two row indices, a "right" tensor [a_l, s, a_r] with one, the
MPS_matrix left_shape / right_shape of MPS_matrix.cpp:111-135):
  * ``TEBD::apply_to_odd_bonds`` (273-318): from the right end, bond
    (v-1, v) for v = N, N-2, ...: theta = reorder(contract(G,
    contract(L[v-1], R, 21), 2032), 0213, rows 2), SVD(D, minS); site v
    becomes V^dagger; (U.S) is rotated to right form by an untruncated SVD
    whose row matrix becomes site v-1 and whose column matrix times its
    values is pushed into L[v-2]; at v = 2 the 1x1 column matrix's trace
    (a phase) rescales site 1;
  * ``TEBD::apply_to_even_bonds`` (320-391): the same from bond (N-2, N-1)
    after an untruncated rotation of site N, then a final rotation of
    site 1 with its phase;
  * ``TEBD::left_canonise`` / ``left_canonise_measure`` (395-547): the
    sweep from site 1 with SVD(chi, minS) and SquareRescale(1), the
    entropy of the bond at v = N/2, the phase of the last 1x1 Vd, and the
    overlap with the initial state;
  * ``TEBD::evolve`` (672-791) for Trotter orders 1, 2 and 4.
``trotter_operators`` extends the iTEBD restatement with order 4
(TEBD_routines.cpp:55-81): Suzuki's fourth-order product with
p = 1 / (4 - 4^(1/3)), whose values the reference writes out as f1 = p,
f3 = 1 - 4p and f1_3 = 1 - 3p.

Parity pins (tests/test_oracle_ftebd.py): without truncation the evolved
state equals the dense state vector under the same products of bond
exponentials (scipy.linalg.expm of the bond Hamiltonian), norm is kept,
every stored site is left- (or right-) canonical, the fourth-order
coefficients satisfy 4p + (1 - 4p) = 1 and 4p^3 + (1 - 4p)^3 = 0, and a
fourth-order step approaches exp(-i tau H) faster than a second-order one.
"""
from __future__ import annotations

import numpy as np

from .mpx import MPX, contract, reorder, mpx_svd, diagonal_mpx
from .qn import MPXIndex
from . import tebd as itebd
from .dmrg import entropy as _entropy


def suzuki4_coefficients():
    """p, 1 - 4p, 1 - 3p with p = 1 / (4 - 4^(1/3)) (Suzuki 1990): the
    f1, f3, f1_3 of TrotterDecomposition (TEBD_routines.cpp:57-59)."""
    p = 1.0 / (4.0 - 4.0 ** (1.0 / 3.0))
    return p, 1.0 - 4.0 * p, 1.0 - 3.0 * p


def trotter_operators(W: MPX, tau: float, order: int):
    """TrotterDecomposition (TEBD_routines.cpp:14-85): the bond operators
    and the order in which one time step applies them (odd bonds at even
    positions of the sequence)."""
    Hodd, Heven = itebd.odd_bond_hamiltonian(W), itebd.even_bond_hamiltonian(W)
    G = itebd.bond_evolution_operator
    if order == 1:
        return [G(Hodd, tau), G(Heven, tau)], [0, 1]
    if order == 2:
        return [G(Hodd, 0.5 * tau), G(Heven, tau), G(Hodd, tau)], [0, 1, 2]
    if order == 4:
        f1, f3, f13 = suzuki4_coefficients()
        ops = [G(Hodd, 0.5 * tau * f1), G(Heven, tau * f1), G(Hodd, tau * f1), G(Hodd, 0.5 * tau * f13),
               G(Heven, tau * f3)]
        return ops, [0, 1, 2, 1, 3, 4, 3, 1, 2, 1, 0]
    raise ValueError("Only 1st, 2nd and 4th order decompositions are supported")


def right_shape(A: MPX) -> MPX:
    """MPS_matrix::right_shape (MPS_matrix.cpp:124-135) of a left-shaped
    site: [s, a_l, a_r] -> [a_l, s, a_r], one row index."""
    return reorder(A, False, [1, 0, 2], 1)


def left_shape(A: MPX) -> MPX:
    """MPS_matrix::left_shape (MPS_matrix.cpp:111-122) of a right-shaped
    site: [a_l, s, a_r] -> [s, a_l, a_r], two row indices."""
    return reorder(A, False, [1, 0, 2], 2)


def _diag(rules, index: MPXIndex, values):
    """MPX_matrix(basis, index, values) (MPX_matrix.cpp:58-63)."""
    return diagonal_mpx(rules, index, index, values)


def _trace(A: MPX):
    return complex(A.matrix.diagonal().sum())


def _theta(G, L, R):
    return reorder(contract(G, False, contract(L, False, R, False, [(2, 1)]), False, [(2, 0), (3, 2)]),
                   False, [0, 2, 1, 3], 2)


def _times_values(A: MPX, k: int, values):
    d = diagonal_mpx(A.rules, A.indices[k], A.indices[k], values)
    return contract(A, False, d, False, [(k, 0)])


class FiniteTEBD:
    """``left[v]`` / ``right[v]`` (v = 1..N) hold the site tensors; the
    state starts left canonical (FiniteMPS::makeLC) and is left canonical
    after every evolve step."""

    def __init__(self, W: MPX, sites, dt: float, order: int):
        self.N = len(sites)
        if self.N < 2 or self.N % 2:
            raise ValueError("an even number of vertices >= 2")
        self.left = [None] + [s.copy() for s in sites]
        self.right = [None] * (self.N + 1)
        self.initial = [None] + [s.copy() for s in sites]
        self.ops, self.seq = trotter_operators(W, dt, order)
        self.order = order
        self.dt = dt
        self.step = 0
        self.max_truncation = 0.0
        self.records = []

    # -- TEBD_routines.cpp:273-318 -------------------------------------------
    def apply_to_odd_bonds(self, G, D, minS):
        N = self.N
        R = self.left[N]
        for v in range(N, 1, -2):
            d = mpx_svd(_theta(G, self.left[v - 1], R), D, minS)
            self.max_truncation = max(self.max_truncation, d["truncation"])
            self.right[v] = d["Vd"]
            rot = mpx_svd(right_shape(_times_values(d["U"], 2, d["values"])))
            self.max_truncation = max(self.max_truncation, rot["truncation"])
            if v == 2:
                rot["Vd"].rescale(_trace(rot["U"]))
            self.right[v - 1] = rot["Vd"]
            if v > 2:
                R = contract(self.left[v - 2], False, _times_values(rot["U"], 1, rot["values"]), False,
                             [(2, 0)])

    # -- TEBD_routines.cpp:320-391 -------------------------------------------
    def apply_to_even_bonds(self, G, D, minS):
        N = self.N
        rot = mpx_svd(right_shape(self.left[N]))
        self.right[N] = rot["Vd"]
        R = contract(self.left[N - 1], False, _times_values(rot["U"], 1, rot["values"]), False, [(2, 0)])
        if N > 2:
            for v in range(N - 1, 1, -2):
                d = mpx_svd(_theta(G, self.left[v - 1], R), D, minS)
                self.max_truncation = max(self.max_truncation, d["truncation"])
                self.right[v] = d["Vd"]
                rot = mpx_svd(right_shape(_times_values(d["U"], 2, d["values"])))
                self.max_truncation = max(self.max_truncation, rot["truncation"])
                self.right[v - 1] = rot["Vd"]
                if v > 2:
                    R = contract(self.left[v - 2], False, _times_values(rot["U"], 1, rot["values"]), False,
                                 [(2, 0)])
        d = mpx_svd(right_shape(R))
        d["Vd"].rescale(_trace(d["U"]))
        self.right[1] = d["Vd"]

    # -- TEBD_routines.cpp:395-547 -------------------------------------------
    def left_canonise(self, chi=0, minS=0.0, measure=False):
        N = self.N
        rec = {"step": self.step, "time": self.step * self.dt}
        first = mpx_svd(left_shape(self.right[1]), chi, minS)
        vals = _square_rescale(first)
        if measure and N == 2:
            rec["truncation"] = self.max_truncation
            rec["entropy"] = _entropy(vals)
        self.left[1] = first["U"]
        overlap = None
        if measure:
            overlap = contract(self.initial[1], True, first["U"], False, [(0, 0), (1, 1)])
        Vd = contract(_diag(first["Vd"].rules, first["Vd"].indices[0], vals), False, first["Vd"], False, [(1, 0)])
        for v in range(2, N + 1):
            X = contract(Vd, False, self.right[v], False, [(1, 0)])     # [k, s, a_r]
            d = mpx_svd(left_shape(X), chi, minS)
            vals = _square_rescale(d)
            if measure and v == N // 2:
                rec["truncation"] = self.max_truncation
                rec["entropy"] = _entropy(vals)
            Vd = contract(_diag(d["Vd"].rules, d["Vd"].indices[0], vals), False, d["Vd"], False, [(1, 0)])
            if v == N:
                d["U"].rescale(_trace(Vd))
            if measure:
                overlap = contract(self.initial[v], True, contract(overlap, False, d["U"], False, [(1, 1)]),
                                   False, [(0, 1), (1, 0)])
            self.left[v] = d["U"]
        if measure:
            ov = _trace(overlap)
            rec["overlap"] = ov
            rec["abs_overlap"] = abs(ov)
            self.records.append(rec)

    # -- TEBD_routines.cpp:672-791 -------------------------------------------
    def _bonds(self, k, D, minS):
        G = self.ops[self.seq[k]]
        if k % 2 == 0:
            self.apply_to_odd_bonds(G, D, minS)
        else:
            self.apply_to_even_bonds(G, D, minS)

    def evolve(self, steps, D=0, minS=0.0, interval=1):
        N = self.N
        if self.order == 1:
            for _ in range(steps):
                self.step += 1
                self._bonds(0, D, minS)
                if N > 2:
                    self.left_canonise()
                    self._bonds(1, D, minS)
                self.left_canonise(measure=self.step % interval == 0)
                self.max_truncation = 0.0
        elif self.order == 2:
            if steps > 0:
                self._bonds(0, D, minS)
                self.left_canonise()
                for _ in range(steps - 1):
                    self.step += 1
                    self._bonds(1, D, minS)
                    self.left_canonise()
                    if self.step % interval == 0:
                        self._bonds(0, D, minS)
                        self.left_canonise(measure=True)
                        self._bonds(0, D, minS)
                        self.left_canonise()
                    else:
                        self._bonds(2, D, minS)
                        self.left_canonise()
                    self.max_truncation = 0.0
                self.step += 1
                self._bonds(1, D, minS)
                self.left_canonise()
                self._bonds(0, D, minS)
                self.left_canonise(measure=self.step % interval == 0)
        elif self.order == 4:
            for _ in range(steps):
                self.step += 1
                for k in range(10):
                    self._bonds(k, D, minS)
                    self.left_canonise()
                self._bonds(10, D, minS)
                self.left_canonise(measure=self.step % interval == 0)
                self.max_truncation = 0.0
        else:
            raise ValueError("order 1, 2 or 4")
        return self.records


def _square_rescale(d):
    """MPXDecomposition::SquareRescale(1.0) (MPX_matrix.hpp:160-164)."""
    v = np.asarray(d["values"], dtype=float)
    s = float(np.sum(v * v))
    d["truncation"] = d["truncation"] / 1.0
    return v * np.sqrt(1.0 / s)


def product_state(rules, phys, levels):
    """Left-canonical product MPS with site v in basis level levels[v-1]:
    [s, a_l, a_r] with 1x1 bonds carrying the accumulated charge (the
    finite analogue of MakeProductStateUnitCell)."""
    phys = np.asarray(phys, dtype=np.int64).reshape(-1, len(rules))
    d = phys.shape[0]
    q = rules.zero(1).reshape(1, -1)
    sites = []
    for lv in levels:
        qn = rules.add(q, phys[lv].reshape(1, -1))
        idx = [MPXIndex(True, phys, True), MPXIndex(True, q), MPXIndex(False, qn)]
        T = np.zeros((d, 1, 1), complex)
        T[lv, 0, 0] = 1.0
        sites.append(MPX.from_dense_tensor(rules, idx, 2, T))
        q = qn
    return sites


def state_vector(sites, left_form=True):
    """Dense amplitude psi[s1, ..., sN] of an MPS given as left-shaped
    ([s, a_l, a_r]) or right-shaped ([a_l, s, a_r]) sites."""
    psi = None
    for A in sites:
        T = A.to_dense_tensor()
        if not left_form:
            T = T.transpose(1, 0, 2)
        if psi is None:
            psi = T[:, 0, :]                         # [s1, a]
        else:
            psi = np.einsum("...a,sab->...sb", psi, T)
    return psi[..., 0].reshape(-1)

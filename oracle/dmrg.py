"""Two-vertex DMRG restatement (oracle; test infrastructure only).

Follows DMRG_routines.cpp:
  * ``TwoVertexComponents`` ctor (779-833): LeftPart = W.LB reordered
    (03214, 3 row indices), RightPart = W.RB reordered (12403, 3 row
    indices); the allowed (QN-consistent) compressed vector index in loop
    order a2 > sigma2 > a0 > sigma1 (808-824);
  * ``TwoVertexMPOMPSMultiply`` (898-974): scatter into V, contract
    LeftPart.V (pairs 30,41), ShiftNumRowIndices(2), contract with
    RightPart (pairs 20,31,42), reshape to a vector, extract allowed; the
    excited-state projector terms (947-973), out += weight (P . V) conj(P)
    with the product taken without conjugation (pinned: with P = conj(psi0)
    the lowest eigenvalue is the first excited one of the dense H_eff);
  * ``TwoVertexComponents::HED`` (835-896): dense route (n <= 1000,
    SPARSE_THRESHOLD, sparse_interface.hpp:16) or ARPACK znaupd with
    nev=1, which="SR", ncv=20, tol 0.1 / maxiter 2 for non-converged
    steps (arpack_interface.cpp:28-39) and the ncv+=3 restart on info=1
    (arpack_interface.hpp:190-218), followed by the phase gauge (870-890);
  * ``MakeLeftBlock``/``MakeRightBlock`` (594-611), initial and dummy
    blocks (583-653), ``TwoVertexInitialWavefunction`` (655-691),
    ``MakePrediction`` (693-707), ``CheckConvergence`` (769-771),
    ``SuperBlock::initialise``/``grow_two_vertex`` (309-374).

The ARPACK here is scipy's bundled ARPACK-NG (same znaupd algorithm the
reference links).  Parity pins (tests/test_oracle_dmrg.py):
  * matvec == brute-force dense einsum of L, W, W, R on the allowed
    subspace (tiny random QN tensors, U(1) and Z2, conj and non-trivial
    target charges);
  * Hermiticity of H_eff for physical models;
  * iDMRG energies without truncation == exact diagonalisation of the
    2N-site open chain (scipy sparse eigsh) for Heisenberg;
  * the reference's own spin_half_example fixture (growth-step energies
    -1,-3,-5,-7,-9 and entropies, examples/user_defined/spin_half_example/
    README.md:31-40);
  * Heisenberg energy per site approaching 1/4 - ln 2 (Hulthen 1938).
ARPACK's loose per-step iterates (tol 0.1) are parity unpinned.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .qn import ChargeRules, MPXIndex
from .mpx import (MPX, contract, contract_to_sparse, reorder, reshape_rows,
                  reshape_to_vector, mpx_svd, diagonal_mpx, EPS)

SPARSE_THRESHOLD = 1000


def fix_complex_precision(c):
    """sparse_interface.cpp:28-32."""
    if c.real != 0 and abs(c.imag / c.real) < EPS:
        return complex(c.real, 0.0)
    if c.imag != 0 and abs(c.real / c.imag) < EPS:
        return complex(0.0, c.imag)
    return c


class TwoVertexComponents:
    """DMRG_routines.hpp:189-214 / DMRG_routines.cpp:779-833."""

    def __init__(self, LB: MPX, W: MPX, RB: MPX, projectors=None):
        """``projectors``: (P, weight) pairs, P the projector tensor's
        entries in the compressed order (ProjectorTensors, 781-806)."""
        self.LB, self.W, self.RB = LB, W, RB
        self.projectors = [(np.asarray(P, dtype=np.complex128), float(w)) for P, w in (projectors or [])]
        rules = W.rules
        self.rules = rules
        lp = contract(W, False, LB, False, [(1, 3)]).remove_dummy_indices([3, 5, 6])
        self.LeftPart = reorder(lp, False, [0, 3, 2, 1, 4], 3)
        rp = contract(W, False, RB, False, [(3, 2)]).remove_dummy_indices([4, 5, 7])
        self.RightPart = reorder(rp, False, [1, 2, 4, 0, 3], 3)
        s1 = W.indices[2].with_direction(True)
        a0 = LB.indices[5].with_direction(True)
        s2 = W.indices[2].with_direction(True)
        a2 = RB.indices[4].with_direction(False)
        self.indices = [s1, a0, s2, a2]
        d1, n0, d2, n2 = s1.size, a0.size, s2.size, a2.size
        self.vrows = d1 * n0
        self.vcols = d2 * n2
        self.length = n0 * d1 * n2 * d2
        # allowed: sigma1 == a2 - sigma2 - a0   (808-824), ascending flat order
        A2 = a2.charges[None, None, None, :, :]
        S2 = s2.charges[None, None, :, None, :]
        A0 = a0.charges[None, :, None, None, :]
        S1 = s1.charges[:, None, None, None, :]
        need = rules.sub(rules.sub(np.broadcast_to(A2, (1, 1, d2, n2, len(rules))), S2), A0)
        ok = np.all(np.broadcast_to(S1, (d1, n0, d2, n2, len(rules))) == need, axis=-1)
        # ok has axes (s1,a0,s2,a2); flat index = s1 + d1*(a0 + n0*(s2 + d2*a2))
        flat_idx = np.flatnonzero(ok.reshape(-1, order="F"))
        self.allowed = flat_idx.astype(np.int64)
        self.rows_and_cols = np.stack([self.allowed % self.vrows, self.allowed // self.vrows], axis=1)

    @property
    def n(self):
        return len(self.allowed)

    # -- TwoVertexMPOMPSMultiply (898-974) --------------------------------------
    def matvec(self, vin):
        vin = np.asarray(vin, dtype=np.complex128)
        V = sp.csc_matrix((vin, (self.rows_and_cols[:, 0], self.rows_and_cols[:, 1])),
                          shape=(self.vrows, self.vcols))
        Vector = MPX(self.rules, self.indices, 2, V)
        M = contract(self.LeftPart, False, Vector, False, [(3, 0), (4, 1)]).shift_num_row_indices(2)
        SM = contract_to_sparse(M, False, self.RightPart, False, [(2, 0), (3, 1), (4, 2)])
        vec = reshape_rows(SM, self.length)
        dense = np.asarray(vec.toarray()).reshape(-1)
        out = dense[self.allowed]
        # projector terms (947-973): s = contract_to_sparse(P, Vector, all
        # four index pairs) without conjugation, out += weight s conj(P)
        for P, w in self.projectors:
            out = out + w * np.sum(P * vin) * np.conj(P)
        return out

    def dense_matrix(self):
        """Dense H_eff on the allowed subspace, column by column through the
        reference-order matvec (the contract_conditional route builds the
        same matrix, MPX_matrix.hpp:199-277)."""
        n = self.n
        H = np.zeros((n, n), dtype=np.complex128)
        e = np.zeros(n, dtype=np.complex128)
        for j in range(n):
            e[:] = 0
            e[j] = 1
            H[:, j] = self.matvec(e)
        return H

    # -- HED (835-896) -----------------------------------------------------
    def hed(self, guess=None, converge=True, tol=None):
        """Lowest ("SR") eigenpair; returns (eigenvalue, vector of length n
        in compressed order)."""
        n = self.n
        if n <= SPARSE_THRESHOLD and not self.projectors:      # dense route without projectors (852)
            H = self.dense_matrix()
            w, v = np.linalg.eigh(0.5 * (H + H.conj().T))
            vec = v[:, 0]
            vec = np.where(np.abs(vec) >= EPS, vec, 0)        # TranslateRows cut
            return float(w[0]), vec
        val, vec = arpack_lowest(self.matvec, n, guess, converge, tol)
        return val, gauge_fix(vec)


class StagewiseHeff:
    """H_eff . v on dense tensors in the four-stage order L . psi, W on
    sigma1, W on sigma2, . R (BLAS tensordots), restricted to the allowed
    entries of ``comp``: an independent restatement of
    TwoVertexMPOMPSMultiply (DMRG_routines.cpp:898-974), pinned against the
    reference-order matvec in tests/test_oracle_dmrg.py and fast enough for
    converged eigensolves at d = 64."""

    def __init__(self, LB: MPX, W: MPX, RB: MPX, comp: TwoVertexComponents):
        self.L = LB.to_dense_tensor()[0, :, 0, :, 0, :]        # (a0', b, a0)
        self.R = RB.to_dense_tensor()[:, 0, :, 0, :, 0]        # (a2', b, a2)
        self.W = W.to_dense_tensor()                           # (s', bl, s, br)
        self.d = self.W.shape[0]
        self.chiL, self.chiR = self.L.shape[0], self.R.shape[0]
        self.allowed = comp.allowed
        self.length = comp.length
        self.n = comp.n

    def matvec(self, v):
        psi = np.zeros(self.length, np.complex128)
        psi[self.allowed] = v
        psi = psi.reshape((self.d, self.chiL, self.d, self.chiR), order="F")   # (s1, a0, s2, a2)
        t1 = np.tensordot(self.L, psi, axes=([2], [1]))                         # (a0', b, s1, s2, a2)
        t2 = np.tensordot(t1, self.W, axes=([1, 2], [1, 2]))                    # (a0', s2, a2, s1', b1)
        t3 = np.tensordot(t2, self.W, axes=([4, 1], [1, 2]))                    # (a0', a2, s1', s2', b2)
        o = np.tensordot(t3, self.R, axes=([1, 4], [2, 1]))                     # (a0', s1', s2', a2')
        o = np.transpose(o, (1, 0, 2, 3))
        return o.reshape(-1, order="F")[self.allowed]


def gauge_fix(vec):
    """Phase gauge of DMRG_routines.cpp:870-890: largest |x| made real
    positive, entries below eps*max dropped, FixComplexPrecision."""
    i = int(np.argmax(np.abs(vec)))
    mx = abs(vec[i])
    ph = mx / vec[i]
    out = vec * ph
    out = np.where(np.abs(out) > EPS * mx, out, 0)
    return np.array([fix_complex_precision(complex(c)) for c in out])


def arpack_lowest(mv, n, guess=None, converge=True, tol=None, ncv=20):
    """arpack_eigs with nev=1, which="SR" (arpack_interface.hpp:174-237,
    arpack_interface.cpp:28-39)."""
    op = spla.LinearOperator((n, n), matvec=mv, dtype=np.complex128)
    maxiter = 50 if converge else 2
    if tol is None:
        tol = 0.0 if converge else 0.1
    ncv = min(ncv, n)
    v0 = None if guess is None else np.asarray(guess, dtype=np.complex128)
    if v0 is not None and not np.any(v0):
        v0 = None                                   # info=-9 path: random start
    while True:
        try:
            w, v = spla.eigs(op, k=1, which="SR", ncv=ncv, tol=tol, maxiter=maxiter, v0=v0)
            break
        except spla.ArpackNoConvergence as e:       # info=1: ncv += 3
            if len(e.eigenvectors):
                v0 = e.eigenvectors[:, 0]
            ncv = min(ncv + 3, n - 1)
    return float(w[0].real), v[:, 0]


# ---------------------------------------------------------------------------
# blocks
# ---------------------------------------------------------------------------

def extract_mpo_block(W: MPX, rows_range, cols_range):
    """MPO_matrix::ExtractMPOBlock (MPO_matrix.cpp:69-85)."""
    d = W.indices[0].size
    r0, r1 = rows_range
    c0, c1 = cols_range
    m = W.matrix.tocsr()[r0 * d:(r1 + 1) * d, :].tocsc()[:, c0 * d:(c1 + 1) * d]
    idx = [W.indices[0], MPXIndex(True, W.indices[1].charges[r0:r1 + 1]),
           W.indices[2], MPXIndex(False, W.indices[3].charges[c0:c1 + 1])]
    return MPX(W.rules, idx, 2, m)


def left_open_bc(W):
    w = W.indices[1].size
    return extract_mpo_block(W, (w - 1, w - 1), (0, W.indices[3].size - 1))


def right_open_bc(W):
    return extract_mpo_block(W, (0, W.indices[1].size - 1), (0, 0))


def make_initial_left_block(LeftH, A):
    """DMRG_routines.cpp:583-587."""
    return contract(A, True, contract(LeftH, False, A, False, [(2, 0)]), False, [(0, 0)])


def make_initial_right_block(RightH, B):
    """DMRG_routines.cpp:588-592."""
    return contract(B, True, contract(RightH, False, B, False, [(2, 1)]), False, [(1, 0)])


def make_left_block(LB, W, A):
    """DMRG_routines.cpp:594-602."""
    t = contract(LB, False, A, False, [(5, 1)])
    t = contract(W, False, t, False, [(1, 3), (2, 5)])
    t = contract(A, True, t, False, [(0, 0), (1, 3)])
    return t.move_dummy_indices([2, 0, 3, 1, 4, 5])


def make_right_block(RB, W, B):
    """DMRG_routines.cpp:603-611."""
    t = contract(RB, False, B, False, [(4, 2)])
    t = contract(W, False, t, False, [(2, 6), (3, 2)])
    t = contract(B, True, t, False, [(1, 0), (2, 2)])
    return t.move_dummy_indices([0, 2, 1, 3, 5, 4])


def make_dummy_left_block(W, target):
    """DMRG_routines.cpp:613-632."""
    rules = W.rules
    z = rules.zero(1)
    idx = [MPXIndex(False, z), MPXIndex(True, z), MPXIndex(True, z),
           MPXIndex(False, W.indices[1].charges), MPXIndex(True, z), MPXIndex(False, z)]
    w = W.indices[1].size
    m = sp.csc_matrix(([1.0], ([w - 1], [0])), shape=(w, 1))
    return MPX(rules, idx, 4, m)


def make_dummy_right_block(W, target):
    """DMRG_routines.cpp:634-653."""
    rules = W.rules
    z = rules.zero(1)
    t = np.asarray(target, dtype=np.int64).reshape(1, -1)
    idx = [MPXIndex(False, t), MPXIndex(True, z), MPXIndex(True, W.indices[3].charges),
           MPXIndex(False, z), MPXIndex(True, t), MPXIndex(False, z)]
    w = W.indices[3].size
    m = sp.csc_matrix(([1.0], ([0], [0])), shape=(w, 1))
    return MPX(rules, idx, 3, m)


def two_vertex_initial_hamiltonian(LeftH, RightH):
    """DMRG_routines.cpp:655-658."""
    return reorder(contract(LeftH, False, RightH, False, [(3, 1)]), False, [0, 3, 2, 4, 1, 5], 2)


def two_vertex_initial_wavefunction(LeftH, RightH, target):
    """DMRG_routines.cpp:660-691: H2.Eigs(TargetSector,1,"SR") on the block
    of columns whose outgoing state equals the target."""
    H2 = two_vertex_initial_hamiltonian(LeftH, RightH)
    rules = H2.rules
    t = np.asarray(target, dtype=np.int64).reshape(-1)
    cols = [c for c in range(H2.cols) if np.array_equal(H2.outgoing_col_state(c), t)]
    Hb = H2.matrix[cols, :][:, cols].toarray()
    w, v = np.linalg.eigh(0.5 * (Hb + Hb.conj().T))
    vec = np.zeros(H2.cols, dtype=np.complex128)
    vec[cols] = np.where(np.abs(v[:, 0]) > EPS / len(cols), v[:, 0], 0)
    d = LeftH.indices[0].size
    phys = LeftH.indices[0].charges
    z = rules.zero(1)
    idx = [MPXIndex(True, phys, True), MPXIndex(True, rules.sub(t, t).reshape(1, -1)),
           MPXIndex(True, phys, True), MPXIndex(False, t.reshape(1, -1))]
    m = sp.csc_matrix(vec.reshape(-1, 1))
    psi = MPX(rules, idx, 2, reshape_rows(m, d))
    return float(w[0]), psi


def psi_from_vector(comp: TwoVertexComponents, vec):
    """DMRG_routines.cpp:765-766: the eigenvector (compressed order) as the
    (d*chiL) x (d*chiR) MPX wavefunction."""
    full = sp.csc_matrix((np.asarray(vec), (comp.allowed, np.zeros(comp.n, np.int64))),
                         shape=(comp.length, 1))
    return MPX(comp.rules, comp.indices, 2, reshape_rows(full, comp.vrows))


def make_prediction(U: MPX, values, Vd: MPX, prev_lambda):
    """DMRG_routines.cpp:693-707; returns (guess sparse vector, auxiliary)."""
    rules = U.rules
    svals = diagonal_mpx(rules, U.indices[2], U.indices[2], values)
    LambdaLB = reorder(contract(U, False, svals, False, [(2, 0)]), False, [1, 0, 2], 1)
    ALambdaR = reorder(contract(svals, False, Vd, False, [(1, 0)]), False, [1, 0, 2], 2)
    rot = mpx_svd(ALambdaR)
    aux = contract_to_sparse(diagonal_mpx(rules, rot["Vd"].indices[0], rot["Vd"].indices[0], rot["values"]),
                             False, rot["Vd"], False, [(1, 0)])
    inv_prev = diagonal_mpx(rules, Vd.indices[2], U.indices[1], prev_lambda, inverse=True)
    g = contract(contract(ALambdaR, False, inv_prev, False, [(2, 0)]), False, LambdaLB, False, [(2, 0)])
    return reshape_to_vector(g), aux


def check_convergence(aux, prev_lambda):
    """DMRG_routines.cpp:769-771: 1 - sum of singular values of
    Auxiliary * diag(previous lambda)."""
    m = (aux @ sp.diags(np.asarray(prev_lambda, dtype=np.complex128))).toarray()
    return 1.0 - float(np.sum(np.linalg.svd(m, compute_uv=False)))


def make_l_finite_prediction(U: MPX, values, A: MPX):
    """MakeLFinitePrediction (DMRG_routines.cpp:709-712): the guess of a
    step to the left, A . (U . diag(values)) with A the stored left matrix
    of the vertex that joins the centre; a sparse vector of the full
    (d chiL d chiR) index in the mixed radix."""
    lam = diagonal_mpx(U.rules, U.indices[2], U.indices[2], values)
    return reshape_to_vector(contract(A, False, contract(U, False, lam, False, [(2, 0)]), False, [(2, 1)]))


def make_r_finite_prediction(U: MPX, values, Vd: MPX, B: MPX):
    """MakeRFinitePrediction (DMRG_routines.cpp:713-716): the guess of a
    step to the right, reorder(diag(values) . Vd, 102, 2) . B with B the
    stored right matrix of the vertex that joins the centre."""
    lam = diagonal_mpx(U.rules, U.indices[2], U.indices[2], values)
    t = reorder(contract(lam, False, Vd, False, [(1, 0)]), False, [1, 0, 2], 2)
    return reshape_to_vector(contract(t, False, B, False, [(2, 0)]))


def check_convergence_finite(guess, U: MPX, values, Vd: MPX):
    """CheckConvergence(Prediction, MPSDecomposition) (DMRG_routines.cpp:
    773-776): 1 - |trace(reshape(guess)^dagger (U diag(values) Vd))|, i.e.
    1 - |<guess|psi>| with psi the truncated, rescaled decomposition."""
    lam = diagonal_mpx(U.rules, Vd.indices[0], Vd.indices[0], values)
    psi = contract_to_sparse(contract(U, False, lam, False, [(2, 0)]), False, Vd, False, [(2, 0)])
    g = np.asarray(guess.toarray()).reshape(-1)
    x = np.asarray(psi.toarray()).reshape(-1, order="F")
    return 1.0 - abs(np.vdot(g, x))


def entropy(values):
    """sparse_interface.cpp:76-85."""
    v = np.asarray(values, dtype=np.float64)
    v = v[v != 0]
    return float(-np.sum(v * v * np.log(v * v)))


class IDMRG:
    """SuperBlock/iDMRG growth (DMRG_routines.cpp:309-374,
    iDMRG_routines.cpp:19-33).  ``converge`` selects the eigensolver
    setting of the growth steps (the reference passes 0)."""

    def __init__(self, W: MPX, target=None, converge=False):
        self.W = W
        self.rules = W.rules
        self.target = np.zeros(len(W.rules), np.int64) if target is None else np.asarray(target, np.int64)
        self.converge = converge
        self.size = 0
        self.history = []
        # block store of the finite algorithm (the reference's .BLOCK and
        # .MPS_matrix files, BlocksStructure DMRG_routines.cpp:57-186):
        # lblk[k] / rblk[k] hold k vertices, lmps[k] / rmps[k] are the MPS
        # matrices of vertex k counted from the left / right end
        self.lblk, self.rblk, self.lmps, self.rmps = {}, {}, {}, {}
        self.nleft = self.nright = 0

    def _svd(self, psi, chi, trunc):
        r = mpx_svd(psi, chi, trunc)
        vals = np.asarray(r["values"], dtype=np.float64)
        sq = float(np.sum(vals * vals))
        vals = vals * np.sqrt(1.0 / sq)                      # SquareRescale(1.0)
        return r["U"], vals, r["Vd"], r["truncation"]

    def initialise(self, chi, trunc):
        W = self.W
        if not W.is_consistent():
            raise ValueError("Hamiltonian MPO is malformed")
        self.LeftH = left_open_bc(W)
        self.RightH = right_open_bc(W)
        E, psi = two_vertex_initial_wavefunction(self.LeftH, self.RightH, self.target)
        bond = (W.indices[0].size if (chi < 1 and trunc <= 0.0) else chi)
        U, vals, Vd, tr = self._svd(psi, bond, trunc)
        self.U, self.values, self.Vd, self.trunc = U, vals, Vd, tr
        self.LB = make_initial_left_block(self.LeftH, U)
        self.RB = make_initial_right_block(self.RightH, Vd)
        self.size = 2
        self.lblk[1], self.rblk[1] = self.LB, self.RB
        self.lmps[1], self.rmps[1] = U, Vd
        self.prev_lambda = np.array([1.0])
        self.guess, self.aux = make_prediction(U, vals, Vd, self.prev_lambda)
        fid = check_convergence(self.aux, self.prev_lambda)
        rec = dict(N=2, energy=E, energy_per_vertex=E / 2.0, entropy=entropy(vals),
                   truncation=tr, fidelity=fid, chi=len(vals))
        self.history.append(rec)
        return rec

    def grow(self, chi, trunc=0.0):
        if self.size < 2:
            return self.initialise(chi, trunc)
        if self.size > 2:
            self.LB = make_left_block(self.LB, self.W, self.U)
            self.RB = make_right_block(self.RB, self.W, self.Vd)
            self.lblk[self.size // 2], self.rblk[self.size // 2] = self.LB, self.RB
        self.size += 2
        self.nleft = self.nright = self.size // 2 - 1
        self.prev_lambda = self.values
        comp = TwoVertexComponents(self.LB, self.W, self.RB)
        g = np.asarray(self.guess.toarray()).reshape(-1)[comp.allowed]
        E, vec = comp.hed(guess=g, converge=self.converge)
        psi = psi_from_vector(comp, vec)
        U, vals, Vd, tr = self._svd(psi, chi, trunc)
        self.U, self.values, self.Vd, self.trunc = U, vals, Vd, tr
        self.lmps[self.size // 2], self.rmps[self.size // 2] = U, Vd
        self.guess, self.aux = make_prediction(U, vals, Vd, self.prev_lambda)
        fid = check_convergence(self.aux, self.prev_lambda)
        rec = dict(N=self.size, energy=E, energy_per_vertex=E / self.size, entropy=entropy(vals),
                   truncation=tr, fidelity=fid, chi=len(vals), n=comp.n)
        self.history.append(rec)
        return rec

    def run(self, steps, chi, trunc=0.0):
        for _ in range(steps):
            self.grow(chi, trunc)
        return self.history

    # -- finite sweeps (DMRG_routines.cpp:376-444, fDMRG_routines.cpp:19-66)
    def _finite_solve(self, guess, chi, trunc):
        comp = TwoVertexComponents(self.lblk_at(self.nleft), self.W, self.rblk_at(self.nright))
        g = np.asarray(guess.toarray()).reshape(-1)[comp.allowed]
        E, vec = comp.hed(guess=g, converge=self.converge)
        U, vals, Vd, tr = self._svd(psi_from_vector(comp, vec), chi, trunc)
        return E, U, vals, Vd, tr, comp.n

    def lblk_at(self, k):
        return make_dummy_left_block(self.W, self.target) if k == 0 else self.lblk[k]

    def rblk_at(self, k):
        return make_dummy_right_block(self.W, self.target) if k == 0 else self.rblk[k]

    def move_right(self, chi, trunc=0.0):
        """SuperBlock::move_right_two_vertex (DMRG_routines.cpp:376-409)."""
        if self.nright < 1:
            raise ValueError("At right end of superblock already...")
        if self.nleft < 1:
            newleft = make_initial_left_block(self.LeftH, self.U)
        else:
            newleft = make_left_block(self.lblk[self.nleft], self.W, self.U)
        self.nleft += 1
        self.nright -= 1
        self.lblk[self.nleft] = newleft
        self.prev_lambda = self.values
        guess = make_r_finite_prediction(self.U, self.values, self.Vd, self.rmps[self.nright + 1])
        t = -0.0 if self.nright <= 1 else trunc
        return self._finish_move(guess, chi, t)

    def move_left(self, chi, trunc=0.0):
        """SuperBlock::move_left_two_vertex (DMRG_routines.cpp:411-444)."""
        if self.nleft < 1:
            raise ValueError("At left end of superblock already...")
        if self.nright < 1:
            newright = make_initial_right_block(self.RightH, self.Vd)
        else:
            newright = make_right_block(self.rblk[self.nright], self.W, self.Vd)
        self.nright += 1
        self.nleft -= 1
        self.rblk[self.nright] = newright
        self.prev_lambda = self.values
        guess = make_l_finite_prediction(self.U, self.values, self.lmps[self.nleft + 1])
        t = -0.0 if self.nleft <= 1 else trunc
        return self._finish_move(guess, chi, t)

    def _finish_move(self, guess, chi, trunc):
        E, U, vals, Vd, tr, n = self._finite_solve(guess, chi, trunc)
        self.U, self.values, self.Vd, self.trunc = U, vals, Vd, tr
        self.lmps[self.nleft + 1], self.rmps[self.nright + 1] = U, Vd
        fid = check_convergence_finite(guess, U, vals, Vd)
        return dict(N=self.size, left=self.nleft, right=self.nright, energy=E,
                    energy_per_vertex=E / self.size, entropy=entropy(vals), truncation=tr,
                    fidelity=fid, chi=len(vals), n=n)

    def finite_sweeps(self, num_sweeps, chi, trunc=0.0):
        """FiniteDMRG::run (fDMRG_routines.cpp:19-66): each sweep moves right
        to the end, left to the other end and right back to the middle;
        returns (every step's record, the records taken at the midpoint,
        i.e. the rows the reference pushes to Energies.dat)."""
        steps, mid = [], []
        if self.size == 2:
            return steps, mid
        for _ in range(num_sweeps):
            for _ in range(self.nright):
                steps.append(self.move_right(chi, trunc))
            for _ in range(self.nleft):
                r = self.move_left(chi, trunc)
                steps.append(r)
                if self.nleft == self.nright:
                    mid.append(r)
            # for (r = right_size(); r > left_size(); --r): both sides move
            for _ in range((self.nright - self.nleft + 1) // 2):
                r = self.move_right(chi, trunc)
                steps.append(r)
                if self.nleft == self.nright:
                    mid.append(r)
        return steps, mid

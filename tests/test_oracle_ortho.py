"""Oracle pins for the unit-cell orthogonalisation and the energy per vertex
(measurement.cpp:180-246, 371-397, 405-629; MPX_matrix.cpp:829-913;
UnitCell.hpp:42-48), each against an independent formulation:
  * matvecs against the explicit transfer matrices built with Kronecker
    products (vec(A^H v A) = (A^T kron A^H) vec(v), vec(conj(A) v A^T) =
    (A kron conj(A)) vec(v), column-major vec);
  * the leading eigenvalue against numpy.linalg.eig of that matrix;
  * the new lambdas of real cells against the Schmidt values of the
    textbook formula s^2 = eig(sqrt(l) r sqrt(l)) from the dense fixed points;
  * the output cell's canonical conditions (left fixed point identity, right
    fixed point Lambda0^2, eigenvalue 1);
  * iTwoVertexEnergy against the expectation values of the full two-site
    bond Hamiltonians on a four-vertex reduced density matrix.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import dmrg, models, ortho, tebd


def _idmrg_cell(model="heisenberg", chi=10, steps=6):
    if model == "heisenberg":
        rules, phys, W = models.heisenberg_mpo()
    else:
        rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=3, Jperp=0.3)
    g = dmrg.IDMRG(W, converge=True)
    for _ in range(steps):
        g.grow(chi)
    mats, lams = ortho.unit_cell_from_decomposition(g.U, g.values, g.Vd, g.prev_lambda)
    return W, mats, lams, g


def _kron_left(cell):
    n = cell[0].shape[1]
    T = np.eye(n * n, dtype=complex)
    for A in cell:
        T = sum(np.kron(A[s].T, A[s].conj().T) for s in range(A.shape[0])) @ T
    return T


def _kron_right(cell):
    n = cell[-1].shape[2]
    T = np.eye(n * n, dtype=complex)
    for A in reversed(cell):
        T = sum(np.kron(A[s], A[s].conj()) for s in range(A.shape[0])) @ T
    return T


@pytest.mark.parametrize("model", ["heisenberg", "coupled"])
def test_transfer_matvecs_and_leading_value(model):
    W, mats, lams, g = _idmrg_cell(model, chi=8, steps=4)
    cell = [m.to_dense_tensor() for m in mats]
    n = cell[0].shape[1]
    rng = np.random.default_rng(5)
    v = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    TL, TR = _kron_left(cell), _kron_right(cell)
    assert np.allclose(ortho.left_apply(cell, v).reshape(-1, order="F"), TL @ v.reshape(-1, order="F"),
                       atol=1e-12)
    assert np.allclose(ortho.right_apply(cell, v).reshape(-1, order="F"), TR @ v.reshape(-1, order="F"),
                       atol=1e-12)
    lead = np.max(np.abs(np.linalg.eigvals(TL)))
    lamL, ML, _ = ortho.transfer_leading(mats, 0)
    lamR, MR, _ = ortho.transfer_leading(mats, 1)
    assert abs(lamL - lead) <= 1e-10 * lead and abs(lamR - lead) <= 1e-10 * lead
    # the completed matrices are Hermitian fixed points of their maps
    assert np.allclose(ML, ML.conj().T, atol=1e-13) and np.allclose(MR, MR.conj().T, atol=1e-13)
    assert np.allclose(ortho.left_apply(cell, ML), lamL * ML, atol=1e-10 * np.abs(ML).max())
    assert np.allclose(ortho.right_apply(cell, MR), lamR * MR, atol=1e-10 * np.abs(MR).max())


@pytest.mark.parametrize("model", ["heisenberg", "coupled"])
def test_orthogonalise_schmidt_values_and_canonical_form(model):
    W, mats, lams, g = _idmrg_cell(model, chi=10, steps=6)
    cell = [m.to_dense_tensor() for m in mats]
    # textbook Schmidt values at the cell boundary from the dense fixed points
    TL, TR = _kron_left(cell), _kron_right(cell)
    n = cell[0].shape[1]
    wl, vl = np.linalg.eig(TL)
    wr, vr = np.linalg.eig(TR)
    l = vl[:, np.argmax(np.abs(wl))].reshape(n, n, order="F")
    rT = vr[:, np.argmax(np.abs(wr))].reshape(n, n, order="F")
    l = l / np.trace(l)
    r = rT.T / np.trace(rT)                    # sum A r A^H = lambda r
    sl = sla.sqrtm(0.5 * (l + l.conj().T))
    s2 = np.sort(np.linalg.eigvalsh(sl @ (0.5 * (r + r.conj().T)) @ sl))[::-1]
    s2 = s2[s2 > 1e-14 * s2[0]]
    s2 = s2 / s2.sum()
    new, nl, info = ortho.orthogonalise(mats, lams)
    got = np.sort(np.asarray(nl[0]) ** 2)[::-1]
    got = got[got > 1e-14 * got[0]]
    assert len(got) == len(s2)
    assert np.allclose(got, s2, rtol=1e-8, atol=1e-12)
    assert abs(info["entropy"] - dmrg.entropy(nl[0])) < 1e-14
    # canonical conditions of the output cell; X^-1 divides by square roots
    # of fixed-point eigenvalues down to 100 eps of the largest (the
    # reference's cut), which amplifies eps-level errors to ~1e-8
    c2 = [m.to_dense_tensor() for m in new]
    m = c2[0].shape[1]
    assert np.allclose(ortho.left_apply(c2, np.eye(m)), np.eye(m), atol=1e-7)
    R = np.diag(np.asarray(nl[0]) ** 2)
    A = np.einsum("pab,qbc->pqac", c2[0], c2[1]).reshape(-1, m, m)
    assert np.allclose(np.einsum("sab,bc,sdc->ad", A, R, A.conj()), R, atol=1e-7)
    # the middle lambda is carried over
    assert np.array_equal(nl[1], lams[1])


def _bond(W, odd):
    H = (tebd.odd_bond_hamiltonian(W) if odd else tebd.even_bond_hamiltonian(W)).to_dense_tensor()
    d = H.shape[0]
    return H.reshape(d, d, d, d)


@pytest.mark.parametrize("model", ["heisenberg", "coupled"])
def test_energy_per_vertex_against_bond_hamiltonians(model):
    W, mats, lams, g = _idmrg_cell(model, chi=10, steps=6)
    new, nl, info = ortho.orthogonalise(mats, lams)
    E, terms = ortho.i_two_vertex_energy(W, new, nl)
    A0, A1 = [m.to_dense_tensor() for m in new]
    l2 = np.asarray(nl[0]) ** 2
    T2 = np.einsum("pab,qbc->pqac", A0, A1)
    e01 = np.einsum("PQac,PQpq,pqac,c->", T2.conj(), _bond(W, True), T2, l2)
    T4 = np.einsum("pab,qbc,rcd,sde->pqrsae", A0, A1, A0, A1)
    e12 = np.einsum("PQRSae,QRqr,PqrSae,e->", T4.conj(), _bond(W, False), T4, l2)
    assert abs(E - 0.5 * (e01 + e12)) <= 1e-12
    assert abs(E.imag) <= 1e-12
    if model == "heisenberg":
        # the thermodynamic value 1/4 - ln 2 to the accuracy of chi = 10
        assert abs(E.real - (0.25 - np.log(2.0))) < 2e-3


def test_sqrt_dr_rejects_mixed_signs():
    from oracle.mpx import MPX
    from oracle.qn import ChargeRules, MPXIndex
    import scipy.sparse as sp
    rules = ChargeRules([0])
    q = np.array([[0], [0], [1]])
    M = MPX(rules, [MPXIndex(True, q), MPXIndex(False, q)], 1, sp.csc_matrix(np.diag([1.0, -1.0, 2.0])))
    assert ortho.sqrt_dr(M) is None
    M = MPX(rules, [MPXIndex(True, q), MPXIndex(False, q)], 1, sp.csc_matrix(np.diag([-1.0, -4.0, 1e-20])))
    roots, R = ortho.sqrt_dr(M)
    assert np.allclose(sorted(roots), [1.0, 2.0]) and R.indices[0].size == 2

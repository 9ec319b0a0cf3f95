"""Oracle pins for the iTEBD restatement (TEBD_routines.cpp:14-85, 233-266,
826-880): bond operators against scipy.linalg.expm of the dense bond
Hamiltonian, unitarity, and an untruncated gate step that preserves the
norm and is exactly the matrix exponential acting on the two-site state."""
import numpy as np
import scipy.linalg as sla

from oracle import models, tebd
from oracle import mpx as O


def _bond_dense(Hb):
    m = Hb.matrix.toarray()
    return m


def test_bond_hamiltonians_sum_to_two_site_h():
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=1, Jperp=0.7)
    Ho = tebd.odd_bond_hamiltonian(W).matrix.toarray()
    He = tebd.even_bond_hamiltonian(W).matrix.toarray()
    d = len(E)
    # rows (s1', s2'), cols (s1, s2); odd bond = H_V x 1 + 1 x H_V + J sig x sig
    Hv = np.diag(E)
    exp_odd = np.kron(np.eye(d), Hv) + np.kron(Hv, np.eye(d)) + 0.7 * np.kron(sig, sig)
    exp_even = 0.7 * np.kron(sig, sig)
    # mixed radix: row = s1' + d*s2'  -> kron(A for s2', B for s1')
    assert np.allclose(Ho, exp_odd, atol=1e-13)
    assert np.allclose(He, exp_even, atol=1e-13)


def test_evolution_operator_is_expm_and_unitary():
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=2, Jperp=0.4)
    Hb = tebd.odd_bond_hamiltonian(W)
    G = tebd.bond_evolution_operator(Hb, 0.13)
    g = G.matrix.toarray()
    assert np.allclose(g, sla.expm(-1j * 0.13 * Hb.matrix.toarray()), atol=1e-12)
    assert np.allclose(g.conj().T @ g, np.eye(g.shape[0]), atol=1e-12)
    assert G.is_consistent()


def test_untruncated_gate_step_is_exact():
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=3, Jperp=0.5)
    G = tebd.bond_evolution_operator(tebd.odd_bond_hamiltonian(W), 0.2)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 0)
    nA0, nA1, nl0, tr = tebd.apply_and_decompose(G, A0, A1, l0, D=0, minS=0.0)
    assert tr == 0.0
    assert abs(np.sum(nl0 ** 2) - 1.0) < 1e-12           # unitary gate keeps the norm
    # two-site state U . A1' equals G applied to the product state |0 0>
    two = O.contract(nA1, False, nA0, False, [(2, 1)]).to_dense_tensor()[:, 0, :, 0]  # (s1', s2')
    d = len(E)
    psi0 = np.zeros(d * d, complex)
    psi0[0] = 1.0
    exp = (G.matrix.toarray() @ psi0).reshape(d, d, order="F")     # (s1', s2')
    assert np.allclose(two, exp, atol=1e-12)

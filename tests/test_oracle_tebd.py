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


def _even_bond_error(order, dt, ref, cell0, W, T=0.4, D=64):
    ops = tebd.trotter_operators(W, dt, order)
    c, _ = tebd.evolve(ops, order, list(cell0), int(round(T / dt)), D, 1e-12)
    a, b = np.sort(ref[3])[::-1], np.sort(c[3])[::-1]
    n = min(len(a), len(b))
    return np.abs(a[:n] - b[:n]).max()


def test_fourth_order_itebd_converges_at_fourth_order():
    """Fourth order (TEBD_routines.cpp:52-85, 194-224): the even-bond
    spectrum (which the closing odd half step leaves alone) against a run
    with dt = 0.025: halving dt from 0.2 to 0.1 divides the error by more
    than 10 (2^4 = 16; second order gives ~4), and at dt = 0.1 the fourth
    order is more than 3 times closer than the second."""
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=2, Jperp=0.4)
    cell0 = list(tebd.product_state_cell(rules, phys, 0))
    ref, _ = tebd.evolve(tebd.trotter_operators(W, 0.025, 4), 4, list(cell0), 16, 64, 1e-12)
    e4a = _even_bond_error(4, 0.2, ref, cell0, W)
    e4b = _even_bond_error(4, 0.1, ref, cell0, W)
    e2b = _even_bond_error(2, 0.1, ref, cell0, W)
    assert e4a / e4b > 10.0, (e4a, e4b)
    assert e4b < e2b / 3.0, (e4b, e2b)


def test_inversion_symmetric_orthogonalisation_is_left_canonical():
    """OrthogonaliseInversionSymmetric (measurement.cpp:131-178) of an
    evolved iTEBD cell: the left transfer map of the new cell fixes the
    identity with eigenvalue 1, and the new lambdas have unit square sum."""
    from oracle import ortho
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=2, Jperp=0.4)
    cell0 = list(tebd.product_state_cell(rules, phys, 0))
    c, _ = tebd.evolve(tebd.trotter_operators(W, 0.1, 2), 2, cell0, 3, 32, 1e-12)
    new, nl, info = ortho.orthogonalise_inversion_symmetric([c[0], c[1]], [c[2], c[3]])
    dense = [m.to_dense_tensor() for m in new]
    m = dense[0].shape[1]
    assert np.allclose(ortho.left_apply(dense, np.eye(m)), np.eye(m), atol=1e-8)
    assert abs(np.sum(np.asarray(nl[0]) ** 2) - 1.0) < 1e-12
    assert np.array_equal(nl[1], c[3])


def _level_projector(rules, phys, k):
    """|k><k| as a one-vertex MPO operator (MPO bond dimension 1)."""
    from oracle.mpx import MPX
    from oracle.qn import MPXIndex
    d = phys.shape[0]
    z = np.zeros((1, len(rules.rules)), dtype=np.int64)
    T = np.zeros((d, 1, d, 1), complex)
    T[k, 0, k, 0] = 1.0
    idx = [MPXIndex(True, phys, True), MPXIndex(True, z), MPXIndex(False, phys, True), MPXIndex(False, z)]
    return MPX.from_dense_tensor(rules, idx, 2, T)


def test_measured_itebd_closed_form():
    """Measurement steps (TEBD_routines.cpp:99-231): without the inter-chain
    coupling the product state of level 0 is an eigenstate, so every
    measurement sees the overlap per two-vertex cell exp(-2 i E_0 t), zero
    entropy and occupation 1 of level 0 -- at orders 1, 2 and 4."""
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=2, Jperp=0.0)
    cell0 = list(tebd.product_state_cell(rules, phys, 0))
    P0 = _level_projector(rules, phys, 0)
    dt = 0.1
    for order in (1, 2, 4):
        ops = tebd.trotter_operators(W, dt, order)
        _, _, recs = tebd.evolve_measured(ops, order, list(cell0), 4, 16, 1e-12, 2, [P0])
        assert [r["step"] for r in recs] == [2, 4]
        for r in recs:
            assert abs(r["overlap"] - np.exp(-2j * E[0] * r["step"] * dt)) <= 1e-12, (order, r)
            assert abs(r["entropy"]) <= 1e-12
            assert abs(r["values"][0] - 1.0) <= 1e-12


def test_overlap_of_a_canonical_cell_with_itself_is_one():
    from oracle import ortho
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=2, Jperp=0.4)
    c, _ = tebd.evolve(tebd.trotter_operators(W, 0.1, 2), 2, list(tebd.product_state_cell(rules, phys, 0)), 3, 16,
                       1e-12)
    new, nl, info = ortho.orthogonalise_inversion_symmetric([c[0], c[1]], [c[2], c[3]])
    assert abs(ortho.overlap(new, new) - 1.0) <= 1e-10

"""Oracle pins for the finite-chain TEBD restatement (TEBD_routines.cpp:14-85,
273-547, 672-791): without truncation the evolved MPS is the dense state
vector under the same products of bond exponentials (scipy.linalg.expm),
phase included; sites are left canonical after each step; the recorded
entropy and overlap equal those of the dense state; the fourth-order
coefficients satisfy Suzuki's conditions and the fourth-order product is
closer to exp(-i t H) than the second-order one."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import ftebd, models, tebd


def _apply_bond(psi, G, i, d, N):
    """G (d^2 x d^2, rows s_i' + d s_{i+1}') on sites i, i+1 (0-based) of
    psi[s_1, ..., s_N]."""
    G4 = G.reshape(d, d, d, d, order="F")                 # [a', b', a, b]
    t = psi.reshape([d] * N)
    t = np.tensordot(G4, t, axes=([2, 3], [i, i + 1]))    # [a', b', rest...]
    t = np.moveaxis(t, [0, 1], [i, i + 1])
    return t.reshape(-1)


def _model(d=4, seed=5, J=0.6):
    rules, phys, W, E, sig = models.coupled_chain_model(d=d, seed=seed, Jperp=J)
    return rules, phys, W


def _dense_sequence(W, tau, order, steps, N, d):
    ops, seq = ftebd.trotter_operators(W, tau, order)
    mats = [o.matrix.toarray() for o in ops]
    odd = [i for i in range(0, N - 1, 2)]
    even = [i for i in range(1, N - 1, 2)]

    def sweep(k):
        bonds = odd if k % 2 == 0 else even
        return [(mats[seq[k]], b) for b in bonds]

    if order == 1:
        plan = [g for _ in range(steps) for k in (0, 1) for g in sweep(k)]
    elif order == 2:
        plan = sweep(0) + [g for _ in range(steps - 1) for k in (1, 2) for g in sweep(k)] + sweep(1) + sweep(0)
    else:
        plan = [g for _ in range(steps) for k in range(11) for g in sweep(k)]
    return plan


@pytest.mark.parametrize("order,N", [(1, 4), (2, 4), (4, 4), (2, 6), (1, 2), (2, 2)])
def test_untruncated_finite_tebd_is_exact(order, N):
    rules, phys, W = _model()
    d = phys.shape[0]
    levels = [0, 1, 2, 3, 0, 1][:N]
    sites = ftebd.product_state(rules, phys, levels)
    tau, steps = 0.07, 3
    run = ftebd.FiniteTEBD(W, sites, tau, order)
    recs = run.evolve(steps, D=0, minS=0.0, interval=1)
    psi = np.zeros(d ** N, complex)
    psi[np.ravel_multi_index(levels, [d] * N)] = 1.0
    psi0 = psi.copy()
    for G, b in _dense_sequence(W, tau, order, steps, N, d):
        psi = _apply_bond(psi, G, b, d, N)
    got = ftebd.state_vector(run.left[1:])
    assert np.allclose(got, psi, atol=1e-12), np.max(np.abs(got - psi))
    # left canonical sites
    for A in run.left[1:]:
        T = A.to_dense_tensor()                           # [s, a_l, a_r]
        M = T.reshape(-1, T.shape[2])
        assert np.allclose(M.conj().T @ M, np.eye(M.shape[1]), atol=1e-12)
    # the last record: entropy of bond N/2 and overlap with the initial state
    last = recs[-1]
    assert last["step"] == steps and abs(last["time"] - steps * tau) < 1e-15
    cut = N // 2
    sv = np.linalg.svd(psi.reshape(d ** cut, -1), compute_uv=False)
    p = sv[sv > 1e-15] ** 2
    assert abs(last["entropy"] - float(-np.sum(p * np.log(p)))) < 1e-10
    assert abs(last["overlap"] - np.vdot(psi0, psi)) < 1e-12
    assert last["truncation"] < 1e-20


def test_fourth_order_coefficients():
    p, q, r = ftebd.suzuki4_coefficients()
    assert abs(4 * p + q - 1.0) < 1e-15
    assert abs(4 * p ** 3 + q ** 3) < 1e-15
    assert abs(r - (p + q)) < 1e-15
    # the decimal values TrotterDecomposition writes out (TEBD_routines.cpp:57-59)
    assert abs(p - 0.41449077179437573714) < 1e-15
    assert abs(q + 0.65796308717750294857) < 1e-15


def test_fourth_order_beats_second_order():
    rules, phys, W = _model(d=2, seed=7, J=0.9)
    d, N, t = 2, 6, 0.4
    ham = np.zeros((d ** N, d ** N), complex)
    Ho = tebd.odd_bond_hamiltonian(W).matrix.toarray()
    He = tebd.even_bond_hamiltonian(W).matrix.toarray()
    for b in range(N - 1):
        Hb = Ho if b % 2 == 0 else He
        for c in range(d ** N):
            e = np.zeros(d ** N, complex)
            e[c] = 1.0
            ham[:, c] += _apply_bond(e, Hb, b, d, N)
    levels = [0, 1, 0, 1, 1, 0]
    psi0 = np.zeros(d ** N, complex)
    psi0[np.ravel_multi_index(levels, [d] * N)] = 1.0
    exact = sla.expm(-1j * t * ham) @ psi0
    err = {}
    for order, steps in ((2, 4), (4, 4)):
        run = ftebd.FiniteTEBD(W, ftebd.product_state(rules, phys, levels), t / steps, order)
        run.evolve(steps)
        err[order] = np.linalg.norm(ftebd.state_vector(run.left[1:]) - exact)
    assert err[4] < 0.2 * err[2], err


def test_truncated_run_records_truncation():
    rules, phys, W = _model(d=4, seed=9, J=1.5)
    sites = ftebd.product_state(rules, phys, [0, 3, 1, 2, 0, 3])
    run = ftebd.FiniteTEBD(W, sites, 0.3, 2)
    recs = run.evolve(3, D=3, minS=0.0, interval=1)
    assert len(recs) == 3
    assert recs[-1]["truncation"] > 1e-8
    for A in run.left[1:]:
        assert A.indices[2].size <= 4 and A.indices[1].size <= 4
    assert abs(np.linalg.norm(ftebd.state_vector(run.left[1:])) - 1.0) < 1e-12

"""GPU: adjoint of the block-sparse H_eff contraction (no reference counterpart).

* dE/dconj(psi) == (H psi - E psi)/<psi|psi> (oracle matvec), 1e-12;
* dE/dW against central finite differences of the oracle's
  E(W) = <psi|H(W)|psi>/<psi|psi> (step 1e-5, FP64) on sampled MPO entries,
  relative 1e-6;
* Hellmann-Feynman: for the H_eff ground state, dE/dJ_perp from the adjoint
  equals <psi|dH_eff/dJ_perp|psi> and the central difference of the lowest
  eigenvalue; dE/dE_n (all levels) equals the diagonal MPO block gradient.
"""
import numpy as np
import pytest

from conftest import to_lib, rel_err
from oracle import models, dmrg

pytestmark = pytest.mark.gpu


def _setup(d=8, chi=16, seed=2, Jperp=0.2):
    rules, phys, W, E, sig = models.coupled_chain_model(d=d, seed=seed, Jperp=Jperp)
    idm = dmrg.IDMRG(W, converge=True)
    for _ in range(3):
        idm.grow(chi)
    LB = dmrg.make_left_block(idm.LB, W, idm.U)
    RB = dmrg.make_right_block(idm.RB, W, idm.Vd)
    return rules, phys, W, E, sig, LB, RB


def _energy(LB, W, RB, psi):
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    return (np.vdot(psi, comp.matvec(psi)) / np.vdot(psi, psi)).real


def test_vjp_psi_and_W_fd(ctx):
    print("case vjp_fd", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig, LB, RB = _setup()
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    rng = np.random.default_rng(0)
    psi = rng.standard_normal(comp.n) + 1j * rng.standard_normal(comp.n)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    m = W.matrix.tocsc()
    m.sort_indices()
    dpsi, dW = H.vjp(psi, m.nnz)
    n2 = np.vdot(psi, psi).real
    Hpsi = comp.matvec(psi)
    En = np.vdot(psi, Hpsi).real / n2
    assert rel_err(dpsi, (Hpsi - En * psi) / n2) <= 1e-12
    h = 1e-5
    picks = rng.choice(m.nnz, size=min(12, m.nnz), replace=False)
    for k in picks:
        Wp, Wm = W.copy(), W.copy()
        Wp.matrix = m.copy(); Wp.matrix.data[k] += h
        Wm.matrix = m.copy(); Wm.matrix.data[k] -= h
        fd = (_energy(LB, Wp, RB, psi) - _energy(LB, Wm, RB, psi)) / (2 * h)
        # E is holomorphic in each W entry; the real-step difference is dE/dW_k
        assert abs(dW[k].real - fd) <= 1e-6 * max(1.0, abs(fd)), (k, dW[k], fd)


def _model_W(d, J, E=None):
    rules, phys, W, E0, sig = models.coupled_chain_model(d=d, seed=2, Jperp=J, E=E)
    return W, E0


def test_hellmann_feynman_couplings(ctx):
    print("case hellmann_feynman", flush=True)
    import chainamps_amd as cam
    J = 0.2
    rules, phys, W, E, sig, LB, RB = _setup(Jperp=J)
    d = len(E)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    E0, psi, info = H.lowest(converge=True)
    m = W.matrix.tocsc()
    m.sort_indices()
    dpsi, dW = H.vjp(psi, m.nnz)
    # at the ground state dE/dconj(psi) vanishes
    assert np.linalg.norm(dpsi) <= 1e-9
    col_of = np.repeat(np.arange(m.shape[1]), np.diff(m.indptr))
    row_of = m.indices
    bl, br = row_of // d, col_of // d
    # chain rule: only W[1,0] = J*sigma depends on J
    dWdJ = np.where((bl == 1) & (br == 0), m.data / J, 0)
    dEdJ = np.sum(dW * dWdJ)
    assert abs(dEdJ.imag) <= 1e-10
    dEdJ = dEdJ.real
    # Hellmann-Feynman expectation <psi|dH_eff/dJ|psi>: E(psi; J) is
    # quadratic in J at fixed psi, so the central difference is exact
    h = 1e-3
    hf = (_energy(LB, _model_W(d, J + h)[0], RB, psi) - _energy(LB, _model_W(d, J - h)[0], RB, psi)) / (2 * h)
    assert abs(dEdJ - hf) <= 1e-9 * max(1.0, abs(hf)), (dEdJ, hf)
    # central finite difference of the lowest eigenvalue (environments fixed)
    hh = 1e-5

    def lowest(Jv):
        Wj, _ = _model_W(d, Jv)
        return cam.Heff(to_lib(ctx, LB), to_lib(ctx, Wj), to_lib(ctx, RB)).lowest(converge=True)[0]
    fd = (lowest(J + hh) - lowest(J - hh)) / (2 * hh)
    assert abs(dEdJ - fd) <= 1e-6 * max(1.0, abs(fd)), (dEdJ, fd)
    # the library's chain rule over the coupling generators
    from chainamps_amd import workloads
    csc = (m.shape[0], m.shape[1], m.indptr.astype(np.int64), m.indices.astype(np.int64), m.data)
    g, dpsi2 = H.coupling_grad(psi, m.nnz, workloads.coupled_chain_generators(csc, d, seed=2), want_dpsi=True)
    assert len(g) == d + 1
    assert abs(g[0] - dEdJ) <= 1e-12 * max(1.0, abs(dEdJ)), (g[0], dEdJ)
    assert abs(g[0].real - fd) <= 1e-6 * max(1.0, abs(fd))
    assert np.max(np.abs(dpsi2 - dpsi)) <= 1e-11      # same vjp, run again
    # dE/dE_n for every level: chain rule through W[2,0] = diag(E), against
    # the (exact) central difference of E(psi; E_n)
    for nlev in range(d):
        k = np.where((bl == 2) & (br == 0) & (row_of % d == nlev) & (col_of % d == nlev))[0]
        assert len(k) == 1
        Ep, Em = E.copy(), E.copy()
        Ep[nlev] += h
        Em[nlev] -= h
        ref = (_energy(LB, _model_W(d, J, Ep)[0], RB, psi) - _energy(LB, _model_W(d, J, Em)[0], RB, psi)) / (2 * h)
        assert abs(dW[k[0]].real - ref) <= 1e-9 * max(1.0, abs(ref)), (nlev, dW[k[0]], ref)
        assert abs(g[1 + nlev] - dW[k[0]]) <= 1e-12 * max(1.0, abs(dW[k[0]]))

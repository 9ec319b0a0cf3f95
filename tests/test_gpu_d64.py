"""GPU parity at d = 64, the local dimension of the coupled-chain
configurations (levels E_n sorted uniform[0,8], Z2 parity n mod 2,
Gaussian inter-chain operator between opposite parities).

* matvec (TwoVertexMPOMPSMultiply, DMRG_routines.cpp:898-974) at chi = 24:
  allowed map bit-exact, FP64 within 1e-12 and FP32 within 2e-5 of the
  oracle's reference-order matvec, unsliced and sliced;
* lowest eigenpair at chi = 16 (environments of the oracle's first growth
  step, n = 524288): FP64 and FP32-mixed device eigensolves within 1e-10 of
  the oracle's (scipy eigsh on the oracle's stage-wise matvec), residual
  against the oracle matvec;
* adjoint at the same H_eff: dE/dpsi, dE/dJ_perp against Hellmann-Feynman
  and the central difference of the lowest eigenvalue (1e-6), dE/dE_n along
  random directions and per level, dE/dW on random entries against central
  differences (step 1e-5);
* iTEBD quench J_perp 0 -> 0.1 from the product of level 0, second order,
  dt = 0.01, chi = 16: overlap with the initial cell, entropy, level-0 and
  level-1 occupations and <sigma^2> at measurement steps within 1e-8 of the
  oracle (<sigma> itself vanishes identically in the Z2-symmetric state).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import to_lib, rel_err
from oracle import models, dmrg, tebd

pytestmark = pytest.mark.gpu

D = 64


def _random_env(chi, seed=5):
    rng = np.random.default_rng(seed)
    rules, phys, W, E, sig = models.coupled_chain_model(d=D, seed=seed, Jperp=0.1)
    aq = (np.arange(chi) % 2).reshape(-1, 1)
    bq = W.indices[1].charges
    LB = models.random_left_block(rules, aq, bq, rng)
    RB = models.random_right_block(rules, aq, bq, rng)
    return LB, W, RB


def test_d64_matvec_parity(ctx):
    print("case d64 matvec chi=24", flush=True)
    import chainamps_amd as cam
    LB, W, RB = _random_env(24)
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    assert H.n == comp.n == (24 * D) ** 2 // 2
    assert np.array_equal(H.allowed(), comp.allowed)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(comp.n) + 1j * rng.standard_normal(comp.n)
    ref = comp.matvec(v)
    assert rel_err(H.apply(v), ref) <= 1e-12
    assert rel_err(H.apply(v, cam.CAM_FP32), ref) <= 2e-5
    H.set_slices(2, 2)
    assert rel_err(H.apply(v), ref) <= 1e-12


@pytest.fixture(scope="module")
def grown(ctx):
    """H_eff of the second growth step of the d=64 chain (J_perp = 0.1):
    environments from the oracle's first step kept at chi = 16."""
    J = 0.1
    rules, phys, W, E, sig = models.coupled_chain_model(d=D, seed=2, Jperp=J)
    idm = dmrg.IDMRG(W, converge=True)
    idm.grow(16)
    LB, RB = idm.LB, idm.RB
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    sw = dmrg.StagewiseHeff(LB, W, RB, comp)
    return dict(J=J, E=E, W=W, LB=LB, RB=RB, comp=comp, sw=sw)


def _oracle_lowest(sw):
    op = spla.LinearOperator((sw.n, sw.n), matvec=sw.matvec, dtype=np.complex128)
    v0 = np.random.default_rng(7).standard_normal(sw.n) + 0j
    w, v = spla.eigsh(op, k=1, which="SA", tol=1e-14, v0=v0)
    return float(w[0]), v[:, 0]


def test_d64_lowest_fp64_and_mixed(ctx, grown):
    print("case d64 lowest chi=16", flush=True)
    import chainamps_amd as cam
    comp, sw = grown["comp"], grown["sw"]
    H = cam.Heff(to_lib(ctx, grown["LB"]), to_lib(ctx, grown["W"]), to_lib(ctx, grown["RB"]))
    print("  oracle eigsh ...", flush=True)
    E_ref, v_ref = _oracle_lowest(sw)
    print(f"  oracle E0 = {E_ref:.14f}", flush=True)
    for mixed in (False, True):
        E, vec, info = H.lowest(converge=True, mixed=mixed)
        assert abs(E - E_ref) <= 1e-10 * max(1.0, abs(E_ref)), (mixed, E, E_ref)
        r = comp.matvec(vec) - E * vec
        assert np.linalg.norm(r) <= 1e-8 * abs(E) * np.linalg.norm(vec), mixed
        # the same state up to a phase
        assert abs(abs(np.vdot(v_ref, vec)) / np.linalg.norm(vec) - 1.0) <= 1e-9, mixed
        if mixed:
            assert info["matvecs_fp32"] > 0


def _energy(sw, psi):
    return (np.vdot(psi, sw.matvec(psi)) / np.vdot(psi, psi)).real


def _sw_for(grown, J=None, E=None):
    rules, phys, W, E0, sig = models.coupled_chain_model(d=D, seed=2, Jperp=grown["J"] if J is None else J, E=E)
    return dmrg.StagewiseHeff(grown["LB"], W, grown["RB"], grown["comp"]), W


def test_d64_adjoint(ctx, grown):
    print("case d64 adjoint chi=16", flush=True)
    import chainamps_amd as cam
    W, LB, RB, J, Elev = grown["W"], grown["LB"], grown["RB"], grown["J"], grown["E"]
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    m = W.matrix.tocsc()
    m.sort_indices()
    # dE/dpsi and dE/dW for a random psi against the oracle
    rng = np.random.default_rng(9)
    psi = rng.standard_normal(H.n) + 1j * rng.standard_normal(H.n)
    dpsi, dW = H.vjp(psi, m.nnz)
    n2 = np.vdot(psi, psi).real
    Hpsi = grown["sw"].matvec(psi)
    En = np.vdot(psi, Hpsi).real / n2
    assert rel_err(dpsi, (Hpsi - En * psi) / n2) <= 1e-12
    h = 1e-5
    for k in rng.choice(m.nnz, size=6, replace=False):
        Wp, Wm = W.copy(), W.copy()
        Wp.matrix = m.copy(); Wp.matrix.data[k] += h
        Wm.matrix = m.copy(); Wm.matrix.data[k] -= h
        comp = grown["comp"]
        fd = (_energy(dmrg.StagewiseHeff(LB, Wp, RB, comp), psi) -
              _energy(dmrg.StagewiseHeff(LB, Wm, RB, comp), psi)) / (2 * h)
        assert abs(dW[k].real - fd) <= 1e-6 * max(1.0, abs(fd)), (k, dW[k], fd)
    # at the ground state
    E0, psi0, info = H.lowest(converge=True)
    dpsi, dW = H.vjp(psi0, m.nnz)
    assert np.linalg.norm(dpsi) <= 1e-9 * max(1.0, abs(E0))
    col_of = np.repeat(np.arange(m.shape[1]), np.diff(m.indptr))
    row_of = m.indices
    bl, br = row_of // D, col_of // D
    dEdJ = np.sum(dW * np.where((bl == 1) & (br == 0), m.data / J, 0))
    assert abs(dEdJ.imag) <= 1e-10
    dEdJ = dEdJ.real
    hj = 1e-3        # E(psi; J) is linear in J at fixed psi: exact difference
    hf = (_energy(_sw_for(grown, J=J + hj)[0], psi0) - _energy(_sw_for(grown, J=J - hj)[0], psi0)) / (2 * hj)
    assert abs(dEdJ - hf) <= 1e-9 * max(1.0, abs(hf)), (dEdJ, hf)

    def lowest(Jv):
        Wj = _sw_for(grown, J=Jv)[1]
        return cam.Heff(to_lib(ctx, LB), to_lib(ctx, Wj), to_lib(ctx, RB)).lowest(converge=True)[0]
    fd = (lowest(J + h) - lowest(J - h)) / (2 * h)
    assert abs(dEdJ - fd) <= 1e-6 * max(1.0, abs(fd)), (dEdJ, fd)
    # dE/dE_n: E(psi; E) is affine in the level energies, so the gradient
    # along any direction u is the exact central difference
    lev = np.zeros(D)
    for nlev in range(D):
        k = np.where((bl == 2) & (br == 0) & (row_of % D == nlev) & (col_of % D == nlev))[0]
        assert len(k) == 1
        lev[nlev] = dW[k[0]].real
    for trial in range(3):
        u = rng.standard_normal(D)
        ref = (_energy(_sw_for(grown, E=Elev + hj * u)[0], psi0) -
               _energy(_sw_for(grown, E=Elev - hj * u)[0], psi0)) / (2 * hj)
        assert abs(lev @ u - ref) <= 1e-9 * max(1.0, abs(ref)), (trial, lev @ u, ref)
    for nlev in (0, 1, 31, 63):
        u = np.zeros(D)
        u[nlev] = 1.0
        ref = (_energy(_sw_for(grown, E=Elev + hj * u)[0], psi0) -
               _energy(_sw_for(grown, E=Elev - hj * u)[0], psi0)) / (2 * hj)
        assert abs(lev[nlev] - ref) <= 1e-9 * max(1.0, abs(ref)), (nlev, lev[nlev], ref)


def _one_vertex_op(rules, phys, M):
    """A one-vertex operator (MPO bond dimension 1) from a d x d matrix."""
    from oracle.mpx import MPX
    from oracle.qn import MPXIndex
    z = np.zeros((1, len(rules.rules)), dtype=np.int64)
    T = np.zeros((D, 1, D, 1), complex)
    T[:, 0, :, 0] = M
    idx = [MPXIndex(True, phys, True), MPXIndex(True, z), MPXIndex(False, phys, True), MPXIndex(False, z)]
    return MPX.from_dense_tensor(rules, idx, 2, T)


def test_d64_itebd_quench(ctx):
    print("case d64 itebd quench", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=D, seed=0, Jperp=0.1)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 0)
    P0 = _one_vertex_op(rules, phys, np.diag((np.arange(D) == 0).astype(float)))
    P1 = _one_vertex_op(rules, phys, np.diag((np.arange(D) == 1).astype(float)))
    S2 = _one_vertex_op(rules, phys, sig @ sig)
    ops_ = [P0, P1, S2]
    dt, steps, chi, interval = 0.01, 10, 16, 5
    ops = tebd.trotter_operators(W, dt, 2)
    print("  oracle gates done", flush=True)
    cell, tr, recs = tebd.evolve_measured(ops, 2, [A0, A1, l0, l1], steps, chi, 1e-12, interval, ops_)
    print("  oracle evolution done", flush=True)
    dev = cam.ITEBD(to_lib(ctx, W), to_lib(ctx, A0), to_lib(ctx, A1), l0, l1, dt, 2)
    got = dev.evolve_measured(steps, chi, 1e-12, interval, [to_lib(ctx, o) for o in ops_])
    assert [g["step"] for g in got] == [r["step"] for r in recs] == [5, 10]
    for g, r in zip(got, recs):
        assert abs(g["overlap"] - r["overlap"]) <= 1e-8, (g, r)
        assert abs(g["entropy"] - r["entropy"]) <= 1e-8
        for a, b in zip(g["values"], r["values"]):
            assert abs(a - b) <= 1e-8, (g["values"], r["values"])
    # the quench moves weight out of level 0 and builds entanglement
    assert recs[-1]["values"][0].real < 1.0 - 1e-6 and recs[-1]["entropy"] > 1e-6

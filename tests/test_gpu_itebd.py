"""GPU: device iTEBD evolution (TrotterDecomposition + iTEBD::evolve,
TEBD_routines.cpp:14-185) against the oracle's evolve restatement.

From the same product state, with the same gates and truncation, the bond
weights agree step by step (1e-10), and so do the truncations and the
gauge-invariant two-site tensor A1 (x) A0 over the inner bond.  Without
truncation the evolution is unitary: sum lambda^2 stays 1.
"""
import numpy as np
import pytest

from conftest import to_lib, from_lib, rel_err
from oracle import models, tebd
from oracle import mpx as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [1, 2])
def test_itebd_evolve_vs_oracle(ctx, order):
    print(f"case itebd_evolve order={order}", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=7, Jperp=0.4)
    dt, D, minS = 0.05, 12, 1e-12
    ops = tebd.trotter_operators(W, dt, order)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 0)
    dev = cam.ITEBD(to_lib(ctx, W), to_lib(ctx, A0), to_lib(ctx, A1), l0, l1, dt, order)
    cell = [A0, A1, l0, l1]
    for chunk in range(3):
        cell, tr = tebd.evolve(ops, order, cell, 2, D, minS)
        rec = dev.evolve(2, D, minS)
        assert rec["steps"] == 2 * (chunk + 1)
        assert abs(rec["time"] - dt * rec["steps"]) < 1e-14
        for k in (0, 1):
            got = dev.lam(k)
            assert len(got) == len(cell[2 + k]), (chunk, k)
            assert np.allclose(got, cell[2 + k], rtol=1e-10, atol=1e-12), (chunk, k)
        assert abs(rec["truncation"] - tr) <= 1e-10
        # singular vectors carry a free phase per bond state (rocSOLVER and
        # LAPACK choose differently), so compare the two-site reduced density
        # matrix sum_{a,b} T[s0,a,s1,b] conj(T[s0',a,s1',b]) of T = A0 A1,
        # which is invariant under any unitary on either bond
        g0, g1 = from_lib(dev.tensor(0), rules), from_lib(dev.tensor(1), rules)
        T = O.contract(g0, False, g1, False, [(2, 1)]).to_dense_tensor()
        Tr = O.contract(cell[0], False, cell[1], False, [(2, 1)]).to_dense_tensor()
        rho = np.einsum("iajb,kalb->ijkl", T, T.conj())
        rho_ref = np.einsum("iajb,kalb->ijkl", Tr, Tr.conj())
        assert rel_err(rho, rho_ref) <= 1e-9, chunk


def test_itebd_unitary_without_truncation(ctx):
    print("case itebd_unitary", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=3, seed=8, Jperp=0.3)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 0)
    dev = cam.ITEBD(to_lib(ctx, W), to_lib(ctx, A0), to_lib(ctx, A1), l0, l1, 0.02, 2)
    # one second-order step is three gates: the bond dimension stays <= 27
    rec = dev.evolve(1, 0, 0.0)
    assert rec["truncation"] == 0.0
    assert abs(np.sum(dev.lam(0) ** 2) - 1.0) < 1e-12
    assert rec["gate_flops"] > 0


def test_itebd_order_four_matches_oracle(ctx):
    """Fourth order (11 pieces per step, OrthogonaliseInversionSymmetric
    after every even-bond piece; TEBD_routines.cpp:52-85, 194-224): the
    cell's two lambdas follow the oracle (1e-8), and so do the values of
    the inversion-symmetric canonisation of the final cell."""
    print("case itebd order 4", flush=True)
    import chainamps_amd as cam
    from oracle import ortho
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=2, Jperp=0.4)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 0)
    dt, D, steps = 0.1, 24, 2
    ops = tebd.trotter_operators(W, dt, 4)
    cell, tr = tebd.evolve(ops, 4, [A0, A1, l0, l1], steps, D, 1e-12)
    dev = cam.ITEBD(to_lib(ctx, W), to_lib(ctx, A0), to_lib(ctx, A1), l0, l1, dt, 4)
    rec = dev.evolve(steps, D, 1e-12)
    for k in (0, 1):
        got, want = np.sort(dev.lam(k))[::-1], np.sort(cell[2 + k])[::-1]
        assert got.shape == want.shape, k
        assert np.allclose(got, want, rtol=1e-8, atol=1e-11), k
    # OrthogonaliseInversionSymmetric alone, on the oracle's final cell
    mats, lams = [cell[0], cell[1]], [cell[2], cell[3]]
    new, nl, info = ortho.orthogonalise_inversion_symmetric(mats, lams)
    c = cam.UnitCell.create(ctx, [to_lib(ctx, m) for m in mats], lams)
    o, dinfo = c.orthogonalise(inversion_symmetric=True)
    assert np.allclose(np.sort(o.lam(0)), np.sort(nl[0]), rtol=1e-8, atol=1e-11)
    assert abs(dinfo["entropy"] - info["entropy"]) < 1e-9


@pytest.mark.parametrize("order", [1, 2, 4])
def test_itebd_measurements_match_oracle(ctx, order):
    """Measurement steps (TEBD_routines.cpp:99-231): canonisation and
    do_measurements at steps 2 and 4 -- the overlap with the initial cell
    (a bra != ket transfer-matrix eigenvalue, 1e-8), entropy and truncation
    (1e-8) and the occupation of level 0 (OneVertexMeasurement, 1e-8) -- and
    the closed form exp(-2 i E_0 t) without the coupling (1e-10)."""
    print(f"case itebd measurements order {order}", flush=True)
    import chainamps_amd as cam
    from test_oracle_tebd import _level_projector
    for Jp in (0.0, 0.4):
        rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=2, Jperp=Jp)
        A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 0)
        P0 = _level_projector(rules, phys, 0)
        dt = 0.1
        ops = tebd.trotter_operators(W, dt, order)
        cell, tr, recs = tebd.evolve_measured(ops, order, [A0, A1, l0, l1], 4, 16, 1e-12, 2, [P0])
        dev = cam.ITEBD(to_lib(ctx, W), to_lib(ctx, A0), to_lib(ctx, A1), l0, l1, dt, order)
        got = dev.evolve_measured(4, 16, 1e-12, 2, [to_lib(ctx, P0)])
        assert [g["step"] for g in got] == [r["step"] for r in recs]
        for g, r in zip(got, recs):
            assert abs(g["overlap"] - r["overlap"]) <= 1e-8, (Jp, g, r)
            assert abs(g["entropy"] - r["entropy"]) <= 1e-8
            assert abs(g["values"][0] - r["values"][0]) <= 1e-8
            if Jp == 0.0:
                assert abs(g["overlap"] - np.exp(-2j * E[0] * g["time"])) <= 1e-10

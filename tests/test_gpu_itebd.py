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


def test_itebd_rejects_order_four(ctx):
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=3, seed=8, Jperp=0.3)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 0)
    with pytest.raises(cam.CamError):
        cam.ITEBD(to_lib(ctx, W), to_lib(ctx, A0), to_lib(ctx, A1), l0, l1, 0.02, 4)

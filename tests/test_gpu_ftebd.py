"""GPU: device finite-chain TEBD (TEBD::evolve with apply_to_odd/even_bonds and
left_canonise, TEBD_routines.cpp:273-791) against the oracle's restatement.

Site tensors carry a free phase per bond state (rocSOLVER and LAPACK choose
differently), so the comparison is on what the reference fixes: the full
state vector (its global phase is set by the traces the reference folds into
the end sites), the recorded entropy of bond N/2, the overlap with the
initial state and the largest truncation of each step.  Without truncation
the oracle equals the dense product of bond exponentials (pinned in
test_oracle_ftebd.py); with truncation both keep the same Schmidt values.
"""
import numpy as np
import pytest

from conftest import to_lib, from_lib
from oracle import ftebd, models

pytestmark = pytest.mark.gpu


def _run(ctx, order, N, D, steps, dt=0.07, seed=5, J=0.6, interval=1):
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=seed, Jperp=J)
    levels = [0, 1, 2, 3, 0, 1, 3, 2][:N]
    sites = ftebd.product_state(rules, phys, levels)
    ref = ftebd.FiniteTEBD(W, sites, dt, order)
    ref_recs = ref.evolve(steps, D=D, minS=0.0, interval=interval)
    dev = cam.FTEBD(to_lib(ctx, W), [to_lib(ctx, s) for s in sites], dt, order)
    dev.evolve(steps, D=D, minS=0.0, interval=interval)
    got_sites = [from_lib(dev.site(v), rules) for v in range(1, N + 1)]
    return ref, ref_recs, dev, dev.records(), got_sites


@pytest.mark.parametrize("order,N", [(1, 4), (2, 4), (4, 4), (2, 6), (1, 2), (2, 2)])
def test_ftebd_untruncated_vs_oracle(ctx, order, N):
    print(f"case ftebd untruncated order={order} N={N}", flush=True)
    ref, rrec, dev, drec, sites = _run(ctx, order, N, D=0, steps=3)
    psi_ref = ftebd.state_vector(ref.left[1:])
    psi = ftebd.state_vector(sites)
    assert np.allclose(psi, psi_ref, atol=1e-11), np.max(np.abs(psi - psi_ref))
    assert len(drec) == len(rrec) == 3
    for a, b in zip(drec, rrec):
        assert a["step"] == b["step"] and abs(a["time"] - b["time"]) < 1e-15
        assert abs(a["entropy"] - b["entropy"]) < 1e-10
        assert abs(a["overlap"] - b["overlap"]) < 1e-11
        assert abs(a["abs_overlap"] - abs(b["overlap"])) < 1e-11
        assert a["truncation"] < 1e-20
    # left canonical on the device too
    for A in sites:
        T = A.to_dense_tensor()
        M = T.reshape(-1, T.shape[2])
        assert np.allclose(M.conj().T @ M, np.eye(M.shape[1]), atol=1e-12)


@pytest.mark.parametrize("order", [2, 4])
def test_ftebd_truncated_vs_oracle(ctx, order):
    print(f"case ftebd truncated order={order}", flush=True)
    ref, rrec, dev, drec, sites = _run(ctx, order, 6, D=3, steps=3, dt=0.3, seed=9, J=1.5)
    psi_ref = ftebd.state_vector(ref.left[1:])
    psi = ftebd.state_vector(sites)
    assert np.allclose(psi, psi_ref, atol=1e-9), np.max(np.abs(psi - psi_ref))
    for a, b in zip(drec, rrec):
        assert abs(a["truncation"] - b["truncation"]) <= 1e-10 + 1e-8 * b["truncation"]
        assert abs(a["entropy"] - b["entropy"]) < 1e-9
    assert rrec[-1]["truncation"] > 1e-8


def test_ftebd_measurement_interval(ctx):
    print("case ftebd interval", flush=True)
    ref, rrec, dev, drec, sites = _run(ctx, 2, 4, D=0, steps=5, interval=2)
    assert [r["step"] for r in drec] == [r["step"] for r in rrec] == [2, 4]
    assert np.allclose(ftebd.state_vector(sites), ftebd.state_vector(ref.left[1:]), atol=1e-11)


def test_ftebd_rejects_bad_input(ctx):
    print("case ftebd errors", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=1, Jperp=0.3)
    sites = ftebd.product_state(rules, phys, [0, 1, 2])
    with pytest.raises(cam.CamError, match="even number"):
        cam.FTEBD(to_lib(ctx, W), [to_lib(ctx, s) for s in sites], 0.1, 2)
    sites = ftebd.product_state(rules, phys, [0, 1])
    with pytest.raises(cam.CamError, match="4th order"):
        cam.FTEBD(to_lib(ctx, W), [to_lib(ctx, s) for s in sites], 0.1, 3)

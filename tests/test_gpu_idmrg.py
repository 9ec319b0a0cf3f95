"""GPU: device iDMRG growth (iDMRG::run / SuperBlock::grow_two_vertex,
DMRG_routines.cpp:309-374) against the oracle iDMRG.

With converged eigensolves in both, energies, entanglement entropies and
kept bond dimensions agree step by step (1e-10); the reference's own
spin_half_example fixture (README.md:31-40) is reproduced with the
reference's growth-step eigensolver setting (converge=0).
"""
import json
import os

import numpy as np
import pytest

from conftest import to_lib
from oracle.qn import ChargeRules
from oracle import models, dmrg

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_idmrg_heisenberg_vs_oracle(ctx):
    print("case idmrg_heisenberg", flush=True)
    import chainamps_amd as cam
    rules, phys, W = models.heisenberg_mpo()
    ref = dmrg.IDMRG(W, converge=True)
    dev = cam.IDMRG(to_lib(ctx, W), converge=True)
    for step in range(8):
        r = ref.grow(16, 1e-12)
        g = dev.grow(16, 1e-12)
        assert g["N"] == r["N"]
        assert abs(g["energy"] - r["energy"]) <= 1e-10 * max(1, abs(r["energy"])), (step, g, r)
        assert g["chi"] == r["chi"], (step, g["chi"], r["chi"])
        assert abs(g["entropy"] - r["entropy"]) <= 1e-9, (step, g["entropy"], r["entropy"])
        assert np.allclose(np.sort(dev.values()), np.sort(ref.values), atol=1e-9)
        # 1 - fidelity of the prediction (CheckConvergence), the Energies.dat column
        assert abs(g["fidelity"] - r["fidelity"]) <= 1e-8, (step, g["fidelity"], r["fidelity"])


def test_idmrg_spin_half_fixture(ctx):
    print("case idmrg_spin_half_fixture", flush=True)
    import chainamps_amd as cam
    g = json.load(open(os.path.join(GOLD, "spin_half_example.json")))
    rules = ChargeRules([0])
    phys = np.array([[b[1]] for b in g["basis"]])
    d = len(phys)
    ops = {}
    for name, trip in g["operators"].items():
        m = np.zeros((d, d), complex)
        for r_, c_, re_, im_ in trip:
            m[r_, c_] = re_ + 1j * im_
        ops[name] = m
    local = [(c["value"], ops[c["ops"][0]]) for c in g["couplings"] if len(c["ops"]) == 1]
    coup = [(c["value"], ops[c["ops"][0]], ops[c["ops"][1]]) for c in g["couplings"] if len(c["ops"]) == 2]
    W = models.general_hmpo(rules, phys, np.zeros((d, d)), local, coup)
    dev = cam.IDMRG(to_lib(ctx, W), converge=False)
    for (index, energy, per_vertex, entropy) in g["growth_steps"]:
        rec = dev.grow(g["chi"])
        assert abs(rec["energy"] - energy) < 1e-10, (index, rec, energy)
        assert abs(rec["energy_per_vertex"] - per_vertex) < 1e-10
        assert abs(rec["entropy"] - entropy) < 1e-9, (index, rec["entropy"], entropy)


def test_idmrg_coupled_chain_mixed(ctx):
    """Z2 d=8 coupled chain: FP32-mixed eigensolves reach the FP64 energies."""
    print("case idmrg_mixed", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=8, seed=1, Jperp=0.3)
    a = cam.IDMRG(to_lib(ctx, W), converge=True, mixed=False)
    b = cam.IDMRG(to_lib(ctx, W), converge=True, mixed=True)
    for step in range(5):
        ra, rb = a.grow(24, 1e-12), b.grow(24, 1e-12)
        assert abs(ra["energy"] - rb["energy"]) <= 1e-10 * max(1, abs(ra["energy"])), (step, ra, rb)
    assert rb["matvecs_fp32"] > 0


def test_idmrg_growth_setting_mixed_vs_fp64_d64(ctx, capfd):
    """The production growth step (converge=0: one Krylov cycle of ncv 20
    at tol 0.1, the predicted start vector) in FP32-mixed precision against
    FP64, on the d=64 Z2 coupled chain at chi = 64.  From N = 4 on the two
    QN blocks of theta are 2048 x 2048 with at most 64 values kept, so the
    leading-values SVD runs inside the driver (checked in its trace).  The
    FP32 Krylov basis differs from the FP64 one by rounding, so the loose
    solves agree to their tolerance's second order, not to 1e-10: energies
    within 1e-7 relative, entropies within 1e-5, the same kept bond."""
    print("case idmrg growth setting mixed vs fp64, d=64 chi=64", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=64, seed=0, Jperp=0.1)
    a = cam.IDMRG(to_lib(ctx, W), converge=False, mixed=False)
    b = cam.IDMRG(to_lib(ctx, W), converge=False, mixed=True)
    old = os.environ.get("CAM_SVD_TRACE")
    os.environ["CAM_SVD_TRACE"] = "1"
    try:
        for step in range(4):
            ra, rb = a.grow(64, 1e-12), b.grow(64, 1e-12)
            print(f"  N={ra['N']} E {ra['energy']:.12f} / {rb['energy']:.12f} S {ra['entropy']:.9f} / "
                  f"{rb['entropy']:.9f} chi {ra['chi']}/{rb['chi']} fp32 matvecs {rb['matvecs_fp32']}", flush=True)
            assert abs(ra["energy"] - rb["energy"]) <= 1e-7 * max(1.0, abs(ra["energy"])), (step, ra, rb)
            assert abs(ra["entropy"] - rb["entropy"]) <= 1e-5, (step, ra, rb)
            assert ra["chi"] == rb["chi"]
    finally:
        if old is None:
            del os.environ["CAM_SVD_TRACE"]
        else:
            os.environ["CAM_SVD_TRACE"] = old
    err = capfd.readouterr().err
    assert "leading" in err, err[-2000:]
    assert rb["matvecs_fp32"] > 0
    a.close()
    b.close()

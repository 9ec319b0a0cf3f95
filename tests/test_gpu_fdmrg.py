"""GPU: finite DMRG sweeps (SuperBlock::move_right/left_two_vertex,
DMRG_routines.cpp:376-444; FiniteDMRG::run, fDMRG_routines.cpp:19-66) with
the finite predictions (MakeL/RFinitePrediction, 709-716) and the finite
CheckConvergence (773-776), all on the device.

* the reference's own fixture: fDMRG_DRV -B 50 -F 2 on the ten-vertex
  spin_half_example (README.md:29-40), midpoint energies -9 and entropies
  to 1e-10;
* step by step against the oracle's sweep: centre position, energy,
  entropy and truncation (1e-10), 1 - |<prediction|psi>| (1e-8), on the
  spin-half chain (dense route) and on a truncated d=8 coupled chain
  (Lanczos route, converged eigensolves, FP64 and mixed precision);
* growth after a sweep and a move past an end are errors.
"""
import json
import os

import numpy as np
import pytest

from conftest import to_lib
from oracle import models, dmrg
from test_oracle_dmrg import spin_half_example_mpo

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _compare(dev_steps, ref_steps, tol_e=1e-10):
    assert len(dev_steps) == len(ref_steps)
    for k, (g, r) in enumerate(zip(dev_steps, ref_steps)):
        print(f"  step {k} ({g['left']},{g['right']}) E {g['energy']:.14f} / {r['energy']:.14f} chi {g['chi']}/"
              f"{r['chi']} S {g['entropy']:.12f}/{r['entropy']:.12f} tr {g['truncation']:.3e}/{r['truncation']:.3e}"
              f" fid {g['fidelity']:.3e}/{r['fidelity']:.3e}", flush=True)
        assert (g["left"], g["right"]) == (r["left"], r["right"]), k
        # without truncation the kept count includes the numerically zero
        # singular values of the block SVDs, which depend on rounding (the
        # entropy, compared below, fixes every weight above them)
        if r["truncation"] > 0:
            assert g["chi"] == r["chi"], (k, g["chi"], r["chi"])
        assert abs(g["energy"] - r["energy"]) <= tol_e * max(1.0, abs(r["energy"])), (k, g["energy"], r["energy"])
        assert abs(g["entropy"] - r["entropy"]) <= 1e-10, (k, g["entropy"], r["entropy"])
        assert abs(g["truncation"] - r["truncation"]) <= 1e-10, (k, g["truncation"], r["truncation"])
        assert abs(g["fidelity"] - r["fidelity"]) <= 1e-8, (k, g["fidelity"], r["fidelity"])


def test_spin_half_example_sweeps(ctx):
    print("case fdmrg spin_half_example", flush=True)
    import chainamps_amd as cam
    g = json.load(open(os.path.join(GOLD, "spin_half_example.json")))
    W = spin_half_example_mpo(g)
    ref = dmrg.IDMRG(W, converge=False)
    dev = cam.IDMRG(to_lib(ctx, W), converge=False)
    for _ in g["growth_steps"]:
        ref.grow(g["chi"])
        dev.grow(g["chi"])
    steps, mid = dev.finite_sweeps(g["sweeps"], g["chi"])
    rsteps, rmid = ref.finite_sweeps(g["sweeps"], g["chi"])
    assert len(mid) == len(g["finite_sweep_rows"])
    for r, (index, energy, per_vertex, entropy, trunc, fid) in zip(mid, g["finite_sweep_rows"]):
        assert (r["left"], r["right"]) == (4, 4)
        assert abs(r["energy"] - energy) < 1e-10
        assert abs(r["energy_per_vertex"] - per_vertex) < 1e-10
        assert abs(r["entropy"] - entropy) < 1e-10
    _compare(steps, rsteps)
    with pytest.raises(cam.CamError):
        dev.grow(g["chi"])
    dev.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_truncated_coupled_chain_sweeps(ctx, mixed):
    """A d=8 Z2 coupled chain of 8 vertices with chi=10: truncated finite
    steps on the Lanczos route; converged eigensolves so that the device and
    the oracle agree step by step (the reference's loose tol-0.1 iterates
    are not comparable between solvers)."""
    print(f"case fdmrg coupled chain mixed={mixed}", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=8, seed=3, Jperp=0.3)
    chi, trunc = 10, 1e-10
    ref = dmrg.IDMRG(W, converge=True)
    dev = cam.IDMRG(to_lib(ctx, W), converge=True, mixed=mixed)
    for _ in range(4):
        ref.grow(chi, trunc)
        dev.grow(chi, trunc)
    steps, mid = dev.finite_sweeps(1, chi, trunc)
    rsteps, rmid = ref.finite_sweeps(1, chi, trunc)
    assert max(r["n"] for r in rsteps) > 1000          # the Lanczos route is taken
    assert any(r["truncation"] > 0 for r in rsteps)
    _compare(steps, rsteps, tol_e=1e-10)
    assert len(mid) == len(rmid) == 2
    dev.close()


def test_single_moves_and_end_errors(ctx):
    """cam_idmrg_move one step at a time reaches the right end, refuses to
    go further, and the blocks it leaves are those of the new centre."""
    print("case fdmrg single moves", flush=True)
    import chainamps_amd as cam
    rules, phys, W = models.heisenberg_mpo()
    dev = cam.IDMRG(to_lib(ctx, W), converge=True)
    for _ in range(3):
        dev.grow(16)
    for k in range(2):
        r = dev.move(True, 16)
        assert (r["left"], r["right"]) == (3 + k, 1 - k)
    with pytest.raises(cam.CamError):
        dev.move(True, 16)
    # centre at the right end: a left block of four vertices (bond 16) and
    # the dummy right block (bond 1)
    LB, RB = dev.blocks()
    assert LB.shape()[0] == 6 and RB.shape()[0] == 6
    assert 1 < LB.indices()[5].charges.shape[0] <= 16
    assert RB.indices()[4].charges.shape[0] == 1
    dev.close()

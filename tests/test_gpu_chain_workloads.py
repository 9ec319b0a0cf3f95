"""GPU: the plain iDMRG configuration at its stated size and the U(1)
batching workload, against the oracle where it can run and through
size-independent properties at full size.

* Heisenberg chain (Jxy = Jz = 1, U(1), MPO bond dimension 5), chi = 100,
  20 iDMRG iterations: energy, energy per vertex, truncation, entropy and
  kept bond dimension of every growth step against the oracle's run (both
  with converged eigensolves, 1e-10); the iDMRG driver's measurement after
  the two warm-up steps (Orthogonalise + iTwoVertexEnergy of the central
  unit cell, iDMRG_DRV.cpp:46-82) approaches the Bethe-ansatz energy per
  vertex 1/4 - ln 2 (Hulthen 1938);
* batching workload (21 U(1) sectors, d = 32, w = 6, seed 0) at chi = 96
  with the same generator: FP64 matvec within 1e-12, FP32 within 2e-5 of
  the oracle's reference-order matvec, sliced too;
* the same workload at chi = 1024: linearity, FP32 against FP64 and one
  a0' row of the output against an independent host contraction (1e-12).
"""
import numpy as np
import pytest

from conftest import rel_err, to_lib
from oracle import models, dmrg

pytestmark = pytest.mark.gpu

HULTHEN = 0.25 - np.log(2.0)


def test_heisenberg_chi100_20_iterations(ctx):
    print("case heisenberg chi=100 x 20", flush=True)
    import chainamps_amd as cam
    rules, phys, W = models.heisenberg_mpo()
    ref = dmrg.IDMRG(W, converge=True)
    dev = cam.IDMRG(to_lib(ctx, W), converge=True)
    Wt = to_lib(ctx, W)
    epv = []
    for step in range(20):
        r = ref.grow(100)
        g = dev.grow(100)
        assert g["N"] == r["N"] == 2 * (step + 1)
        assert abs(g["energy"] - r["energy"]) <= 1e-10 * max(1.0, abs(r["energy"])), (step, g, r)
        assert abs(g["energy_per_vertex"] - r["energy_per_vertex"]) <= 1e-10, step
        assert abs(g["truncation"] - r["truncation"]) <= 1e-10, (step, g["truncation"], r["truncation"])
        assert abs(g["entropy"] - r["entropy"]) <= 1e-9, step
        if r["truncation"] > 0:
            assert g["chi"] == r["chi"], (step, g["chi"], r["chi"])
        if step >= 2:
            # iDMRG_DRV: Orthogonalise(central cell) and iTwoVertexEnergy
            cell = dev.unitcell()
            oc, info = cell.orthogonalise()
            e, _ = oc.energy(Wt)
            epv.append(e.real)
            oc.close()
            cell.close()
        print(f"  step {step} N={g['N']} chi={g['chi']} E/N={g['energy_per_vertex']:.12f} trunc={g['truncation']:.3e}"
              + (f" orthogonalised E/vertex={epv[-1]:.12f}" if step >= 2 else ""), flush=True)
    assert ref.history[-1]["truncation"] > 0          # chi = 100 truncates by N = 40
    assert abs(epv[-1] - HULTHEN) <= 1e-4, (epv[-1], HULTHEN)
    assert abs(epv[-1] - HULTHEN) < abs(epv[0] - HULTHEN)
    dev.close()


def _batching_oracle(chi):
    from chainamps_amd import workloads
    from oracle.mpx import MPX
    from oracle.qn import ChargeRules, MPXIndex
    rules, *parts = workloads.u1_batching_dense(chi)
    R = ChargeRules(rules)
    return [MPX.from_dense_tensor(R, [MPXIndex(i.inward, np.asarray(i.charges), i.physical) for i in idx], nr, T)
            for idx, T, nr in parts]


def test_u1_batching_chi96_vs_oracle(ctx):
    print("case u1 batching chi=96 vs oracle", flush=True)
    import chainamps_amd as cam
    from chainamps_amd import workloads
    LB, W, RB = _batching_oracle(96)
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    H = cam.Heff(*workloads.u1_batching_inputs(ctx, chi=96))
    assert H.n == comp.n
    assert np.array_equal(H.allowed(), comp.allowed)
    rng = np.random.default_rng(6)
    v = rng.standard_normal(comp.n) + 1j * rng.standard_normal(comp.n)
    ref = comp.matvec(v)
    assert rel_err(H.apply(v), ref) <= 1e-12
    assert rel_err(H.apply(v, cam.CAM_FP32), ref) <= 2e-5
    H.set_slices(4, 2)
    assert rel_err(H.apply(v), ref) <= 1e-12


def test_u1_batching_chi1024_full_size(ctx):
    """chi = 1024 (the workload the benchmark's batching entry times)."""
    print("case u1 batching chi=1024 full size", flush=True)
    import torch
    import chainamps_amd as cam
    from chainamps_amd import workloads
    from conftest import from_lib
    chi = 1024
    L, W, R = workloads.u1_batching_inputs(ctx, chi=chi)
    H = cam.Heff(L, W, R)
    n = H.n
    print(f"  n = {n}, {H.flops:.3e} flops per matvec", flush=True)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Hx, Hy = H.apply(x), H.apply(y)                  # the first call autotunes
    al, be = 0.7 + 0.2j, -1.3 + 0.4j
    assert rel_err(H.apply(al * x + be * y), al * Hx + be * Hy) <= 1e-13
    e32 = rel_err(H.apply(x, cam.CAM_FP32), Hx)
    print(f"  fp32 vs fp64 {e32:.3e}; launch fp64 {H.get_launch(0)}", flush=True)
    assert e32 <= 2e-5
    # one a0' row against numpy tensordots over the dense L, W, R and psi
    d = 32
    Ld = from_lib(L, None).to_dense_tensor()[0, :, 0, :, 0, :]
    Rd = from_lib(R, None).to_dense_tensor()[:, 0, :, 0, :, 0]
    Wd = from_lib(W, None).to_dense_tensor()
    allowed = H.allowed()
    psi = np.zeros(d * chi * d * chi, np.complex128)
    psi[allowed] = x
    psi = psi.reshape((d, chi, d, chi), order="F")
    a0p = 517
    t1 = np.tensordot(Ld[a0p], psi, axes=([1], [1]))
    del psi
    t2 = np.tensordot(t1, Wd, axes=([0, 1], [1, 2]))
    t3 = np.tensordot(t2, Wd, axes=([3, 0], [1, 2]))
    row = np.tensordot(t3, Rd, axes=([0, 3], [2, 1]))
    sel = np.nonzero((allowed // d) % chi == a0p)[0]
    f = allowed[sel]
    s1 = f % d
    rest = f // (d * chi)
    got, want = Hx[sel], row[s1, rest % d, rest // d]
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"  row a0'={a0p}: {len(sel)} entries, relative error {err:.3e}", flush=True)
    assert len(sel) > 0 and err <= 1e-12
    H.close()

"""GPU: gate construction (TEBD_routines.cpp:14-85, 826-880) against the oracle.

* ExtractMPOBlock (MPO_matrix.cpp:69-85): identical sub-MPO, bit-exact;
* odd / even bond Hamiltonians: identical index charges, values within
  1e-14 (they are sums of products of MPO entries);
* exp(-i tau H_bond) per charge block (Exponentiate over SparseMatrix::HED):
  within 1e-12 of the oracle, whose own pin is scipy.linalg.expm; unitary;
  the Trotter pieces of order 2 compose to the right total time.
"""
import numpy as np
import pytest

from conftest import to_lib, from_lib, rel_err
from oracle import models, tebd
from oracle.dmrg import extract_mpo_block as ref_extract

pytestmark = pytest.mark.gpu


def _models():
    rules, phys, W, E, sig = models.coupled_chain_model(d=6, seed=5, Jperp=0.4)
    yield "chain", rules, W
    rules, phys, W = models.heisenberg_mpo()
    yield "heisenberg", rules, W


@pytest.mark.parametrize("which", ["chain", "heisenberg"])
def test_extract_and_bond_hamiltonians(ctx, which):
    print(f"case bond_hamiltonians {which}", flush=True)
    import chainamps_amd as cam
    rules, W = [(r, w) for (n, r, w) in _models() if n == which][0]
    Wd = to_lib(ctx, W)
    w = W.indices[1].size
    for rr, cc in [((w - 1, w - 1), (0, w - 1)), ((0, w - 1), (0, 0)), ((1, w - 2), (0, 0))]:
        got = from_lib(cam.extract_mpo_block(Wd, rr, cc), rules)
        ref = ref_extract(W, rr, cc)
        assert np.array_equal(got.to_dense_tensor(), ref.to_dense_tensor())
        for a, b in zip(got.indices, ref.indices):
            assert np.array_equal(a.charges, b.charges) and a.inward == b.inward
    for odd, ref in [(True, tebd.odd_bond_hamiltonian(W)), (False, tebd.even_bond_hamiltonian(W))]:
        got = from_lib(cam.bond_hamiltonian(Wd, odd), rules)
        assert got.numrows == ref.numrows == 2
        assert len(got.indices) == len(ref.indices) == 6
        for a, b in zip(got.indices, ref.indices):
            assert np.array_equal(a.charges, b.charges) and a.inward == b.inward
        assert rel_err(got.to_dense_tensor(), ref.to_dense_tensor()) <= 1e-14


@pytest.mark.parametrize("which", ["chain", "heisenberg"])
def test_bond_evolution_operator(ctx, which):
    print(f"case bond_evolution {which}", flush=True)
    import chainamps_amd as cam
    rules, W = [(r, w) for (n, r, w) in _models() if n == which][0]
    Wd = to_lib(ctx, W)
    for odd in (True, False):
        Href = tebd.odd_bond_hamiltonian(W) if odd else tebd.even_bond_hamiltonian(W)
        Hd = cam.bond_hamiltonian(Wd, odd)
        for tau in (0.01, 0.37):
            G = from_lib(cam.bond_evolution_operator(Hd, tau), rules)
            Gref = tebd.bond_evolution_operator(Href, tau)
            assert len(G.indices) == 4 and G.numrows == 2
            g, gr = G.matrix.toarray(), Gref.matrix.toarray()
            assert rel_err(g, gr) <= 1e-12, (odd, tau, rel_err(g, gr))
            assert np.allclose(g.conj().T @ g, np.eye(g.shape[0]), atol=1e-12)
    # second-order Trotter pieces: G(tau/2) G(tau/2) = G(tau)
    Hd = cam.bond_hamiltonian(Wd, True)
    half = from_lib(cam.bond_evolution_operator(Hd, 0.05), rules).matrix.toarray()
    full = from_lib(cam.bond_evolution_operator(Hd, 0.1), rules).matrix.toarray()
    assert rel_err(half @ half, full) <= 1e-12

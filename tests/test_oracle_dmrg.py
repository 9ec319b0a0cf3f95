"""Oracle pins for the two-vertex H_eff and the iDMRG growth loop.

* matvec (DMRG_routines.cpp:898-974, reference contraction order) against
  an independent stage-wise dense contraction of L, W, W, R (U(1) with w=6,
  Z2 coupled chain, Heisenberg environments);
* allowed_indices order (DMRG_routines.cpp:808-824) against a literal
  nested-loop restatement;
* iDMRG growth energies without truncation == exact diagonalisation of the
  open 2N-site Heisenberg chain;
* the reference's own fixture examples/user_defined/spin_half_example
  (README.md:31-40): growth-step energies and entanglement entropies.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.qn import ChargeRules, MPXIndex
from oracle import models, dmrg

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def dense_heff_apply(LB, W, RB, comp, v):
    L = LB.to_dense_tensor()[0, :, 0, :, 0, :]
    R = RB.to_dense_tensor()[:, 0, :, 0, :, 0]
    Wd = W.to_dense_tensor()
    d, chiL, chiR = Wd.shape[0], L.shape[0], R.shape[0]
    psi = np.zeros(comp.length, complex)
    psi[comp.allowed] = v
    psi = psi.reshape((d, chiL, d, chiR), order="F")
    t1 = np.einsum("xby,sytw->bsxtw", L, psi)
    t2 = np.einsum("pbsc,bsxtw->cpxtw", Wd, t1)
    t3 = np.einsum("qcte,cpxtw->epxqw", Wd, t2)
    o = np.einsum("zew,epxqw->pxqz", R, t3)
    return o.reshape(-1, order="F")[comp.allowed]


def _envs(kind, seed=0):
    rng = np.random.default_rng(seed)
    if kind == "u1":
        rules = ChargeRules([0])
        qs, sizes = models.sector_profile(30, -3, 3, sigma=1.5)
        aq = models.bond_charges_from_profile(qs, sizes)
        phys = rng.integers(-1, 2, 6).reshape(-1, 1)
        bq = np.array([0, 1, -1, 2, -2, 0]).reshape(-1, 1)
        W = models.random_qn_mpo(rules, phys, bq, rng)
    else:
        rules, phys, W, E, sig = models.coupled_chain_model(d=8, seed=seed)
        aq = (np.arange(20) % 2).reshape(-1, 1)
        bq = W.indices[1].charges
    LB = models.random_left_block(rules, aq, bq, rng)
    RB = models.random_right_block(rules, aq, bq, rng)
    return LB, W, RB


@pytest.mark.parametrize("kind", ["u1", "z2"])
def test_matvec_vs_dense(kind):
    LB, W, RB = _envs(kind)
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    rng = np.random.default_rng(1)
    v = rng.standard_normal(comp.n) + 1j * rng.standard_normal(comp.n)
    out = comp.matvec(v)
    ref = dense_heff_apply(LB, W, RB, comp, v)
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
    # the oracle's tensordot restatement (used for large eigensolves) agrees
    sw = dmrg.StagewiseHeff(LB, W, RB, comp).matvec(v)
    assert np.linalg.norm(sw - ref) <= 1e-13 * np.linalg.norm(ref)


def test_allowed_indices_literal_loop():
    LB, W, RB = _envs("u1", seed=2)
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    r = W.rules
    s1, a0, s2, a2 = comp.indices
    allowed = []
    for ca2 in range(a2.size):
        for cs2 in range(s2.size):
            acc2 = r.sub(a2.charges[ca2], s2.charges[cs2])
            for ra0 in range(a0.size):
                acc3 = r.sub(acc2, a0.charges[ra0])
                for rs1 in range(s1.size):
                    if np.array_equal(s1.charges[rs1], acc3):
                        allowed.append(rs1 + s1.size * ra0 + s1.size * a0.size * (cs2 + s2.size * ca2))
    assert np.array_equal(np.array(allowed), comp.allowed)


def heis_ed(N):
    sx = np.array([[0, 1], [1, 0]]) / 2
    sy = np.array([[0, -1j], [1j, 0]]) / 2
    sz = np.diag([0.5, -0.5])
    H = sp.csr_matrix((2 ** N, 2 ** N), dtype=complex)

    def op(o, i):
        m = sp.identity(1)
        for j in range(N):
            m = sp.kron(m, o if j == i else sp.identity(2))
        return m
    for i in range(N - 1):
        for o in (sx, sy, sz):
            H = H + op(o, i) @ op(o, i + 1)
    return spla.eigsh(H, k=1, which="SA")[0][0]


def test_idmrg_heisenberg_matches_exact_diagonalisation():
    rules, phys, W = models.heisenberg_mpo()
    idm = dmrg.IDMRG(W, converge=True)
    for N in (2, 4, 6, 8, 10):
        rec = idm.grow(1024)          # no truncation at these sizes
        assert rec["N"] == N
        assert abs(rec["energy"] - heis_ed(N)) < 1e-10


def spin_half_example_mpo(g):
    """MakeGeneralHMPO of the fixture's model (example_input couplings)."""
    rules = ChargeRules([0])
    phys = np.array([[b[1]] for b in g["basis"]])
    d = len(phys)
    ops = {}
    for name, trip in g["operators"].items():
        m = np.zeros((d, d), complex)
        for r_, c_, re_, im_ in trip:
            m[r_, c_] = re_ + 1j * im_
        ops[name] = m
    local, coup = [], []
    for c in g["couplings"]:
        if len(c["ops"]) == 1:
            local.append((c["value"], ops[c["ops"][0]]))
        else:
            coup.append((c["value"], ops[c["ops"][0]], ops[c["ops"][1]]))
    return models.general_hmpo(rules, phys, np.zeros((d, d)), local, coup)


def test_spin_half_example_fixture():
    g = json.load(open(os.path.join(GOLD, "spin_half_example.json")))
    W = spin_half_example_mpo(g)
    assert W.dims() == [2, 11, 2, 11] and W.is_consistent()
    idm = dmrg.IDMRG(W, converge=False)
    for (index, energy, per_vertex, entropy) in g["growth_steps"]:
        rec = idm.grow(g["chi"])
        assert abs(rec["energy"] - energy) < 1e-12
        assert abs(rec["energy_per_vertex"] - per_vertex) < 1e-12
        assert abs(rec["entropy"] - entropy) < 1e-12


def test_spin_half_example_finite_sweeps():
    """fDMRG_DRV -B 50 -F 2 on the ten vertices (README.md:29-40): the
    midpoint rows of the two sweeps (iterations 5-8) reproduce the
    README's energies and entanglement entropies (1e-12).  Every finite
    step's prediction is exact here (no truncation), so 1 - |<guess|psi>|
    is 0 to rounding; the README's Fidelity column reads exactly 1 for these
    rows, a value the CheckConvergence formula (DMRG_routines.cpp:773-776)
    cannot take for an exact prediction -- parity unpinned (see the
    oracle's check_convergence_finite)."""
    g = json.load(open(os.path.join(GOLD, "spin_half_example.json")))
    W = spin_half_example_mpo(g)
    idm = dmrg.IDMRG(W, converge=False)
    for _ in g["growth_steps"]:
        idm.grow(g["chi"])
    steps, mid = idm.finite_sweeps(g["sweeps"], g["chi"])
    assert len(steps) == 2 * 2 * (10 - 2) and len(mid) == len(g["finite_sweep_rows"])
    for r, (index, energy, per_vertex, entropy, trunc, fid) in zip(mid, g["finite_sweep_rows"]):
        assert (r["left"], r["right"]) == (4, 4)
        assert abs(r["energy"] - energy) < 1e-12
        assert abs(r["energy_per_vertex"] - per_vertex) < 1e-12
        assert abs(r["entropy"] - entropy) < 1e-12
        assert r["truncation"] == trunc == 0.0
    assert all(abs(r["fidelity"]) < 1e-12 for r in steps)
    # the centre visits both ends: (8,0) and (0,8)
    assert {(r["left"], r["right"]) for r in steps} >= {(8, 0), (0, 8)}


def test_finite_sweeps_match_exact_diagonalisation():
    """Without truncation every finite step of an N = 8 Heisenberg chain
    returns the exact open-chain ground energy (1e-10); with a small bond
    dimension the energies stay above it (variational) and the truncated
    steps report a non-zero discarded weight."""
    rules, phys, W = models.heisenberg_mpo()
    e8 = heis_ed(8)
    idm = dmrg.IDMRG(W, converge=True)
    for _ in range(4):
        idm.grow(1024)
    steps, mid = idm.finite_sweeps(1, 1024)
    assert len(mid) == 2
    for r in steps:
        assert abs(r["energy"] - e8) < 1e-10, r
    small = dmrg.IDMRG(W, converge=True)
    for _ in range(4):
        small.grow(4)
    steps, _ = small.finite_sweeps(1, 4, 1e-14)
    assert all(r["energy"] >= e8 - 1e-10 for r in steps)
    assert any(r["truncation"] > 0 for r in steps)


def test_heff_hermitian_for_physical_model():
    rules, phys, W = models.heisenberg_mpo()
    idm = dmrg.IDMRG(W, converge=True)
    for _ in range(4):
        idm.grow(12)
    comp = dmrg.TwoVertexComponents(dmrg.make_left_block(idm.LB, W, idm.U), W,
                                    dmrg.make_right_block(idm.RB, W, idm.Vd))
    H = comp.dense_matrix()
    assert np.allclose(H, H.conj().T, atol=1e-12)


def test_prediction_fidelity_limits():
    """CheckConvergence (DMRG_routines.cpp:769-771): while the bond is not
    truncated the prediction is exact and 1 - fidelity is 0; once it is
    truncated 1 - fidelity is positive and falls as iDMRG converges to its
    fixed point."""
    rules, phys, W = models.heisenberg_mpo()
    ref = dmrg.IDMRG(W, converge=True)
    fids = [ref.grow(16, 1e-12)["fidelity"] for _ in range(14)]
    assert abs(fids[0]) < 1e-12 and abs(fids[1]) < 1e-12
    assert all(f > 0 for f in fids[2:])
    assert fids[-1] < 0.2 * fids[2]


def test_projector_terms_shift_the_ground_state():
    """Projector terms (DMRG_routines.cpp:947-973): with P = conj(psi0) and
    weight w the operator is H_eff + w |psi0><psi0|, so its lowest
    eigenvalue is the first excited one of H_eff (dense eigh), and the
    matvec equals the dense matrix plus that rank-one term.  The eigensolve
    takes the ARPACK route even below SPARSE_THRESHOLD (DMRG_routines.cpp:852)."""
    rules, phys, W = models.heisenberg_mpo()
    idm = dmrg.IDMRG(W, converge=True)
    for _ in range(4):
        idm.grow(12)
    LB = dmrg.make_left_block(idm.LB, W, idm.U)
    RB = dmrg.make_right_block(idm.RB, W, idm.Vd)
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    Hd = comp.dense_matrix()
    w, v = np.linalg.eigh(0.5 * (Hd + Hd.conj().T))
    psi0 = v[:, 0]
    pc = dmrg.TwoVertexComponents(LB, W, RB, projectors=[(np.conj(psi0), 10.0)])
    x = np.random.default_rng(2).standard_normal(comp.n) + 0j
    assert np.allclose(pc.matvec(x), Hd @ x + 10.0 * psi0 * np.vdot(psi0, x), atol=1e-12)
    E, vec = pc.hed(converge=True)
    assert abs(E - w[1]) <= 1e-10 * max(1.0, abs(w[1]))
    assert abs(np.vdot(psi0, vec)) <= 1e-8

"""GPU parity: H_eff matvec (TwoVertexMPOMPSMultiply, DMRG_routines.cpp:898-974).

* allowed_indices (DMRG_routines.cpp:808-824) bit-exact;
* FP64 matvec within 1e-12 relative of the oracle's reference-order
  matvec (LeftPart . V . RightPart);
* FP32 split-complex matvec within 2e-5 relative (FP32 accumulation);
* Hermiticity <x,Hy> == <Hx,y> for a physical model, linearity.
"""
import numpy as np
import pytest

from conftest import to_lib, rel_err
from oracle.qn import ChargeRules, MPXIndex
from oracle import models, dmrg

pytestmark = pytest.mark.gpu


def _heisenberg_env(steps=4, chi=16):
    rules, phys, W = models.heisenberg_mpo()
    idm = dmrg.IDMRG(W, converge=True)
    for _ in range(steps):
        idm.grow(chi)
    LB = dmrg.make_left_block(idm.LB, W, idm.U)
    RB = dmrg.make_right_block(idm.RB, W, idm.Vd)
    return rules, LB, W, RB


def _random_env(seed, chi=40, d=6, w=4, qrange=3, rules=None):
    rng = np.random.default_rng(seed)
    rules = rules or ChargeRules([0])
    qs, sizes = models.sector_profile(chi, -qrange, qrange, sigma=1.5)
    aq = models.bond_charges_from_profile(qs, sizes)
    phys = rng.integers(-1, 2, d).reshape(-1, 1)
    bq = np.array([0, 1, -1, 0, 2, -2][:w]).reshape(-1, 1)
    W = models.random_qn_mpo(rules, phys, bq, rng)
    LB = models.random_left_block(rules, aq, bq, rng)
    RB = models.random_right_block(rules, aq, bq, rng)
    return rules, LB, W, RB


def _z2_env(seed, chi=24, d=8):
    rng = np.random.default_rng(seed)
    rules, phys, W, E, sig = models.coupled_chain_model(d=d, seed=seed)
    aq = (np.arange(chi) % 2).reshape(-1, 1)
    bq = W.indices[1].charges
    LB = models.random_left_block(rules, aq, bq, rng)
    RB = models.random_right_block(rules, aq, bq, rng)
    return rules, LB, W, RB


CASES = {
    "heisenberg": lambda: _heisenberg_env(),
    "random_u1": lambda: _random_env(3),
    "random_u1_w6": lambda: _random_env(4, chi=30, d=7, w=6),
    "z2_chain": lambda: _z2_env(1),
}


@pytest.mark.parametrize("case", list(CASES))
def test_heff_matvec_parity(ctx, case):
    print(f"case heff_matvec {case}", flush=True)
    import chainamps_amd as cam
    rules, LB, W, RB = CASES[case]()
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    assert H.n == comp.n
    assert np.array_equal(H.allowed(), comp.allowed)
    rng = np.random.default_rng(11)
    for trial in range(2):
        v = rng.standard_normal(comp.n) + 1j * rng.standard_normal(comp.n)
        ref = comp.matvec(v)
        out = H.apply(v)
        assert rel_err(out, ref) <= 1e-12, (case, trial)
        out32 = H.apply(v, cam.CAM_FP32)
        assert rel_err(out32, ref) <= 2e-5, (case, trial)


@pytest.mark.parametrize("algo,tile_mode,wg", [(0, 1, 4), (1, 1, 4), (1, 0, 2), (0, 2, 8), (1, 2, 1),
                                               (1, 3, 2), (0, 3, 4)])
@pytest.mark.parametrize("kc", [16, 8])
def test_heff_launch_configurations(ctx, algo, tile_mode, wg, kc):
    """Every launch configuration the autotuner can pick (four or three real
    MFMAs per complex product, each tile policy, grid sizes) gives the same
    matvec to the same bar."""
    print(f"case heff_launch algo={algo} tile_mode={tile_mode} wg={wg} kc={kc}", flush=True)
    import chainamps_amd as cam
    rules, LB, W, RB = CASES["random_u1_w6"]()
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    for prec in (cam.CAM_FP64, cam.CAM_FP32):
        H.set_launch(prec, algo, tile_mode, wg, kc)
        assert H.get_launch(prec) == [(algo, tile_mode, wg, kc)] * 4
    rng = np.random.default_rng(12)
    v = rng.standard_normal(comp.n) + 1j * rng.standard_normal(comp.n)
    ref = comp.matvec(v)
    assert rel_err(H.apply(v), ref) <= 1e-12
    assert rel_err(H.apply(v, cam.CAM_FP32), ref) <= 2e-5


@pytest.mark.parametrize("case", ["random_u1_w6", "z2_chain"])
def test_sliced_matvec(ctx, case):
    """The matvec in row x column slices of the intermediates (a0' rows and
    a2 columns of every sector cut into ranges; the column slices
    accumulate into the output) equals the oracle's to the same bar, for
    slice counts above some sector sizes (empty ranges), in FP64 and FP32,
    and the eigensolve runs on it; the workspace shrinks with the slices."""
    print(f"case heff_sliced {case}", flush=True)
    import chainamps_amd as cam
    rules, LB, W, RB = CASES[case]()
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    assert H.slices()[:2] == (1, 1)
    full = H.slices()[2]
    rng = np.random.default_rng(21)
    v = rng.standard_normal(comp.n) + 1j * rng.standard_normal(comp.n)
    ref = comp.matvec(v)
    for rs, cs in [(2, 1), (1, 2), (3, 2), (8, 8), (1, 16)]:
        H.set_slices(rs, cs)
        r, c, ws = H.slices()
        assert (r, c) == (rs, cs)
        assert ws <= full
        assert rel_err(H.apply(v), ref) <= 1e-12, (rs, cs)
        assert rel_err(H.apply(v, cam.CAM_FP32), ref) <= 2e-5, (rs, cs)
    H.set_workspace_limit(full / 5)
    r, c, ws = H.slices()
    assert ws <= full / 5 and r * c >= 5
    assert rel_err(H.apply(v), ref) <= 1e-12


def test_sliced_eigensolve(ctx):
    """The Lanczos eigensolve on a sliced matvec (a Hermitian H_eff from
    Heisenberg environments) reaches the unsliced lowest eigenvalue."""
    print("case heff_sliced_eigensolve", flush=True)
    import chainamps_amd as cam
    rules, LB, W, RB = CASES["heisenberg"]()
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    E1, _, _ = H.lowest(converge=True)
    H.set_slices(4, 2)
    for mixed in (False, True):
        E, vec, info = H.lowest(converge=True, mixed=mixed)
        assert abs(E - E1) <= 1e-10 * max(1.0, abs(E1)), (mixed, E, E1)


def test_heff_instances_share_workspace(ctx):
    """Two H_eff on one context share the grow-only intermediate buffers: a
    smaller one built first still computes correctly after a larger one has
    grown (re-allocated) them, and so does the larger one afterwards."""
    print("case heff_shared_workspace", flush=True)
    import chainamps_amd as cam
    small = CASES["heisenberg"]()
    big = CASES["random_u1_w6"]()
    rng = np.random.default_rng(13)
    Hs = cam.Heff(*(to_lib(ctx, t) for t in (small[1], small[2], small[3])))
    comp_s = dmrg.TwoVertexComponents(small[1], small[2], small[3])
    vs = rng.standard_normal(comp_s.n) + 1j * rng.standard_normal(comp_s.n)
    assert rel_err(Hs.apply(vs), comp_s.matvec(vs)) <= 1e-12
    Hb = cam.Heff(*(to_lib(ctx, t) for t in (big[1], big[2], big[3])))
    comp_b = dmrg.TwoVertexComponents(big[1], big[2], big[3])
    vb = rng.standard_normal(comp_b.n) + 1j * rng.standard_normal(comp_b.n)
    assert rel_err(Hb.apply(vb), comp_b.matvec(vb)) <= 1e-12
    assert rel_err(Hs.apply(vs), comp_s.matvec(vs)) <= 1e-12
    assert rel_err(Hs.apply(vs, cam.CAM_FP32), comp_s.matvec(vs)) <= 2e-5
    assert rel_err(Hb.apply(vb), comp_b.matvec(vb)) <= 1e-12


def test_heff_hermitian_and_linear(ctx):
    print("case heff_hermitian", flush=True)
    import chainamps_amd as cam
    rules, LB, W, RB = _heisenberg_env(steps=5, chi=24)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    rng = np.random.default_rng(2)
    x = rng.standard_normal(H.n) + 1j * rng.standard_normal(H.n)
    y = rng.standard_normal(H.n) + 1j * rng.standard_normal(H.n)
    Hx, Hy = H.apply(x), H.apply(y)
    assert abs(np.vdot(x, Hy) - np.vdot(Hx, y)) <= 1e-12 * np.linalg.norm(Hx) * np.linalg.norm(y)
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    assert rel_err(H.apply(a * x + b * y), a * Hx + b * Hy) <= 1e-13


def test_heff_flop_count(ctx):
    """The planner's algorithmic flop count equals an independent count of
    8*m*n*k over the QN-allowed block triples of the four stages."""
    print("case heff_flops", flush=True)
    import chainamps_amd as cam
    rules, LB, W, RB = _random_env(8, chi=20, d=5, w=4)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    L = LB.to_dense_tensor()[0, :, 0, :, 0, :]      # (a', b, a)
    R = RB.to_dense_tensor()[:, 0, :, 0, :, 0]      # (a2', b, a2)
    Wd = W.to_dense_tensor()                        # (s', bl, s, br)
    aq = LB.indices[5].charges[:, 0]
    a2q = RB.indices[4].charges[:, 0]
    sq = W.indices[2].charges[:, 0]
    secs = lambda q: {c: int(np.sum(q == c)) for c in np.unique(q)}
    A, A2, S = secs(aq), secs(a2q), secs(sq)
    w = Wd.shape[1]
    flops = 0.0
    # psi blocks (s1, s0, s2) with s3 = s1+s0+s2
    psi = [(s1, s0, s2, s1 + s0 + s2) for s1 in S for s0 in A for s2 in S if (s1 + s0 + s2) in A2]
    nzL = {(b, c0): c0p for b in range(w) for c0 in A for c0p in A
           if np.any(L[np.ix_(aq == c0p, [b], aq == c0)])}
    nzR = {(b, c3): c3p for b in range(w) for c3 in A2 for c3p in A2
           if np.any(R[np.ix_(a2q == c3p, [b], a2q == c3)])}
    nzW = {(bl, br, sp, s_) for bl in range(w) for br in range(w) for sp in S for s_ in S
           if np.any(Wd[np.ix_(sq == sp, [bl], sq == s_, [br])])}
    T1 = set()
    for (s1, s0, s2, s3) in psi:
        for b in range(w):
            if (b, s0) in nzL:
                s0p = nzL[(b, s0)]
                flops += 8 * A[s0p] * A[s0] * (A2[s3] * S[s1] * S[s2])
                T1.add((b, s0p, s3, s1, s2))
    T2 = set()
    for (b, s0p, s3, s1, s2) in T1:
        for (bl, br, sp, s_) in nzW:
            if bl == b and s_ == s1:
                flops += 8 * (A[s0p] * A2[s3]) * S[sp] * S[s1] * S[s2]
                T2.add((br, s0p, s3, sp, s2))
    T3 = set()
    for (b, s0p, s3, s1p, s2) in T2:
        for (bl, br, sp, s_) in nzW:
            if bl == b and s_ == s2:
                flops += 8 * (A[s0p] * A2[s3] * S[s1p]) * S[sp] * S[s2]
                T3.add((br, s0p, s3, s1p, sp))
    for (b, s0p, s3, s1p, s2p) in T3:
        if (b, s3) in nzR:
            s3p = nzR[(b, s3)]
            flops += 8 * A[s0p] * A2[s3p] * A2[s3] * S[s1p] * S[s2p]
    assert H.flops == flops

"""GPU: transfer-matrix fixed points, Orthogonalise and the energy per
vertex through the C ABI (cam_transfer_leading, cam_orthogonalise,
cam_itwovertex_energy, cam_unitcell_from_decomposition / cam_idmrg_unitcell;
measurement.cpp:180-246, 371-397, 405-629, MPX_matrix.cpp:829-913,
UnitCell.hpp:42-48) against the oracle restatement (oracle/ortho.py).

Compared quantities are those independent of the gauge freedom left by the
reference (ARPACK's eigenvector phase, the SVD's vector phases): leading
eigenvalues (relative 1e-12), the new boundary lambdas (relative 1e-8,
absolute 1e-10: a Schmidt value s from the fixed points carries an absolute
error ~ eps / s, so the smallest ones differ between two eigensolvers at
that level), the entropy (1e-10), energies per vertex (1e-10), and the
canonical conditions of the device's output cell (1e-7: X^-1 divides by
square roots of fixed-point eigenvalues down to the reference's 100 eps
cut)."""
import numpy as np
import pytest

from conftest import to_lib, from_lib
from oracle import dmrg, models, ortho
from oracle.qn import ChargeRules, MPXIndex

pytestmark = pytest.mark.gpu


def _idmrg_cell(model, chi, steps):
    if model == "heisenberg":
        rules, phys, W = models.heisenberg_mpo()
    else:
        rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=3, Jperp=0.3)
    g = dmrg.IDMRG(W, converge=True)
    for _ in range(steps):
        g.grow(chi)
    mats, lams = ortho.unit_cell_from_decomposition(g.U, g.values, g.Vd, g.prev_lambda)
    return rules, W, mats, lams, g


def _random_cell(chi_per_charge, seed, d=3):
    """A random complex U(1) cell [A0, A1] whose two boundary indices carry
    the same charges."""
    rng = np.random.default_rng(seed)
    rules = ChargeRules([0])
    phys = np.array([[1], [0], [-1]])[:d]
    q = np.concatenate([np.full(n, c) for c, n in chi_per_charge.items()]).reshape(-1, 1)
    qm = np.concatenate([np.full(n + 1, c) for c, n in chi_per_charge.items()]).reshape(-1, 1)
    A0 = models.random_qn_tensor(rules, [MPXIndex(True, phys, True), MPXIndex(True, q), MPXIndex(False, qm)], 2, rng)
    A1 = models.random_qn_tensor(rules, [MPXIndex(True, phys, True), MPXIndex(True, qm), MPXIndex(False, q)], 2, rng)
    return rules, [A0, A1], [np.ones(q.shape[0]), np.ones(qm.shape[0])]


def _cell(ctx, mats, lams):
    import chainamps_amd as cam
    return cam.UnitCell.create(ctx, [to_lib(ctx, m) for m in mats], lams)


def _check_canonical(new_cell, rules, real):
    """Left fixed point identity; for real cells also the right fixed point
    Lambda0^2.  The reference takes the right fixed point transposed
    (Y Y^H = M_R with sum conj(A) M_R A^T = l M_R, i.e. the transpose of the
    textbook sum A R A^H = l R), which is the same matrix only for real cells,
    so a complex cell's new lambdas are not its right fixed point -- in the
    reference and here alike (the lambdas above match the oracle's)."""
    A0, A1 = [from_lib(new_cell.matrix(k), rules).to_dense_tensor() for k in (0, 1)]
    lam = new_cell.lam(0)
    m = A0.shape[1]
    assert np.allclose(ortho.left_apply([A0, A1], np.eye(m)), np.eye(m), atol=1e-7)
    if not real:
        return
    A = np.einsum("pab,qbc->pqac", A0, A1).reshape(-1, m, m)
    R = np.diag(lam ** 2)
    assert np.allclose(np.einsum("sab,bc,sdc->ad", A, R, A.conj()), R, atol=1e-7)


@pytest.mark.parametrize("case", ["heisenberg", "coupled", "random"])
def test_transfer_leading_values(ctx, case):
    print(f"case ortho transfer leading {case}", flush=True)
    if case == "random":
        rules, mats, lams = _random_cell({-1: 5, 0: 8, 1: 5, 2: 3}, seed=4)
    else:
        rules, W, mats, lams, g = _idmrg_cell(case, chi=10, steps=6)
    cell = _cell(ctx, mats, lams)
    for side in (0, 1):
        want, _, _ = ortho.transfer_leading(mats, side)
        got, info = cell.transfer_leading(side)
        assert info["converged"]
        assert abs(got - want) <= 1e-12 * abs(want), (side, got, want, info)


@pytest.mark.parametrize("case", ["heisenberg", "coupled", "random"])
def test_orthogonalise_matches_oracle(ctx, case):
    print(f"case ortho orthogonalise {case}", flush=True)
    if case == "random":
        rules, mats, lams = _random_cell({-1: 6, 0: 9, 1: 6, 2: 2}, seed=9)
    else:
        rules, W, mats, lams, g = _idmrg_cell(case, chi=12, steps=6)
    new, nl, info = ortho.orthogonalise(mats, lams)
    dev, dinfo = _cell(ctx, mats, lams).orthogonalise()
    assert abs(complex(dinfo["left_re"], dinfo["left_im"]) - info["left_value"]) <= 1e-12 * abs(info["left_value"])
    assert abs(complex(dinfo["right_re"], dinfo["right_im"]) - info["right_value"]) <= 1e-12 * abs(info["right_value"])
    got = np.sort(dev.lam(0))[::-1]
    want = np.sort(np.asarray(nl[0]))[::-1]
    assert got.shape == want.shape
    # lambda_j^2 are eigenvalues of the fixed-point product, which device and
    # oracle both know to about 4e-15 absolute (the transfer eigensolve's
    # residual and the rounding of its products): |d lambda_j| <= 4e-15 /
    # (2 lambda_j) on top of the relative 1e-8 -- the small values are
    # ill-conditioned, and a different summation order in the device's
    # products moves them within this bound
    assert np.all(np.abs(got - want) <= 1e-8 * want + 2e-15 / want), (got, want)
    assert abs(dinfo["entropy"] - info["entropy"]) <= 1e-10
    assert np.array_equal(dev.lam(1), np.asarray(nl[1]))
    _check_canonical(dev, rules, real=case != "random")
    if case != "random":
        e_dev, t_dev = dev.energy(to_lib(ctx, W))
        e_or, t_or = ortho.i_two_vertex_energy(W, new, nl)
        assert abs(e_dev - e_or) <= 1e-10, (e_dev, e_or)


@pytest.mark.parametrize("model", ["heisenberg", "coupled"])
def test_energy_terms_of_the_oracle_cell(ctx, model):
    print(f"case ortho energy {model}", flush=True)
    rules, W, mats, lams, g = _idmrg_cell(model, chi=10, steps=6)
    new, nl, info = ortho.orthogonalise(mats, lams)
    e_or, t_or = ortho.i_two_vertex_energy(W, new, nl)
    e_dev, t_dev = _cell(ctx, new, nl).energy(to_lib(ctx, W))
    assert abs(e_dev - e_or) <= 1e-12
    for a, b in zip(t_dev, [t_or["local"], t_or["bond1"], t_or["bond2"]]):
        assert abs(a - b) <= 1e-12


def test_idmrg_cell_from_the_device_decomposition(ctx):
    """The device iDMRG's unit cell (cam_idmrg_unitcell) orthogonalises to
    the oracle's: same boundary lambdas and energy per vertex."""
    print("case ortho idmrg cell", flush=True)
    import chainamps_amd as cam
    rules, phys, W = models.heisenberg_mpo()
    g = dmrg.IDMRG(W, converge=True)
    dev = cam.IDMRG(to_lib(ctx, W), converge=True)
    for _ in range(7):
        g.grow(12)
        dev.grow(12)
    mats, lams = ortho.unit_cell_from_decomposition(g.U, g.values, g.Vd, g.prev_lambda)
    new, nl, info = ortho.orthogonalise(mats, lams)
    e_or, _ = ortho.i_two_vertex_energy(W, new, nl)
    cell = dev.unitcell()
    assert np.allclose(cell.lam(0), g.prev_lambda, rtol=1e-8, atol=1e-12)
    assert np.allclose(cell.lam(1), g.values, rtol=1e-8, atol=1e-12)
    ocell, dinfo = cell.orthogonalise()
    assert np.allclose(np.sort(ocell.lam(0)), np.sort(nl[0]), rtol=1e-7, atol=1e-10)
    e_dev, _ = ocell.energy(to_lib(ctx, W))
    assert abs(e_dev - e_or) <= 1e-9
    assert abs(e_dev.real - (0.25 - np.log(2.0))) < 2e-3


def test_unitcell_from_decomposition_matches_oracle(ctx):
    print("case ortho cell from decomposition", flush=True)
    import chainamps_amd as cam
    rules, W, mats, lams, g = _idmrg_cell("coupled", chi=8, steps=4)
    cell = cam.UnitCell.from_decomposition(to_lib(ctx, g.U), g.values, to_lib(ctx, g.Vd), g.prev_lambda)
    A1 = from_lib(cell.matrix(1), rules).to_dense_tensor()
    assert np.allclose(A1, mats[1].to_dense_tensor(), rtol=1e-13, atol=1e-15)
    assert np.array_equal(cell.lam(0), lams[0]) and np.array_equal(cell.lam(1), lams[1])


def test_boundary_charge_mismatch_is_an_error(ctx):
    print("case ortho boundary mismatch", flush=True)
    import chainamps_amd as cam
    rules, mats, lams = _random_cell({0: 3, 1: 2}, seed=1)
    # shift the charges of the last outward index: same size, different charges
    A1 = mats[1]
    q = A1.indices[2].charges + 2
    import scipy.sparse as sp
    from oracle.mpx import MPX
    bad = MPX(rules, [A1.indices[0], A1.indices[1], MPXIndex(False, q)], 2, sp.csc_matrix(A1.matrix.shape,
                                                                                          dtype=complex))
    cell = _cell(ctx, [mats[0], bad], lams)
    with pytest.raises(cam.CamError, match="boundary indices"):
        cell.transfer_leading(0)


def test_transfer_apply_at_size(ctx):
    """LeftComponentsMultiply / RightComponentsMultiply on the allowed
    entries in the reference's order, with blocks of 60-101 states (several
    MFMA tiles per block and many K chunks per product), against the oracle's
    dense maps (1e-12 relative)."""
    print("case ortho transfer apply at size", flush=True)
    rules, mats, lams = _random_cell({-1: 60, 0: 100, 1: 60}, seed=12)
    cell = _cell(ctx, mats, lams)
    dense = [m.to_dense_tensor() for m in mats]
    q = mats[0].indices[1].charges
    n = q.shape[0]
    flat, (r, c) = ortho.allowed_entries(q)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(len(flat)) + 1j * rng.standard_normal(len(flat))
    V = np.zeros((n, n), complex)
    V[r, c] = x
    for side, app in ((0, ortho.left_apply), (1, ortho.right_apply)):
        want = app(dense, V)[r, c]
        got = cell.transfer_apply(side, x)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), side


def test_transfer_leading_with_restarts(ctx):
    """A Krylov dimension of 4 forces thick restarts on the Hermitian
    subspace; the leading eigenvalue still matches ARPACK's (1e-12)."""
    print("case ortho transfer leading restarts", flush=True)
    rules, mats, lams = _random_cell({-1: 7, 0: 10, 1: 7, 2: 3}, seed=6)
    cell = _cell(ctx, mats, lams)
    for side in (0, 1):
        want, _, _ = ortho.transfer_leading(mats, side)
        got, info = cell.transfer_leading(side, ncv=4)
        assert info["converged"] and info["restarts"] > 0, info
        assert abs(got - want) <= 1e-12 * abs(want), (side, got, want, info)

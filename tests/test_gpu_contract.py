"""GPU parity: block-sparse contract/reorder vs the oracle (MPX_matrix.cpp:630-791).

Integer/index outputs (index directions, charges, num_row_indices) must be
identical; values within 1e-12 relative (FP64, different summation order).
"""
import numpy as np
import pytest

from conftest import to_lib, from_lib, rel_err
from oracle.qn import ChargeRules, MPXIndex
from oracle import mpx as O
from oracle import models, dmrg

pytestmark = pytest.mark.gpu


def _idx(inward, q, phys=False):
    return MPXIndex(inward, np.asarray(q).reshape(-1, 1), phys)


def _same_indices(A, B):
    assert len(A.indices) == len(B.indices)
    assert A.numrows == B.numrows
    for a, b in zip(A.indices, B.indices):
        assert a.inward == b.inward and a.physical == b.physical
        assert np.array_equal(a.charges, b.charges)


def _cmp(ctx, Mo, T, rules, tol=1e-12):
    Ml = from_lib(T, rules)
    _same_indices(Mo, Ml)
    a = Mo.to_dense_tensor()
    b = Ml.to_dense_tensor()
    assert rel_err(b, a) <= tol


@pytest.mark.parametrize("seed", [0, 1])
def test_contract_mps_chain(ctx, seed):
    print(f"case contract_mps_chain seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    rules = ChargeRules([0])
    qa = rng.integers(-3, 4, 13)
    qs = np.array([1, -1, 1, 3])
    qc = rng.integers(-3, 4, 17)
    qe = rng.integers(-4, 5, 11)
    A = models.random_qn_tensor(rules, [_idx(True, qa), _idx(True, qs, True), _idx(False, qc)], 2, rng)
    B = models.random_qn_tensor(rules, [_idx(True, qc), _idx(True, qs, True), _idx(False, qe)], 2, rng)
    for conjA, conjB, pairs, Aop, Bop in [
        (False, False, [(2, 0)], A, B),
        (True, False, [(0, 0), (1, 1)], A, A),
        (True, False, [(1, 1), (0, 0)], A, A),
        (False, True, [(2, 2)], B, B),
    ]:
        Mo = O.contract(Aop, conjA, Bop, conjB, pairs)
        T = __import__("chainamps_amd").contract(to_lib(ctx, Aop), conjA, to_lib(ctx, Bop), conjB, pairs)
        _cmp(ctx, Mo, T, rules)


def test_contract_z2_and_multicharge(ctx):
    print("case contract_z2_and_multicharge", flush=True)
    import chainamps_amd as cam
    rng = np.random.default_rng(5)
    rules = ChargeRules([2, 0])
    def q(n):
        return np.stack([rng.integers(0, 2, n), rng.integers(-2, 3, n)], axis=1)
    qa, qb, qc, qd = q(7), q(9), q(6), q(5)
    A = models.random_qn_tensor(rules, [MPXIndex(True, qa), MPXIndex(False, qb), MPXIndex(True, qc),
                                        MPXIndex(False, qd)], 2, rng)
    B = models.random_qn_tensor(rules, [MPXIndex(True, qd), MPXIndex(True, qb), MPXIndex(False, qa)], 1, rng)
    for pairs in ([(3, 0), (1, 1)], [(1, 1), (3, 0)], [(3, 0)], [(0, 2), (1, 1), (3, 0)]):
        Mo = O.contract(A, False, B, False, pairs)
        T = cam.contract(to_lib(ctx, A), False, to_lib(ctx, B), False, pairs)
        _cmp(ctx, Mo, T, rules)


def test_reorder_and_metadata(ctx):
    print("case reorder_and_metadata", flush=True)
    import chainamps_amd as cam
    rng = np.random.default_rng(7)
    rules = ChargeRules([0])
    idx = [_idx(True, rng.integers(-2, 3, 5)), _idx(False, rng.integers(-2, 3, 4)),
           _idx(True, rng.integers(-2, 3, 6)), _idx(False, rng.integers(-2, 3, 3)), _idx(True, [0])]
    A = models.random_qn_tensor(rules, idx, 3, rng)
    TA = to_lib(ctx, A)
    for order, nrow, conj in [([4, 3, 2, 1, 0], 2, False), ([1, 0, 3, 2, 4], 3, True),
                              ([0, 1, 2, 3, 4], 1, False)]:
        Mo = O.reorder(A, conj, order, nrow)
        _cmp(ctx, Mo, cam.reorder(TA, conj, order, nrow), rules)
    # round trip CSC exactly (bit-exact values and pattern)
    rows, cols, p, i, x = TA.to_csc()
    m = A.matrix.tocsc()
    m.eliminate_zeros()
    assert np.array_equal(p, m.indptr) and np.array_equal(i, m.indices) and np.array_equal(x, m.data)
    # metadata operations
    Mo = A.copy().remove_dummy_indices([4])
    _cmp(ctx, Mo, to_lib(ctx, A).remove_dummy_indices([4]), rules, tol=0)


def test_zero_flux_violation_is_an_error(ctx):
    print("case zero_flux_violation", flush=True)
    import chainamps_amd as cam
    import scipy.sparse as sp
    from chainamps_amd.capi import IndexDesc
    rules = ChargeRules([0])
    idx = [IndexDesc(True, False, np.array([[0], [1]])), IndexDesc(False, False, np.array([[0], [1]]))]
    m = sp.csc_matrix(np.array([[1.0, 1.0], [0.0, 1.0]], dtype=complex))   # (0,1) breaks flux
    with pytest.raises(cam.CamError, match="zero-flux"):
        cam.Tensor.from_csc(ctx, [0], idx, 1, 2, 2, m.indptr, m.indices, m.data)


def test_block_updates_match_reference_chain(ctx):
    """MakeLeftBlock / MakeRightBlock (DMRG_routines.cpp:594-611) on real
    iDMRG tensors: three contractions + MoveDummyIndices each."""
    print("case block_updates", flush=True)
    import chainamps_amd as cam
    rules, phys, W = models.heisenberg_mpo()
    idm = dmrg.IDMRG(W, converge=True)
    for _ in range(4):
        idm.grow(12)
    LB, RB, U, Vd = idm.LB, idm.RB, idm.U, idm.Vd
    Lo = dmrg.make_left_block(LB, W, U)
    Ro = dmrg.make_right_block(RB, W, Vd)
    tLB, tW, tU = to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, U)
    t = cam.contract(tLB, False, tU, False, [(5, 1)])
    t = cam.contract(tW, False, t, False, [(1, 3), (2, 5)])
    t = cam.contract(tU, True, t, False, [(0, 0), (1, 3)]).move_dummy_indices([2, 0, 3, 1, 4, 5])
    _cmp(ctx, Lo, t, rules)
    tRB, tV = to_lib(ctx, RB), to_lib(ctx, Vd)
    t = cam.contract(tRB, False, tV, False, [(4, 2)])
    t = cam.contract(tW, False, t, False, [(2, 6), (3, 2)])
    t = cam.contract(tV, True, t, False, [(1, 0), (2, 2)]).move_dummy_indices([0, 2, 1, 3, 5, 4])
    _cmp(ctx, Ro, t, rules)

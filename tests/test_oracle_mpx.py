"""Oracle pins: mixed-radix reshape/contract (MPX_matrix.cpp:630-791,
sparse_interface.cpp:966-1102) against numpy.einsum on dense tensors, and
the blockwise SVD truncation rule (sparse_interface.cpp:1248-1406)."""
import numpy as np
import pytest

from oracle.qn import ChargeRules, MPXIndex
from oracle import mpx as O
from oracle import models


def _rand_tensor(rng, rules, spec, numrows):
    idx = [MPXIndex(inw, rng.integers(-2, 3, n).reshape(-1, 1)) for inw, n in spec]
    return models.random_qn_tensor(rules, idx, numrows, rng)


def test_dense_roundtrip_mixed_radix():
    rng = np.random.default_rng(0)
    rules = ChargeRules([0])
    A = _rand_tensor(rng, rules, [(True, 3), (False, 4), (True, 2), (False, 5)], 2)
    T = A.to_dense_tensor()
    # element (i0,i1,i2,i3) sits at row i0 + 3 i1, col i2 + 2 i3
    m = A.matrix.toarray()
    for _ in range(20):
        i = [rng.integers(0, d) for d in T.shape]
        assert m[i[0] + 3 * i[1], i[2] + 2 * i[3]] == T[tuple(i)]
    B = O.MPX.from_dense_tensor(rules, A.indices, 2, T)
    assert (B.matrix != A.matrix).nnz == 0
    assert A.is_consistent()


@pytest.mark.parametrize("seed", range(4))
def test_reorder_vs_transpose(seed):
    rng = np.random.default_rng(seed)
    rules = ChargeRules([0])
    A = _rand_tensor(rng, rules, [(True, 3), (False, 4), (True, 2), (False, 5), (True, 3)], 3)
    order = list(rng.permutation(5))
    nrow = int(rng.integers(1, 5))
    for conj in (False, True):
        B = O.reorder(A, conj, order, nrow)
        exp = A.to_dense_tensor().transpose(order)
        exp = np.conj(exp) if conj else exp
        assert np.allclose(B.to_dense_tensor(), exp, rtol=0, atol=0)
        assert all(B.indices[k].inward == (A.indices[o].inward != conj) for k, o in enumerate(order))


@pytest.mark.parametrize("seed", range(6))
def test_contract_vs_einsum(seed):
    rng = np.random.default_rng(seed)
    rules = ChargeRules([0])
    # A: 4 indices; B: 3 indices; contract 1 or 2 pairs in random order
    qa = [rng.integers(-2, 3, n).reshape(-1, 1) for n in (3, 4, 2, 5)]
    A = models.random_qn_tensor(rules, [MPXIndex(True, qa[0]), MPXIndex(False, qa[1]),
                                        MPXIndex(True, qa[2]), MPXIndex(False, qa[3])], 2, rng)
    conjA = bool(seed % 2)
    # B indices that match A's 1 and 3 (directions chosen to satisfy match)
    dir1 = not (A.indices[1].inward != conjA)  # B must be "ingoing after conj" opposite of A
    B = models.random_qn_tensor(rules, [MPXIndex(not A.indices[3].inward != conjA, qa[3]),
                                        MPXIndex(True, rng.integers(-2, 3, 4).reshape(-1, 1)),
                                        MPXIndex(not A.indices[1].inward != conjA, qa[1])], 1, rng)
    pairs = [(3, 0), (1, 2)] if seed % 3 else [(1, 2), (3, 0)]
    C = O.contract(A, conjA, B, False, pairs)
    TA = A.to_dense_tensor()
    TA = np.conj(TA) if conjA else TA
    TB = B.to_dense_tensor()
    exp = np.einsum("abcd,dxb->acx", TA, TB)
    assert np.allclose(C.to_dense_tensor(), exp, atol=1e-12)
    assert C.numrows == 2 and C.is_consistent()


def test_truncation_rule_cases():
    # D cap
    v = [1.0, 0.5, 0.25, 0.125]
    assert O.truncation_rule(v, 2, 0.0)[0] == 2
    assert O.truncation_rule(v, 0, 0.0)[0] == 4
    # tail truncation: discard while cumulative squared tail < max_trunc
    # tail weights 0.015625, +0.0625 ; max 0.05 -> drop only the last
    assert O.truncation_rule(v, 0, 0.05)[0] == 3
    # degeneracy extension beyond D
    w = [1.0, 0.5, 0.49999, 0.2]
    assert O.truncation_rule(w, 2, 0.0)[0] == 3
    L, kept, disc = O.truncation_rule(w, 2, 0.0)
    assert abs(kept - (1 + 0.25 + 0.49999 ** 2)) < 1e-15 and abs(disc - 0.04) < 1e-15


def test_svd_blockwise_vs_dense():
    rng = np.random.default_rng(3)
    rules = ChargeRules([0])
    A = _rand_tensor(rng, rules, [(True, 4), (True, 6), (False, 7), (True, 3)], 2)
    r = O.mpx_svd(A)
    dense = A.matrix.toarray()
    s_exp = np.linalg.svd(dense, compute_uv=False)
    s_exp = np.sort(s_exp[s_exp > 1e-12])[::-1]
    assert np.allclose(r["values"][:len(s_exp)], s_exp, atol=1e-12)
    U = r["U"].matrix.toarray()
    Vd = r["Vd"].matrix.toarray()
    rec = U @ np.diag(r["values"]) @ Vd
    assert np.allclose(rec, dense, atol=1e-12)
    assert np.allclose(U.conj().T @ U, np.eye(U.shape[1]), atol=1e-12)
    assert r["U"].is_consistent() and r["Vd"].is_consistent()
    # new bond charges: ingoing row state of each U column
    for c in range(U.shape[1]):
        row = int(r["U"].matrix.indices[r["U"].matrix.indptr[c]])
        assert np.array_equal(A.ingoing_row_state(row), r["U"].indices[-1].charges[c])

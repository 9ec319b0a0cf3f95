"""Oracle pins: QN charge algebra (states.hpp:176-205, states.cpp:13-27,
191-252) and MPXIndex combine/match (MPXIndex.cpp:47-96)."""
import numpy as np
import pytest

from oracle.qn import ChargeRules, MPXIndex, combine, match, qn_combinations


def test_zn_wrap_and_u1():
    r = ChargeRules([3, 0, 1])
    a = np.array([2, -5, 0])
    b = np.array([2, 7, 0])
    assert list(r.add(a, b)) == [1, 2, 0]          # Z3: 4 mod 3; U(1) plain; trivial 0
    assert list(r.sub(a - a, b)) == [1, -7, 0]     # Z3: -2 -> 1 (while temp<0 temp+=N)
    assert list(r.neg(np.array([1, 3, 0]))) == [2, -3, 0]
    with pytest.raises(ValueError):
        r.normalise([[-1, 0, 0]])
    assert list(r.normalise([[5, -2, 9]])[0]) == [2, -2, 0]


def test_combine_order_brute_force():
    r = ChargeRules([0])
    a = MPXIndex(True, [[1], [-1], [3]])
    b = MPXIndex(False, [[2], [0]])
    c = combine(r, a, b)
    # b outer, a inner; directions differ -> a - b ; result direction of a
    exp = [ai - bj for bj in (2, 0) for ai in (1, -1, 3)]
    assert list(c.charges[:, 0]) == exp and c.inward
    d = combine(r, a, MPXIndex(True, [[2], [0]]))
    assert list(d.charges[:, 0]) == [ai + bj for bj in (2, 0) for ai in (1, -1, 3)]


def test_match_rules():
    a = MPXIndex(False, [[1], [2]])
    b = MPXIndex(True, [[1], [2]])
    assert match(a, False, b, False)
    assert not match(a, False, b, True)             # conj flips B to outgoing
    assert match(a, True, b, True)
    assert not match(a, False, MPXIndex(True, [[1], [3]]), False)   # charges differ
    assert not match(a, False, MPXIndex(True, [[1], [2]], True), False)  # physical flag


def test_qn_combinations_spin_half():
    r = ChargeRules([0])
    inv, ordered = qn_combinations(r, [[1], [-1]], True)
    assert [q for q, _ in inv] == [(0,)]
    assert sorted(q for q, _ in ordered) == [(-2,), (2,)]
    # ordered pairs come as (x, ..., -x) so position l and ncomb-1-l are negatives
    qs = [q[0] for q, _ in ordered]
    assert qs[0] == -qs[-1]


def test_qn_combinations_z2():
    r = ChargeRules([2])
    inv, ordered = qn_combinations(r, [[0], [1]], True)
    assert sorted(q for q, _ in inv) == [(0,), (1,)] and ordered == []

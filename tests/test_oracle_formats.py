"""Oracle pins for the reference's binary formats (MPX_matrix.cpp:219-247,
594-612; MPXIndex.cpp:8-24, 98-120; sparse_interface.cpp:727-835;
states.cpp:119-189; UnitCell.cpp:70-145): a literal byte string assembled
field by field from the reference's writer for a small tensor, and round
trips of tensors (including an empty one) and of a unit cell."""
import io
import struct

import numpy as np
import scipy.sparse as sp

from oracle import formats, models, tebd
from oracle.mpx import MPX
from oracle.qn import ChargeRules, MPXIndex


def _small():
    rules = ChargeRules([0])
    phys = np.array([[1], [-1]])
    idx = [MPXIndex(True, phys, True), MPXIndex(False, np.array([[1], [-1]]))]
    m = sp.csc_matrix(np.array([[1.0, 0.0], [0.0, 0.5j]]))
    return rules, phys, MPX(rules, idx, 1, m)


def test_literal_bytes_of_a_small_tensor():
    rules, phys, A = _small()
    got = formats.to_bytes(formats.write_mpx, A, 1)
    want = b"".join([
        struct.pack("<Q", 2),                       # number of indices (size_t)
        struct.pack("<q", 1),                       # number of row indices (Sparseint)
        struct.pack("<??", True, True), struct.pack("<Q", 2),            # physical: no charges stored
        struct.pack("<??", False, False), struct.pack("<Q", 2), struct.pack("<hh", 1, -1),
        struct.pack("<qqqq", 2, -1, 2, 2),          # nzmax, nz (CSC), rows, cols
        struct.pack("<qqq", 0, 1, 2),               # column pointers
        struct.pack("<qq", 0, 1),                   # row indices
        struct.pack("<dddd", 1.0, 0.0, 0.0, 0.5),   # complex values
    ])
    assert got == want


def test_round_trip_tensors():
    rules, phys, W, E, sig = models.coupled_chain_model(d=6, seed=2, Jperp=0.4)
    cases = [(rules, phys, W), (rules, phys, tebd.odd_bond_hamiltonian(W)), _small()]
    for rules_, phys_, A in cases:
        b = formats.to_bytes(formats.write_mpx, A, len(rules_.rules))
        B = formats.read_mpx(io.BytesIO(b), rules_, phys_)
        assert (B.matrix != A.matrix).nnz == 0
        assert B.numrows == A.numrows
        for a, c in zip(A.indices, B.indices):
            assert a.inward == c.inward and a.physical == c.physical
            assert np.array_equal(a.charges, c.charges)
    # an empty matrix stores nzmax = 0 and nothing after the column pointers
    rules, phys, A = _small()
    Z = MPX(rules, A.indices, 1, sp.csc_matrix((2, 2), dtype=complex))
    b = formats.to_bytes(formats.write_mpx, Z, 1)
    assert b.endswith(struct.pack("<qqqq", 0, -1, 2, 2) + struct.pack("<qqq", 0, 0, 0))
    assert formats.read_mpx(io.BytesIO(b), rules, phys).matrix.nnz == 0


def test_round_trip_unitcell():
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=1, Jperp=0.3)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 1)
    b = formats.to_bytes(formats.write_unitcell, rules.rules, phys, E, [A0, A1], [l0, l1])
    r, ch, en, mats, lams = formats.read_unitcell(io.BytesIO(b), rules)
    assert tuple(r) == rules.rules and np.array_equal(ch, phys) and np.array_equal(en, E)
    assert len(mats) == 2 and all((m.matrix != a.matrix).nnz == 0 for m, a in zip(mats, [A0, A1]))
    assert all(np.array_equal(x, y) for x, y in zip(lams, [l0, l1]))

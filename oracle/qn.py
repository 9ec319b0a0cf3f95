"""QN charge algebra and tensor indices (oracle; test infrastructure only).

Restates:
  * ``State`` charge arithmetic: states.hpp:176-205 (+= / -=), states.cpp:13-27
    (constructor normalisation: Z_N charges reduced mod N, N==1 forced to 0,
    N<=0 is a plain integer "momentum-like" U(1) charge).
  * ``MPXIndex``: MPXIndex.hpp:17-65 (direction, physical flag, per-entry
    charges), ``match`` MPXIndex.cpp:47-72, ``combine`` MPXIndex.cpp:74-96
    (b outer, a inner: first index fastest; a-b when directions differ).
  * ``QNCombinations``: states.cpp:191-252 (used to size MPO channels).

Charges are stored as int64 numpy arrays of shape (n, nrules); the reference
stores ``short`` (QuantumNumberInt, common_defs.hpp:20).  Values here never
leave the short range in any workload used by the tests.

Parity pins: tests/test_oracle_qn.py checks the Z_N wrap, U(1) arithmetic,
the combine ordering against a brute-force double loop, and the
QNCombinations of the spin-1/2 basis {+1,-1} (involution {0}; ordered
pairs {+2,-2} or {-2,+2}).
"""
from __future__ import annotations

import numpy as np


class ChargeRules:
    """The list of N's defining each charge (states.hpp:32-34)."""

    def __init__(self, rules):
        self.rules = tuple(int(r) for r in rules)

    def __len__(self):
        return len(self.rules)

    def __eq__(self, other):
        return isinstance(other, ChargeRules) and self.rules == other.rules

    def __hash__(self):
        return hash(self.rules)

    def __repr__(self):
        return f"ChargeRules({list(self.rules)})"

    # --- construction normalisation: states.cpp:13-27 -----------------------
    def normalise(self, q):
        q = np.array(q, dtype=np.int64).reshape(-1, len(self.rules)).copy()
        for i, n in enumerate(self.rules):
            if n > 1:
                if np.any(q[:, i] < 0):
                    raise ValueError(f"Z_{n} charges can't be negative")
                q[:, i] %= n
            elif n == 1:
                q[:, i] = 0
        return q

    # --- arithmetic: states.hpp:176-205 --------------------------------------
    def add(self, a, b):
        a = np.asarray(a, dtype=np.int64)
        b = np.asarray(b, dtype=np.int64)
        out = a + b
        for i, n in enumerate(self.rules):
            if n > 0:  # Z_N like (N==1 gives 0)
                out[..., i] = out[..., i] % n
        return out

    def sub(self, a, b):
        a = np.asarray(a, dtype=np.int64)
        b = np.asarray(b, dtype=np.int64)
        out = a - b
        for i, n in enumerate(self.rules):
            if n > 0:
                # "while (temp<0) temp+=N; temp % N" == python modulo for N>0
                out[..., i] = out[..., i] % n
        return out

    def neg(self, a):
        """operator-(State) = Identity - a (states.hpp:217-219)."""
        a = np.asarray(a, dtype=np.int64)
        return self.sub(np.zeros_like(a), a)

    def zero(self, n=1):
        return np.zeros((n, len(self.rules)), dtype=np.int64)


class MPXIndex:
    """Directed tensor index with per-entry charges (MPXIndex.hpp:17-65).

    ``charges`` has shape (size, nrules).  Physical indices in the reference
    point at the model Basis; here they simply carry its charges and the
    ``physical`` flag (which ``match`` compares).
    """

    __slots__ = ("inward", "physical", "charges")

    def __init__(self, inward, charges, physical=False):
        self.inward = bool(inward)
        self.physical = bool(physical)
        self.charges = np.asarray(charges, dtype=np.int64)
        if self.charges.ndim == 1:
            self.charges = self.charges.reshape(-1, 1)

    @property
    def size(self):
        return self.charges.shape[0]

    @property
    def outgoing(self):
        return not self.inward

    def flip(self):
        """MPXIndex::flip (MPXIndex.hpp:57)."""
        return MPXIndex(not self.inward, self.charges, self.physical)

    def with_direction(self, inward):
        return MPXIndex(inward, self.charges, self.physical)

    def __repr__(self):
        return (f"MPXIndex({'in' if self.inward else 'out'}, size={self.size}"
                f"{', phys' if self.physical else ''})")


def match(a: MPXIndex, conj_a: bool, b: MPXIndex, conj_b: bool) -> bool:
    """MPXIndex.cpp:47-72: same type, A outgoing / B ingoing after conj,
    same size; charges equal (checked in the reference's debug build)."""
    if a.physical != b.physical:
        return False
    if (a.outgoing != conj_a) != (b.inward != conj_b):
        return False
    if a.size != b.size:
        return False
    return bool(np.array_equal(a.charges, b.charges))


def combine(rules: ChargeRules, a: MPXIndex, b: MPXIndex) -> MPXIndex:
    """MPXIndex.cpp:74-96: for b (outer) for a (inner): a-b if directions
    differ else a+b; result takes a's direction."""
    na, nb = a.size, b.size
    ca = np.tile(a.charges, (nb, 1))            # a fastest
    cb = np.repeat(b.charges, na, axis=0)
    if a.inward != b.inward:
        s = rules.sub(ca, cb)
    else:
        s = rules.add(ca, cb)
    return MPXIndex(a.inward, s, False)


def qn_combinations(rules: ChargeRules, states, minus: bool):
    """QNCombinations::getQNCombinations (states.cpp:191-252).

    Returns (involution_pairs, ordered_pairs): lists of (charge_tuple,
    [(l, r), ...]).
    """
    states = np.asarray(states, dtype=np.int64).reshape(-1, len(rules))
    unsorted = []  # list of [charge_tuple, pairs]
    for li in range(states.shape[0]):
        for ri in range(states.shape[0]):
            st = rules.sub(states[li], states[ri]) if minus else rules.add(states[li], states[ri])
            key = tuple(int(x) for x in st)
            for ent in unsorted:
                if ent[0] == key:
                    ent[1].append((li, ri))
                    break
            else:
                unsorted.append([key, [(li, ri)]])
    inv = []
    rest = []
    for ent in unsorted:
        neg = tuple(int(x) for x in rules.neg(np.array(ent[0])))
        if neg == ent[0]:
            inv.append((ent[0], ent[1]))
        else:
            rest.append(ent)
    ordered = []  # deque semantics of states.cpp:231-250
    if rest:
        front = [rest.pop(0)]
        back = []
        i = 0
        while i < len(rest):
            neg = tuple(int(x) for x in rules.neg(np.array(rest[i][0])))
            if neg == front[0][0]:
                back.append(rest.pop(i))
                if not rest:
                    break
                front.insert(0, rest.pop(0))
                i = 0
            else:
                i += 1
        ordered = [tuple(e) for e in front] + [tuple(e) for e in back]
        ordered = [(e[0], e[1]) for e in ordered]
    return inv, ordered

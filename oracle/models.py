"""Model MPOs and synthetic workloads (oracle; test infrastructure only).

MPO convention (MPO_matrix.hpp:27-68; make_model.hpp:471-621): indices
[sigma' (in, physical), b_left (in), sigma (out, physical), b_right (out)],
two row indices, so the CSC row is sigma' + d*b_left and the column is
sigma + d*b_right; the entry is <sigma'|O_{b_left,b_right}|sigma>.  The MPO
is lower triangular: W[0,0]=I, W[w-1,w-1]=I, W[w-1,0]=on-site Hamiltonian,
and a bond term J*A(x)B is W[w-1,j]=J*A, W[j,0]=B.  The left open boundary
takes row w-1 and the right open boundary column 0 (DMRG_routines.hpp:218-222).
Bond charges follow from the zero-flux rule q(sigma')+q(b_l)=q(sigma)+q(b_r).

Workloads:
  * ``heisenberg_mpo``: spin-1/2 antiferromagnetic Heisenberg chain,
    U(1) Sz (basis {up:+1, down:-1}), w=5 (1 identity, 3 channels S+,S-,Sz,
    1 start row).
  * ``general_hmpo``: restatement of MakeGeneralHMPO (make_model.hpp:471-621)
    for user-defined vertices (used with the reference's spin_half_example
    input files as a fixture).
  * ``coupled_chain_model``: chains truncated to d levels, E_n sorted
    uniform[0,8], Z2 parity n mod 2, Gaussian symmetric inter-chain operator
    nonzero only between opposite parities, nearest-neighbour J_perp.  MPO
    in the continuumIsing::MakeHamiltonian layout
    (continuum_Ising_generator.hpp:869-905): [I; J*K; H_V, K, I] with
    bond charges (0, 1, 0) under Z2.
  * ``sector_profile`` / ``random_env``: the QN-blocked synthetic batching
    workload (21 U(1) sectors, Gaussian size profile).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .qn import ChargeRules, MPXIndex, qn_combinations
from .mpx import MPX


def mpo_from_blocks(rules, phys_charges, bond_charges, blocks):
    """Assemble an MPO MPX from dense d x d operator blocks keyed (bl, br)."""
    phys_charges = np.asarray(phys_charges, dtype=np.int64).reshape(-1, len(rules))
    bond_charges = np.asarray(bond_charges, dtype=np.int64).reshape(-1, len(rules))
    d = phys_charges.shape[0]
    w = bond_charges.shape[0]
    rr, cc, xx = [], [], []
    for (bl, br), op in blocks.items():
        op = np.asarray(op, dtype=np.complex128)
        nz = np.nonzero(op)
        for sp_, s_ in zip(*nz):
            rr.append(sp_ + d * bl)
            cc.append(s_ + d * br)
            xx.append(op[sp_, s_])
    m = sp.csc_matrix((xx, (rr, cc)), shape=(d * w, d * w))
    idx = [MPXIndex(True, phys_charges, True), MPXIndex(True, bond_charges),
           MPXIndex(False, phys_charges, True), MPXIndex(False, bond_charges)]
    W = MPX(rules, idx, 2, m)
    if not W.is_consistent():
        raise ValueError("MPO blocks violate the zero-flux rule")
    return W


def op_charge_change(rules, phys_charges, op):
    """q(sigma') - q(sigma) of a charge-homogeneous operator."""
    nz = np.argwhere(np.abs(op) > 0)
    if len(nz) == 0:
        return np.zeros(len(rules), dtype=np.int64)
    diffs = {tuple(rules.sub(phys_charges[a], phys_charges[b])) for a, b in nz}
    if len(diffs) != 1:
        raise ValueError("operator is not charge-homogeneous")
    return np.array(next(iter(diffs)), dtype=np.int64)


def lower_triangular_mpo(rules, phys_charges, h_loc, terms):
    """Build W with channels j=1..len(terms): terms = [(J, A, B)] meaning
    J * A(left site) (x) B(right site); each A, B charge-homogeneous."""
    phys_charges = np.asarray(phys_charges, dtype=np.int64).reshape(-1, len(rules))
    d = phys_charges.shape[0]
    w = len(terms) + 2
    blocks = {(0, 0): np.eye(d), (w - 1, w - 1): np.eye(d)}
    if h_loc is not None and np.any(h_loc != 0):
        blocks[(w - 1, 0)] = np.asarray(h_loc)
    bq = [np.zeros(len(rules), np.int64)]
    for j, (J, A, B) in enumerate(terms, start=1):
        blocks[(j, 0)] = np.asarray(B)
        blocks[(w - 1, j)] = J * np.asarray(A)
        # W[j,0]=B: q(s')+q(b_j)=q(s)  ->  q(b_j) = -dq(B)
        bq.append(rules.neg(op_charge_change(rules, phys_charges, np.asarray(B))))
    bq.append(np.zeros(len(rules), np.int64))
    return mpo_from_blocks(rules, phys_charges, np.array(bq), blocks)


# --------------------------------------------------------------------------
# spin-1/2 Heisenberg (U(1) Sz)
# --------------------------------------------------------------------------

def spin_half_ops():
    sp_ = np.array([[0, 1], [0, 0]], dtype=np.complex128)   # S+ : down->up
    sm = sp_.T.copy()
    sz = np.diag([0.5, -0.5]).astype(np.complex128)
    return sp_, sm, sz


def heisenberg_mpo(J=1.0, Jz=None):
    """H = sum_i J/2 (S+S- + S-S+) + Jz Sz Sz, basis index 0=up(+1), 1=down(-1)."""
    Jz = J if Jz is None else Jz
    rules = ChargeRules([0])
    phys = np.array([[1], [-1]])
    sp_, sm, sz = spin_half_ops()
    terms = [(0.5 * J, sm, sp_), (0.5 * J, sp_, sm), (Jz, sz, sz)]
    return rules, phys, lower_triangular_mpo(rules, phys, None, terms)


# --------------------------------------------------------------------------
# MakeGeneralHMPO restatement (make_model.hpp:471-621)
# --------------------------------------------------------------------------

def general_hmpo(rules, phys_charges, vertex_h, local_terms, couplings):
    """``local_terms``: [(value, op)]; ``couplings``: [(value, op1, op2)]
    (value * op1(left) (x) op2(right)).  Channel layout, bond charges and
    the split of each operator into charge-difference components follow
    make_model.hpp:489-620."""
    phys_charges = np.asarray(phys_charges, dtype=np.int64).reshape(-1, len(rules))
    d = phys_charges.shape[0]
    inv, ordered = qn_combinations(rules, phys_charges, True)
    ncomb = len(inv) + len(ordered)
    nc = len(couplings)
    w = 2 + nc * ncomb
    last = w - 1
    H = np.zeros((d * w, d * w), dtype=np.complex128)
    for i in range(d):
        H[i, i] = 1.0
        H[i + last * d, i + last * d] = 1.0
    if vertex_h is not None:
        H[last * d:last * d + d, 0:d] += np.asarray(vertex_h)
    for val, op in local_terms:
        H[last * d:last * d + d, 0:d] += val * np.asarray(op)

    def offset_for(diff, is_row):
        neg = tuple(int(x) for x in rules.neg(np.array(diff)))
        if neg == tuple(diff):
            for l, (q, _) in enumerate(inv):
                if q == tuple(diff):
                    return l
        else:
            for l, (q, _) in enumerate(ordered):
                if q == neg:
                    return (len(inv) + l) if is_row else (ncomb - l - 1)
        return 0

    for b, (val, op1, op2) in enumerate(couplings):
        op1 = np.asarray(op1, dtype=np.complex128)
        op2 = np.asarray(op2, dtype=np.complex128)
        same = op1 is op2 or np.array_equal(op1, op2)
        row_off = (1 + b * ncomb) * d
        col_off = (1 + b * ncomb) * d
        for i, c in np.argwhere(op1 != 0):
            diff = tuple(int(x) for x in rules.sub(phys_charges[i], phys_charges[c]))
            l = offset_for(diff, True)
            H[i + row_off + l * d, c] += op1[i, c]
            if same:
                H[i + last * d, c + col_off + offset_for(diff, False) * d] += val * op1[i, c]
        if not same:
            for i, c in np.argwhere(op2 != 0):
                diff = tuple(int(x) for x in rules.sub(phys_charges[i], phys_charges[c]))
                H[i + last * d, c + col_off + offset_for(diff, False) * d] += val * op2[i, c]
    s = [np.zeros(len(rules), np.int64)]
    for _ in couplings:
        s += [np.array(q) for q, _ in inv]
        s += [np.array(q) for q, _ in ordered]
    s.append(np.zeros(len(rules), np.int64))
    bq = np.array(s, dtype=np.int64).reshape(-1, len(rules))
    idx = [MPXIndex(True, phys_charges, True), MPXIndex(True, bq),
           MPXIndex(False, phys_charges, True), MPXIndex(False, bq)]
    m = sp.csc_matrix(H)
    m.eliminate_zeros()
    return MPX(rules, idx, 2, m)


def read_triplet_operator(path, d):
    """User operator file: a name line, then 'row col value' triplets with
    value either a real or '(re,im)' (make_model.hpp user-model reader)."""
    op = np.zeros((d, d), dtype=np.complex128)
    with open(path) as f:
        lines = f.read().splitlines()
    for line in lines[1:]:
        parts = line.split()
        if len(parts) < 3:
            continue
        r, c, v = int(parts[0]), int(parts[1]), parts[2]
        if v.startswith("("):
            re_, im_ = v.strip("()").split(",")
            op[r, c] = complex(float(re_), float(im_))
        else:
            op[r, c] = float(v)
    return op


# --------------------------------------------------------------------------
# d-level Z2 coupled-chain model
# --------------------------------------------------------------------------

def coupled_chain_spectrum(d=64, seed=0):
    rng = np.random.default_rng(seed)
    E = np.sort(rng.uniform(0.0, 8.0, d))
    parity = (np.arange(d) % 2).astype(np.int64)
    sig = rng.standard_normal((d, d))
    sig = 0.5 * (sig + sig.T)
    mask = (parity[:, None] != parity[None, :])
    sig = np.where(mask, sig, 0.0)
    return E, parity, sig


def coupled_chain_model(d=64, seed=0, Jperp=0.1, E=None, sig=None):
    """Returns (rules, phys charges, W, E, sigma)."""
    E0, parity, sig0 = coupled_chain_spectrum(d, seed)
    E = E0 if E is None else np.asarray(E, dtype=np.float64)
    sig = sig0 if sig is None else np.asarray(sig, dtype=np.float64)
    rules = ChargeRules([2])
    phys = parity.reshape(-1, 1)
    blocks = {(0, 0): np.eye(d), (2, 2): np.eye(d), (2, 0): np.diag(E),
              (1, 0): Jperp * sig, (2, 1): sig}
    bq = np.array([[0], [1], [0]])
    W = mpo_from_blocks(rules, phys, bq, {k: v for k, v in blocks.items() if np.any(v != 0)})
    return rules, phys, W, E, sig


# --------------------------------------------------------------------------
# QN-blocked synthetic batching workload
# --------------------------------------------------------------------------

def sector_profile(chi=1024, qmin=-10, qmax=10, sigma=4.0):
    """Sector sizes from chi*exp(-q^2/(2 sigma^2))/sum, rounded by largest
    remainder so they sum to chi."""
    qs = np.arange(qmin, qmax + 1)
    raw = chi * np.exp(-qs.astype(float) ** 2 / (2 * sigma * sigma))
    raw = raw / raw.sum() * chi
    base = np.floor(raw).astype(np.int64)
    rem = chi - int(base.sum())
    order = np.argsort(-(raw - base), kind="stable")
    base[order[:rem]] += 1
    return qs, base


def bond_charges_from_profile(qs, sizes):
    return np.repeat(qs, sizes).reshape(-1, 1).astype(np.int64)


def random_qn_mpo(rules, phys, bond_q, rng):
    """Random complex N(0,1) entries on every QN-allowed MPO entry."""
    d = phys.shape[0]
    w = bond_q.shape[0]
    blocks = {}
    for bl in range(w):
        for br in range(w):
            need = rules.sub(rules.add(phys, bond_q[br]), bond_q[bl])  # q(s') = q(s)+q(br)-q(bl)
            mask = np.array([[np.array_equal(phys[a], need[b]) for b in range(d)] for a in range(d)])
            if mask.any():
                x = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
                blocks[(bl, br)] = np.where(mask, x, 0)
    return mpo_from_blocks(rules, phys, bond_q, blocks)


def _dense_mask_fill(rules, indices, rng, density=1.0):
    """Random complex N(0,1) on every zero-flux entry of a dense tensor over
    ``indices`` (optionally thinned to ``density``)."""
    dims = [ix.size for ix in indices]
    flux = np.zeros(tuple(dims) + (len(rules),), dtype=np.int64)
    for k, ix in enumerate(indices):
        shape = [1] * len(dims) + [len(rules)]
        shape[k] = ix.size
        q = ix.charges.reshape(shape)
        flux = rules.add(flux, q) if ix.inward else rules.sub(flux, q)
    mask = np.all(flux == 0, axis=-1)
    if density < 1.0:
        mask &= rng.random(mask.shape) < density
    x = rng.standard_normal(mask.shape) + 1j * rng.standard_normal(mask.shape)
    return np.where(mask, x, 0)


def random_left_block(rules, a_q, b_q, rng, density=1.0):
    """Left block with the index layout of MakeDummyLeftBlock /
    MakeLeftBlock (DMRG_routines.cpp:594-632):
    [dummy(out), a'(in), dummy(in), b(out), dummy(in), a(out)]."""
    z = rules.zero(1)
    idx = [MPXIndex(False, z), MPXIndex(True, a_q), MPXIndex(True, z),
           MPXIndex(False, b_q), MPXIndex(True, z), MPXIndex(False, a_q)]
    T = _dense_mask_fill(rules, idx, rng, density)
    return MPX.from_dense_tensor(rules, idx, 3, T)


def random_right_block(rules, a_q, b_q, rng, target=None, density=1.0):
    """Right block with the layout of MakeDummyRightBlock / MakeRightBlock
    (DMRG_routines.cpp:603-653):
    [a'(out), dummy(in,T), b(in), dummy(out), a(in), dummy(out,T)]."""
    z = rules.zero(1)
    t = z if target is None else np.asarray(target, dtype=np.int64).reshape(1, -1)
    idx = [MPXIndex(False, a_q), MPXIndex(True, t), MPXIndex(True, b_q),
           MPXIndex(False, z), MPXIndex(True, a_q), MPXIndex(False, t)]
    T = _dense_mask_fill(rules, idx, rng, density)
    return MPX.from_dense_tensor(rules, idx, 2, T)


def random_qn_tensor(rules, indices, numrows, rng, density=1.0):
    T = _dense_mask_fill(rules, indices, rng, density)
    return MPX.from_dense_tensor(rules, indices, numrows, T)

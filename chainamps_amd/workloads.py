"""Synthetic QN-blocked workloads for the H_eff benchmark (host side).

Builds, directly as CSC arrays for ``Tensor.from_csc``, the tensors of the
two-vertex effective Hamiltonian in the reference's index conventions:

* the d-level Z2 coupled-chain MPO (levels E_n sorted uniform[0,8], parity
  n mod 2, Gaussian symmetric inter-chain operator between opposite
  parities, coupling J_perp) in the lower-triangular layout
  [I; J*K; H_V, K, I] of continuumIsing::MakeHamiltonian
  (continuum_Ising_generator.hpp:869-905), indices
  [sigma'(in), b_l(in), sigma(out), b_r(out)] (MPO_matrix.hpp);
* left/right environment blocks with the index layout of MakeLeftBlock /
  MakeRightBlock (DMRG_routines.cpp:594-653), random complex normal on every
  QN-allowed entry;
* the 21-sector U(1) batching workload (Gaussian sector profile, d charges
  uniform in [-3,3], w=6 MPO with QN-conserving random entries).
"""
from __future__ import annotations

import numpy as np

from .capi import IndexDesc, Tensor


def _csc_from_dense(T, numrows):
    """Dense tensor (C-order axes = index order) -> CSC of the mixed radix
    (first index fastest)."""
    dims = T.shape
    rows = int(np.prod(dims[:numrows]))
    flat = np.asarray(T, dtype=np.complex128).transpose(list(range(T.ndim))[::-1]).reshape(-1)
    m = flat.reshape((rows, -1), order="F")
    cols = m.shape[1]
    nzr, nzc = np.nonzero(m.T)                 # column-major order: by column then row
    p = np.zeros(cols + 1, np.int64)
    np.add.at(p, nzr + 1, 1)
    p = np.cumsum(p)
    return rows, cols, p, nzc.astype(np.int64), m[nzc, nzr]


def tensor_from_dense(ctx, rules, indices, numrows, T):
    rows, cols, p, i, x = _csc_from_dense(T, numrows)
    return Tensor.from_csc(ctx, rules, indices, numrows, rows, cols, p, i, x)


def _flux_mask(rules, indices, dims):
    """True where sum_in q - sum_out q == 0 (zero-flux rule)."""
    nr = len(rules)
    flux = np.zeros(tuple(dims) + (nr,), dtype=np.int64)
    for k, ix in enumerate(indices):
        shape = [1] * len(dims) + [nr]
        shape[k] = dims[k]
        q = np.asarray(ix.charges, dtype=np.int64).reshape(shape)
        flux = flux + q if ix.inward else flux - q
    for r, n in enumerate(rules):
        if n > 0:
            flux[..., r] %= n
    return np.all(flux == 0, axis=-1)


def coupled_chain_spectrum(d=64, seed=0):
    rng = np.random.default_rng(seed)
    E = np.sort(rng.uniform(0.0, 8.0, d))
    parity = (np.arange(d) % 2).astype(np.int64)
    sig = rng.standard_normal((d, d))
    sig = 0.5 * (sig + sig.T)
    sig = np.where(parity[:, None] != parity[None, :], sig, 0.0)
    return E, parity, sig


def coupled_chain_mpo(ctx, d=64, seed=0, Jperp=0.1):
    """Returns (Tensor W, rules, physical charges, bond charges)."""
    E, parity, sig = coupled_chain_spectrum(d, seed)
    rules = [2]
    phys = parity.reshape(-1, 1)
    bq = np.array([[0], [1], [0]])
    T = np.zeros((d, 3, d, 3), complex)
    T[:, 0, :, 0] = np.eye(d)
    T[:, 2, :, 2] = np.eye(d)
    T[:, 2, :, 0] = np.diag(E)
    T[:, 1, :, 0] = Jperp * sig
    T[:, 2, :, 1] = sig
    idx = [IndexDesc(True, True, phys), IndexDesc(True, False, bq),
           IndexDesc(False, True, phys), IndexDesc(False, False, bq)]
    return tensor_from_dense(ctx, rules, idx, 2, T), rules, phys, bq


def product_state_cell_dense(rules, phys, level=0):
    """Two-vertex unit cell of the product state of basis level `level`
    (MakeProductStateUnitCell, UnitCell.cpp:190-210): [(index descriptors,
    dense A[s, a_l, a_r]) for the two vertices] with 1 x 1 bonds; the left
    bond of A0 carries charge 0, its right bond the level's charge, A1's
    right bond twice it (Z_N charges wrap)."""
    phys = np.asarray(phys, dtype=np.int64).reshape(-1, len(rules))
    q = phys[level]
    q2 = np.array([(2 * q[i]) % n if n > 0 else 2 * q[i] for i, n in enumerate(rules)])
    z = np.zeros((1, len(rules)), np.int64)
    d = phys.shape[0]

    def site(ql, qr):
        T = np.zeros((d, 1, 1), complex)
        T[level, 0, 0] = 1.0
        return [IndexDesc(True, True, phys), IndexDesc(True, False, ql), IndexDesc(False, False, qr)], T
    return [site(z, q.reshape(1, -1)), site(q.reshape(1, -1), q2.reshape(1, -1))]


def product_state_cell(ctx, rules, phys, level=0):
    """(A0, A1, lambda0, lambda1) of product_state_cell_dense on the device."""
    (i0, T0), (i1, T1) = product_state_cell_dense(rules, phys, level)
    return (tensor_from_dense(ctx, rules, i0, 2, T0), tensor_from_dense(ctx, rules, i1, 2, T1),
            np.ones(1), np.ones(1))


def coupled_chain_generators(csc, d=64, seed=0):
    """Coupling generators of the coupled-chain MPO of coupled_chain_mpo, in
    the form Heff.coupling_grad takes: [dW/dJperp, dW/dE_0, ..., dW/dE_{d-1}]
    as (entries, coefs) over W's stored entries, csc = W.to_csc() (row
    s' + d*b_left, column s + d*b_right).  W must have been built with
    Jperp != 0 so that its (1, 0) channel is stored."""
    rows, cols, p, i, x = csc
    E, parity, sig = coupled_chain_spectrum(d, seed)
    col = np.repeat(np.arange(cols), np.diff(p))
    bl, sl = i // d, i % d
    br, sr = col // d, col % d
    jk = np.nonzero((bl == 1) & (br == 0))[0]
    if len(jk) == 0:
        raise ValueError("coupled_chain_generators: W has no stored Jperp channel (built with Jperp = 0?)")
    gens = [(jk, sig[sl[jk], sr[jk]].astype(np.complex128))]
    for n in range(d):
        k = np.nonzero((bl == 2) & (br == 0) & (sl == n) & (sr == n))[0]
        gens.append((k, np.ones(len(k), np.complex128)))
    return gens


def _hermitian_channels(T, mask, rows, cols):
    """Make every MPO channel b of a random environment a Hermitian
    (a' x a) matrix, so that H_eff is Hermitian for a Hermitian MPO whose
    channels are self-conjugate (the Z2 coupled chain)."""
    M = np.moveaxis(T.reshape(rows), [0, 1, 2], [1, 0, 2])      # (b, a', a)
    M = 0.5 * (M + np.conj(np.swapaxes(M, 1, 2)))
    return np.where(mask, np.moveaxis(M, [0, 1, 2], [1, 0, 2]).reshape(cols), 0)


def left_block_dense(rules, aq, bq, rng, hermitian=False):
    """(index descriptors, dense tensor) of a random left block
    [dummy(out), a'(in), dummy(in), b(out), dummy(in), a(out)]."""
    z = np.zeros((1, len(rules)), np.int64)
    idx = [IndexDesc(False, False, z), IndexDesc(True, False, aq), IndexDesc(True, False, z),
           IndexDesc(False, False, bq), IndexDesc(True, False, z), IndexDesc(False, False, aq)]
    dims = [1, len(aq), 1, len(bq), 1, len(aq)]
    mask = _flux_mask(rules, idx, dims)
    T = np.where(mask, rng.standard_normal(dims) + 1j * rng.standard_normal(dims), 0)
    if hermitian:
        T = _hermitian_channels(T, mask, (len(aq), len(bq), len(aq)), dims)
    return idx, T


def right_block_dense(rules, aq, bq, rng, hermitian=False):
    """(index descriptors, dense tensor) of a random right block
    [a'(out), dummy(in), b(in), dummy(out), a(in), dummy(out)]."""
    z = np.zeros((1, len(rules)), np.int64)
    idx = [IndexDesc(False, False, aq), IndexDesc(True, False, z), IndexDesc(True, False, bq),
           IndexDesc(False, False, z), IndexDesc(True, False, aq), IndexDesc(False, False, z)]
    dims = [len(aq), 1, len(bq), 1, len(aq), 1]
    mask = _flux_mask(rules, idx, dims)
    T = np.where(mask, rng.standard_normal(dims) + 1j * rng.standard_normal(dims), 0)
    if hermitian:
        T = _hermitian_channels(T, mask, (len(aq), len(bq), len(aq)), dims)
    return idx, T


def random_left_block(ctx, rules, aq, bq, rng, hermitian=False):
    idx, T = left_block_dense(rules, aq, bq, rng, hermitian)
    return tensor_from_dense(ctx, rules, idx, 3, T)


def random_right_block(ctx, rules, aq, bq, rng, hermitian=False):
    idx, T = right_block_dense(rules, aq, bq, rng, hermitian)
    return tensor_from_dense(ctx, rules, idx, 2, T)


def coupled_chain_heff_inputs(ctx, chi=500, d=64, seed=0, Jperp=0.1, hermitian=False):
    """Z2 environments with two equal parity sectors of chi/2 states;
    hermitian=True gives Hermitian environment channels (an H_eff the
    eigensolver can be timed on)."""
    rng = np.random.default_rng(seed + 1)
    W, rules, phys, bq = coupled_chain_mpo(ctx, d, seed, Jperp)
    aq = (np.arange(chi) % 2).reshape(-1, 1)
    L = random_left_block(ctx, rules, aq, bq, rng, hermitian)
    R = random_right_block(ctx, rules, aq, bq, rng, hermitian)
    return L, W, R


def sector_profile(chi=1024, qmin=-10, qmax=10, sigma=4.0):
    qs = np.arange(qmin, qmax + 1)
    raw = chi * np.exp(-qs.astype(float) ** 2 / (2 * sigma * sigma))
    raw = raw / raw.sum() * chi
    base = np.floor(raw).astype(np.int64)
    rem = chi - int(base.sum())
    order = np.argsort(-(raw - base), kind="stable")
    base[order[:rem]] += 1
    return qs, base


def u1_batching_dense(chi=1024, d=32, w=6, seed=0):
    """The batching workload on the host: 21 U(1) sectors (charges -10..10,
    Gaussian sizes, sigma 4), d charges uniform in [-3,3], MPO bond charges
    (0,+1,-1,+2,-2,0)[:w], all blocks complex N(0,1) on QN-allowed entries.
    Returns rules and (index descriptors, dense tensor, row indices) of L,
    W and R."""
    rng = np.random.default_rng(seed)
    rules = [0]
    qs, sizes = sector_profile(chi)
    aq = np.repeat(qs, sizes).reshape(-1, 1)
    phys = rng.integers(-3, 4, d).reshape(-1, 1)
    bq = np.array([0, 1, -1, 2, -2, 0][:w]).reshape(-1, 1)
    idx = [IndexDesc(True, True, phys), IndexDesc(True, False, bq),
           IndexDesc(False, True, phys), IndexDesc(False, False, bq)]
    dims = [d, w, d, w]
    mask = _flux_mask(rules, idx, dims)
    T = np.where(mask, rng.standard_normal(dims) + 1j * rng.standard_normal(dims), 0)
    li, lt = left_block_dense(rules, aq, bq, rng)
    ri, rt = right_block_dense(rules, aq, bq, rng)
    return rules, (li, lt, 3), (idx, T, 2), (ri, rt, 2)


def u1_batching_inputs(ctx, chi=1024, d=32, w=6, seed=0):
    """The batching workload (u1_batching_dense) as device tensors L, W, R."""
    rules, *parts = u1_batching_dense(chi, d, w, seed)
    return tuple(tensor_from_dense(ctx, rules, idx, nr, T) for idx, T, nr in parts)

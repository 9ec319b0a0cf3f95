"""GPU: blockwise SVD truncation of large QN blocks.

Blocks with min(rows, cols) >= 1536 of which at most D values can be kept
are decomposed for their leading values only (randomized subspace iteration
with a Ritz-residual acceptance test); the truncation rule, the discarded
weight and the new bond are those of the full decomposition as long as the
computed values reach below the kept set, otherwise the block is decomposed
in full.  The blocks here are built as Q1 diag(s) Q2^H with prescribed s, so
the exact singular values are known: the kept values must match them to
1e-13 (absolute, |theta|_F = O(1)), the discarded weight must equal the
prescribed tail, and U^H theta V must be diag(values).  The same checks run
with CAM_SVD_EXACT=1 (every block decomposed in full).
"""
import os

import numpy as np
import pytest

from conftest import from_lib
from oracle.qn import ChargeRules
from oracle import mpx as O

pytestmark = pytest.mark.gpu


def _theta(ctx, spectra, m, seed):
    import chainamps_amd as cam
    from chainamps_amd.capi import IndexDesc
    from chainamps_amd.workloads import tensor_from_dense
    rng = np.random.default_rng(seed)
    q = (np.arange(2 * m) % 2).reshape(-1, 1)
    A = np.zeros((2 * m, 2 * m), complex)
    for p, s in enumerate(spectra):
        rows = np.nonzero(q[:, 0] == p)[0]
        q1, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
        q2, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
        A[np.ix_(rows, rows)] = (q1 * s) @ q2.conj().T
    idx = [IndexDesc(True, False, q), IndexDesc(False, False, q)]
    return tensor_from_dense(ctx, [2], idx, 1, A), A


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("kind,D,trunc", [("geometric", 150, 0.0), ("geometric", 120, 1e-7), ("slow", 100, 0.0)])
def test_large_block_truncation(ctx, kind, D, trunc, exact):
    print(f"case svd_large {kind} D={D} trunc={trunc} exact={exact}", flush=True)
    import chainamps_amd as cam
    m = 2048
    i = np.arange(m)
    if kind == "geometric":
        spectra = [0.98 ** i, 0.9 * 0.975 ** i]
    else:
        spectra = [(1.0 + i) ** -0.6, 0.95 * (1.5 + i) ** -0.6]
    norm = np.sqrt(sum(np.sum(s ** 2) for s in spectra))
    spectra = [s / norm for s in spectra]
    T, A = _theta(ctx, spectra, m, seed=5)
    allv = np.sort(np.concatenate(spectra))[::-1]
    length, kept, disc = O.truncation_rule(list(allv), D, trunc)
    old = os.environ.get("CAM_SVD_EXACT")
    os.environ["CAM_SVD_EXACT"] = "1" if exact else "0"
    try:
        U, vals, Vd, d = cam.svd_truncate(T, D, trunc)
    finally:
        if old is None:
            del os.environ["CAM_SVD_EXACT"]
        else:
            os.environ["CAM_SVD_EXACT"] = old
    assert len(vals) == length
    assert np.max(np.abs(vals - allv[:length])) <= 1e-13
    assert abs(d - disc) <= 1e-12
    rules = ChargeRules([2])
    Um = from_lib(U, rules).matrix.toarray()
    Vm = from_lib(Vd, rules).matrix.toarray()
    assert np.allclose(Um.conj().T @ Um, np.eye(length), atol=1e-12)
    assert np.allclose(Vm @ Vm.conj().T, np.eye(length), atol=1e-12)
    proj = Um.conj().T @ A @ Vm.conj().T
    assert np.allclose(proj, np.diag(vals), atol=1e-12)


def test_kept_count_hint_reverification(ctx, capfd):
    """A block whose leading-values decomposition was accepted on the first
    (last kept count + margin) triplets, but of which the truncation keeps
    more: the kept triplets beyond the verified prefix must be re-verified
    (or the block decomposed in full), so the result stays the exact one.
    Three decompositions of one shape on one context: the split of the kept
    values between the two blocks swaps in the second."""
    import chainamps_amd as cam
    print("case svd kept-count hint re-verification", flush=True)
    m, D = 2048, 150
    i = np.arange(m)
    splits = [(0.45 * 0.98 ** i, 0.98 ** i), (0.98 ** i, 0.45 * 0.98 ** i), (0.98 ** i, 0.45 * 0.98 ** i)]
    old = os.environ.get("CAM_SVD_TRACE")
    os.environ["CAM_SVD_TRACE"] = "1"
    try:
        for call, spectra in enumerate(splits):
            norm = np.sqrt(sum(np.sum(s ** 2) for s in spectra))
            spectra = [s / norm for s in spectra]
            T, A = _theta(ctx, spectra, m, seed=7 + call)
            allv = np.sort(np.concatenate(spectra))[::-1]
            length, kept, disc = O.truncation_rule(list(allv), D, 0.0)
            U, vals, Vd, d = cam.svd_truncate(T, D, 0.0)
            err = capfd.readouterr().err
            kept0 = int(np.sum(spectra[0] >= allv[length - 1]))
            print(f"  call {call}: block 0 keeps {kept0} of {length}; "
                  f"re-verified: {'re-verified' in err}", flush=True)
            if call == 1:
                # the hint from call 0 (block 0 kept few) is too small now
                assert "re-verified" in err or "full decomposition" in err
            assert len(vals) == length
            assert np.max(np.abs(vals - allv[:length])) <= 1e-13
            assert abs(d - disc) <= 1e-12
            rules = ChargeRules([2])
            Um = from_lib(U, rules).matrix.toarray()
            Vm = from_lib(Vd, rules).matrix.toarray()
            proj = Um.conj().T @ A @ Vm.conj().T
            assert np.allclose(proj, np.diag(vals), atol=1e-12)
    finally:
        if old is None:
            del os.environ["CAM_SVD_TRACE"]
        else:
            os.environ["CAM_SVD_TRACE"] = old


def test_itebd_block_shape_32768(ctx, capfd):
    """The SVD shape of the d=64 chi=1024 iTEBD quench: two Z2 blocks of
    32768 x 32768 keeping D = 1024 values (leading-values path, k = 1536).
    theta = A . B is formed on the device from factors with a prescribed
    spectrum, A = Q1 diag(s) and B = Q2^H per block (Q1, Q2 random with
    orthonormal columns, rank 1792 per block, geometric spectra that
    interleave between the blocks), so the exact values, the discarded
    weight and the singular vectors' defining relations are known."""
    import chainamps_amd as cam
    from chainamps_amd.capi import IndexDesc
    print("case svd itebd block 2 x 32768^2 keep 1024", flush=True)
    m, r, D = 32768, 1792, 1024
    rng = np.random.default_rng(17)
    j = np.arange(r)
    spectra = [0.996 ** j, 0.93 * 0.9955 ** j]
    norm = np.sqrt(sum(np.sum(s ** 2) for s in spectra))
    spectra = [s / norm for s in spectra]

    def orthonormal_columns(rows, cols):
        q, _ = np.linalg.qr(rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols)))
        return q
    Q1 = [orthonormal_columns(m, r) for _ in range(2)]
    Q2 = [orthonormal_columns(m, r) for _ in range(2)]
    print("  factors drawn", flush=True)
    qrow = (np.arange(2 * m) % 2).reshape(-1, 1)
    qb = (np.arange(2 * r) % 2).reshape(-1, 1)
    # A (rows 2m, bond 2r): column c of parity g holds Q1_g[:, c // 2] s_g[c // 2] on the rows of parity g
    XA = np.empty((2 * r, m), complex)
    IA = np.empty((2 * r, m), np.int64)
    XB = np.empty((2 * m, r), complex)
    IB = np.empty((2 * m, r), np.int64)
    for g in range(2):
        XA[g::2] = (Q1[g] * spectra[g]).T
        IA[g::2] = 2 * np.arange(m) + g
        XB[g::2] = Q2[g].conj()
        IB[g::2] = 2 * np.arange(r) + g
    A = cam.Tensor.from_csc(ctx, [2], [IndexDesc(True, False, qrow), IndexDesc(False, False, qb)], 1,
                            2 * m, 2 * r, np.arange(2 * r + 1, dtype=np.int64) * m, IA.reshape(-1),
                            XA.reshape(-1))
    B = cam.Tensor.from_csc(ctx, [2], [IndexDesc(True, False, qb), IndexDesc(False, False, qrow)], 1,
                            2 * r, 2 * m, np.arange(2 * m + 1, dtype=np.int64) * r, IB.reshape(-1),
                            XB.reshape(-1))
    del XA, IA, XB, IB
    T = cam.contract(A, False, B, False, [(1, 0)])
    A.close()
    B.close()
    print("  theta formed on the device", flush=True)
    allv = np.sort(np.concatenate(spectra))[::-1]
    length, kept, disc = O.truncation_rule(list(allv), D, 0.0)
    old = os.environ.get("CAM_SVD_TRACE")
    os.environ["CAM_SVD_TRACE"] = "1"
    try:
        U, vals, Vd, d = cam.svd_truncate(T, D, 0.0)
    finally:
        if old is None:
            del os.environ["CAM_SVD_TRACE"]
        else:
            os.environ["CAM_SVD_TRACE"] = old
    T.close()
    err = capfd.readouterr().err
    print("  " + "\n  ".join(l for l in err.splitlines() if l.startswith("svd_truncate")), flush=True)
    assert "leading 1536" in err
    assert len(vals) == length
    assert np.max(np.abs(vals - allv[:length])) <= 1e-13
    assert abs(d - disc) <= 1e-12
    rules = ChargeRules([2])
    Um = from_lib(U, rules).matrix.toarray()
    Vm = from_lib(Vd, rules).matrix.toarray()
    U.close()
    Vd.close()
    assert np.allclose(Um.conj().T @ Um, np.eye(length), atol=1e-12)
    assert np.allclose(Vm @ Vm.conj().T, np.eye(length), atol=1e-12)
    # U^H theta V^H = diag(values), block by block from the factors
    proj = np.zeros((length, length), complex)
    for g in range(2):
        rows = slice(g, 2 * m, 2)
        proj += ((Um[rows].conj().T @ Q1[g]) * spectra[g]) @ (Q2[g].conj().T @ Vm[:, rows].conj().T)
    assert np.allclose(proj, np.diag(vals), atol=1e-12)

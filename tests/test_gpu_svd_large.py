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

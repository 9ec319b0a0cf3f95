"""CPU: the oracle's C++ reference-order matvec (oracle/refmatvec.cpp), the
CPU baseline bench.py times, and the closed-form flop count bench.py uses
for the d=64 Z2 coupled chain.

* refmatvec --check: its two CSC sparse products in the reference order
  (threaded and not, both routes of the row-order rule) against a dense
  seven-index contraction of the same tensors, relative 1e-13;
* the closed form 12 chi^3 d^2 + 20 chi^2 d^3 against a count of 8 m n k
  over the QN-allowed block triples of the oracle's tensors (two parity
  sectors of chi/2, d/2 levels per parity).
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def refmatvec():
    sys.path.insert(0, ROOT)
    from oracle import build as ob
    return str(ob.build_refmatvec())


@pytest.mark.parametrize("chi,d,threads,seed", [(8, 4, 1, 0), (6, 6, 3, 1), (10, 8, 4, 2), (4, 16, 2, 3)])
def test_refmatvec_against_dense(refmatvec, chi, d, threads, seed):
    r = subprocess.run([refmatvec, "--chi", str(chi), "--d", str(d), "--threads", str(threads),
                        "--seed", str(seed), "--check"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout)
    assert rec["n"] == chi * chi * d * d // 2
    assert rec["check_rel_err"] <= 1e-13, rec
    # the reference order's work: LeftPart has chi/2 (d + 2) entries per column
    assert rec["nz_left_part"] == chi * d * (chi // 2) * (d + 2)
    # the multiply-add count bench.py extrapolates with
    sys.path.insert(0, ROOT)
    import bench
    assert rec["complex_macs"] == bench.ref_macs(chi, d)


def test_refmatvec_rejects_bad_arguments(refmatvec):
    r = subprocess.run([refmatvec, "--chi", "7"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


def _block_flops(LB, W, RB):
    """Sum of 8 m n k over nonzero QN block triples of the four stages,
    counted from the dense tensors."""
    L = LB.to_dense_tensor()[0, :, 0, :, 0, :]
    R = RB.to_dense_tensor()[:, 0, :, 0, :, 0]
    Wd = W.to_dense_tensor()
    aq = LB.indices[5].charges[:, 0]
    a2q = RB.indices[4].charges[:, 0]
    sq = W.indices[2].charges[:, 0]
    add = lambda a, b: (a + b) % 2
    secs = lambda q: {c: int(np.sum(q == c)) for c in np.unique(q)}
    A, A2, S = secs(aq), secs(a2q), secs(sq)
    w = Wd.shape[1]
    nzL = {(b, c0): c0p for b in range(w) for c0 in A for c0p in A
           if np.any(L[np.ix_(aq == c0p, [b], aq == c0)])}
    nzR = {(b, c3): c3p for b in range(w) for c3 in A2 for c3p in A2
           if np.any(R[np.ix_(a2q == c3p, [b], a2q == c3)])}
    nzW = {(bl, br, sp, s_) for bl in range(w) for br in range(w) for sp in S for s_ in S
           if np.any(Wd[np.ix_(sq == sp, [bl], sq == s_, [br])])}
    flops = 0.0
    psi = [(s1, s0, s2, add(add(s1, s0), s2)) for s1 in S for s0 in A for s2 in S if add(add(s1, s0), s2) in A2]
    T1 = set()
    for (s1, s0, s2, s3) in psi:
        for b in range(w):
            if (b, s0) in nzL:
                s0p = nzL[(b, s0)]
                flops += 8 * A[s0p] * A[s0] * A2[s3] * S[s1] * S[s2]
                T1.add((b, s0p, s3, s1, s2))
    T2 = set()
    for (b, s0p, s3, s1, s2) in T1:
        for (bl, br, sp, s_) in nzW:
            if bl == b and s_ == s1:
                flops += 8 * A[s0p] * A2[s3] * S[sp] * S[s1] * S[s2]
                T2.add((br, s0p, s3, sp, s2))
    T3 = set()
    for (b, s0p, s3, s1p, s2) in T2:
        for (bl, br, sp, s_) in nzW:
            if bl == b and s_ == s2:
                flops += 8 * A[s0p] * A2[s3] * S[s1p] * S[sp] * S[s2]
                T3.add((br, s0p, s3, s1p, sp))
    for (b, s0p, s3, s1p, s2p) in T3:
        if (b, s3) in nzR:
            flops += 8 * A[s0p] * A2[nzR[(b, s3)]] * A2[s3] * S[s1p] * S[s2p]
    return flops


@pytest.mark.parametrize("chi,d", [(8, 8), (12, 16)])
def test_closed_form_flops(chi, d):
    sys.path.insert(0, ROOT)
    import bench
    from oracle import models
    rules, phys, W, E, sig = models.coupled_chain_model(d=d, seed=0, Jperp=0.1)
    rng = np.random.default_rng(1)
    aq = (np.arange(chi) % 2).reshape(-1, 1)
    bq = W.indices[1].charges
    LB = models.random_left_block(rules, aq, bq, rng)
    RB = models.random_right_block(rules, aq, bq, rng)
    assert _block_flops(LB, W, RB) == bench.matvec_flops_z2(chi, d)

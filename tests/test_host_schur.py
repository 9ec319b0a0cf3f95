"""Host code: the complex Schur decomposition and reordering used by the
largest-magnitude transfer-matrix eigensolver (chainamps_amd/csrc/schur.cpp),
built with g++ and checked against numpy: A = Z T Z^H, Z unitary, T upper
triangular, diag(T) = numpy.linalg.eigvals(A), leading magnitudes descending.
Cases: random complex, real nonsymmetric, Hessenberg-like, defective Jordan
blocks, diagonal with repeated values, 1 x 1, and a transfer-matrix-like
spectrum (one dominant real eigenvalue)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("schur") / "schur_check"
    subprocess.run([gxx, "-O2", "-std=c++17", str(ROOT / "tests/host/schur_check.cpp"),
                    str(ROOT / "chainamps_amd/csrc/schur.cpp"), "-o", str(out)], check=True)
    return out


def _cases():
    rng = np.random.default_rng(11)
    out = []
    for n in (1, 2, 3, 5, 8, 13, 20, 31, 40):
        out.append(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        out.append(rng.standard_normal((n, n)) + 0j)
    J = np.diag(np.full(6, 2.0 + 1.0j)) + np.diag(np.ones(5), 1)
    out.append(J)
    out.append(np.diag([1.0, 3.0, 3.0, -2.0, 0.5, 3.0]) + 0j)
    # transfer-matrix-like: dominant 1, the rest inside |z| < 0.9
    n = 20
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    d = np.concatenate([[1.0], 0.9 * rng.uniform(0, 1, n - 1) * np.exp(2j * np.pi * rng.uniform(0, 1, n - 1))])
    out.append(Q @ np.diag(d) @ np.linalg.inv(Q))
    return out


def test_schur_against_numpy(exe):
    cases = _cases()
    lines = []
    for A in cases:
        n = A.shape[0]
        lines.append(f"{n} {max(1, n // 2)}")
        lines.extend(f"{float(z.real)!r} {float(z.imag)!r}" for z in A.reshape(-1, order="F"))
    r = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    tok = r.stdout.split()
    p = 0
    for A in cases:
        n, ok = int(tok[p]), int(tok[p + 1])
        p += 2
        assert ok == 1
        vals = np.array(tok[p:p + 4 * n * n], dtype=float)
        p += 4 * n * n
        T = (vals[0:2 * n * n:2] + 1j * vals[1:2 * n * n:2]).reshape(n, n, order="F")
        Z = (vals[2 * n * n::2] + 1j * vals[2 * n * n + 1::2]).reshape(n, n, order="F")
        scale = max(1.0, np.abs(A).max())
        assert np.abs(Z @ T @ Z.conj().T - A).max() <= 1e-12 * scale * n
        assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= 1e-13 * n
        assert np.abs(np.tril(T, -1)).max(initial=0.0) == 0.0
        ev = np.linalg.eigvals(A)
        # same multiset (defective blocks are ill-conditioned: eps^(1/6) spread)
        tol = 1e-2 if A is cases[-3] else 1e-9 * scale
        got = np.diag(T).copy()
        for e in ev:
            k = np.argmin(np.abs(got - e))
            assert abs(got[k] - e) <= tol
            got[k] = np.inf
        lead = np.abs(np.diag(T))[:max(1, n // 2)]
        assert np.all(np.diff(lead) <= 1e-12 * scale)
        assert abs(lead[0] - np.abs(ev).max()) <= 1e-9 * scale

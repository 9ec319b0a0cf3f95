import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) -- parity tests of the HIP path")


@pytest.fixture(scope="session")
def ctx():
    import chainamps_amd as cam
    c = cam.Context(0)
    yield c
    c.close()


def to_lib(ctx, M):
    """oracle MPX -> device Tensor through the C ABI."""
    import chainamps_amd as cam
    from chainamps_amd.capi import IndexDesc
    m = M.matrix.tocsc()
    m.sort_indices()
    idx = [IndexDesc(ix.inward, ix.physical, ix.charges) for ix in M.indices]
    return cam.Tensor.from_csc(ctx, list(M.rules.rules), idx, M.numrows, m.shape[0], m.shape[1],
                               m.indptr.astype(np.int64), m.indices.astype(np.int64), m.data)


def from_lib(T, rules):
    """device Tensor -> oracle MPX."""
    import scipy.sparse as sp
    from oracle.mpx import MPX
    from oracle.qn import MPXIndex
    rows, cols, p, i, x = T.to_csc()
    idx = [MPXIndex(d.inward, d.charges, d.physical) for d in T.indices()]
    return MPX(rules, idx, T.numrows, sp.csc_matrix((x, i, p), shape=(rows, cols)))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(np.linalg.norm(b), 1e-300)
    return float(np.linalg.norm(a - b) / den)

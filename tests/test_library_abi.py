"""CPU: the C-ABI library loads and exports every symbol include/*.h
declares; the build entry point exists.  No compute calls (no GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(cam_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("cam_ctx_create", "cam_tensor_from_csc", "cam_contract", "cam_heff_create", "cam_heff_apply",
                 "cam_heff_lowest", "cam_svd_truncate", "cam_tebd_gate", "cam_heff_vjp", "cam_idmrg_grow"):
        assert must in names


def test_library_exports_all_declared_symbols():
    from chainamps_amd.capi import library_path
    path = library_path()
    if not path.exists():
        pytest.skip("library not built in this checkout (python -m chainamps_amd.build)")
    lib = ctypes.CDLL(str(path))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    lib.cam_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.cam_version()


def test_python_bindings_cover_the_header():
    from chainamps_amd import capi
    bound = set(capi._SIGS)
    unbound = [n for n in declared_symbols() if n not in bound and n not in ("cam_last_error", "cam_version")]
    assert not unbound, unbound


def test_synthetic_generator_matches_oracle_models():
    """chainamps_amd.workloads (used by bench.py) builds the same tensors as
    the oracle's model generators: coupled-chain MPO and the flux masks."""
    from chainamps_amd import workloads
    from oracle import models
    E, par, sig = workloads.coupled_chain_spectrum(16, seed=3)
    E2, par2, sig2 = models.coupled_chain_spectrum(16, seed=3)
    assert np.array_equal(E, E2) and np.array_equal(par, par2) and np.array_equal(sig, sig2)
    # CSC of a dense tensor in the reference mixed radix
    rules, phys, W, _, _ = models.coupled_chain_model(d=6, seed=0, Jperp=0.1)
    T = W.to_dense_tensor()
    rows, cols, p, i, x = workloads._csc_from_dense(T, 2)
    m = W.matrix.tocsc()
    assert (rows, cols) == m.shape
    assert np.array_equal(p, m.indptr) and np.array_equal(i, m.indices) and np.array_equal(x, m.data)
    qs, sz = workloads.sector_profile(1024)
    qs2, sz2 = models.sector_profile(1024)
    assert np.array_equal(sz, sz2) and sz.sum() == 1024 and len(qs) == 21


def test_coupled_chain_generators():
    """The coupling generators handed to cam_heff_coupling_grad are
    dW/dJperp and dW/dE_n of the coupled-chain MPO: W(J, E) - W(0, E) = J G_J
    and W(J, E + u) - W(J, E) = sum_n u_n G_{E_n} entry by entry."""
    from chainamps_amd import workloads
    from oracle import models
    d, seed, J = 8, 5, 0.37
    rules, phys, W, E, sig = models.coupled_chain_model(d=d, seed=seed, Jperp=J)
    m = W.matrix.tocsc()
    m.sort_indices()
    csc = (m.shape[0], m.shape[1], m.indptr.astype(np.int64), m.indices.astype(np.int64), m.data)
    gens = workloads.coupled_chain_generators(csc, d, seed=seed)
    assert len(gens) == d + 1

    def dense(gen):
        out = np.zeros(m.nnz, complex)
        np.add.at(out, gen[0], gen[1])
        return out
    # W linear in J: the J generator times J is the whole (1, 0) channel
    W0 = models.coupled_chain_model(d=d, seed=seed, Jperp=1e-300)[2].matrix.tocsc()
    W0.sort_indices()
    assert np.allclose(J * dense(gens[0]), m.data - W0.data, rtol=0, atol=1e-15)
    u = np.random.default_rng(1).standard_normal(d)
    Wu = models.coupled_chain_model(d=d, seed=seed, Jperp=J, E=E + u)[2].matrix.tocsc()
    Wu.sort_indices()
    step = sum(u[n] * dense(gens[1 + n]) for n in range(d))
    assert np.allclose(step, Wu.data - m.data, rtol=0, atol=1e-14)
    with pytest.raises(ValueError):
        W00 = models.coupled_chain_model(d=d, seed=seed, Jperp=0.0)[2].matrix.tocsc()
        W00.sort_indices()
        workloads.coupled_chain_generators(
            (W00.shape[0], W00.shape[1], W00.indptr.astype(np.int64), W00.indices.astype(np.int64), W00.data),
            d, seed=seed)


@pytest.mark.parametrize("rules,level", [([2], 0), ([2], 1), ([0], 3)])
def test_product_state_cell_matches_oracle(rules, level):
    """chainamps_amd.workloads (the bench's iTEBD start) builds the oracle's
    product-state unit cell: charges, directions and entries."""
    from chainamps_amd import workloads
    from oracle import tebd
    from oracle.qn import ChargeRules
    d = 6
    phys = (np.arange(d) % 2).reshape(-1, 1) if rules[0] == 2 else (np.arange(d) - 2).reshape(-1, 1)
    got = workloads.product_state_cell_dense(rules, phys, level)
    want = tebd.product_state_cell(ChargeRules(rules), phys, level)[:2]
    for (idx, T), W in zip(got, want):
        assert np.array_equal(T, W.to_dense_tensor())
        for a, b in zip(idx, W.indices):
            assert bool(a.inward) == bool(b.inward) and bool(a.physical) == bool(b.physical)
            assert np.array_equal(np.asarray(a.charges).reshape(-1, 1), np.asarray(b.charges).reshape(-1, 1))

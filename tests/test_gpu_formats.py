"""GPU: the reference's binary formats through the C ABI (cam_tensor_store /
cam_tensor_load, cam_unitcell_store / cam_unitcell_load; MPX_matrix.cpp:
219-247, 594-612, UnitCell.cpp:70-145).  Files written by the library are
byte-for-byte the oracle's restatement of the reference writer; files
written by the oracle load back to the same tensors; a unit cell loaded from
disk drives the device iTEBD exactly as the in-memory one."""
import numpy as np
import pytest

from conftest import to_lib, from_lib
from oracle import formats, models, tebd

pytestmark = pytest.mark.gpu


def _cases():
    rules, phys, W, E, sig = models.coupled_chain_model(d=6, seed=2, Jperp=0.4)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 1)
    return rules, phys, E, [W, tebd.odd_bond_hamiltonian(W), A0, A1]


def test_tensor_store_matches_reference_layout(ctx, tmp_path):
    print("case formats tensor store", flush=True)
    rules, phys, E, tensors = _cases()
    for k, A in enumerate(tensors):
        path = tmp_path / f"t{k}.MPX_matrix"
        to_lib(ctx, A).store(path)
        assert path.read_bytes() == formats.to_bytes(formats.write_mpx, A, len(rules.rules)), k


def test_tensor_load_of_oracle_files(ctx, tmp_path):
    print("case formats tensor load", flush=True)
    import chainamps_amd as cam
    rules, phys, E, tensors = _cases()
    for k, A in enumerate(tensors):
        path = tmp_path / f"o{k}.MPS_matrix"
        path.write_bytes(formats.to_bytes(formats.write_mpx, A, len(rules.rules)))
        T = cam.Tensor.load(ctx, path, rules.rules, phys)
        B = from_lib(T, rules)
        assert (B.matrix != A.matrix).nnz == 0, k
        assert B.numrows == A.numrows
        for a, c in zip(A.indices, B.indices):
            assert a.inward == c.inward and a.physical == c.physical
            assert np.array_equal(np.asarray(a.charges), np.asarray(c.charges))


def test_unitcell_round_trip_and_resume(ctx, tmp_path):
    print("case formats unitcell", flush=True)
    import chainamps_amd as cam
    rules, phys, W, E, sig = models.coupled_chain_model(d=4, seed=7, Jperp=0.4)
    A0, A1, l0, l1 = tebd.product_state_cell(rules, phys, 0)
    # the reference-format cell written by the oracle, loaded by the library
    path = tmp_path / "cell.UNITCELL"
    path.write_bytes(formats.to_bytes(formats.write_unitcell, rules.rules, phys, E, [A0, A1], [l0, l1]))
    r, ch, en, mats, lams = cam.load_unitcell(ctx, path)
    assert tuple(r) == rules.rules and np.array_equal(ch, phys) and np.array_equal(en, E)
    # written back by the library: the same bytes
    out = tmp_path / "again.UNITCELL"
    cam.store_unitcell(out, r, ch, en, mats, lams)
    assert out.read_bytes() == path.read_bytes()
    # resume: iTEBD from the loaded cell follows the oracle from the original
    dt, D = 0.05, 12
    ops = tebd.trotter_operators(W, dt, 2)
    dev = cam.ITEBD(to_lib(ctx, W), mats[0], mats[1], lams[0], lams[1], dt, 2)
    rec = dev.evolve(3, D, 1e-12)
    cell, tr = tebd.evolve(ops, 2, [A0, A1, l0, l1], 3, D, 1e-12)
    for k in (0, 1):
        assert np.allclose(dev.lam(k), cell[2 + k], rtol=1e-10, atol=1e-12)
    assert abs(rec["truncation"] - tr) <= 1e-10

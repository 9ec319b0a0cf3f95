"""GPU: the no-allocation matvec entry point (cam_heff_apply_blocked_noalloc,
no counterpart in the reference) against the allocating one it mirrors
(cam_heff_apply_blocked, FP64).

Two calls in a row on one device, the second with a smaller input, each
bit-identical to the allocating entry point (same kernels, same plans) and
within 1e-12 of the oracle's reference-order matvec; an H_eff whose
intermediates exceed the 256 MiB scratch array is rejected with an error,
and runs once sliced so that they fit.  Device vectors are allocated by the
test itself (hipMalloc through ctypes), as a caller of the entry point
would.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import rel_err, to_lib
from oracle import models, dmrg

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    lib = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return lib


class DevVec:
    """n complex128 on the device (the caller's buffer)."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(16 * max(n, 1))) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, 16 * self.n, _H2D) == 0

    def download(self):
        a = np.empty(self.n, np.complex128)
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, 16 * self.n, _D2H) == 0
        return a

    @property
    def ptr(self):
        return self.p.value

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _u1_env(seed, chi=30, d=7, w=6):
    from oracle.qn import ChargeRules
    rng = np.random.default_rng(seed)
    rules = ChargeRules([0])
    qs, sizes = models.sector_profile(chi, -3, 3, sigma=1.5)
    aq = models.bond_charges_from_profile(qs, sizes)
    phys = rng.integers(-1, 2, d).reshape(-1, 1)
    bq = np.array([0, 1, -1, 0, 2, -2][:w]).reshape(-1, 1)
    W = models.random_qn_mpo(rules, phys, bq, rng)
    return models.random_left_block(rules, aq, bq, rng), W, models.random_right_block(rules, aq, bq, rng)


def _z2_env(seed, chi, d):
    rng = np.random.default_rng(seed)
    rules, phys, W, E, sig = models.coupled_chain_model(d=d, seed=seed)
    aq = (np.arange(chi) % 2).reshape(-1, 1)
    bq = W.indices[1].charges
    return models.random_left_block(rules, aq, bq, rng), W, models.random_right_block(rules, aq, bq, rng)


def _both(hip, H, v):
    """(allocating, no-allocation) outputs for the reference-order vector v,
    in the block layout, and the no-allocation one in reference order.  The
    device-pointer entry points are ordered on the context's stream, so the
    context is synchronised before the buffers are read."""
    n = H.n
    dref, dx, d1, d2 = (DevVec(hip, n) for _ in range(4))
    try:
        dref.upload(v)
        H.to_blocked(dref.ptr, dx.ptr)
        H.apply_device(dx.ptr, d1.ptr)
        H.apply_device_noalloc(dx.ptr, d2.ptr)
        H.from_blocked(d2.ptr, dref.ptr)
        H.ctx.sync()
        return d1.download(), d2.download(), dref.download()
    finally:
        for b in (dref, dx, d1, d2):
            b.free()


def test_noalloc_matches_allocating_entry_point(ctx):
    import chainamps_amd as cam
    hip = _hip()
    rng = np.random.default_rng(31)
    for case, env in (("z2 d=16 chi=40", lambda: _z2_env(2, 40, 16)), ("u1 chi=30 d=7 w=6", lambda: _u1_env(4))):
        print(f"case noalloc {case}", flush=True)
        LB, W, RB = env()
        comp = dmrg.TwoVertexComponents(LB, W, RB)
        H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
        v = rng.standard_normal(H.n) + 1j * rng.standard_normal(H.n)
        y1, y2, out = _both(hip, H, v)
        assert np.array_equal(y1, y2), case
        assert rel_err(out, comp.matvec(v)) <= 1e-12, case
        H.close()


def test_noalloc_rejects_oversized_and_runs_sliced(ctx):
    print("case noalloc oversized d=64 chi=64", flush=True)
    import chainamps_amd as cam
    hip = _hip()
    LB, W, RB = _z2_env(3, 64, 64)
    H = cam.Heff(to_lib(ctx, LB), to_lib(ctx, W), to_lib(ctx, RB))
    assert H.slices()[2] > 256 * 2 ** 20
    n = H.n
    v = np.random.default_rng(5).standard_normal(n) + 0j
    dx, dy = DevVec(hip, n), DevVec(hip, n)
    try:
        with pytest.raises(cam.CamError, match="scratch"):
            H.apply_device_noalloc(dx.ptr, dy.ptr)
    finally:
        dx.free()
        dy.free()
    H.set_workspace_limit(200 * 2 ** 20)
    assert H.slices()[2] <= 256 * 2 ** 20
    y1, y2, out = _both(hip, H, v)
    assert np.array_equal(y1, y2)
    H.close()

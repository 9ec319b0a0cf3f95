"""ctypes bindings of include/chainamps_amd.h.

Mirrors the reference operator interface for the hot path:
``contract`` (MPX_matrix.cpp:744-791), ``reorder`` (630-638),
``Heff`` = TwoVertexComponents (DMRG_routines.cpp:779-974) with its matvec
and lowest-eigenpair solve (replacing arpack_eigs, arpack_interface.hpp:85),
``svd_truncate`` = MPX_matrix::SVD (MPX_matrix.cpp:157-192) and
``tebd_gate`` = iTEBD::apply_and_decompose (TEBD_routines.cpp:233-266).
Errors raise CamError carrying cam_last_error() (the reference prints and
exits).
"""
from __future__ import annotations

import ctypes as C
import os
from collections import namedtuple
from pathlib import Path

import numpy as np

CAM_FP64 = 0
CAM_FP32 = 1
CAM_MEM_HOST = 0
CAM_MEM_DEVICE = 1

_LIB = None
_HERE = Path(__file__).resolve().parent


def library_path() -> Path:
    return Path(os.environ.get("CAM_LIBRARY", _HERE / "lib" / "libchainamps_amd.so"))


class CamError(RuntimeError):
    pass


class _IndexDesc(C.Structure):
    _fields_ = [("inward", C.c_int32), ("physical", C.c_int32), ("size", C.c_int64),
                ("charges", C.POINTER(C.c_int16))]


IndexDesc = namedtuple("IndexDesc", "inward physical charges")

_p = C.c_void_p
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)

_SIGS = {
    "cam_ctx_create": [C.c_int, C.POINTER(_p)],
    "cam_ctx_destroy": [_p],
    "cam_ctx_sync": [_p],
    "cam_ctx_stream": [_p, C.POINTER(_p)],
    "cam_tensor_from_csc": [_p, C.c_int, C.POINTER(C.c_int16), C.c_int, C.POINTER(_IndexDesc), C.c_int,
                            C.c_int64, C.c_int64, _i64p, _i64p, _dp, C.POINTER(_p)],
    "cam_tensor_shape": [_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "cam_tensor_index": [_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _i64p,
                         C.POINTER(C.c_int16)],
    "cam_tensor_nnz": [_p, _i64p],
    "cam_tensor_to_csc": [_p, C.c_int64, _i64p, _i64p, _dp, _i64p],
    "cam_tensor_destroy": [_p],
    "cam_tensor_shift_num_row_indices": [_p, C.c_int],
    "cam_tensor_remove_dummy_indices": [_p, C.c_int, _i64p],
    "cam_tensor_move_dummy_indices": [_p, C.c_int, _i64p],
    "cam_tensor_rescale": [_p, C.c_double, C.c_double],
    "cam_contract": [_p, _p, C.c_int, _p, C.c_int, C.c_int, _i64p, C.POINTER(_p)],
    "cam_reorder": [_p, _p, C.c_int, C.c_int, _i64p, C.c_int, C.POINTER(_p)],
    "cam_heff_create": [_p, _p, _p, _p, C.POINTER(_p), _i64p],
    "cam_heff_destroy": [_p],
    "cam_heff_allowed_map": [_p, _i64p],
    "cam_heff_flops": [_p, _dp],
    "cam_heff_apply": [_p, _p, _p, C.c_int, C.c_int],
    "cam_heff_apply_blocked": [_p, _p, _p, C.c_int],
    "cam_heff_apply_blocked32": [_p, _p, _p],
    "cam_heff_apply_blocked_noalloc": [_p, _p, _p],
    "cam_heff_set_projectors": [_p, C.c_int, C.POINTER(_dp), _dp],
    "cam_ctx_release_workspaces": [_p],
    "cam_heff_profile": [_p, C.c_int, C.c_int, _dp, _dp],
    "cam_heff_profile_device": [_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, _dp, _dp],
    "cam_heff_set_slices": [_p, C.c_int, C.c_int],
    "cam_heff_set_workspace_limit": [_p, C.c_double, C.c_int],
    "cam_heff_get_slices": [_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp],
    "cam_heff_set_launch": [_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "cam_heff_get_launch": [_p, C.c_int, C.POINTER(C.c_int)],
    "cam_heff_stage_bytes": [_p, _dp],
    "cam_heff_to_blocked": [_p, _p, _p],
    "cam_heff_from_blocked": [_p, _p, _p],
    "cam_heff_lowest": [_p, _p, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _p, _dp],
    "cam_heff_lowest_blocked": [_p, _p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp],
    "cam_svd_truncate": [_p, _p, C.c_int64, C.c_double, C.POINTER(_p), C.POINTER(_p), _dp, C.c_int64,
                         _i64p, _dp],
    "cam_tebd_gate": [_p, _p, _p, _p, _dp, C.c_int64, C.c_int64, C.c_double, C.POINTER(_p),
                      C.POINTER(_p), _dp, C.c_int64, _i64p, _dp],
    "cam_heff_vjp": [_p, _p, _p, _p, C.c_int64],
    "cam_heff_coupling_grad": [_p, _p, C.c_int64, C.c_int, _i64p, _i64p, _dp, _dp, _p],
}


def load_library():
    """Load the HIP library; raises CamError when it is missing (there is
    no CPU fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not path.exists():
        raise CamError(f"chainamps_amd native library not built: {path} "
                       "(run `python -m chainamps_amd.build`)")
    lib = C.CDLL(str(path))
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.cam_last_error.restype = C.c_char_p
    lib.cam_last_error.argtypes = []
    lib.cam_version.restype = C.c_char_p
    lib.cam_version.argtypes = []
    _LIB = lib
    return lib


def _check(rc):
    if rc != 0:
        raise CamError(_LIB.cam_last_error().decode())


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Context:
    """cam_ctx: one device, one stream."""

    def __init__(self, device=0):
        lib = load_library()
        h = _p()
        _check(lib.cam_ctx_create(device, C.byref(h)))
        self.h = h

    def release_workspaces(self):
        """Free the grow-only Lanczos and H_eff workspaces of the context."""
        _check(_LIB.cam_ctx_release_workspaces(self.h))

    def sync(self):
        _check(_LIB.cam_ctx_sync(self.h))

    def stream(self):
        s = _p()
        _check(_LIB.cam_ctx_stream(self.h, C.byref(s)))
        return s.value

    def close(self):
        if getattr(self, "h", None):
            _LIB.cam_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tensor:
    """Device QN-blocked tensor (cam_tensor)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    @staticmethod
    def from_csc(ctx, rules, indices, numrows, rows, cols, p, i, x):
        """indices: list of IndexDesc(inward, physical, charges[size, nrules]);
        x: complex128 array."""
        rules = np.ascontiguousarray(rules, dtype=np.int16)
        nr = len(rules)
        descs = (_IndexDesc * max(len(indices), 1))()
        keep = []
        for k, ix in enumerate(indices):
            ch = np.ascontiguousarray(np.asarray(ix.charges, dtype=np.int16).reshape(-1, nr))
            keep.append(ch)
            descs[k].inward = int(bool(ix.inward))
            descs[k].physical = int(bool(ix.physical))
            descs[k].size = ch.shape[0]
            descs[k].charges = _ptr(ch, C.c_int16)
        p = np.ascontiguousarray(p, dtype=np.int64)
        i = np.ascontiguousarray(i, dtype=np.int64)
        xv = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64)
        if i.size == 0:
            i = np.zeros(1, np.int64)
            xv = np.zeros(2, np.float64)
        h = _p()
        _check(_LIB.cam_tensor_from_csc(ctx.h, nr, _ptr(rules, C.c_int16), len(indices), descs, numrows,
                                         rows, cols, _ptr(p, C.c_int64), _ptr(i, C.c_int64),
                                         _ptr(xv, C.c_double), C.byref(h)))
        return Tensor(ctx, h)

    def shape(self):
        n, r, nr = C.c_int(), C.c_int(), C.c_int()
        _check(_LIB.cam_tensor_shape(self.h, C.byref(n), C.byref(r), C.byref(nr)))
        return n.value, r.value, nr.value

    def indices(self):
        n, _, nr = self.shape()
        out = []
        for k in range(n):
            inw, ph, sz = C.c_int32(), C.c_int32(), C.c_int64()
            _check(_LIB.cam_tensor_index(self.h, k, C.byref(inw), C.byref(ph), C.byref(sz), None))
            ch = np.zeros((sz.value, nr), np.int16)
            _check(_LIB.cam_tensor_index(self.h, k, None, None, None, _ptr(ch, C.c_int16)))
            out.append(IndexDesc(bool(inw.value), bool(ph.value), ch.astype(np.int64)))
        return out

    @property
    def numrows(self):
        return self.shape()[1]

    def stored(self):
        n = C.c_int64()
        _check(_LIB.cam_tensor_nnz(self.h, C.byref(n)))
        return n.value

    def to_csc(self):
        """Returns (rows, cols, p, i, x complex)."""
        idx = self.indices()
        nrow = self.numrows
        rows = int(np.prod([ix.charges.shape[0] for ix in idx[:nrow]], dtype=np.int64))
        cols = int(np.prod([ix.charges.shape[0] for ix in idx[nrow:]], dtype=np.int64))
        nnz = C.c_int64()
        _check(_LIB.cam_tensor_to_csc(self.h, 0, None, None, None, C.byref(nnz)))
        cap = max(nnz.value, 1)
        p = np.zeros(cols + 1, np.int64)
        i = np.zeros(cap, np.int64)
        x = np.zeros(2 * cap, np.float64)
        _check(_LIB.cam_tensor_to_csc(self.h, cap, _ptr(p, C.c_int64), _ptr(i, C.c_int64),
                                      _ptr(x, C.c_double), C.byref(nnz)))
        n = nnz.value
        return rows, cols, p, i[:n], x[:2 * n].view(np.complex128)

    def shift_num_row_indices(self, n):
        _check(_LIB.cam_tensor_shift_num_row_indices(self.h, n))
        return self

    def remove_dummy_indices(self, which):
        w = np.ascontiguousarray(which, dtype=np.int64)
        _check(_LIB.cam_tensor_remove_dummy_indices(self.h, len(w), _ptr(w, C.c_int64)))
        return self

    def move_dummy_indices(self, order):
        o = np.ascontiguousarray(order, dtype=np.int64)
        _check(_LIB.cam_tensor_move_dummy_indices(self.h, len(o), _ptr(o, C.c_int64)))
        return self

    def rescale(self, factor):
        f = complex(factor)
        _check(_LIB.cam_tensor_rescale(self.h, f.real, f.imag))
        return self

    def store(self, path):
        """MPX_matrix::store: the reference's .MPX_matrix / .MPS_matrix format."""
        _check(_LIB.cam_tensor_store(self.h, os.fsencode(path)))

    @staticmethod
    def load(ctx, path, rules, phys_charges):
        """load_MPX_matrix: physical indices take the basis charges
        phys_charges (d x nrules)."""
        rules = np.ascontiguousarray(rules, dtype=np.int16)
        ch = np.ascontiguousarray(np.asarray(phys_charges, dtype=np.int16).reshape(-1, len(rules)))
        h = _p()
        _check(_LIB.cam_tensor_load(ctx.h, os.fsencode(path), len(rules), _ptr(rules, C.c_int16), ch.shape[0],
                                    _ptr(ch, C.c_int16), C.byref(h)))
        return Tensor(ctx, h)

    def close(self):
        if getattr(self, "h", None):
            _LIB.cam_tensor_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def contract(A: Tensor, conjA: bool, B: Tensor, conjB: bool, pairs) -> Tensor:
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int64).reshape(-1))
    h = _p()
    _check(_LIB.cam_contract(A.ctx.h, A.h, int(conjA), B.h, int(conjB), len(pr) // 2,
                             _ptr(pr, C.c_int64), C.byref(h)))
    return Tensor(A.ctx, h)


def reorder(A: Tensor, conjA: bool, order, numrows) -> Tensor:
    o = np.ascontiguousarray(order, dtype=np.int64)
    h = _p()
    _check(_LIB.cam_reorder(A.ctx.h, A.h, int(conjA), len(o), _ptr(o, C.c_int64), numrows, C.byref(h)))
    return Tensor(A.ctx, h)


class Heff:
    """TwoVertexComponents on the device."""

    def __init__(self, L: Tensor, W: Tensor, R: Tensor):
        self.ctx = L.ctx
        self._keep = (L, W, R)
        h = _p()
        n = C.c_int64()
        _check(_LIB.cam_heff_create(L.ctx.h, L.h, W.h, R.h, C.byref(h), C.byref(n)))
        self.h = h
        self.n = n.value

    def allowed(self):
        out = np.zeros(max(self.n, 1), np.int64)
        _check(_LIB.cam_heff_allowed_map(self.h, _ptr(out, C.c_int64)))
        return out[:self.n]

    @property
    def flops(self):
        f = C.c_double()
        _check(_LIB.cam_heff_flops(self.h, C.byref(f)))
        return f.value

    def apply(self, vin, precision=CAM_FP64):
        vin = np.ascontiguousarray(vin, dtype=np.complex128)
        out = np.zeros(self.n, np.complex128)
        _check(_LIB.cam_heff_apply(self.h, vin.ctypes.data, out.ctypes.data, CAM_MEM_HOST, precision))
        return out

    def apply_device(self, din, dout, precision=CAM_FP64, blocked=True):
        """din/dout: device pointers (ints) of n complex128."""
        if blocked:
            _check(_LIB.cam_heff_apply_blocked(self.h, din, dout, precision))
        else:
            _check(_LIB.cam_heff_apply(self.h, din, dout, CAM_MEM_DEVICE, precision))

    def set_projectors(self, projectors):
        """(P, weight) pairs: P the projector tensor's entries in the
        reference compressed order (TwoVertexComponents' ProjectorTensors)."""
        vecs = [np.ascontiguousarray(P, dtype=np.complex128) for P, _ in projectors]
        for v in vecs:
            if v.shape != (self.n,):
                raise ValueError("projector length differs from the compressed vector length")
        w = np.ascontiguousarray([float(wt) for _, wt in projectors], dtype=np.float64)
        ptrs = (_dp * max(len(vecs), 1))(*[_ptr(v.view(np.float64), C.c_double) for v in vecs])
        _check(_LIB.cam_heff_set_projectors(self.h, len(vecs), ptrs, _ptr(w, C.c_double) if len(w) else None))
        self._projectors = vecs

    def apply_device_noalloc(self, din, dout):
        """FP64 matvec on device block-layout vectors with the intermediates
        in the library's per-device scratch array (no device allocation)."""
        _check(_LIB.cam_heff_apply_blocked_noalloc(self.h, din, dout))

    def apply_device32(self, din, dout):
        """din/dout: device pointers (ints) of n complex64 (split-complex
        FP32) in the block layout: the mixed-precision solver's matvec."""
        _check(_LIB.cam_heff_apply_blocked32(self.h, din, dout))

    def profile(self, precision=CAM_FP64, reps=5):
        """Average ms per stage (L, W, W, R) from HIP events and the
        per-stage algorithmic flops."""
        ms = np.zeros(4)
        fl = np.zeros(4)
        _check(_LIB.cam_heff_profile(self.h, precision, reps, _ptr(ms, C.c_double), _ptr(fl, C.c_double)))
        return ms, fl

    def profile_device(self, din, dout, precision=CAM_FP64, reps=5):
        """As profile(), on the caller's device block-layout vectors."""
        ms = np.zeros(4)
        fl = np.zeros(4)
        _check(_LIB.cam_heff_profile_device(self.h, precision, reps, din, dout, _ptr(ms, C.c_double),
                                            _ptr(fl, C.c_double)))
        return ms, fl

    def set_slices(self, row_slices, col_slices):
        """Run the matvec in row_slices x col_slices slices of the
        intermediates (1 x 1: unsliced)."""
        _check(_LIB.cam_heff_set_slices(self.h, int(row_slices), int(col_slices)))

    def set_workspace_limit(self, nbytes, precision=CAM_FP64):
        """Fewest slices whose intermediates fit nbytes at `precision`."""
        _check(_LIB.cam_heff_set_workspace_limit(self.h, float(nbytes), int(precision)))

    def slices(self):
        """(row slices, column slices, workspace bytes)."""
        r, c, b = C.c_int(), C.c_int(), C.c_double()
        _check(_LIB.cam_heff_get_slices(self.h, C.byref(r), C.byref(c), C.byref(b)))
        return r.value, c.value, b.value

    def stage_bytes(self):
        """Algorithmic HBM bytes of each FP64 stage (L, W, W, R)."""
        b = np.zeros(4)
        _check(_LIB.cam_heff_stage_bytes(self.h, _ptr(b, C.c_double)))
        return b.tolist()

    def set_launch(self, precision=CAM_FP64, algo=1, tile_mode=1, wg_per_cu=4, kchunk=16):
        """Fix the stage GEMM launch configuration (disables autotuning)."""
        _check(_LIB.cam_heff_set_launch(self.h, precision, algo, tile_mode, wg_per_cu, kchunk))

    def get_launch(self, precision=CAM_FP64):
        """[(algo, tile_mode, wg_per_cu, kchunk)] for the stages L, W, W, R."""
        cfg = np.zeros(16, np.int32)
        _check(_LIB.cam_heff_get_launch(self.h, precision, _ptr(cfg, C.c_int)))
        return [tuple(int(v) for v in cfg[4 * i:4 * i + 4]) for i in range(4)]

    def to_blocked(self, din, dout):
        _check(_LIB.cam_heff_to_blocked(self.h, din, dout))

    def from_blocked(self, din, dout):
        _check(_LIB.cam_heff_from_blocked(self.h, din, dout))

    def lowest(self, guess=None, converge=True, tol=0.0, ncv=20, mixed=False):
        """Returns (eigenvalue, eigenvector (reference order), info dict)."""
        ev = C.c_double()
        vec = np.zeros(self.n, np.complex128)
        info = np.zeros(8, np.float64)
        g = None
        if guess is not None:
            g = np.ascontiguousarray(guess, dtype=np.complex128)
        _check(_LIB.cam_heff_lowest(self.h, None if g is None else g.ctypes.data, int(converge), tol, ncv,
                                    int(mixed), C.byref(ev), vec.ctypes.data, _ptr(info, C.c_double)))
        return ev.value, vec, dict(matvecs_fp64=int(info[0]), matvecs_fp32=int(info[1]),
                                   restarts=int(info[2]), residual=float(info[3]))

    def lowest_device(self, dx, have_start=False, converge=True, tol=0.0, ncv=20, mixed=False):
        """Eigensolve on a device block-layout vector (pointer `dx`, n
        complex128), overwritten by the eigenvector; returns (E, info)."""
        ev = C.c_double()
        info = np.zeros(8, np.float64)
        _check(_LIB.cam_heff_lowest_blocked(self.h, C.c_void_p(dx), int(have_start), int(converge), tol, ncv,
                                            int(mixed), C.byref(ev), _ptr(info, C.c_double)))
        return ev.value, dict(matvecs_fp64=int(info[0]), matvecs_fp32=int(info[1]),
                              restarts=int(info[2]), residual=float(info[3]))

    def vjp(self, psi, nnzW):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        dpsi = np.zeros(self.n, np.complex128)
        dW = np.zeros(max(nnzW, 1), np.complex128)
        _check(_LIB.cam_heff_vjp(self.h, psi.ctypes.data, dpsi.ctypes.data, dW.ctypes.data, nnzW))
        return dpsi, dW[:nnzW]

    def coupling_grad(self, psi, nnzW, generators, want_dpsi=False):
        """dE/d(theta_c) for W linear in couplings theta_c: `generators` is a
        list of (entries, coefs) pairs, dW/dtheta_c over W's stored entries
        (CSC order) with complex coefficients.  Returns the complex array of
        dE/dtheta_c (imaginary parts vanish for a Hermitian MPO) and, when
        asked, dE/dconj(psi)."""
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        ptr = np.zeros(len(generators) + 1, np.int64)
        ent, cof = [], []
        for c, (e, v) in enumerate(generators):
            e = np.asarray(e, np.int64).ravel()
            v = np.broadcast_to(np.asarray(v, np.complex128), e.shape)
            ptr[c + 1] = ptr[c] + len(e)
            ent.append(e)
            cof.append(v)
        ent = np.ascontiguousarray(np.concatenate(ent) if ent else np.zeros(1, np.int64))
        cof = np.ascontiguousarray(np.concatenate(cof) if cof else np.zeros(1, np.complex128))
        if len(ent) == 0:
            ent, cof = np.zeros(1, np.int64), np.zeros(1, np.complex128)
        dE = np.zeros(max(len(generators), 1), np.complex128)
        dpsi = np.zeros(self.n, np.complex128) if want_dpsi else None
        _check(_LIB.cam_heff_coupling_grad(self.h, psi.ctypes.data, nnzW, len(generators), _ptr(ptr, C.c_int64),
                                           _ptr(ent, C.c_int64), _ptr(cof.view(np.float64), C.c_double),
                                           _ptr(dE.view(np.float64), C.c_double),
                                           dpsi.ctypes.data if want_dpsi else None))
        dE = dE[:len(generators)]
        return (dE, dpsi) if want_dpsi else dE

    def close(self):
        if getattr(self, "h", None):
            _LIB.cam_heff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def svd_truncate(theta: Tensor, D=0, max_trunc=0.0):
    """Returns (U, values, Vd, discarded)."""
    ctx = theta.ctx
    cap = 1 << 22
    vals = np.zeros(cap, np.float64)
    nv = C.c_int64()
    disc = C.c_double()
    U, V = _p(), _p()
    _check(_LIB.cam_svd_truncate(ctx.h, theta.h, int(D), float(max_trunc), C.byref(U), C.byref(V),
                                 _ptr(vals, C.c_double), cap, C.byref(nv), C.byref(disc)))
    return Tensor(ctx, U), vals[:nv.value].copy(), Tensor(ctx, V), disc.value


def tebd_gate(G: Tensor, A0: Tensor, A1: Tensor, lambda0, D=0, minS=0.0):
    """Returns (A0new, A1new, lambda1, truncation)."""
    ctx = G.ctx
    l0 = np.ascontiguousarray(lambda0, dtype=np.float64)
    cap = 1 << 22
    l1 = np.zeros(cap, np.float64)
    n1 = C.c_int64()
    tr = C.c_double()
    a, b = _p(), _p()
    _check(_LIB.cam_tebd_gate(ctx.h, G.h, A0.h, A1.h, _ptr(l0, C.c_double), len(l0), int(D), float(minS),
                              C.byref(a), C.byref(b), _ptr(l1, C.c_double), cap, C.byref(n1), C.byref(tr)))
    return Tensor(ctx, a), Tensor(ctx, b), l1[:n1.value].copy(), tr.value


_SIGS_GATES = {
    "cam_extract_mpo_block": [_p, _p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.POINTER(_p)],
    "cam_bond_hamiltonian": [_p, _p, C.c_int, C.POINTER(_p)],
    "cam_bond_evolution_operator": [_p, _p, C.c_double, C.POINTER(_p)],
}
_SIGS.update(_SIGS_GATES)


def extract_mpo_block(W: Tensor, rows, cols) -> Tensor:
    """MPO_matrix::ExtractMPOBlock: rows=(r0, r1), cols=(c0, c1) inclusive."""
    h = _p()
    _check(_LIB.cam_extract_mpo_block(W.ctx.h, W.h, int(rows[0]), int(rows[1]), int(cols[0]), int(cols[1]),
                                      C.byref(h)))
    return Tensor(W.ctx, h)


def bond_hamiltonian(W: Tensor, odd=True) -> Tensor:
    """MakeOddBondHamiltonian / MakeEvenBondHamiltonian."""
    h = _p()
    _check(_LIB.cam_bond_hamiltonian(W.ctx.h, W.h, int(bool(odd)), C.byref(h)))
    return Tensor(W.ctx, h)


def bond_evolution_operator(H: Tensor, timestep) -> Tensor:
    """MakeBondEvolutionOperator: exp(-i timestep H) per charge block."""
    h = _p()
    _check(_LIB.cam_bond_evolution_operator(H.ctx.h, H.h, float(timestep), C.byref(h)))
    return Tensor(H.ctx, h)


_SIGS_IDMRG = {
    "cam_idmrg_create": [_p, _p, C.POINTER(C.c_int16), C.c_int, C.c_int, C.c_int, C.POINTER(_p)],
    "cam_idmrg_grow": [_p, C.c_int64, C.c_double, _dp],
    "cam_idmrg_values": [_p, _dp, C.c_int64, _i64p],
    "cam_idmrg_destroy": [_p],
    "cam_idmrg_move": [_p, C.c_int, C.c_int64, C.c_double, _dp],
    "cam_fdmrg_sweeps": [_p, C.c_int64, C.c_int64, C.c_double, _dp, C.POINTER(C.c_int32), C.c_int64, _i64p],
    "cam_idmrg_blocks": [_p, C.POINTER(_p), C.POINTER(_p)],
}
_SIGS.update(_SIGS_IDMRG)

IDMRG_FIELDS = ("N", "energy", "energy_per_vertex", "entropy", "truncation", "chi", "n",
                "matvecs_fp64", "matvecs_fp32", "residual", "seconds", "heff_flops",
                "eig_seconds", "svd_seconds", "block_seconds", "fidelity", "guess_dropped", "left", "right")


def _idmrg_record(rec):
    d = dict(zip(IDMRG_FIELDS, np.asarray(rec).tolist()))
    for k in ("N", "chi", "n", "matvecs_fp64", "matvecs_fp32", "left", "right"):
        d[k] = int(d[k])
    d["guess_dropped"] = bool(d["guess_dropped"])
    return d


class IDMRG:
    """Device iDMRG growth (iDMRG::run / SuperBlock::grow_two_vertex)."""

    def __init__(self, W: Tensor, target=None, converge=False, mixed=False, ncv=20):
        self.ctx = W.ctx
        self._W = W
        nr = W.shape()[2]
        t = np.zeros(nr, np.int16) if target is None else np.ascontiguousarray(target, dtype=np.int16)
        h = _p()
        _check(_LIB.cam_idmrg_create(W.ctx.h, W.h, _ptr(t, C.c_int16), int(converge), int(mixed), ncv,
                                     C.byref(h)))
        self.h = h

    def unitcell(self):
        """UnitCell(getCentralDecomposition(), getPreviousLambda()) of the
        last growth step (iDMRG_DRV.cpp:70), as a UnitCell."""
        h = _p()
        _check(_LIB.cam_idmrg_unitcell(self.h, C.byref(h)))
        return UnitCell(self.ctx, h)

    def grow(self, chi, truncation=0.0):
        rec = np.zeros(len(IDMRG_FIELDS))
        _check(_LIB.cam_idmrg_grow(self.h, int(chi), float(truncation), _ptr(rec, C.c_double)))
        return _idmrg_record(rec)

    def move(self, right, chi, truncation=0.0):
        """One finite step: SuperBlock::move_right_two_vertex (right=True)
        or move_left_two_vertex (DMRG_routines.cpp:376-444)."""
        rec = np.zeros(len(IDMRG_FIELDS))
        _check(_LIB.cam_idmrg_move(self.h, int(bool(right)), int(chi), float(truncation), _ptr(rec, C.c_double)))
        return _idmrg_record(rec)

    def finite_sweeps(self, num_sweeps, chi, truncation=0.0):
        """FiniteDMRG::run (fDMRG_routines.cpp:19-66): (every step's record,
        the records the reference writes to Energies.dat)."""
        n = C.c_int64()
        _check(_LIB.cam_fdmrg_sweeps(self.h, int(num_sweeps), int(chi), float(truncation), None, None, 0,
                                     C.byref(n)))
        recs = np.zeros((max(n.value, 1), len(IDMRG_FIELDS)))
        mid = np.zeros(max(n.value, 1), np.int32)
        _check(_LIB.cam_fdmrg_sweeps(self.h, int(num_sweeps), int(chi), float(truncation), _ptr(recs, C.c_double),
                                     _ptr(mid, C.c_int32), n.value, C.byref(n)))
        steps = [_idmrg_record(r) for r in recs[:n.value]]
        return steps, [r for r, m in zip(steps, mid[:n.value]) if m]

    def blocks(self):
        """The left and right blocks of the current centre."""
        l, r = _p(), _p()
        _check(_LIB.cam_idmrg_blocks(self.h, C.byref(l), C.byref(r)))
        return Tensor(self.ctx, l), Tensor(self.ctx, r)

    def values(self):
        n = C.c_int64()
        _check(_LIB.cam_idmrg_values(self.h, None, 0, C.byref(n)))
        v = np.zeros(max(n.value, 1))
        _check(_LIB.cam_idmrg_values(self.h, _ptr(v, C.c_double), len(v), C.byref(n)))
        return v[:n.value]

    def close(self):
        if getattr(self, "h", None):
            _LIB.cam_idmrg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_SIGS_ITEBD = {
    "cam_itebd_create": [_p, _p, _p, _p, _dp, C.c_int64, _dp, C.c_int64, C.c_double, C.c_int, C.POINTER(_p)],
    "cam_itebd_evolve": [_p, C.c_int64, C.c_int64, C.c_double, _dp],
    "cam_itebd_evolve_measured": [_p, C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int, C.POINTER(_p), _dp,
                                  C.c_int64, _i64p],
    "cam_itebd_lambda": [_p, C.c_int, _dp, C.c_int64, _i64p],
    "cam_itebd_tensor": [_p, C.c_int, C.POINTER(_p)],
    "cam_itebd_destroy": [_p],
}
_SIGS.update(_SIGS_ITEBD)

ITEBD_FIELDS = ("steps", "time", "truncation", "entropy", "chi", "seconds", "gate_flops")

_SIGS_FTEBD = {
    "cam_ftebd_create": [_p, _p, C.POINTER(_p), C.c_int64, C.c_double, C.c_int, C.POINTER(_p)],
    "cam_ftebd_evolve": [_p, C.c_int64, C.c_int64, C.c_double, C.c_int64],
    "cam_ftebd_records": [_p, _dp, C.c_int64, _i64p],
    "cam_ftebd_site": [_p, C.c_int64, C.POINTER(_p)],
    "cam_ftebd_destroy": [_p],
}
_SIGS.update(_SIGS_FTEBD)
FTEBD_FIELDS = ("step", "time", "truncation", "entropy", "overlap_re", "overlap_im", "abs_overlap")


class ITEBD:
    """Device iTEBD (TrotterDecomposition + iTEBD::evolve, orders 1, 2 and 4)."""

    def __init__(self, W: Tensor, A0: Tensor, A1: Tensor, lambda0, lambda1, dt, order=2):
        self.ctx = W.ctx
        l0 = np.ascontiguousarray(lambda0, dtype=np.float64)
        l1 = np.ascontiguousarray(lambda1, dtype=np.float64)
        h = _p()
        _check(_LIB.cam_itebd_create(W.ctx.h, W.h, A0.h, A1.h, _ptr(l0, C.c_double), len(l0),
                                     _ptr(l1, C.c_double), len(l1), float(dt), int(order), C.byref(h)))
        self.h = h

    def evolve(self, steps, D=0, minS=0.0):
        rec = np.zeros(7)
        _check(_LIB.cam_itebd_evolve(self.h, int(steps), int(D), float(minS), _ptr(rec, C.c_double)))
        d = dict(zip(ITEBD_FIELDS, rec.tolist()))
        d["steps"] = int(d["steps"])
        d["chi"] = int(d["chi"])
        return d

    def evolve_measured(self, steps, D=0, minS=0.0, interval=1, measured=()):
        """iTEBD::evolve with measurement steps: one dict per measurement
        (step, time, truncation, entropy, abs_overlap, overlap, values)."""
        ops = list(measured)
        hs = (_p * max(len(ops), 1))(*[o.h for o in ops])
        n = C.c_int64()
        per = 7 + 2 * len(ops)
        cap = int(steps) // max(int(interval), 1) + 1
        rec = np.zeros(cap * per)
        _check(_LIB.cam_itebd_evolve_measured(self.h, int(steps), int(D), float(minS), int(interval), len(ops), hs,
                                              _ptr(rec, C.c_double), cap, C.byref(n)))
        out = []
        for k in range(min(n.value, cap)):
            r = rec[k * per:(k + 1) * per]
            out.append(dict(step=int(r[0]), time=r[1], truncation=r[2], entropy=r[3], abs_overlap=r[4],
                            overlap=complex(r[5], r[6]),
                            values=[complex(r[7 + 2 * j], r[8 + 2 * j]) for j in range(len(ops))]))
        return out

    def lam(self, k):
        n = C.c_int64()
        _check(_LIB.cam_itebd_lambda(self.h, int(k), None, 0, C.byref(n)))
        v = np.zeros(max(n.value, 1))
        _check(_LIB.cam_itebd_lambda(self.h, int(k), _ptr(v, C.c_double), len(v), C.byref(n)))
        return v[:n.value]

    def tensor(self, k) -> Tensor:
        h = _p()
        _check(_LIB.cam_itebd_tensor(self.h, int(k), C.byref(h)))
        return Tensor(self.ctx, h)

    def close(self):
        if getattr(self, "h", None):
            _LIB.cam_itebd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FTEBD:
    """Device finite-chain TEBD (TEBD::evolve with apply_to_odd/even_bonds and
    left_canonise, Trotter orders 1, 2, 4) on a left-canonical MPS given as
    a list of site tensors [s, a_l (in), a_r (out)]."""

    def __init__(self, W: Tensor, sites, dt, order=2):
        self.ctx = W.ctx
        arr = (_p * len(sites))(*[t.h for t in sites])
        h = _p()
        _check(_LIB.cam_ftebd_create(W.ctx.h, W.h, arr, len(sites), float(dt), int(order), C.byref(h)))
        self.h = h
        self.N = len(sites)

    def evolve(self, steps, D=0, minS=0.0, interval=1):
        _check(_LIB.cam_ftebd_evolve(self.h, int(steps), int(D), float(minS), int(interval)))

    def records(self):
        n = C.c_int64()
        _check(_LIB.cam_ftebd_records(self.h, None, 0, C.byref(n)))
        v = np.zeros(max(7 * n.value, 1))
        _check(_LIB.cam_ftebd_records(self.h, _ptr(v, C.c_double), n.value, C.byref(n)))
        out = []
        for i in range(n.value):
            d = dict(zip(FTEBD_FIELDS, v[7 * i:7 * i + 7].tolist()))
            d["step"] = int(d["step"])
            d["overlap"] = complex(d.pop("overlap_re"), d.pop("overlap_im"))
            out.append(d)
        return out

    def site(self, v) -> Tensor:
        h = _p()
        _check(_LIB.cam_ftebd_site(self.h, int(v), C.byref(h)))
        return Tensor(self.ctx, h)

    def close(self):
        if getattr(self, "h", None):
            _LIB.cam_ftebd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_SIGS_STORAGE = {
    "cam_tensor_store": [_p, C.c_char_p],
    "cam_tensor_load": [_p, C.c_char_p, C.c_int, C.POINTER(C.c_int16), C.c_int64, C.POINTER(C.c_int16),
                        C.POINTER(_p)],
    "cam_unitcell_store": [C.c_char_p, C.c_int, C.POINTER(C.c_int16), C.c_int64, C.POINTER(C.c_int16), _dp, C.c_int,
                           C.POINTER(_p), C.POINTER(_dp), _i64p],
    "cam_unitcell_load": [_p, C.c_char_p, C.POINTER(_p)],
    "cam_unitcell_basis": [_p, C.POINTER(C.c_int), C.POINTER(C.c_int16), _i64p, C.POINTER(C.c_int16), _dp,
                           C.c_int64],
    "cam_unitcell_size": [_p, C.POINTER(C.c_int)],
    "cam_unitcell_matrix": [_p, C.c_int, C.POINTER(_p)],
    "cam_unitcell_lambda": [_p, C.c_int, _dp, C.c_int64, _i64p],
    "cam_unitcell_destroy": [_p],
}
_SIGS.update(_SIGS_STORAGE)


def store_unitcell(path, rules, phys_charges, energies, matrices, lambdas):
    """UnitCell::store: basis (rules, charges d x nrules, energies), the MPS
    matrices and their lambdas, in the reference's .UNITCELL format."""
    rules = np.ascontiguousarray(rules, dtype=np.int16)
    ch = np.ascontiguousarray(np.asarray(phys_charges, dtype=np.int16).reshape(-1, len(rules)))
    en = np.ascontiguousarray(energies, dtype=np.float64)
    lam = [np.ascontiguousarray(l, dtype=np.float64) for l in lambdas]
    n = len(matrices)
    mh = (_p * max(n, 1))(*[m.h for m in matrices])
    lp = (_dp * max(n, 1))(*[_ptr(l, C.c_double) for l in lam])
    ls = np.array([len(l) for l in lam] or [0], dtype=np.int64)
    _check(_LIB.cam_unitcell_store(os.fsencode(path), len(rules), _ptr(rules, C.c_int16), ch.shape[0],
                                   _ptr(ch, C.c_int16), _ptr(en, C.c_double), n, mh, lp, _ptr(ls, C.c_int64)))


def load_unitcell(ctx, path):
    """load_UnitCell_binary: returns (rules, charges, energies, matrices,
    lambdas)."""
    h = _p()
    _check(_LIB.cam_unitcell_load(ctx.h, os.fsencode(path), C.byref(h)))
    try:
        nr, d = C.c_int(), C.c_int64()
        _check(_LIB.cam_unitcell_basis(h, C.byref(nr), None, C.byref(d), None, None, 0))
        rules = np.zeros(nr.value, np.int16)
        ch = np.zeros((d.value, nr.value), np.int16)
        en = np.zeros(d.value)
        _check(_LIB.cam_unitcell_basis(h, None, _ptr(rules, C.c_int16), None, _ptr(ch, C.c_int16),
                                       _ptr(en, C.c_double), d.value))
        n = C.c_int()
        _check(_LIB.cam_unitcell_size(h, C.byref(n)))
        mats, lams = [], []
        for k in range(n.value):
            t = _p()
            _check(_LIB.cam_unitcell_matrix(h, k, C.byref(t)))
            mats.append(Tensor(ctx, t))
            m = C.c_int64()
            _check(_LIB.cam_unitcell_lambda(h, k, None, 0, C.byref(m)))
            v = np.zeros(max(m.value, 1))
            _check(_LIB.cam_unitcell_lambda(h, k, _ptr(v, C.c_double), len(v), C.byref(m)))
            lams.append(v[:m.value])
        return rules, ch, en, mats, lams
    finally:
        _LIB.cam_unitcell_destroy(h)


_SIGS_ORTHO = {
    "cam_unitcell_create": [_p, C.c_int, C.POINTER(_p), C.POINTER(_dp), _i64p, C.POINTER(_p)],
    "cam_unitcell_from_decomposition": [_p, _p, _dp, C.c_int64, _p, _dp, C.c_int64, C.POINTER(_p)],
    "cam_idmrg_unitcell": [_p, C.POINTER(_p)],
    "cam_transfer_apply": [_p, C.c_int, _dp, _dp],
    "cam_transfer_leading": [_p, C.c_int, C.c_double, C.c_int, _dp, _dp],
    "cam_orthogonalise": [_p, C.POINTER(_p), _dp],
    "cam_orthogonalise_inversion_symmetric": [_p, C.POINTER(_p), _dp],
    "cam_itwovertex_energy": [_p, _p, _dp, _dp],
}
_SIGS.update(_SIGS_ORTHO)

ORTHO_FIELDS = ("left_re", "left_im", "right_re", "right_im", "entropy", "chi", "left_matvecs", "right_matvecs",
                "left_residual", "right_residual", "kept", "seconds")


class UnitCell:
    """A unit cell on the device (UnitCell): MPS matrices [s, a_l, a_r] and
    the lambda on the left of each."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h

    @classmethod
    def create(cls, ctx, matrices, lambdas):
        """UnitCell(spectrum, matrices, lambdas) (UnitCell.hpp:41)."""
        n = len(matrices)
        hs = (_p * n)(*[m.h for m in matrices])
        ls = [np.ascontiguousarray(l, dtype=np.float64) for l in lambdas]
        lp = (_dp * n)(*[_ptr(l, C.c_double) for l in ls])
        sz = np.array([len(l) for l in ls], np.int64)
        h = _p()
        _check(_LIB.cam_unitcell_create(ctx.h, n, hs, lp, _ptr(sz, C.c_int64), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_decomposition(cls, U, values, Vd, previous):
        """UnitCell(MPSDecomposition, PreviousLambda) (UnitCell.hpp:42-48)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        p = np.ascontiguousarray(previous, dtype=np.float64)
        h = _p()
        _check(_LIB.cam_unitcell_from_decomposition(U.ctx.h, U.h, _ptr(v, C.c_double), len(v), Vd.h,
                                                    _ptr(p, C.c_double), len(p), C.byref(h)))
        return cls(U.ctx, h)

    def size(self):
        n = C.c_int()
        _check(_LIB.cam_unitcell_size(self.h, C.byref(n)))
        return n.value

    def matrix(self, k) -> Tensor:
        t = _p()
        _check(_LIB.cam_unitcell_matrix(self.h, int(k), C.byref(t)))
        return Tensor(self.ctx, t)

    def lam(self, k):
        m = C.c_int64()
        _check(_LIB.cam_unitcell_lambda(self.h, int(k), None, 0, C.byref(m)))
        v = np.zeros(max(m.value, 1))
        _check(_LIB.cam_unitcell_lambda(self.h, int(k), _ptr(v, C.c_double), len(v), C.byref(m)))
        return v[:m.value]

    def transfer_apply(self, side, vec):
        """Left/RightComponentsMultiply on the allowed entries in the
        reference's order (c outer, r inner)."""
        x = np.ascontiguousarray(vec, dtype=np.complex128)
        y = np.zeros_like(x)
        _check(_LIB.cam_transfer_apply(self.h, int(side), _ptr(x.view(np.float64), C.c_double),
                                       _ptr(y.view(np.float64), C.c_double)))
        return y

    def transfer_leading(self, side, tol=0.0, ncv=0):
        """Leading eigenvalue of the left (side 0) or right (side 1)
        transfer map (TransferMatrixComponents::LeftED / RightED)."""
        val = np.zeros(2)
        info = np.zeros(4)
        _check(_LIB.cam_transfer_leading(self.h, int(side), float(tol), int(ncv), _ptr(val, C.c_double),
                                         _ptr(info, C.c_double)))
        return complex(val[0], val[1]), dict(matvecs=int(info[0]), restarts=int(info[1]), residual=info[2],
                                             converged=bool(info[3]))

    def orthogonalise(self, inversion_symmetric=False):
        """Orthogonalise (measurement.cpp:180-246), or
        OrthogonaliseInversionSymmetric (131-178): (new cell, info)."""
        h = _p()
        info = np.zeros(len(ORTHO_FIELDS))
        f = _LIB.cam_orthogonalise_inversion_symmetric if inversion_symmetric else _LIB.cam_orthogonalise
        _check(f(self.h, C.byref(h), _ptr(info, C.c_double)))
        return UnitCell(self.ctx, h), dict(zip(ORTHO_FIELDS, info.tolist()))

    def energy(self, W: Tensor):
        """iTwoVertexEnergy (measurement.cpp:371-397): (energy per vertex,
        [on-vertex, bond in the cell, bond to the next cell])."""
        e = np.zeros(2)
        t = np.zeros(6)
        _check(_LIB.cam_itwovertex_energy(W.h, self.h, _ptr(e, C.c_double), _ptr(t, C.c_double)))
        return complex(e[0], e[1]), [complex(t[2 * k], t[2 * k + 1]) for k in range(3)]

    def close(self):
        if getattr(self, "h", None):
            _LIB.cam_unitcell_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

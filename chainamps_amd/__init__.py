"""chainamps_amd: MI355X-native ChainAMPS hot path (Python bindings).

The numerical work runs in ``lib/libchainamps_amd.so`` (hand-written HIP
kernels for gfx950 behind the C ABI in ``include/chainamps_amd.h``).  This
package only binds that ABI with ctypes; importing it never falls back to a
CPU implementation -- a missing library is an ImportError-like failure at
first use (``load_library``).
"""
from .capi import (CamError, Context, Tensor, Heff, IDMRG, ITEBD, FTEBD, contract, reorder, svd_truncate,
                   tebd_gate, extract_mpo_block, bond_hamiltonian, bond_evolution_operator,
                   store_unitcell, load_unitcell, UnitCell, load_library, library_path, CAM_FP64, CAM_FP32)

__all__ = ["CamError", "Context", "Tensor", "Heff", "IDMRG", "ITEBD", "FTEBD", "contract", "reorder", "svd_truncate",
           "store_unitcell", "load_unitcell", "UnitCell",
           "tebd_gate", "extract_mpo_block", "bond_hamiltonian", "bond_evolution_operator",
           "load_library", "library_path", "CAM_FP64", "CAM_FP32"]

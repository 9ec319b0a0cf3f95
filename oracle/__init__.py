"""CPU oracle for the chainamps_amd hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, on the CPU, the reference
algorithms of ChainAMPS (andyjaj/chainamps v1.2.2) that the MI355X library
replaces: quantum-number (QN) charge algebra, the mixed-radix MPX_matrix
tensor with its reshape/contract semantics, the two-vertex effective
Hamiltonian matvec in the reference's contraction order, the blockwise SVD
truncation rule, the ARPACK-driven eigensolve and the iDMRG / iTEBD steps
built on them.  Every function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed
CPU baseline).  The product path (``chainamps_amd``) never imports it.

Parity status of each module is recorded in its own header.
"""

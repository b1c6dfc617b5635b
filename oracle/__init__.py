"""CPU oracle for the dense statevector hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker
(or as the timed CPU baseline), never as the thing measured or shipped.

Contents
--------
``sv_oracle``   NumPy restatement of the reference semantics: gate tables,
                the per-gate gather kernel, the eager matrix path, the
                transpiler decompositions, Pauli-string algebra, the energy
                function, and c128 adjoint / parameter-shift / finite-difference
                gradients.  Every function cites the reference file:line it
                follows.
``sv_oracle_c`` ctypes binding of ``c/sv_oracle.c``, a C + OpenMP
                restatement of the gate semantics (op tables, k-qubit
                matrices) for checks at larger sizes and the CPU baseline;
                pinned against ``sv_oracle`` in ``tests/test_oracle_c.py``.

Pinning (what keeps the oracle honest; each check lives in
``tests/test_oracle.py``):
  * gate tables            closed forms (unitarity, known matrices, e.g. sx^2 = x)
  * gather kernel          brute-force Kronecker products for n <= 8 and the
                           eager moveaxis/dot path (two independent formulations)
  * transpiler             literal matrix equality of every decomposition
  * Pauli terms            brute-force Kronecker of 2x2 Pauli matrices
  * energy / gradients     central finite differences, parameter shift and
                           the adjoint sweep must agree (three formulations);
                           the reference's 2-qubit VQE_QASM fixture
  * Trotter generator      scipy.linalg.expm
"""

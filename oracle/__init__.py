"""CPU oracle for the Kore solve path.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker (or as the
timed CPU baseline).  The product path (``kore_amd``) never imports it and
fails loudly when its HIP library is missing.

Modules
-------
layout        l-major block permutation and block-tridiagonal extraction
              (bin/utils.py:26-44, 174-183; bin/assemble.py:606-813).
blocktri      FP64 block-Thomas LU and solve: the exact sparse direct solve
              that MUMPS provides to bin/solve.py:123 and :227.
bcr           FP64 block cyclic reduction, the GPU factorisation's algorithm,
              used to check the device kernels level by level.
krylov_schur  shift-and-invert Krylov-Schur (Stewart 2001), the SLEPc
              default EPS solver that bin/solve.py:91-123 configures.

How each function is pinned is stated in its module header.
"""

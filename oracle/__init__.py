"""CPU oracle for the MPS backend.

TEST INFRASTRUCTURE ONLY.  This package restates, in numpy / torch-CPU, the
algorithm of the reference's custom MPS simulator backend
(test/backend/source/*.cuh, *.cu) and of its reverse-gradient engine
(test/tools/reverse_mps.py).  It exists so that tests, the smoke check and the
CPU baseline leg of the benchmark can check the HIP product path against it.
The product path never imports it.

Modules
  rng        -- SplitMix64 uniform stream and the per-call seed mixers
  mps_ref    -- explicit-SVD MPS: gates, rank rule/guard, prefix-DAG energies,
                perfect sampling, shot estimator, observe/sample semantics
  dense      -- dense FP64 statevector reference (<= 24 qubits)
  reverse_ref-- bond-doubling reverse-mode MPS (torch autograd on the CPU)

Every function cites the reference file:line it follows.  How each function
is pinned (golden vectors, closed forms, dense statevector, brute force) is
listed in tests/test_oracle.py; no function of this oracle is left unpinned.
"""

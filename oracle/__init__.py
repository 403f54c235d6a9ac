"""CPU oracle: a plain-numpy float64 restatement of the reference algorithms on the hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``phast_amd``) may import,
call or execute anything under ``oracle/``.  The only permitted users are ``tests/``,
``__graft_entry__.smoke()`` (as the checker of one small GPU invocation) and the
``cpu_baseline`` leg of ``bench.py`` (as the timed CPU port).  The product path runs on
the HIP library and fails loudly when that library is missing.

Every function cites the reference file:line it restates.  The arithmetic follows the
reference op-for-op (gather -> contraction -> elementwise -> scatter-add) so that it is
also the timed "reference CPU path".

Modules
-------
``oracle.q4``     native-Q4 explicit phase-field path (2D).
``oracle.hex8``   periodic hex8 voxel operator, Jacobi PCG and box-KKT damage (3D).
``oracle.geometry`` voxel microstructure builders (pores, particles).

Parity pinning of each function (closed forms, invariants, brute-force dense
assembly, the reference's own recorded vectors) lives in ``tests/test_oracle_*.py``.
"""

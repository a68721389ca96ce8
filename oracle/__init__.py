"""CPU oracle for the finite-size DMRG hot path (TEST INFRASTRUCTURE ONLY).

This package is a numpy/scipy float64 restatement of the reference's one-site
finite DMRG (``algorithms/DMRG_anyH.py:18-101`` and the methods it calls in
``library/MPSClass.py``, ``library/TensorBasicModule.py``,
``library/HamiltonianModule.py`` and ``library/Parameters.py``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or the timed CPU
baseline).  The product engine under ``t-nalg_amd/`` never imports it and
fails loudly when its HIP library is missing.

Modules
-------
tensor_ops   restatement of the tensor primitives (TBM) used on the path.
model        spin operators and coupling-list generators (HAM, PAR).
mps_ref      the reference's MPS state, per-coupling-term effective-operator
             cache, ``opt_env`` aggregation, matrix-free ``I - tau*Heff`` and
             ``eigsh`` local solve, observables and entanglement (MPS).
dmrg_ref     the sweep driver ``dmrg_finite_size`` (DRV).
block_dmrg   an algebraically identical variant with block-aggregated
             environments (O(1) transfers per step) used as the fast CPU
             checker at larger sizes.
ed           sparse exact diagonalisation, used to pin the above.

How each function is pinned (not by retyping its formula):
* tensor_ops.bound_vec_*, absorb_*, transfer_matrix_mps: the reference's own
  docstring known-answer examples (TBM:497-504,536-543,579-586,644-652,
  690-698,760-770).
* model: closed-form row counts/orders and the coupling conventions fixed by
  exact diagonalisation (a wrong coefficient changes E0).
* mps_ref/dmrg_ref: (i) converged energies, bond energies, Schmidt values and
  entropy of the reference's own saved runs (data_dmrg/*.pr, numbers quoted in
  the tests), (ii) exact diagonalisation E0 for N<=12, (iii) brute-force
  P^T H P effective Hamiltonians for N<=8, (iv) chi >= 2^(N/2) reproduces ED.
* block_dmrg: agrees with mps_ref (same seed) and with ED.
* ed: Lanczos-free dense diagonalisation for N<=8 and the free-fermion closed
  form of the XX chain.
"""

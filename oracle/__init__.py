"""CPU oracle for the closed-shell CCSD hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy FP64, the reference algorithm of CrawfordGroup/pycc for the
spin-adapted RHF-CCSD residual, its Jacobi + DIIS iteration, the HBAR blocks and the Lambda
equations (the backward). Every function cites the reference ``file:line`` it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it, and only as the checker (or the timed CPU baseline). The product (``pycc_amd``) never imports
it: the GPU path fails loudly when its HIP library is missing instead of falling back here.

How each part is pinned (so that a dropped term, a wrong sign or a transposed operand fails):

* ``ccsd_ref``  -- residual, energy, DIIS, solve loop (``ccwfn.py``, ``utils.py``).
  Pinned by ``brute`` (exact similarity transform e^{-T} H e^{T} in the determinant space of
  tiny systems: every r1/r2 element, every term), by ``so_ccsd`` (independent spin-orbital
  Stanton-Gauss CCSD: same energy/amplitudes), by exact two-electron FCI for ``o = 1``, by the
  MP2 closed form at iteration 0, and by the reference's own known answer for H2O/STO-3G
  (``tests/test_002_ccsd_energy.py:30``) through ``gaussian`` (a psi4-free integral + RHF front end).
* ``lambda_ref`` -- HBAR (``cchbar.py``) and Lambda (``cclambda.py``).
  Pinned by finite differences of the pinned forward residual (the Lambda residual is the gradient
  of the CC Lagrangian), by the brute-force left-hand similarity transform, and by the reference's
  H2O/STO-3G pseudoenergy (``tests/test_003_ccsd_lambda.py:50``).
"""

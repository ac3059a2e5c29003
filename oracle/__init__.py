"""CPU oracle for the closed-shell CCSD hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy FP64, the reference algorithm of CrawfordGroup/pycc for the
spin-adapted RHF-CCSD residual, its Jacobi + DIIS iteration, the HBAR blocks, the Lambda
equations (the backward), the densities and the Jacobian-vector product. Every function cites the reference
``file:line`` it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it, and only as the checker (or the timed CPU baseline). The product (``pycc_amd``) never imports
it: the GPU path fails loudly when its HIP library is missing instead of falling back here.

How each part is pinned (``tests/test_oracle_pins.py``), so that a dropped term, a wrong sign or a
transposed operand fails:

* ``ccsd_ref`` -- residual, energy, DIIS, solve loop (``ccwfn.py``, ``utils.py``).
  Pinned by ``brute`` (the exact similarity transform e^{-T} H e^{T} in the determinant space of
  tiny systems: every r1/r2 element at random amplitudes and a non-diagonal Fock; a planted sign
  error is shown to fail it), by exact two-electron FCI for ``o = 1``, by the MP2 closed form at
  iteration 0, by invariants (pair symmetry, vanishing residual at convergence, DIIS coefficients
  summing to 1, orbital-order invariance), and by the reference's own known answer for H2O/STO-3G
  (``tests/test_002_ccsd_energy.py:31``) through ``gaussian`` (a psi4-free integral + RHF front end).
* ``lambda_ref`` -- HBAR (``cchbar.py``) and Lambda (``cclambda.py``).
  Pinned by central differences of the pinned forward residual (the Lambda residual is the gradient
  of the CC Lagrangian at any t and l) and by the reference's H2O/STO-3G pseudoenergy
  (``tests/test_003_ccsd_lambda.py:49``).
* ``density_ref`` -- the one- and two-particle densities (``ccdensity.py:335-931``, the integral
  half of the vJp). Pinned by central differences of the pinned Lagrangian with respect to Fock
  elements and 8-fold symmetric ERI perturbations at arbitrary t and l, by L = <D, F> + <Gamma, ERI>
  (L is linear in H), and by compute_energy = E_CCSD at convergence for a canonical reference.
* ``lambda_ref.ccsd_jacobian`` -- the Jacobian-vector product (``ccderiv.py:149-180``).
  Pinned by central differences of the pinned forward residual along random directions, and by
  the transpose identity <c, J x> = <J^T c, x> against the pinned Lambda (vector-Jacobian) residual.
"""

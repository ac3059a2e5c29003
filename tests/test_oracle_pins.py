"""The CPU oracle, pinned against the reference's own known answers and against exact results.

* H2O/STO-3G frozen-core CCSD energy and Lambda pseudoenergy: pycc's ``test_002_ccsd_energy.py:31``
  and ``test_003_ccsd_lambda.py:49`` (psi4/CFOUR constants, 1e-11), reproduced through the
  psi4-free integral front end in ``oracle.gaussian``.
* H2O/cc-pVDZ (nmo = 24, d shells as solid harmonics), frozen core (``test_002_ccsd_energy.py:38``,
  ``test_003_ccsd_lambda.py:61``) and all electrons (``test_030_sp.py:30,39``): CCSD energy and
  Lambda pseudoenergy at 1e-11. The C two-electron integrals behind it equal the pure-Python ones
  (STO-3G, and random quartets over s, p and d shells).
* every r1/r2 element at arbitrary amplitudes and a non-diagonal Fock: the exact similarity
  transform ``e^{-T} H e^{T}`` in the determinant space (``oracle.brute``).
* o = 1: CCSD equals two-electron FCI.
* iteration 0 is MP2 (``ccwfn.py:260``; ``mpwfn.py:112``).
* the Lambda residual is the gradient of the CC Lagrangian E + l1.r1 + 1/2 l2.r2 (central
  differences of the pinned forward residual), at arbitrary t and l.
"""
import numpy as np
import pytest

from oracle.brute import hbar_projections, two_electron_fci, BruteHbar
from oracle.ccsd_ref import OracleCCSD, OracleDIIS
from oracle.lambda_ref import OracleHBAR, OracleLambda
from pycc_amd.synth import synthetic_df, DenseHamiltonian, dirac_eri
from conftest import random_amplitudes, reference_constant


def _ham(no, nv, seed, t2max=0.2, offdiag=0.05):
    sd = synthetic_df(no, nv, seed=seed, t2max=t2max)
    H = DenseHamiltonian(sd.fock, dirac_eri(sd.B))
    if offdiag:
        X = np.random.default_rng(seed + 100).standard_normal(H.F.shape) * offdiag
        H.F = H.F + X + X.T
    return H


@pytest.fixture(scope="module")
def water():
    from oracle.gaussian import water_sto3g_hamiltonian
    F, ERI, escf, nocc = water_sto3g_hamiltonian()
    return DenseHamiltonian(F, ERI), nocc


def test_water_sto3g_ccsd_matches_reference_constant(water):
    H, nocc = water
    cc = OracleCCSD(H, no=nocc - 1, nfzc=1)          # psi4 freeze_core=true freezes O 1s
    e = cc.solve_cc(1e-12, 1e-12, 75)
    ref, tol = reference_constant("water_sto3g_frozen_core", "ccsd_correlation_energy")
    assert abs(e - ref) < tol     # pycc tests/test_002_ccsd_energy.py:31


def test_water_sto3g_lambda_matches_reference_constant(water):
    H, nocc = water
    cc = OracleCCSD(H, no=nocc - 1, nfzc=1)
    cc.solve_cc(1e-12, 1e-12, 75)
    lam = OracleLambda(cc, OracleHBAR(cc))
    lecc = lam.solve_lambda(1e-12, 1e-12)
    ref, tol = reference_constant("water_sto3g_frozen_core", "lambda_pseudoenergy")
    assert abs(lecc - ref) < tol  # pycc tests/test_003_ccsd_lambda.py:49


@pytest.fixture(scope="module")
def water_dz():
    from oracle.gaussian import water_hamiltonian
    F, ERI, escf, nocc = water_hamiltonian("cc-pvdz")[:4]
    return DenseHamiltonian(F, ERI), nocc


@pytest.mark.parametrize("nfzc,key", [(1, "water_ccpvdz_frozen_core"), (0, "water_ccpvdz_all_electron")])
def test_water_ccpvdz_matches_reference_constants(water_dz, nfzc, key):
    H, nocc = water_dz
    assert H.F.shape == (24, 24)
    cc = OracleCCSD(H, no=nocc - nfzc, nfzc=nfzc)
    e = cc.solve_cc(1e-12, 1e-12, 100)
    ref, tol = reference_constant(key, "ccsd_correlation_energy")
    assert abs(e - ref) < tol, (e, ref)      # test_002_ccsd_energy.py:38 / test_030_sp.py:30
    lecc = OracleLambda(cc, OracleHBAR(cc)).solve_lambda(1e-12, 1e-12)
    ref, tol = reference_constant(key, "lambda_pseudoenergy")
    assert abs(lecc - ref) < tol, (lecc, ref)  # test_003_ccsd_lambda.py:61 / test_030_sp.py:39


def test_c_eri_equals_python_eri():
    """md_eri.c against the pure-Python McMurchie-Davidson primitives: the whole STO-3G ERI tensor, and
    random quartets of a basis with s, p and two-primitive d shells on all three centres."""
    from oracle import gaussian as g
    atoms = g.water_geometry()
    eri_py = g.ao_integrals(atoms)[3]
    eri_c = g.ao_integrals_shells(atoms, g.STO3G, spherical=False)[3]
    assert np.abs(eri_c - eri_py).max() < 1e-14
    basis = {"O": [(2, [1.185, 0.4], [0.7, 0.4]), (1, [1.046], [1.0]), (0, [5.0], [1.0])],
             "H": [(0, [0.4446], [1.0]), (2, [0.727], [1.0])]}
    eri = g.ao_integrals_shells(atoms, basis, spherical=False)[3]
    funcs = []
    for sym, xyz in atoms:
        for l, exps, coefs in basis[sym]:
            e, c = g._shell_coefs(l, exps, coefs)
            funcs += [g._ShellFunction(np.asarray(xyz, float), lmn, e, c) for lmn in g._cart_comps(l)]
    rng = np.random.default_rng(0)
    for i, j, k, l in rng.integers(0, len(funcs), (150, 4)):
        f, h, u, w = funcs[i], funcs[j], funcs[k], funcs[l]
        val = sum(ca * cb * cc * cd * g._eri_prim(a, f.lmn, f.A, b, h.lmn, h.A, c, u.lmn, u.A, d, w.lmn, w.A)
                  for ca, a in zip(f.c, f.exps) for cb, b in zip(h.c, h.exps)
                  for cc, c in zip(u.c, u.exps) for cd, d in zip(w.c, w.exps))
        assert abs(val - eri[i, j, k, l]) < 1e-14, (i, j, k, l)


@pytest.mark.parametrize("no,nv,seed", [(2, 2, 1), (2, 3, 2), (1, 4, 3)])
def test_residual_equals_exact_similarity_transform(no, nv, seed):
    H = _ham(no, nv, seed)
    cc = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, seed, scale=0.1)
    e, r1, r2 = hbar_projections(H.F, H.ERI, no, t1, t2)
    R1, R2 = cc.residuals(H.F, t1, t2)
    assert abs(e - cc.energy(t1, t2)) < 1e-13
    assert np.abs(r1 - R1).max() < 1e-13
    assert np.abs(r2 - R2).max() < 1e-13


def test_brute_force_detects_planted_sign_error():
    """A planted mistake in one term must be caught by the brute-force pin."""
    no, nv = 2, 3
    H = _ham(no, nv, 4)
    cc = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, 4, scale=0.1)
    _, r1, r2 = hbar_projections(H.F, H.ERI, no, t1, t2)
    orig = cc.build_Wmbje
    cc.build_Wmbje = lambda *a: -orig(*a)
    R1, R2 = cc.residuals(H.F, t1, t2)
    assert np.abs(r2 - R2).max() > 1e-4


def test_two_electron_ccsd_is_fci():
    sd = synthetic_df(1, 7, seed=3, t2max=0.1)
    H = DenseHamiltonian(sd.fock, dirac_eri(sd.B))
    cc = OracleCCSD(H, 1)
    e = cc.solve_cc(1e-12, 1e-12)
    assert abs(e - two_electron_fci(H.F, H.ERI)) < 1e-11


def test_iteration_zero_is_mp2():
    no, nv = 3, 7
    H = _ham(no, nv, 9, offdiag=0.0)
    cc = OracleCCSD(H, no)
    o, v = cc.o, cc.v
    t2 = H.ERI[o, o, v, v] / cc.Dijab
    emp2 = np.sum(t2 * (2 * H.ERI[o, o, v, v] - H.ERI[o, o, v, v].swapaxes(2, 3)))
    assert abs(cc.energy() - emp2) < 1e-14


def test_residual_pair_symmetry_and_convergence():
    no, nv = 3, 6
    H = _ham(no, nv, 11, offdiag=0.0, t2max=0.05)
    cc = OracleCCSD(H, no)
    e = cc.solve_cc(1e-11, 1e-11)
    assert e is not None
    r1, r2 = cc.residuals(H.F, cc.t1, cc.t2)
    assert np.abs(r2 - r2.transpose(1, 0, 3, 2)).max() < 1e-14
    assert np.abs(r1).max() < 1e-9 and np.abs(r2).max() < 1e-9


def test_solve_returns_none_without_convergence():
    H = _ham(2, 4, 12, offdiag=0.0)
    cc = OracleCCSD(H, 2)
    assert cc.solve_cc(1e-14, 1e-14, maxiter=2) is None     # ccwfn.py:268-319 falls through


def test_diis_coefficients_sum_to_one():
    rng = np.random.default_rng(0)
    t1, t2 = rng.standard_normal((2, 3)), rng.standard_normal((2, 2, 3, 3))
    d = OracleDIIS(t1, t2, 4)
    for _ in range(6):
        t1 = t1 + 0.1 * rng.standard_normal(t1.shape)
        t2 = t2 + 0.1 * rng.standard_normal(t2.shape)
        d.add_error_vector(t1, t2)
        t1, t2 = d.extrapolate(t1, t2)
        assert abs(d.last_coefficients.sum() - 1.0) < 1e-12
    assert len(d.diis_errors) == 4


def test_energy_invariant_to_orbital_order_within_spaces():
    no, nv = 3, 5
    H = _ham(no, nv, 13, offdiag=0.0, t2max=0.05)
    e0 = OracleCCSD(H, no).solve_cc(1e-11, 1e-11)
    perm = np.concatenate([np.array([2, 0, 1]), no + np.array([4, 1, 3, 0, 2])])
    H2 = DenseHamiltonian(H.F[np.ix_(perm, perm)], H.ERI[np.ix_(perm, perm, perm, perm)])
    e1 = OracleCCSD(H2, no).solve_cc(1e-11, 1e-11)
    assert abs(e0 - e1) < 1e-10


def test_lambda_residual_is_lagrangian_gradient():
    no, nv = 2, 3
    H = _ham(no, nv, 7)
    cc = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, 3, scale=0.1)
    l1, l2 = random_amplitudes(no, nv, 4, scale=0.3)
    lam = OracleLambda(cc, OracleHBAR(cc, t1, t2))
    rl1, rl2 = lam.residuals(H.F, t1, t2, l1, l2)

    def lag(a1, a2):
        r1, r2 = cc.residuals(H.F, a1, a2)
        return cc.energy(a1, a2) + np.sum(l1 * r1) + 0.5 * np.sum(l2 * r2)

    h = 1e-5
    g1 = np.zeros_like(t1)
    for idx in np.ndindex(t1.shape):
        p, m = t1.copy(), t1.copy()
        p[idx] += h; m[idx] -= h
        g1[idx] = (lag(p, t2) - lag(m, t2)) / (2 * h)
    g2 = np.zeros_like(t2)
    for idx in np.ndindex(t2.shape):
        p, m = t2.copy(), t2.copy()
        p[idx] += h; m[idx] -= h
        g2[idx] = (lag(t1, p) - lag(t1, m)) / (2 * h)
    assert np.abs(g1 - rl1).max() < 1e-8
    assert np.abs(g2 + g2.transpose(1, 0, 3, 2) - rl2).max() < 1e-8


def test_jacobian_is_residual_derivative():
    """ccsd_jacobian (ccderiv.py:149-180) = the directional derivative of the pinned residual."""
    from oracle.lambda_ref import ccsd_jacobian
    no, nv = 2, 4
    H = _ham(no, nv, 21)
    cc = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, 5, scale=0.1)
    x1, x2 = random_amplitudes(no, nv, 6, scale=1.0)
    j1, j2 = ccsd_jacobian(cc, OracleHBAR(cc, t1, t2), x1, x2, t2=t2)
    h = 1e-5
    p1, p2 = cc.residuals(H.F, t1 + h * x1, t2 + h * x2)
    m1, m2 = cc.residuals(H.F, t1 - h * x1, t2 - h * x2)
    fd1, fd2 = (p1 - m1) / (2 * h), (p2 - m2) / (2 * h)
    assert np.abs(fd1 - j1).max() < 1e-8 * max(1.0, np.abs(j1).max())
    assert np.abs(fd2 - (j2 + j2.transpose(1, 0, 3, 2))).max() < 1e-8 * max(1.0, np.abs(fd2).max())


def test_jacobian_transpose_is_the_lambda_vjp():
    """<c, J x> = <J^T c, x> with J^T c the source-free Lambda residual at l1 = c1, l2 = c2 + c2^T
    (t2 part halved): the JVP and the vJp are adjoint."""
    from oracle.lambda_ref import ccsd_jacobian
    no, nv = 3, 5
    H = _ham(no, nv, 22)
    cc = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, 7, scale=0.1)
    x1, x2 = random_amplitudes(no, nv, 8, scale=1.0)
    rng = np.random.default_rng(9)
    c1, c2 = rng.standard_normal((no, nv)), rng.standard_normal((no, no, nv, nv))
    h = OracleHBAR(cc, t1, t2)
    j1, j2 = ccsd_jacobian(cc, h, x1, x2, t2=t2)
    lhs = np.sum(c1 * j1) + np.sum(c2 * (j2 + j2.transpose(1, 0, 3, 2)))
    lam = OracleLambda(cc, h)
    r1, r2 = lam.residuals_with(h, t2, c1, c2 + c2.transpose(1, 0, 3, 2))
    g1 = r1 - 2.0 * h.Hov
    g2 = 0.5 * (r2 - 2.0 * H.L[:no, :no, no:, no:])
    rhs = np.sum(g1 * x1) + np.sum(g2 * x2)
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def _lagrangian(H, no, t1, t2, l1, l2):
    cc = OracleCCSD(H, no)
    r1, r2 = cc.residuals(H.F, t1, t2)
    return cc.energy(t1, t2) + np.sum(l1 * r1) + 0.5 * np.sum(l2 * r2)


def test_densities_are_lagrangian_derivatives():
    """ccdensity's spatial builders (ccdensity.py:335-931) are dL/dF and dL/d<pq|rs> of the
    Lagrangian at any t and l (central differences; the ERI perturbation keeps the 8-fold symmetry),
    and L = <D, F> + <Gamma, ERI> (L is linear in H)."""
    from oracle.density_ref import OracleDensity
    no, nv = 2, 4
    H = _ham(no, nv, 23, offdiag=0.05)
    nmo = no + nv
    t1, t2 = random_amplitudes(no, nv, 11, scale=0.1)
    l1, l2 = random_amplitudes(no, nv, 12, scale=0.3)
    D, G = OracleDensity(OracleCCSD(H, no), t1, t2, l1, l2).gradient_densities()
    L0 = _lagrangian(H, no, t1, t2, l1, l2)
    assert abs(L0 - (np.sum(D * H.F) + np.sum(G * H.ERI))) < 1e-12 * max(1.0, abs(L0))
    h = 1e-5
    rng = np.random.default_rng(13)
    for _ in range(12):
        p, q = rng.integers(0, nmo, 2)
        dF = np.zeros_like(H.F)
        dF[p, q] = 1.0
        Hp, Hm = DenseHamiltonian(H.F + h * dF, H.ERI), DenseHamiltonian(H.F - h * dF, H.ERI)
        fd = (_lagrangian(Hp, no, t1, t2, l1, l2) - _lagrangian(Hm, no, t1, t2, l1, l2)) / (2 * h)
        assert abs(fd - D[p, q]) < 1e-8
    for _ in range(24):
        p, q, r, s = rng.integers(0, nmo, 4)
        dE = np.zeros((nmo,) * 4)
        for idx in [(p, q, r, s), (r, q, p, s), (p, s, r, q), (r, s, p, q),
                    (q, p, s, r), (s, p, q, r), (q, r, s, p), (s, r, q, p)]:
            dE[idx] = 1.0
        Hp, Hm = DenseHamiltonian(H.F, H.ERI + h * dE), DenseHamiltonian(H.F, H.ERI - h * dE)
        fd = (_lagrangian(Hp, no, t1, t2, l1, l2) - _lagrangian(Hm, no, t1, t2, l1, l2)) / (2 * h)
        assert abs(fd - np.sum(G * dE)) < 1e-8


def test_density_energy_at_convergence():
    """compute_energy (ccdensity.py:113-211) reproduces E_CCSD for a canonical reference (f_ov = 0)
    once t and l are converged (r = 0, so L = E)."""
    from oracle.density_ref import OracleDensity
    no, nv = 3, 6
    H = _ham(no, nv, 24, offdiag=0.0, t2max=0.05)
    cc = OracleCCSD(H, no)
    e = cc.solve_cc(1e-12, 1e-12)
    lam = OracleLambda(cc, OracleHBAR(cc))
    lam.solve_lambda(1e-12, 1e-12)
    assert abs(OracleDensity(cc, cc.t1, cc.t2, lam.l1, lam.l2).energy() - e) < 1e-10

"""A small psi4-free Gaussian-integral + RHF front end (oracle -- test infrastructure).

Purpose: reproduce the MO Hamiltonian psi4 hands pycc for H2O in STO-3G and cc-pVDZ
(``pycc/tests/conftest.py:28-36`` SCF options; geometry ``pycc/data/molecules.py:42-46``), so the
oracle can be checked against the reference's own known answers (``tests/test_002_ccsd_energy.py:31,38``,
``tests/test_003_ccsd_lambda.py:49,61``, ``tests/test_030_sp.py:30,39``).

Integrals use the McMurchie-Davidson scheme (Hermite expansion coefficients E, Hermite Coulomb
integrals R with the Boys function), written from the published recurrences
(McMurchie & Davidson, J. Comput. Phys. 26, 218 (1978); Helgaker, Jorgensen, Olsen, "Molecular
Electronic-Structure Theory", ch. 9). ``ao_integrals`` / ``water_sto3g_hamiltonian`` take Cartesian
s/p shells in pure Python (all STO-3G needs). ``ao_integrals_shells`` / ``water_hamiltonian`` also take
d shells (cc-pVDZ): the one-electron integrals come from the same Python primitives, the ERIs from
``md_eri.c`` (compiled by ``build_md_eri`` into ``oracle/_build``), and each d shell is reduced to its
five real solid harmonics (psi4 uses spherical functions for the correlation-consistent sets: nmo = 24 for
water). Energies do not depend on how a shell's functions are normalised or ordered, only on the
space they span, so the plain combinations 2z^2 - x^2 - y^2, xz, yz, x^2 - y^2, xy are used.
"""
from __future__ import annotations

import itertools
import math
from functools import lru_cache

import numpy as np
from scipy.special import gammainc, gamma

BOHR_ANGSTROM = 0.52917721067      # CODATA 2014, the value psi4 1.x uses

# STO-3G (Hehre, Stewart, Pople 1969), as distributed in the EMSL/psi4 library.
_STO3G_1S = ([3.42525091, 0.62391373, 0.16885540], [0.15432897, 0.53532814, 0.44463454])
STO3G = {
    "H": [(0, [3.42525091, 0.62391373, 0.16885540], [0.15432897, 0.53532814, 0.44463454])],
    "O": [(0, [130.7093200, 23.8088610, 6.4436083], [0.15432897, 0.53532814, 0.44463454]),
          (0, [5.0331513, 1.1695961, 0.3803890], [-0.09996723, 0.39951283, 0.70011547]),
          (1, [5.0331513, 1.1695961, 0.3803890], [0.15591627, 0.60768372, 0.39195739])],
}
ZNUC = {"H": 1.0, "O": 8.0}

# cc-pVDZ (Dunning, J. Chem. Phys. 90, 1007 (1989)), as distributed in the EMSL/psi4 library
# (segmented form: the two contracted O s functions carry the 0.3023 primitive too).
_O_S = [11720.0, 1759.0, 400.8, 113.7, 37.03, 13.27, 5.025, 1.013, 0.3023]
CCPVDZ = {
    "H": [(0, [13.01, 1.962, 0.4446], [0.019685, 0.137977, 0.478148]),
          (0, [0.1220], [1.0]),
          (1, [0.7270], [1.0])],
    "O": [(0, _O_S, [0.000710, 0.005470, 0.027837, 0.104800, 0.283062, 0.448719, 0.270952, 0.015458,
                     -0.002585]),
          (0, _O_S, [-0.000160, -0.001263, -0.006267, -0.025716, -0.070924, -0.165411, -0.116955,
                     0.557368, 0.572759]),
          (0, [0.3023], [1.0]),
          (1, [17.70, 3.854, 1.046], [0.043018, 0.228913, 0.508728]),
          (1, [0.2753], [1.0]),
          (2, [1.185], [1.0])],
}
BASIS_SETS = {"sto-3g": STO3G, "cc-pvdz": CCPVDZ}


def water_geometry(r_oh_ang=1.1, angle_deg=104.0):
    """Z-matrix ``O / H 1 r / H 1 r 2 angle`` in bohr (orientation does not change energies)."""
    r = r_oh_ang / BOHR_ANGSTROM
    th = math.radians(angle_deg)
    return [("O", np.zeros(3)),
            ("H", np.array([0.0, 0.0, r])),
            ("H", np.array([r * math.sin(th), 0.0, r * math.cos(th)]))]


def _df(n):  # double factorial, (-1)!! = 1
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2)))


class Shell:
    def __init__(self, center, l, exps, coefs):
        self.center = np.asarray(center, float)
        self.l = l
        self.exps = np.asarray(exps, float)
        self.coefs = np.asarray(coefs, float)


class BasisFunction:
    """A normalised contracted Cartesian Gaussian x^i y^j z^k sum_p c_p N_p exp(-a_p r^2)."""

    def __init__(self, center, lmn, exps, coefs):
        self.A = np.asarray(center, float)
        self.lmn = tuple(lmn)
        self.exps = np.asarray(exps, float)
        L = sum(lmn)
        i, j, k = lmn
        prim_norm = ((2.0 * self.exps / math.pi) ** 0.75 * (4.0 * self.exps) ** (L / 2.0)
                     / math.sqrt(_df(2 * i - 1) * _df(2 * j - 1) * _df(2 * k - 1)))
        self.c = np.asarray(coefs, float) * prim_norm
        # contraction normalisation (psi4 normalises contracted functions)
        s = 0.0
        for ca, a in zip(self.c, self.exps):
            for cb, b in zip(self.c, self.exps):
                s += ca * cb * _overlap_prim(a, self.lmn, self.A, b, self.lmn, self.A)
        self.c /= math.sqrt(s)


def build_basis(atoms, basis=STO3G):
    funcs = []
    for sym, xyz in atoms:
        for l, exps, coefs in basis[sym]:
            comps = [(l - a - b, a, b) for a in range(l + 1) for b in range(l + 1 - a)] if l else [(0, 0, 0)]
            if l == 1:
                comps = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
            for lmn in comps:
                funcs.append(BasisFunction(xyz, lmn, exps, coefs))
    return funcs


# ---- McMurchie-Davidson building blocks -------------------------------------------------------
def _E(i, j, t, Qx, a, b):
    """Hermite expansion coefficient E_t^{ij} for 1D Gaussians with exponents a, b and A-B = Qx."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (_E(i - 1, j, t - 1, Qx, a, b) / (2 * p) - q * Qx / a * _E(i - 1, j, t, Qx, a, b)
                + (t + 1) * _E(i - 1, j, t + 1, Qx, a, b))
    return (_E(i, j - 1, t - 1, Qx, a, b) / (2 * p) + q * Qx / b * _E(i, j - 1, t, Qx, a, b)
            + (t + 1) * _E(i, j - 1, t + 1, Qx, a, b))


def _boys(n, T):
    if T < 1e-12:
        return 1.0 / (2 * n + 1) - T / (2 * n + 3)
    return gammainc(n + 0.5, T) * gamma(n + 0.5) / (2.0 * T ** (n + 0.5))


def _R(t, u, v, n, p, PC):
    """Hermite Coulomb integral R^n_{tuv}(p, PC)."""
    X, Y, Z = PC
    if t < 0 or u < 0 or v < 0:
        return 0.0
    if t == u == v == 0:
        return (-2.0 * p) ** n * _boys(n, p * (X * X + Y * Y + Z * Z))
    if t > 0:
        return (t - 1) * _R(t - 2, u, v, n + 1, p, PC) + X * _R(t - 1, u, v, n + 1, p, PC)
    if u > 0:
        return (u - 1) * _R(t, u - 2, v, n + 1, p, PC) + Y * _R(t, u - 1, v, n + 1, p, PC)
    return (v - 1) * _R(t, u, v - 2, n + 1, p, PC) + Z * _R(t, u, v - 1, n + 1, p, PC)


def _overlap_prim(a, lmn1, A, b, lmn2, B):
    s = 1.0
    for d in range(3):
        s *= _E(lmn1[d], lmn2[d], 0, A[d] - B[d], a, b)
    return s * (math.pi / (a + b)) ** 1.5


def _kinetic_prim(a, lmn1, A, b, lmn2, B):
    l, m, n = lmn2
    t0 = b * (2 * (l + m + n) + 3) * _overlap_prim(a, lmn1, A, b, lmn2, B)
    t1 = -2 * b * b * (_overlap_prim(a, lmn1, A, b, (l + 2, m, n), B)
                       + _overlap_prim(a, lmn1, A, b, (l, m + 2, n), B)
                       + _overlap_prim(a, lmn1, A, b, (l, m, n + 2), B))
    t2 = -0.5 * (l * (l - 1) * _overlap_prim(a, lmn1, A, b, (l - 2, m, n), B)
                 + m * (m - 1) * _overlap_prim(a, lmn1, A, b, (l, m - 2, n), B)
                 + n * (n - 1) * _overlap_prim(a, lmn1, A, b, (l, m, n - 2), B))
    return t0 + t1 + t2


def _nuclear_prim(a, lmn1, A, b, lmn2, B, C):
    p = a + b
    P = (a * A + b * B) / p
    PC = P - C
    val = 0.0
    for t in range(lmn1[0] + lmn2[0] + 1):
        Ex = _E(lmn1[0], lmn2[0], t, A[0] - B[0], a, b)
        for u in range(lmn1[1] + lmn2[1] + 1):
            Ey = _E(lmn1[1], lmn2[1], u, A[1] - B[1], a, b)
            for v in range(lmn1[2] + lmn2[2] + 1):
                Ez = _E(lmn1[2], lmn2[2], v, A[2] - B[2], a, b)
                val += Ex * Ey * Ez * _R(t, u, v, 0, p, PC)
    return 2.0 * math.pi / p * val


def _eri_prim(a, l1, A, b, l2, B, c, l3, C, d, l4, D):
    p = a + b
    q = c + d
    alpha = p * q / (p + q)
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    PQ = P - Q
    E1 = [[_E(l1[k], l2[k], t, A[k] - B[k], a, b) for t in range(l1[k] + l2[k] + 1)] for k in range(3)]
    E2 = [[_E(l3[k], l4[k], t, C[k] - D[k], c, d) for t in range(l3[k] + l4[k] + 1)] for k in range(3)]
    val = 0.0
    for t, ex in enumerate(E1[0]):
        for u, ey in enumerate(E1[1]):
            for v, ez in enumerate(E1[2]):
                e1 = ex * ey * ez
                for tau, fx in enumerate(E2[0]):
                    for nu, fy in enumerate(E2[1]):
                        for phi, fz in enumerate(E2[2]):
                            val += (e1 * fx * fy * fz * (-1) ** (tau + nu + phi)
                                    * _R(t + tau, u + nu, v + phi, 0, alpha, PQ))
    return val * 2.0 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))


def _contract2(f, g, prim):
    return sum(ca * cb * prim(a, f.lmn, f.A, b, g.lmn, g.A)
               for ca, a in zip(f.c, f.exps) for cb, b in zip(g.c, g.exps))


def ao_integrals(atoms, basis=STO3G):
    """(S, T, V, ERI_chemist, E_nuc) in the AO basis."""
    funcs = build_basis(atoms, basis)
    n = len(funcs)
    S = np.zeros((n, n)); T = np.zeros((n, n)); V = np.zeros((n, n))
    for i, j in itertools.product(range(n), repeat=2):
        f, g = funcs[i], funcs[j]
        S[i, j] = _contract2(f, g, _overlap_prim)
        T[i, j] = _contract2(f, g, _kinetic_prim)
        V[i, j] = sum(-ZNUC[sym] * _contract2(f, g, lambda a, l1, A, b, l2, B: _nuclear_prim(a, l1, A, b, l2, B, xyz))
                      for sym, xyz in atoms)
    eri = np.zeros((n, n, n, n))
    for i, j, k, l in itertools.product(range(n), repeat=4):
        if j > i or l > k or (i * (i + 1) // 2 + j) < (k * (k + 1) // 2 + l):
            continue
        f, g, h, w = funcs[i], funcs[j], funcs[k], funcs[l]
        val = 0.0
        for ca, a in zip(f.c, f.exps):
            for cb, b in zip(g.c, g.exps):
                for cc, c in zip(h.c, h.exps):
                    for cd, d in zip(w.c, w.exps):
                        val += ca * cb * cc * cd * _eri_prim(a, f.lmn, f.A, b, g.lmn, g.A,
                                                             c, h.lmn, h.A, d, w.lmn, w.A)
        for (p, q, r, s) in [(i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                             (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)]:
            eri[p, q, r, s] = val
    enuc = 0.0
    for (s1, x1), (s2, x2) in itertools.combinations(atoms, 2):
        enuc += ZNUC[s1] * ZNUC[s2] / np.linalg.norm(x1 - x2)
    return S, T, V, eri, enuc


# ---- shells with d functions (ERIs in C) -----------------------------------------------------------
def _cart_comps(l):
    """Cartesian components of a shell, i descending then j descending (the order md_eri.c uses)."""
    return [(i, j, l - i - j) for i in range(l, -1, -1) for j in range(l - i, -1, -1)]


class _ShellFunction:
    """One Cartesian component with the shell's coefficients (radial normalisation of the x^l
    component, contraction normalised for it); used for the one-electron integrals."""

    def __init__(self, A, lmn, exps, c):
        self.A, self.lmn, self.exps, self.c = A, lmn, exps, c


def _shell_coefs(l, exps, coefs):
    exps = np.asarray(exps, float)
    c = np.asarray(coefs, float) * ((2.0 * exps / math.pi) ** 0.75 * (4.0 * exps) ** (l / 2.0)
                                    / math.sqrt(_df(2 * l - 1)))
    lmn = (l, 0, 0)
    s = sum(ca * cb * _overlap_prim(a, lmn, np.zeros(3), b, lmn, np.zeros(3))
            for ca, a in zip(c, exps) for cb, b in zip(c, exps))
    return exps, c / math.sqrt(s)


_LIB = None


def _md_lib():
    """ctypes handle of md_eri.c (compiled into oracle/_build by build_md_eri when missing)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    import ctypes
    import os
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libmderi.so")
    if not os.path.exists(out):
        build_md_eri()
    lib = ctypes.CDLL(out)
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    lib.md_eri.argtypes = [ctypes.c_int, ip, dp, ip, ip, dp, dp, ip, ctypes.c_int, dp]
    lib.md_eri.restype = ctypes.c_int
    _LIB = lib
    return lib


def build_md_eri():
    """Compile md_eri.c with gcc into oracle/_build/libmderi.so."""
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "_build", "libmderi.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.check_call(["gcc", "-O2", "-shared", "-fPIC", "-o", out + ".tmp", os.path.join(here, "md_eri.c"), "-lm"])
    os.replace(out + ".tmp", out)
    return out


def _sph_transform(shell_l):
    """Cartesian -> (spherical for d) basis transformation matrix, columns = final functions."""
    d = np.zeros((6, 5))
    xx, xy, xz, yy, yz, zz = range(6)
    d[[xx, yy, zz], 0] = [-1.0, -1.0, 2.0]      # 2z^2 - x^2 - y^2
    d[xz, 1] = 1.0
    d[yz, 2] = 1.0
    d[[xx, yy], 3] = [1.0, -1.0]                # x^2 - y^2
    d[xy, 4] = 1.0
    blocks = [np.eye(3 if l == 1 else 1) if l < 2 else d for l in shell_l]
    ncart = sum(b.shape[0] for b in blocks)
    nsph = sum(b.shape[1] for b in blocks)
    T = np.zeros((ncart, nsph))
    r = c = 0
    for b in blocks:
        T[r:r + b.shape[0], c:c + b.shape[1]] = b
        r += b.shape[0]
        c += b.shape[1]
    return T


def ao_integrals_shells(atoms, basis=CCPVDZ, spherical=True):
    """(S, T, V, ERI_chemist, E_nuc) for shells up to d; d shells reduced to 5 solid harmonics."""
    import ctypes
    shells = []
    for sym, xyz in atoms:
        for l, exps, coefs in basis[sym]:
            e, c = _shell_coefs(l, exps, coefs)
            shells.append((l, np.asarray(xyz, float), e, c))
    funcs = [_ShellFunction(A, lmn, e, c) for l, A, e, c in shells for lmn in _cart_comps(l)]
    n = len(funcs)
    S = np.zeros((n, n)); T = np.zeros((n, n)); V = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            f, g = funcs[i], funcs[j]
            S[i, j] = S[j, i] = _contract2(f, g, _overlap_prim)
            T[i, j] = T[j, i] = _contract2(f, g, _kinetic_prim)
            V[i, j] = V[j, i] = sum(
                -ZNUC[sym] * _contract2(f, g, lambda a, l1, A, b, l2, B: _nuclear_prim(a, l1, A, b, l2, B, xyz))
                for sym, xyz in atoms)
    ls = np.array([sh[0] for sh in shells], dtype=np.int32)
    xyz = np.ascontiguousarray(np.concatenate([sh[1] for sh in shells]), dtype=np.float64)
    nprim = np.array([len(sh[2]) for sh in shells], dtype=np.int32)
    first = np.concatenate([[0], np.cumsum(nprim)[:-1]]).astype(np.int32)
    exps = np.ascontiguousarray(np.concatenate([sh[2] for sh in shells]), dtype=np.float64)
    coef = np.ascontiguousarray(np.concatenate([sh[3] for sh in shells]), dtype=np.float64)
    ncomp = np.array([(l + 1) * (l + 2) // 2 for l in ls], dtype=np.int32)
    bf0 = np.concatenate([[0], np.cumsum(ncomp)[:-1]]).astype(np.int32)
    eri = np.zeros((n, n, n, n))
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    rc = _md_lib().md_eri(len(shells), ls.ctypes.data_as(ip), xyz.ctypes.data_as(dp), nprim.ctypes.data_as(ip),
                          first.ctypes.data_as(ip), exps.ctypes.data_as(dp), coef.ctypes.data_as(dp),
                          bf0.ctypes.data_as(ip), n, eri.ctypes.data_as(dp))
    if rc != 0:
        raise ValueError("md_eri: angular momentum above 2")
    if spherical:
        X = _sph_transform(ls)
        S, T, V = (X.T @ M @ X for M in (S, T, V))
        eri = np.einsum('pqrs,pi,qj,rk,sl->ijkl', eri, X, X, X, X, optimize=True)
    enuc = 0.0
    for (s1, x1), (s2, x2) in itertools.combinations(atoms, 2):
        enuc += ZNUC[s1] * ZNUC[s2] / np.linalg.norm(x1 - x2)
    return S, T, V, eri, enuc


def water_hamiltonian(basis="cc-pvdz", r_oh_ang=1.1, angle_deg=104.0):
    """(F_mo, ERI_mo Dirac, E_scf, nocc, h_mo, eri_mo chemist, E_nuc) for water; the default geometry
    is pycc's "H2O" molecule (O-H 1.1 A, H-O-H 104 degrees)."""
    atoms = water_geometry(r_oh_ang, angle_deg)
    S, T, V, eri, enuc = ao_integrals_shells(atoms, BASIS_SETS[basis.lower()])
    nocc = 5
    E, C, eps, F_ao = rhf(S, T + V, eri, nocc)
    F_mo = C.T @ F_ao @ C
    h_mo = C.T @ (T + V) @ C
    eri_mo = np.einsum('pqrs,pi,qj,rk,sl->ijkl', eri, C, C, C, C, optimize=True)
    return F_mo, np.ascontiguousarray(eri_mo.transpose(0, 2, 1, 3)), E + enuc, nocc, h_mo, eri_mo, enuc


def rhf(S, Hcore, eri, nocc, conv=1e-13, maxiter=200):
    """Closed-shell RHF with Pulay DIIS; returns (E_elec, C, eps, F_ao)."""
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    e, Cp = np.linalg.eigh(X.T @ Hcore @ X)
    C = X @ Cp
    D = C[:, :nocc] @ C[:, :nocc].T
    focks, errs = [], []
    E_old = 0.0
    for _ in range(maxiter):
        J = np.einsum('pqrs,rs->pq', eri, D)
        K = np.einsum('prqs,rs->pq', eri, D)
        F = Hcore + 2 * J - K
        E = float(np.sum(D * (Hcore + F)))
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        focks.append(F); errs.append(err)
        focks, errs = focks[-8:], errs[-8:]
        n = len(focks)
        B = -np.ones((n + 1, n + 1)); B[-1, -1] = 0
        for a in range(n):
            for b in range(n):
                B[a, b] = np.sum(errs[a] * errs[b])
        rhs = np.zeros(n + 1); rhs[-1] = -1
        cvec = np.linalg.solve(B, rhs)
        Fd = sum(cvec[k] * focks[k] for k in range(n))
        e, Cp = np.linalg.eigh(X.T @ Fd @ X)
        C = X @ Cp
        D_new = C[:, :nocc] @ C[:, :nocc].T
        if abs(E - E_old) < conv and np.abs(D_new - D).max() < conv and np.abs(err).max() < 1e-11:
            D = D_new
            break
        D, E_old = D_new, E
    J = np.einsum('pqrs,rs->pq', eri, D)
    K = np.einsum('prqs,rs->pq', eri, D)
    F = Hcore + 2 * J - K
    e, Cp = np.linalg.eigh(X.T @ F @ X)
    C = X @ Cp
    D = C[:, :nocc] @ C[:, :nocc].T
    F = Hcore + 2 * np.einsum('pqrs,rs->pq', eri, D) - np.einsum('prqs,rs->pq', eri, D)
    E = float(np.sum(D * (Hcore + F)))
    return E, C, e, F


def water_sto3g_hamiltonian():
    """(F_mo, ERI_mo Dirac, E_scf, nocc) for pycc's "H2O" STO-3G molecule."""
    atoms = water_geometry()
    S, T, V, eri, enuc = ao_integrals(atoms)
    nocc = 5
    E, C, eps, F_ao = rhf(S, T + V, eri, nocc)
    F_mo = C.T @ F_ao @ C
    eri_mo = np.einsum('pqrs,pi,qj,rk,sl->ijkl', eri, C, C, C, C, optimize=True)
    return F_mo, np.ascontiguousarray(eri_mo.transpose(0, 2, 1, 3)), E + enuc, nocc

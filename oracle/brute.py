"""Brute-force pins for the CCSD oracle on tiny systems (oracle -- test infrastructure).

Everything here is computed from first principles in the N-electron determinant space, with no
use of the CC working equations:

* :func:`hbar_projections` builds e^{-T} H e^{T} |0> exactly (T is nilpotent, so the exponential
  series terminates) for the closed-shell spin-adapted cluster operator
  ``T = sum_ia t_ia E_ai + 1/2 sum_ijab t_ijab E_ai E_bj`` (``E_pq = sum_s a+_{p s} a_{q s}``) and
  projects it onto ``|Phi_i^a> = a+_{a alpha} a_{i alpha}|0>`` and
  ``|Phi_ij^ab> = a+_{a alpha} a+_{b beta} a_{j beta} a_{i alpha}|0>``.
  These projections are the spin-adapted residuals r1, r2 for any amplitudes, and
  ``<0|e^{-T} H e^{T}|0> - E_ref`` is the CC energy functional.
* :func:`hbar_left_projections` gives the left-hand counterpart used to pin the Lambda residual.
* :func:`two_electron_fci` is the exact singlet ground state of a two-electron system (o = 1),
  which CCSD reproduces exactly.

The one-electron operator is recovered from the Fock matrix as
``h = F - sum_k (2 <pk|qk> - <pk|kq>)`` over the occupied orbitals, so that the normal-ordered
Hamiltonian seen by the CC equations is exactly (F, ERI).
"""
from __future__ import annotations

import itertools

import numpy as np


def core_hamiltonian(F, ERI, no):
    o = slice(0, no)
    return F - (2.0 * np.einsum('pkqk->pq', ERI[:, o, :, o]) - np.einsum('pkkq->pq', ERI[:, o, o, :]))


class DeterminantSpace:
    """All determinants of ``nel`` electrons in ``2 * nmo`` spin orbitals (index 2p = alpha, 2p+1 = beta)."""

    def __init__(self, nmo, nel):
        self.nso = 2 * nmo
        self.dets = [sum(1 << k for k in occ) for occ in itertools.combinations(range(self.nso), nel)]
        self.index = {d: n for n, d in enumerate(self.dets)}

    @staticmethod
    def _annihilate(det, p):
        if not (det >> p) & 1:
            return None, 0
        sign = -1 if bin(det & ((1 << p) - 1)).count('1') % 2 else 1
        return det ^ (1 << p), sign

    @staticmethod
    def _create(det, p):
        if (det >> p) & 1:
            return None, 0
        sign = -1 if bin(det & ((1 << p) - 1)).count('1') % 2 else 1
        return det | (1 << p), sign

    def apply_string(self, ops, det):
        """Apply a product of operators, rightmost first; ops = [(p, +1 create | -1 annihilate), ...]."""
        sign = 1
        for p, kind in reversed(ops):
            det, s = (self._create if kind > 0 else self._annihilate)(det, p)
            if det is None:
                return None, 0
            sign *= s
        return det, sign

    def operator_matrix(self, terms):
        """Matrix of sum_k c_k * op-string_k in this space; terms = [(coef, ops), ...]."""
        n = len(self.dets)
        M = np.zeros((n, n))
        for col, det in enumerate(self.dets):
            for coef, ops in terms:
                if coef == 0.0:
                    continue
                out, s = self.apply_string(ops, det)
                if out is not None:
                    M[self.index[out], col] += coef * s
        return M


def _hamiltonian_terms(h, ERI):
    nmo = h.shape[0]
    terms = []
    for p, q in itertools.product(range(nmo), repeat=2):
        for s in (0, 1):
            terms.append((h[p, q], [(2 * p + s, 1), (2 * q + s, -1)]))
    for p, q, r, t in itertools.product(range(nmo), repeat=4):
        v = ERI[p, q, r, t]
        if v == 0.0:
            continue
        for s1 in (0, 1):
            for s2 in (0, 1):
                # 1/2 <pq|rt> a+_p a+_q a_t a_r with spins (p,r)=s1, (q,t)=s2
                terms.append((0.5 * v, [(2 * p + s1, 1), (2 * q + s2, 1), (2 * t + s2, -1), (2 * r + s1, -1)]))
    return terms


def _cluster_terms(t1, t2, no):
    nv = t1.shape[1]
    terms = []
    for i, a in itertools.product(range(no), range(nv)):
        for s in (0, 1):
            terms.append((t1[i, a], [(2 * (no + a) + s, 1), (2 * i + s, -1)]))
    for i, j, a, b in itertools.product(range(no), range(no), range(nv), range(nv)):
        for s1 in (0, 1):
            for s2 in (0, 1):
                # 1/2 t_ijab E_ai E_bj = 1/2 t a+_{a s1} a_{i s1} a+_{b s2} a_{j s2}
                terms.append((0.5 * t2[i, j, a, b], [(2 * (no + a) + s1, 1), (2 * i + s1, -1),
                                                     (2 * (no + b) + s2, 1), (2 * j + s2, -1)]))
    return terms


def _exp_apply(T, vec, sign=1.0):
    out = vec.copy()
    term = vec.copy()
    for k in range(1, 64):
        term = sign * (T @ term) / k
        if not np.any(term):
            break
        out += term
    return out


class BruteHbar:
    """Dense H, T and the reference determinant for one small closed-shell system."""

    def __init__(self, F, ERI, no):
        nmo = F.shape[0]
        self.no, self.nv, self.nmo = no, nmo - no, nmo
        self.space = DeterminantSpace(nmo, 2 * no)
        self.h = core_hamiltonian(F, ERI, no)
        self.E_ref = float(sum(self.h[i, i] + F[i, i] for i in range(no)))
        self.H = self.space.operator_matrix(_hamiltonian_terms(self.h, ERI))
        self.ref = self.space.index[sum(1 << k for k in range(2 * no))]

    def _det(self, ops):
        d, s = self.space.apply_string(ops, self.space.dets[self.ref])
        return self.space.index[d], s

    def single(self, i, a):
        return self._det([(2 * (self.no + a), 1), (2 * i, -1)])

    def double(self, i, j, a, b):
        return self._det([(2 * (self.no + a), 1), (2 * (self.no + b) + 1, 1), (2 * j + 1, -1), (2 * i, -1)])

    def cluster(self, t1, t2):
        return self.space.operator_matrix(_cluster_terms(t1, t2, self.no))

    def hbar_ket(self, t1, t2):
        """e^{-T} H e^{T} |0> as a vector."""
        T = self.cluster(t1, t2)
        e0 = np.zeros(len(self.space.dets)); e0[self.ref] = 1.0
        return _exp_apply(T, self.H @ _exp_apply(T, e0, 1.0), -1.0)


def hbar_projections(F, ERI, no, t1, t2, brute=None):
    """(E_cc, r1, r2) from the exact similarity transform; see the module docstring."""
    bh = brute or BruteHbar(F, ERI, no)
    psi = bh.hbar_ket(t1, t2)
    nv = bh.nv
    r1 = np.zeros((no, nv))
    r2 = np.zeros((no, no, nv, nv))
    for i, a in itertools.product(range(no), range(nv)):
        k, s = bh.single(i, a)
        r1[i, a] = s * psi[k]
    for i, j, a, b in itertools.product(range(no), range(no), range(nv), range(nv)):
        k, s = bh.double(i, j, a, b)
        r2[i, j, a, b] = s * psi[k]
    return float(psi[bh.ref]) - bh.E_ref, r1, r2


def two_electron_fci(F, ERI):
    """Exact correlation energy of a two-electron singlet (o = 1): E_FCI - E_ref."""
    nmo = F.shape[0]
    h = core_hamiltonian(F, ERI, 1)
    # symmetric spatial functions C_pq = C_qp; (H C)_pq = h C + C h^T + sum_rs <pq|rs> C_rs
    pairs = [(p, q) for p in range(nmo) for q in range(p, nmo)]
    n = len(pairs)
    Hm = np.zeros((n, n))
    basis = []
    for p, q in pairs:
        C = np.zeros((nmo, nmo))
        if p == q:
            C[p, p] = 1.0
        else:
            C[p, q] = C[q, p] = 1.0 / np.sqrt(2.0)
        basis.append(C)
    for k, Ck in enumerate(basis):
        HC = h @ Ck + Ck @ h.T + np.einsum('pqrs,rs->pq', ERI, Ck)
        for l, Cl in enumerate(basis):
            Hm[l, k] = np.sum(Cl * HC)
    E_ref = 2.0 * h[0, 0] + ERI[0, 0, 0, 0]
    return float(np.linalg.eigvalsh(Hm)[0] - E_ref)

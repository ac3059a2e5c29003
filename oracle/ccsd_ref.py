"""Numpy FP64 restatement of pycc's spin-adapted RHF-CCSD solver (oracle -- test infrastructure).

Follows, term by term and in the same order:

* ``pycc/ccwfn.py:195-211``   denominators and MP1 guess
* ``pycc/ccwfn.py:216-319``   ``solve_cc`` (Jacobi update, 2-norm "rms", convergence test, DIIS
  placement after the test, return ``None`` when ``maxiter`` is exhausted)
* ``pycc/ccwfn.py:321-372``   ``residuals``
* ``pycc/ccwfn.py:432-761``   ``build_tau``, ``build_Fae/Fmi/Fme/Wmnij/Wmbej/Wmbje/Zmbij``, ``r_T1``
* ``pycc/ccwfn.py:764-791``   ``r_T2`` with the final P(ij,ab) symmetrisation
* ``pycc/ccwfn.py:886-944``   ``_r_T2_ccsd``
* ``pycc/ccwfn.py:1122-1162`` ``cc_energy``
* ``pycc/utils.py:257-361``   ``helper_diis``

Input is a pycc-layout Hamiltonian: ``H.F`` (nmo, nmo), ``H.ERI[p,q,r,s] = <pq|rs>`` and
``H.L = 2 ERI - ERI.swapaxes(2, 3)``; a frozen core is a slice offset (``wavefunction.py:304-312``).
Contractions use ``numpy.einsum(optimize=True)`` in place of ``opt_einsum.contract``
(``pycc/device.py:64-68``), which is not installed; both dispatch pairwise products to BLAS.
"""
from __future__ import annotations

import numpy as np


def contract(subscripts, *operands):
    return np.einsum(subscripts, *operands, optimize=True)


class OracleCCSD:
    """Spatial-orbital closed-shell CCSD, restated from ``pycc.ccwfn.CCwfn`` (``ccwfn.py:32``)."""

    def __init__(self, H, no: int, nfzc: int = 0):
        # Frozen core as a slice offset, as in wavefunction.py:304-312.
        self.H = H
        nmo = H.F.shape[0]
        self.nfzc = nfzc
        self.no = no
        self.nv = nmo - no - nfzc
        self.nmo = nmo
        self.o = slice(nfzc, nfzc + no)
        self.v = slice(nfzc + no, nmo)
        o, v = self.o, self.v
        # ccwfn.py:195-198
        eps_o, eps_v = H.eps[o], H.eps[v]
        self.Dijab = (eps_o.reshape(-1, 1, 1, 1) + eps_o.reshape(-1, 1, 1)
                      - eps_v.reshape(-1, 1) - eps_v)
        self.Dia = eps_o.reshape(-1, 1) - eps_v
        # ccwfn.py:210-211
        self.t1 = np.zeros((self.no, self.nv))
        self.t2 = H.ERI[o, o, v, v].copy() / self.Dijab
        self.history = []          # (niter, ecc, ediff, rms) rows as printed by solve_cc
        self.niter = 0

    # ---- intermediates -------------------------------------------------------------------
    def build_tau(self, t1, t2, fact1=1.0, fact2=1.0):
        # ccwfn.py:455
        return fact1 * t2 + fact2 * contract('ia,jb->ijab', t1, t1)

    def build_Fae(self, o, v, F, L, t1, t2):
        # ccwfn.py:494-497
        Fae = F[v, v].copy()
        Fae = Fae - 0.5 * contract('me,ma->ae', F[o, v], t1)
        Fae = Fae + contract('mf,mafe->ae', t1, L[o, v, v, v])
        Fae = Fae - contract('mnaf,mnef->ae', self.build_tau(t1, t2, 1.0, 0.5), L[o, o, v, v])
        return Fae

    def build_Fmi(self, o, v, F, L, t1, t2):
        # ccwfn.py:530-533
        Fmi = F[o, o].copy()
        Fmi = Fmi + 0.5 * contract('ie,me->mi', t1, F[o, v])
        Fmi = Fmi + contract('ne,mnie->mi', t1, L[o, o, o, v])
        Fmi = Fmi + contract('inef,mnef->mi', self.build_tau(t1, t2, 1.0, 0.5), L[o, o, v, v])
        return Fmi

    def build_Fme(self, o, v, F, L, t1):
        # ccwfn.py:563-564
        Fme = F[o, v].copy()
        Fme = Fme + contract('nf,mnef->me', t1, L[o, o, v, v])
        return Fme

    def build_Wmnij(self, o, v, ERI, t1, t2):
        # ccwfn.py:596-603
        Wmnij = ERI[o, o, o, o].copy()
        Wmnij = Wmnij + contract('je,mnie->mnij', t1, ERI[o, o, o, v])
        Wmnij = Wmnij + contract('ie,mnej->mnij', t1, ERI[o, o, v, o])
        Wmnij = Wmnij + contract('ijef,mnef->mnij', self.build_tau(t1, t2), ERI[o, o, v, v])
        return Wmnij

    def build_Wmbej(self, o, v, ERI, L, t1, t2):
        # ccwfn.py:641-645
        Wmbej = ERI[o, v, v, o].copy()
        Wmbej = Wmbej + contract('jf,mbef->mbej', t1, ERI[o, v, v, v])
        Wmbej = Wmbej - contract('nb,mnej->mbej', t1, ERI[o, o, v, o])
        Wmbej = Wmbej - contract('jnfb,mnef->mbej', self.build_tau(t1, t2, 0.5, 1.0), ERI[o, o, v, v])
        Wmbej = Wmbej + 0.5 * contract('njfb,mnef->mbej', t2, L[o, o, v, v])
        return Wmbej

    def build_Wmbje(self, o, v, ERI, t1, t2):
        # ccwfn.py:680-683
        Wmbje = -1.0 * ERI[o, v, o, v].copy()
        Wmbje = Wmbje - contract('jf,mbfe->mbje', t1, ERI[o, v, v, v])
        Wmbje = Wmbje + contract('nb,mnje->mbje', t1, ERI[o, o, o, v])
        Wmbje = Wmbje + contract('jnfb,mnfe->mbje', self.build_tau(t1, t2, 0.5, 1.0), ERI[o, o, v, v])
        return Wmbje

    def build_Zmbij(self, o, v, ERI, t1, t2):
        # ccwfn.py:715
        return contract('mbef,ijef->mbij', ERI[o, v, v, v], self.build_tau(t1, t2))

    # ---- residuals -----------------------------------------------------------------------
    def r_T1(self, o, v, F, ERI, L, t1, t2, Fae, Fme, Fmi):
        # ccwfn.py:754-760
        r_T1 = F[v, o].swapaxes(0, 1).copy()
        r_T1 = r_T1 + contract('ie,ae->ia', t1, Fae)
        r_T1 = r_T1 - contract('mi,ma->ia', Fmi, t1)
        r_T1 = r_T1 + contract('imae,me->ia', (2.0 * t2 - t2.swapaxes(2, 3)), Fme)
        r_T1 = r_T1 + contract('nf,nafi->ia', t1, L[o, v, v, o])
        r_T1 = r_T1 + contract('mief,maef->ia', (2.0 * t2 - t2.swapaxes(2, 3)), ERI[o, v, v, v])
        r_T1 = r_T1 - contract('mnae,nmei->ia', t2, L[o, o, v, o])
        return r_T1

    def r_T2_half(self, o, v, F, ERI, t1, t2, Fae, Fme, Fmi, Wmnij, Wmbej, Wmbje, Zmbij):
        """``_r_T2_ccsd`` (ccwfn.py:922-940), before the P(ij,ab) symmetrisation."""
        r_T2 = 0.5 * ERI[v, v, o, o].swapaxes(0, 2).swapaxes(1, 3).copy()
        r_T2 = r_T2 + contract('ijae,eb->ijab', t2, Fae.T)
        tmp = contract('bm,me->be', t1.T, Fme)
        r_T2 = r_T2 - 0.5 * contract('ijae,eb->ijab', t2, tmp.T)
        r_T2 = r_T2 - contract('imab,mj->ijab', t2, Fmi)
        tmp = contract('je,em->jm', t1, Fme.T)
        r_T2 = r_T2 - 0.5 * contract('imab,mj->ijab', t2, tmp.T)
        tau = self.build_tau(t1, t2)
        r_T2 = r_T2 + 0.5 * contract('mnab,mnij->ijab', tau, Wmnij)
        r_T2 = r_T2 + 0.5 * contract('ijef,abef->ijab', tau, ERI[v, v, v, v])
        r_T2 = r_T2 - contract('ma,mbij->ijab', t1, Zmbij)
        r_T2 = r_T2 + contract('imae,mbej->ijab', (t2 - t2.swapaxes(2, 3)), Wmbej)
        r_T2 = r_T2 + contract('imae,mbej->ijab', t2, (Wmbej + Wmbje.swapaxes(2, 3)))
        r_T2 = r_T2 + contract('mjae,mbie->ijab', t2, Wmbje)
        tmp = contract('ei,am->aemi', t1.T, t1.T)
        r_T2 = r_T2 - contract('imea,mbej->ijab', tmp.swapaxes(0, 3).swapaxes(1, 2), ERI[o, v, v, o])
        r_T2 = r_T2 - contract('imeb,maje->ijab', tmp.swapaxes(0, 3).swapaxes(1, 2), ERI[o, v, o, v])
        r_T2 = r_T2 + contract('ie,abej->ijab', t1, ERI[v, v, v, o])
        r_T2 = r_T2 - contract('ma,mbij->ijab', t1, ERI[o, v, o, o])
        return r_T2

    def r_T2(self, *args):
        # ccwfn.py:788-790
        r_T2 = self.r_T2_half(*args)
        return r_T2 + r_T2.swapaxes(0, 1).swapaxes(2, 3)

    def residuals(self, F, t1, t2):
        # ccwfn.py:341-372
        o, v = self.o, self.v
        ERI, L = self.H.ERI, self.H.L
        Fae = self.build_Fae(o, v, F, L, t1, t2)
        Fmi = self.build_Fmi(o, v, F, L, t1, t2)
        Fme = self.build_Fme(o, v, F, L, t1)
        Wmnij = self.build_Wmnij(o, v, ERI, t1, t2)
        Wmbej = self.build_Wmbej(o, v, ERI, L, t1, t2)
        Wmbje = self.build_Wmbje(o, v, ERI, t1, t2)
        Zmbij = self.build_Zmbij(o, v, ERI, t1, t2)
        r1 = self.r_T1(o, v, F, ERI, L, t1, t2, Fae, Fme, Fmi)
        r2 = self.r_T2(o, v, F, ERI, t1, t2, Fae, Fme, Fmi, Wmnij, Wmbej, Wmbje, Zmbij)
        return r1, r2

    def cc_energy(self, o, v, F, L, t1, t2):
        # ccwfn.py:1160-1161
        ecc = 2.0 * contract('ia,ia->', F[o, v], t1)
        ecc = ecc + contract('ijab,ijab->', self.build_tau(t1, t2), L[o, o, v, v])
        return float(ecc)

    def energy(self, t1=None, t2=None):
        t1 = self.t1 if t1 is None else t1
        t2 = self.t2 if t2 is None else t2
        return self.cc_energy(self.o, self.v, self.H.F, self.H.L, t1, t2)

    # ---- solver --------------------------------------------------------------------------
    def solve_cc(self, e_conv=1e-7, r_conv=1e-7, maxiter=100, max_diis=8, start_diis=1, verbose=False):
        """ccwfn.py:216-319 (spatial, non-local branch)."""
        o, v = self.o, self.v
        F, L = self.H.F, self.H.L
        Dia, Dijab = self.Dia, self.Dijab
        ecc = self.cc_energy(o, v, F, L, self.t1, self.t2)          # ccwfn.py:260
        self.history = [(0, ecc, -ecc, None)]
        diis = OracleDIIS(self.t1, self.t2, max_diis)                # ccwfn.py:266
        for niter in range(1, maxiter + 1):                          # ccwfn.py:268
            ecc_last = ecc
            r1, r2 = self.residuals(F, self.t1, self.t2)             # ccwfn.py:272
            self.t1 = self.t1 + r1 / Dia                             # ccwfn.py:281-284
            self.t2 = self.t2 + r2 / Dijab
            rms = contract('ia,ia->', r1 / Dia, r1 / Dia) + contract('ijab,ijab->', r2 / Dijab, r2 / Dijab)
            rms = float(np.sqrt(rms))
            ecc = self.cc_energy(o, v, F, L, self.t1, self.t2)       # ccwfn.py:286
            ediff = ecc - ecc_last
            self.history.append((niter, ecc, ediff, rms))
            self.niter = niter
            if verbose:
                print("Iter %3d: CC Ecorr = %.15f  dE = % .5E  rms = % .5E" % (niter, ecc, ediff, rms))
            if (abs(ediff) < e_conv) and abs(rms) < r_conv:          # ccwfn.py:295
                self.ecc = ecc
                return ecc
            diis.add_error_vector(self.t1, self.t2)                  # ccwfn.py:317-319
            if niter >= start_diis:
                self.t1, self.t2 = diis.extrapolate(self.t1, self.t2)
        return None


class OracleDIIS:
    """``pycc.utils.helper_diis`` (utils.py:257-361), restated."""

    def __init__(self, t1, t2, max_diis):
        # utils.py:272-282
        self.oldt1 = t1.copy()
        self.oldt2 = t2.copy()
        self.diis_vals_t1 = [t1.copy()]
        self.diis_vals_t2 = [t2.copy()]
        self.diis_errors = []
        self.diis_size = 0
        self.max_diis = max_diis
        self.last_coefficients = None

    def add_error_vector(self, t1, t2):
        # utils.py:284-295
        self.diis_vals_t1.append(t1.copy())
        self.diis_vals_t2.append(t2.copy())
        error_t1 = (self.diis_vals_t1[-1] - self.oldt1).ravel()
        error_t2 = (self.diis_vals_t2[-1] - self.oldt2).ravel()
        self.diis_errors.append(np.concatenate((error_t1, error_t2)))
        self.oldt1 = t1.copy()
        self.oldt2 = t2.copy()

    def extrapolate(self, t1, t2):
        # utils.py:313-361
        if self.max_diis == 0:
            return t1, t2
        if len(self.diis_errors) > self.max_diis:
            del self.diis_vals_t1[0]
            del self.diis_vals_t2[0]
            del self.diis_errors[0]
        self.diis_size = len(self.diis_errors)
        B = np.zeros((self.diis_size + 1, self.diis_size + 1)) - 1.0
        B[-1, -1] = 0
        for n1, e1 in enumerate(self.diis_errors):
            B[n1, n1] = np.dot(e1, e1)
            for n2, e2 in enumerate(self.diis_errors):
                if n1 >= n2:
                    continue
                B[n1, n2] = np.dot(e1, e2)
                B[n2, n1] = B[n1, n2]
        B[:-1, :-1] /= np.abs(B[:-1, :-1]).max()
        resid = np.zeros(self.diis_size + 1)
        resid[-1] = -1
        ci = np.linalg.solve(B, resid)
        self.last_coefficients = ci[:-1].copy()
        t1 = np.zeros_like(self.oldt1)
        t2 = np.zeros_like(self.oldt2)
        for num in range(self.diis_size):
            t1 += ci[num] * self.diis_vals_t1[num + 1]
            t2 += ci[num] * self.diis_vals_t2[num + 1]
        self.oldt1 = t1.copy()
        self.oldt2 = t2.copy()
        return t1, t2

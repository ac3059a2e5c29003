"""Numpy FP64 restatement of pycc's spatial CCSD HBAR and Lambda equations (oracle -- test infrastructure).

Follows, term by term:

* ``pycc/cchbar.py:86-96``  block order; CCSD branches of ``build_Hov`` (:152-153), ``build_Hvv``
  (:206-209), ``build_Hoo`` (:269-272), ``build_Hoooo`` (:331-337), ``build_Hvvvv`` (:395-401),
  ``build_Hvovv`` (:455-456), ``build_Hooov`` (:505-506), ``build_Hovvo`` (:563-568),
  ``build_Hovov`` (:625-629), ``build_Hvvvo`` (:682-695), ``build_Hovoo`` (:806-819)
* ``pycc/cclambda.py:60-67``   guess ``l1 = 2 t1``, ``l2 = 2 (2 t2 - t2^T)``
* ``pycc/cclambda.py:69-200``  ``solve_lambda`` (Jacobi update, DIIS after the convergence test)
* ``pycc/cclambda.py:202-256`` ``residuals`` (HBAR rebuilt from arbitrary t1, t2)
* ``pycc/cclambda.py:258-306`` ``build_Goo``, ``build_Gvv``
* ``pycc/cclambda.py:345-369`` ``r_L1``; ``:468-496`` ``r_L2``; ``:570`` ``pseudoenergy``
* ``pycc/ccderiv.py:149-180`` ``_ccsd_jacobian`` (the Jacobian-vector product on HBAR)
"""
from __future__ import annotations

import numpy as np

from .ccsd_ref import OracleCCSD, OracleDIIS, contract


class OracleHBAR:
    """``pycc.cchbar.cchbar`` spatial CCSD blocks (cchbar.py:54-98)."""

    def __init__(self, cc: OracleCCSD, t1=None, t2=None, F=None):
        t1 = cc.t1 if t1 is None else t1
        t2 = cc.t2 if t2 is None else t2
        F = cc.H.F if F is None else F
        ERI, L = cc.H.ERI, cc.H.L
        o, v = cc.o, cc.v
        self.cc = cc
        self.Hov = self.build_Hov(o, v, F, L, t1)
        self.Hvv = self.build_Hvv(o, v, F, L, t1, t2)
        self.Hoo = self.build_Hoo(o, v, F, L, t1, t2)
        self.Hoooo = self.build_Hoooo(o, v, ERI, t1, t2)
        self.Hvvvv = self.build_Hvvvv(o, v, ERI, t1, t2)
        self.Hvovv = self.build_Hvovv(o, v, ERI, t1)
        self.Hooov = self.build_Hooov(o, v, ERI, t1)
        self.Hovvo = self.build_Hovvo(o, v, ERI, L, t1, t2)
        self.Hovov = self.build_Hovov(o, v, ERI, t1, t2)
        self.Hvvvo = self.build_Hvvvo(o, v, ERI, L, self.Hov, self.Hvvvv, t1, t2, F)
        self.Hovoo = self.build_Hovoo(o, v, ERI, L, self.Hov, self.Hoooo, t1, t2)

    def tau(self, t1, t2):
        return self.cc.build_tau(t1, t2)

    def build_Hov(self, o, v, F, L, t1):
        Hov = F[o, v].copy()                                                      # cchbar.py:152
        Hov = Hov + contract('nf,mnef->me', t1, L[o, o, v, v])                    # :153
        return Hov

    def build_Hvv(self, o, v, F, L, t1, t2):
        Hvv = F[v, v].copy()                                                      # :206
        Hvv = Hvv - contract('me,ma->ae', F[o, v], t1)                            # :207
        Hvv = Hvv + contract('mf,amef->ae', t1, L[v, o, v, v])                    # :208
        Hvv = Hvv - contract('mnfa,mnfe->ae', self.tau(t1, t2), L[o, o, v, v])    # :209
        return Hvv

    def build_Hoo(self, o, v, F, L, t1, t2):
        Hoo = F[o, o].copy()                                                      # :269
        Hoo = Hoo + contract('ie,me->mi', t1, F[o, v])                            # :270
        Hoo = Hoo + contract('ne,mnie->mi', t1, L[o, o, o, v])                    # :271
        Hoo = Hoo + contract('inef,mnef->mi', self.tau(t1, t2), L[o, o, v, v])    # :272
        return Hoo

    def build_Hoooo(self, o, v, ERI, t1, t2):
        Hoooo = ERI[o, o, o, o].copy()                                            # :331
        tmp = contract('je,mnie->mnij', t1, ERI[o, o, o, v])                      # :332
        Hoooo = Hoooo + (tmp + tmp.swapaxes(0, 1).swapaxes(2, 3))                 # :333
        Hoooo = Hoooo + contract('ijef,mnef->mnij', self.tau(t1, t2), ERI[o, o, v, v])  # :337
        return Hoooo

    def build_Hvvvv(self, o, v, ERI, t1, t2):
        Hvvvv = ERI[v, v, v, v].copy()                                            # :395
        tmp = contract('mb,amef->abef', t1, ERI[v, o, v, v])                      # :396
        Hvvvv = Hvvvv - (tmp + tmp.swapaxes(0, 1).swapaxes(2, 3))                 # :397
        Hvvvv = Hvvvv + contract('mnab,mnef->abef', self.tau(t1, t2), ERI[o, o, v, v])  # :401
        return Hvvvv

    def build_Hvovv(self, o, v, ERI, t1):
        Hvovv = ERI[v, o, v, v].copy()                                            # :455
        Hvovv = Hvovv - contract('na,nmef->amef', t1, ERI[o, o, v, v])            # :456
        return Hvovv

    def build_Hooov(self, o, v, ERI, t1):
        Hooov = ERI[o, o, o, v].copy()                                            # :505
        Hooov = Hooov + contract('if,nmef->mnie', t1, ERI[o, o, v, v])            # :506
        return Hooov

    def build_Hovvo(self, o, v, ERI, L, t1, t2):
        Hovvo = ERI[o, v, v, o].copy()                                            # :563
        Hovvo = Hovvo + contract('jf,mbef->mbej', t1, ERI[o, v, v, v])            # :564
        Hovvo = Hovvo - contract('nb,mnej->mbej', t1, ERI[o, o, v, o])            # :565
        Hovvo = Hovvo - contract('jnfb,mnef->mbej', self.tau(t1, t2), ERI[o, o, v, v])  # :567
        Hovvo = Hovvo + contract('njfb,mnef->mbej', t2, L[o, o, v, v])            # :568
        return Hovvo

    def build_Hovov(self, o, v, ERI, t1, t2):
        Hovov = ERI[o, v, o, v].copy()                                            # :625
        Hovov = Hovov + contract('jf,bmef->mbje', t1, ERI[v, o, v, v])            # :626
        Hovov = Hovov - contract('nb,mnje->mbje', t1, ERI[o, o, o, v])            # :627
        Hovov = Hovov - contract('jnfb,nmef->mbje', self.tau(t1, t2), ERI[o, o, v, v])  # :629
        return Hovov

    def build_Hvvvo(self, o, v, ERI, L, Hov, Hvvvv, t1, t2, F):
        Hvvvo = ERI[v, v, v, o].copy()                                            # :682
        Hvvvo = Hvvvo - contract('me,miab->abei', Hov, t2)                        # :683
        Hvvvo = Hvvvo + contract('if,abef->abei', t1, Hvvvv)                      # :684
        Hvvvo = Hvvvo + contract('mnab,mnei->abei', self.tau(t1, t2), ERI[o, o, v, o])  # :685
        Hvvvo = Hvvvo - contract('imfa,bmfe->abei', t2, ERI[v, o, v, v])          # :686
        Hvvvo = Hvvvo - contract('imfb,amef->abei', t2, ERI[v, o, v, v])          # :687
        Hvvvo = Hvvvo + contract('mifb,amef->abei', t2, L[v, o, v, v])            # :688
        tmp = ERI[v, o, v, o].copy()                                              # :689
        tmp = tmp - contract('infa,mnfe->amei', t2, ERI[o, o, v, v])              # :690
        Hvvvo = Hvvvo - contract('mb,amei->abei', t1, tmp)                        # :691
        tmp = ERI[v, o, o, v].copy()                                              # :692
        tmp = tmp - contract('infb,mnef->bmie', t2, ERI[o, o, v, v])              # :693
        tmp = tmp + contract('nifb,mnef->bmie', t2, L[o, o, v, v])                # :694
        Hvvvo = Hvvvo - contract('ma,bmie->abei', t1, tmp)                        # :695
        return Hvvvo

    def build_Hovoo(self, o, v, ERI, L, Hov, Hoooo, t1, t2):
        Hovoo = ERI[o, v, o, o].copy()                                            # :806
        Hovoo = Hovoo + contract('me,ijeb->mbij', Hov, t2)                        # :807
        Hovoo = Hovoo - contract('nb,mnij->mbij', t1, Hoooo)                      # :808
        Hovoo = Hovoo + contract('ijef,mbef->mbij', self.tau(t1, t2), ERI[o, v, v, v])  # :809
        Hovoo = Hovoo - contract('ineb,nmje->mbij', t2, ERI[o, o, o, v])          # :810
        Hovoo = Hovoo - contract('jneb,mnie->mbij', t2, ERI[o, o, o, v])          # :811
        Hovoo = Hovoo + contract('njeb,mnie->mbij', t2, L[o, o, o, v])            # :812
        tmp = ERI[o, v, o, v].copy()                                              # :813
        tmp = tmp - contract('infb,mnfe->mbie', t2, ERI[o, o, v, v])              # :814
        Hovoo = Hovoo + contract('je,mbie->mbij', t1, tmp)                        # :815
        tmp = ERI[v, o, o, v].copy()                                              # :816
        tmp = tmp - contract('jnfb,mnef->bmje', t2, ERI[o, o, v, v])              # :817
        tmp = tmp + contract('njfb,mnef->bmje', t2, L[o, o, v, v])                # :818
        Hovoo = Hovoo + contract('ie,bmje->mbij', t1, tmp)                        # :819
        return Hovoo


class OracleLambda:
    """``pycc.cclambda.cclambda`` spatial CCSD branch (cclambda.py:42-570)."""

    def __init__(self, cc: OracleCCSD, hbar: OracleHBAR):
        self.cc = cc
        self.hbar = hbar
        self.l1 = 2.0 * cc.t1                                       # cclambda.py:66
        self.l2 = 2.0 * (2.0 * cc.t2 - cc.t2.swapaxes(2, 3))        # cclambda.py:67
        self.history = []

    @staticmethod
    def build_Goo(t2, l2):
        return contract('mjab,ijab->mi', t2, l2)                    # cclambda.py:281

    @staticmethod
    def build_Gvv(t2, l2):
        return -1.0 * contract('ijeb,ijab->ae', t2, l2)             # cclambda.py:306

    @staticmethod
    def r_L1(l1, l2, h, Gvv, Goo):
        r_l1 = 2.0 * h.Hov.copy()                                                     # :348
        r_l1 = r_l1 + contract('ie,ea->ia', l1, h.Hvv)                                # :354
        r_l1 = r_l1 - contract('ma,im->ia', l1, h.Hoo)                                # :355
        r_l1 = r_l1 + contract('imef,efam->ia', l2, h.Hvvvo)                          # :356
        r_l1 = r_l1 - contract('mnae,iemn->ia', l2, h.Hovoo)                          # :357
        r_l1 = r_l1 + contract('me,ieam->ia', l1, (2.0 * h.Hovvo - h.Hovov.swapaxes(2, 3)))  # :358
        r_l1 = r_l1 - 2.0 * contract('ef,eifa->ia', Gvv, h.Hvovv)                     # :366
        r_l1 = r_l1 + contract('ef,eiaf->ia', Gvv, h.Hvovv)                           # :367
        r_l1 = r_l1 - 2.0 * contract('mn,mina->ia', Goo, h.Hooov)                     # :368
        r_l1 = r_l1 + contract('mn,imna->ia', Goo, h.Hooov)                           # :369
        return r_l1

    @staticmethod
    def r_L2(o, v, l1, l2, L, h, Gvv, Goo):
        r_l2 = L[o, o, v, v].copy()                                                   # :468
        r_l2 = r_l2 + 2.0 * contract('ia,jb->ijab', l1, h.Hov)                        # :476
        r_l2 = r_l2 - contract('ja,ib->ijab', l1, h.Hov)                              # :477
        r_l2 = r_l2 + 2.0 * contract('ie,ejab->ijab', l1, h.Hvovv)                    # :478
        r_l2 = r_l2 - contract('ie,ejba->ijab', l1, h.Hvovv)                          # :479
        r_l2 = r_l2 - 2.0 * contract('mb,jima->ijab', l1, h.Hooov)                    # :480
        r_l2 = r_l2 + contract('mb,ijma->ijab', l1, h.Hooov)                          # :481
        r_l2 = r_l2 + contract('ijeb,ea->ijab', l2, h.Hvv)                            # :486
        r_l2 = r_l2 - contract('mjab,im->ijab', l2, h.Hoo)                            # :487
        r_l2 = r_l2 + 0.5 * contract('mnab,ijmn->ijab', l2, h.Hoooo)                  # :488
        r_l2 = r_l2 + 0.5 * contract('ijef,efab->ijab', l2, h.Hvvvv)                  # :489
        r_l2 = r_l2 + contract('mjeb,ieam->ijab', l2, (2.0 * h.Hovvo - h.Hovov.swapaxes(2, 3)))  # :490
        r_l2 = r_l2 - contract('mibe,jema->ijab', l2, h.Hovov)                        # :491
        r_l2 = r_l2 - contract('mieb,jeam->ijab', l2, h.Hovvo)                        # :492
        r_l2 = r_l2 + contract('ae,ijeb->ijab', Gvv, L[o, o, v, v])                   # :493
        r_l2 = r_l2 - contract('mi,mjab->ijab', Goo, L[o, o, v, v])                   # :494
        r_l2 = r_l2 + r_l2.swapaxes(0, 1).swapaxes(2, 3)                              # :496
        return r_l2

    def residuals_with(self, hbar, t2, l1, l2):
        Goo = self.build_Goo(t2, l2)
        Gvv = self.build_Gvv(t2, l2)
        r1 = self.r_L1(l1, l2, hbar, Gvv, Goo)
        r2 = self.r_L2(self.cc.o, self.cc.v, l1, l2, self.cc.H.L, hbar, Gvv, Goo)
        return r1, r2

    def residuals(self, F, t1, t2, l1, l2):
        """cclambda.py:202-256: rebuild HBAR from (F, t1, t2), then r_L1/r_L2."""
        hbar = OracleHBAR(self.cc, t1, t2, F)
        return self.residuals_with(hbar, t2, l1, l2)

    def pseudoenergy(self, l2):
        o, v = self.cc.o, self.cc.v
        return float(0.5 * contract('ijab,ijab->', self.cc.H.ERI[o, o, v, v], l2))   # :570

    def solve_lambda(self, e_conv=1e-7, r_conv=1e-7, maxiter=100, max_diis=8, start_diis=1):
        """cclambda.py:69-200 (spatial CCSD branch)."""
        cc = self.cc
        Dia, Dijab = cc.Dia, cc.Dijab
        lecc = self.pseudoenergy(self.l2)
        self.history = [(0, lecc, -lecc, None)]
        diis = OracleDIIS(self.l1, self.l2, max_diis)
        for niter in range(1, maxiter + 1):
            lecc_last = lecc
            r1, r2 = self.residuals_with(self.hbar, cc.t2, self.l1, self.l2)
            self.l1 = self.l1 + r1 / Dia
            self.l2 = self.l2 + r2 / Dijab
            rms = contract('ia,ia->', r1 / Dia, r1 / Dia)
            rms += contract('ijab,ijab->', r2 / Dijab, r2 / Dijab)
            rms = float(np.sqrt(rms))
            lecc = self.pseudoenergy(self.l2)
            ediff = lecc - lecc_last
            self.history.append((niter, lecc, ediff, rms))
            if (abs(ediff) < e_conv) and abs(rms) < r_conv:
                self.lecc = lecc
                return lecc
            diis.add_error_vector(self.l1, self.l2)
            if niter >= start_diis:
                self.l1, self.l2 = diis.extrapolate(self.l1, self.l2)
        return None


def ccsd_jacobian(cc: OracleCCSD, h: OracleHBAR, X1, X2, t2=None):
    """``CCderiv._ccsd_jacobian`` (pycc/ccderiv.py:149-180): (HBAR . X) on singles and (un-symmetrised)
    doubles. The derivative of the residual along X is (j1, j2 + j2^T)."""
    o, v = cc.o, cc.v
    t2 = cc.t2 if t2 is None else t2
    L = cc.H.L
    Hvvmvo = 2.0 * h.Hvovv - h.Hvovv.swapaxes(2, 3)                                   # :163
    Hooomv = 2.0 * h.Hooov - h.Hooov.swapaxes(0, 1)                                   # :164
    r1 = contract('ie,ae->ia', X1, h.Hvv) - contract('ma,mi->ia', X1, h.Hoo)          # :165
    r1 = r1 + 2.0 * contract('me,maei->ia', X1, h.Hovvo) - contract('me,maie->ia', X1, h.Hovov)   # :166
    r1 = r1 + contract('me,miea->ia', h.Hov, (2.0 * X2 - X2.swapaxes(0, 1)))         # :167
    r1 = r1 + contract('imef,amef->ia', X2, Hvvmvo)                                   # :168
    r1 = r1 - contract('mnae,mnie->ia', X2, Hooomv)                                   # :169
    Zvv = contract('amef,mf->ae', Hvvmvo, X1) - contract('mnef,mnaf->ae', L[o, o, v, v], X2)     # :170
    Zoo = -1.0 * contract('mnie,ne->mi', Hooomv, X1) - contract('mnef,inef->mi', L[o, o, v, v], X2)  # :171
    r2 = contract('ie,abej->ijab', X1, h.Hvvvo) - contract('ma,mbij->ijab', X1, h.Hovoo)          # :172
    r2 = r2 + contract('mi,mjab->ijab', Zoo, t2) + contract('ae,ijeb->ijab', Zvv, t2)             # :173
    r2 = r2 + contract('ijeb,ae->ijab', X2, h.Hvv) - contract('mjab,mi->ijab', X2, h.Hoo)         # :174
    r2 = r2 + 0.5 * contract('mnab,mnij->ijab', X2, h.Hoooo) + 0.5 * contract('ijef,abef->ijab', X2, h.Hvvvv)  # :175
    r2 = r2 - contract('imeb,maje->ijab', X2, h.Hovov) - contract('imea,mbej->ijab', X2, h.Hovvo)  # :176
    r2 = r2 + 2.0 * contract('miea,mbej->ijab', X2, h.Hovvo) - contract('miea,mbje->ijab', X2, h.Hovov)  # :177
    return r1, r2


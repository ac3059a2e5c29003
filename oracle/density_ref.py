"""Numpy FP64 restatement of pycc's spatial CCSD one- and two-particle densities (oracle -- test
infrastructure).

Follows the CCSD branches of ``pycc/ccdensity.py``:

* ``build_Doo`` :335, ``build_Dvv`` :373, ``build_Dvo`` :410, ``build_Dov`` :415 (one-particle)
* ``build_Doooo`` :563, ``build_Dvvvv`` :587, ``build_Dooov`` :611, ``build_Dvvvo`` :682,
  ``build_Dovov`` :753, ``build_Doovv`` :784 (two-particle)
* ``compute_energy`` :113 (E = F.D + the six two-particle blocks) and ``gradient_densities`` :267
  (full D and the four-fold symmetrised Gamma with contract(D, F) + contract(Gamma, ERI) = E)

These are the derivatives of the CC Lagrangian E + l1.r1 + 1/2 l2.r2 with respect to the Fock
matrix and the two-electron integrals at fixed t and l.
"""
from __future__ import annotations

import numpy as np

from .ccsd_ref import OracleCCSD, contract


class OracleDensity:
    def __init__(self, cc: OracleCCSD, t1, t2, l1, l2):
        self.cc = cc
        tau = cc.build_tau(t1, t2)
        self.Doo = self.build_Doo(t1, t2, l1, l2)
        self.Dvv = self.build_Dvv(t1, t2, l1, l2)
        self.Dvo = l1.T.copy()                                                         # :413
        self.Dov = self.build_Dov(t1, t2, l1, l2, tau)
        self.Doooo = contract('ijef,klef->ijkl', tau, l2)                              # :585
        self.Dvvvv = contract('mnab,mncd->abcd', tau, l2)                              # :609
        self.Dooov = self.build_Dooov(t1, t2, l1, l2, tau)
        self.Dvvvo = self.build_Dvvvo(t1, t2, l1, l2, tau)
        self.Dovov = self.build_Dovov(t1, t2, l1, l2, tau)
        self.Doovv = self.build_Doovv(t1, t2, l1, l2, tau)

    @staticmethod
    def Goo(t2, l2):
        return contract('mjab,ijab->mi', t2, l2)                                       # cclambda.py:281

    @staticmethod
    def Gvv(t2, l2):
        return -1.0 * contract('ijeb,ijab->ae', t2, l2)                                # cclambda.py:306

    @staticmethod
    def build_Doo(t1, t2, l1, l2):
        return -contract('ie,je->ij', t1, l1) - contract('imef,jmef->ij', t2, l2)      # :364

    @staticmethod
    def build_Dvv(t1, t2, l1, l2):
        return contract('mb,ma->ab', t1, l1) + contract('mnbe,mnae->ab', t2, l2)       # :401

    @staticmethod
    def build_Dov(t1, t2, l1, l2, tau):
        Dov = 2.0 * t1.copy()                                                          # :454
        Dov = Dov + 2.0 * contract('me,imae->ia', l1, t2)                              # :456
        Dov = Dov - contract('me,miae->ia', l1, tau)                                   # :457
        tmp = contract('mnef,inef->mi', l2, t2)                                        # :458
        Dov = Dov - contract('mi,ma->ia', tmp, t1)                                     # :459
        tmp = contract('mnef,mnaf->ea', l2, t2)                                        # :460
        Dov = Dov - contract('ea,ie->ia', tmp, t1)                                     # :461
        return Dov

    def build_Dooov(self, t1, t2, l1, l2, tau):
        tmp = 2.0 * tau - tau.swapaxes(2, 3)                                           # :645
        D = -1.0 * contract('ke,ijea->ijka', l1, tmp)                                  # :646
        D = D - contract('ie,jkae->ijka', t1, l2)                                      # :647
        Goo = self.Goo(t2, l2)                                                         # :651
        D = D - 2.0 * contract('ik,ja->ijka', Goo, t1)                                 # :652
        D = D + contract('jk,ia->ijka', Goo, t1)                                       # :653
        tmp = contract('jmaf,kmef->jake', t2, l2)                                      # :654
        D = D - 2.0 * contract('jake,ie->ijka', tmp, t1)                               # :655
        D = D + contract('iake,je->ijka', tmp, t1)                                     # :656
        tmp = contract('ijef,kmef->ijkm', t2, l2)                                      # :658
        D = D + contract('ijkm,ma->ijka', tmp, t1)                                     # :659
        tmp = contract('mjaf,kmef->jake', t2, l2)                                      # :660
        D = D + contract('jake,ie->ijka', tmp, t1)                                     # :661
        tmp = contract('imea,kmef->iakf', t2, l2)                                      # :662
        D = D + contract('iakf,jf->ijka', tmp, t1)                                     # :663
        tmp = contract('kmef,jf->kmej', l2, t1)                                        # :668
        tmp = contract('kmej,ie->kmij', tmp, t1)                                       # :669
        D = D + contract('kmij,ma->ijka', tmp, t1)                                     # :670
        return D

    def build_Dvvvo(self, t1, t2, l1, l2, tau):
        tmp = 2.0 * tau - tau.swapaxes(2, 3)                                           # :716
        D = contract('mc,miab->abci', l1, tmp)                                         # :717
        D = D + contract('ma,imbc->abci', t1, l2)                                      # :718
        Gvv = self.Gvv(t2, l2)                                                         # :722
        D = D - 2.0 * contract('ca,ib->abci', Gvv, t1)                                 # :723
        D = D + contract('cb,ia->abci', Gvv, t1)                                       # :724
        tmp = contract('imbe,nmce->ibnc', t2, l2)                                      # :725
        D = D + 2.0 * contract('ibnc,na->abci', tmp, t1)                               # :726
        D = D - contract('ianc,nb->abci', tmp, t1)                                     # :727
        tmp = contract('nmab,nmce->abce', t2, l2)                                      # :729
        D = D - contract('abce,ie->abci', tmp, t1)                                     # :730
        tmp = contract('niae,nmce->iamc', t2, l2)                                      # :731
        D = D - contract('iamc,mb->abci', tmp, t1)                                     # :732
        tmp = contract('mibe,nmce->ibnc', t2, l2)                                      # :733
        D = D - contract('ibnc,na->abci', tmp, t1)                                     # :734
        tmp = contract('nmce,ie->nmci', l2, t1)                                        # :739
        tmp = contract('nmci,na->amci', tmp, t1)                                       # :740
        D = D - contract('amci,mb->abci', tmp, t1)                                     # :741
        return D

    @staticmethod
    def build_Dovov(t1, t2, l1, l2, tau):
        D = -1.0 * contract('ia,jb->iajb', t1, l1)                                     # :775
        D = D - contract('mibe,jmea->iajb', tau, l2)                                   # :779
        D = D - contract('imbe,mjea->iajb', t2, l2)                                    # :780
        return D

    def build_Doovv(self, t1, t2, l1, l2, tau):
        tau_spinad = 2.0 * tau - tau.swapaxes(2, 3)                                    # :830
        D = 4.0 * contract('ia,jb->ijab', t1, l1)                                      # :860
        D = D + 2.0 * tau_spinad                                                       # :861
        D = D + l2                                                                     # :862
        tmp1 = 2.0 * t2 - t2.swapaxes(2, 3)                                            # :864
        tmp2 = 2.0 * contract('me,jmbe->jb', l1, tmp1)                                 # :865
        D = D + 2.0 * contract('jb,ia->ijab', tmp2, t1)                                # :866
        D = D - contract('ja,ib->ijab', tmp2, t1)                                      # :867
        tmp2 = 2.0 * contract('ijeb,me->ijmb', tmp1, l1)                               # :868
        D = D - contract('ijmb,ma->ijab', tmp2, t1)                                    # :869
        tmp2 = 2.0 * contract('jmba,me->jeba', tau_spinad, l1)                         # :870
        D = D - contract('jeba,ie->ijab', tmp2, t1)                                    # :871
        D = D + 4.0 * contract('imae,mjeb->ijab', t2, l2)                              # :876
        D = D - 2.0 * contract('mjbe,imae->ijab', tau, l2)                             # :877
        tmp_oooo = contract('ijef,mnef->ijmn', t2, l2)                                 # :879
        D = D + contract('ijmn,mnab->ijab', tmp_oooo, t2)                              # :880
        tmp1 = contract('njbf,mnef->jbme', t2, l2)                                     # :881
        D = D + contract('jbme,miae->ijab', tmp1, t2)                                  # :882
        tmp1 = contract('imfb,mnef->ibne', t2, l2)                                     # :883
        D = D + contract('ibne,njae->ijab', tmp1, t2)                                  # :884
        Gvv = self.Gvv(t2, l2)                                                         # :885
        D = D + 4.0 * contract('eb,ijae->ijab', Gvv, tau)                              # :886
        D = D - 2.0 * contract('ea,ijbe->ijab', Gvv, tau)                              # :887
        Goo = self.Goo(t2, l2)                                                         # :888
        D = D - 4.0 * contract('jm,imab->ijab', Goo, tau)                              # :889
        D = D + 2.0 * contract('jm,imba->ijab', Goo, tau)                              # :890
        tmp1 = contract('inaf,mnef->iame', t2, l2)                                     # :891
        D = D - 4.0 * contract('iame,mjbe->ijab', tmp1, tau)                           # :892
        D = D + 2.0 * contract('ibme,mjae->ijab', tmp1, tau)                           # :893
        D = D + 4.0 * contract('jbme,imae->ijab', tmp1, t2)                            # :894
        D = D - 2.0 * contract('jame,imbe->ijab', tmp1, t2)                            # :895
        tmp = contract('nb,ijmn->ijmb', t1, tmp_oooo)                                  # :898
        D = D + contract('ma,ijmb->ijab', t1, tmp)                                     # :899
        tmp = contract('ie,mnef->mnif', t1, l2)                                        # :900
        tmp = contract('jf,mnif->mnij', t1, tmp)                                       # :901
        D = D + contract('mnij,mnab->ijab', tmp, t2)                                   # :902
        tmp = contract('ie,mnef->mnif', t1, l2)                                        # :903
        tmp = contract('mnif,njbf->mijb', tmp, t2)                                     # :904
        D = D + contract('ma,mijb->ijab', t1, tmp)                                     # :905
        tmp = contract('jf,mnef->mnej', t1, l2)                                        # :906
        tmp = contract('mnej,miae->njia', tmp, t2)                                     # :907
        D = D + contract('nb,njia->ijab', t1, tmp)                                     # :908
        tmp = contract('je,mnef->mnjf', t1, l2)                                        # :909
        tmp = contract('mnjf,imfb->njib', tmp, t2)                                     # :910
        D = D + contract('na,njib->ijab', t1, tmp)                                     # :911
        tmp = contract('if,mnef->mnei', t1, l2)                                        # :912
        tmp = contract('mnei,njae->mija', tmp, t2)                                     # :913
        D = D + contract('mb,mija->ijab', t1, tmp)                                     # :914
        tmp = contract('jf,mnef->mnej', t1, l2)                                        # :919
        tmp = contract('ie,mnej->mnij', t1, tmp)                                       # :920
        tmp = contract('nb,mnij->mbij', t1, tmp)                                       # :921
        D = D + contract('ma,mbij->ijab', t1, tmp)                                     # :922
        return D

    def energy(self, F=None, ERI=None):
        """compute_energy (:158-193): F_ij D_ij + F_ab D_ab + the six two-particle blocks."""
        cc = self.cc
        o, v = cc.o, cc.v
        F = cc.H.F if F is None else F
        ERI = cc.H.ERI if ERI is None else ERI
        e1 = contract('ij,ij->', F[o, o], self.Doo) + contract('ab,ab->', F[v, v], self.Dvv)
        e2 = (0.5 * contract('ijkl,ijkl->', ERI[o, o, o, o], self.Doooo)
              + 0.5 * contract('abcd,abcd->', ERI[v, v, v, v], self.Dvvvv)
              + contract('ijka,ijka->', ERI[o, o, o, v], self.Dooov)
              + contract('abci,abci->', ERI[v, v, v, o], self.Dvvvo)
              + contract('iajb,iajb->', ERI[o, v, o, v], self.Dovov)
              + 0.5 * contract('ijab,ijab->', ERI[o, o, v, v], self.Doovv))
        return float(e1 + e2)

    def gradient_densities(self):
        """gradient_densities (:315-334): full-MO D and the four-fold symmetrised Gamma."""
        cc = self.cc
        o, v = cc.o, cc.v
        nmo = cc.H.F.shape[0]
        D = np.zeros((nmo, nmo))
        D[o, o] = self.Doo; D[v, v] = self.Dvv; D[o, v] = self.Dov; D[v, o] = self.Dvo
        G = np.zeros((nmo,) * 4)
        G[o, o, o, o] = 0.5 * self.Doooo
        G[v, v, v, v] = 0.5 * self.Dvvvv
        G[o, v, o, v] = self.Dovov
        G[o, o, v, v] = 0.25 * self.Doovv
        G[v, v, o, o] = 0.25 * self.Doovv.transpose(2, 3, 0, 1)
        G[o, o, o, v] = 0.5 * self.Dooov
        G[o, v, o, o] = 0.5 * self.Dooov.transpose(2, 3, 0, 1)
        G[v, v, v, o] = 0.5 * self.Dvvvo
        G[v, o, v, v] = 0.5 * self.Dvvvo.transpose(2, 3, 0, 1)
        G = 0.25 * (G + G.transpose(1, 0, 3, 2) + G.transpose(2, 3, 0, 1) + G.transpose(3, 2, 1, 0))
        return D, G

"""CC one- and two-particle densities on the GPU (counterpart of ``pycc.ccdensity.ccdensity``,
spatial CCSD).

``ccdensity(ccwfn, cclambda)`` builds the density blocks of ``pycc/ccdensity.py:58-111`` from the
current amplitudes (``Doo``, ``Dvv``, ``Dov``, ``Dvo``, ``Doooo``, ``Dooov``, ``Dvvvo``, ``Dovov``,
``Doovv``; ``Dvvvv`` only with ``store_vvvv=True``, it is v^4). They are the derivatives of the CC
Lagrangian E + l1.r1 + 1/2 l2.r2 with respect to the Fock matrix and the two-electron integrals at
fixed amplitudes. ``compute_energy`` follows ``ccdensity.py:113-211``; ``gradient_densities``
follows ``ccdensity.py:267-334`` (full nmo^2 D and nmo^4 Gamma -- for small systems).
"""
from __future__ import annotations

import time

from . import _lib
from .device import dptr
from .exceptions import PyCCError


def _blocks(cc, t1, t2, l1, l2, big=True, vvvv=False):
    torch = cc._torch
    o, v, nmo = cc.no, cc.nv, cc.nmo
    f64, dev = torch.float64, cc.device1
    out = {"D": torch.empty((nmo, nmo), dtype=f64, device=dev),
           "Doooo": torch.empty((o, o, o, o), dtype=f64, device=dev),
           "Dooov": torch.empty((o, o, o, v), dtype=f64, device=dev),
           "Dovov": torch.empty((o, v, o, v), dtype=f64, device=dev),
           "Doovv": torch.empty((o, o, v, v), dtype=f64, device=dev),
           "Dvvvo": torch.empty((v, v, v, o), dtype=f64, device=dev) if big else None,
           "Dvvvv": torch.empty((v, v, v, v), dtype=f64, device=dev) if vvvv else None}
    ptr = lambda x: dptr(x) if x is not None else None
    _lib.call("ccx_densities", cc._ctx.h, dptr(t1), dptr(t2), dptr(l1), dptr(l2), dptr(out["D"]),
              dptr(out["Doooo"]), dptr(out["Dooov"]), dptr(out["Dovov"]), dptr(out["Doovv"]),
              ptr(out["Dvvvo"]), ptr(out["Dvvvv"]), cc._stream())
    return out


class ccdensity(object):
    """Densities of a converged (or any) CCSD / Lambda pair (ccdensity.py:58-111)."""

    def __init__(self, ccwfn, cclambda, onlyone=False, store_vvvv=False):
        time_init = time.time()
        ccwfn._require_dp("ccdensity")
        self.ccwfn = ccwfn
        self.cclambda = cclambda
        self.contract = ccwfn.contract
        self.onlyone = onlyone
        o, v = ccwfn.o, ccwfn.v
        self._t = (ccwfn.t1.clone(), ccwfn.t2.clone(), cclambda.l1.clone(), cclambda.l2.clone())
        b = _blocks(ccwfn, *self._t, big=True, vvvv=store_vvvv)
        D = b["D"]
        self.Doo, self.Dvv = D[o, o].contiguous(), D[v, v].contiguous()
        self.Dov, self.Dvo = D[o, v].contiguous(), D[v, o].contiguous()
        self._D = D
        if not onlyone:
            self.Doooo, self.Dooov, self.Dovov = b["Doooo"], b["Dooov"], b["Dovov"]
            self.Doovv, self.Dvvvo, self.Dvvvv = b["Doovv"], b["Dvvvo"], b["Dvvvv"]
        if ccwfn.verbose:
            print("CC density constructed in %.3f seconds." % (time.time() - time_init))

    def compute_energy(self):
        """ccdensity.py:113-211 (two-particle part included unless onlyone)."""
        cc = self.ccwfn
        if self.onlyone:
            raise PyCCError("compute_energy needs the two-particle density (onlyone=False)")
        t1, t2, l1, l2 = self._t
        _lib.call("ccx_density_energy", cc._ctx.h, dptr(t1), dptr(t2), dptr(l2), dptr(cc.H.F), dptr(self._D),
                  dptr(self.Doooo), dptr(self.Dooov), dptr(self.Dovov), dptr(self.Doovv), dptr(self.Dvvvo),
                  dptr(cc._e_dev[:1]), cc._stream())
        self.ecc = float(cc._e_dev[0].item())
        return self.ecc

    def gradient_densities(self):
        """ccdensity.py:267-334: full-MO D (nmo x nmo) and the four-fold symmetrised Gamma (nmo^4,
        Dirac order) with contract(D, F) + contract(Gamma, ERI) = the Lagrangian. Needs
        ``store_vvvv=True`` (Gamma holds the v^4 block)."""
        if self.onlyone or self.Dvvvv is None:
            raise PyCCError("gradient_densities needs the full two-particle density (store_vvvv=True)")
        torch = self.ccwfn._torch
        o, v, nmo = self.ccwfn.o, self.ccwfn.v, self.ccwfn.nmo
        G = torch.zeros((nmo,) * 4, dtype=torch.float64, device=self._D.device)
        G[o, o, o, o] = 0.5 * self.Doooo
        G[v, v, v, v] = 0.5 * self.Dvvvv
        G[o, v, o, v] = self.Dovov
        G[o, o, v, v] = 0.25 * self.Doovv
        G[v, v, o, o] = 0.25 * self.Doovv.permute(2, 3, 0, 1)
        G[o, o, o, v] = 0.5 * self.Dooov
        G[o, v, o, o] = 0.5 * self.Dooov.permute(2, 3, 0, 1)
        G[v, v, v, o] = 0.5 * self.Dvvvo
        G[v, o, v, v] = 0.5 * self.Dvvvo.permute(2, 3, 0, 1)
        G = 0.25 * (G + G.permute(1, 0, 3, 2) + G.permute(2, 3, 0, 1) + G.permute(3, 2, 1, 0))
        return self._D.clone(), G

"""Lambda equations on the GPU (counterpart of ``pycc.cclambda.cclambda``, spatial CCSD).

``cclambda(ccwfn, hbar).solve_lambda(e_conv, r_conv, maxiter, max_diis, start_diis)`` follows
``pycc/cclambda.py:69-200``: guess ``l1 = 2 t1``, ``l2 = 2 (2 t2 - t2^T)`` (:66-67), Jacobi update
``l += r/D``, the 2-norm "rms", the pseudoenergy ``1/2 <ij|ab> l2_ijab`` (:570) as the energy,
convergence when both ``|dE| < e_conv`` and ``rms < r_conv``, DIIS after the convergence test, and
``None`` when ``maxiter`` is exhausted. ``residuals(F, t1, t2, l1, l2)`` rebuilds HBAR from the given
amplitudes, as ``cclambda.residuals`` (:202-256) does.

The contractions run in the library (``ccx_lambda_residual``); amplitudes, denominators, DIIS and
norms are FP64 device tensors.
"""
from __future__ import annotations

import ctypes
import math
import time

from . import _lib
from .ccwfn import _iteration, _title
from .device import dptr


class cclambda(object):
    """Lambda amplitudes of a ``CCwfn`` on its HBAR (cclambda.py:42-200)."""

    def __init__(self, ccwfn, hbar):
        ccwfn._require_dp("cclambda", mixed_ok=True)
        self.ccwfn = ccwfn
        self.hbar = hbar
        self.contract = ccwfn.contract
        self.l1 = (2.0 * ccwfn.t1).contiguous()                                      # :66
        self.l2 = (2.0 * (2.0 * ccwfn.t2 - ccwfn.t2.swapaxes(2, 3))).contiguous()     # :67
        self.history = []

    def build_Goo(self, t2, l2):
        """G_mi = t_mjab l_ijab (cclambda.py:281)."""
        return self.contract('mjab,ijab->mi', t2, l2)

    def build_Gvv(self, t2, l2):
        """G_ae = -t_ijeb l_ijab (cclambda.py:306)."""
        return self.contract('ijeb,ijab->ae', t2, l2, alpha=-1.0)

    def pseudoenergy(self, o=None, v=None, ERI=None, l2=None):
        """1/2 <ij|ab> l2_ijab (cclambda.py:547-570). The integrals are the context's; ``o``, ``v`` and
        ``ERI`` are accepted for the reference signature and not used."""
        return self.ccwfn.pseudoenergy(self.l2 if l2 is None else l2)

    def residuals(self, F, t1, t2, l1, l2):
        """r_L1, r_L2 at (l1, l2) on HBAR rebuilt from (F, t1, t2) (cclambda.py:202-256)."""
        cc = self.ccwfn
        F = cc._as_dev(F, (cc.nmo, cc.nmo))
        t1 = cc._as_dev(t1, (cc.no, cc.nv))
        t2 = cc._as_dev(t2, (cc.no, cc.no, cc.nv, cc.nv))
        cc._hbar_for(F, t1, t2, getattr(self.hbar, "stored", False))
        return cc._lambda_residual(l1, l2, with_source=True)

    def solve_lambda(self, e_conv=1e-7, r_conv=1e-7, maxiter=100, max_diis=8, start_diis=1):
        """cclambda.py:69-200: returns the pseudoenergy, or None when ``maxiter`` iterations do not
        converge."""
        cc = self.ccwfn
        h = cc._ctx.h
        stream = cc._stream
        t0 = time.time()
        self.hbar.activate()
        lecc = self.pseudoenergy(l2=self.l2)
        name = "Lambda-amplitudes (%s)" % cc.model
        self.history = [(0, lecc, -lecc, None)]
        if cc.verbose:
            print(_title(name))
            print(_iteration(0, energy=lecc, de=-lecc, e_label="LCC PseudoE"))
        cap = max(1, max_diis) + max(1, start_diis)
        _lib.call("ccx_diis_reset", h, dptr(self.l1), dptr(self.l2), int(max_diis), int(cap), stream())
        coef = (ctypes.c_double * 64)()
        e_dev = cc._e_dev
        for niter in range(1, maxiter + 1):
            lecc_last = lecc
            self.hbar.activate()
            r1, r2 = cc._lambda_residual(self.l1, self.l2, with_source=True)
            _lib.call("ccx_update", h, dptr(self.l1), dptr(self.l2), dptr(r1), dptr(r2),
                      dptr(e_dev[1:]), stream())
            del r1, r2
            rms = math.sqrt(float(e_dev[1].item()))
            lecc = self.pseudoenergy(l2=self.l2)
            ediff = lecc - lecc_last
            self.history.append((niter, lecc, ediff, rms))
            if cc.verbose:
                print(_iteration(niter, energy=lecc, de=ediff, rms=rms, e_label="LCC PseudoE"))
            if abs(ediff) < e_conv and abs(rms) < r_conv:
                if cc.verbose:
                    print("%s converged in %.3f seconds." % (name, time.time() - t0))
                self.lecc = lecc
                return lecc
            _lib.call("ccx_diis_add", h, dptr(self.l1), dptr(self.l2), stream())
            if niter >= start_diis:
                _lib.call("ccx_diis_extrapolate", h, dptr(self.l1), dptr(self.l2), coef, stream())
        return None

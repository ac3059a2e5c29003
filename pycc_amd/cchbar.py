"""Similarity-transformed Hamiltonian on the GPU (counterpart of ``pycc.cchbar.cchbar``).

``cchbar(ccwfn)`` builds the spatial CCSD HBAR blocks of ``pycc/cchbar.py:86-96`` at the current
amplitudes of a ``CCwfn`` (FP64, inside the library context). By default only the blocks of size
o^2 v^2 and below are held, and the Lambda equations evaluate every term that needs Hvvvv, Hvovv or
Hvvvo in factored form from the packed <ab|ef> the context already holds; ``store_vvvv=True`` also
stores those three blocks, as the reference does (Hvvvv is v^4: 174 GB in FP64 at v = 384).

Blocks are read as attributes (``hbar.Hov`` ... ``hbar.Hovoo``), each a fresh FP64 device tensor in
the reference layout.
"""
from __future__ import annotations

import time

from . import _lib
from .device import dptr
from .exceptions import PyCCError

BLOCKS = ("Hov", "Hvv", "Hoo", "Hoooo", "Hvvvv", "Hvovv", "Hooov", "Hovvo", "Hovov", "Hvvvo", "Hovoo")
LARGE = ("Hvvvv", "Hvovv", "Hvvvo")


class cchbar(object):
    """HBAR of a ``CCwfn`` (cchbar.py:54-98, spatial CCSD builders :128-823)."""

    def __init__(self, ccwfn, store_vvvv=False):
        time_init = time.time()
        ccwfn._require_dp("cchbar" + (" with the stored v^4 block" if store_vvvv else ""), mixed_ok=not store_vvvv)
        self.ccwfn = ccwfn
        self.contract = ccwfn.contract
        self.o, self.v = ccwfn.o, ccwfn.v
        self.no, self.nv = ccwfn.no, ccwfn.nv
        self.stored = bool(store_vvvv)
        # the amplitudes this HBAR belongs to (the solver may update ccwfn.t1/t2 in place later)
        self._F = ccwfn.H.F.clone()
        self._t1 = ccwfn.t1.clone()
        self._t2 = ccwfn.t2.clone()
        self.activate()
        if ccwfn.verbose:
            print("HBAR constructed in %.3f seconds." % (time.time() - time_init))

    def activate(self):
        """Make this HBAR the context's current one (rebuilt if another replaced it)."""
        self.ccwfn._hbar_for(self._F, self._t1, self._t2, self.stored, owner=self)

    def _shape(self, name):
        o, v = self.no, self.nv
        dims = {"o": o, "v": v}
        return tuple(dims[c] for c in name[1:].lower())

    def block(self, name):
        if name not in BLOCKS:
            raise PyCCError("unknown HBAR block %r" % name)
        if name in LARGE and not self.stored:
            raise PyCCError("%s is not held by a factored HBAR: build cchbar(ccwfn, store_vvvv=True)" % name)
        self.activate()
        torch = self.ccwfn._torch
        out = torch.empty(self._shape(name), dtype=torch.float64, device=self.ccwfn.device1)
        _lib.call("ccx_hbar_block", self.ccwfn._ctx.h, BLOCKS.index(name), dptr(out), self.ccwfn._stream())
        return out

    def __getattr__(self, name):
        if name in BLOCKS:
            return self.block(name)
        raise AttributeError(name)

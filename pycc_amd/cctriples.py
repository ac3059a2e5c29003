"""The (T) correction and the batched T3 builders of closed-shell CCSD(T) (counterpart of
``pycc/cctriples.py``).

``t_tjl(ccwfn)`` (``cctriples.py:177-240``), ``t_vikings(ccwfn)`` (``:243``) and
``t_vikings_inverted(ccwfn)`` (``:311``) are three formulations of one quantity, E(T) at the
wavefunction's current amplitudes; all three evaluate it through ``ccx_triples`` (one v^3 block of
T3 numerators per occupied triple i >= j >= k, formed tile by tile in LDS and never stored), in FP64
except in the MIXED_F16X3 mode, whose GEMMs are FP16x3 split products.

``t3c_ijk``, ``t3c_abc``, ``t3d_ijk`` and ``t3d_abc`` (``cctriples.py:27-173``) return one batch of
T3 amplitudes, the connected or the disconnected part, for fixed (i, j, k) or fixed (a, b, c), with the
reference's arguments and layouts: ``o`` and ``v`` are the occupied and virtual slices of the Fock
matrix ``F``, ``t2`` is [i][j][a][b], ``Wvvvo``/``Wovoo``/``Woovv`` are <pq|rs> blocks (any strided
device views, e.g. slices of the full ERI tensor) and ``contract`` is the wavefunction's device
contraction backend (``ccwfn.contract``). Each term runs on the library's FP64 MFMA contraction
kernel, accumulated in place in the output (beta = 1), and ``WithDenom=True`` divides by the
orbital-energy denominator in one kernel (``ccx_t3_denominator``).
"""
from __future__ import annotations

import ctypes

from . import _lib
from .device import ContractionBackend, current_stream_handle, dptr
from .exceptions import PyCCError


def t_tjl(ccwfn) -> float:
    """E(T) (``pycc/cctriples.py:177-240``) from ``ccwfn.t1``, ``ccwfn.t2`` and ``ccwfn.H.F``."""
    return ccwfn.triples_energy()


def t_vikings(ccwfn) -> float:
    """E(T) as ``pycc/cctriples.py:243`` returns it (the same value as :func:`t_tjl`)."""
    return ccwfn.triples_energy()


def t_vikings_inverted(ccwfn) -> float:
    """E(T) as ``pycc/cctriples.py:311`` returns it (the same value as :func:`t_tjl`)."""
    return ccwfn.triples_energy()


# The twelve connected terms of cctriples.py:50-62 (fixed i, j, k) and :83-95 (fixed a, b, c), each as
# (sign, subscripts, W block, its fixed index, t2 slice): a W block of +1 is Wvvvo, of -1 Wovoo.
_C_IJK = [(+1, 'bae,ce->abc', 'i', ('k', 'j')), (+1, 'cae,be->abc', 'i', ('j', 'k')),
          (+1, 'ace,be->abc', 'k', ('j', 'i')), (+1, 'bce,ae->abc', 'k', ('i', 'j')),
          (+1, 'cbe,ae->abc', 'j', ('i', 'k')), (+1, 'abe,ce->abc', 'j', ('k', 'i')),
          (-1, 'mc,mab->abc', ('j', 'k'), 'i'), (-1, 'mb,mac->abc', ('k', 'j'), 'i'),
          (-1, 'mb,mca->abc', ('i', 'j'), 'k'), (-1, 'ma,mcb->abc', ('j', 'i'), 'k'),
          (-1, 'ma,mbc->abc', ('k', 'i'), 'j'), (-1, 'mc,mba->abc', ('i', 'k'), 'j')]
_C_ABC = [(+1, 'ei,kje->ijk', ('b', 'a'), 'c'), (+1, 'ei,jke->ijk', ('c', 'a'), 'b'),
          (+1, 'ek,jie->ijk', ('a', 'c'), 'b'), (+1, 'ek,ije->ijk', ('b', 'c'), 'a'),
          (+1, 'ej,ike->ijk', ('c', 'b'), 'a'), (+1, 'ej,kie->ijk', ('a', 'b'), 'c'),
          (-1, 'mjk,im->ijk', 'c', ('a', 'b')), (-1, 'mkj,im->ijk', 'b', ('a', 'c')),
          (-1, 'mij,km->ijk', 'b', ('c', 'a')), (-1, 'mji,km->ijk', 'a', ('c', 'b')),
          (-1, 'mki,jm->ijk', 'a', ('b', 'c')), (-1, 'mik,jm->ijk', 'c', ('b', 'a'))]


def _check(contract, F, o, v, *tensors):
    import torch
    if not isinstance(contract, ContractionBackend):
        raise PyCCError("the T3 builders run on the device contraction backend (ccwfn.contract)")
    if not isinstance(o, slice) or not isinstance(v, slice):
        raise PyCCError("o and v must be slices of the Fock matrix")
    dev = contract.device1
    for t in (F,) + tensors:
        if not torch.is_tensor(t) or t.dtype != torch.float64 or t.device != dev:
            raise PyCCError("the T3 builders take float64 tensors on %s" % dev)
    if F.dim() != 2 or F.shape[0] != F.shape[1] or F.stride(1) != 1 or F.stride(0) != F.shape[1]:
        raise PyCCError("F must be a contiguous square Fock matrix")
    nmo = F.shape[0]
    o0, o1, _ = o.indices(nmo)
    v0, v1, _ = v.indices(nmo)
    if o.step not in (None, 1) or v.step not in (None, 1) or o1 < o0 or v1 < v0:
        raise PyCCError("o and v must be contiguous slices")
    return o0, o1 - o0, v0, v1 - v0


def _range(x, n, what):
    x = int(x)
    if not 0 <= x < n:
        raise PyCCError("%s = %d outside [0, %d)" % (what, x, n))
    return x


def _out(contract, n):
    import torch
    return torch.zeros((n, n, n), dtype=torch.float64, device=contract.device1)


def _divide(F, p, q, r, e0, n, occ_fixed, t3):
    _lib.call("ccx_t3_denominator", dptr(F), F.shape[0], p, q, r, e0, n, 1 if occ_fixed else 0, dptr(t3),
              ctypes.c_void_p(current_stream_handle()))
    return t3


def t3c_ijk(o, v, i, j, k, t2, Wvvvo, Wovoo, F, contract, WithDenom=True):
    """Connected T3 amplitudes for fixed i, j, k, shape (nv, nv, nv) (``pycc/cctriples.py:27-72``):
    -P(ijk/abc) [t2_ijae W_bcek - t2_imab W_mcjk], divided by D_ijkabc when ``WithDenom``."""
    o0, no, v0, nv = _check(contract, F, o, v, t2, Wvvvo, Wovoo)
    ix = {'i': _range(i, no, 'i'), 'j': _range(j, no, 'j'), 'k': _range(k, no, 'k')}
    t3 = _out(contract, nv)
    for sign, sub, w, t in _C_IJK:
        if sign > 0:
            A, B = Wvvvo[:, :, :, ix[w]], t2[ix[t[0]], ix[t[1]]]
        else:
            A, B = Wovoo[:, :, ix[w[0]], ix[w[1]]], t2[ix[t]]
        contract(sub, A, B, alpha=float(sign), out=t3, beta=1.0)
    if WithDenom:
        _divide(F, o0 + ix['i'], o0 + ix['j'], o0 + ix['k'], v0, nv, True, t3)
    return t3


def t3c_abc(o, v, a, b, c, t2, Wvvvo, Wovoo, F, contract, WithDenom=True):
    """Connected T3 amplitudes for fixed a, b, c, shape (no, no, no) (``pycc/cctriples.py:75-105``):
    the same amplitudes as :func:`t3c_ijk`, batched over i, j, k."""
    o0, no, v0, nv = _check(contract, F, o, v, t2, Wvvvo, Wovoo)
    ix = {'a': _range(a, nv, 'a'), 'b': _range(b, nv, 'b'), 'c': _range(c, nv, 'c')}
    t3 = _out(contract, no)
    for sign, sub, w, t in _C_ABC:
        if sign > 0:
            A, B = Wvvvo[ix[w[0]], ix[w[1]]], t2[:, :, ix[t]]
        else:
            A, B = Wovoo[:, ix[w]], t2[:, :, ix[t[0]], ix[t[1]]]
        contract(sub, A, B, alpha=float(sign), out=t3, beta=1.0)
    if WithDenom:
        _divide(F, v0 + ix['a'], v0 + ix['b'], v0 + ix['c'], o0, no, False, t3)
    return t3


def t3d_ijk(o, v, i, j, k, t1, t2, Woovv, F, contract, WithDenom=True):
    """Disconnected T3 amplitudes for fixed i, j, k, shape (nv, nv, nv) (``pycc/cctriples.py:108-147``):
    W_ijab t1_kc + W_ikac t1_jb + W_jkbc t1_ia + t2_ijab f_kc + t2_ikac f_jb + t2_jkbc f_ia."""
    o0, no, v0, nv = _check(contract, F, o, v, t1, t2, Woovv)
    i, j, k = _range(i, no, 'i'), _range(j, no, 'j'), _range(k, no, 'k')
    Fov = F[o, v]
    t3 = _out(contract, nv)
    for X, Y in ((Woovv, t1), (t2, Fov)):
        contract('ab,c->abc', X[i, j], Y[k], out=t3, beta=1.0)
        contract('ac,b->abc', X[i, k], Y[j], out=t3, beta=1.0)
        contract('bc,a->abc', X[j, k], Y[i], out=t3, beta=1.0)
    if WithDenom:
        _divide(F, o0 + i, o0 + j, o0 + k, v0, nv, True, t3)
    return t3


def t3d_abc(o, v, a, b, c, t1, t2, Woovv, F, contract, WithDenom=True):
    """Disconnected T3 amplitudes for fixed a, b, c, shape (no, no, no) (``pycc/cctriples.py:149-173``):
    the same amplitudes as :func:`t3d_ijk`, batched over i, j, k."""
    o0, no, v0, nv = _check(contract, F, o, v, t1, t2, Woovv)
    a, b, c = _range(a, nv, 'a'), _range(b, nv, 'b'), _range(c, nv, 'c')
    Fov = F[o, v]
    t3 = _out(contract, no)
    for X, Y in ((Woovv, t1), (t2, Fov)):
        contract('ij,k->ijk', X[:, :, a, b], Y[:, c], out=t3, beta=1.0)
        contract('ik,j->ijk', X[:, :, a, c], Y[:, b], out=t3, beta=1.0)
        contract('jk,i->ijk', X[:, :, b, c], Y[:, a], out=t3, beta=1.0)
    if WithDenom:
        _divide(F, v0 + a, v0 + b, v0 + c, o0, no, False, t3)
    return t3

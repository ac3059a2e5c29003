"""Numpy FP64 restatement of pycc's closed-shell (T) correction (oracle -- test infrastructure).

* ``pycc/cctriples.py:27-72``   ``t3c_ijk``: connected T3 numerator for fixed i, j, k
* ``pycc/cctriples.py:108-146`` ``t3d_ijk``: disconnected numerator (<ij|ab> t1_kc + t2_ijab f_kc, ...)
* ``pycc/cctriples.py:64-70, 97-103`` ``denominator``: D_ijkabc over an i, j, k or an a, b, c batch
* ``pycc/cctriples.py:177-240`` ``t_tjl``: Lee-Rendell sum over i >= j >= k and a >= b >= c with
  the X3 / Y3 / Z3 intermediates and the (1 + d_ab + d_ac + d_bc) and (2 - d_ij - d_ik - d_jk)
  weights

The a, b, c loops of the reference are written as masks over whole v^3 arrays. Pinned by the
reference's own (T) energies for water (``pycc/tests/test_005_ccsd_t_energy.py:33,44``, STO-3G and
cc-pVDZ, frozen core) through the psi4-free front end, and by ``t_full`` below (the textbook
full-loop form over all i, j, k) on tiny systems; the a, b, c batch of ``denominator`` equals its
i, j, k batch element for element (tests/test_triples_oracle.py).
"""
from __future__ import annotations

import numpy as np


def contract(subscripts, *operands):
    return np.einsum(subscripts, *operands, optimize=True)


def t3c_ijk(i, j, k, t2, Wvvvo, Wovoo):
    """cctriples.py:50-62 (WithDenom=False)."""
    t3 = contract('bae,ce->abc', Wvvvo[:, :, :, i], t2[k, j])
    t3 += contract('cae,be->abc', Wvvvo[:, :, :, i], t2[j, k])
    t3 += contract('ace,be->abc', Wvvvo[:, :, :, k], t2[j, i])
    t3 += contract('bce,ae->abc', Wvvvo[:, :, :, k], t2[i, j])
    t3 += contract('cbe,ae->abc', Wvvvo[:, :, :, j], t2[i, k])
    t3 += contract('abe,ce->abc', Wvvvo[:, :, :, j], t2[k, i])
    t3 -= contract('mc,mab->abc', Wovoo[:, :, j, k], t2[i])
    t3 -= contract('mb,mac->abc', Wovoo[:, :, k, j], t2[i])
    t3 -= contract('mb,mca->abc', Wovoo[:, :, i, j], t2[k])
    t3 -= contract('ma,mcb->abc', Wovoo[:, :, j, i], t2[k])
    t3 -= contract('ma,mbc->abc', Wovoo[:, :, k, i], t2[j])
    t3 -= contract('mc,mba->abc', Wovoo[:, :, i, k], t2[j])
    return t3


def t3d_ijk(i, j, k, t1, t2, Woovv, Fov):
    """cctriples.py:132-137 (WithDenom=False)."""
    t3 = contract('ab,c->abc', Woovv[i, j], t1[k])
    t3 += contract('ac,b->abc', Woovv[i, k], t1[j])
    t3 += contract('bc,a->abc', Woovv[j, k], t1[i])
    t3 += contract('ab,c->abc', t2[i, j], Fov[k])
    t3 += contract('ac,b->abc', t2[i, k], Fov[j])
    t3 += contract('bc,a->abc', t2[j, k], Fov[i])
    return t3


def denominator(Fo, Fv, i=None, j=None, k=None, a=None, b=None, c=None):
    """D_ijkabc = (f_ii + f_jj + f_kk) - (f_aa + f_bb + f_cc) over one batch: fixed (i, j, k) and all
    a, b, c (cctriples.py:64-70, shape (nv, nv, nv)), or fixed (a, b, c) and all i, j, k (:97-103,
    shape (no, no, no)); each sum of three added left to right, as the reference's broadcasts add it.
    t_tjl and t_full divide by the first form."""
    if i is not None:
        return (Fo[i] + Fo[j] + Fo[k]) - (Fv.reshape(-1, 1, 1) + Fv.reshape(-1, 1) + Fv)
    return (Fo.reshape(-1, 1, 1) + Fo.reshape(-1, 1) + Fo) - (Fv[a] + Fv[b] + Fv[c])


def _blocks(F, ERI, no, nfzc):
    nmo = F.shape[0]
    o = slice(nfzc, nfzc + no)
    v = slice(nfzc + no, nmo)
    return o, v, ERI[v, v, v, o], ERI[o, v, o, o], ERI[o, o, v, v], F[o, v], np.diag(F)[o], np.diag(F)[v]


def t_tjl(F, ERI, t1, t2, no, nfzc=0):
    """cctriples.py:177-240: E(T) from the CCSD amplitudes (pycc layouts: F (nmo, nmo), ERI Dirac
    <pq|rs>, t1 [i][a], t2 [i][j][a][b] over the active orbitals after ``nfzc`` frozen ones)."""
    o, v, Wvvvo, Wovoo, Woovv, Fov, Fo, Fv = _blocks(F, ERI, no, nfzc)
    nv = Fv.shape[0]
    a, b, c = np.meshgrid(np.arange(nv), np.arange(nv), np.arange(nv), indexing='ij')
    norm = 1.0 + (a == b) + (a == c) + (b == c)                         # cctriples.py:210-213
    tri = (a >= b) & (b >= c)                                            # cctriples.py:231-233
    ET = 0.0
    for i in range(no):
        for j in range(i + 1):
            for k in range(j + 1):
                W3 = t3c_ijk(i, j, k, t2, Wvvvo, Wovoo)
                V3 = (t3d_ijk(i, j, k, t1, t2, Woovv, Fov) + W3) / norm
                X3 = (W3 * V3 + W3.swapaxes(1, 2) * V3.swapaxes(1, 2)
                      + W3.swapaxes(0, 1) * V3.swapaxes(0, 1)
                      + W3.swapaxes(0, 1).swapaxes(1, 2) * V3.swapaxes(0, 1).swapaxes(1, 2)
                      + W3.swapaxes(0, 1).swapaxes(0, 2) * V3.swapaxes(0, 1).swapaxes(0, 2)
                      + W3.swapaxes(0, 2) * V3.swapaxes(0, 2))
                Y3 = V3 + V3.swapaxes(0, 1).swapaxes(1, 2) + V3.swapaxes(0, 1).swapaxes(0, 2)
                Z3 = V3.swapaxes(1, 2) + V3.swapaxes(0, 1) + V3.swapaxes(0, 2)
                denom = denominator(Fo, Fv, i=i, j=j, k=k)
                Wcyc = W3 + W3.transpose(1, 2, 0) + W3.transpose(2, 0, 1)      # W_abc + W_bca + W_cab
                Wacy = W3.transpose(0, 2, 1) + W3.transpose(1, 0, 2) + W3.transpose(2, 1, 0)
                term = ((Y3 - 2.0 * Z3) * Wcyc + (Z3 - 2.0 * Y3) * Wacy + 3.0 * X3) / denom
                ET += float(term[tri].sum()) * (2.0 - ((i == j) + (i == k) + (j == k)))
    return ET


def t_full(F, ERI, t1, t2, no, nfzc=0):
    """The textbook closed-shell (T) energy over all i, j, k and a, b, c:
    E = sum (4 W_abc + W_bca + W_cab)(V_abc - V_cba) / (3 D), W = t3c_ijk, V = W + t3d_ijk
    (Rendell, Lee, Komornicki, Chem. Phys. Lett. 178, 462 (1991)); an independent check of t_tjl's
    triangular sums and weights."""
    o, v, Wvvvo, Wovoo, Woovv, Fov, Fo, Fv = _blocks(F, ERI, no, nfzc)
    ET = 0.0
    for i in range(no):
        for j in range(no):
            for k in range(no):
                W = t3c_ijk(i, j, k, t2, Wvvvo, Wovoo)
                V = W + t3d_ijk(i, j, k, t1, t2, Woovv, Fov)
                D = denominator(Fo, Fv, i=i, j=j, k=k)
                Q = 4.0 * W + W.transpose(1, 2, 0) + W.transpose(2, 0, 1)
                ET += float((Q * (V - V.transpose(2, 1, 0)) / (3.0 * D)).sum())
    return ET

"""The (T) correction of closed-shell CCSD(T) (counterpart of ``pycc/cctriples.py``).

``t_tjl(ccwfn)`` (``cctriples.py:177-240``), ``t_vikings(ccwfn)`` (``:243``) and
``t_vikings_inverted(ccwfn)`` (``:311``) are three formulations of one quantity, E(T) at the
wavefunction's current amplitudes; all three evaluate it through ``ccx_triples`` (one v^3 block of
T3 numerators per occupied triple i >= j >= k, never the full o^3 v^3 array), in FP64 except in
the MIXED_F16X3 mode, whose GEMMs are FP16x3 split products.
"""
from __future__ import annotations


def t_tjl(ccwfn) -> float:
    """E(T) (``pycc/cctriples.py:177-240``) from ``ccwfn.t1``, ``ccwfn.t2`` and ``ccwfn.H.F``."""
    return ccwfn.triples_energy()


def t_vikings(ccwfn) -> float:
    """E(T) as ``pycc/cctriples.py:243`` returns it (the same value as :func:`t_tjl`)."""
    return ccwfn.triples_energy()


def t_vikings_inverted(ccwfn) -> float:
    """E(T) as ``pycc/cctriples.py:311`` returns it (the same value as :func:`t_tjl`)."""
    return ccwfn.triples_energy()

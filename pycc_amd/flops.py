"""Algorithmic work of one spin-adapted CCSD iteration, as the reference computes it.

``f_ref(o, v)`` sums, over every contraction of pycc's spatial CCSD residual and energy
(``pycc/ccwfn.py:455`` tau x8, Fae :495-497, Fmi :531-533, Fme :564, Wmnij :597-603,
Wmbej :642-645, Wmbje :681-683, Zmbij :715, r_T1 :755-760, r_T2 :923-940, energy :1160-1161),
2 * prod(extents) for each two-operand contraction and prod(extents) for each outer product.
This is the count TFLOP/s figures are quoted against, whatever algorithm runs.

``f_kernel(o, v)`` is what this package's kernels actually execute for the ladder/Z GEMMs
(tau+- packing over i<=j, e<=f, a<=b) and the six ring GEMM units.
"""
from __future__ import annotations

# (subscripts, count per iteration); outer products have no summed index
REFERENCE_SITES = [
    ("ia,jb->ijab", 8),                                                     # build_tau x8
    ("me,ma->ae", 1), ("mf,mafe->ae", 1), ("mnaf,mnef->ae", 1),            # Fae
    ("ie,me->mi", 1), ("ne,mnie->mi", 1), ("inef,mnef->mi", 1),            # Fmi
    ("nf,mnef->me", 1),                                                     # Fme
    ("je,mnie->mnij", 1), ("ie,mnej->mnij", 1), ("ijef,mnef->mnij", 1),    # Wmnij
    ("jf,mbef->mbej", 1), ("nb,mnej->mbej", 1), ("jnfb,mnef->mbej", 1), ("njfb,mnef->mbej", 1),
    ("jf,mbfe->mbje", 1), ("nb,mnje->mbje", 1), ("jnfb,mnfe->mbje", 1),    # Wmbje
    ("mbef,ijef->mbij", 1),                                                 # Zmbij
    ("ie,ae->ia", 1), ("mi,ma->ia", 1), ("imae,me->ia", 1), ("nf,nafi->ia", 1),
    ("mief,maef->ia", 1), ("mnae,nmei->ia", 1),                             # r_T1
    ("ijae,eb->ijab", 2), ("bm,me->be", 1), ("imab,mj->ijab", 2), ("je,em->jm", 1),
    ("mnab,mnij->ijab", 1), ("ijef,abef->ijab", 1), ("ma,mbij->ijab", 2),
    ("imae,mbej->ijab", 2), ("mjae,mbie->ijab", 1), ("ei,am->aemi", 1),
    ("imea,mbej->ijab", 1), ("imeb,maje->ijab", 1), ("ie,abej->ijab", 1),   # r_T2
    ("ia,ia->", 1), ("ijab,ijab->", 1),                                     # cc_energy
]

OCC = set("ijmn")


def _site_flops(sub, o, v):
    lhs, out = sub.split("->")
    a, b = lhs.split(",")
    idx = set(a) | set(b)
    prod = 1
    for c in idx:
        prod *= o if c in OCC else v
    summed = (set(a) & set(b)) - set(out)
    return prod if not summed else 2 * prod


def f_ref(o: int, v: int) -> float:
    return float(sum(n * _site_flops(s, o, v) for s, n in REFERENCE_SITES))


def f_ladder_kernel(o: int, v: int) -> float:
    """Packed ladder + Z GEMMs: 2 M N K over the (anti)symmetrised pairs."""
    npP, npM = v * (v + 1) // 2, v * (v - 1) // 2
    noP, noM = o * (o + 1) // 2, o * (o - 1) // 2
    return 2.0 * noP * (npP + o * v) * npP + 2.0 * noM * (npM + o * v) * npM


def f_ring_kernel(o: int, v: int) -> float:
    """Six ov x ov x ov GEMM units: J.[S1a|S1b] (2), K.A1 (1), [U|T].[Wej;Wje] (2), Tx.Wje (1)."""
    ov = o * v
    return 6 * 2.0 * ov ** 3


def f_kernel(o: int, v: int) -> float:
    return f_ladder_kernel(o, v) + f_ring_kernel(o, v)


# Sites of F_ref that the kernels evaluate differently: the ladder (:931) and Zmbij (:715) as the
# packed ladder GEMM, the six o^3 v^3 W / ring contractions (:644, :645, :683, :933-935) as the ring
# GEMM units, and the t1 t1 <ov|vo> / <ov|ov> terms (:937-938), factored as t_ma (t_ie <mb|ej>)
# (4 o^3 v^2 each).
_REPLACED = {"ijef,abef->ijab", "mbef,ijef->mbij", "jnfb,mnef->mbej", "njfb,mnef->mbej", "jnfb,mnfe->mbje",
             "imae,mbej->ijab", "mjae,mbie->ijab", "imea,mbej->ijab", "imeb,maje->ijab"}
# the two o^4 v^2 hole-hole terms (:603, :930), which run in the ring / ladder precision
_HOLE_HOLE = {"ijef,mnef->mnij", "mnab,mnij->ijab"}


def f_rest(o: int, v: int) -> float:
    """Flops of every contraction besides the ladder, Z and the ring units, as the kernels run them."""
    rest = sum(n * _site_flops(s, o, v) for s, n in REFERENCE_SITES if s not in _REPLACED)
    return float(rest + 2 * 4.0 * o ** 3 * v ** 2)


def f_hole_hole(o: int, v: int) -> float:
    return float(sum(n * _site_flops(s, o, v) for s, n in REFERENCE_SITES if s in _HOLE_HOLE))


def f_executed(o: int, v: int) -> float:
    """Flops one residual evaluation executes (each product counted once, whatever its precision)."""
    return f_kernel(o, v) + f_rest(o, v)

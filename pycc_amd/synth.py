"""Seeded synthetic closed-shell Hamiltonians for the CCSD solver.

The generator draws, in this order, from ``numpy.random.default_rng(seed)``:

1. occupied orbital energies ``eps_o ~ U[-1.5, -0.4]`` (sorted),
2. virtual orbital energies ``eps_v ~ U[0.2, 3.0]`` (sorted),
3. ``X ~ N(0, 1)`` of shape ``(naux, nmo, nmo)``, ``naux = 2 * nmo`` by default.

The density-fitting factors are ``B = (triu(X) + triu(X, 1)^T) * exp(-|p - q| / nmo)``, symmetric
in ``pq``. The chemist integrals are ``(pq|rs) = sum_Q B_Qpq B_Qrs`` (8-fold symmetric and positive
semidefinite), the Dirac integrals ``<pq|rs> = (pr|qs)`` and the Fock matrix is ``diag(eps)``.
``B`` is scaled by one factor so that the largest MP1 doubles amplitude
``|<ij|ab> / (e_i + e_j - e_a - e_b)|`` equals ``t2max`` (0.03 by default).

The dense helpers (:func:`dirac_eri`, :class:`DenseHamiltonian`) materialise the full ``nmo^4``
array with the layout pycc's ``Hamiltonian`` uses (``hamiltonian.py:67-70``: Dirac ``ERI`` and
``L = 2 ERI - ERI.swapaxes(2, 3)``); they are meant for small ``nmo`` only. Large runs hand ``B``
itself to the device library, which builds its integral blocks on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SyntheticDF:
    """Orbital energies and scaled density-fitting factors of one synthetic system."""
    no: int
    nv: int
    seed: int
    eps: np.ndarray      # (nmo,)
    B: np.ndarray        # (naux, nmo, nmo), symmetric in the last two indices
    scale: float         # factor applied to the raw B

    @property
    def nmo(self) -> int:
        return self.no + self.nv

    @property
    def naux(self) -> int:
        return self.B.shape[0]

    @property
    def fock(self) -> np.ndarray:
        return np.diag(self.eps)


def _max_mp1(eps_o, eps_v, Bov, block_bytes=2.5e8):
    """max_ijab |(ia|jb) / (e_i + e_j - e_a - e_b)| computed in row blocks of i."""
    naux, no, nv = Bov.shape
    flat = Bov.reshape(naux, no * nv)
    best = 0.0
    bi = max(1, int(block_bytes // (nv * no * nv * 8)))
    for i0 in range(0, no, bi):
        i1 = min(no, i0 + bi)
        rows = flat[:, i0 * nv:i1 * nv]
        iajb = (rows.T @ flat).reshape(i1 - i0, nv, no, nv)
        d = (eps_o[i0:i1, None, None, None] + eps_o[None, None, :, None]
             - eps_v[None, :, None, None] - eps_v[None, None, None, :])
        best = max(best, float(np.abs(iajb / d).max()))
    return best


def synthetic_df(no: int, nv: int, seed: int, t2max: float = 0.03, naux: int | None = None) -> SyntheticDF:
    """Draw one synthetic closed-shell system (see the module docstring for the recipe)."""
    if no < 1 or nv < 1:
        raise ValueError("need at least one occupied and one virtual orbital (got no=%d, nv=%d)" % (no, nv))
    rng = np.random.default_rng(seed)
    nmo = no + nv
    naux = 2 * nmo if naux is None else int(naux)
    eps_o = np.sort(rng.uniform(-1.5, -0.4, no))
    eps_v = np.sort(rng.uniform(0.2, 3.0, nv))
    X = rng.standard_normal((naux, nmo, nmo))
    idx = np.arange(nmo)
    weight = np.exp(-np.abs(idx[:, None] - idx[None, :]) / nmo)
    B = np.triu(X)
    B += np.swapaxes(np.triu(X, 1), 1, 2)
    del X
    B *= weight
    raw = _max_mp1(eps_o, eps_v, np.ascontiguousarray(B[:, :no, no:]))
    scale = float(np.sqrt(t2max / raw))
    B *= scale
    return SyntheticDF(no=no, nv=nv, seed=seed, eps=np.concatenate([eps_o, eps_v]), B=B, scale=scale)


def chemist_eri(B: np.ndarray) -> np.ndarray:
    """Dense chemist integrals (pq|rs) = sum_Q B_Qpq B_Qrs, shape (nmo,)*4."""
    naux, nmo, _ = B.shape
    flat = B.reshape(naux, nmo * nmo)
    return (flat.T @ flat).reshape(nmo, nmo, nmo, nmo)


def dirac_eri(B: np.ndarray) -> np.ndarray:
    """Dense Dirac integrals <pq|rs> = (pr|qs), C-contiguous, shape (nmo,)*4."""
    return np.ascontiguousarray(chemist_eri(B).transpose(0, 2, 1, 3))


class DenseHamiltonian:
    """A pycc-layout Hamiltonian: ``F`` (nmo, nmo), ``eps``, Dirac ``ERI`` and ``L``.

    Mirrors the attributes pycc's ``Hamiltonian`` exposes (``hamiltonian.py:59-70``):
    ``F``, ``eps = diag(F)``, ``ERI[p, q, r, s] = <pq|rs>`` and ``L = 2 ERI - ERI.swapaxes(2, 3)``.
    """

    def __init__(self, F: np.ndarray, ERI: np.ndarray):
        self.F = np.asarray(F, dtype=np.float64)
        self.eps = np.diagonal(self.F).copy()
        self.ERI = np.asarray(ERI, dtype=np.float64)
        self.L = 2.0 * self.ERI - self.ERI.swapaxes(2, 3)

    @classmethod
    def from_df(cls, sd: SyntheticDF) -> "DenseHamiltonian":
        return cls(sd.fock, dirac_eri(sd.B))

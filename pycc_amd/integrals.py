"""Integral files and the closed-shell MO Hamiltonian built from them.

pycc reads its integrals from a psi4 RHF wavefunction (``pycc/hamiltonian.py:36-70``): the MO Fock
matrix (``hamiltonian.py:59``, the AO Fock in the MO basis, which for RHF orbitals equals
``h + sum_i (2 (pq|ii) - (pi|iq))``), the Dirac-ordered ``ERI`` (``hamiltonian.py:68-69``) and the
occupations (``wavefunction.py:304-305``: ``no`` = doubly occupied minus frozen core). psi4 is not
part of this package, so the same inputs come from files instead:

* FCIDUMP (Knowles & Handy, Comput. Phys. Commun. 54, 75 (1989)): a namelist header (NORB, NELEC,
  MS2, ORBSYM, ISYM) and lines ``value i j k l`` holding the chemist-ordered ``(ij|kl)`` (8-fold
  symmetry, one representative per class), ``h_ij`` (k = l = 0) and the core energy (all zero).
* ``.npz``: arrays ``h`` (norb x norb), ``eri`` (chemist (pq|rs), norb^4) and scalars ``nelec``,
  ``ecore``.

``MOHamiltonian`` carries ``F``, ``ERI`` (Dirac), ``h``, ``ecore``, ``nocc`` and ``escf`` and is
accepted by ``CCwfn`` like pycc's Hamiltonian (``CCwfn(H, nfzc=...)`` takes ``nocc - nfzc`` active
occupied orbitals). Everything here is host-side numpy; the integrals go to the device when a
``CCwfn`` is built.
"""
from __future__ import annotations

import io
import re
from dataclasses import dataclass, field

import numpy as np

from .exceptions import PyCCError

_HEADER_KEYS = ("NORB", "NELEC", "MS2", "ISYM")


@dataclass
class FCIDump:
    norb: int
    nelec: int
    ms2: int
    h: np.ndarray            # (norb, norb) MO core Hamiltonian
    eri: np.ndarray          # (norb,)*4 chemist (pq|rs)
    ecore: float = 0.0
    orbsym: list = field(default_factory=list)
    isym: int = 1


def _parse_header(text):
    """NORB / NELEC / MS2 / ISYM / ORBSYM from the namelist text between &FCI and &END (or /)."""
    vals = {}
    for key in _HEADER_KEYS:
        m = re.search(r"\b%s\s*=\s*(-?\d+)" % key, text, re.I)
        if m:
            vals[key] = int(m.group(1))
    m = re.search(r"\bORBSYM\s*=\s*([-\d,\s]+?)(?=,?\s*[A-Z&/]|$)", text, re.I)
    orbsym = [int(x) for x in re.findall(r"-?\d+", m.group(1))] if m else []
    if "NORB" not in vals or "NELEC" not in vals:
        raise PyCCError("FCIDUMP header lacks NORB or NELEC")
    return vals["NORB"], vals["NELEC"], vals.get("MS2", 0), vals.get("ISYM", 1), orbsym


def read_fcidump(path) -> FCIDump:
    """Read an FCIDUMP file (restricted orbitals). Fortran ``D`` exponents are accepted."""
    with open(path) as fh:
        text = fh.read()
    m = re.search(r"&END|^\s*/\s*$", text, re.I | re.M)
    if not text.lstrip().upper().startswith("&FCI") or m is None:
        raise PyCCError("%s: not an FCIDUMP file (no &FCI ... &END header)" % path)
    norb, nelec, ms2, isym, orbsym = _parse_header(text[:m.start()])
    if ms2 != 0:
        raise PyCCError("FCIDUMP with MS2=%d: only closed-shell (MS2=0) integrals are supported" % ms2)
    body = text[m.end():].replace("D", "E").replace("d", "e")
    data = np.loadtxt(io.StringIO(body), ndmin=2) if body.strip() else np.zeros((0, 5))
    if data.shape[1] != 5:
        raise PyCCError("%s: integral lines must have 5 fields (value i j k l)" % path)
    val = data[:, 0]
    idx = data[:, 1:].astype(np.int64)
    if idx.min(initial=0) < 0 or idx.max(initial=0) > norb:
        raise PyCCError("%s: orbital index outside 1..NORB=%d" % (path, norb))
    i, j, k, l = idx.T
    two = (i > 0) & (j > 0) & (k > 0) & (l > 0)
    one = (i > 0) & (j > 0) & (k == 0) & (l == 0)
    core = (i == 0) & (j == 0) & (k == 0) & (l == 0)
    if not np.all(two | one | core):
        raise PyCCError("%s: lines other than (ij|kl), h_ij and the core energy (e.g. orbital energies) "
                        "are not supported" % path)
    eri = np.zeros((norb,) * 4)
    p, q, r, s = i[two] - 1, j[two] - 1, k[two] - 1, l[two] - 1
    v = val[two]
    for a, b, c, d in ((p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                       (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
        eri[a, b, c, d] = v
    h = np.zeros((norb, norb))
    h[i[one] - 1, j[one] - 1] = val[one]
    h[j[one] - 1, i[one] - 1] = val[one]
    ecore = float(val[core].sum())
    return FCIDump(norb, nelec, ms2, h, eri, ecore, orbsym, isym)


def write_fcidump(path, h, eri, nelec, ecore=0.0, ms2=0, orbsym=None, isym=1, tol=1e-14):
    """Write h (norb x norb) and chemist eri (norb^4) as an FCIDUMP file: one representative
    (i >= j, k >= l, ij >= kl) per 8-fold class, entries with |value| <= tol omitted."""
    h = np.asarray(h, dtype=np.float64)
    eri = np.asarray(eri, dtype=np.float64)
    norb = h.shape[0]
    if h.shape != (norb, norb) or eri.shape != (norb,) * 4:
        raise PyCCError("h must be (norb, norb) and eri (norb,)*4")
    orbsym = list(orbsym) if orbsym is not None else [1] * norb
    lines = ["&FCI NORB=%d,NELEC=%d,MS2=%d," % (norb, nelec, ms2),
             " ORBSYM=" + ",".join(str(x) for x in orbsym) + ",",
             " ISYM=%d," % isym, "&END"]
    ij = [(a, b) for a in range(norb) for b in range(a + 1)]
    for x, (a, b) in enumerate(ij):
        for (c, d) in ij[:x + 1]:
            v = eri[a, b, c, d]
            if abs(v) > tol:
                lines.append("%23.16e %4d %4d %4d %4d" % (v, a + 1, b + 1, c + 1, d + 1))
    for (a, b) in ij:
        if abs(h[a, b]) > tol:
            lines.append("%23.16e %4d %4d %4d %4d" % (h[a, b], a + 1, b + 1, 0, 0))
    lines.append("%23.16e %4d %4d %4d %4d" % (ecore, 0, 0, 0, 0))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def load_npz(path) -> FCIDump:
    """``.npz`` with ``h``, ``eri`` (chemist), ``nelec`` and optionally ``ecore`` (no pickles)."""
    with np.load(path, allow_pickle=False) as z:
        for key in ("h", "eri", "nelec"):
            if key not in z:
                raise PyCCError("%s lacks array %r" % (path, key))
        h = np.asarray(z["h"], dtype=np.float64)
        eri = np.asarray(z["eri"], dtype=np.float64)
        nelec = int(z["nelec"])
        ecore = float(z["ecore"]) if "ecore" in z else 0.0
    norb = h.shape[0]
    if h.shape != (norb, norb) or eri.shape != (norb,) * 4:
        raise PyCCError("%s: h must be (norb, norb) and eri (norb,)*4" % path)
    return FCIDump(norb, nelec, 0, h, eri, ecore)


class MOHamiltonian:
    """Closed-shell MO Hamiltonian (the quantities pycc's ``Hamiltonian`` holds, hamiltonian.py:59-70).

    ``F = h + 2 J - K`` over the ``nocc = nelec / 2`` lowest orbitals, ``ERI[p,q,r,s] = <pq|rs> =
    (pr|qs)``, ``escf = ecore + sum_i (h_ii + F_ii)``. The orbitals must be ordered occupied first (as
    an RHF calculation writes them)."""

    def __init__(self, h, eri_chem, nelec, ecore=0.0):
        h = np.asarray(h, dtype=np.float64)
        eri_chem = np.asarray(eri_chem, dtype=np.float64)
        if nelec % 2:
            raise PyCCError("closed-shell CCSD needs an even number of electrons (NELEC=%d)" % nelec)
        self.nocc = nelec // 2
        self.nmo = h.shape[0]
        if self.nocc > self.nmo:
            raise PyCCError("NELEC=%d does not fit in %d orbitals" % (nelec, self.nmo))
        o = slice(0, self.nocc)
        J = np.einsum("pqii->pq", eri_chem[:, :, o, o])
        K = np.einsum("piiq->pq", eri_chem[:, o, o, :])
        self.h = h
        self.ecore = float(ecore)
        self.F = h + 2.0 * J - K
        self.ERI = np.ascontiguousarray(eri_chem.transpose(0, 2, 1, 3))
        self.escf = self.ecore + float(np.trace(h[o, o]) + np.trace(self.F[o, o]))

    @classmethod
    def from_fcidump(cls, path):
        d = read_fcidump(path)
        return cls(d.h, d.eri, d.nelec, d.ecore)

    @classmethod
    def from_npz(cls, path):
        d = load_npz(path)
        return cls(d.h, d.eri, d.nelec, d.ecore)

"""Integral files (pycc_amd.integrals): FCIDUMP / .npz readers and writers and the closed-shell MO
Hamiltonian built from them (the quantities pycc's Hamiltonian takes from psi4,
pycc/hamiltonian.py:59-70). CPU only.

Pins: byte-exact round trips of random 8-fold symmetric integrals; a hand-written FCIDUMP (Fortran D
exponents, '/' terminator, one representative per symmetry class); and water STO-3G, whose MO core
Hamiltonian and integrals come from the oracle's psi4-free front end: F = h + 2J - K from the file
equals the RHF Fock matrix in the MO basis and escf equals the RHF energy.
"""
import numpy as np
import pytest

from pycc_amd import PyCCError
from pycc_amd.integrals import MOHamiltonian, load_npz, read_fcidump, write_fcidump


def _sym8(rng, n):
    x = rng.standard_normal((n,) * 4)
    x = x + x.transpose(1, 0, 2, 3)
    x = x + x.transpose(0, 1, 3, 2)
    return x + x.transpose(2, 3, 0, 1)


def test_fcidump_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 6
    eri = _sym8(rng, n)
    h = rng.standard_normal((n, n))
    h = h + h.T
    path = tmp_path / "FCIDUMP"
    write_fcidump(path, h, eri, nelec=4, ecore=1.25)
    d = read_fcidump(path)
    assert (d.norb, d.nelec, d.ms2, d.ecore) == (n, 4, 0, 1.25)
    assert np.array_equal(d.h, h) and np.array_equal(d.eri, eri)
    assert d.orbsym == [1] * n


def test_fcidump_hand_written(tmp_path):
    text = """ &FCI NORB=  2,NELEC=2,MS2=0,
  ORBSYM=1,1,
  ISYM=1
 /
  0.5D+00   1   1   1   1
  0.25D+00  2   1   1   1
  0.125d0   2   1   2   1
 -1.0D+00   1   1   0   0
  0.1D+00   2   1   0   0
  0.75      0   0   0   0
"""
    path = tmp_path / "FCIDUMP"
    path.write_text(text)
    d = read_fcidump(path)
    assert d.norb == 2 and d.nelec == 2 and d.ecore == 0.75
    assert d.h[0, 0] == -1.0 and d.h[0, 1] == d.h[1, 0] == 0.1
    assert d.eri[0, 0, 0, 0] == 0.5
    for idx in [(1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1)]:
        assert d.eri[idx] == 0.25
    for idx in [(1, 0, 1, 0), (0, 1, 1, 0), (1, 0, 0, 1), (0, 1, 0, 1)]:
        assert d.eri[idx] == 0.125
    assert d.eri[1, 1, 0, 0] == 0.0


def test_fcidump_errors(tmp_path):
    p = tmp_path / "bad"
    p.write_text(" &FCI NORB=2,NELEC=2,MS2=2,\n &END\n 1.0 1 1 1 1\n")
    with pytest.raises(PyCCError):
        read_fcidump(p)
    p.write_text(" &FCI NORB=2,NELEC=2,MS2=0,\n &END\n 1.0 3 1 1 1\n")
    with pytest.raises(PyCCError):
        read_fcidump(p)
    p.write_text("not an fcidump\n")
    with pytest.raises(PyCCError):
        read_fcidump(p)
    with pytest.raises(PyCCError):
        MOHamiltonian(np.eye(3), np.zeros((3,) * 4), nelec=3)


def test_npz_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    n = 4
    eri = _sym8(rng, n)
    h = np.diag(np.arange(n, dtype=float))
    path = tmp_path / "ints.npz"
    np.savez(path, h=h, eri=eri, nelec=np.array(2), ecore=np.array(0.5))
    d = load_npz(path)
    assert np.array_equal(d.h, h) and np.array_equal(d.eri, eri) and d.nelec == 2 and d.ecore == 0.5


def _water_mo():
    from oracle.gaussian import ao_integrals, rhf, water_geometry
    S, T, V, eri, enuc = ao_integrals(water_geometry())
    E, C, eps, F_ao = rhf(S, T + V, eri, 5)
    h_mo = C.T @ (T + V) @ C
    eri_mo = np.einsum('pqrs,pi,qj,rk,sl->ijkl', eri, C, C, C, C, optimize=True)
    return h_mo, eri_mo, enuc, C.T @ F_ao @ C, E + enuc


def test_water_sto3g_hamiltonian_from_fcidump(tmp_path):
    h_mo, eri_mo, enuc, F_mo, escf = _water_mo()
    path = tmp_path / "FCIDUMP"
    write_fcidump(path, h_mo, eri_mo, nelec=10, ecore=enuc)
    H = MOHamiltonian.from_fcidump(path)
    assert H.nocc == 5 and H.nmo == 7
    assert np.abs(H.F - F_mo).max() < 1e-10
    assert abs(H.escf - escf) < 1e-10
    assert np.abs(H.ERI - eri_mo.transpose(0, 2, 1, 3)).max() < 1e-14

"""GPU CCSD from integral files: water STO-3G written to an FCIDUMP (MO integrals from the oracle's
psi4-free front end), read back through pycc_amd.integrals and solved on the GPU with one frozen core
orbital, reproduces the reference's CCSD correlation energy (pycc tests/test_002_ccsd_energy.py:31)."""
import numpy as np
import pytest

from conftest import reference_constant

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_water_sto3g_ccsd_from_fcidump(tmp_path):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycc_amd
    from pycc_amd.integrals import MOHamiltonian, write_fcidump
    from oracle.gaussian import ao_integrals, rhf, water_geometry
    print("water STO-3G from FCIDUMP", flush=True)
    S, T, V, eri, enuc = ao_integrals(water_geometry())
    E, C, eps, F_ao = rhf(S, T + V, eri, 5)
    h_mo = C.T @ (T + V) @ C
    eri_mo = np.einsum('pqrs,pi,qj,rk,sl->ijkl', eri, C, C, C, C, optimize=True)
    path = tmp_path / "FCIDUMP"
    write_fcidump(path, h_mo, eri_mo, nelec=10, ecore=enuc)
    H = MOHamiltonian.from_fcidump(path)
    cc = pycc_amd.CCwfn(H, nfzc=1, precision='DP', verbose=False)
    assert cc.no == 4
    e = cc.solve_cc(1e-12, 1e-12, 75)
    ref, tol = reference_constant("water_sto3g_frozen_core", "ccsd_correlation_energy")
    assert abs(e - ref) < tol, (e, ref)

"""GPU CCSD from integral files. Water STO-3G and cc-pVDZ (MO integrals from the oracle's psi4-free
front end) are written to FCIDUMP files, read back through pycc_amd.integrals and solved on the GPU.
The results reproduce the reference's constants: STO-3G frozen core (pycc
tests/test_002_ccsd_energy.py:31); cc-pVDZ frozen core (test_002_ccsd_energy.py:38,
test_003_ccsd_lambda.py:61); cc-pVDZ all electrons (test_030_sp.py:30,39), for the CCSD energy and
the Lambda pseudoenergy. The mixed modes are held to the 1 microhartree bar on cc-pVDZ."""
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


@pytest.fixture(scope="module")
def water_dz_fcidump(tmp_path_factory):
    from pycc_amd.integrals import write_fcidump
    from oracle.gaussian import water_hamiltonian
    F, ERI, escf, nocc, h_mo, eri_mo, enuc = water_hamiltonian("cc-pvdz")
    path = tmp_path_factory.mktemp("dz") / "FCIDUMP"
    write_fcidump(path, h_mo, eri_mo, nelec=2 * nocc, ecore=enuc)
    return path, escf


@pytest.mark.parametrize("nfzc,key", [(1, "water_ccpvdz_frozen_core"), (0, "water_ccpvdz_all_electron")])
def test_water_ccpvdz_ccsd_and_lambda_on_gpu(water_dz_fcidump, nfzc, key):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycc_amd
    from pycc_amd.integrals import MOHamiltonian
    path, escf = water_dz_fcidump
    print("water cc-pVDZ from FCIDUMP, nfzc", nfzc, flush=True)
    H = MOHamiltonian.from_fcidump(path)
    assert H.nmo == 24 and abs(H.escf - escf) < 1e-10
    cc = pycc_amd.CCwfn(H, nfzc=nfzc, precision='DP', verbose=False)
    assert cc.no == 5 - nfzc and cc.nv == 19
    e = cc.solve_cc(1e-12, 1e-12, 100)
    ref, tol = reference_constant(key, "ccsd_correlation_energy")
    assert abs(e - ref) < tol, (e, ref)
    lecc = pycc_amd.cclambda(cc, pycc_amd.cchbar(cc)).solve_lambda(1e-12, 1e-12, 100)
    ref, tol = reference_constant(key, "lambda_pseudoenergy")
    assert abs(lecc - ref) < tol, (lecc, ref)


@pytest.mark.parametrize("precision", ["MIXED", "MIXED_BF16X3", "MIXED_F16X3"])
def test_water_ccpvdz_mixed_within_one_microhartree(water_dz_fcidump, precision):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycc_amd
    from pycc_amd.integrals import MOHamiltonian
    path, _ = water_dz_fcidump
    print("water cc-pVDZ", precision, flush=True)
    cc = pycc_amd.CCwfn(MOHamiltonian.from_fcidump(path), nfzc=1, precision=precision, verbose=False)
    # thresholds above each mode's noise floor: the BF16x3 ladder (16-bit split products) moves the
    # energy by ~1e-8 Eh from one iteration to the next once converged
    e_conv, r_conv = {"MIXED": (1e-10, 1e-6), "MIXED_BF16X3": (1e-7, 1e-5), "MIXED_F16X3": (1e-9, 1e-6)}[precision]
    e = cc.solve_cc(e_conv, r_conv, 100)
    print("last iterations (n, E, dE, norm):", cc.history[-3:])
    assert e is not None
    ref, _ = reference_constant("water_ccpvdz_frozen_core", "ccsd_correlation_energy")
    print("%s energy error %.3e Eh" % (precision, e - ref))
    assert abs(e - ref) < 1e-6


"""GPU (T) correction (``ccx_triples``, ``pycc/cctriples.py:177-240``) against the oracle
(``oracle/triples_ref.py``, pinned by the reference's water (T) energies) and against those energies
themselves.

Bars: the reference's water (T) energies (``pycc/tests/test_005_ccsd_t_energy.py:33,44``) to 1e-11
Eh through a GPU CCSD(T) solve; the oracle's E(T) at arbitrary amplitudes and a non-diagonal Fock to
1e-12 relative, for dense and density-fitted contexts and odd and even sizes; E(T) from a MIXED
context equals the FP64 context's (its (T) path is FP64), and from a MIXED_F16X3 context (FP16x3
slot GEMMs) agrees to 1e-8 relative (2e-9 measured); a MIXED_F16X3 CCSD(T) solve of water cc-pVDZ
gives E(T) within 1e-8 Eh and the total within 1 uEh of the reference's constants."""
import numpy as np
import pytest

from conftest import random_amplitudes, reference_constant
from oracle.triples_ref import t_tjl
from pycc_amd.synth import synthetic_df, DenseHamiltonian, dirac_eri

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def water_dz_path(tmp_path_factory):
    from pycc_amd.integrals import write_fcidump
    from oracle.gaussian import water_hamiltonian
    _, _, _, nocc, h_mo, eri_mo, enuc = water_hamiltonian("cc-pvdz")
    path = tmp_path_factory.mktemp("dz_t") / "FCIDUMP"
    write_fcidump(path, h_mo, eri_mo, nelec=2 * nocc, ecore=enuc)
    return path


def _ham(no, nv, seed, offdiag=0.05):
    sd = synthetic_df(no, nv, seed=seed, t2max=0.1)
    H = DenseHamiltonian(sd.fock, dirac_eri(sd.B))
    X = np.random.default_rng(seed + 100).standard_normal(H.F.shape) * offdiag
    H.F = H.F + X + X.T
    H.eps = np.diagonal(H.F).copy()
    return sd, H


@pytest.mark.parametrize("basis,key", [("sto-3g", "water_sto3g_frozen_core"),
                                       ("cc-pvdz", "water_ccpvdz_frozen_core")])
def test_water_ccsd_t_matches_reference_constants(tmp_path, basis, key):
    _need_gpu()
    import pycc_amd
    from pycc_amd.integrals import MOHamiltonian, write_fcidump
    if basis == "sto-3g":
        from oracle.gaussian import ao_integrals, rhf, water_geometry
        S, T, V, eri, enuc = ao_integrals(water_geometry())
        _, C, _, _ = rhf(S, T + V, eri, 5)
        h_mo = C.T @ (T + V) @ C
        eri_mo = np.einsum('pqrs,pi,qj,rk,sl->ijkl', eri, C, C, C, C, optimize=True)
        nocc = 5
    else:
        from oracle.gaussian import water_hamiltonian
        _, _, _, nocc, h_mo, eri_mo, enuc = water_hamiltonian("cc-pvdz")
    path = tmp_path / "FCIDUMP"
    write_fcidump(path, h_mo, eri_mo, nelec=2 * nocc, ecore=enuc)
    print("water", basis, "CCSD(T) from FCIDUMP", flush=True)
    cc = pycc_amd.CCwfn(MOHamiltonian.from_fcidump(path), nfzc=1, model='CCSD(T)', precision='DP',
                        verbose=False)
    e = cc.solve_cc(1e-12, 1e-12, 100)
    eccsd, tol = reference_constant(key, "ccsd_correlation_energy")
    et_ref, tol_t = reference_constant(key, "triples_correction")
    assert abs(cc.et - et_ref) < tol_t, (cc.et, et_ref)
    assert abs(e - (eccsd + et_ref)) < tol + tol_t, (e, eccsd + et_ref)
    assert pycc_amd.cctriples.t_tjl(cc) == cc.et


@pytest.mark.parametrize("no,nv,seed", [(1, 5, 1), (2, 3, 2), (3, 17, 3), (5, 24, 4), (6, 33, 5), (10, 12, 6)])
def test_triples_matches_oracle(no, nv, seed):
    _need_gpu()
    import pycc_amd
    print("triples dense no=%d nv=%d seed=%d" % (no, nv, seed), flush=True)
    _, H = _ham(no, nv, seed)
    t1, t2 = random_amplitudes(no, nv, seed, scale=0.1)
    cc = pycc_amd.CCwfn(H, no=no, precision='DP', verbose=False)
    et = cc.triples_energy(t1, t2)
    ref = t_tjl(H.F, H.ERI, t1, t2, no)
    assert abs(et - ref) < 1e-12 * max(abs(ref), 1e-3), (et, ref)


# o >= 9: several k chunks per (i, j); v = 50: 4^3 tiles of 16 (orbits of one, three and six tiles,
# the last tile partly past v); v = 3, 2: one partial tile, rows shorter than a 4-float vector;
# v = 8, 24, 48: whole 8^3 tiles, Y in the GEMM's 8 x 8-block layout (the path of v a multiple of
# 8; one tile, then orbits of one, three and six tiles); v = 9, 17: a last tile of one element
@pytest.mark.parametrize("no,nv,seed", [(4, 21, 7), (9, 13, 8), (6, 50, 10), (2, 3, 11), (3, 2, 12),
                                        (4, 48, 13), (3, 8, 14), (10, 24, 15), (5, 9, 16), (3, 17, 17)])
def test_triples_density_fitted_and_mixed_contexts(no, nv, seed):
    _need_gpu()
    import pycc_amd
    print("triples DF / mixed no=%d nv=%d" % (no, nv), flush=True)
    sd = synthetic_df(no, nv, seed=seed, t2max=0.1)
    H = DenseHamiltonian(sd.fock, dirac_eri(sd.B))
    t1, t2 = random_amplitudes(no, nv, seed, scale=0.1)
    ref = t_tjl(H.F, H.ERI, t1, t2, no)
    df = pycc_amd.CCwfn(sd, no=no, precision='DP', verbose=False).triples_energy(t1, t2)
    assert abs(df - ref) < 1e-12 * abs(ref), (df, ref)
    dp = pycc_amd.CCwfn(H, no=no, precision='DP', verbose=False).triples_energy(t1, t2)
    mx = pycc_amd.CCwfn(H, no=no, precision="MIXED", verbose=False).triples_energy(t1, t2)
    assert abs(mx - dp) < 1e-14 * abs(dp), (mx, dp)
    x3 = pycc_amd.CCwfn(H, no=no, precision="MIXED_F16X3", verbose=False).triples_energy(t1, t2)
    # FP16x3 products carry ~2^-22 relative error each; over a few dozen terms (v < 8) it does not
    # average out (3e-8 to 5e-8 measured at v = 2, 3, before and after the K pairing)
    assert abs(x3 - dp) < (1e-7 if nv < 8 else 1e-8) * abs(dp), (x3, dp)


@pytest.mark.parametrize("precision", ["DP", "MIXED_F16X3"])
def test_triples_in_ranges_of_i_equal_one_call(precision):
    _need_gpu()
    import pycc_amd
    no, nv, seed = 9, 13, 9
    print("triples by ranges of i, %s no=%d nv=%d" % (precision, no, nv), flush=True)
    _, H = _ham(no, nv, seed)
    t1, t2 = random_amplitudes(no, nv, seed, scale=0.1)
    cc = pycc_amd.CCwfn(H, no=no, precision=precision, verbose=False)
    whole = cc.triples_energy(t1, t2)
    seen = []
    parts = cc.triples_energy(t1, t2, i_chunk=4, progress=lambda d, n: seen.append((d, n)))
    assert parts == whole, (parts, whole)
    assert seen == [(4, 9), (8, 9), (9, 9)]


def test_water_ccsd_t_mixed_f16x3_within_one_microhartree(water_dz_path):
    _need_gpu()
    import pycc_amd
    from pycc_amd.integrals import MOHamiltonian
    print("water cc-pVDZ CCSD(T) MIXED_F16X3", flush=True)
    cc = pycc_amd.CCwfn(MOHamiltonian.from_fcidump(water_dz_path), nfzc=1, model='CCSD(T)',
                        precision='MIXED_F16X3', verbose=False)
    e = cc.solve_cc(1e-9, 1e-6, 100)
    eccsd, _ = reference_constant("water_ccpvdz_frozen_core", "ccsd_correlation_energy")
    et_ref, _ = reference_constant("water_ccpvdz_frozen_core", "triples_correction")
    print("E(T) error %.3e Eh, total error %.3e Eh" % (cc.et - et_ref, e - eccsd - et_ref))
    assert abs(cc.et - et_ref) < 1e-8
    assert abs(e - (eccsd + et_ref)) < 1e-6

"""GPU parity of the HIP CCSD path against the CPU oracle (FP64: bit-level agreement bar).

Bars (from the north star): plain FP64 reproduces the oracle to 1e-10 Eh in the energy and 1e-8
max-abs in T1/T2; residual elements agree to 1e-11 absolute here (sizes are small).
"""
import math

import numpy as np
import pytest

from conftest import random_amplitudes, reference_constant
from oracle.ccsd_ref import OracleCCSD
from pycc_amd.synth import synthetic_df, DenseHamiltonian, dirac_eri

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _ham(no, nv, seed, t2max=0.05, offdiag=0.0):
    sd = synthetic_df(no, nv, seed=seed, t2max=t2max)
    H = DenseHamiltonian(sd.fock, dirac_eri(sd.B))
    if offdiag:
        X = np.random.default_rng(seed + 100).standard_normal(H.F.shape) * offdiag
        H.F = H.F + X + X.T
    return sd, H


@pytest.fixture(scope="module")
def pycc_amd():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycc_amd
    return pycc_amd


# ---- the contraction engine (ContractionBackend / ccx_contract) ------------------------------
CASES = [
    ("ij,jk->ik", [(37, 53), (53, 29)]),
    ("ijab,mb->ijam", [(3, 4, 17, 19), (6, 19)]),
    ("jnfb,mnef->mbej", [(4, 5, 9, 11), (4, 5, 13, 9)]),
    ("mbef,ijef->mbij", [(3, 7, 8, 8), (4, 5, 8, 8)]),
    ("imae,mbej->ijab", [(5, 4, 6, 7), (4, 9, 7, 6)]),
    ("ie,abej->ijab", [(5, 7), (6, 8, 7, 5)]),
    ("mnab,mnij->ijab", [(4, 4, 9, 9), (4, 4, 5, 5)]),
    ("ijab,ijab->ab", [(3, 4, 5, 6), (3, 4, 5, 6)]),
    ("ia,jb->ijab", [(4, 6), (5, 7)]),
    ("pq,qr->pr", [(300, 700), (700, 260)]),          # several tiles, split-K
    ("k,kn->n", [(5000,), (5000, 3)]),                # long K, tiny output
]


@pytest.mark.parametrize("sub,shapes", CASES)
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_contract_engine_matches_numpy(pycc_amd, sub, shapes, dtype):
    rng = np.random.default_rng(len(sub))
    A = rng.standard_normal(shapes[0])
    B = rng.standard_normal(shapes[1])
    ref = np.einsum(sub, A, B)
    dt = getattr(torch, dtype)
    mgr = pycc_amd.DeviceManager('GPU', 'DP')
    out = mgr.contract(sub, torch.tensor(A, dtype=dt, device="cuda"), torch.tensor(B, dtype=dt, device="cuda"))
    tol = 1e-12 if dtype == "float64" else 2e-5
    scale = max(1.0, float(np.abs(ref).max()))
    assert np.abs(out.cpu().double().numpy() - ref).max() < tol * scale * math.sqrt(A.size / max(1, ref.size) + 1)


def test_contract_engine_strided_views_and_beta(pycc_amd):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 9, 5))
    B = rng.standard_normal((5, 9, 4))
    C0 = rng.standard_normal((4, 6))
    At = torch.tensor(A, device="cuda").permute(2, 0, 1)       # non-contiguous view, labels 'kij'
    Bt = torch.tensor(B, device="cuda")
    Ct = torch.tensor(C0, device="cuda")
    mgr = pycc_amd.DeviceManager('GPU', 'DP')
    mgr.contract("kij,kjl->li", At, Bt, alpha=0.5, out=Ct, beta=-2.0)
    ref = 0.5 * np.einsum("ijk,kjl->li", A, B) - 2.0 * C0
    assert np.abs(Ct.cpu().numpy() - ref).max() < 1e-12


# ---- the residual --------------------------------------------------------------------------------
@pytest.mark.parametrize("no,nv,seed,offdiag", [(1, 5, 1, 0.05), (2, 3, 2, 0.05), (4, 12, 3, 0.05),
                                                (5, 19, 4, 0.0), (3, 33, 5, 0.02)])
def test_residual_dp_matches_oracle(pycc_amd, no, nv, seed, offdiag):
    sd, H = _ham(no, nv, seed, t2max=0.1, offdiag=offdiag)
    ref = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, seed, scale=0.05)
    r1_ref, r2_ref = ref.residuals(H.F, t1, t2)
    cc = pycc_amd.CCwfn(H, no=no, precision='DP', verbose=False)
    r1, r2 = cc.residuals(H.F, t1, t2)
    print("case no=%d nv=%d: max|r1-ref| %.2e  max|r2-ref| %.2e" % (
        no, nv, np.abs(r1.cpu().numpy() - r1_ref).max(), np.abs(r2.cpu().numpy() - r2_ref).max()))
    assert np.abs(r1.cpu().numpy() - r1_ref).max() < 1e-11
    assert np.abs(r2.cpu().numpy() - r2_ref).max() < 1e-11
    # energy functional at the same amplitudes (ccwfn.py:1160-1161)
    assert abs(cc.cc_energy(t1=t1, t2=t2) - ref.energy(t1, t2)) < 1e-12


@pytest.mark.parametrize("no,nv,seed", [(1, 5, 1), (3, 1, 2), (2, 3, 3)])
def test_f16x3_residual_edge_shapes(pycc_amd, no, nv, seed):
    """MIXED_F16X3 with empty packed parts: o = 1 has no (i < j) pairs (no tau- rows), v = 1 no (e < f)
    pairs (no LM columns but the Z ones). Bar: the FP16x3 residual tolerance, relative to max|r|."""
    print("f16x3 edge", no, nv, flush=True)
    sd, H = _ham(no, nv, seed, t2max=0.1, offdiag=0.05)
    ref = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, seed, scale=0.05)
    r1_ref, r2_ref = ref.residuals(H.F, t1, t2)
    cc = pycc_amd.CCwfn(H, no=no, precision='MIXED_F16X3', verbose=False)
    r1, r2 = cc.residuals(H.F, t1, t2)
    rel1 = np.abs(r1.cpu().numpy() - r1_ref).max() / np.abs(r1_ref).max()
    rel2 = np.abs(r2.cpu().numpy() - r2_ref).max() / np.abs(r2_ref).max()
    print("rel err r1 %.2e r2 %.2e" % (rel1, rel2))
    assert rel1 < 3e-5 and rel2 < 3e-5


def test_df_integrals_match_dense(pycc_amd):
    no, nv = 4, 21
    sd, H = _ham(no, nv, 7, t2max=0.1)
    t1, t2 = random_amplitudes(no, nv, 7, scale=0.05)
    dense = pycc_amd.CCwfn(H, no=no, verbose=False)
    df = pycc_amd.CCwfn(sd, no=no, verbose=False)
    a1, a2 = dense.residuals(H.F, t1, t2)
    b1, b2 = df.residuals(H.F, t1, t2)
    assert (a1 - b1).abs().max().item() < 1e-12
    assert (a2 - b2).abs().max().item() < 1e-12


@pytest.mark.parametrize("no,nv,seed", [(5, 19, 0), (8, 40, 1)])
def test_solve_cc_dp_matches_oracle(pycc_amd, no, nv, seed):
    sd, H = _ham(no, nv, seed, t2max=0.03)
    ref = OracleCCSD(H, no)
    e_ref = ref.solve_cc(1e-10, 1e-10, 100, 8, 1)
    cc = pycc_amd.CCwfn(H, no=no, precision='DP', verbose=False)
    e = cc.solve_cc(1e-10, 1e-10, 100, 8, 1)
    assert e is not None and e_ref is not None
    print("solve no=%d nv=%d: E %.14f oracle %.14f (iters %d/%d)" % (no, nv, e, e_ref, cc.niter, ref.niter))
    assert abs(e - e_ref) < 1e-10
    assert cc.niter == ref.niter
    assert np.abs(cc.t1.cpu().numpy() - ref.t1).max() < 1e-8
    assert np.abs(cc.t2.cpu().numpy() - ref.t2).max() < 1e-8


def test_solve_cc_dp_matches_oracle_o16_v128(pycc_amd):
    """The FP64 benchmark configuration o=16, v=128 (nmo = 144, seeded synthetic DF Hamiltonian with
    max |MP2 t2| = 0.03): the GPU solve (from the DF factors) against the oracle on the same dense
    integrals, to 1e-10 Eh in the energy and 1e-8 max-abs in T1 and T2, in the same number of
    iterations. The oracle needs ~30-40 s of host time here; it prints each iteration."""
    no, nv = 16, 128
    print("o=16 v=128 FP64 solve vs oracle", flush=True)
    sd = synthetic_df(no, nv, seed=0)
    cc = pycc_amd.CCwfn(sd, no=no, precision='DP', verbose=False)
    e = cc.solve_cc(1e-10, 1e-10, 100, 8, 1)
    print("GPU: E %.14f in %d iterations" % (e, cc.niter), flush=True)
    H = DenseHamiltonian(sd.fock, dirac_eri(sd.B))
    ref = OracleCCSD(H, no)
    e_ref = ref.solve_cc(1e-10, 1e-10, 100, 8, 1, verbose=True)
    assert e is not None and e_ref is not None
    d1 = np.abs(cc.t1.cpu().numpy() - ref.t1).max()
    d2 = np.abs(cc.t2.cpu().numpy() - ref.t2).max()
    print("E %.14f oracle %.14f, iterations %d / %d, max|dT1| %.2e max|dT2| %.2e" % (
        e, e_ref, cc.niter, ref.niter, d1, d2))
    assert abs(e - e_ref) < 1e-10
    assert cc.niter == ref.niter
    assert d1 < 1e-8 and d2 < 1e-8


def test_water_sto3g_reference_constant_on_gpu(pycc_amd):
    from oracle.gaussian import water_sto3g_hamiltonian
    F, ERI, escf, nocc = water_sto3g_hamiltonian()
    H = DenseHamiltonian(F, ERI)
    cc = pycc_amd.CCwfn(H, no=nocc - 1, nfzc=1, verbose=False)
    e = cc.solve_cc(1e-12, 1e-12, 75)
    ref, tol = reference_constant("water_sto3g_frozen_core", "ccsd_correlation_energy")
    assert abs(e - ref) < tol      # pycc tests/test_002_ccsd_energy.py:31


def test_solve_returns_none_when_not_converged(pycc_amd):
    sd, H = _ham(2, 6, 3)
    cc = pycc_amd.CCwfn(H, no=2, verbose=False)
    assert cc.solve_cc(1e-14, 1e-14, maxiter=2) is None


# Low-precision contractions leave a noise floor in the Jacobi step norm, so the mixed solves use
# r_conv above that floor; the energy bar is the north star's 1 microhartree.
@pytest.mark.parametrize("precision,tol_r,r_conv", [("MIXED", 1e-5, 1e-7), ("MIXED_BF16X3", 3e-4, 1e-6),
                                                     ("MIXED_F16X3", 3e-5, 1e-7)])
def test_mixed_precision_residual_and_energy(pycc_amd, precision, tol_r, r_conv):
    no, nv = 6, 30
    sd, H = _ham(no, nv, 11, t2max=0.03)
    ref = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, 11, scale=0.02)
    r1_ref, r2_ref = ref.residuals(H.F, t1, t2)
    cc = pycc_amd.CCwfn(H, no=no, precision=precision, verbose=False)
    r1, r2 = cc.residuals(H.F, t1, t2)
    rel2 = np.abs(r2.cpu().numpy() - r2_ref).max() / np.abs(r2_ref).max()
    rel1 = np.abs(r1.cpu().numpy() - r1_ref).max() / np.abs(r1_ref).max()
    print("%s residual rel err r1 %.2e r2 %.2e" % (precision, rel1, rel2))
    assert rel2 < tol_r and rel1 < tol_r
    ref = OracleCCSD(H, no)
    e_ref = ref.solve_cc(1e-10, 1e-10)
    e = cc.solve_cc(1e-9, r_conv)
    assert e is not None
    print("%s energy error %.3e Eh, max|dT1| %.2e max|dT2| %.2e, iterations %d (FP64 oracle %d)" % (
        precision, e - e_ref, np.abs(cc.t1.cpu().numpy() - ref.t1).max(),
        np.abs(cc.t2.cpu().numpy() - ref.t2).max(), cc.niter, ref.niter))
    assert abs(e - e_ref) < 1e-6


@pytest.mark.parametrize("precision,tol_r", [("MIXED", 1e-5), ("MIXED_BF16X3", 3e-4), ("MIXED_F16X3", 3e-5)])
def test_mixed_residual_large_pair_count(pycc_amd, precision, tol_r):
    """o = 24 gives 300 occupied pairs: the ladder takes the 256 x 256 BF16x3 tiles (with partial edge
    tiles in M and N) and the hole-hole terms the FP32 split-K path."""
    no, nv = 24, 40
    print("large pair count", precision, no, nv, flush=True)
    sd, H = _ham(no, nv, 17, t2max=0.03)
    ref = OracleCCSD(H, no)
    t1, t2 = random_amplitudes(no, nv, 17, scale=0.02)
    r1_ref, r2_ref = ref.residuals(H.F, t1, t2)
    cc = pycc_amd.CCwfn(H, no=no, precision=precision, verbose=False)
    r1, r2 = cc.residuals(H.F, t1, t2)
    rel2 = np.abs(r2.cpu().numpy() - r2_ref).max() / np.abs(r2_ref).max()
    rel1 = np.abs(r1.cpu().numpy() - r1_ref).max() / np.abs(r1_ref).max()
    print("%s residual rel err r1 %.2e r2 %.2e" % (precision, rel1, rel2))
    assert rel2 < tol_r and rel1 < tol_r


def test_f16x3_residual_matches_dp_on_a_full_grid(pycc_amd):
    """o = 24, v = 256: the :930 hole-hole split GEMM has 3 x 256 = 768 tiles, so it adds into the FP64
    Hh straight from the kernel (no split-K partials); the ring, ladder and W paths run at a size where
    every split GEMM takes the 256 x 256 LDS-DMA kernel. The FP64 residual of the same library (pinned
    to the oracle at smaller sizes) is the reference; the bar is the FP16x3 residual tolerance."""
    no, nv = 24, 256
    print("f16x3 vs dp residual", no, nv, flush=True)
    sd = synthetic_df(no, nv, seed=21)
    t1, t2 = random_amplitudes(no, nv, 21, scale=0.02)
    out = {}
    for precision in ("DP", "MIXED_F16X3"):
        cc = pycc_amd.CCwfn(sd, no=no, precision=precision, verbose=False)
        r1, r2 = cc.residuals(cc.H.F, t1, t2)
        out[precision] = (r1.cpu().numpy(), r2.cpu().numpy())
        del cc, r1, r2
        torch.cuda.empty_cache()
    (a1, a2), (b1, b2) = out["DP"], out["MIXED_F16X3"]
    rel1 = np.abs(b1 - a1).max() / np.abs(a1).max()
    rel2 = np.abs(b2 - a2).max() / np.abs(a2).max()
    print("residual rel err r1 %.2e r2 %.2e" % (rel1, rel2))
    assert np.isfinite(b2).all() and rel1 < 3e-5 and rel2 < 3e-5


def test_contract_f32_large_tiles(pycc_amd):
    """8200 x 8200 output: the FP32 plain GEMM takes its 256 x 256 tiles (1089 tiles, partial edge
    tiles in M and N), with a K that is not a multiple of the 32-wide k tile."""
    print("f32 256-tile GEMM", flush=True)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((8200, 72))
    B = rng.standard_normal((72, 8200))
    mgr = pycc_amd.DeviceManager('GPU', 'DP')
    for sub, a, b in (("ij,jk->ik", A, B), ("ji,jk->ik", A.T.copy(), B)):
        out = mgr.contract(sub, torch.tensor(a, dtype=torch.float32, device="cuda"),
                           torch.tensor(b, dtype=torch.float32, device="cuda"))
        ref = np.einsum(sub, a, b)
        assert np.abs(out.cpu().double().numpy() - ref).max() < 2e-4 * np.abs(ref).max()

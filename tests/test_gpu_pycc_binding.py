"""The pycc-side binding (``pycc_amd/ccx_backend.py``, the file a pycc maintainer copies into pycc)
driven the way pycc's own ``solve_cc`` drives ``residuals`` and ``cc_energy`` (``ccwfn.py:268-295``):
in 'DP' with float64 tensors (residuals equal the oracle's to 1e-11) and in 'SP' with the float32
tensors pycc holds in that mode (``device.py:147-151``): the binding converts them, the residuals
come back as float32, and pycc's float32 Jacobi loop converges to the FP64 energy within 1e-6 Eh."""
import numpy as np
import pytest

from conftest import random_amplitudes
from oracle.ccsd_ref import OracleCCSD
from pycc_amd.synth import synthetic_df, DenseHamiltonian, dirac_eri

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pycc_amd import ccx_backend, _lib
    ccx_backend.load(_lib.LIB_PATH)
    return ccx_backend


def _system(no=4, nv=14, seed=7):
    sd = synthetic_df(no, nv, seed=seed, t2max=0.05)
    return DenseHamiltonian(sd.fock, dirac_eri(sd.B))


def test_binding_dp_residual_equals_oracle(backend):
    no, nv = 4, 14
    H = _system(no, nv)
    ctx = backend.CCXContext(no, nv, "DP", H.F, H.ERI)
    t1, t2 = random_amplitudes(no, nv, 3)
    F = torch.tensor(H.F, device="cuda")
    r1, r2 = ctx.residuals(F, torch.tensor(t1, device="cuda"), torch.tensor(t2, device="cuda"))
    assert r1.dtype == torch.float64 and r2.dtype == torch.float64
    q1, q2 = OracleCCSD(H, no).residuals(H.F, t1, t2)
    assert np.abs(r1.cpu().numpy() - q1).max() < 1e-11
    assert np.abs(r2.cpu().numpy() - q2).max() < 1e-11
    e = ctx.energy(F, torch.tensor(t1, device="cuda"), torch.tensor(t2, device="cuda"))
    assert abs(float(e) - OracleCCSD(H, no).energy(t1, t2)) < 1e-12


def test_binding_sp_takes_float32_tensors_through_pycc_loop(backend):
    no, nv = 4, 14
    H = _system(no, nv)
    e64 = OracleCCSD(H, no).solve_cc(1e-11, 1e-11, 200)
    ctx = backend.CCXContext(no, nv, "SP", H.F, H.ERI)
    f32 = torch.float32
    F = torch.tensor(H.F, dtype=f32, device="cuda")              # pycc's real_dtype in 'SP'
    eps = torch.diagonal(F)
    Dia = eps[:no].reshape(-1, 1) - eps[no:]
    Dijab = eps[:no].reshape(-1, 1, 1, 1) + eps[:no].reshape(-1, 1, 1) - eps[no:].reshape(-1, 1) - eps[no:]
    t1 = torch.zeros((no, nv), dtype=f32, device="cuda")
    t2 = torch.tensor(H.ERI[:no, :no, no:, no:], dtype=f32, device="cuda") / Dijab      # ccwfn.py:210-211
    ecc = float(ctx.energy(F, t1, t2))
    for niter in range(1, 200):                                   # ccwfn.py:268-295, without DIIS
        last = ecc
        r1, r2 = ctx.residuals(F, t1, t2)
        assert r1.dtype == f32 and r2.dtype == f32
        t1 = t1 + r1 / Dia
        t2 = t2 + r2 / Dijab
        ecc = float(ctx.energy(F, t1, t2))
        if abs(ecc - last) < 1e-9:
            break
    print("SP binding: E %.10f vs FP64 %.10f after %d iterations" % (ecc, e64, niter))
    assert abs(ecc - e64) < 1e-6

"""GPU tests of the FP16x3 split GEMM (ccx_gemm_f16x3, the GEMM of the MIXED_F16X3 mode) against
FP64 products of the same FP32 inputs.

Each operand is carried as hi + lo FP16 with a power-of-two scale (22 significant bits) and
accumulated in FP32, so the bar is set relative to |A| |B|: max |C - C_ref| <= 1e-6 max (|A| |B|),
about four FP32 roundings of the largest term. Shapes cover both kernels (M < 256: 128 x 128
tiles with register staging, down to M = 1; M >= 256: 256 x 256 tiles fed by LDS DMA, 256 x 320
from N = 8192, with ragged last tiles),
split K (few tiles, long K),
ragged M / N edges, K not a multiple of 32, transposed operands, alpha / beta and operands many
decades apart in magnitude.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mgr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycc_amd
    return pycc_amd.DeviceManager('GPU', 'MIXED_F16X3')


def _check(mgr, A, B, alpha=1.0, beta=0.0, C0=None, ta=False, tb=False):
    Ad = torch.tensor(A.T.copy() if ta else A, dtype=torch.float32, device="cuda")
    Bd = torch.tensor(B.T.copy() if tb else B, dtype=torch.float32, device="cuda")
    Ad = Ad.T if ta else Ad
    Bd = Bd.T if tb else Bd
    out = None if C0 is None else torch.tensor(C0, dtype=torch.float32, device="cuda")
    C = mgr.contract.matmul_f16x3(Ad, Bd, out=out, alpha=alpha, beta=beta).cpu().double().numpy()
    A32 = A.astype(np.float32).astype(np.float64)
    B32 = B.astype(np.float32).astype(np.float64)
    ref = alpha * (A32 @ B32)
    if C0 is not None:
        ref = ref + beta * C0.astype(np.float32).astype(np.float64)
    scale = np.abs(alpha) * (np.abs(A32) @ np.abs(B32)).max()
    err = np.abs(C - ref).max() / scale
    return err


@pytest.mark.parametrize("M,N,K", [(100, 70, 33), (300, 517, 1000), (257, 1030, 2048), (512, 256, 64),
                                   (300, 260, 50000), (100, 90, 40000), (48, 3000, 384), (64, 70, 33),
                                   (1, 5, 40), (48, 384, 100000), (300, 9000, 200), (512, 8200, 64),
                                   (260, 16500, 1000)])
def test_split_gemm_shapes(mgr, M, N, K):
    print("f16x3 gemm", M, N, K, flush=True)
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    err = _check(mgr, A, B)
    print("max rel err %.3e" % err)
    assert err < 1e-6


@pytest.mark.parametrize("ta,tb", [(True, False), (False, True), (True, True)])
def test_split_gemm_transposes(mgr, ta, tb):
    print("f16x3 gemm transposes", ta, tb, flush=True)
    rng = np.random.default_rng(7)
    A = rng.standard_normal((300, 170))
    B = rng.standard_normal((170, 290))
    assert _check(mgr, A, B, ta=ta, tb=tb) < 1e-6


def test_split_gemm_scales_alpha_beta(mgr):
    """Operands near 1e-7 and 1e5 (power-of-two scaling keeps them in FP16 range), alpha and beta."""
    print("f16x3 gemm scales", flush=True)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((280, 400)) * 1e-7
    B = rng.standard_normal((400, 300)) * 1e5
    C0 = rng.standard_normal((280, 300)) * 1e-3
    assert _check(mgr, A, B, alpha=-0.5, beta=2.0, C0=C0) < 1e-6
    # a wide dynamic range inside one operand: small entries keep their absolute accuracy
    A2 = rng.standard_normal((260, 300)) * np.logspace(-12, 0, 300)[None, :]
    B2 = rng.standard_normal((300, 260))
    assert _check(mgr, A2, B2) < 1e-6


def test_split_gemm_empty_k(mgr):
    """K = 0: C = beta C."""
    C0 = np.arange(12.0).reshape(3, 4)
    A = np.zeros((3, 0))
    B = np.zeros((0, 4))
    Ad = torch.zeros((3, 0), dtype=torch.float32, device="cuda")
    Bd = torch.zeros((0, 4), dtype=torch.float32, device="cuda")
    out = torch.tensor(C0, dtype=torch.float32, device="cuda")
    C = mgr.contract.matmul_f16x3(Ad, Bd, out=out, beta=0.5).cpu().numpy()
    assert np.array_equal(C, 0.5 * C0)
    assert A.size == 0 and B.size == 0

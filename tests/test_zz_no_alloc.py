"""ccx_contract_noalloc against ccx_contract, through the C ABI.

ccx_contract_noalloc runs the same contraction planner as ccx_contract (ContractionBackend.__call__,
pycc/device.py:64-86), with its split-K partial sums in the library's fixed per-GPU device array
instead of an allocated workspace. Each case makes two calls in a row, the second with a smaller
input, and compares them with ccx_contract on the same inputs:

* thin outputs with a long K, whose split plan fits the fixed array: the same partial sums in the
  same order, so the results are bit-identical;
* a 1024 x 1024 output whose split plan (16 partial matrices, 128 MiB) exceeds the 64 MiB array:
  the planner takes fewer splits, so the results agree to rounding (1e-12 relative).
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pycc_amd import _lib
    return _lib.load()


def _i64(vals):
    return (ctypes.c_int64 * len(vals))(*vals)


def _run(lib, name, sub, A, B, C, alpha, beta):
    rc = getattr(lib, name)(sub.encode(), alpha,
                            ctypes.c_void_p(A.data_ptr()), _i64(A.shape), _i64(A.stride()),
                            ctypes.c_void_p(B.data_ptr()), _i64(B.shape), _i64(B.stride()),
                            beta, ctypes.c_void_p(C.data_ptr()), _i64(C.shape), _i64(C.stride()), None)
    assert rc == 0, (name, lib.ccx_last_error())


def _both(lib, sub, shape_a, shape_b, shape_c, seed, beta=0.0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn(shape_a, dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn(shape_b, dtype=torch.float64, device="cuda", generator=g)
    C0 = torch.randn(shape_c, dtype=torch.float64, device="cuda", generator=g)
    ref, out = C0.clone(), C0.clone()
    _run(lib, "ccx_contract", sub, A, B, ref, 0.75, beta)
    _run(lib, "ccx_contract_noalloc", sub, A, B, out, 0.75, beta)
    torch.cuda.synchronize()
    exact = torch.einsum(sub, A, B) * 0.75 + beta * C0
    return ref, out, exact


@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_noalloc_split_plan_that_fits_is_bit_identical(lib, beta):
    # 48 x 48 output, K = 200000: one 64 x 64 tile, 256 K splits (4.7 MB of partials); then smaller
    for sub, sa, sb, sc, seed in (("ik,jk->ij", (48, 200000), (48, 200000), (48, 48), 1),
                                  ("ik,jk->ij", (40, 50000), (40, 50000), (40, 40), 2)):
        print("ccx_contract_noalloc %s A%s B%s beta %.1f" % (sub, sa, sb, beta), flush=True)
        ref, out, exact = _both(lib, sub, sa, sb, sc, seed, beta)
        assert torch.equal(ref, out)
        assert float((out - exact).abs().max()) < 1e-10 * max(1.0, float(exact.abs().max()))


def test_noalloc_split_plan_beyond_the_array_takes_fewer_splits(lib):
    # 1024 x 1024 output, K = 40960: 16 splits would need 128 MiB; the fixed array holds 64 MiB
    for sub, sa, sb, sc, seed in (("ik,kj->ij", (1024, 40960), (40960, 1024), (1024, 1024), 3),
                                  ("ik,kj->ij", (512, 8192), (8192, 768), (512, 768), 4)):
        print("ccx_contract_noalloc %s A%s B%s" % (sub, sa, sb), flush=True)
        ref, out, exact = _both(lib, sub, sa, sb, sc, seed, 0.5)
        scale = max(1.0, float(exact.abs().max()))
        assert float((out - ref).abs().max()) < 1e-12 * scale
        assert float((out - exact).abs().max()) < 1e-10 * scale


def test_noalloc_rejects_bad_arguments_like_contract(lib):
    assert lib.ccx_contract_noalloc(b"ab,bc", 1.0, None, None, None, None, None, None, 0.0, None, None, None,
                                    None) == -1

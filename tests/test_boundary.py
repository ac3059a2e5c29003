"""The C-ABI boundary and the host-side API, without computing on a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ccx.h")
LIB = os.path.join(ROOT, "pycc_amd", "libccx.so")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\s*\*)\s+(ccx_\w+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points():
    names = _declared()
    for n in ["ccx_create", "ccx_destroy", "ccx_set_integrals_dense", "ccx_set_integrals_df", "ccx_residual",
              "ccx_energy", "ccx_update", "ccx_diis_reset", "ccx_diis_add", "ccx_diis_extrapolate",
              "ccx_contract", "ccx_triples", "ccx_last_error"]:
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_loads_and_exports_every_declared_symbol():
    import torch  # noqa: F401  (the library binds torch's HIP runtime)
    lib = ctypes.CDLL(LIB)
    for name in _declared():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_binding_signatures_cover_the_header():
    from pycc_amd import _lib
    assert set(_declared()) == set(_lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_invalid_arguments_return_error_codes_without_a_gpu():
    from pycc_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.ccx_create(0, 5, 0, ctypes.byref(h)) == -1           # CCX_ERR_ARG
    assert b"no >= 1" in lib.ccx_last_error()
    assert lib.ccx_create(2, 5, 7, ctypes.byref(h)) == -1
    assert lib.ccx_residual(None, None, None, None, None, None, None) == -1
    assert lib.ccx_triples(None, None, None, None, None, None) == -1
    assert lib.ccx_triples_range(None, None, None, None, 0, 1, None, None) == -1
    assert b"null context" in lib.ccx_last_error()
    assert lib.ccx_contract(b"ab,bc", 1.0, None, None, None, None, None, None, 0.0, None, None, None, None) == -1
    assert lib.ccx_t3_denominator(None, 5, 0, 0, 0, 2, 3, 1, None, None) == -1
    assert b"null fock" in lib.ccx_last_error()
    fake = ctypes.c_void_p(16)                                          # never dereferenced: rejected first
    assert lib.ccx_t3_denominator(fake, 5, 0, 0, 5, 2, 3, 1, fake, None) == -1
    assert b"fixed orbital outside" in lib.ccx_last_error()
    assert lib.ccx_t3_denominator(fake, 5, 0, 1, 1, 3, 3, 1, fake, None) == -1
    assert b"running orbitals outside" in lib.ccx_last_error()


def test_device_manager_validates_keywords():
    from pycc_amd import DeviceManager, InvalidKeywordError
    with pytest.raises(InvalidKeywordError):
        DeviceManager(device='TPU')
    with pytest.raises(InvalidKeywordError):
        DeviceManager(device='GPU', precision='HALF')
    with pytest.raises(InvalidKeywordError):
        DeviceManager(device='CPU')                                   # no CPU fallback here


def test_gpu_request_without_gpu_fails_loudly():
    import torch
    from pycc_amd import DeviceManager, PyCCError
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(PyCCError):
        DeviceManager(device='GPU', precision='DP')


def test_ccwfn_rejects_unknown_kwargs_and_models():
    from pycc_amd import CCwfn, PyCCError, InvalidKeywordError
    with pytest.raises(PyCCError):
        CCwfn(None, no=1, bogus=1)
    with pytest.raises(InvalidKeywordError):
        CCwfn(None, no=1, model='CC3')
    assert 'CCSD(T)' in CCwfn.VALID_MODELS


class _RecordingLib:
    """Stands in for libccx in the pycc binding: records every call, computes nothing."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def fn(*args):
            self.calls.append(name)
            return 0
        return fn


def test_pycc_binding_refuses_bad_operands_before_calling_the_library():
    """pycc's 'SP' hands float32 tensors to the binding (pycc/device.py:147-151); the library reads
    FP64 only. The binding converts float32 / float64 CUDA tensors and refuses anything else
    (host memory, integer dtypes, wrong shapes) before any library call."""
    import torch
    from pycc_amd import ccx_backend as cb
    ctx = object.__new__(cb.CCXContext)
    ctx.lib, ctx.h = _RecordingLib(), ctypes.c_void_p(1)
    ctx.no, ctx.nv, ctx.nmo = 2, 3, 5
    F = torch.zeros(5, 5, dtype=torch.float32)
    t1 = torch.zeros(2, 3, dtype=torch.float32)
    t2 = torch.zeros(2, 2, 3, 3, dtype=torch.float32)
    for call in (ctx.residuals, ctx.energy, ctx.triples):
        with pytest.raises(TypeError, match="on the GPU"):
            call(F, t1, t2)                        # host tensors
    with pytest.raises(TypeError, match="float32 or float64"):
        cb.as_fp64(torch.zeros(2, 3, dtype=torch.int32), "t1")
    with pytest.raises(TypeError, match="torch tensor"):
        cb.as_fp64([[0.0] * 3] * 2, "t1")
    assert ctx.lib.calls == []
    ctx.h = None
    with pytest.raises(ValueError, match="no libccx mode"):
        cb.CCXContext(2, 3, "HALF", None, None, lib=_RecordingLib())
    assert set(cb.CCXContext.MODES) == {"DP", "SP", "F16X3"}


def test_t3_builders_validate_arguments_without_a_gpu():
    """pycc_amd.cctriples' batched builders refuse what the device path cannot take (host tensors, a
    contraction callable other than the device backend, non-slice o / v, out-of-range indices) before
    any library call."""
    import numpy as np
    import torch
    from pycc_amd import PyCCError, cctriples
    from pycc_amd.device import ContractionBackend
    backend = ContractionBackend(torch.device("cuda:0"))               # constructing it touches no GPU
    no, nv = 2, 3
    o, v = slice(0, no), slice(no, no + nv)
    F = torch.zeros((no + nv, no + nv), dtype=torch.float64)
    t2 = torch.zeros((no, no, nv, nv), dtype=torch.float64)
    W = torch.zeros((nv, nv, nv, no), dtype=torch.float64)
    U = torch.zeros((no, nv, no, no), dtype=torch.float64)
    with pytest.raises(PyCCError, match="device contraction backend"):
        cctriples.t3c_ijk(o, v, 0, 0, 0, t2, W, U, F, np.einsum)
    with pytest.raises(PyCCError, match="slices"):
        cctriples.t3c_ijk([0, 1], v, 0, 0, 0, t2, W, U, F, backend)
    with pytest.raises(PyCCError, match="float64 tensors on cuda:0"):      # host tensors
        cctriples.t3c_abc(o, v, 0, 0, 0, t2, W, U, F, backend)
    with pytest.raises(PyCCError, match="float64 tensors on cuda:0"):
        cctriples.t3d_ijk(o, v, 0, 0, 0, t2[0].float(), t2, t2, F, backend)

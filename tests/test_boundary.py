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

"""Device / precision policy and the contraction backend (counterpart of ``pycc/device.py``).

``DeviceManager(device, precision)`` validates the keywords like ``pycc/device.py:113-133`` and
resolves the compute device. It differs from the reference in two deliberate ways:

* the compute device is an AMD GPU and there is no CPU fallback: ``device='CPU'`` is not offered
  (pycc's own numpy path is the CPU implementation) and a missing GPU raises ``PyCCError``;
* ``precision`` selects the MI355X precision mode of the contraction kernels: ``'DP'`` (FP64
  everywhere), ``'MIXED'`` (FP32 MFMA ladder + ring GEMMs, everything else FP64; ``'SP'`` is
  accepted as its alias), ``'MIXED_BF16X3'`` (as MIXED, ladder as BF16x3 split products) or
  ``'MIXED_F16X3'`` (the ladder, the ring GEMMs and the other large contractions as FP16x3 split
  products with power-of-two operand scales).

``ContractionBackend.__call__(subscripts, A, B)`` mirrors ``pycc/device.py:64-86`` for two torch
operands on the GPU and runs on the library's MFMA contraction kernel (``ccx_contract``).
"""
from __future__ import annotations

import ctypes

from . import _lib
from .exceptions import InvalidKeywordError, PyCCError

PRECISION_MODES = {"DP": _lib.MODE_DP, "SP": _lib.MODE_MIXED, "MIXED": _lib.MODE_MIXED,
                   "MIXED_BF16X3": _lib.MODE_MIXED_BF16X3, "MIXED_F16X3": _lib.MODE_MIXED_F16X3}


def _torch():
    import torch
    return torch


def current_stream_handle():
    """hipStream_t of torch's current stream, as an integer for the C ABI."""
    return _torch().cuda.current_stream().cuda_stream


def dptr(t):
    """Device pointer of a contiguous torch tensor."""
    return ctypes.c_void_p(t.data_ptr())


class ContractionBackend(object):
    """Two-operand einsum on the GPU through the library's MFMA contraction kernel."""

    def __init__(self, device1):
        self.device1 = device1

    def __call__(self, subscripts, *operands, alpha=1.0, out=None, beta=0.0):
        torch = _torch()
        if len(operands) != 2:
            raise PyCCError("the device contraction backend takes exactly two operands")
        A, B = operands
        if A.dtype != B.dtype or A.dtype not in (torch.float64, torch.float32):
            raise PyCCError("operands must both be float64 or both float32")
        A = A if A.device == self.device1 else A.to(self.device1)
        B = B if B.device == self.device1 else B.to(self.device1)
        lhs, rhs = subscripts.replace(" ", "").split("->")
        la, lb = lhs.split(",")
        ext = {}
        for lab, t in ((la, A), (lb, B)):
            if len(lab) != t.dim():
                raise PyCCError("subscripts %r do not match operand rank" % subscripts)
            for c, n in zip(lab, t.shape):
                if ext.setdefault(c, n) != n:
                    raise PyCCError("extent mismatch for index %r" % c)
        shape = [ext[c] for c in rhs]
        if out is None:
            out = torch.zeros(shape, dtype=A.dtype, device=self.device1)
            beta = 0.0
        i64 = lambda xs: (ctypes.c_int64 * max(1, len(xs)))(*xs)
        fn = "ccx_contract" if A.dtype == torch.float64 else "ccx_contract_f32"
        _lib.call(fn, subscripts.encode(), alpha, dptr(A), i64(A.shape), i64(A.stride()),
                  dptr(B), i64(B.shape), i64(B.stride()), beta, dptr(out), i64(out.shape),
                  i64(out.stride()), ctypes.c_void_p(current_stream_handle()))
        return out

    def matmul_f16x3(self, A, B, out=None, alpha=1.0, beta=0.0):
        """alpha A @ B (+ beta out) for 2-D float32 tensors through the FP16x3 split GEMM
        (ccx_gemm_f16x3): ~22-bit operands, FP32 accumulation. Transposed views are passed as
        transposes, other strides are made contiguous."""
        torch = _torch()
        if A.dim() != 2 or B.dim() != 2 or A.dtype != torch.float32 or B.dtype != torch.float32:
            raise PyCCError("matmul_f16x3 takes two 2-D float32 tensors")
        if A.shape[1] != B.shape[0]:
            raise PyCCError("inner extents differ: %s @ %s" % (tuple(A.shape), tuple(B.shape)))
        A = A.to(self.device1)
        B = B.to(self.device1)

        def op(X):
            if X.stride(1) == 1 and X.stride(0) >= X.shape[1]:
                return X, 0, X.stride(0)
            if X.stride(0) == 1 and X.stride(1) >= X.shape[0]:
                return X, 1, X.stride(1)
            X = X.contiguous()
            return X, 0, X.shape[1]

        A, ta, lda = op(A)
        B, tb, ldb = op(B)
        M, K, N = A.shape[0], A.shape[1], B.shape[1]
        if out is None:
            out = torch.zeros((M, N), dtype=torch.float32, device=self.device1)
            beta = 0.0
        elif out.shape != (M, N) or out.stride(1) != 1 or out.dtype != torch.float32:
            raise PyCCError("out must be a row-major float32 (M, N) tensor")
        _lib.call("ccx_gemm_f16x3", ta, tb, M, N, K, alpha, dptr(A), lda, dptr(B), ldb, beta, dptr(out),
                  out.stride(0), ctypes.c_void_p(current_stream_handle()))
        return out


class DeviceManager(object):
    """Owns the device/precision policy and the contraction backend (pycc/device.py:89-173)."""

    VALID_DEVICE = ['GPU']
    VALID_PRECISION = list(PRECISION_MODES)

    def __init__(self, device: str = 'GPU', precision: str = 'DP') -> None:
        if not isinstance(precision, str) or precision.upper() not in self.VALID_PRECISION:
            raise InvalidKeywordError('precision', precision, self.VALID_PRECISION)
        self.precision = precision.upper()
        if not isinstance(device, str) or device.upper() not in self.VALID_DEVICE:
            raise InvalidKeywordError('device', device, self.VALID_DEVICE)
        self.device = device.upper()
        torch = _torch()
        if not torch.cuda.is_available():
            raise PyCCError("device='GPU' needs an AMD GPU visible to torch; none was found "
                            "(there is no CPU fallback)")
        self.device0 = torch.device('cpu')
        self.device1 = torch.device('cuda:0')
        self.mode = PRECISION_MODES[self.precision]
        self.real_dtype = torch.float64          # amplitudes, residuals, DIIS are FP64 in every mode
        _lib.load()
        self.contract = ContractionBackend(self.device1)

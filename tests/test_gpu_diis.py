"""The device DIIS (``ccx_diis_reset/add/extrapolate``, ``pycc/utils.py:257-361``) against the same
mathematical pins as the oracle's: the constrained least-squares coefficients over the newest
vectors the drop rule keeps (``start_diis`` 1 and later), and the exact solution of a linear
fixed-point map once n + 1 error vectors span the space (Anderson acceleration = GMRES). The
vectors are arbitrary; the context only supplies the (no, nv) shapes."""
import ctypes

import numpy as np
import pytest

import test_diis_pins as pins
from pycc_amd.synth import synthetic_df

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _device_diis_class(no, nv):
    import pycc_amd
    from pycc_amd import _lib
    from pycc_amd.device import dptr
    cc = pycc_amd.CCwfn(synthetic_df(no, nv, seed=0), no=no, precision="DP", verbose=False)

    class DeviceDIIS:
        """helper_diis's interface (add_error_vector, extrapolate, last_coefficients) over the C ABI."""

        def __init__(self, t1, t2, max_diis):
            self.max_diis = max_diis
            self.last_coefficients = None
            d1, d2 = self._dev(t1, t2)
            _lib.call("ccx_diis_reset", cc._ctx.h, dptr(d1), dptr(d2), int(max_diis), 32, cc._stream())
            torch.cuda.synchronize()

        @staticmethod
        def _dev(t1, t2):
            d1 = torch.tensor(np.asarray(t1).reshape(no, nv), dtype=torch.float64, device="cuda")
            d2 = torch.tensor(np.asarray(t2).reshape(no, no, nv, nv), dtype=torch.float64, device="cuda")
            return d1, d2

        def add_error_vector(self, t1, t2):
            d1, d2 = self._dev(t1, t2)
            _lib.call("ccx_diis_add", cc._ctx.h, dptr(d1), dptr(d2), cc._stream())
            torch.cuda.synchronize()

        def extrapolate(self, t1, t2):
            d1, d2 = self._dev(t1, t2)
            coef = (ctypes.c_double * 64)(*([float("nan")] * 64))
            _lib.call("ccx_diis_extrapolate", cc._ctx.h, dptr(d1), dptr(d2), coef, cc._stream())
            torch.cuda.synchronize()
            c = np.array(coef[:])
            self.last_coefficients = c[~np.isnan(c)]
            return d1.cpu().numpy().reshape(np.shape(t1)), d2.cpu().numpy().reshape(np.shape(t2))

    return cc, DeviceDIIS


@pytest.fixture(scope="module")
def wide_diis():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert pins.WIDE == (6, 36)                  # no = 1, nv = 6: t1 has 6 numbers, t2 has 36
    cc, cls = _device_diis_class(1, 6)
    yield cls
    del cc


@pytest.mark.parametrize("max_diis,start_diis", [(8, 1), (3, 1), (1, 1), (3, 2), (3, 6), (2, 8)])
def test_device_diis_coefficients_and_window(wide_diis, max_diis, start_diis):
    assert pins.run_window_check(wide_diis, max_diis, start_diis) < 1e-9


def test_device_diis_solves_linear_fixed_point_exactly():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert (pins.N1, pins.N2) == (2, 4)          # no = 1, nv = 2
    cc, cls = _device_diis_class(1, 2)
    assert pins.linear_fixed_point_error(cls) < 1e-10
    del cc

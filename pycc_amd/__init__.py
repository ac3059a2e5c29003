"""MI355X-native closed-shell CCSD (the spatial RHF-CCSD hot path of CrawfordGroup/pycc).

Public API (names as exported by ``pycc/__init__.py``):

* ``CCwfn`` / ``ccwfn`` -- amplitude solver: ``residuals``, ``cc_energy``, ``solve_cc``;
* ``DeviceManager``, ``ContractionBackend`` -- device/precision policy and the GPU einsum;
* ``PyCCError``, ``PyCCWarning``, ``InvalidKeywordError``, ``DeviceError``.

The arithmetic runs in ``libccx.so`` (HIP, gfx950), declared in ``include/ccx.h``.
"""
from .exceptions import PyCCError, PyCCWarning, InvalidKeywordError, DeviceError
from .device import DeviceManager, ContractionBackend
from .ccwfn import CCwfn, ccwfn
from . import synth

__all__ = ["CCwfn", "ccwfn", "DeviceManager", "ContractionBackend", "PyCCError", "PyCCWarning",
           "InvalidKeywordError", "DeviceError", "synth"]

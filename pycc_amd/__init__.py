"""MI355X-native closed-shell CCSD (the spatial RHF-CCSD hot path of CrawfordGroup/pycc).

Public API (names as exported by ``pycc/__init__.py``):

* ``CCwfn`` / ``ccwfn`` -- amplitude solver: ``residuals``, ``cc_energy``, ``solve_cc``;
* ``cchbar``, ``cclambda`` -- HBAR and the Lambda equations (``solve_lambda``, ``residuals``,
  ``pseudoenergy``); ``autograd.residual`` -- the residual as a ``torch.autograd.Function`` whose
  backward is the Lambda-equation vector-Jacobian product; ``ccdensity`` -- the one- and
  two-particle densities (the integral half of the vJp);
* ``cctriples`` -- the (T) correction (``t_tjl``, ``t_vikings``, ``t_vikings_inverted``);
* ``DeviceManager``, ``ContractionBackend`` -- device/precision policy and the GPU einsum;
* ``integrals`` -- FCIDUMP / ``.npz`` integral files and ``MOHamiltonian`` (the psi4-free input);
* ``PyCCError``, ``PyCCWarning``, ``InvalidKeywordError``, ``DeviceError``.

The arithmetic runs in ``libccx.so`` (HIP, gfx950), declared in ``include/ccx.h``.
"""
from .exceptions import PyCCError, PyCCWarning, InvalidKeywordError, DeviceError
from .device import DeviceManager, ContractionBackend
from .ccwfn import CCwfn, ccwfn
from .cchbar import cchbar
from .cclambda import cclambda
from .ccdensity import ccdensity
from . import cctriples
from . import synth
from . import integrals

__all__ = ["CCwfn", "ccwfn", "cchbar", "cclambda", "ccdensity", "DeviceManager", "ContractionBackend", "PyCCError", "PyCCWarning",
           "InvalidKeywordError", "DeviceError", "cctriples", "synth", "integrals"]

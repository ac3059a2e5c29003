"""Closed-shell CCSD amplitude solver on the GPU (counterpart of ``pycc.ccwfn.CCwfn``).

``CCwfn(H, no=..., device='GPU', precision='DP').solve_cc(e_conv, r_conv, maxiter, max_diis,
start_diis)`` reproduces ``pycc/ccwfn.py:216-319`` for the spin-adapted CCSD model: MP1 guess,
Jacobi update ``t += r/D``, the 2-norm "rms", convergence when both ``|dE| < e_conv`` and
``rms < r_conv``, DIIS after the convergence test, and ``None`` when ``maxiter`` is exhausted.
``residuals(F, t1, t2)`` and ``cc_energy(o, v, F, L, t1, t2)`` keep the reference signatures.
With ``model='CCSD(T)'`` the converged energy includes the (T) correction (``ccwfn.py:298-310``,
``triples_energy`` / ``pycc_amd.cctriples.t_tjl``).

The Hamiltonian ``H`` is either

* a pycc-layout dense Hamiltonian (attributes ``F`` (nmo, nmo) and Dirac ``ERI`` (nmo,)*4, as
  ``pycc/hamiltonian.py:59-70`` builds them), or
* density-fitting factors (attributes ``B`` (naux, nmo, nmo) and ``eps`` or ``fock``), in which case
  the integral blocks are generated on the GPU and no nmo^4 array exists anywhere.

A frozen core is a slice offset (``nfzc``), as in ``pycc/wavefunction.py:304-312``.
Amplitudes, residuals, denominators, DIIS, norms and energies are FP64 torch tensors on the GPU in
every precision mode; ``precision`` only changes the arithmetic of the ladder and ring GEMMs.
"""
from __future__ import annotations

import ctypes
import weakref
import math
import time

import numpy as np

from . import _lib
from .device import DeviceManager, current_stream_handle, dptr
from .exceptions import InvalidKeywordError, PyCCError


def _title(text):
    return "\n" + text


def _iteration(niter, energy=None, de=None, rms=None, e_label="Ecorr", note=None):
    cols = []
    if energy is not None:
        cols.append("%s = %.15f" % (e_label, energy))
    if de is not None:
        cols.append("dE = % .5E" % de)
    if rms is not None:
        cols.append("rms = % .5E" % rms)
    if note is not None:
        cols.append(note)
    return "Iter %3d: %s" % (niter, "  ".join(cols))


class _Ctx:
    """Owns one ``ccx_ctx``."""

    def __init__(self, no, nv, mode):
        h = ctypes.c_void_p()
        _lib.call("ccx_create", no, nv, mode, ctypes.byref(h))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                _lib.load().ccx_destroy(h)
            except Exception:
                pass
            self.h = None


class CCwfn(object):
    """Spin-adapted closed-shell CCSD on one AMD GPU (see the module docstring)."""

    VALID_MODELS = ['CCSD', 'CCSD(T)']

    def __init__(self, H, no=None, nfzc=0, model='CCSD', device='GPU', precision='DP',
                 verbose=True, **kwargs):
        time_init = time.time()
        if kwargs:
            raise PyCCError("Unexpected keyword argument(s): %s" % sorted(kwargs))   # wavefunction.py:181
        model = str(model).upper()
        if model not in self.VALID_MODELS:
            raise InvalidKeywordError('model', model, self.VALID_MODELS)
        self.model = model
        self.verbose = verbose
        mgr = DeviceManager(device=device, precision=precision)
        self.device_manager = mgr
        self.device, self.precision = mgr.device, mgr.precision
        self.device0, self.device1 = mgr.device0, mgr.device1
        self.contract = mgr.contract
        import torch
        self._torch = torch

        if no is None:
            no = getattr(H, "no", None)
        if no is None and getattr(H, "nocc", None) is not None:
            no = int(H.nocc) - int(nfzc)      # doubly occupied minus frozen core (wavefunction.py:305)
        if no is None:
            raise PyCCError("the number of active occupied orbitals `no` is required")
        self.nfzc = int(nfzc)
        self.no = int(no)
        self.e_conv, self.r_conv, self.maxiter = 1e-7, 1e-7, 100

        dev = self.device1
        f64 = torch.float64
        if hasattr(H, "B"):
            B = H.B
            fock = H.fock if hasattr(H, "fock") else np.diag(np.asarray(H.eps))
            fock = np.asarray(fock, dtype=np.float64)
            nmo_all = fock.shape[0]
            act = slice(self.nfzc, nmo_all)
            F = torch.as_tensor(np.ascontiguousarray(fock[act, act]), device=dev)
            if isinstance(B, torch.Tensor):
                Bt = B.to(dev, f64)[:, act, act].contiguous()
            else:
                Bt = torch.as_tensor(np.asarray(B, dtype=np.float64), device=dev)[:, act, act].contiguous()
            self._source = ("df", Bt)
        elif hasattr(H, "ERI") and hasattr(H, "F"):
            Fh = H.F.detach().cpu().numpy() if hasattr(H.F, "detach") else np.asarray(H.F)
            nmo_all = Fh.shape[0]
            act = slice(self.nfzc, nmo_all)
            F = torch.as_tensor(np.ascontiguousarray(Fh[act, act], dtype=np.float64), device=dev)
            E = H.ERI
            E = E if isinstance(E, torch.Tensor) else torch.as_tensor(np.asarray(E, dtype=np.float64))
            E = E.to(dev, f64)[act, act, act, act].contiguous()
            self._source = ("dense", E)
        else:
            raise PyCCError("H must provide (F, ERI) or density-fitting factors (B, eps/fock)")

        self.nmo = F.shape[0]
        self.nv = self.nmo - self.no
        if self.no < 1 or self.nv < 1:
            raise PyCCError("need at least one active occupied and one virtual orbital")
        self.o = slice(0, self.no)
        self.v = slice(self.no, self.nmo)
        self.H = _DeviceHamiltonian(F)

        self._ctx = _Ctx(self.no, self.nv, mgr.mode)
        stream = ctypes.c_void_p(current_stream_handle())
        kind, src = self._source
        if kind == "df":
            _lib.call("ccx_set_integrals_df", self._ctx.h, dptr(F), dptr(src), int(src.shape[0]), stream)
        else:
            _lib.call("ccx_set_integrals_dense", self._ctx.h, dptr(F), dptr(src), stream)
        del src
        self._source = None

        # denominators (ccwfn.py:195-198) and the MP1 guess (ccwfn.py:210-211)
        eps = torch.diagonal(F).clone()
        eps_o, eps_v = eps[self.o], eps[self.v]
        self.H.eps = eps
        self.Dia = eps_o.reshape(-1, 1) - eps_v
        self._e_dev = torch.zeros(2, dtype=f64, device=dev)     # device scalars: energy, rms^2
        self.t1 = torch.zeros((self.no, self.nv), dtype=f64, device=dev)
        self.t2 = self._mp1_guess()
        self.ecc = None
        self.niter = 0
        if self.verbose:
            print("CCwfn built in %.3f seconds." % (time.time() - time_init))

    # ---- helpers --------------------------------------------------------------------------
    @property
    def Dijab(self):
        """Doubles denominators (ccwfn.py:195-197), formed on request (o^2 v^2 FP64)."""
        e = self.H.eps
        eo, ev = e[self.o], e[self.v]
        return eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev

    def _stream(self):
        return ctypes.c_void_p(current_stream_handle())

    def _mp1_guess(self):
        # t2 = <ij|ab>/Dijab: r2 at t = 0 is exactly <ij|ab> (every other term is at least linear in
        # t), so one residual evaluation plus a Jacobi step from zero gives the MP1 amplitudes.
        torch = self._torch
        t1 = torch.zeros((self.no, self.nv), dtype=torch.float64, device=self.device1)
        t2 = torch.zeros((self.no, self.no, self.nv, self.nv), dtype=torch.float64, device=self.device1)
        r1, r2 = self.residuals(self.H.F, t1, t2)
        r1.zero_()          # t1 starts at zero (ccwfn.py:210) whatever f_vo is
        self._update(t1, t2, r1, r2)
        return t2

    def _update(self, t1, t2, r1, r2):
        _lib.call("ccx_update", self._ctx.h, dptr(t1), dptr(t2), dptr(r1), dptr(r2),
                  dptr(self._e_dev[1:]), self._stream())
        return math.sqrt(float(self._e_dev[1].item()))

    # ---- reference API --------------------------------------------------------------------
    def residuals(self, F, t1, t2):
        """r1, r2 = <mu|e^{-T} H e^{T}|0> (ccwfn.py:321-372); r2 symmetrised as ccwfn.py:790."""
        torch = self._torch
        F = self._as_dev(F, (self.nmo, self.nmo))
        t1 = self._as_dev(t1, (self.no, self.nv))
        t2 = self._as_dev(t2, (self.no, self.no, self.nv, self.nv))
        r1 = torch.empty_like(t1)
        r2 = torch.empty_like(t2)
        _lib.call("ccx_residual", self._ctx.h, dptr(F), dptr(t1), dptr(t2), dptr(r1), dptr(r2),
                  self._stream())
        return r1, r2

    def cc_energy(self, o=None, v=None, F=None, L=None, t1=None, t2=None):
        """E = 2 f_ia t_ia + tau_ijab L_ijab (ccwfn.py:1160-1161); L is held by the device."""
        F = self.H.F if F is None else self._as_dev(F, (self.nmo, self.nmo))
        t1 = self.t1 if t1 is None else self._as_dev(t1, (self.no, self.nv))
        t2 = self.t2 if t2 is None else self._as_dev(t2, (self.no, self.no, self.nv, self.nv))
        _lib.call("ccx_energy", self._ctx.h, dptr(F), dptr(t1), dptr(t2), dptr(self._e_dev[:1]),
                  self._stream())
        return float(self._e_dev[0].item())

    def solve_cc(self, e_conv=1e-7, r_conv=1e-7, maxiter=100, max_diis=8, start_diis=1):
        """ccwfn.py:216-319 (spatial CCSD branch): returns the correlation energy, or None when
        ``maxiter`` iterations do not converge."""
        t0 = time.time()
        name = "T-amplitudes (%s)" % self.model
        for step in self.iterate(e_conv, r_conv, maxiter, max_diis, start_diis):
            if step["converged"]:
                ecc = step["ecc"]
                if self.verbose:
                    print("%s converged in %.3f seconds." % (name, time.time() - t0))
                if self.model == 'CCSD(T)':                                 # ccwfn.py:298-310
                    et = self.triples_energy()
                    self.et = et
                    if self.verbose:
                        print("E(CCSD) = %20.15f" % ecc)
                        print("E(T)    = %20.15f" % et)
                    ecc = ecc + et
                    self.ecc = ecc
                elif self.verbose:
                    print("E(%s) = %20.15f" % (self.model, ecc))
                return ecc
        return None

    def iterate(self, e_conv=1e-7, r_conv=1e-7, maxiter=100, max_diis=8, start_diis=1):
        """Generator over the iterations of ``solve_cc`` (ccwfn.py:260-319); yields one dict per
        iteration (niter, ecc, ediff, rms, converged) after its convergence test, so a caller can
        time or stop the solve between iterations. Stops after the converged iteration."""
        self.e_conv, self.r_conv, self.maxiter = e_conv, r_conv, maxiter
        F = self.H.F
        ecc = self.cc_energy()                                              # ccwfn.py:260
        name = "T-amplitudes (%s)" % self.model
        self.history = [(0, ecc, -ecc, None)]
        if self.verbose:
            print(_title(name))
            print(_iteration(0, energy=ecc, de=-ecc, e_label="CC Ecorr", note="MP2"))
        cap = max(1, max_diis) + max(1, start_diis)
        _lib.call("ccx_diis_reset", self._ctx.h, dptr(self.t1), dptr(self.t2), int(max_diis), int(cap),
                  self._stream())                                           # ccwfn.py:266
        coef = (ctypes.c_double * 64)()
        for niter in range(1, maxiter + 1):
            ecc_last = ecc
            r1, r2 = self.residuals(F, self.t1, self.t2)                    # ccwfn.py:272
            rms = self._update(self.t1, self.t2, r1, r2)                    # ccwfn.py:281-284
            del r1, r2
            ecc = self.cc_energy()                                          # ccwfn.py:286
            ediff = ecc - ecc_last
            self.history.append((niter, ecc, ediff, rms))
            self.niter = niter
            if self.verbose:
                print(_iteration(niter, energy=ecc, de=ediff, rms=rms, e_label="CC Ecorr"))
            conv = abs(ediff) < e_conv and abs(rms) < r_conv                # ccwfn.py:295
            if conv:
                self.ecc = ecc
            yield {"niter": niter, "ecc": ecc, "ediff": ediff, "rms": rms, "converged": conv}
            if conv:
                return
            _lib.call("ccx_diis_add", self._ctx.h, dptr(self.t1), dptr(self.t2), self._stream())
            if niter >= start_diis:                                         # ccwfn.py:317-319
                _lib.call("ccx_diis_extrapolate", self._ctx.h, dptr(self.t1), dptr(self.t2), coef,
                          self._stream())

    # ---- backward: HBAR, Lambda residual, vector-Jacobian product ----------------------------
    def _require_dp(self, what, mixed_ok=False):
        """HBAR, the Lambda residual and the vJp also run in a MIXED_F16X3 context (``mixed_ok``;
        their large contractions as FP16x3 split products, HBAR held in FP64, its v^4 block never
        stored); the Jacobian and the densities run in a DP context only."""
        if self.precision == 'DP' or (mixed_ok and self.precision == 'MIXED_F16X3'):
            return
        raise PyCCError("%s runs in FP64: build the CCwfn with precision='DP'%s (this one is %s)"
                        % (what, " (or 'MIXED_F16X3')" if mixed_ok else "", self.precision))

    def _hbar_for(self, F, t1, t2, stored=False, owner=None):
        """Make the context's HBAR the one at (F, t1, t2), building it unless it already is.

        The context holds one HBAR. It is identified by the tensors it was built from (held here, so
        their memory cannot be reused) and their version counters, so an in-place change of t1, t2
        or F forces a rebuild. A stored HBAR also serves factored requests."""
        self._require_dp("HBAR" + (" with the stored v^4 block" if stored else ""), mixed_ok=not stored)
        key = tuple((x.data_ptr(), x._version, tuple(x.shape)) for x in (F, t1, t2))
        cur = getattr(self, "_hbar_src", None)
        # the owner (a cchbar, which refers back to this object) is held weakly: no reference cycle
        # keeps the context, and its device memory, alive after the last user reference goes
        cur_owner = cur["owner"]() if cur is not None and cur["owner"] is not None else None
        if cur is not None and cur["key"] == key and (cur["stored"] or not stored) and cur_owner is owner:
            return
        self._hbar_src = None
        _lib.call("ccx_hbar_build", self._ctx.h, dptr(F), dptr(t1), dptr(t2),
                  _lib.HBAR_STORED if stored else 0, self._stream())
        self._hbar_src = {"key": key, "stored": bool(stored),
                          "owner": weakref.ref(owner) if owner is not None else None, "refs": (F, t1, t2)}

    def _lambda_residual(self, l1, l2, with_source=True):
        """r_L1, r_L2 (cclambda.py:345-369, :468-496) on the context's current HBAR."""
        torch = self._torch
        l1 = self._as_dev(l1, (self.no, self.nv))
        l2 = self._as_dev(l2, (self.no, self.no, self.nv, self.nv))
        r1 = torch.empty_like(l1)
        r2 = torch.empty_like(l2)
        _lib.call("ccx_lambda_residual", self._ctx.h, dptr(l1), dptr(l2), dptr(r1), dptr(r2),
                  1 if with_source else 0, self._stream())
        return r1, r2

    def vjp(self, t1, t2, c1, c2, F=None, store_vvvv=False):
        """(g1, g2) = the vector-Jacobian product of ``residuals(F, t1, t2)`` with cotangent
        (c1, c2): g = sum_k c_k . d r_k / dt, with g2 pair-symmetrised (the t2 parameterisation is
        pair-symmetric). It is the Lambda residual without source terms at l1 = c1, l2 = c2 + c2^T
        (cclambda.py:345-369, :468-496); F is held constant."""
        torch = self._torch
        F = self.H.F if F is None else self._as_dev(F, (self.nmo, self.nmo))
        t1 = self._as_dev(t1, (self.no, self.nv))
        t2 = self._as_dev(t2, (self.no, self.no, self.nv, self.nv))
        c1 = self._as_dev(c1, (self.no, self.nv))
        c2 = self._as_dev(c2, (self.no, self.no, self.nv, self.nv))
        self._hbar_for(F, t1, t2, store_vvvv)
        g1 = torch.empty_like(c1)
        g2 = torch.empty_like(c2)
        _lib.call("ccx_vjp", self._ctx.h, None, None, None, dptr(c1), dptr(c2), dptr(g1), dptr(g2),
                  _lib.VJP_REUSE_HBAR, self._stream())
        return g1, g2

    def jvp(self, t1, t2, x1, x2, F=None, store_vvvv=False):
        """(j1, j2) = the derivative of ``residuals(F, t1, t2)`` along (x1, x2): the CCSD Jacobian
        applied on HBAR (``CCderiv._ccsd_jacobian``, ccderiv.py:149-180, doubles symmetrised). Only the
        pair-symmetric part of x2 acts; F is held constant."""
        self._require_dp("the Jacobian (jvp)")
        torch = self._torch
        F = self.H.F if F is None else self._as_dev(F, (self.nmo, self.nmo))
        t1 = self._as_dev(t1, (self.no, self.nv))
        t2 = self._as_dev(t2, (self.no, self.no, self.nv, self.nv))
        x1 = self._as_dev(x1, (self.no, self.nv))
        x2 = self._as_dev(x2, (self.no, self.no, self.nv, self.nv))
        self._hbar_for(F, t1, t2, store_vvvv)
        j1 = torch.empty_like(x1)
        j2 = torch.empty_like(x2)
        _lib.call("ccx_jvp", self._ctx.h, None, None, None, dptr(x1), dptr(x2), dptr(j1), dptr(j2),
                  _lib.VJP_REUSE_HBAR, self._stream())
        return j1, j2

    def integral_vjp(self, t1, t2, c1, c2, vvvv=False, fock_only=False):
        """The integral half of the vector-Jacobian product of ``residuals(F, t1, t2)`` with cotangent
        (c1, c2): d(c1.r1 + c2.r2)/dF and d(...)/d<pq|rs> at fixed amplitudes, as the density blocks of
        ``ccdensity.py:335-931`` at l = (c1, c2 + c2^T) with the energy part (l = 0) removed. Returns a
        dict of device tensors: ``D`` (nmo x nmo, Fock gradient) and ``Doooo``, ``Dooov``, ``Dovov``,
        ``Doovv``, ``Dvvvo`` (and ``Dvvvv`` with ``vvvv=True``) in the reference's density layout.
        ``fock_only=True`` returns only ``D``."""
        from .ccdensity import _blocks
        self._require_dp("the integral vJp")
        t1 = self._as_dev(t1, (self.no, self.nv))
        t2 = self._as_dev(t2, (self.no, self.no, self.nv, self.nv))
        c1 = self._as_dev(c1, (self.no, self.nv))
        c2 = self._as_dev(c2, (self.no, self.no, self.nv, self.nv))
        l2 = (c2 + c2.permute(1, 0, 3, 2)).contiguous()
        big = not fock_only
        with_l = _blocks(self, t1, t2, c1, l2, big=big, vvvv=vvvv and big)
        # l = 0 part; Dvvvo and Dvvvv have an l factor in every term (ccdensity.py:609, :716-741)
        z1, z2 = self._torch.zeros_like(t1), self._torch.zeros_like(t2)
        energy_part = _blocks(self, t1, t2, z1, z2, big=False, vvvv=False)
        del z1, z2
        keys = ["D"] if fock_only else [k for k, x in with_l.items() if x is not None]
        out = {}
        for k in keys:
            if energy_part.get(k) is not None:
                with_l[k].sub_(energy_part[k])
            out[k] = with_l[k]
        return out

    def energy_densities(self, t1, t2, vvvv=False, fock_only=False):
        """dE/dF and dE/d<pq|rs> of ``cc_energy`` at (t1, t2): the density blocks of
        ``ccdensity.py:335-931`` at l = 0 (same layout as ``integral_vjp``)."""
        from .ccdensity import _blocks
        self._require_dp("the energy densities")
        t1 = self._as_dev(t1, (self.no, self.nv))
        t2 = self._as_dev(t2, (self.no, self.no, self.nv, self.nv))
        z1, z2 = self._torch.zeros_like(t1), self._torch.zeros_like(t2)
        b = _blocks(self, t1, t2, z1, z2, big=not fock_only, vvvv=vvvv and not fock_only)
        return {"D": b["D"]} if fock_only else {k: x for k, x in b.items() if x is not None}

    def energy_gradient(self, t1=None, t2=None, F=None):
        """(dE/dt1, dE/dt2) of ``cc_energy`` = (2 Hov, L): the source terms of the Lambda equations
        (cclambda.py:348, :468)."""
        torch = self._torch
        F = self.H.F if F is None else self._as_dev(F, (self.nmo, self.nmo))
        t1 = self.t1 if t1 is None else self._as_dev(t1, (self.no, self.nv))
        t2 = self.t2 if t2 is None else self._as_dev(t2, (self.no, self.no, self.nv, self.nv))
        self._hbar_for(F, t1, t2)
        z1 = torch.zeros_like(t1)
        z2 = torch.zeros_like(t2)
        r1, r2 = self._lambda_residual(z1, z2, with_source=True)
        return r1, r2.mul_(0.5)

    def pseudoenergy(self, l2):
        """1/2 <ij|ab> l2_ijab (cclambda.py:570)."""
        l2 = self._as_dev(l2, (self.no, self.no, self.nv, self.nv))
        _lib.call("ccx_pseudoenergy", self._ctx.h, dptr(l2), dptr(self._e_dev[:1]), self._stream())
        return float(self._e_dev[0].item())

    def triples_energy(self, t1=None, t2=None, F=None, i_chunk=None, progress=None):
        """The (T) correction E(T) at (F, t1, t2) on the GPU (``pycc/cctriples.py:177-240``,
        ``t_tjl``): FP64, except that MIXED_F16X3 runs its GEMMs as FP16x3 split products (E(T)
        within ~1e-9 relative of FP64).

        With ``i_chunk`` the occupied triples run in ranges of ``i_chunk`` values of the largest
        index (``ccx_triples_range``), the same value bit for bit; ``progress(i_done, no)`` is called
        after each range."""
        F = self.H.F if F is None else self._as_dev(F, (self.nmo, self.nmo))
        t1 = self.t1 if t1 is None else self._as_dev(t1, (self.no, self.nv))
        t2 = self.t2 if t2 is None else self._as_dev(t2, (self.no, self.no, self.nv, self.nv))
        e = self._e_dev[:1]
        if i_chunk is None:
            _lib.call("ccx_triples", self._ctx.h, dptr(F), dptr(t1), dptr(t2), dptr(e), self._stream())
        else:
            if int(i_chunk) < 1:
                raise PyCCError("i_chunk must be >= 1")
            e.zero_()
            for i0 in range(0, self.no, int(i_chunk)):
                i1 = min(self.no, i0 + int(i_chunk))
                _lib.call("ccx_triples_range", self._ctx.h, dptr(F), dptr(t1), dptr(t2), int(i0), int(i1),
                          dptr(e), self._stream())
                if progress is not None:
                    progress(i1, self.no)
        return float(self._e_dev[0].item())

    # ---- measurement hooks -----------------------------------------------------------------
    def timing_start(self):
        """Start HIP-event timing of the ladder / ring GEMMs and whole residuals (ccx_timing_*)."""
        _lib.call("ccx_timing_start", self._ctx.h)

    def timing_read(self):
        """{'ladder'|'ring'|'residual': (launches, ms, algorithmic flops)} since timing_start."""
        out = (ctypes.c_double * 9)()
        _lib.call("ccx_timing_read", self._ctx.h, out, 9)
        return {k: (int(out[3 * i]), out[3 * i + 1], out[3 * i + 2])
                for i, k in enumerate(("ladder", "ring", "residual"))}

    def device_bytes(self):
        return int(_lib.load().ccx_device_bytes(self._ctx.h))

    def _as_dev(self, x, shape):
        torch = self._torch
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x, dtype=np.float64))
        x = x.to(self.device1, torch.float64)
        if tuple(x.shape) != tuple(shape):
            raise PyCCError("expected shape %s, got %s" % (tuple(shape), tuple(x.shape)))
        return x.contiguous()


class _DeviceHamiltonian:
    """The active-space Fock matrix on the device (the ERIs live inside the library context)."""

    def __init__(self, F):
        self.F = F
        self.eps = None


def ccwfn(*args, **kwargs):
    """Lower-case alias, as exported by pycc (``pycc/__init__.py``)."""
    return CCwfn(*args, **kwargs)

"""Differentiable frozen-potential propagation: loss = J(T).d after nsteps
ETRS steps under A(t) = d sum_j c_j sin(j w t), with gradients with respect to
the coefficients c_j and the scalar potential V(r) from the library's
discrete adjoint (checkpointed, band-sliced replay).

    prop = propagation_adjoint(ham, nst=64, nsteps=400, dt=0.015, ncoef=32,
                               omega=0.1, direction=(1, 0, 0), checkpoint_every=20)
    loss = td_current_loss(coef, vloc, prop, psi0)    # torch scalar
    loss.backward()                                    # coef.grad, vloc.grad

The forward is real_time::propagate's loop (propagate.hpp:93-131) around
etrs (etrs.hpp:22-82) with observables::current (current.hpp:25-53) at the
end.  The reference has no adjoint.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .hamiltonian import _orbital_args, _stream_ptr


class propagation_adjoint:
    """Owns an inqamd_td_t bound to a ks_hamiltonian."""

    def __init__(self, ham, nst, nsteps, dt, ncoef, omega, direction=(1.0, 0.0, 0.0), order=4,
                 occ=2.0, checkpoint_every=20, band_slice=0, store_powers=True, commutator=None):
        """commutator: the loss's [V_nl, r] form, "inq" or "corrected"
        (default: the Hamiltonian's position_commutator)."""
        self.ham = ham
        if commutator is None:
            commutator = ham.position_commutator
        self.commutator = commutator
        d = _lib.TDDesc()
        d.nst = int(nst)
        d.nsteps = int(nsteps)
        d.dt = float(dt)
        d.order = int(order)
        d.occ = float(occ)
        d.ncoef = int(ncoef)
        d.omega = float(omega)
        d.direction = (ctypes.c_double * 3)(*[float(x) for x in direction])
        d.checkpoint_every = int(checkpoint_every)
        d.band_slice = int(band_slice)
        d.store_powers = 1 if store_powers else 0
        d.commutator = _lib.commutator_code(commutator)
        self.desc = d
        t = ctypes.c_void_p()
        ham.handle.call("inqamd_td_create", ctypes.byref(d), ctypes.byref(t))
        self._t = t
        self.ncoef = int(ncoef)
        self.nst = int(nst)

    @property
    def memory_bytes(self):
        return int(_lib.lib().inqamd_td_memory(self._t))

    @property
    def forward_count(self):
        """Forwards run so far; the backward differentiates the last one."""
        return int(_lib.lib().inqamd_td_forward_count(self._t))

    def _set_stream(self):
        h = self.ham.handle
        _lib.check(h._lib.inqamd_set_stream(h.h, _stream_ptr()), h.h)

    def forward(self, coef, psi0, psi_T=None):
        """Propagate psi0 (device [.., nst]); returns (loss, psi_T or None)."""
        p, nst, ld = _orbital_args(psi0)
        if nst != self.nst:
            raise ValueError(f"psi0 has {nst} bands, the propagation {self.nst}")
        c = np.ascontiguousarray(np.asarray(coef, float).reshape(-1))
        if c.size != self.ncoef:
            raise ValueError(f"{c.size} coefficients, expected {self.ncoef}")
        q = None
        if psi_T is not None:
            q, _, _ = _orbital_args(psi_T)
        loss = ctypes.c_double(0.0)
        self._set_stream()
        h = self.ham.handle
        _lib.check(h._lib.inqamd_td_forward(self._t, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            p, ld, q, ctypes.byref(loss)), h.h)
        self._keep = c
        return loss.value, psi_T

    def backward(self, dloss=1.0, want_lambda0=False):
        """(dL/dc [numpy], dL/dV [device, N], lambda0 or None) at the last forward."""
        dev = torch.device("cuda", torch.cuda.current_device())
        gV = torch.empty(int(np.prod(self.ham.n)), dtype=torch.float64, device=dev)
        gc = np.zeros(max(self.ncoef, 1))
        lam = None
        if want_lambda0:
            lam = torch.empty(tuple(self.ham.n) + (self.nst,), dtype=torch.complex128, device=dev)
        self._set_stream()
        h = self.ham.handle
        _lib.check(h._lib.inqamd_td_backward(
            self._t, ctypes.c_double(dloss), gc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            ctypes.c_void_p(gV.data_ptr()),
            ctypes.c_void_p(lam.data_ptr()) if lam is not None else None), h.h)
        return gc[: self.ncoef], gV, lam

    def close(self):
        if getattr(self, "_t", None):
            _lib.lib().inqamd_td_destroy(self._t)
            self._t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TDCurrentLoss(torch.autograd.Function):
    """forward: RT propagation -> J(T).d; backward: the adjoint -> dL/dc, dL/dV."""

    @staticmethod
    def forward(ctx, coef, vloc, prop, psi0):
        prop.ham.set_scalar_potential(vloc.reshape(-1))
        loss, _ = prop.forward(coef.detach().cpu().numpy(), psi0)
        ctx.prop = prop
        ctx.forward_id = prop.forward_count
        ctx.cinfo = (coef.dtype, coef.device)
        ctx.vshape = vloc.shape
        return torch.tensor(loss, dtype=torch.float64, device=coef.device)

    @staticmethod
    def backward(ctx, g):
        # the propagation object keeps only its last forward's checkpoints
        if ctx.prop.forward_count != ctx.forward_id:
            raise RuntimeError(
                "td_current_loss: another forward ran on this propagation object after this "
                "loss's forward, so its checkpoints are gone; give each live loss its own "
                "propagation_adjoint or call backward before the next forward")
        gc, gV, _ = ctx.prop.backward(float(g))
        dt, dev = ctx.cinfo
        return (torch.from_numpy(np.array(gc)).to(device=dev, dtype=dt), gV.reshape(ctx.vshape),
                None, None)


def td_current_loss(coef, vloc, prop, psi0):
    return TDCurrentLoss.apply(coef, vloc, prop, psi0)

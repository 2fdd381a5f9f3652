"""Load-path loss and its adjoint gradient (pfsf_load_path_*), with a
torch.autograd.Function front end.

The island's load is ramped in `steps` equal increments; every state is
solved with the multigrid PCG solver and the loss sums
softplus(crushing flux - rcomp) over the free voxels of every state.  The
backward pass runs one adjoint solve per state with the same operator,
re-solving states from checkpoints kept every `checkpoint_every` steps.

With damage=True the damage of each state degrades the next state's
conductivity (kappa_t = kappa g(d), g(d) = (1 - 1e-3)(1 - d)^2 + 1e-3): after
every state the elastic energy density psi raises the history H = max(H, psi)
and one FP64 phase-field step advances d; the gradient then also runs
backwards through d and H.
"""
from __future__ import annotations

import ctypes as C

from . import capi
from .device import _dims, _lib, _p, _stream


def _desc(vtype, kappa, source, rcomp, steps, checkpoint_every, tol, max_iters, damage):
    return capi.LoadPathDesc(_dims(vtype), _p(vtype).value, _p(kappa).value, _p(source).value,
                             _p(rcomp).value, steps, checkpoint_every, tol, max_iters,
                             _stream().value, 1 if damage else 0)


def load_path_loss(vtype, kappa, source, rcomp, steps=50, checkpoint_every=10, tol=1e-6,
                   max_iters=2000, fp64=False, damage=False):
    """Loss of the load path (FP32 forward, or FP64 throughout); returns LoadPathInfo."""
    d = _desc(vtype, kappa, source, rcomp, steps, checkpoint_every, tol, max_iters, damage)
    info = capi.LoadPathInfo()
    capi.check("pfsf_load_path_loss",
               _lib().pfsf_load_path_loss(C.byref(d), 1 if fp64 else 0, C.byref(info)))
    return info


def load_path_grad(vtype, kappa, source, rcomp, steps=50, checkpoint_every=10, tol=1e-6,
                   max_iters=2000, damage=False):
    """(LoadPathInfo, dL/dkappa as a float64 device tensor shaped like kappa)."""
    import torch

    g = torch.empty(kappa.shape, dtype=torch.float64, device=kappa.device)
    d = _desc(vtype, kappa, source, rcomp, steps, checkpoint_every, tol, max_iters, damage)
    info = capi.LoadPathInfo()
    capi.check("pfsf_load_path_grad", _lib().pfsf_load_path_grad(C.byref(d), _p(g),
                                                                  C.byref(info)))
    return info, g


def _function():
    import torch

    class LoadPathLoss(torch.autograd.Function):
        """L(kappa) of the load path; backward returns dL/dkappa (adjoint)."""

        @staticmethod
        def forward(ctx, kappa, vtype, source, rcomp, steps=50, checkpoint_every=10, tol=1e-6,
                    damage=False):
            kappa = kappa.detach().contiguous()
            info = load_path_loss(vtype, kappa, source, rcomp, steps, checkpoint_every, tol,
                                  damage=damage)
            ctx.save_for_backward(kappa, vtype, source, rcomp)
            ctx.cfg = (steps, checkpoint_every, tol)
            ctx.damage = damage
            return torch.tensor(info.loss, dtype=torch.float64, device=kappa.device)

        @staticmethod
        def backward(ctx, grad_out):
            kappa, vtype, source, rcomp = ctx.saved_tensors
            _, g = load_path_grad(vtype, kappa, source, rcomp, *ctx.cfg, damage=ctx.damage)
            return (grad_out * g).to(kappa.dtype), None, None, None, None, None, None, None

    return LoadPathLoss


_FN = None


def loss(kappa, vtype, source, rcomp, steps=50, checkpoint_every=10, tol=1e-6, damage=False):
    """Differentiable in kappa: L.backward() fills kappa.grad with dL/dkappa."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(kappa, vtype, source, rcomp, steps, checkpoint_every, tol, damage)

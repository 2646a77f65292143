"""torch.autograd binding of the Hellmann-Feynman gradient kernels.

``exciton_energies(theta, base, n_exc)`` returns the lowest n_exc exciton
energies of ``base`` with the parameters theta = (hopping scale s, scissor,
r0, eps_env, exchange field h) applied, as a differentiable torch tensor;
``spectrum_misfit(theta, base, target)`` returns the spectrum misfit
L = sum_w (Im eps(w) - target(w))^2 with its gradient from
``bse_grad_spectrum``.
Forward runs the FP64 GPU path (bse_build + bse_eigh); backward contracts the
incoming gradient with dE/dtheta from ``bse_grad_energies`` (no derivative
matrix is stored).  torch is used for the autograd graph only; all compute
goes through the C ABI.
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .system import BSESystem


def apply_theta(base: BSESystem, theta) -> BSESystem:
    """System with theta = (s, scissor, r0, eps_env, h) applied to ``base``:
    H(R != 0) -> s H(R), h diag(+1, -1, ...) added to H(0)."""
    s, scissor, r0, eps_env, h = (float(x) for x in theta)
    model = base.model.scaled(hopping_scale=s, exchange_field=h)
    return base.with_params(model=model, scissor=scissor, r0=r0, eps_env=eps_env)


def default_theta(base: BSESystem):
    return np.array([1.0, base.scissor, base.r0, base.eps_env, 0.0])


def energies_and_gradients(ctx: _abi.Context, base: BSESystem, theta, n_exc: int,
                           precision: str = "fp64"):
    """(E[:n_exc], dE/dtheta (n_exc x 5)) at theta."""
    s = apply_theta(base, theta)
    ctx.set_system(s).build(precision)
    ctx.eigh(n_exc, refine=(precision != "fp64"))
    E = ctx.eigenvalues()[:n_exc]
    g = ctx.grad_energies(n_exc)
    # the kernel differentiates w.r.t. a relative scale of the loaded H(R != 0)
    # blocks (already multiplied by s): dE/ds = (dE/dlambda) / s
    g[:, 0] = g[:, 0] / float(theta[0])
    return E, g


def exciton_energies(theta, base: BSESystem, n_exc: int = 8, ctx: _abi.Context | None = None,
                     precision: str = "fp64"):
    import torch

    class _ExcitonEnergies(torch.autograd.Function):
        @staticmethod
        def forward(fctx, th):
            c = ctx if ctx is not None else _abi.Context(0)
            E, g = energies_and_gradients(c, base, th.detach().cpu().numpy().astype(np.float64),
                                          n_exc, precision)
            fctx.save_for_backward(torch.as_tensor(g, dtype=th.dtype))
            return torch.as_tensor(E, dtype=th.dtype)

        @staticmethod
        def backward(fctx, grad_E):
            (g,) = fctx.saved_tensors
            return grad_E.to(g.dtype) @ g

    return _ExcitonEnergies.apply(theta)


def spectrum_misfit_and_gradient(ctx: _abi.Context, base: BSESystem, theta, target,
                                 omegas=None, eta=None):
    """(L, dL/dtheta) of L = sum_w (Im eps(w) - target(w))^2 at theta (FP64 path,
    all eigenvectors)."""
    s = apply_theta(base, theta)
    om = s.omegas() if omegas is None else np.asarray(omegas, dtype=np.float64)
    et = s.eta if eta is None else float(eta)
    ctx.set_system(s).build("fp64")
    ctx.eigh(-1)
    L, g = ctx.grad_spectrum(om, et, target)
    g[0] = g[0] / float(theta[0])
    return L, g


def spectrum_misfit(theta, base: BSESystem, target, ctx: _abi.Context | None = None,
                    omegas=None, eta=None):
    """Differentiable spectrum misfit L(theta) (torch scalar)."""
    import torch

    class _SpectrumMisfit(torch.autograd.Function):
        @staticmethod
        def forward(fctx, th):
            c = ctx if ctx is not None else _abi.Context(0)
            L, g = spectrum_misfit_and_gradient(c, base, th.detach().cpu().numpy().astype(np.float64),
                                                target, omegas, eta)
            fctx.save_for_backward(torch.as_tensor(g, dtype=th.dtype))
            return torch.as_tensor(L, dtype=th.dtype)

        @staticmethod
        def backward(fctx, grad_L):
            (g,) = fctx.saved_tensors
            return grad_L.to(g.dtype) * g

    return _SpectrumMisfit.apply(theta)

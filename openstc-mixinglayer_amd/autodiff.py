"""The differentiated mixing-layer run as a torch.autograd.Function.

``momentum_thickness(amp, solver)`` maps the amplitudes of the Fourier-mode
velocity perturbation of a differentiation case (``cases.fourier_case``) to
the momentum thickness theta after ``case["n_steps"]`` SSP-RK3 steps.  The
forward pass is the library's checkpointed forward sweep on the GPU
(ostc_adjoint_forward), the backward pass its hand-written reverse sweep
(ostc_adjoint_backward), so ``theta.backward()`` fills ``amp.grad`` with
d theta / d amp without torch tracing anything.  torch is only the
interface: the amplitudes and theta are float64 host tensors (32 numbers),
the work runs on the GPU inside the library.  torch's own GPU runtime is not
used: this torch build ships a HIP runtime of its own, and a process that
has loaded the library's runtime does not get a device through torch's.  The reference has no automatic differentiation;
this is the build's own wrapper around its reverse mode.
"""
from __future__ import annotations

import importlib

import numpy as np


def _pkg():
    return importlib.import_module(__name__.rsplit(".", 1)[0])


def _function():
    import torch

    class MomentumThickness(torch.autograd.Function):
        @staticmethod
        def forward(ctx, amp, solver):
            case = solver.case
            a = amp.detach().to("cpu", dtype=torch.float64).numpy().copy()
            solver.setup_mixing_layer()
            solver.perturb_fourier(0, case["modes"], a)
            solver.initialize()
            theta = solver.adjoint_forward(case["n_steps"], case.get("checkpoint_every", 10), case["theta"])
            ctx.solver, ctx.a = solver, a
            return torch.tensor(theta, dtype=torch.float64, device=amp.device)

        @staticmethod
        def backward(ctx, grad_theta):
            g = ctx.solver.adjoint_backward(ctx.solver.case["modes"], ctx.a)
            g = torch.from_numpy(np.ascontiguousarray(g)).to(grad_theta.device)
            return g * grad_theta, None

    return MomentumThickness


_FN = None


def momentum_thickness(amp, solver):
    """theta(amp) for a Solver built from a differentiation case; differentiable
    with respect to ``amp`` (1-D float64 tensor of the mode amplitudes)."""
    global _FN
    if _FN is None:
        _FN = _function()
    if amp.dim() != 1 or amp.numel() != len(solver.case["modes"]["amp"]):
        raise ValueError("amp must hold one amplitude per Fourier mode of the case")
    if amp.device.type != "cpu":
        raise ValueError("amp must be a host tensor (the library runs the sweeps on the GPU itself)")
    return _FN.apply(amp, solver)

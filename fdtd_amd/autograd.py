"""torch.autograd binding of the adjoint, for checking gradients with torch.

`DesignObjective.apply(eps_r, sim, terms, n_steps, checkpoint_every)` returns
J = sum_t w_t |DFT(probe_t, freq_t)|^2 of an engine run with the design
region's relative permittivity set to `eps_r` (a 1-D tensor over the design
cells, x-fastest); its backward pass is the engine's discrete adjoint
(fdtd_adjoint_gradient), not a tape of the forward.  The objective is the one
the C ABI's fdtd_objective describes; the reference has no counterpart.

When eps_r requires a gradient the forward pass is the adjoint call itself
(one forward run, then the reverse sweep), and backward only scales the
stored gradient: the forward is not run twice.  Without a gradient it is a
plain forward run.
"""
import numpy as np
import torch

from . import api


class DesignObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r, sim, terms, n_steps, checkpoint_every=64):
        if not isinstance(sim, api.Simulation):
            raise TypeError("sim must be an fdtd_amd.api.Simulation with a design region")
        eps = np.ascontiguousarray(eps_r.detach().to("cpu", torch.float32).numpy())
        sim.set_design_eps(eps)
        if eps_r.requires_grad:
            J, g = sim.adjoint_gradient(list(terms), int(n_steps), int(checkpoint_every))
            ctx.grad = torch.from_numpy(g).to(device=eps_r.device, dtype=torch.float64)
        else:
            sim.run(n_steps)
            X = sim.dft()
            J = float(sum(w * abs(X[p, f]) ** 2 for p, f, w in terms))
            ctx.grad = None
        ctx.out_dtype = eps_r.dtype
        return torch.tensor(J, dtype=torch.float64, device=eps_r.device)

    @staticmethod
    def backward(ctx, grad_J):
        return (grad_J * ctx.grad).to(ctx.out_dtype), None, None, None, None


def design_objective(eps_r, sim, terms, n_steps, checkpoint_every=64):
    """J(eps_r) with the adjoint as its gradient (see DesignObjective)."""
    return DesignObjective.apply(eps_r, sim, terms, n_steps, checkpoint_every)

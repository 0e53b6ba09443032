"""torch.autograd wrapper of the adjoint transient.

`TransientObjective.apply(values, ctx_args...)` returns J for every instance;
its backward calls the hand-written adjoint kernel (kg_adjoint_backward),
which reuses the stored LU factors for transposed solves and recomputes
states from checkpoints.
"""
import numpy as np
import torch

from . import _lib
from .solver import Context, SolverConfig


class TransientObjective(torch.autograd.Function):
    """J(values) = sum_s w_s c.x_s for a fixed-step linear transient.

    values: [n_instances, n_elements] primary element values (R, C, L, ...)
    on any device; the parameter records are rebuilt from `params_template`
    with these values in slot 0.
    """

    @staticmethod
    def forward(fctx, values, ctx: Context, elements, params_template, n_unknowns, h, nsteps, cvec,
                method=_lib.TRAP, precision=_lib.FP64, checkpoint_every=64):
        vals = values.detach().to("cpu", torch.float64).numpy()
        params = np.array(params_template, dtype=np.float64, copy=True)
        params[:, :, 0] = vals
        ctx.load(elements, params, n_unknowns)
        J, st = ctx.adjoint_forward(h, nsteps, cvec, method=method, precision=precision,
                                    checkpoint_every=checkpoint_every, config=SolverConfig(compat=False))
        fctx.kg = ctx
        fctx.device = values.device
        return torch.as_tensor(J, dtype=values.dtype, device=values.device)

    @staticmethod
    def backward(fctx, gJ):
        g, st = fctx.kg.adjoint_backward()
        g = torch.as_tensor(g, dtype=gJ.dtype, device=fctx.device)
        return (gJ[:, None] * g,) + (None,) * 10

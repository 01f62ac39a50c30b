"""torch.autograd.Function over the HIP path: forward = RESIDUALS at the points x, backward =
J^T grad from the adjoint Jacobian (the reverse-mode trace), so residual vectors of optical
systems can sit inside a torch graph.

The reference exposes the same two operations to Python as TraceEngine.run(RESIDUALS) and the
adjoint Jacobian (python/src/bind_trace.cpp:411-623, src/optim/adjoint_jacobian.cpp:370); the
torch wrapper is new. Values cross the C-ABI as float64 host arrays; the result tensors are
placed on x's device."""

import numpy as np
import torch

from .batch import residual_rows


class TraceResiduals(torch.autograd.Function):
    """r = TraceResiduals.apply(x, engine, base, settings, variables)

    x: [runs][P] float64 tensor of variable values; base: SystemBatch with one variant (shared
    template) or one per run; returns r [runs][R] float64 (trace rows, then the settings'
    constraint rows). backward: grad_x[i] = J_i^T grad_r[i] with the FP64 adjoint Jacobian
    J_i [P][R] at x_i."""

    @staticmethod
    def forward(ctx, x, engine, base, settings, variables):
        if x.dtype != torch.float64:
            raise TypeError("x must be float64")
        xn = np.ascontiguousarray(x.detach().to("cpu").numpy())
        runs = xn.shape[0]
        if base.variants not in (1, runs):
            raise ValueError("base must have one variant or one per run")
        pts = base.take(np.zeros(runs, dtype=np.int64)) if base.variants == 1 else base.copy()
        variables.apply(pts, xn)
        r, _, _ = engine.residuals(pts, settings)
        ctx.save_for_backward(x)
        ctx.args = (engine, base, settings, variables)
        return torch.from_numpy(r).to(x.device)

    @staticmethod
    def backward(ctx, grad_r):
        (x,) = ctx.saved_tensors
        engine, base, settings, variables = ctx.args
        xn = np.ascontiguousarray(x.detach().to("cpu").numpy())
        J, _, _, _ = engine.jacobian(base, settings, variables, xn, method="adjoint",
                                     precision="fp64")
        gr = grad_r.detach().to("cpu", torch.float64).numpy()
        assert gr.shape == (xn.shape[0], residual_rows(base, settings))
        gx = np.matmul(J, gr[:, :, None])[:, :, 0]
        return torch.from_numpy(gx).to(x.device), None, None, None, None


def trace_residuals(x, engine, base, settings, variables):
    """Functional form of TraceResiduals."""
    return TraceResiduals.apply(x, engine, base, settings, variables)

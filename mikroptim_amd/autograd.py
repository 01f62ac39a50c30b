"""torch.autograd.Function over the HIP path: forward = RESIDUALS at the points x, backward =
J^T grad_r by the reverse-mode sweep seeded with grad_r (mko_vjp_device), so residual vectors of
optical systems can sit inside a torch graph.

The reference exposes the two operations to Python as TraceEngine.run(RESIDUALS) and the adjoint
Jacobian (python/src/bind_trace.cpp:411-623, src/optim/adjoint_jacobian.cpp:370); the torch
wrapper is new. For CUDA (HIP) tensors everything stays in device memory: the template batch is
uploaded once per forward, the points x are applied on the device (mko_point_batch_device), the
residuals are written into a torch tensor (mko_residuals_device) and the backward forms J^T grad
without J (mko_vjp_device), all on torch's current stream. CPU tensors, constraint rows and glass
variables take the host route: residuals and the FP64 adjoint Jacobian through the engine, the
product in numpy."""

import ctypes as C

import numpy as np
import torch

from .batch import residual_rows
from .engine import library

_dp = C.POINTER(C.c_double)


def _ptr(t, ctype=C.c_double):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(ctype))


def _check(rc):
    if rc != 0:
        raise RuntimeError(library().mko_last_error().decode())


class _DevicePoints:
    """Per-run points of a template batch in device memory (torch tensors): the template's
    per-variant arrays uploaded, the variables' type / surface arrays uploaded, and the point
    batch written by mko_point_batch_device."""

    def __init__(self, base, variables, x, stream):
        dev = x.device
        runs = x.shape[0]
        f64 = dict(dtype=torch.float64, device=dev)
        self.tmpl = {}
        for name in ("curvature", "thickness", "index", "pupil_radius", "field_angle",
                     "field_target", "reference_curvature"):
            a = getattr(base, name)
            if a is not None:
                self.tmpl[name] = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
        self.vtype = torch.from_numpy(np.ascontiguousarray(variables.type, dtype=np.int32)).to(dev)
        self.vsurf = torch.from_numpy(
            np.ascontiguousarray(variables.batch_surface, dtype=np.int32)).to(dev)
        S, L, F = base.surfaces, base.wavelengths, base.fields
        self.pts = {"curvature": torch.empty((runs, S), **f64),
                    "thickness": torch.empty((runs, S), **f64),
                    "index": torch.empty((runs, S, L), **f64),
                    "pupil_radius": torch.empty((runs,), **f64),
                    "field_angle": torch.empty((runs, F), **f64),
                    "field_target": torch.empty((runs, F), **f64),
                    "reference_curvature": (torch.empty((runs, F), **f64)
                                            if base.reference_curvature is not None else None)}
        ctmpl = base.c_struct({k: v.data_ptr() for k, v in self.tmpl.items()})
        self.vars_host = variables.c_struct()
        vdev = variables.c_struct()
        vdev.type = _ptr(self.vtype, C.c_int32)
        vdev.batch_surface = _ptr(self.vsurf, C.c_int32)
        outs = [self.pts[k] for k in ("curvature", "thickness", "index", "pupil_radius",
                                      "field_angle", "field_target", "reference_curvature")]
        _check(library().mko_point_batch_device(
            C.byref(ctmpl), int(base.variants > 1), runs, 1, C.byref(vdev),
            C.c_void_p(x.data_ptr()), None, None,
            *[C.c_void_p(o.data_ptr()) if o is not None else None for o in outs], None,
            C.c_void_p(stream)))
        self.batch = base.c_struct({k: v.data_ptr() for k, v in self.pts.items() if v is not None})
        self.batch.variants = runs


class TraceResiduals(torch.autograd.Function):
    """r = TraceResiduals.apply(x, engine, base, settings, variables)

    x: [runs][P] float64 tensor of variable values; base: SystemBatch with one variant (shared
    template) or one per run; returns r [runs][R] float64 on x's device (trace rows, then the
    settings' constraint rows). backward: grad_x[i] = J_i^T grad_r[i], J_i [R][P] the FP64
    adjoint Jacobian at x_i."""

    @staticmethod
    def forward(ctx, x, engine, base, settings, variables):
        if x.dtype != torch.float64:
            raise TypeError("x must be float64")
        runs = x.shape[0]
        if base.variants not in (1, runs):
            raise ValueError("base must have one variant or one per run")
        ctx.args = (engine, base, settings, variables)
        device = (x.is_cuda and settings.extra_rows == 0 and not variables.has_glass and
                  settings.precision == 0)
        ctx.device = device
        if device:
            xc = x.detach().contiguous()
            stream = torch.cuda.current_stream(x.device).cuda_stream
            pts = _DevicePoints(base, variables, xc, stream)
            R = residual_rows(base, settings)
            r = torch.empty((runs, R), dtype=torch.float64, device=x.device)
            cs = settings.c_struct()
            _check(library().mko_residuals_device(C.byref(pts.batch), C.byref(cs), _ptr(r), R,
                                                  None, None, C.c_void_p(stream)))
            ctx.pts, ctx.cs, ctx.R = pts, cs, R
            ctx.save_for_backward(x)
            return r
        xn = np.ascontiguousarray(x.detach().to("cpu").numpy())
        pts = base.take(np.zeros(runs, dtype=np.int64)) if base.variants == 1 else base.copy()
        variables.apply(pts, xn)
        r, _, _ = engine.residuals(pts, settings)
        ctx.save_for_backward(x)
        return torch.from_numpy(r).to(x.device)

    @staticmethod
    def backward(ctx, grad_r):
        (x,) = ctx.saved_tensors
        engine, base, settings, variables = ctx.args
        runs, P = x.shape
        if ctx.device:
            lib = library()
            g = grad_r.detach().to(torch.float64).contiguous()
            stream = torch.cuda.current_stream(x.device).cuda_stream
            out = torch.empty((runs, P), dtype=torch.float64, device=x.device)
            nbytes = lib.mko_vjp_workspace(runs, base.fields, base.wavelengths, P)
            ws = torch.empty((max(int(nbytes), 1),), dtype=torch.uint8, device=x.device)
            _check(lib.mko_vjp_device(C.byref(ctx.pts.batch), C.byref(ctx.cs),
                                      C.byref(ctx.pts.vars_host), C.c_void_p(g.data_ptr()),
                                      ctx.R, C.c_void_p(out.data_ptr()),
                                      C.c_void_p(ws.data_ptr()), nbytes, C.c_void_p(stream)))
            return out, None, None, None, None
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

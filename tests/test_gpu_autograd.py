"""The torch.autograd.Function over the HIP path on device tensors: forward = RESIDUALS at the
points x applied on the device (bitwise the engine's RESIDUALS of the same points), backward =
J^T grad formed by the reverse sweep seeded with the cotangents (mko_vjp_device), never forming
J. Checked against the stored FP64 adjoint Jacobian's product J^T grad per run (kAdjointCpuGpu,
1e-10 relative): on the 24-surface NA 0.95 objective (16 runs, 17920 ray rows, 48 variables, hard
failure) and on the singlet in soft failure mode with failing rays (penalty rows)."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from test_gpu_parity import ADJ_TOL

pytestmark = pytest.mark.gpu


def _check_vjp(engine, base, s, v, x0, seed):
    import torch
    from mikroptim_amd.autograd import trace_residuals

    x = torch.tensor(x0, dtype=torch.float64, device="cuda", requires_grad=True)
    r = trace_residuals(x, engine, base, s, v)
    assert r.is_cuda
    pts = base.take(np.zeros(len(x0), dtype=np.int64))
    v.apply(pts, x0)
    res, _, _ = engine.residuals(pts, s)
    assert np.array_equal(r.detach().cpu().numpy(), res)
    gen = np.random.default_rng(seed)
    grad = gen.standard_normal(res.shape)
    r.backward(torch.tensor(grad, device="cuda"))
    gx = x.grad.cpu().numpy()
    J, _, _, _ = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    ref = np.einsum("npr,nr->np", J, grad)
    rel = np.linalg.norm(gx - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() <= ADJ_TOL, rel.max()
    return gx, ref


def test_vjp_on_the_objective(engine):
    b, v = systems.objective(variants=64, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    _check_vjp(engine, b.take([0]), s, v, v.extract(b)[:16], seed=1)


def test_vjp_soft_failure_singlet(engine):
    b, v = systems.singlet(16)
    s = systems.singlet_settings(abi.FAILURE_SOFT)
    x0 = v.extract(b)[:8].copy()
    x0[:, 0] = np.linspace(0.19, 0.24, 8)  # front radius near the pupil radius: rays fail softly
    pts = b.take(np.zeros(8, dtype=np.int64))
    v.apply(pts, x0)
    _, _, st = engine.residuals(pts, s)
    assert (st != 0).any()
    _check_vjp(engine, b.take([0]), s, v, x0, seed=2)


def test_vjp_per_run_template_and_run_independence(engine):
    """One template per run; a run's product does not depend on the other runs."""
    import torch
    from mikroptim_amd.autograd import trace_residuals

    b, v = systems.objective(variants=64, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    s.rings, s.spokes = 4, 8
    base = b.take(np.arange(6))
    x0 = v.extract(b)[:6]
    gx, _ = _check_vjp(engine, base, s, v, x0, seed=3)
    x = torch.tensor(x0[2:3], dtype=torch.float64, device="cuda", requires_grad=True)
    r = trace_residuals(x, engine, base.take([2]), s, v)
    grad = np.random.default_rng(3).standard_normal((6, r.shape[1]))[2:3]
    r.backward(torch.tensor(grad, device="cuda"))
    assert np.array_equal(x.grad.cpu().numpy()[0], gx[2])

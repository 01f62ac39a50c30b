"""Soft failure mode with the mixed-precision adjoint (the reference CLI's default failure mode,
src/cli/optimize_command.cpp:1566-1568). A run whose point has no failed ray has zero penalty
residuals and zero penalty derivatives, so its normal equations are hard mode's and come from the
fused kernel. A run with failed rays takes the FP64 stored-J path, which has the penalty rows
(adjoint_jacobian.cpp). Checked bit for bit against those two paths called directly, and the LM
with the device control against the host control and against the reference's own control
(through HipLmBackend)."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from oracle_lib import reference
from test_gpu_configs import U_BF, U_F

pytestmark = pytest.mark.gpu


def _mixed_batch():
    """Six singlet runs; runs 1 and 4 with a 60 mm pupil, so their outer rays miss the first
    surface (radius 50 mm) and fail softly."""
    b, v = systems.singlet(6, seed=4)
    b.pupil_radius[[1, 4]] = 60.0
    return b, v


def test_soft_mixed_normal_equations_follow_the_failed_rays(engine):
    b, v = _mixed_batch()
    soft = systems.singlet_settings(abi.FAILURE_SOFT)
    hard = systems.singlet_settings(abi.FAILURE_HARD)
    x = v.extract(b)
    _, f, _ = engine.merit(b, soft)
    assert f[1] > 0 and f[4] > 0 and np.count_nonzero(f) == 2
    _, _, a, g = engine.jacobian(b, soft, v, x, method="adjoint", precision="mixed",
                                 want_jacobian=False)
    clean = np.flatnonzero(f == 0)
    failed = np.flatnonzero(f > 0)
    _, _, ah, gh = engine.jacobian(b.take(clean), hard, v, x[clean], method="adjoint",
                                   precision="mixed", want_jacobian=False)
    assert np.array_equal(a[clean], ah) and np.array_equal(g[clean], gh)
    _, _, a64, g64 = engine.jacobian(b.take(failed), soft, v, x[failed], method="adjoint",
                                     precision="fp64", want_jacobian=False)
    assert np.array_equal(a[failed], a64) and np.array_equal(g[failed], g64)


def test_soft_mixed_lm_device_control_equals_host_control(engine):
    b, v = _mixed_batch()
    s = systems.singlet_settings(abi.FAILURE_SOFT)
    x0 = v.extract(b)
    kw = dict(max_iterations=10, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rd = engine.optimize(b, s, v, x0, abi.lm_options(**kw), precision="mixed")
    rh = engine.optimize(b, s, v, x0, abi.lm_options(host_control=True, **kw), precision="mixed")
    assert np.array_equal(rd.x, rh.x) and np.array_equal(rd.merit, rh.merit)
    assert rd.status == rh.status and rd.cycles == rh.cycles
    assert [r.penalty_rounds for r in rd.runs] == [r.penalty_rounds for r in rh.runs]
    ref = reference()
    if ref is not None:
        x, runs, cycles = ref.optimize_lm_hip_backend(engine, b, s, v, x0, abi.lm_options(**kw),
                                                      "mixed")
        assert np.array_equal(x, rh.x) and cycles == rh.cycles
        assert [r.merit for r in runs] == list(rh.merit)


def test_soft_mixed_failure_free_objective_is_hard_mode(engine):
    """On a failure-free workload the soft-mode LM with the mixed adjoint takes the same steps as
    the hard-mode LM (same normal equations; the penalty rows stay zero)."""
    b, v = systems.objective(variants=16, seed=2)
    base, x0 = b.take([0]), v.extract(b)[:3]
    kw = dict(max_iterations=4, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    soft = systems.objective_settings(abi.FAILURE_SOFT)
    soft.rings, soft.spokes = 8, 16
    hard = systems.objective_settings(abi.FAILURE_HARD)
    hard.rings, hard.spokes = 8, 16
    rs = engine.optimize(base, soft, v, x0, abi.lm_options(**kw), precision="mixed")
    rh = engine.optimize(base, hard, v, x0, abi.lm_options(**kw), precision="mixed")
    assert all(r.failed_rays == 0 for r in rs.runs)
    assert np.array_equal(rs.x, rh.x) and np.array_equal(rs.merit, rh.merit)


def test_soft_split_with_constraint_rows(engine):
    """Soft split with extrema and paraxial constraint rows: the failing runs' normal equations
    are the FP64 soft path's bit for bit; the clean runs' (fused trace rows, constraint rows by
    central differences under the soft settings) are the FP64 soft path's within the
    mixed-precision bound of their trace rows."""
    b, v = _mixed_batch()
    s = systems.singlet_settings(abi.FAILURE_SOFT)
    s.constraints = systems.singlet_constraints()
    x = v.extract(b)
    _, f, _ = engine.merit(b, s)
    failed, clean = np.flatnonzero(f > 0), np.flatnonzero(f == 0)
    assert len(failed) == 2
    _, _, a, g = engine.jacobian(b, s, v, x, method="adjoint", precision="mixed",
                                 want_jacobian=False)
    J, r0, a64, g64 = engine.jacobian(b, s, v, x, method="adjoint", precision="fp64")
    assert np.array_equal(a[failed], a64[failed]) and np.array_equal(g[failed], g64[failed])
    nJ = np.linalg.norm(J.reshape(len(x), -1), axis=1)[clean]
    nr = np.linalg.norm(r0, axis=1)[clean]
    eps_a = 3 * U_BF ** 2 + 512 * U_F + 2 * U_F
    assert np.all(np.linalg.norm(a[clean] - a64[clean], axis=(1, 2)) <= eps_a * nJ ** 2)
    assert np.all(np.linalg.norm(g[clean] - g64[clean], axis=1) <= 4 * U_F * nJ * nr)


def test_soft_graph_cycles_equal_eager_cycles(engine):
    """The captured soft-mode cycle (device split of the Jacobian list, fused kernel for the runs
    without failed rays, FP64 stored J for the others) takes the eager cycle's steps bit for bit,
    through the penalty rounds, and replays while runs have failed rays."""
    from mikroptim_amd.engine import TraceEngine

    b, v = _mixed_batch()
    s = systems.singlet_settings(abi.FAILURE_SOFT)
    x0 = v.extract(b)
    kw = dict(max_iterations=8, jacobian_method=abi.JACOBIAN_ADJOINT)
    g = engine.lm(b, s, v, x0, abi.lm_options(**kw), precision="mixed")
    other = TraceEngine(0)
    try:
        e = other.lm(b, s, v, x0, abi.lm_options(graph=False, **kw), precision="mixed")
        while True:
            more_g, more_e = g.cycle(), e.cycle()
            assert more_g == more_e
            rg, re_ = g.results(), e.results()
            assert np.array_equal(rg.x, re_.x) and np.array_equal(rg.merit, re_.merit)
            if not more_g:
                break
        assert g.graph_cycles() > 0 and e.graph_cycles() == 0
        assert any(r.penalty_rounds > 0 for r in rg.runs)
        e.close()
    finally:
        g.close()
        other.close()

"""The autotuner (mko_engine_autotune): it picks a valid launch shape for the MERIT trace and the
fused normal equations from timings on the caller's points, caches the choice per shape, and the
choice changes no FP64 result (tuple packs are bitwise equal to unpacked traces); LM sessions keep
the tuning the engine had at their creation; optimize_groups(autotune=True) restores the engine's
own tuning afterwards."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems

pytestmark = pytest.mark.gpu


def _group(index, variants=6, failure=abi.FAILURE_HARD):
    b, v = systems.uneven_topology(index, variants=variants)
    s = systems.uneven_settings(failure)
    return b.take([0]), s, v, v.extract(b)


def test_autotune_choice_is_valid_and_cached(engine):
    base, s, v, x0 = _group(1)
    before = engine.tuning()
    try:
        pack, waves = engine.autotune(base, s, v, x0)
        assert pack in (32, 64, 128, 256)
        assert waves in (1, 2, 4)
        assert engine.tuning() == (pack, waves)
        engine.tuning(pack_rays=0, fused_waves=0)
        # the same shape again: the cached choice, nothing timed
        assert engine.autotune(base, s, v, x0) == (pack, waves)
    finally:
        engine.tuning(*before)


def test_autotuned_fp64_results_equal_untuned(engine):
    base, s, v, x0 = _group(2)
    opt = abi.lm_options(max_iterations=5, jacobian_method=abi.JACOBIAN_ADJOINT)
    before = engine.tuning()
    ref_m = engine.point_merit(base, s, v, x0)
    ref = engine.optimize(base, s, v, x0, opt, precision="fp64")
    try:
        engine.autotune(base, s, v, x0)
        m = engine.point_merit(base, s, v, x0)
        out = engine.optimize(base, s, v, x0, opt, precision="fp64")
    finally:
        engine.tuning(*before)
    for a, b in zip(m, ref_m):
        assert np.array_equal(a, b)
    assert np.array_equal(out.x, ref.x) and np.array_equal(out.merit, ref.merit)


def test_sessions_keep_their_creation_tuning(engine):
    """A session created under one tuning cycles under it after the engine's tuning changed: its
    mixed-precision steps equal those of a session that ran with that tuning throughout."""
    base, s, v, x0 = _group(4)
    opt = abi.lm_options(max_iterations=4, jacobian_method=abi.JACOBIAN_ADJOINT, graph=False)
    before = engine.tuning()
    try:
        engine.tuning(pack_rays=256, fused_waves=2)
        ref = engine.optimize(base, s, v, x0, opt, precision="mixed")
        sess = engine.lm(base, s, v, x0, opt, precision="mixed")
        try:
            engine.tuning(pack_rays=32, fused_waves=4)
            while sess.cycle():
                pass
            out = sess.results()
        finally:
            sess.close()
    finally:
        engine.tuning(*before)
    assert np.array_equal(out.x, ref.x) and np.array_equal(out.merit, ref.merit)


def test_optimize_groups_autotune(engine):
    groups = [_group(i, variants=4) for i in (0, 3, 6)]
    opt = abi.lm_options(max_iterations=6, jacobian_method=abi.JACOBIAN_ADJOINT)
    before = engine.tuning()
    ref = engine.optimize_groups(groups, opt, precision="fp64")
    out = engine.optimize_groups(groups, opt, precision="fp64", autotune=True)
    assert engine.tuning() == before
    for a, b in zip(out, ref):
        assert np.array_equal(a.x, b.x) and np.array_equal(a.merit, b.merit)
    mixed = engine.optimize_groups(groups, opt, precision="mixed", autotune=True)
    for (base, s, v, x0), r in zip(groups, mixed):
        m0, _, _ = engine.point_merit(base, s, v, x0)
        assert np.all(r.merit <= m0)

"""The LM cycle as a captured hipGraph (mko_lm_options.host_control = 0, device control with the
fused mixed-precision adjoint): bit-for-bit equal to the same LM launched kernel by kernel, cycle
by cycle, with the program's own eager engine work and synchronizations between the replays;
the graph is really replayed (mko_lm_graph_cycles), and LMs it does not cover run eagerly."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd.engine import TraceEngine

pytestmark = pytest.mark.gpu


def _objective(runs=4, per_run_template=False):
    b, v = systems.objective(variants=16, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    s.rings, s.spokes = 4, 6
    x0 = v.extract(b)[:runs]
    base = b.take(np.arange(runs)) if per_run_template else b.take([0])
    return base, v, s, x0


def _state(sess):
    r = sess.results()
    return r.x.copy(), r.merit.copy(), list(r.status), r.iterations.copy(), r.cycles


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2] and np.array_equal(a[3], b[3]) and a[4] == b[4]


@pytest.mark.parametrize("per_run_template", [False, True])
def test_graph_cycles_equal_eager_cycles(engine, per_run_template):
    base, v, s, x0 = _objective(4, per_run_template)
    kw = dict(max_iterations=12, jacobian_method=abi.JACOBIAN_ADJOINT)
    g = engine.lm(base, s, v, x0, abi.lm_options(**kw), precision="mixed")
    other = TraceEngine(0)
    try:
        e = other.lm(base, s, v, x0, abi.lm_options(graph=False, **kw), precision="mixed")
        side_b, _ = systems.singlet(8, seed=3)
        side_s = systems.singlet_settings()
        cycles = 0
        while True:
            more_g = g.cycle()
            more_e = e.cycle()
            assert more_g == more_e
            # eager work on the graph's engine between replays (its buffers may grow)
            engine.merit(side_b, side_s)
            engine.residuals(side_b.take(np.arange(4)), side_s)
            _same(_state(g), _state(e))
            cycles += 1
            if not more_g:
                break
        assert cycles > 3
        assert g.graph_cycles() > 0 and e.graph_cycles() == 0
        e.close()
    finally:
        other.close()
        g.close()


def test_graph_lm_equals_host_control(engine):
    base, v, s, x0 = _objective(3)
    kw = dict(max_iterations=8, jacobian_method=abi.JACOBIAN_ADJOINT, threads=4)
    rd = engine.optimize(base, s, v, x0, abi.lm_options(**kw), precision="mixed")
    rh = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw),
                         precision="mixed")
    assert np.array_equal(rd.x, rh.x) and np.array_equal(rd.merit, rh.merit)
    assert rd.status == rh.status and rd.cycles == rh.cycles


def test_graph_kernel_timing_counts_only_launches_with_work(engine):
    base, v, s, x0 = _objective(2)
    sess = engine.lm(base, s, v, x0, abi.lm_options(max_iterations=6,
                                                     jacobian_method=abi.JACOBIAN_ADJOINT),
                     precision="mixed")
    sess.cycle()
    engine.kernel_timing(True)
    n = 0
    while sess.cycle():
        n += 1
    kt = engine.kernel_timing(False)
    assert sess.graph_cycles() > 0
    res = sess.results()
    sess.close()
    ms, launches = kt["adjoint"]
    assert launches <= n + 1 and (launches == 0 or ms > 0.0)
    assert kt["trace"][1] <= n + 1
    assert all(r.iterations <= 6 for r in res.runs)


@pytest.mark.parametrize("case", ["fd", "rows", "soft"])
def test_lms_outside_the_graph_run_eagerly(engine, case):
    """FD Jacobians and constraint rows keep the eager cycle; soft failure mode with the mixed
    adjoint replays the captured cycle (its Jacobian list split on the device)."""
    from mikroptim_amd import batch as bt
    base, v, s, x0 = _objective(2)
    method, precision = abi.JACOBIAN_ADJOINT, "mixed"
    if case == "fd":
        method, precision = abi.JACOBIAN_FD, "fp64"
    elif case == "soft":
        s.failure = abi.FAILURE_SOFT
    else:
        s.constraints = [bt.focal_length(50.0, 1e-3)]
    sess = engine.lm(base, s, v, x0, abi.lm_options(max_iterations=3, jacobian_method=method),
                     precision=precision)
    while sess.cycle():
        pass
    if case == "soft":
        assert sess.graph_cycles() > 0
    else:
        assert sess.graph_cycles() == 0
    sess.close()


def test_uneven_groups_concurrent_equal_sequential(engine):
    """optimize_groups begins every group's cycle before waiting for any: the captured cycles of
    the groups run concurrently on their own streams; each group's result equals optimize() on
    it bit for bit, graph groups (hard and soft failure mode) and eager groups (constraint rows)
    alike."""
    from mikroptim_amd import batch as bt
    groups = []
    for i in range(6):
        b, v = systems.uneven_topology(i, variants=3)
        # group 4 in soft failure mode (captured cycle with the device split); group 5 with a
        # constraint row: outside the graph, cycled eagerly
        s = systems.uneven_settings(abi.FAILURE_SOFT if i == 4 else abi.FAILURE_HARD)
        if i == 5:
            s.constraints = [bt.focal_length(50.0, 1e-3)]
        groups.append((b.take([0]), s, v, v.extract(b)))
    opt = abi.lm_options(max_iterations=6, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    together = engine.optimize_groups(groups, opt, precision="mixed")
    for (base, s, v, x0), r in zip(groups, together):
        alone = engine.optimize(base, s, v, x0, opt, precision="mixed")
        assert np.array_equal(r.x, alone.x) and np.array_equal(r.merit, alone.merit)
        assert r.status == alone.status and r.cycles == alone.cycles
    # the eligible groups' cycles were graph replays
    sessions = [engine.lm(b, s, v, x0, opt, precision="mixed") for b, s, v, x0 in groups]
    active = list(sessions)
    while active:
        begun = [(sess, sess.cycle_begin()) for sess in active]
        active = [sess for sess, rc in begun if rc != 0 and sess.cycle_end()]
    for (base, s, v, x0), sess in zip(groups, sessions):
        assert (sess.graph_cycles() > 0) == (not s.constraints)
        sess.close()

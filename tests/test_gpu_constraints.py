"""Parity of the constraint rows on the HIP path (C-ABI of libmikroptim_hip.so) with the CPU
oracle: surface extrema of the MERIT / RESIDUALS traces, the rows in residuals, merit, FD and
adjoint Jacobians, the fused mixed-precision normal equations and the LM.

Rows that need no extrema are computed on the device with the oracle's operation order and
correctly rounded division / square root, so they are compared bit for bit (values and their
central differences). Rows from extrema inherit the trace's tolerance (1e-12 mm per ray
height); their central differences divide that by the FD step (~6e-6 relative)."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd import batch as bt
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

H_TOL = 1e-12    # mm, largest ray height (kCpuGpuFp64 per ray)
C_TOL = 1e-13    # smallest cos I


def _singlet(n=6, seed=7, failure=abi.FAILURE_HARD):
    b, v = systems.singlet(n, seed=seed)
    s = systems.singlet_settings(failure)
    s.constraints = systems.singlet_constraints()
    return b, v, s


def _objective(n=3):
    b, v = systems.objective(variants=16, seed=2)
    b = b.take(np.arange(n))
    s = systems.objective_settings(abi.FAILURE_HARD)
    s.rings, s.spokes = 4, 6
    # limits just inside the variant-0 extrema, so the rows from extrema are active
    s.constraints = systems.objective_constraints(b, oracle().surface_extrema(b, s, threads=8))
    return b, v, s


def _split(s):
    E = sum(not c.needs_extrema for c in s.constraints)
    return E, len(s.constraints)


def _rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


# ---- surface extrema -----------------------------------------------------------------------

@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_extrema_match_oracle(engine, failure):
    b, _ = systems.singlet(8, seed=5)
    b.pupil_radius[3] = 1e4    # no valid ray: neutral extrema
    b.pupil_radius[5] = 110.0  # outer rings fail
    s = systems.singlet_settings(failure)
    e = engine.surface_extrema(b, s)
    eo = oracle().surface_extrema(b, s)
    assert np.all(e[3, :, 0] == 0.0) and np.all(e[3, :, 1] == 1.0)
    assert np.max(np.abs(e[..., 0] - eo[..., 0])) <= H_TOL
    assert np.max(np.abs(e[..., 1] - eo[..., 1])) <= C_TOL


def test_extrema_objective_and_df64(engine):
    b, _, s = _objective(3)
    e = engine.surface_extrema(b, s)
    eo = oracle().surface_extrema(b, s, threads=8)
    assert np.max(np.abs(e[..., 0] - eo[..., 0])) <= H_TOL
    assert np.max(np.abs(e[..., 1] - eo[..., 1])) <= C_TOL
    s.precision = abi.PRECISION_DF64
    ed = engine.surface_extrema(b, s)
    assert np.max(np.abs(ed[..., 0] - eo[..., 0])) <= 1e-9
    assert np.max(np.abs(ed[..., 1] - eo[..., 1])) <= 1e-10


# ---- rows in residuals and merit ---------------------------------------------------------------

@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_residual_rows_match_oracle(engine, failure):
    b, _, s = _singlet(8, failure=failure)
    plain = systems.singlet_settings(failure)
    res, opd, st = engine.residuals(b, s)
    res0, opd0, st0 = engine.residuals(b, plain)
    T = res0.shape[1]
    assert np.array_equal(res[:, :T], res0) and np.array_equal(st, st0)
    rows = res[:, T:]
    ro = oracle().constraint_rows(b, s)
    E, _ = _split(s)
    assert np.array_equal(rows[:, :E], ro[:, :E])          # paraxial / geometric rows
    assert np.max(np.abs(rows[:, E:] - ro[:, E:])) <= 1e-10  # rows from extrema


@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_merit_with_rows(engine, failure):
    b, _, s = _singlet(8, failure=failure)
    m, f, h = engine.merit(b, s)
    mo, fo, ho = oracle().merit(b, s)
    assert np.array_equal(f, fo) and np.array_equal(h, ho)
    assert _rel(m, mo) <= 1e-9
    res, _, _ = engine.residuals(b, s)
    assert _rel(m, (res ** 2).sum(1)) <= 1e-13
    extra = np.linspace(0.5, 2.0, b.variants)
    me, _, _ = engine.merit(b, s, extra=extra)
    assert _rel(me, m + extra) <= 1e-14


# ---- rows in the Jacobians ---------------------------------------------------------------------

@pytest.mark.parametrize("method", ["fd", "adjoint"])
@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_jacobian_rows_match_oracle(engine, method, failure):
    b, v, s = _singlet(6, failure=failure)
    x0 = v.extract(b)[:3]
    base = b.take([0])
    J, r0, a, g = engine.jacobian(base, s, v, x0, method=method)
    fn = oracle().fd_jacobian if method == "fd" else oracle().adjoint_jacobian
    Jo, r0o, _ = fn(base, s, v, x0)
    E, n = _split(s)
    T = J.shape[2] - n
    assert np.array_equal(J[:, :, T:T + E], Jo[:, :, T:T + E])
    assert np.array_equal(r0[:, T:T + E], r0o[:, T:T + E])
    assert np.max(np.abs(r0[:, T + E:] - r0o[:, T + E:])) <= 1e-10
    # derivatives of rows from extrema: value differences / FD step
    err = np.abs(J[:, :, T + E:] - Jo[:, :, T + E:])
    assert err.max() <= 1e-5 * max(1.0, np.abs(Jo[:, :, T + E:]).max())
    ao, go = oracle().normal_equations(J, r0)
    assert np.array_equal(a, ao) and np.array_equal(g, go)


def test_objective_jacobian_rows(engine):
    b, v, s = _objective(2)
    x0 = v.extract(b)[:2]
    base = b.take([0])
    for method in ("fd", "adjoint"):
        J, r0, _, _ = engine.jacobian(base, s, v, x0, method=method)
        fn = oracle().fd_jacobian if method == "fd" else oracle().adjoint_jacobian
        Jo, r0o, _ = fn(base, s, v, x0, threads=8)
        E, n = _split(s)
        T = J.shape[2] - n
        assert np.array_equal(J[:, :, T:T + E], Jo[:, :, T:T + E])
        rows_t = J[:, :, T + E:]
        scale = max(1.0, np.abs(Jo[:, :, T + E:]).max())
        assert np.abs(rows_t - Jo[:, :, T + E:]).max() <= 1e-5 * scale
        assert np.abs(r0[:, T:] - r0o[:, T:]).max() <= 1e-9
        assert np.count_nonzero(r0o[:, T + E:]) > 0  # the traced rows are active


def test_mixed_normal_equations_with_rows(engine):
    """Fused mixed-precision A, g (no J formed; rows added by their own normal equations) against
    the FP64 normal equations of the stored J with the same rows."""
    b, v, s = _objective(2)
    x0 = v.extract(b)[:2]
    base = b.take([0])
    _, _, a, g = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    _, _, am, gm = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed",
                                   want_jacobian=False)
    for i in range(2):
        assert np.linalg.norm(am[i] - a[i]) <= 1e-5 * np.linalg.norm(a[i])
        assert np.linalg.norm(gm[i] - g[i]) <= 1e-7 * np.linalg.norm(g[i])


# ---- LM with rows ------------------------------------------------------------------------------

def _same_lm(a, b):
    assert np.array_equal(a.x, b.x)
    assert np.array_equal(a.merit, b.merit)
    assert a.status == b.status
    assert np.array_equal(a.iterations, b.iterations)
    assert a.cycles == b.cycles


@pytest.mark.parametrize("method", [abi.JACOBIAN_FD, abi.JACOBIAN_ADJOINT])
def test_lm_with_rows_device_equals_host_control(engine, method):
    b, v, s = _singlet(6)
    x0 = v.extract(b)[:6]
    kw = dict(max_iterations=10, threads=4, jacobian_method=method)
    rd = engine.optimize(b.take([0]), s, v, x0, abi.lm_options(**kw))
    rh = engine.optimize(b.take([0]), s, v, x0, abi.lm_options(host_control=True, **kw))
    _same_lm(rd, rh)


def test_lm_with_rows_follows_oracle(engine):
    b, v, s = _singlet(4, seed=3)
    x0 = v.extract(b)[:3]
    base = b.take([0])
    # merits after 10 cycles: FD 2.1e-8, adjoint 8.0e-10 relative measured (the Jacobians'
    # device-host differences, the rows from extrema to 1e-10, carried along the LM)
    for method, rtol in ((abi.JACOBIAN_FD, 1e-6), (abi.JACOBIAN_ADJOINT, 1e-7)):
        opt = abi.lm_options(max_iterations=10, threads=4, jacobian_method=method)
        res = engine.optimize(base, s, v, x0, opt)
        _, runs_o, _ = oracle().optimize_lm(base, s, v, x0, opt)
        mo = np.array([r.merit for r in runs_o])
        assert np.all(np.abs(res.merit / mo - 1.0) <= rtol)
        pts = base.take(np.zeros(3, dtype=np.int64))
        v.apply(pts, x0)
        m0, _, _ = engine.merit(pts, s)
        assert np.all(res.merit < m0)


def test_lm_with_rows_mixed_objective(engine):
    b, v, s = _objective(2)
    x0 = v.extract(b)[:2]
    base = b.take([0])
    pts = base.take(np.zeros(2, dtype=np.int64))
    v.apply(pts, x0)
    m0, _, _ = engine.merit(pts, s)
    kw = dict(max_iterations=3, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rd = engine.optimize(base, s, v, x0, abi.lm_options(**kw), precision="mixed")
    rh = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw),
                         precision="mixed")
    _same_lm(rd, rh)
    assert np.all(rd.merit < m0)


def test_row_differences_on_candidate_rays_equal_full_traces(engine):
    """The adjoint paths difference the rows from extrema over the 2P + 1 points of each run; at
    the difference points only the candidate rays (heights / cos I within 1e-3 of the run's
    extrema, or not ending with status Ok) are traced. Stored-J and mixed normal equations equal
    those of tracing every ray bit for bit, on the objective and the singlet in soft mode with
    failed rays."""
    sb, sv, ss = _singlet(6, failure=abi.FAILURE_SOFT)
    sb.pupil_radius[1] = 110.0  # outer rings miss: failed rays in the soft trace
    cases = [(*_objective(2), 2), (sb, sv, ss, 3)]
    for b, v, s, n in cases:
        base, x0 = b.take([0]), v.extract(b)[:n]
        if b is sb:
            base, x0 = b.take(np.arange(n)), v.extract(b)[:n]
        before = engine.row_trace_stats()
        J, r0, a, g = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
        after = engine.row_trace_stats()
        assert after[0] > before[0] and after[1] == before[1]  # candidates used, check passed
        _, _, am, gm = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed",
                                       want_jacobian=False)
        engine.tuning(full_row_traces=True)
        try:
            Jf, r0f, af, gf = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
            _, _, amf, gmf = engine.jacobian(base, s, v, x0, method="adjoint",
                                             precision="mixed", want_jacobian=False)
        finally:
            engine.tuning(full_row_traces=False)
        assert np.array_equal(J, Jf) and np.array_equal(r0, r0f)
        assert np.array_equal(a, af) and np.array_equal(g, gf)
        assert np.array_equal(am, amf) and np.array_equal(gm, gmf)


# ---- sine condition and distortion (real rays at lambda_ref) ----------------------------------

def _image_rows(n=6, seed=7):
    b, v = systems.singlet(n, seed=seed)
    s = systems.singlet_settings(abi.FAILURE_HARD)
    s.constraints = systems.singlet_image_constraints()
    return b, v, s


def test_image_rows_match_oracle(engine):
    """The zone rays of the sine condition are traced with the oracle's operations (rounded one
    by one, correctly rounded divisions and roots): bit for bit. The chief rays of the distortion
    rows start with sin / cos of the field angle, which the device's library may round
    differently by an ulp: 1e-12 relative."""
    b, _, s = _image_rows()
    res, _, _ = engine.residuals(b, s)
    T = res.shape[1] - len(s.constraints)
    rows = res[:, T:]
    ro = oracle().constraint_rows(b, s)
    assert np.array_equal(rows[:, :2], ro[:, :2]) and np.count_nonzero(ro[:, :2]) > 0
    assert np.all(rows[:, 2] == 0.0)  # axis field
    assert np.abs(rows[:, 3:] - ro[:, 3:]).max() <= 1e-12 * np.abs(ro[:, 3:]).max()
    assert np.count_nonzero(ro[:, 3:]) > 0
    m, _, _ = engine.merit(b, s)
    assert _rel(m, oracle().merit(b, s)[0]) <= 1e-9


def test_image_rows_concentric_systems(engine):
    """OSC = 0 up to rounding for systems concentric to the stop, every zone; their distortion
    with the image plane shifted by delta from the focus is delta / f."""
    zones = (0.25, 0.5, 0.70710678118654752440, 1.0)
    for a, gaps, n in ((4.0, [10.0, 3.0], [1.0, 1.6, 1.0]), (5.0, [10.0], [1.0, 1.7]),
                       (4.0, [8.0, 2.0, 3.0], [1.0, 1.75, 1.45, 1.0])):
        b = systems.concentric(a, gaps, n, 20.0)
        s = systems.ResidualSettings(rings=3, spokes=4, failure=abi.FAILURE_HARD,
                                     constraints=[bt.sine_condition(z) for z in zones])
        res, _, _ = engine.residuals(b, s)
        assert np.all(np.abs(res[0, -4:]) <= 1e-14)
    gaps, n, fields = [10.0, 3.0], [1.0, 1.6, 1.0], (0.0, 4.0, 12.0)
    f = 1.0 / oracle().paraxial(systems.concentric(4.0, gaps, n, 20.0, fields), 0)[4]
    for delta in (0.05, -0.2):
        b = systems.concentric(4.0, gaps, n, f - sum(gaps) + delta, fields)
        b.field_target[0] = f * np.tan(np.radians(fields))
        s = systems.ResidualSettings(rings=3, spokes=4, failure=abi.FAILURE_HARD,
                                     constraints=[bt.distortion(0.0, k) for k in range(3)])
        res, _, _ = engine.residuals(b, s)
        assert res[0, -3] == 0.0
        assert np.all(np.abs(res[0, -2:] - abs(delta / f)) <= 1e-12)


@pytest.mark.parametrize("method", ["fd", "adjoint"])
def test_image_rows_in_jacobians(engine, method):
    """The rows' derivatives are central differences of the rows at the FD pattern's points, as
    in the reference (extra_rows_cpu): the sine-condition columns bit for bit, the distortion
    columns within the chief ray's ulp over the FD step."""
    b, v, s = _image_rows()
    x0 = v.extract(b)[:3]
    base = b.take([0])
    J, r0, a, g = engine.jacobian(base, s, v, x0, method=method)
    fn = oracle().fd_jacobian if method == "fd" else oracle().adjoint_jacobian
    Jo, r0o, _ = fn(base, s, v, x0)
    T = J.shape[2] - len(s.constraints)
    assert np.array_equal(J[:, :, T:T + 2], Jo[:, :, T:T + 2])
    dist = np.abs(J[:, :, T + 2:] - Jo[:, :, T + 2:])
    assert dist.max() <= 1e-6 * max(1.0, np.abs(Jo[:, :, T + 2:]).max())
    assert np.count_nonzero(Jo[:, :, T:T + 2]) > 0


def test_lm_with_image_rows(engine):
    """The LM with sine-condition and distortion rows: device control equals host control bit for
    bit, the merit decreases, and it follows the oracle's LM."""
    b, v, s = _image_rows(4, seed=3)
    for c in s.constraints:
        c.weight = 1e4
    x0 = v.extract(b)[:3]
    base = b.take([0])
    kw = dict(max_iterations=8, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rd = engine.optimize(base, s, v, x0, abi.lm_options(**kw))
    rh = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw))
    _same_lm(rd, rh)
    pts = base.take(np.zeros(3, dtype=np.int64))
    v.apply(pts, x0)
    m0, _, _ = engine.merit(pts, s)
    assert np.all(rd.merit < m0)
    _, runs_o, _ = oracle().optimize_lm(base, s, v, x0, abi.lm_options(**kw))
    # 2.7e-9 relative measured (the distortion rows agree to 1e-12, the sine rows bit for bit)
    assert np.all(np.abs(rd.merit / np.array([r.merit for r in runs_o]) - 1.0) <= 1e-7)


def test_device_entry_points_reject_rows(engine):
    """The raw device entry points trace ray rows only and say so (no kernel is launched)."""
    import ctypes as C

    from mikroptim_amd.engine import library

    b, _, s = _singlet(1)
    cb, cs = b.c_struct(), s.c_struct()
    rc = library().mko_residuals_device(C.byref(cb), C.byref(cs), None, 0, None, None, None)
    assert rc != 0 and b"constraint rows" in library().mko_last_error()
    rc = library().mko_merit_device(C.byref(cb), C.byref(cs), None, None, None, None, None, None,
                                    None)
    assert rc != 0

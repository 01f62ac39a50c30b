"""Pins the oracle's constraint rows (ExtraResiduals / TraceExtraResiduals of the reference):
surface extrema, the paraxial y-nu trace, the paraxial and mechanical row formulas, and their
place in residuals, merit, FD / adjoint Jacobians and the LM. Checked bit for bit against the
reference's own CPU code (oracle/_ref, else the committed fixtures) and against closed forms:
the reference's own extrema tests (tests/test_batch_merit.cpp:154-290 restated on this
repository's singlet), thick-lens optics, the sag of a sphere, and linear rows whose derivative
is known exactly."""

import math
import os

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd import batch as bt
from mikroptim_amd.batch import SystemBatch, layout
from oracle_lib import oracle, reference

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = np.load(os.path.join(HERE, "golden", "reference_outputs.npz"))


def _singlet_with(fields_deg, variants=1, seed=1):
    b, v = systems.singlet(variants, seed=seed)
    c = SystemBatch(b.type, b.wavelength_nm, b.wavelength_weight, np.ones(len(fields_deg)),
                    [abi.PUPIL_AXIS if a == 0.0 else abi.PUPIL_HALF for a in fields_deg],
                    variants, b.reference_wavelength_nm)
    for name in ("curvature", "thickness", "index", "pupil_radius"):
        setattr(c, name, getattr(b, name).copy())
    c.paraxial_index = b.paraxial_index.copy()
    th = np.radians(fields_deg)
    c.field_angle[:] = th
    c.field_target[:] = 100.0 * np.tan(th)
    return c, v


# ---- surface extrema (tests/test_batch_merit.cpp:201-290) ------------------------------------

def test_extrema_of_a_collimated_axial_beam_match_closed_forms():
    b, _ = _singlet_with([0.0])
    s = systems.singlet_settings()
    e = oracle().surface_extrema(b, s)[0]
    u, _ = oracle().gauss_legendre(s.rings)
    h = b.pupil_radius[0] * math.sqrt(u[-1])  # outermost ring
    c = b.curvature[0, 1]
    assert abs(e[0, 0] - h) <= 1e-12 and e[0, 1] == 1.0
    assert abs(e[1, 0] - h) <= 1e-12  # rays parallel to the axis reach surface 1 at that height
    assert abs(e[1, 1] - math.sqrt(1.0 - (c * h) ** 2)) <= 1e-14
    assert 0.9 * h < e[2, 0] < h  # the biconvex lens makes the beam converge


def test_extrema_with_field_angles():
    b, _ = _singlet_with([0.0, 1.0, 2.0])
    s = systems.singlet_settings()
    e = oracle().surface_extrema(b, s)[0]
    assert abs(e[0, 1] - math.cos(math.radians(2.0))) <= 1e-15
    u, _ = oracle().gauss_legendre(s.rings)
    assert abs(e[0, 0] - b.pupil_radius[0] * math.sqrt(u[-1])) <= 1e-12


def test_failed_rays_do_not_contribute_to_extrema():
    b, _ = systems.singlet(3, seed=5)
    b.pupil_radius[1] = 1e4    # every ray misses the lens
    b.pupil_radius[2] = 110.0  # outer rings miss
    s = systems.singlet_settings(abi.FAILURE_HARD)
    o = oracle()
    e = o.surface_extrema(b, s)
    assert np.all(e[1, :, 0] == 0.0) and np.all(e[1, :, 1] == 1.0)
    _, _, st = o.residuals(b, s)
    offsets, samples, _ = layout(b, s)
    rho, _, _ = o.pupil_samples(s.rings, s.spokes, abi.PUPIL_HALF)
    ok = st[2] == abi.STATUS_OK
    assert 0 < ok.sum() < ok.size
    stop = 0.0
    for f in range(b.fields):
        r, _, _ = o.pupil_samples(s.rings, s.spokes, int(b.pupil_mode[f]))
        for l in range(b.wavelengths):
            o0 = offsets[f] + l * samples[f]
            sel = ok[o0:o0 + samples[f]]
            if sel.any():
                stop = max(stop, float(np.max(b.pupil_radius[2] * r[sel])))
    assert abs(e[2, 0, 0] - stop) <= 1e-12
    assert e[2, 1, 0] < 1.0 / b.curvature[2, 1]  # valid rays meet the sphere below its radius
    # soft mode: the same extrema (only rays with status Ok count)
    soft = systems.singlet_settings(abi.FAILURE_SOFT)
    assert np.array_equal(o.surface_extrema(b, soft), e)


def test_extrema_equal_reference():
    b, _ = systems.singlet(4, seed=7)
    s = systems.singlet_settings()
    e = oracle().surface_extrema(b, s)
    ref = reference()
    assert np.array_equal(e, FIXTURES["singlet_extrema"])
    if ref is not None:
        assert np.array_equal(e, ref.surface_extrema(b, s))


# ---- paraxial y-nu trace (src/paraxial/paraxial.cpp:98-175) ------------------------------------

def test_paraxial_thick_lens_closed_forms():
    b, _ = systems.singlet(3, seed=4)
    for v in range(b.variants):
        A, B, C, D, power, kp = oracle().paraxial(b, v)
        n = b.paraxial_index[v, 1]
        c1, c2 = b.curvature[v, 1], b.curvature[v, 2]
        t0, t1 = b.thickness[v, 0], b.thickness[v, 1]
        p1, p2 = c1 * (n - 1.0), c2 * (1.0 - n)
        phi = p1 + p2 - t1 / n * p1 * p2  # thick-lens power
        assert abs(power - phi) <= 1e-15 * max(1.0, abs(phi)) + 1e-17
        assert abs(A * D - B * C - 1.0) <= 1e-13  # det M = 1 in reduced angles
        # marginal ray height at the last surface: 1 - t1 p1 / n
        assert abs(A - (1.0 - t1 * p1 / n)) <= 1e-14
        # chief ray through the stop centre: height at surface 1 is t0, then through the lens
        assert abs(B - (t0 + t1 * (1.0 - t0 * p1) / n)) <= 1e-13
        petz = p1 / n + c2 * (1.0 - n) / n
        assert abs(kp - (-1.0 * petz)) <= 1e-15


def test_paraxial_equals_reference():
    ref = reference()
    b, _ = systems.singlet(3, seed=4)
    ob, _ = systems.objective(variants=2, seed=2)
    for batch in (b, ob):
        for v in range(batch.variants):
            o = oracle().paraxial(batch, v)
            if ref is not None:
                assert np.array_equal(o, ref.paraxial(batch, v))
    assert np.array_equal(oracle().paraxial(ob, 1), FIXTURES["objective_paraxial_v1"])


# ---- row formulas (paraxial_residuals.cpp:257-339, mechanical_constraints.cpp:58-281) ---------

def _sag(c, h):
    return 0.0 if c == 0.0 else (1.0 - math.sqrt(1.0 - (c * h) ** 2)) / c


def test_rows_match_closed_forms():
    b, _ = systems.singlet(2, seed=9)
    s = systems.singlet_settings()
    s.constraints = systems.singlet_constraints()
    rows = oracle().constraint_rows(b, s)
    ext = oracle().surface_extrema(b, s)
    order = s.row_constraints()
    assert [c.needs_extrema for c in order] == [False] * 7 + [True] * 3
    for v in range(b.variants):
        A, _, C, _, power, kp = oracle().paraxial(b, v)
        t, c = b.thickness[v], b.curvature[v]
        expect = []
        for con in order:
            w = math.sqrt(con.weight)
            if con.type == abi.CON_FOCAL_LENGTH:
                expect.append(w * con.target * (1.0 - con.target * power))
            elif con.type == abi.CON_WORKING_DISTANCE:
                expect.append(w * min(t[-1] - con.target, 0.0))
            elif con.type == abi.CON_PARFOCAL_LENGTH:
                expect.append(w * (con.offset + t[con.surface:].sum() - con.target))
            elif con.type == abi.CON_PARAXIAL_FOCUS:
                expect.append(w * b.pupil_radius[v] * (A + t[-1] * C / b.paraxial_index[v, -1]))
            elif con.type == abi.CON_PETZVAL:
                expect.append(w * max(abs(kp) - con.target, 0.0))
            elif con.type == abi.CON_MAX_SEMI_DIAMETER:
                expect.append(w * max(ext[v, con.surface, 0] + con.offset - con.target, 0.0))
            elif con.type == abi.CON_MIN_CENTER_THICKNESS:
                expect.append(w * max(con.target - t[con.surface:con.next_surface].sum(), 0.0))
            elif con.type == abi.CON_MIN_EDGE_THICKNESS:
                if con.edge_height == abi.EDGE_RAYS:
                    h = max(ext[v, con.surface, 0], ext[v, con.next_surface, 0]) + con.offset
                else:
                    h = max(con.semi_diameter, con.next_semi_diameter) + con.offset
                edge = (t[con.surface:con.next_surface].sum() - _sag(c[con.surface], h)
                        + _sag(c[con.next_surface], h))
                expect.append(w * max(con.target - edge, 0.0))
            elif con.type == abi.CON_MAX_INCIDENCE_ANGLE:
                expect.append(w * max(con.target - ext[v, con.surface, 1], 0.0))
        np.testing.assert_allclose(rows[v], expect, rtol=1e-12, atol=1e-12)
    # the set activates every row kind somewhere (not all zeros)
    assert np.count_nonzero(rows[0]) >= 7


def test_modes_and_shifts():
    b, _ = systems.singlet(1)
    s = systems.singlet_settings()
    wd = b.thickness[0, -1]
    cases = [(bt.working_distance(wd + 1.0, 4.0), -2.0),               # equal: q
             (bt.working_distance(wd + 1.0, 4.0, abi.MODE_MIN), -2.0),  # min, active
             (bt.working_distance(wd - 1.0, 4.0, abi.MODE_MIN), 0.0),   # min, inactive
             (bt.working_distance(wd - 1.0, 4.0, abi.MODE_MAX), 2.0),   # max, active
             (bt.working_distance(wd - 1.0, 4.0, abi.MODE_MAX, shift=0.5), 3.0),
             (bt.working_distance(wd + 1.0, 4.0, abi.MODE_MIN, shift=0.5), -3.0),
             (bt.center_thickness(1, 2, 6.0, 4.0, shift=0.25), 2.0 * 1.25)]
    for con, expect in cases:
        s.constraints = [con]
        r = oracle().constraint_rows(b, s)[0, 0]
        assert abs(r - expect) <= 1e-12, (con, r, expect)


def test_invalid_constraints_are_rejected():
    b, _ = systems.singlet(1)
    s = systems.singlet_settings()
    bad = [bt.Constraint(abi.CON_FOCAL_LENGTH, 1.0, 0.0),                 # f_s = 0
           bt.Constraint(abi.CON_FOCAL_LENGTH, 1.0, 50.0, abi.MODE_MIN),  # equation only
           bt.Constraint(abi.CON_PETZVAL, 1.0, 0.1, abi.MODE_EQUAL),      # upper bound only
           bt.center_thickness(2, 1, 1.0),                                # next before surface
           bt.max_semi_diameter(7, 1.0),                                  # surface out of range
           bt.Constraint(abi.CON_MAX_INCIDENCE_ANGLE, 1.0, 1.5, surface=1),  # cos out of (0, 1)
           bt.working_distance(1.0, -1.0),                                # negative weight
           bt.working_distance(1.0, 1.0, abi.MODE_EQUAL, shift=0.5),      # shift on an equation
           bt.Constraint(5, 1.0, 0.0)]                                    # ghost L1: unsupported
    for con in bad:
        s.constraints = [con]
        with pytest.raises(RuntimeError):
            oracle().constraint_rows(b, s)
    b.paraxial_index = None
    s.constraints = [bt.focal_length(100.0)]
    with pytest.raises(RuntimeError):
        oracle().constraint_rows(b, s)


# ---- rows in residuals, merit and Jacobians ------------------------------------------------------

def test_rows_in_residuals_and_merit():
    b, _ = systems.singlet(3, seed=2)
    s = systems.singlet_settings(abi.FAILURE_SOFT)
    res0, _, _ = oracle().residuals(b, s)
    m0 = oracle().merit(b, s)[0]
    s.constraints = systems.singlet_constraints()
    res, _, _ = oracle().residuals(b, s)
    rows = oracle().constraint_rows(b, s)
    T = res0.shape[1]
    assert np.array_equal(res[:, :T], res0) and np.array_equal(res[:, T:], rows)
    m = oracle().merit(b, s)[0]
    E = sum(not c.needs_extrema for c in s.constraints)
    expect = (m0 + (rows[:, :E] ** 2).sum(axis=1)) + (rows[:, E:] ** 2).sum(axis=1)
    np.testing.assert_allclose(m, expect, rtol=1e-14)


def test_linear_rows_have_exact_derivatives():
    """d/dw of sqrt(w_c)(t_{S-1} - w_s) is sqrt(w_c); d/dt1 of an active min center thickness
    row is -sqrt(w_c): the FD and adjoint Jacobians hold these (linear rows, central steps)."""
    b, v = systems.singlet(1)
    s = systems.singlet_settings()
    s.constraints = [bt.working_distance(90.0, 0.25), bt.center_thickness(1, 2, 8.0, 4.0)]
    x0 = v.extract(b)[:1]
    T = oracle().residuals(b.take([0]), systems.singlet_settings())[0].shape[1]
    for J, _, _ in (oracle().fd_jacobian(b, s, v, x0), oracle().adjoint_jacobian(b, s, v, x0)):
        rows = J[0, :, T:]
        np.testing.assert_allclose(rows[3], [0.5, 0.0], atol=1e-9)   # working distance
        np.testing.assert_allclose(rows[2], [0.0, -2.0], atol=1e-9)  # lens thickness
        np.testing.assert_allclose(rows[:2], 0.0, atol=1e-9)          # curvatures


@pytest.mark.parametrize("method", ["fd", "adjoint"])
def test_jacobian_rows_equal_reference(method):
    b, v = systems.singlet(4, seed=7)
    s = systems.singlet_settings(abi.FAILURE_SOFT)
    s.constraints = systems.singlet_constraints()
    x0 = v.extract(b)[:2]
    fn = "fd_jacobian" if method == "fd" else "adjoint_jacobian"
    J, r0, _ = getattr(oracle(), fn)(b.take([0, 1]), s, v, x0)
    T = J.shape[2] - len(s.constraints)
    assert np.array_equal(J[:, :, T:], FIXTURES[f"singlet_{method}_rows_J"])
    assert np.array_equal(r0[:, T:], FIXTURES[f"singlet_{method}_rows_r0"])
    ref = reference()
    if ref is not None:
        Jr, r0r, _ = getattr(ref, fn)(b.take([0, 1]), s, v, x0)
        assert np.array_equal(J, Jr) and np.array_equal(r0, r0r)


def test_rows_residuals_merit_equal_reference():
    b, _ = systems.singlet(4, seed=7)
    s = systems.singlet_settings(abi.FAILURE_HARD)
    s.constraints = systems.singlet_constraints()
    rows = oracle().constraint_rows(b, s)
    m = oracle().merit(b, s)[0]
    assert np.array_equal(rows, FIXTURES["singlet_rows"])
    assert np.array_equal(m, FIXTURES["singlet_rows_merit"])
    ref = reference()
    if ref is not None:
        assert np.array_equal(rows, ref.constraint_rows(b, s))
        for a, c in zip(oracle().residuals(b, s), ref.residuals(b, s)):
            assert np.array_equal(a, c)
        assert np.array_equal(m, ref.merit(b, s)[0])


@pytest.mark.parametrize("method", [abi.JACOBIAN_FD, abi.JACOBIAN_ADJOINT])
def test_lm_with_constraints_equals_reference(method):
    b, v = systems.singlet(8, seed=3)
    s = systems.singlet_settings(abi.FAILURE_HARD)
    s.constraints = systems.singlet_constraints()
    x0 = v.extract(b)[:2]
    opt = abi.lm_options(max_iterations=8, threads=2, jacobian_method=method)
    xo, ro, co = oracle().optimize_lm(b.take([0]), s, v, x0, opt)
    key = "fd" if method == abi.JACOBIAN_FD else "adjoint"
    assert np.array_equal(xo, FIXTURES[f"singlet_rows_lm_{key}_x"])
    ref = reference()
    if ref is not None:
        xr, rr, cr = ref.optimize_lm(b.take([0]), s, v, x0, opt)
        assert np.array_equal(xo, xr) and co == cr
        for a, c in zip(ro, rr):
            assert (a.merit, a.status, a.iterations, a.accepted) == \
                   (c.merit, c.status, c.iterations, c.accepted)


# ---- sine condition and distortion (paraxial_residuals.cpp:68-91, :290-315) --------------------

SQRT_HALF = 0.70710678118654752440
# the reference's tolerances (tests/tolerances.hpp:456, :578)
SINE_CONDITION_EXACT = 1e-14
CHIEF_RAY_ANALYTIC = 1e-12


def _rows_of(batch, constraints, failure=abi.FAILURE_HARD):
    s = systems.ResidualSettings(rings=3, spokes=4, failure=failure, constraints=constraints)
    return oracle().constraint_rows(batch, s)


def test_sine_condition_concentric_systems_satisfy_it_exactly():
    """Systems concentric to the stop are free of coma: OSC = 0 up to rounding in every zone and
    at large aperture, also with glass in image space
    (SineCondition.ConcentricSystemsSatisfyItExactly, tests/test_sine_condition.cpp:104)."""
    cases = [(4.0, [10.0, 3.0], [1.0, 1.6, 1.0]), (7.0, [10.0, 4.0], [1.0, 1.9, 1.0]),
             (5.0, [10.0], [1.0, 1.7]), (4.0, [8.0, 2.0, 3.0], [1.0, 1.75, 1.45, 1.0])]
    zones = (0.25, 0.5, SQRT_HALF, 1.0)
    for a, gaps, n in cases:
        b = systems.concentric(a, gaps, n, 20.0)
        rows = _rows_of(b, [bt.sine_condition(z) for z in zones])
        assert np.all(np.abs(rows) <= SINE_CONDITION_EXACT), (a, gaps, n, rows)


def test_sine_condition_shifted_stop_violates_it():
    """SineCondition.ShiftedStopViolatesIt (tests/test_sine_condition.cpp:134): the stop 4 mm
    closer to the surfaces breaks the concentricity, OSC != 0; and a heavier weight scales the
    row by its square root."""
    b = systems.concentric(4.0, [10.0, 3.0], [1.0, 1.6, 1.0], 20.0)
    b.thickness[0, 0] = 6.0
    r1 = _rows_of(b, [bt.sine_condition(1.0)])[0, 0]
    r4 = _rows_of(b, [bt.sine_condition(1.0, weight=4.0)])[0, 0]
    assert abs(r1) > 1e-3 and r4 == 2.0 * r1


def test_sine_condition_small_aperture_is_third_order():
    """At small aperture the OSC is third-order coma, -S_II / (2 H) with S_II ~ H^3: it grows
    as the square of the zone, so the zone 1/sqrt(2) has half the rim's OSC
    (SineCondition.SmallApertureMatchesSeidelComa, tests/test_sine_condition.cpp:162, with its
    tolerance 2 x 1e-2, in the form that needs no Seidel sums)."""
    b, _ = systems.singlet(1)
    b.pupil_radius[:] = 0.1
    rim, zone = _rows_of(b, [bt.sine_condition(1.0), bt.sine_condition(SQRT_HALF)])[0]
    assert rim != 0.0 and abs(zone - 0.5 * rim) <= 2.0 * 1e-2 * abs(0.5 * rim)


def test_chief_ray_height_through_plates_is_analytic():
    """Plane plates shift the chief ray only: y = sum t_i tan(theta_i), n_i sin(theta_i) =
    sin(theta) at lambda_ref (ReferencePoint.ChiefRayHeightThroughPlatesIsAnalytic,
    tests/test_reference_point.cpp:120), read through the distortion row with D_max = 0."""
    b, _ = systems.singlet(1)
    b.curvature[:, 1:] = 0.0
    fields = np.radians([5.0, 20.0])
    b.field_angle[0, 1:] = fields
    n = b.paraxial_index[0, 1]
    t = b.thickness[0]
    for k, theta in enumerate(fields, start=1):
        y = t[0] * math.tan(theta) + t[1] * math.tan(math.asin(math.sin(theta) / n)) \
            + t[2] * math.tan(theta)
        h = 0.5 * y
        b.field_target[0, k] = h
        row = _rows_of(b, [bt.distortion(0.0, k)])[0, 0]
        # row = |(y - h) / h| = 1 + (y - 2h) / h
        assert abs(h * (1.0 + row) - y) <= CHIEF_RAY_ANALYTIC * (1.0 + abs(y))


def test_distortion_of_concentric_system_is_analytic():
    """A system concentric to the stop images the chief ray at L tan(theta), L the distance of the
    image plane from the stop; with the image plane in the focus (L = f) it is free of distortion,
    shifted by delta D = delta / f (ReferencePoint.DistortionOfConcentricSystemIsAnalytic,
    tests/test_reference_point.cpp:157). Axis rows and rows of absent fields are 0."""
    gaps, n, fields = [10.0, 3.0], [1.0, 1.6, 1.0], (0.0, 4.0, 12.0)
    probe = systems.concentric(4.0, gaps, n, 20.0, fields)
    f = 1.0 / oracle().paraxial(probe, 0)[4]
    w0 = f - sum(gaps)
    for delta in (0.0, 0.05, -0.2):
        b = systems.concentric(4.0, gaps, n, w0 + delta, fields)
        b.field_target[0] = f * np.tan(np.radians(fields))
        rows = _rows_of(b, [bt.distortion(0.0, k, weight=4.0) for k in range(3)]
                        + [bt.distortion(0.0, 7, weight=4.0)])[0]
        assert rows[0] == 0.0 and rows[3] == 0.0
        for k in (1, 2):
            assert abs(rows[k] - 2.0 * abs(delta / f)) <= 2.0 * CHIEF_RAY_ANALYTIC
        limit = 0.5 * abs(delta / f) + 1e-9
        r = _rows_of(b, [bt.distortion(limit, 2, weight=4.0)])[0, 0]
        assert abs(r - 2.0 * max(0.0, abs(delta / f) - limit)) <= 2.0 * CHIEF_RAY_ANALYTIC


def test_sine_condition_failed_ray_gives_zero_row():
    """A zone ray that misses a surface has no OSC: its row is 0 (the ray rows penalise the
    failure; SineCondition.FailedRayGivesZeroResidual, tests/test_sine_condition.cpp:212)."""
    b, _ = systems.singlet(1)
    b.pupil_radius[:] = 60.0  # the rim ray misses the first lens surface (radius 50 mm)
    rim, zone = _rows_of(b, [bt.sine_condition(1.0), bt.sine_condition(0.1)])[0]
    assert rim == 0.0 and zone != 0.0


def test_image_constraints_equal_reference():
    """Sine-condition and distortion rows, their place in residuals and merit, and their FD /
    adjoint Jacobian columns are the reference's own values bit for bit."""
    b, v = systems.singlet(4, seed=7)
    s = systems.singlet_settings(abi.FAILURE_HARD)
    s.constraints = systems.singlet_image_constraints()
    rows = oracle().constraint_rows(b, s)
    assert np.count_nonzero(rows[:, :2]) == 8 and np.all(rows[:, 2] == 0.0)
    assert np.array_equal(rows, FIXTURES["singlet_image_rows"])
    x0 = v.extract(b)[:2]
    J, r0, _ = oracle().fd_jacobian(b.take([0, 1]), s, v, x0)
    T = J.shape[2] - len(s.constraints)
    assert np.array_equal(J[:, :, T:], FIXTURES["singlet_image_fd_J"])
    ref = reference()
    if ref is not None:
        assert np.array_equal(rows, ref.constraint_rows(b, s))
        assert np.array_equal(oracle().merit(b, s)[0], ref.merit(b, s)[0])
        for fn in ("fd_jacobian", "adjoint_jacobian"):
            Jo, r0o, _ = getattr(oracle(), fn)(b.take([0, 1]), s, v, x0)
            Jr, r0r, _ = getattr(ref, fn)(b.take([0, 1]), s, v, x0)
            assert np.array_equal(Jo, Jr) and np.array_equal(r0o, r0r)


def test_invalid_image_constraints_are_rejected():
    b, _ = systems.singlet(1)
    s = systems.singlet_settings()
    bad = [bt.Constraint(abi.CON_SINE_CONDITION, 1.0, 0.0, abi.MODE_MIN, zone=1.0),  # equation
           bt.Constraint(abi.CON_SINE_CONDITION, 1.0, 0.1, zone=1.0),                # target 0
           bt.sine_condition(0.0), bt.sine_condition(1.5),                            # zone
           bt.Constraint(abi.CON_DISTORTION, 1.0, 0.01, abi.MODE_EQUAL, field=1),     # max only
           bt.distortion(-0.01, 1), bt.distortion(0.01, -1)]
    for con in bad:
        s.constraints = [con]
        with pytest.raises(RuntimeError):
            oracle().constraint_rows(b, s)
    pb, _ = systems.singlet(1)
    pb.source = abi.SOURCE_POINT
    pb.source_point[:] = [0.0, 0.0, -100.0]
    pb.source_na = 0.05
    for con in (bt.sine_condition(1.0), bt.distortion(0.01, 0)):
        s.constraints = [con]
        with pytest.raises(RuntimeError):
            oracle().constraint_rows(pb, s)

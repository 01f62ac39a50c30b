"""Parity of the HIP path (through the C-ABI of libmikroptim_hip.so) with the CPU FP64 oracle on
the same seeded inputs. Tolerances follow tests/tolerances.hpp of the reference:
kCpuGpuFp64 = 1e-12 mm per ray (OPD), kAdjointCpuGpu = 1e-10 relative per Jacobian column,
kMeritVsResidualSum = 1e-13; statuses must be equal."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd.engine import TraceEngine
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

OPD_TOL = 1e-12          # mm per ray (kCpuGpuFp64)
ADJ_TOL = 1e-10          # relative per column (kAdjointCpuGpu)


def _col_rel(J, Jo):
    num = np.linalg.norm(J - Jo, axis=2)
    den = np.linalg.norm(Jo, axis=2)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)


@pytest.fixture(scope="module")
def singlet():
    return systems.singlet(16)


@pytest.fixture(scope="module")
def objective():
    b, v = systems.objective(variants=64, seed=2)
    return b.take(np.arange(6)), v


@pytest.mark.parametrize("failure", [abi.FAILURE_SOFT, abi.FAILURE_HARD])
def test_residuals_singlet(engine, singlet, failure):
    b, _ = singlet
    s = systems.singlet_settings(failure)
    res, opd, st = engine.residuals(b, s)
    ro, opdo, sto = oracle().residuals(b, s)
    assert np.array_equal(st, sto)
    assert np.max(np.abs(opd - opdo)) <= OPD_TOL
    lam = b.reference_wavelength_nm * 1e-6
    assert np.max(np.abs(res - ro)) <= OPD_TOL / lam


def test_residuals_objective(engine, objective):
    b, _ = objective
    s = systems.objective_settings(abi.FAILURE_HARD)
    res, opd, st = engine.residuals(b, s)
    ro, opdo, sto = oracle().residuals(b, s, threads=8)
    assert np.array_equal(st, sto)
    assert (st == 0).all()
    err = np.abs(opd - opdo)
    assert err.max() <= OPD_TOL  # absolute bound per ray


@pytest.mark.parametrize("failure", [abi.FAILURE_SOFT, abi.FAILURE_HARD])
def test_merit_matches_oracle_and_residual_sum(engine, singlet, failure):
    b, _ = singlet
    s = systems.singlet_settings(failure)
    m, f, h = engine.merit(b, s)
    mo, fo, ho = oracle().merit(b, s)
    assert np.array_equal(f, fo) and np.array_equal(h, ho)
    assert np.max(np.abs(m - mo) / mo) <= 1e-9
    res, _, _ = engine.residuals(b, s)
    assert np.max(np.abs(m - (res ** 2).sum(1)) / m) <= 1e-13


def test_merit_extra(engine, singlet):
    b, _ = singlet
    s = systems.singlet_settings()
    extra = np.linspace(0.5, 2.0, b.variants)
    m0, _, _ = engine.merit(b, s)
    m1, _, _ = engine.merit(b, s, extra)
    assert np.array_equal(m1, m0 + extra)


def test_residuals_repeatable_and_tranche_invariant(singlet):
    b, _ = singlet
    s = systems.singlet_settings(abi.FAILURE_SOFT)
    e1 = TraceEngine(0)
    r1 = e1.residuals(b, s)
    r2 = e1.residuals(b, s)
    per_variant = (2 * b.surfaces + b.surfaces * b.wavelengths + 1 + 2 * b.fields) * 8 + 900 * 8 + 450 * 9
    e2 = TraceEngine(0, memory_limit_bytes=3 * per_variant)  # forces tranches of 3 variants
    r3 = e2.residuals(b, s)
    for a, c, d in zip(r1, r2, r3):
        assert np.array_equal(a, c) and np.array_equal(a, d)


@pytest.mark.parametrize("failure", [abi.FAILURE_SOFT, abi.FAILURE_HARD])
def test_adjoint_fp64_matches_oracle(engine, singlet, failure):
    b, v = singlet
    s = systems.singlet_settings(failure)
    x0 = v.extract(b)[:4]
    base = b.take([0])
    J, r0, a, g = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    Jo, r0o, _ = oracle().adjoint_jacobian(base, s, v, x0)
    assert _col_rel(J, Jo).max() <= ADJ_TOL
    assert np.max(np.abs(r0 - r0o)) <= OPD_TOL / (b.reference_wavelength_nm * 1e-6)
    # normal equations of the downloaded J: bitwise the host FMA chain
    ao, go = oracle().normal_equations(J, r0)
    assert np.array_equal(a, ao) and np.array_equal(g, go)


def test_adjoint_fp64_objective(engine, objective):
    b, v = objective
    s = systems.objective_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:2]
    base = b.take([0])
    J, r0, _, _ = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    Jo, r0o, _ = oracle().adjoint_jacobian(base, s, v, x0, threads=8)
    assert _col_rel(J, Jo).max() <= ADJ_TOL


def test_adjoint_mixed_stored_jacobian_is_fp64(engine, objective):
    """With J requested, MIXED computes the stored Jacobian in FP64: same bits as FP64."""
    b, v = objective
    s = systems.objective_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:2]
    base = b.take([0])
    Jm, r0m, am, gm = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed")
    Jd, r0d, ad, gd = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    for x, y in ((Jm, Jd), (r0m, r0d), (am, ad), (gm, gd)):
        assert np.array_equal(x, y)


# Mixed-precision normal equations (fused adjoint kernel) against the FP64 J^T J, J^T r0 of the
# same points. Error sources: J^T J from split-bf16 products (hi*hi + hi*lo + lo*hi, the dropped
# lo*lo term and the rounding of lo are ~2^-17 relative per product) accumulated in float per
# tuple; per-ray gradients staged as float differences to a pilot ray of the wave (2^-24 relative
# to the gradient's variation over the pupil). Observed on the NA 0.95 objective: A 1.2e-6
# (Frobenius, relative), diagonal 6.6e-6, g 4.8e-9.
MIXED_A_TOL = 1e-5       # ||A - A64||_F / ||A64||_F per run
MIXED_DIAG_TOL = 5e-5    # max_p |A_pp / A64_pp - 1|
MIXED_G_TOL = 1e-7       # ||g - g64|| / ||g64|| per run


def _mixed_errors(engine, base, s, v, x0):
    _, _, a, g = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed",
                                 want_jacobian=False)
    _, _, ad, gd = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    ea = np.linalg.norm(a - ad, axis=(1, 2)) / np.linalg.norm(ad, axis=(1, 2))
    diag = np.abs(np.diagonal(a, axis1=1, axis2=2) / np.diagonal(ad, axis1=1, axis2=2) - 1)
    eg = np.linalg.norm(g - gd, axis=1) / np.linalg.norm(gd, axis=1)
    print(f"mixed normal equations: A {ea.max():.2e}, diag {diag.max():.2e}, g {eg.max():.2e}")
    return ea.max(), diag.max(), eg.max(), a


def test_adjoint_mixed_normal_equations_objective(engine, objective):
    """512 rays per tuple (two halves per workgroup), 48 variables."""
    b, v = objective
    s = systems.objective_settings(abi.FAILURE_HARD)
    ea, ed, eg, a = _mixed_errors(engine, b.take([0]), s, v, v.extract(b)[:6])
    assert ea <= MIXED_A_TOL and ed <= MIXED_DIAG_TOL and eg <= MIXED_G_TOL
    # exactly symmetric: one C tile per pair of off-diagonal blocks, mirrored in the finalize
    assert np.array_equal(a, np.swapaxes(a, 1, 2))


def test_adjoint_mixed_normal_equations_singlet(engine, singlet):
    """72 rays per tuple (one partial half), few variables; per-variant templates."""
    b, v = singlet
    s = systems.singlet_settings(abi.FAILURE_HARD)
    ea, ed, eg, _ = _mixed_errors(engine, b.take(np.arange(5)), s, v, v.extract(b)[:5])
    assert ea <= MIXED_A_TOL and ed <= MIXED_DIAG_TOL and eg <= MIXED_G_TOL


def test_adjoint_mixed_soft_mode_without_failures_is_hard_mode(engine, singlet):
    """Soft failure mode at points without failed rays: the penalty rows and their derivatives
    are zero, so MIXED gives hard mode's fused normal equations (runs with failed rays take the
    FP64 path: tests/test_gpu_soft_mixed.py)."""
    b, v = singlet
    s = systems.singlet_settings(abi.FAILURE_SOFT)
    base, x0 = b.take([0]), v.extract(b)[:3]
    _, _, am, gm = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed",
                                   want_jacobian=False)
    _, _, ah, gh = engine.jacobian(base, systems.singlet_settings(abi.FAILURE_HARD), v, x0,
                                   method="adjoint", precision="mixed", want_jacobian=False)
    assert np.array_equal(am, ah) and np.array_equal(gm, gh)


@pytest.mark.parametrize("failure", [abi.FAILURE_SOFT, abi.FAILURE_HARD])
def test_fd_jacobian_matches_oracle(engine, singlet, failure):
    b, v = singlet
    s = systems.singlet_settings(failure)
    x0 = v.extract(b)[:3]
    base = b.take([0])
    J, r0, a, g = engine.jacobian(base, s, v, x0, method="fd")
    Jo, r0o, _ = oracle().fd_jacobian(base, s, v, x0)
    assert _col_rel(J, Jo).max() <= 1e-5
    ao, go = oracle().normal_equations(J, r0)
    assert np.array_equal(a, ao) and np.array_equal(g, go)


def test_lm_fd_and_adjoint_follow_the_oracle(engine, singlet):
    b, v = singlet
    s = systems.singlet_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:4]
    base = b.take([0])
    # after 15 cycles the merits agree to the Jacobians' device-host differences carried along
    # the bending valley: FD columns ~1e-5 relative (the residuals' difference over the central
    # step) -> merits 2.2e-7 measured; adjoint columns ~1e-12 -> 2.0e-11 measured
    for method, rtol in ((abi.JACOBIAN_FD, 1e-5), (abi.JACOBIAN_ADJOINT, 1e-8)):
        opt = abi.lm_options(max_iterations=15, threads=4, jacobian_method=method)
        res = engine.optimize(base, s, v, x0, opt)
        xo, runs_o, cyc_o = oracle().optimize_lm(base, s, v, x0, opt)
        mo = np.array([r.merit for r in runs_o])
        assert np.all(np.abs(res.merit / mo - 1.0) <= rtol)


def test_lm_adjoint_mixed_converges_like_fp64(engine, singlet, objective):
    """Batched LM with the fused mixed-precision normal equations against the FP64 adjoint LM.
    The normal equations differ by ~1e-6 (pinned above); LM amplifies such differences along
    flat valleys (the singlet's bending, the objective's ill-conditioned damped systems), so
    trajectories are not compared point by point. Required: every run's merit decreases, and on
    the singlet the mixed LM ends within 5 % of (or below) the FP64 LM's merit. Both evaluate
    the merit in FP64, so accepted points are compared exactly."""
    b, v = singlet
    s = systems.singlet_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:6]
    base = b.take([0])
    pts = base.take(np.zeros(6, dtype=np.int64))
    v.apply(pts, x0)
    m0, _, _ = engine.merit(pts, s)
    opt = abi.lm_options(max_iterations=8, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rm = engine.optimize(base, s, v, x0, opt, precision="mixed")
    rd = engine.optimize(base, s, v, x0, opt, precision="fp64")
    assert np.all(rm.merit < m0) and np.all(rd.merit < m0)
    assert np.all(rm.merit <= 1.05 * rd.merit)

    b, v = objective
    s = systems.objective_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:4]
    base = b.take([0])
    pts = base.take(np.zeros(4, dtype=np.int64))
    v.apply(pts, x0)
    m0, _, _ = engine.merit(pts, s)
    opt = abi.lm_options(max_iterations=6, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rm = engine.optimize(base, s, v, x0, opt, precision="mixed")
    assert np.all(rm.merit < m0)


def _same_lm(a, b):
    assert np.array_equal(a.x, b.x)
    assert np.array_equal(a.merit, b.merit)
    assert a.status == b.status
    assert np.array_equal(a.iterations, b.iterations)
    for ra, rb in zip(a.runs, b.runs):
        assert (ra.accepted, ra.jacobians, ra.penalty_rounds) == (rb.accepted, rb.jacobians,
                                                                   rb.penalty_rounds)
        assert ra.damping == rb.damping and ra.fail_weight == rb.fail_weight
        assert (ra.failed_rays, ra.hard_failed_rays) == (rb.failed_rays, rb.hard_failed_rays)
    assert a.cycles == b.cycles


@pytest.mark.parametrize("method,failure,per_run,stall", [
    (abi.JACOBIAN_FD, abi.FAILURE_HARD, False, 0),
    (abi.JACOBIAN_ADJOINT, abi.FAILURE_HARD, True, 3),
    (abi.JACOBIAN_ADJOINT, abi.FAILURE_SOFT, False, 0),   # penalty rounds
    (abi.JACOBIAN_FD, abi.FAILURE_SOFT, True, 2),
])
def test_lm_device_control_matches_host_control(engine, singlet, method, failure, per_run, stall):
    """The LM control on the device (step solves, acceptance, convergence, stall test) takes the
    same steps as the host control, bit for bit, given the same normal equations and merits."""
    b, v = singlet
    s = systems.singlet_settings(failure)
    n = 6
    x0 = v.extract(b)[:n]
    base = b.take(np.arange(n)) if per_run else b.take([0])
    kw = dict(max_iterations=12, threads=4, jacobian_method=method, stall_window=stall)
    rd = engine.optimize(base, s, v, x0, abi.lm_options(**kw))
    rh = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw))
    _same_lm(rd, rh)


def test_lm_device_control_mixed_objective(engine, objective):
    """Fused mixed-precision normal equations with the device control against the host control
    on the same normal equations: identical runs."""
    b, v = objective
    s = systems.objective_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:4]
    kw = dict(max_iterations=4, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rd = engine.optimize(b.take([0]), s, v, x0, abi.lm_options(**kw), precision="mixed")
    rh = engine.optimize(b.take([0]), s, v, x0, abi.lm_options(host_control=True, **kw),
                         precision="mixed")
    _same_lm(rd, rh)


def test_lm_session_cycles_match_optimize(engine, singlet):
    """Stepwise LM (one cycle per call, results readable in between) ends where optimize ends."""
    b, v = singlet
    s = systems.singlet_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:5]
    opt = abi.lm_options(max_iterations=8, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    sess = engine.lm(b.take([0]), s, v, x0, opt)
    mid = None
    while sess.cycle():
        if mid is None:
            mid = sess.results()
    end = sess.results()
    sess.close()
    ref = engine.optimize(b.take([0]), s, v, x0, opt)
    _same_lm(end, ref)
    assert np.all(end.merit <= mid.merit)


@pytest.mark.parametrize("method,precision,failure", [
    ("adjoint", "fp64", abi.FAILURE_SOFT), ("adjoint", "mixed", abi.FAILURE_HARD),
    ("fd", "fp64", abi.FAILURE_HARD)])
def test_backend_group_equals_single_runs(engine, singlet, method, precision, failure):
    """LmBackend contract (levenberg_marquardt.hpp:167-170): a run's normal equations and merit
    depend only on its index and point, bit for bit, not on the other runs of the call."""
    b, v = singlet
    s = systems.singlet_settings(failure)
    base = b.take(np.arange(8))
    x = v.extract(b)[:8] * (1 + 1e-3 * np.arange(8))[:, None]
    runs = np.array([1, 4, 6], dtype=np.int32)
    ag, gg = engine.normal_equations(base, s, v, x[runs], runs=runs, method=method,
                                     precision=precision)
    mg, fg, hg = engine.point_merit(base, s, v, x[runs], runs=runs)
    for j, r in enumerate(runs):
        a1, g1 = engine.normal_equations(base, s, v, x[[r]], runs=[r], method=method,
                                         precision=precision)
        m1, f1, h1 = engine.point_merit(base, s, v, x[[r]], runs=[r])
        assert np.array_equal(a1[0], ag[j]) and np.array_equal(g1[0], gg[j])
        assert m1[0] == mg[j] and f1[0] == fg[j] and h1[0] == hg[j]
    # the merit of a point equals MERIT of the batch with the point applied
    pts = base.take(runs)
    v.apply(pts, x[runs])
    m, f, h = engine.merit(pts, s)
    assert np.array_equal(m, mg) and np.array_equal(f, fg) and np.array_equal(h, hg)


def test_autograd_residuals_backward_is_adjoint_vjp(engine, singlet):
    """torch.autograd.Function: forward equals RESIDUALS at the points; backward equals the
    vector-Jacobian product checked against central differences of <v, r(x)> in float64."""
    import torch
    from mikroptim_amd.autograd import trace_residuals

    b, v = singlet
    s = systems.singlet_settings(abi.FAILURE_HARD)
    base = b.take([0])
    x0 = v.extract(b)[:3]
    x = torch.tensor(x0, dtype=torch.float64, device="cuda", requires_grad=True)
    r = trace_residuals(x, engine, base, s, v)
    pts = base.take(np.zeros(3, dtype=np.int64))
    v.apply(pts, x0)
    res, _, _ = engine.residuals(pts, s)
    assert np.array_equal(r.detach().cpu().numpy(), res)
    w = torch.linspace(-1.0, 1.0, r.shape[1], dtype=torch.float64, device="cuda")[None, :]
    (r * w).sum().backward()
    gx = x.grad.cpu().numpy()
    wn = w.cpu().numpy()[0]
    for i in range(3):
        for k in range(v.size):
            h = 1e-6 * max(abs(x0[i, k]), v.typical[k])
            xp, xm = x0[i].copy(), x0[i].copy()
            xp[k] += h
            xm[k] -= h
            pp = base.take([0, 0])
            v.apply(pp, np.stack([xp, xm]))
            rr, _, _ = engine.residuals(pp, s)
            fd = (rr[0] @ wn - rr[1] @ wn) / (2 * h)
            assert abs(gx[i, k] - fd) <= 1e-5 * max(1.0, abs(fd))


# ---- df64 -----------------------------------------------------------------------------------

@pytest.mark.parametrize("op", ["add", "sub", "mul", "mul_fp", "div", "sqrt"])
def test_df64_device_matches_oracle_bitwise(op):
    """The device df64 operations against the oracle's host restatement, bit for bit
    (CudaDf64.DeviceMatchesHostBitwise)."""
    import torch
    from mikroptim_amd.engine import library
    from oracle_lib import DF64_OPS, df64_ops, df64_split

    rng = np.random.default_rng(41)
    n = 100000
    x = rng.uniform(0.5, 2.0, n) * np.exp2(rng.integers(-20, 21, n).astype(np.float64))
    y = rng.uniform(-2.0, 2.0, n) * np.exp2(rng.integers(-20, 21, n).astype(np.float64))
    a, b = df64_split(x), df64_split(y)
    if op == "mul_fp":
        b[:, 1] = 0.0
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.empty_like(ta)
    rc = library().mko_df64_ops_device(DF64_OPS[op], n, ta.data_ptr(), tb.data_ptr(),
                                       out.data_ptr(), None)
    assert rc == 0
    torch.cuda.synchronize()
    dev = out.cpu().numpy()
    host = df64_ops(oracle(), op, a, b)
    assert np.array_equal(dev.view(np.uint32), host.view(np.uint32))


DF64_OPD_TOL = 1e-9        # mm, kDf64VsFp64 (tests/tolerances.hpp:182)
DF64_OPD_REL = 1e-10       # kDf64VsFp64Relative (:189)


@pytest.mark.parametrize("which", ["singlet_soft", "singlet_hard", "objective"])
def test_df64_trace_matches_fp64_oracle(engine, singlet, objective, which):
    """RESIDUALS traced in df64 against the FP64 oracle: equal statuses, OPD within
    max(1e-9 mm, 1e-10 |OPD|) per ray; MERIT in df64 equals the sum of squares of the df64
    residuals (kMeritVsResidualSumDf64 = 2e-13)."""
    if which == "objective":
        b, _ = objective
        s = systems.objective_settings(abi.FAILURE_HARD)
    else:
        b, _ = singlet
        s = systems.singlet_settings(abi.FAILURE_SOFT if which == "singlet_soft"
                                     else abi.FAILURE_HARD)
    s.precision = abi.PRECISION_DF64
    res, opd, st = engine.residuals(b, s)
    s64 = systems.ResidualSettings(s.rings, s.spokes, s.failure, s.fail_weight)
    _, opdo, sto = oracle().residuals(b, s64, threads=8)
    assert np.array_equal(st, sto)
    tol = np.maximum(DF64_OPD_TOL, DF64_OPD_REL * np.abs(opdo))
    assert np.all(np.abs(opd - opdo) <= tol)
    m, _, _ = engine.merit(b, s)
    assert np.max(np.abs(m - (res ** 2).sum(1)) / m) <= 2e-13


def test_uneven_groups_follow_the_oracle(engine):
    """Uneven batches (runs of different topologies): one LM per topology group, each against
    the oracle's LM on the same group (adjoint Jacobian, FP64), merits to 1e-7 after 5 cycles;
    the mixed-precision path decreases every run's merit."""
    groups = []
    for i in range(4):
        b, v = systems.uneven_topology(i, variants=3)
        groups.append((b.take([0]), systems.uneven_settings(), v, v.extract(b)))
    opt = abi.lm_options(max_iterations=5, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    results = engine.optimize_groups(groups, opt)
    mixed = engine.optimize_groups(groups, opt, precision="mixed")
    for (base, s, v, x0), res, rm in zip(groups, results, mixed):
        _, runs_o, _ = oracle().optimize_lm(base, s, v, x0, opt)
        mo = np.array([r.merit for r in runs_o])
        # adjoint LM, 5 cycles: 2.7e-9 relative at most measured over these groups
        assert np.all(np.abs(res.merit / mo - 1.0) <= 1e-7)
        pts = base.take(np.zeros(len(x0), dtype=np.int64))
        v.apply(pts, x0)
        m0, _, _ = engine.merit(pts, s)
        assert np.all(rm.merit < m0)

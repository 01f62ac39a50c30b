"""The HIP path on the benchmark configurations at their stated shapes:

* the singlet (stop + 2 spheres, 3 fields x 3 wavelengths, 6 x 12 Half pupil) at all 1024
  perturbed variants, RESIDUALS and MERIT against the oracle in both failure modes;
* the central-FD Jacobian of the 24-surface NA 0.95 objective (16 x 32 pupil, 17920 ray rows,
  48 variables, hard failure) against the oracle's FD Jacobian on every ray row, and at 4096
  runs: per-run independence of the FD normal equations, and FD columns against the FP64
  adjoint columns within the central difference's rounding bound;
* the mixed-precision normal equations at points the LM has converged towards, against the FP64
  stored-J path under absolute bounds derived from the split-bf16 products and the float
  staging (relative bounds on g fail there: g -> 0 while its rounding error does not).

Tolerances as in test_gpu_parity.py (the reference's tests/tolerances.hpp)."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from oracle_lib import oracle
from test_gpu_limits import _check_jacobians
from test_gpu_parity import OPD_TOL

pytestmark = pytest.mark.gpu

U = 2.0 ** -53       # FP64 unit roundoff
U_F = 2.0 ** -24     # float unit roundoff
U_BF = 2.0 ** -8     # bf16 unit roundoff (8 significand bits)
SAMPLE = np.array([0, 1, 517, 1024, 2047, 3000, 4094, 4095])


@pytest.fixture(scope="module")
def objective_full():
    b, v = systems.objective(variants=4096, seed=2)
    return b, v, systems.objective_settings(abi.FAILURE_HARD)


@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_singlet_all_1024_variants(engine, failure):
    b, _ = systems.singlet(1024)
    s = systems.singlet_settings(failure)
    res, opd, st = engine.residuals(b, s)
    ro, opdo, sto = oracle().residuals(b, s, threads=8)
    assert np.array_equal(st, sto)
    assert np.max(np.abs(opd - opdo)) <= OPD_TOL
    lam = b.reference_wavelength_nm * 1e-6
    assert np.max(np.abs(res - ro)) <= OPD_TOL / lam
    m, f, h = engine.merit(b, s)
    mo, fo, ho = oracle().merit(b, s, threads=8)
    assert np.array_equal(f, fo) and np.array_equal(h, ho)
    assert np.max(np.abs(m / mo - 1.0)) <= 1e-9
    # kMeritVsResidualSum: the same residuals summed in another order
    assert np.max(np.abs(m / np.sum(res * res, axis=1) - 1.0)) <= 1e-13


def test_objective_fd_jacobian_ray_rows_follow_the_oracle(engine):
    """Config 2's FD Jacobian on its 17920 ray rows (and the FP64 adjoint beside it) against the
    oracle's on two runs; the FD bound is the residuals' device-host difference over the step."""
    b, v = systems.objective(variants=64, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    _check_jacobians(engine, b, v, s, runs=2, fd=True)


def _fd_rounding_bound(b, h):
    """Column bound of |J_fd - J| from rounding: each trace's OPL carries at most ~S ulp of the
    largest optical length Lambda, the central difference divides two such errors by 2h, and the
    residual scales s_t sqrt(W_k) have sum_t s_t^2 sum_k W_k = 1 / lambda_ref^2 (normalised
    field and wavelength weights, pupil weights summing to 1 per tuple), so the column norm of
    the rounding part is at most 2 S u Lambda / (lambda_ref h). Lambda: twice the axial length
    times the largest index (the obliquity at NA 0.95 stays below 1 / cos(72 deg) ~ 3.2 on the
    last gap only)."""
    lam = b.reference_wavelength_nm * 1e-6
    axial = np.sum(np.abs(b.thickness), axis=1) + 10.0
    Lambda = 2.0 * np.max(b.index) * axial
    return 2.0 * b.surfaces * U * Lambda[:, None] / (lam * h)


def test_objective_fd_4096_runs_per_run_and_near_adjoint(engine, objective_full):
    b, v, s = objective_full
    base, x = b.take([0]), v.extract(b)
    _, _, a, g = engine.jacobian(base, s, v, x, method="fd", want_jacobian=False)
    assert np.isfinite(a).all() and np.isfinite(g).all()
    Jf, r0f, a_s, g_s = engine.jacobian(base, s, v, x[SAMPLE], method="fd")
    # a run's FD normal equations do not depend on the other runs or on the tranches
    assert np.array_equal(a_s, a[SAMPLE]) and np.array_equal(g_s, g[SAMPLE])
    J, r0, _, _ = engine.jacobian(base, s, v, x[SAMPLE], method="adjoint", precision="fp64")
    lam = b.reference_wavelength_nm * 1e-6
    assert np.max(np.abs(r0f - r0)) <= OPD_TOL / lam
    h = 2.0 ** (-52.0 / 3.0) * np.maximum(np.abs(x[SAMPLE]), v.typical)
    err = np.linalg.norm(Jf - J, axis=2)
    # truncation h^2/6 r''' of the second-order stencil: kAdjointJacobian (1e-6 per column)
    bound = 1e-6 * np.linalg.norm(J, axis=2) + _fd_rounding_bound(b.take(SAMPLE), h)
    assert np.all(err <= bound), float(np.max(err / bound))


def test_mixed_normal_equations_at_converged_points(engine, objective_full):
    """20 LM cycles on a sample of runs, then the mixed normal equations at the final points.

    With u_k = sqrt(W_k) (g_k - c) (c: a per-tuple shift near the mean gradient), the mixed path
    forms C = sum u u^T from bf16 hi + lo splits: the dropped lo*lo and the rounding of lo cost
    at most 3 u_bf^2 |u_a u_b| per product, the float accumulation over n <= 512 rays n u_f
    sum |u_a u_b|, the float staging of g_k - pilot 2 u_f; with Cauchy-Schwarz and ||u|| ~ ||J|| / s:
        ||dA||_F <= (3 u_bf^2 + 512 u_f + 2 u_f) ||J||_F^2  (~7.7e-5).
    g = J^T r is formed in FP64 from the float-staged gradients g_k = pilot + float(g_k - pilot),
    so dg_p = sum_k s^2 W_k dG_kp (OPL_k - piston) with |dG| <= u_f |g_k - pilot|; the pilot is
    one of the wave's rays, so ||s sqrt(W) (g_p - pilot_p)|| ~ sqrt(2) ||J_p||, and
        ||dg|| <= 4 u_f ||J||_F ||r||  (a factor ~3 above sqrt(2) u_f for the pilot's spread)."""
    b, v, s = objective_full
    base, x0 = b.take([0]), v.extract(b)[SAMPLE]
    opt = abi.lm_options(max_iterations=20, jacobian_method=abi.JACOBIAN_ADJOINT)
    res = engine.optimize(base, s, v, x0, opt, precision="mixed")
    m0, _, _ = engine.merit(b.take(SAMPLE), s)
    assert np.all(res.merit < m0)
    xs = res.x
    _, _, a, g = engine.jacobian(base, s, v, xs, method="adjoint", precision="mixed",
                                 want_jacobian=False)
    J, r0, ad, gd = engine.jacobian(base, s, v, xs, method="adjoint", precision="fp64")
    nJ = np.linalg.norm(J.reshape(len(xs), -1), axis=1)
    nr = np.linalg.norm(r0, axis=1)
    eps_a = 3 * U_BF ** 2 + 512 * U_F + 2 * U_F
    da = np.linalg.norm(a - ad, axis=(1, 2))
    dg = np.linalg.norm(g - gd, axis=1)
    assert np.all(da <= eps_a * nJ ** 2), float(np.max(da / (eps_a * nJ ** 2)))
    assert np.all(dg <= 4 * U_F * nJ * nr), float(np.max(dg / (4 * U_F * nJ * nr)))
    # the gradient has shrunk relative to its start, which is what makes relative bounds fail
    _, _, _, g0 = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    assert np.all(np.linalg.norm(gd, axis=1) < np.linalg.norm(g0, axis=1))

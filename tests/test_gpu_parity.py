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
    assert err.max() <= max(OPD_TOL, 1e-10 * 0)  # absolute bound per ray


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


def test_adjoint_mixed_objective(engine, objective):
    b, v = objective
    s = systems.objective_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:2]
    base = b.take([0])
    J, r0, a, g = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed")
    Jo, r0o, _ = oracle().adjoint_jacobian(base, s, v, x0, threads=8)
    rel = _col_rel(J, Jo)
    print("mixed adjoint: max column relative error", rel.max())
    assert rel.max() <= 1e-4
    # r0 comes from the FP64 forward trace in both precisions
    assert np.max(np.abs(r0 - r0o)) <= OPD_TOL / (b.reference_wavelength_nm * 1e-6)


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
    for method in (abi.JACOBIAN_FD, abi.JACOBIAN_ADJOINT):
        opt = abi.lm_options(max_iterations=15, threads=4, jacobian_method=method)
        res = engine.optimize(base, s, v, x0, opt)
        xo, runs_o, cyc_o = oracle().optimize_lm(base, s, v, x0, opt)
        mo = np.array([r.merit for r in runs_o])
        assert np.all(res.merit <= mo * (1 + 1e-6) + 1e-12)  or np.allclose(res.merit, mo, rtol=1e-4)
        assert np.allclose(res.merit, mo, rtol=1e-3)

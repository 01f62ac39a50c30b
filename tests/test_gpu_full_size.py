"""The HIP path at the full workload: 4096 variants x 48 variables of the 24-surface NA 0.95
objective, 5 fields x 7 wavelengths x 16x32 Half pupil (17920 rays per variant), hard failure.
The oracle is too slow for every variant at this size, so the checks are properties that hold at
any size, plus the oracle on a sample of variants spread over the batch:

* MERIT is repeatable bit for bit and equals the sum of squared RESIDUALS rows (the identity
  F = sum r^2, docs/architektur.md:613; the device sums in a fixed tree, numpy pairwise, so the
  two orders differ by at most a few ulp per row);
* results of a variant do not depend on the others (include/mikroptim/optim/
  levenberg_marquardt.hpp:167-170): the sample alone equals the sample inside the full batch bit
  for bit, for MERIT, the mixed normal equations and the LM;
* the FP64 adjoint Jacobian of the sample against the oracle's, column by column;
* the mixed normal equations against the FP64 stored-J path, and the LM's merit never above its
  start. Tolerances as in test_gpu_parity.py."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from oracle_lib import oracle
from test_gpu_parity import ADJ_TOL, MIXED_A_TOL, MIXED_DIAG_TOL, MIXED_G_TOL, _col_rel

pytestmark = pytest.mark.gpu

SAMPLE = np.array([0, 1, 517, 1024, 2047, 3000, 4094, 4095])


@pytest.fixture(scope="module")
def full():
    b, v = systems.objective(variants=4096, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    return b, v, s


def test_full_merit_is_repeatable_residual_sum_and_follows_the_oracle(engine, full):
    b, v, s = full
    m, f, h = engine.merit(b, s)
    m2, f2, h2 = engine.merit(b, s)
    assert np.array_equal(m, m2) and np.array_equal(f, f2) and np.array_equal(h, h2)
    assert (f == 0).all() and (h == 0).all()
    res, _, st = engine.residuals(b, s)
    assert (st == 0).all()
    rsum = np.sum(res * res, axis=1)
    assert np.max(np.abs(m / rsum - 1.0)) <= 1e-12
    sub = b.take(SAMPLE)
    ms, _, _ = engine.merit(sub, s)
    assert np.array_equal(ms, m[SAMPLE])
    mo, _, _ = oracle().merit(sub, s, threads=8)
    assert np.max(np.abs(ms / mo - 1.0)) <= 1e-11


def test_full_normal_equations_are_per_run_and_near_fp64(engine, full):
    b, v, s = full
    base, x = b.take([0]), v.extract(b)
    _, _, a, g = engine.jacobian(base, s, v, x, method="adjoint", precision="mixed",
                                 want_jacobian=False)
    assert np.isfinite(a).all() and np.isfinite(g).all()
    _, _, a_s, g_s = engine.jacobian(base, s, v, x[SAMPLE], method="adjoint", precision="mixed",
                                     want_jacobian=False)
    assert np.array_equal(a_s, a[SAMPLE]) and np.array_equal(g_s, g[SAMPLE])
    J, r0, ad, gd = engine.jacobian(base, s, v, x[SAMPLE], method="adjoint", precision="fp64")
    ea = np.linalg.norm(a_s - ad, axis=(1, 2)) / np.linalg.norm(ad, axis=(1, 2))
    diag = np.abs(np.diagonal(a_s, axis1=1, axis2=2) / np.diagonal(ad, axis1=1, axis2=2) - 1)
    eg = np.linalg.norm(g_s - gd, axis=1) / np.linalg.norm(gd, axis=1)
    assert ea.max() <= MIXED_A_TOL and diag.max() <= MIXED_DIAG_TOL and eg.max() <= MIXED_G_TOL
    Jo, r0o, _ = oracle().adjoint_jacobian(base, s, v, x[SAMPLE], threads=8)
    assert _col_rel(J, Jo).max() <= ADJ_TOL


def test_full_lm_runs_are_independent_and_never_worse(engine, full):
    b, v, s = full
    base, x0 = b.take([0]), v.extract(b)
    opt = abi.lm_options(max_iterations=3, jacobian_method=abi.JACOBIAN_ADJOINT)
    r = engine.optimize(base, s, v, x0, opt, precision="mixed")
    rs = engine.optimize(base, s, v, x0[SAMPLE], opt, precision="mixed")
    assert np.array_equal(r.x[SAMPLE], rs.x) and np.array_equal(r.merit[SAMPLE], rs.merit)
    assert [r.status[i] for i in SAMPLE] == rs.status
    m0, _, _ = engine.merit(b, s)
    assert np.all(r.merit <= m0) and np.mean(r.merit) < np.mean(m0)

"""Glass variables (Buchdahl n0, nu1, nu2 of a medium; VariableSet with VariableType::Glass,
src/optim/variable_set.cpp:145-285, AdjointParameterMap::glass, ray_jacobian.hpp:28-75) in the
oracle, pinned bit for bit against the reference's own CPU code (residuals after apply, FD and
adjoint Jacobians, the LM), the adjoint's index columns against central differences of the
oracle itself, and the Python VariableSet.apply against the oracle's apply."""

import ctypes as C

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd.batch import buchdahl_index, buchdahl_omega
from oracle_lib import oracle, reference


def _need_ref():
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    return ref


def test_buchdahl_closed_form():
    # w = dl / (1 + 2.5 dl), n = n0 + nu1 w + nu2 w^2 + nu3 w^3 at 0.6 um
    dl = 0.6 - 0.546074
    w = dl / (1.0 + 2.5 * dl)
    assert abs(buchdahl_omega(600.0) - w) <= 1e-17
    c = (1.5, -0.08, 0.02, 0.004)
    assert abs(buchdahl_index(c, 600.0) - (1.5 - 0.08 * w + 0.02 * w * w + 0.004 * w ** 3)) <= 1e-15
    assert buchdahl_omega(546.074) == 0.0 and buchdahl_index(c, 546.074) == 1.5
    with pytest.raises(ValueError):
        buchdahl_omega(146.0)


def test_python_apply_equals_oracle_apply():
    b, v = systems.glass_singlet(3)
    x = v.extract(b)
    x[:, 2] += 0.01
    x[:, 3] *= 1.1
    py = b.copy()
    v.apply(py, x)
    orc = b.copy()
    f = oracle().fn("apply")
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    cb, cv = orc.c_struct(), v.c_struct()
    for i in range(orc.variants):
        oracle().check(f(C.byref(cb), C.byref(cv), i, abi.ptr(np.ascontiguousarray(x[i]))))
    for name in ("index", "paraxial_index", "buchdahl", "curvature", "thickness"):
        assert np.array_equal(getattr(py, name), getattr(orc, name)), name


@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_oracle_equals_reference_with_glass_variables(failure):
    ref = _need_ref()
    b, v = systems.glass_singlet()
    s = systems.singlet_settings(failure)
    for a, c in zip(oracle().residuals(b, s), ref.residuals(b, s)):
        assert np.array_equal(a, c)
    base, x0 = b.take([0]), v.extract(b)[:3]
    Ja, ra, _ = oracle().adjoint_jacobian(base, s, v, x0)
    Jr, rr, _ = ref.adjoint_jacobian(base, s, v, x0)
    assert np.array_equal(Ja, Jr) and np.array_equal(ra, rr)
    Fa, fa, _ = oracle().fd_jacobian(base, s, v, x0)
    Fr, fr, _ = ref.fd_jacobian(base, s, v, x0)
    assert np.array_equal(Fa, Fr) and np.array_equal(fa, fr)


def test_glass_columns_are_derivatives():
    """Index columns of the adjoint against central differences of the oracle's residuals in
    n0, nu1, nu2, with steps that move the index by about 1e-7 (nu1 acts through w ~ 0.07, nu2
    through w^2 ~ 0.005 at these wavelengths)."""
    b, v = systems.glass_singlet(1)
    s = systems.singlet_settings(abi.FAILURE_HARD)
    x0 = v.extract(b)[:1]
    J, r0, _ = oracle().adjoint_jacobian(b, s, v, x0)
    for k, h in ((2, 1e-7), (3, 2e-6), (4, 3e-5)):
        bp, bm = b.copy(), b.copy()
        xp, xm = x0.copy(), x0.copy()
        xp[0, k] += h
        xm[0, k] -= h
        v.apply(bp, xp)
        v.apply(bm, xm)
        rp, _, _ = oracle().residuals(bp, s)
        rm, _, _ = oracle().residuals(bm, s)
        fd = (rp[0] - rm[0]) / (2 * h)
        assert np.linalg.norm(J[0, k] - fd) <= 1e-5 * np.linalg.norm(fd)


def test_lm_with_glass_variables_equals_reference():
    ref = _need_ref()
    b, v = systems.glass_singlet()
    s = systems.singlet_settings(abi.FAILURE_HARD)
    base, x0 = b.take([0]), v.extract(b)[:2]
    for method in (abi.JACOBIAN_FD, abi.JACOBIAN_ADJOINT):
        opt = abi.lm_options(max_iterations=8, jacobian_method=method)
        xa, ra, _ = oracle().optimize_lm(base, s, v, x0, opt)
        xr, rr, _ = ref.optimize_lm(base, s, v, x0, opt)
        assert np.array_equal(xa, xr)
        assert [r.merit for r in ra] == [r.merit for r in rr]
        assert [r.iterations for r in ra] == [r.iterations for r in rr]


def test_glass_variables_need_coefficients():
    b, v = systems.glass_singlet(2)
    b.buchdahl = None
    s = systems.singlet_settings()
    with pytest.raises(Exception):
        oracle().adjoint_jacobian(b.take([0]), s, v, v.extract(systems.glass_singlet(2)[0])[:1])
    b, v = systems.glass_singlet(2)
    v.component = np.array([0, 0, 0, 1, 3, 0], dtype=np.int32)
    with pytest.raises(Exception):
        oracle().fd_jacobian(b.take([0]), s, v, v.extract(b)[:1])

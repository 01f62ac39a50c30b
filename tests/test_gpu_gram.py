"""Normal equations on FP64 MFMA (MKO_PRECISION_FP64_MFMA, mko_gram_mfma_device) against the
host's FMA chain (mko_gram_device, bitwise normal_equations_host): both sum the same FP64
products (the MFMA unit has been measured to give the chain's bits, but its order is not a
documented contract), so they may differ by rounding only, within the summation bound
|dG| <= 2 R u ||J||_F^2 (R rows, u = 2^-53) in the Frobenius norm; the LM on them takes device
and host control bit for bit. The 24-surface objective (48 variables, 17920 rows, FD and adjoint
Jacobians) and the uneven workload's 96-variable topology (3 and 9 tiles per wave)."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def _near_chain(engine, base, s, v, x0, method):
    J, r0, ad, gd = engine.jacobian(base, s, v, x0, method=method, precision="fp64")
    _, _, am, gm = engine.jacobian(base, s, v, x0, method=method, precision="fp64_mfma",
                                   want_jacobian=False)
    R = J.shape[2]
    nJ = np.linalg.norm(J.reshape(len(x0), -1), axis=1)
    nr = np.linalg.norm(r0, axis=1)
    assert np.all(np.linalg.norm(am - ad, axis=(1, 2)) <= 2 * R * U * nJ ** 2)
    assert np.all(np.linalg.norm(gm - gd, axis=1) <= 2 * R * U * nJ * nr)
    assert np.array_equal(am, np.swapaxes(am, 1, 2))
    return am, ad


@pytest.mark.parametrize("method", ["adjoint", "fd"])
def test_mfma_normal_equations_objective(engine, method):
    b, v = systems.objective(variants=64, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    _near_chain(engine, b.take([0]), s, v, v.extract(b)[:4], method)


def test_mfma_normal_equations_96_variables(engine):
    b, v = systems.uneven_topology(0, variants=3, surfaces=32, variables=96)
    s = systems.uneven_settings(abi.FAILURE_HARD)
    _near_chain(engine, b.take([0]), s, v, v.extract(b), "adjoint")


def test_mfma_lm_device_equals_host_control(engine):
    b, v = systems.objective(variants=16, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    s.rings, s.spokes = 4, 8
    base, x0 = b.take([0]), v.extract(b)[:4]
    kw = dict(max_iterations=8, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rd = engine.optimize(base, s, v, x0, abi.lm_options(**kw), precision="fp64_mfma")
    rh = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw),
                         precision="fp64_mfma")
    assert np.array_equal(rd.x, rh.x) and np.array_equal(rd.merit, rh.merit)
    pts = base.take(np.zeros(4, dtype=np.int64))
    v.apply(pts, x0)
    m0, _, _ = engine.merit(pts, s)
    assert np.all(rd.merit < m0)

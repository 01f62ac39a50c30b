"""Glass variables on the HIP path: the point batch sets the Buchdahl coefficients and the
glass surfaces' indices on the device, the FD Jacobian differences those traces, the FP64 adjoint
carries the index adjoints into the (n0, nu1, nu2) columns, the fused mixed-precision normal
equations carry them too (their reverse sweep starts from a MERIT trace's stored ray states), and
the LM runs with them, device control equal to host control bit for bit. Tolerances as in
test_gpu_parity.py / test_gpu_limits.py / test_gpu_configs.py."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from oracle_lib import oracle
from test_gpu_configs import U_BF, U_F
from test_gpu_limits import _check_jacobians, _check_trace

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_glass_traces_and_jacobians_match_oracle(engine, failure):
    b, v = systems.glass_singlet()
    s = systems.singlet_settings(failure)
    _check_trace(engine, b, s)
    _check_jacobians(engine, b, v, s, runs=3)


def test_glass_mixed_precision_near_fp64(engine):
    """The fused normal equations with glass columns against the FP64 stored-J path, under the
    absolute bounds of the split-bf16 products and the float staging."""
    b, v = systems.glass_singlet()
    s = systems.singlet_settings(abi.FAILURE_HARD)
    base, x0 = b.take([0]), v.extract(b)[:3]
    _, _, am, gm = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed",
                                   want_jacobian=False)
    J, r0, ad, gd = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    assert not np.array_equal(am, ad)  # the fused path ran, not the FP64 one
    nJ = np.linalg.norm(J.reshape(len(x0), -1), axis=1)
    nr = np.linalg.norm(r0, axis=1)
    eps_a = 3 * U_BF ** 2 + 512 * U_F + 2 * U_F
    assert np.all(np.linalg.norm(am - ad, axis=(1, 2)) <= eps_a * nJ ** 2)
    assert np.all(np.linalg.norm(gm - gd, axis=1) <= 4 * U_F * nJ * nr)
    # the glass columns' own diagonal entries
    glass = np.flatnonzero(np.asarray(v.type) == abi.VAR_GLASS)
    dd = np.abs(am[:, glass, glass] / ad[:, glass, glass] - 1)
    assert dd.max() <= 5e-5


@pytest.mark.parametrize("method", [abi.JACOBIAN_FD, abi.JACOBIAN_ADJOINT])
def test_glass_lm_device_equals_host(engine, method):
    b, v = systems.glass_singlet()
    s = systems.singlet_settings(abi.FAILURE_HARD)
    base, x0 = b.take([0]), v.extract(b)[:3]
    kw = dict(max_iterations=10, threads=4, jacobian_method=method)
    for precision in ("fp64", "mixed"):
        rd = engine.optimize(base, s, v, x0, abi.lm_options(**kw), precision=precision)
        rh = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw),
                             precision=precision)
        assert np.array_equal(rd.x, rh.x) and np.array_equal(rd.merit, rh.merit)
        assert rd.status == rh.status and rd.cycles == rh.cycles


def test_glass_lm_follows_oracle(engine):
    b, v = systems.glass_singlet()
    s = systems.singlet_settings(abi.FAILURE_HARD)
    base, x0 = b.take([0]), v.extract(b)[:3]
    pts = base.take(np.zeros(3, dtype=np.int64))
    v.apply(pts, x0)
    m0, _, _ = engine.merit(pts, s)
    # merits after 10 cycles: FD 1.3e-7, adjoint 1.8e-12 relative measured
    for method, rtol in ((abi.JACOBIAN_FD, 1e-5), (abi.JACOBIAN_ADJOINT, 1e-9)):
        opt = abi.lm_options(max_iterations=10, threads=4, jacobian_method=method)
        res = engine.optimize(base, s, v, x0, opt)
        _, runs_o, _ = oracle().optimize_lm(base, s, v, x0, opt)
        assert np.all(np.abs(res.merit / np.array([r.merit for r in runs_o]) - 1.0) <= rtol)
        assert np.all(res.merit < m0)

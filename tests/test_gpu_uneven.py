"""The uneven workload at the edges of its stated range: a topology of 32 refracting surfaces with
96 variables (31 curvatures, 31 inner thicknesses, the working distance and the Buchdahl
coefficients of the glass media), hard failure.

* FP64 adjoint columns against the oracle's (kAdjointCpuGpu), glass columns included;
* the fused mixed-precision normal equations (index adjoints in the reverse sweep, 3 x 3 MFMA
  tiles) against the FP64 stored-J path, under the absolute bounds of test_gpu_configs.py;
* the LM control on the device (two variables per lane past 64) against the host control, bit
  for bit, on the FP64 path; and the captured mixed-precision cycle against the eager one.

A group's LM follows the oracle's on its own (the reference holds one topology per batch), as
test_gpu_parity.py::test_uneven_groups_follow_the_oracle checks for drawn groups."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd.engine import TraceEngine
from oracle_lib import oracle
from test_gpu_configs import U_BF, U_F
from test_gpu_parity import ADJ_TOL, _col_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wide():
    b, v = systems.uneven_topology(0, variants=4, surfaces=32, variables=96)
    assert v.size == 96 and v.has_glass
    s = systems.uneven_settings(abi.FAILURE_HARD)
    return b, v, s


def test_96_variables_fp64_adjoint_follows_the_oracle(engine, wide):
    b, v, s = wide
    base, x0 = b.take([0]), v.extract(b)[:2]
    J, r0, _, _ = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    Jo, r0o, _ = oracle().adjoint_jacobian(base, s, v, x0, threads=8)
    assert _col_rel(J, Jo).max() <= ADJ_TOL


def test_96_variables_fused_normal_equations_near_fp64(engine, wide):
    b, v, s = wide
    base, x0 = b.take([0]), v.extract(b)
    _, _, a, g = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed",
                                 want_jacobian=False)
    J, r0, ad, gd = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    nJ = np.linalg.norm(J.reshape(len(x0), -1), axis=1)
    nr = np.linalg.norm(r0, axis=1)
    eps_a = 3 * U_BF ** 2 + 512 * U_F + 2 * U_F
    assert np.all(np.linalg.norm(a - ad, axis=(1, 2)) <= eps_a * nJ ** 2)
    assert np.all(np.linalg.norm(g - gd, axis=1) <= 4 * U_F * nJ * nr)
    # every column of the fused path, the glass ones included, carries its share
    diag = np.abs(np.diagonal(a, axis1=1, axis2=2) / np.diagonal(ad, axis1=1, axis2=2) - 1)
    assert diag.max() <= 5e-5


def test_96_variables_device_control_equals_host_control(engine, wide):
    b, v, s = wide
    base, x0 = b.take([0]), v.extract(b)
    kw = dict(max_iterations=6, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rd = engine.optimize(base, s, v, x0, abi.lm_options(**kw), precision="fp64")
    rh = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw),
                         precision="fp64")
    assert np.array_equal(rd.x, rh.x) and np.array_equal(rd.merit, rh.merit)
    assert rd.status == rh.status and np.array_equal(rd.iterations, rh.iterations)
    m0, _, _ = engine.merit(b, s)
    assert np.all(rd.merit < m0)


def test_96_variables_graph_cycles_equal_eager_cycles(engine, wide):
    b, v, s = wide
    base, x0 = b.take([0]), v.extract(b)
    kw = dict(max_iterations=6, jacobian_method=abi.JACOBIAN_ADJOINT)
    g = engine.lm(base, s, v, x0, abi.lm_options(**kw), precision="mixed")
    other = TraceEngine(0)
    try:
        e = other.lm(base, s, v, x0, abi.lm_options(graph=False, **kw), precision="mixed")
        while True:
            more_g, more_e = g.cycle(), e.cycle()
            assert more_g == more_e
            rg, re_ = g.results(), e.results()
            assert np.array_equal(rg.x, re_.x) and np.array_equal(rg.merit, re_.merit)
            if not more_g:
                break
        assert g.graph_cycles() > 0 and e.graph_cycles() == 0
        m0, _, _ = engine.merit(b, s)
        assert np.all(g.results().merit < m0)
        e.close()
    finally:
        g.close()
        other.close()

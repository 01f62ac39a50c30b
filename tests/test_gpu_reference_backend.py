"""The reference's own LM control driving the HIP library: optimize_lm_rounds
(src/optim/levenberg_marquardt.cpp:891) with HipLmBackend, the reference-side adapter that
answers mko::LmBackend's two pure virtuals (include/mikroptim/optim/levenberg_marquardt.hpp:
171-182) with mko_engine_normal_equations and mko_engine_point_merit (oracle/ref_driver.cpp).
This is the drop-in boundary as the reference's code uses it. The library's own host control
(host_control = 1) restates the same control logic over the same two calls, so both take the same
steps bit for bit. Needs the reference build (oracle/_ref)."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from oracle_lib import reference

pytestmark = pytest.mark.gpu


def _same(ref_result, lib_result):
    x, runs, cycles = ref_result
    assert np.array_equal(x, lib_result.x)
    assert [r.merit for r in runs] == list(lib_result.merit)
    assert [abi.LM_STATUS_NAMES[r.status] for r in runs] == lib_result.status
    assert [r.iterations for r in runs] == list(lib_result.iterations)
    assert [r.penalty_rounds for r in runs] == [r.penalty_rounds for r in lib_result.runs]
    assert cycles == lib_result.cycles


@pytest.mark.parametrize("method,precision,failure", [
    (abi.JACOBIAN_ADJOINT, "fp64", abi.FAILURE_HARD),
    (abi.JACOBIAN_ADJOINT, "fp64", abi.FAILURE_SOFT),
    (abi.JACOBIAN_FD, "fp64", abi.FAILURE_HARD),
    (abi.JACOBIAN_ADJOINT, "mixed", abi.FAILURE_HARD),
])
def test_reference_control_over_hip_backend_singlet(engine, method, precision, failure):
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    b, v = systems.singlet(8)
    s = systems.singlet_settings(failure)
    base, x0 = b.take([0]), v.extract(b)[:4]
    kw = dict(max_iterations=12, jacobian_method=method, threads=4)
    got = ref.optimize_lm_hip_backend(engine, base, s, v, x0, abi.lm_options(**kw), precision)
    lib = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw),
                          precision=precision)
    _same(got, lib)


def test_reference_control_over_hip_backend_objective(engine):
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    b, v = systems.objective(variants=16, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    s.rings, s.spokes = 8, 16
    base, x0 = b.take(np.arange(3)), v.extract(b)[:3]
    kw = dict(max_iterations=6, jacobian_method=abi.JACOBIAN_ADJOINT, threads=4)
    got = ref.optimize_lm_hip_backend(engine, base, s, v, x0, abi.lm_options(**kw), "mixed")
    lib = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw),
                          precision="mixed")
    _same(got, lib)

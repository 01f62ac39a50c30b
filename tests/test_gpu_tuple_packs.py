"""Tuple packs (tuples of at most 128 rays sharing a 256-lane workgroup in segments of 32, 64 or
128 lanes): RESIDUALS (rows, OPD, status), MERIT (merit, failed and hard-failed counts) and the
LM that runs on the packed MERIT trace's stored ray states are bit for bit those of one tuple per
workgroup, for every pack size (CudaTuplePacks.ResultsAreBitIdenticalForAllPackSizes,
tests/test_cuda_tuple_packs.cpp:61 of the reference). The singlet (6-ray Axis and 72-ray Half
tuples) in both failure modes, with failing rays in soft mode, and a topology of the uneven
workload (48-ray tuples)."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems

pytestmark = pytest.mark.gpu

PACKS = (32, 64, 128, 0)


def _outputs(engine, b, s, pack):
    engine.tuning(pack_rays=pack)
    try:
        res, opd, st = engine.residuals(b, s)
        m, f, h = engine.merit(b, s)
    finally:
        engine.tuning(pack_rays=0)
    return res, opd, st, m, f, h


def _workloads():
    b, _ = systems.singlet(12)
    x = b.curvature.copy()
    x[6:, 1] = np.linspace(0.19, 0.24, 6)  # front radius near the pupil radius: soft failures
    b.curvature[:] = x
    yield b, systems.singlet_settings(abi.FAILURE_SOFT)
    yield b, systems.singlet_settings(abi.FAILURE_HARD)
    u, _ = systems.uneven_topology(3, variants=5)
    yield u, systems.uneven_settings(abi.FAILURE_HARD)


@pytest.mark.parametrize("case", range(3))
def test_packed_traces_equal_unpacked_for_every_pack_size(engine, case):
    b, s = list(_workloads())[case]
    ref = _outputs(engine, b, s, 256)
    if case == 0:
        assert (ref[2] != 0).any()  # soft failures present
    for pack in PACKS:
        out = _outputs(engine, b, s, pack)
        for x, y in zip(out, ref):
            assert np.array_equal(x, y), pack


def test_lm_on_packed_states_equals_unpacked(engine):
    """The mixed-precision LM (device control, captured cycle) reads the MERIT trace's stored ray
    states; with packed traces it takes the same steps bit for bit."""
    b, v = systems.uneven_topology(5, variants=6)
    s = systems.uneven_settings(abi.FAILURE_HARD)
    base, x0 = b.take([0]), v.extract(b)
    opt = abi.lm_options(max_iterations=6, jacobian_method=abi.JACOBIAN_ADJOINT)
    runs = {}
    for pack in (256, 0):
        engine.tuning(pack_rays=pack)
        try:
            runs[pack] = engine.optimize(base, s, v, x0, opt, precision="mixed")
        finally:
            engine.tuning(pack_rays=0)
    assert np.array_equal(runs[256].x, runs[0].x)
    assert np.array_equal(runs[256].merit, runs[0].merit)


def test_tuning_rejects_unknown_shapes(engine):
    with pytest.raises(Exception):
        engine.tuning(pack_rays=48)
    with pytest.raises(Exception):
        engine.tuning(fused_waves=3)
    assert engine.tuning() == (0, 0)

"""Parity of the HIP path with the oracle at the edges of what a batch holds and on the source and
reference-sphere kinds the other GPU tests leave out: point sources (Axis / Half / Full pupils),
the image reference spheres (exit pupil, infinite), 64 surfaces with 64 variables (the fused
normal equations' and the device LM control's maximum), tuples of 1024 and 2048 rays (4 and 8
rays per lane), tuples of 160 and 256 spokes, and 32 fields x 32 wavelengths. The oracle equals
the reference's own CPU code bit for bit on the same workloads (tests/test_oracle.py). Tolerances
as in test_gpu_parity.py."""

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd.batch import layout
from oracle_lib import oracle
from test_gpu_parity import ADJ_TOL, MIXED_A_TOL, MIXED_DIAG_TOL, MIXED_G_TOL, OPD_TOL, _col_rel

pytestmark = pytest.mark.gpu


def _check_trace(engine, b, s):
    """OPD per ray within OPD_TOL once the tuple's common offset is removed; that offset is the
    piston, a weighted mean over up to 2048 rays summed in another order on the device (a tree)
    than on the host (sequential), so it may differ by up to 2 n u max|OPL| (n rays, u = 2^-53)."""
    res, opd, st = engine.residuals(b, s)
    ro, opdo, sto = oracle().residuals(b, s, threads=8)
    assert np.array_equal(st, sto)
    offsets, samples, _ = layout(b, s)
    d = opd - opdo
    for f, (off, n) in enumerate(zip(offsets, samples)):
        for l in range(b.wavelengths):
            seg = slice(off + l * n, off + (l + 1) * n)
            ok = sto[:, seg] == 0
            for v in range(b.variants):
                dv = d[v, seg][ok[v]]
                if dv.size == 0:
                    continue
                common = 0.5 * (dv.max() + dv.min())
                assert np.max(np.abs(dv - common)) <= OPD_TOL
                # |OPL| <= source distance + axial length + a margin for the obliquity (mm)
                opl_max = 2.0 * (abs(b.source_point[2]) + np.sum(b.thickness[v]) + 10.0)
                assert abs(common) <= max(OPD_TOL, 2 * n * 2.0 ** -53 * opl_max)
    m, f, h = engine.merit(b, s)
    mo, fo, ho = oracle().merit(b, s, threads=8)
    assert np.array_equal(f, fo) and np.array_equal(h, ho)
    assert np.max(np.abs(m / mo - 1.0)) <= 1e-11
    return st


def _check_jacobians(engine, b, v, s, runs=2, fd=True):
    base, x0 = b.take([0]), v.extract(b)[:runs]
    J, r0, _, _ = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    Jo, r0o, _ = oracle().adjoint_jacobian(base, s, v, x0, threads=8)
    assert _col_rel(J, Jo).max() <= ADJ_TOL
    if fd:
        # FD columns inherit the residuals' device-host difference (OPD_TOL / lambda_ref per
        # row) divided by the central step h_k = 2^(-52/3) max(|x_k|, typical_k)
        Jf, _, _, _ = engine.jacobian(base, s, v, x0, method="fd")
        Jfo, _, _ = oracle().fd_jacobian(base, s, v, x0, threads=8)
        h = 2.0 ** (-52.0 / 3.0) * np.maximum(np.abs(x0), v.typical)
        bound = (OPD_TOL / (b.reference_wavelength_nm * 1e-6)) / h[:, :, None]
        assert np.all(np.abs(Jf - Jfo) <= bound + 1e-5 * np.abs(Jfo))


def _check_mixed(engine, b, v, s, runs=2):
    base, x0 = b.take([0]), v.extract(b)[:runs]
    _, _, a, g = engine.jacobian(base, s, v, x0, method="adjoint", precision="mixed",
                                 want_jacobian=False)
    _, _, ad, gd = engine.jacobian(base, s, v, x0, method="adjoint", precision="fp64")
    ea = np.linalg.norm(a - ad, axis=(1, 2)) / np.linalg.norm(ad, axis=(1, 2))
    diag = np.abs(np.diagonal(a, axis1=1, axis2=2) / np.diagonal(ad, axis1=1, axis2=2) - 1)
    eg = np.linalg.norm(g - gd, axis=1) / np.linalg.norm(gd, axis=1)
    assert ea.max() <= MIXED_A_TOL and diag.max() <= MIXED_DIAG_TOL and eg.max() <= MIXED_G_TOL


# ---- point sources --------------------------------------------------------------------------

POINTS = {"axis": (0.0, 0.0, -300.0), "half": (0.0, 2.0, -300.0), "full": (1.0, 1.5, -300.0)}


@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
@pytest.mark.parametrize("point", sorted(POINTS))
def test_point_source(engine, point, failure):
    b, v = systems.point_singlet(POINTS[point])
    s = systems.singlet_settings(failure)
    _check_trace(engine, b, s)
    _check_jacobians(engine, b, v, s)
    if failure == abi.FAILURE_HARD and point != "axis":
        _check_mixed(engine, b, v, s)


# ---- image reference spheres ------------------------------------------------------------------

@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
@pytest.mark.parametrize("kind", [abi.REF_IMAGE_EXIT_PUPIL, abi.REF_IMAGE_INFINITE])
def test_image_reference_spheres(engine, kind, failure):
    from test_oracle import _image_reference_singlet
    b, v = _image_reference_singlet(kind)
    s = systems.singlet_settings(failure)
    _check_trace(engine, b, s)
    _check_jacobians(engine, b, v, s)
    if failure == abi.FAILURE_HARD:
        _check_mixed(engine, b, v, s)


# ---- 64 surfaces, 64 variables ------------------------------------------------------------------

def _relay():
    b, v = systems.long_relay()
    return b, v, systems.ResidualSettings(rings=16, spokes=32, failure=abi.FAILURE_HARD)


def test_relay_64_surfaces_trace_and_jacobians(engine):
    b, v, s = _relay()
    st = _check_trace(engine, b, s)
    assert (st == 0).all()
    _check_jacobians(engine, b, v, s, runs=1)


def test_relay_64_variables_fused_normal_equations(engine):
    b, v, s = _relay()
    _check_mixed(engine, b, v, s, runs=2)


def test_relay_64_variables_lm_device_equals_host(engine):
    b, v, s = _relay()
    base, x0 = b.take([0]), v.extract(b)[:2]
    kw = dict(max_iterations=3, threads=4, jacobian_method=abi.JACOBIAN_ADJOINT)
    rd = engine.optimize(base, s, v, x0, abi.lm_options(**kw), precision="mixed")
    rh = engine.optimize(base, s, v, x0, abi.lm_options(host_control=True, **kw),
                         precision="mixed")
    assert np.array_equal(rd.x, rh.x) and np.array_equal(rd.merit, rh.merit)
    assert rd.status == rh.status and rd.cycles == rh.cycles


# ---- big tuples: 4 and 8 rays per lane ------------------------------------------------------------

@pytest.mark.parametrize("spokes", [64, 128])
def test_big_tuples(engine, spokes):
    b, v = systems.point_singlet(POINTS["full"], variants=2)
    s = systems.ResidualSettings(rings=16, spokes=spokes, failure=abi.FAILURE_SOFT)
    _check_trace(engine, b, s)
    _check_jacobians(engine, b, v, s, runs=1, fd=False)


# ---- more spokes than the trace kernel's sample table holds ---------------------------------------

@pytest.mark.parametrize("rings,spokes", [(2, 160), (1, 256)])
def test_many_spokes(engine, rings, spokes):
    """Tuples of more than 128 spokes: the trace kernel forms each ray's pupil sample on its own
    instead of reading the workgroup's LDS table of rings and spokes. 256 spokes divide the fused
    kernel's 256-ray chunk (a lane keeps its spoke from chunk to chunk), 160 do not."""
    b, v = systems.singlet(variants=2)
    s = systems.ResidualSettings(rings=rings, spokes=spokes, failure=abi.FAILURE_HARD)
    _check_trace(engine, b, s)
    _check_jacobians(engine, b, v, s, runs=2, fd=False)
    _check_mixed(engine, b, v, s, runs=2)


# ---- 32 fields x 32 wavelengths --------------------------------------------------------------------

@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_32_fields_32_wavelengths(engine, failure):
    b, v = systems.many_tuples()
    s = systems.ResidualSettings(rings=2, spokes=3, failure=failure)
    _check_trace(engine, b, s)
    _check_jacobians(engine, b, v, s, runs=1)

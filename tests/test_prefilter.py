"""Pins the FP32 prefilter: the oracle's float restatement of the per-variant stats against the
reference's own CPU float path (prefilter_stats_cpu, bit for bit through oracle/_ref, else the
committed fixtures), the library's decision (mko_prefilter_decide, host code) against the
reference's prefilter_decide bit for bit, and the properties the decision rests on
(docs of the reference: assumption (A) and the merit bounds): certain-failed <= FP64 failed <=
certain-failed + uncertain, L_v <= F_FP64 <= U_v, and no variant rejected for a reason its FP64
values do not have (the acceptance criterion of CudaPrefilter.AcceptanceNoFalseRejections)."""

import ctypes as C
import os

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd.batch import rays_per_variant
from mikroptim_amd.engine import Prefilter, prefilter_decide
from oracle_lib import oracle, reference

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = np.load(os.path.join(HERE, "golden", "reference_outputs.npz"))


def stats(lib, b, s, eta=0.15, threads=8):
    out = Prefilter(b.variants)
    f = lib.fn("prefilter_stats")
    args = [C.byref(b.c_struct()), C.byref(s.c_struct()), C.c_double(eta), abi.ptr(out.merit),
            abi.ptr(out.optical_length), abi.ptr(out.failed, C.c_int64),
            abi.ptr(out.uncertain_ok, C.c_int64), abi.ptr(out.uncertain_failed, C.c_int64)]
    if lib.prefix == "orc_":
        args.append(threads)
    f.argtypes = [C.c_void_p] * 2 + [C.c_double] + [C.c_void_p] * 5 + (
        [C.c_int] if lib.prefix == "orc_" else [])
    lib.check(f(*args))
    return out


def ref_decide(ref, b, s, st, o, extra=None):
    V = b.variants
    lo, up = np.zeros(V), np.zeros(V)
    rej = np.zeros(V, dtype=np.uint8)
    rm, rv = C.c_double(0), C.c_int(0)
    opts = np.array([o.max_failed_fraction, o.max_merit, o.max_merit_factor, o.boundary,
                     o.opl_error_factor, o.merit_relative])
    ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.float64)
    f = ref.fn("prefilter_decide")
    f.argtypes = [C.c_void_p] * 14
    ref.check(f(C.byref(b.c_struct()), C.byref(s.c_struct()), abi.ptr(opts), abi.ptr(st.merit),
                abi.ptr(st.optical_length), abi.ptr(st.failed, C.c_int64),
                abi.ptr(st.uncertain_ok, C.c_int64), abi.ptr(st.uncertain_failed, C.c_int64),
                abi.ptr(ex), abi.ptr(lo), abi.ptr(up), abi.ptr(rej, C.c_uint8), C.byref(rm),
                C.byref(rv)))
    return lo, up, rej, rm.value, rv.value


def exact_reject(merit, failed, R, o, extra=None):
    """prefilter_decide_exact (src/batch/batch_prefilter.cpp:331-368) on FP64 values."""
    e = np.zeros_like(merit) if extra is None else extra
    F = merit + e
    too_many = failed.astype(float) > o.max_failed_fraction * float(R)
    rej = np.where(too_many, abi.PREFILTER_FAILED_RAYS, 0).astype(np.uint8)
    rej |= np.where(F > o.max_merit, abi.PREFILTER_MERIT_ABSOLUTE, 0).astype(np.uint8)
    cand = (~too_many) & (failed < R)
    if np.isfinite(o.max_merit_factor) and cand.any():
        best = F[cand].min()
        rej |= np.where(F > o.max_merit_factor * best, abi.PREFILTER_MERIT_RELATIVE,
                        0).astype(np.uint8)
    return rej


def criteria_sets(merit, failed, R):
    """The reference acceptance test's kinds of criteria: failed fraction thresholds, merit
    quantiles, multiples of the best merit, combinations."""
    q = np.quantile(merit, [0.25, 0.5, 0.9])
    qf = np.quantile(failed / R, 0.9)
    o = abi.prefilter_options
    return [o(max_failed_fraction=0.01), o(max_failed_fraction=0.1), o(max_failed_fraction=0.5),
            o(max_failed_fraction=float(qf)), o(max_failed_fraction=1.0, max_merit=float(q[0])),
            o(max_failed_fraction=1.0, max_merit=float(q[1])),
            o(max_failed_fraction=1.0, max_merit=float(q[2])),
            o(max_failed_fraction=1.0, max_merit_factor=2.0),
            o(max_failed_fraction=1.0, max_merit_factor=10.0),
            o(max_failed_fraction=1.0, max_merit_factor=100.0),
            o(max_failed_fraction=1.0, max_merit_factor=1e4),
            o(max_failed_fraction=0.1, max_merit=float(q[2]), max_merit_factor=100.0)]


@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_stats_equal_reference(failure):
    b, _ = systems.singlet_multistart(150, seed=11)
    s = systems.singlet_settings(failure)
    st = stats(oracle(), b, s)
    assert st.failed.sum() > 0 and (st.uncertain_ok + st.uncertain_failed).sum() > 0
    key = "prefilter_multistart"
    for name in ("merit", "optical_length", "failed", "uncertain_ok", "uncertain_failed"):
        assert np.array_equal(getattr(st, name), FIXTURES[f"{key}_{name}"]), name
    ref = reference()
    if ref is not None:
        rs = stats(ref, b, s)
        for name in ("merit", "optical_length", "failed", "uncertain_ok", "uncertain_failed"):
            assert np.array_equal(getattr(st, name), getattr(rs, name)), name


def test_stats_equal_reference_objective_and_image_sphere():
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    b, _ = systems.objective(variants=3, seed=2)
    s = systems.objective_settings()
    s.rings, s.spokes = 4, 6
    for x, y in zip(vars(stats(oracle(), b, s)).values(), vars(stats(ref, b, s)).values()):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    from test_oracle import _image_reference_singlet
    for kind in (abi.REF_IMAGE_EXIT_PUPIL, abi.REF_IMAGE_INFINITE):
        bi, _ = _image_reference_singlet(kind)
        bi.pupil_radius[:] = [5.0, 30.0, 48.0]
        si = systems.singlet_settings(abi.FAILURE_HARD)
        for x, y in zip(vars(stats(oracle(), bi, si)).values(), vars(stats(ref, bi, si)).values()):
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_bounds_and_assumption_a():
    """certain-failed <= FP64 failed <= certain-failed + uncertain, and L <= F_FP64 <= U."""
    b, _ = systems.singlet_multistart(600, seed=11)
    s = systems.singlet_settings(abi.FAILURE_HARD)
    st = stats(oracle(), b, s)
    m, f, _ = oracle().merit(b, s, threads=8)
    unc = st.uncertain_ok + st.uncertain_failed
    assert np.all(st.failed <= f) and np.all(f <= st.failed + unc)
    r = prefilter_decide(b, s, st)
    assert np.all(r.merit_lower <= m)
    fin = np.isfinite(r.merit_upper)
    assert fin.sum() > 0.8 * b.variants and np.all(m[fin] <= r.merit_upper[fin])


@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_decision_equals_reference(failure):
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    b, _ = systems.singlet_multistart(300, seed=13)
    s = systems.singlet_settings(failure)
    st = stats(oracle(), b, s)
    R = rays_per_variant(b, s)
    extra = np.linspace(0.0, 3.0, b.variants)
    for o in criteria_sets(st.merit, st.failed, R):
        for ex in (None, extra):
            r = prefilter_decide(b, s, st, o, ex)
            lo, up, rej, rm, rv = ref_decide(ref, b, s, st, o, ex)
            assert np.array_equal(r.merit_lower, lo) and np.array_equal(r.merit_upper, up)
            assert np.array_equal(r.reject, rej)
            assert r.reference_merit == rm and r.reference_variant == rv
            assert r.kept == int((rej == 0).sum())


@pytest.mark.parametrize("failure", [abi.FAILURE_HARD, abi.FAILURE_SOFT])
def test_no_false_rejections(failure):
    """No variant is rejected for a reason that its FP64 values do not have; most FP64
    rejections are found."""
    b, _ = systems.singlet_multistart(800, seed=17)
    s = systems.singlet_settings(failure)
    st = stats(oracle(), b, s)
    m, f, _ = oracle().merit(b, s, threads=8)
    R = rays_per_variant(b, s)
    found = truth_total = 0
    for o in criteria_sets(m, f, R):
        r = prefilter_decide(b, s, st, o)
        truth = exact_reject(m, f, R, o)
        assert np.all((r.reject & ~truth) == 0)
        found += int(np.count_nonzero(r.reject))
        truth_total += int(np.count_nonzero(truth))
    assert truth_total > 0 and found >= 0.5 * truth_total


def test_invalid_options_are_rejected():
    b, _ = systems.singlet(2)
    s = systems.singlet_settings()
    st = stats(oracle(), b, s)
    from mikroptim_amd.engine import HipError
    for o in (abi.prefilter_options(max_failed_fraction=-1.0), abi.prefilter_options(boundary=-0.1),
              abi.prefilter_options(merit_relative=1.0),
              abi.prefilter_options(opl_error_factor=float("inf"))):
        with pytest.raises(HipError):
            prefilter_decide(b, s, st, o)

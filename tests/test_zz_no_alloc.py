"""mko_merit_device_noalloc (MERIT with the tuple partials in the library's static device array)
against mko_merit_device (caller's workspace) through the C-ABI, on a device-resident batch: two
calls in a row, the second with a smaller batch, each bit for bit equal to the workspace entry
point; and the size limit rejected before any device work."""

import ctypes as C

import numpy as np
import pytest

from mikroptim_amd import abi, systems
from mikroptim_amd.engine import library

pytestmark = pytest.mark.gpu

MAX_TUPLES = 1 << 19  # the static array's capacity in (variant, field, wavelength) tuples


def _device_batch(b, torch):
    """mko_batch whose per-variant arrays are device tensors (kept alive in the returned list)."""
    keep, ptrs = [], {}
    for name in b.PER_VARIANT:
        a = getattr(b, name)
        if a is None:
            continue
        t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")
        keep.append(t)
        ptrs[name] = t.data_ptr()
    return b.c_struct(ptrs), keep


def _dp(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))


def _ip64(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_int64))


def _merit_both(b, s, torch):
    lib = library()
    cb, keep = _device_batch(b, torch)
    cs = s.c_struct()
    V, T = b.variants, b.fields * b.wavelengths
    out = {}
    for name in ("workspace", "noalloc"):
        merit = torch.full((V,), float("nan"), dtype=torch.float64, device="cuda")
        failed = torch.full((V,), -1, dtype=torch.int64, device="cuda")
        hard = torch.full((V,), -1, dtype=torch.int64, device="cuda")
        extra = torch.from_numpy(np.linspace(0.25, 1.0, V)).to("cuda")
        if name == "workspace":
            partial = torch.empty(V * T, dtype=torch.float64, device="cuda")
            counts = torch.empty(2 * V * T, dtype=torch.int32, device="cuda")
            rc = lib.mko_merit_device(C.byref(cb), C.byref(cs), _dp(extra), _dp(merit),
                                      _ip64(failed), _ip64(hard), _dp(partial),
                                      C.cast(C.c_void_p(counts.data_ptr()),
                                             C.POINTER(C.c_int)), None)
        else:
            rc = lib.mko_merit_device_noalloc(C.byref(cb), C.byref(cs), _dp(extra), _dp(merit),
                                              _ip64(failed), _ip64(hard), None)
        assert rc == 0, lib.mko_last_error().decode()
        torch.cuda.synchronize()
        out[name] = (merit.cpu().numpy(), failed.cpu().numpy(), hard.cpu().numpy())
    del keep
    return out


def test_noalloc_merit_equals_workspace_merit_on_two_calls(engine):
    torch = pytest.importorskip("torch")
    # first call: 512 objective variants (17 920 tuples), then 16 singlet variants with failures
    b1, _ = systems.objective(variants=512, seed=2)
    s1 = systems.objective_settings(abi.FAILURE_HARD)
    b2, _ = systems.singlet(16)
    s2 = systems.singlet_settings(abi.FAILURE_SOFT)
    for b, s in ((b1, s1), (b2, s2)):
        r = _merit_both(b, s, torch)
        for x, y in zip(r["workspace"], r["noalloc"]):
            assert np.array_equal(x, y)
        assert np.isfinite(r["noalloc"][0]).all() and (r["noalloc"][1] >= 0).all()
        # and the host layer's merit of the same batch (its own buffers) plus the extra term
        m, f, h = engine.merit(b, s)
        extra = np.linspace(0.25, 1.0, b.variants)
        assert np.array_equal(r["noalloc"][0], m + extra)
        assert np.array_equal(r["noalloc"][1], f) and np.array_equal(r["noalloc"][2], h)


def test_noalloc_merit_rejects_batches_beyond_the_scratch():
    torch = pytest.importorskip("torch")  # noqa: F841
    b, _ = systems.singlet(1)
    s = systems.singlet_settings(abi.FAILURE_HARD)
    cb = b.c_struct()
    cb.variants = MAX_TUPLES // (b.fields * b.wavelengths) + 1
    cs = s.c_struct()
    rc = library().mko_merit_device_noalloc(C.byref(cb), C.byref(cs), None, None, None, None,
                                            None)
    assert rc == -1  # MKO_ERR_INVALID: nothing launched
    assert "at most 524288" in library().mko_last_error().decode()

"""Python face of libmikroptim_hip.so, shaped like the reference's Python module
(python/src/bind_trace.cpp:411-623 TraceEngine / SystemBatch, python/src/bind_optimize.cpp:476
optimize): TraceEngine.residuals / merit (TraceMode RESIDUALS / MERIT), jacobian (FD or adjoint
Jacobian with normal equations) and optimize (batched Levenberg-Marquardt on the GPU backend).

The HIP library is required: importing works without it (so CPU-only tooling can load the
package), but every call raises if libmikroptim_hip.so is missing or has no device."""

import ctypes as C
import os

import numpy as np

from . import abi
from .batch import rays_per_variant, residual_rows

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmikroptim_hip.so")

_dp = C.POINTER(C.c_double)
_lib = None

# every symbol include/mikroptim_hip.h declares
EXPORTS = [
    "mko_last_error", "mko_residuals_device", "mko_merit_device", "mko_merit_device_noalloc",
    "mko_point_batch_device",
    "mko_fd_difference_device", "mko_adjoint_jacobian_device", "mko_adjoint_normal_workspace",
    "mko_adjoint_normal_device", "mko_vjp_workspace", "mko_vjp_device", "mko_gram_device",
    "mko_gram_mfma_device",
    "mko_df64_ops_device",
    "mko_engine_create", "mko_engine_destroy", "mko_engine_set_tuning", "mko_engine_get_tuning",
    "mko_engine_autotune", "mko_engine_row_trace_stats",
    "mko_engine_residuals", "mko_engine_merit",
    "mko_engine_surface_extrema", "mko_engine_prefilter", "mko_prefilter_decide",
    "mko_engine_jacobian", "mko_engine_kernel_timing", "mko_engine_normal_equations",
    "mko_engine_point_merit", "mko_optimize_lm", "mko_lm_create", "mko_lm_cycle",
    "mko_lm_results", "mko_lm_destroy", "mko_lm_graph_cycles", "mko_lm_cycle_begin",
    "mko_lm_cycle_end",
]


class HipError(RuntimeError):
    pass


def library():
    """The loaded libmikroptim_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipError(f"{LIB_PATH} is missing: build it first (python -c "
                       "'import __graft_entry__; __graft_entry__.build()')")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 (same SONAME); loading torch
    # first makes this library bind to that runtime, so torch tensors and this library share the
    # device context (the torch.autograd.Function path depends on it).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    bp, sp, vp = C.POINTER(abi.Batch), C.POINTER(abi.ResidualSettings), C.POINTER(abi.Variables)
    sig = {
        "mko_last_error": ([], C.c_char_p),
        "mko_residuals_device": ([bp, sp, _dp, C.c_int64, _dp, C.POINTER(C.c_uint8), C.c_void_p],
                                 C.c_int),
        "mko_merit_device": ([bp, sp, _dp, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _dp,
                              C.POINTER(C.c_int), C.c_void_p], C.c_int),
        "mko_merit_device_noalloc": ([bp, sp, _dp, _dp, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64), C.c_void_p], C.c_int),
        "mko_gram_device": ([C.c_void_p, _dp, C.c_int, C.c_int, C.c_int64, _dp, _dp,
                             C.c_void_p], C.c_int),
        "mko_adjoint_jacobian_device": ([bp, sp, vp, C.c_int, C.c_void_p, _dp,
                                         C.POINTER(C.c_uint8), C.c_void_p], C.c_int),
        "mko_adjoint_normal_workspace": ([C.c_int, C.c_int, C.c_int, C.c_int], C.c_size_t),
        "mko_adjoint_normal_device": ([bp, sp, vp, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_size_t, C.c_void_p], C.c_int),
        "mko_gram_mfma_device": ([C.c_void_p, _dp, C.c_int, C.c_int, C.c_int64, _dp, _dp,
                                  C.c_void_p], C.c_int),
        "mko_vjp_workspace": ([C.c_int, C.c_int, C.c_int, C.c_int], C.c_size_t),
        "mko_vjp_device": ([bp, sp, vp, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_size_t,
                            C.c_void_p], C.c_int),
        "mko_point_batch_device": ([bp, C.c_int, C.c_int, C.c_int, vp, C.c_void_p, C.c_void_p,
                                    C.c_void_p] + [C.c_void_p] * 8 + [C.c_void_p], C.c_int),
        "mko_df64_ops_device": ([C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p], C.c_int),
        "mko_engine_create": ([C.c_int, C.c_size_t], C.c_void_p),
        "mko_engine_destroy": ([C.c_void_p], None),
        "mko_engine_set_tuning": ([C.c_void_p, C.POINTER(abi.Tuning)], C.c_int),
        "mko_engine_get_tuning": ([C.c_void_p, C.POINTER(abi.Tuning)], C.c_int),
        "mko_engine_autotune": ([C.c_void_p, bp, sp, vp, C.c_int, _dp, C.POINTER(abi.Tuning)],
                                C.c_int),
        "mko_engine_row_trace_stats": ([C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
                                       C.c_int),
        "mko_engine_residuals": ([C.c_void_p, bp, sp, _dp, _dp, C.POINTER(C.c_uint8)], C.c_int),
        "mko_engine_merit": ([C.c_void_p, bp, sp, _dp, _dp, C.POINTER(C.c_int64),
                              C.POINTER(C.c_int64)], C.c_int),
        "mko_engine_surface_extrema": ([C.c_void_p, bp, sp, _dp], C.c_int),
        "mko_engine_prefilter": ([C.c_void_p, bp, sp, C.POINTER(abi.PrefilterOptions), _dp,
                                  C.POINTER(abi.PrefilterResult)], C.c_int),
        "mko_prefilter_decide": ([bp, sp, C.POINTER(abi.PrefilterOptions), _dp,
                                  C.POINTER(abi.PrefilterResult)], C.c_int),
        "mko_engine_jacobian": ([C.c_void_p, bp, sp, vp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp,
                                 _dp, _dp], C.c_int),
        "mko_engine_kernel_timing": ([C.c_void_p, C.c_int, _dp, C.POINTER(C.c_int64)], C.c_int),
        "mko_engine_normal_equations": ([C.c_void_p, bp, sp, vp, C.c_int, C.POINTER(C.c_int32),
                                         _dp, C.c_int, C.c_int, _dp, _dp], C.c_int),
        "mko_engine_point_merit": ([C.c_void_p, bp, sp, vp, C.c_int, C.POINTER(C.c_int32), _dp,
                                    _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
        "mko_optimize_lm": ([C.c_void_p, bp, sp, vp, C.c_int, _dp, C.POINTER(abi.LmOptions),
                             C.c_int, _dp, C.POINTER(abi.LmRun), C.POINTER(C.c_int)], C.c_int),
        "mko_lm_create": ([C.c_void_p, bp, sp, vp, C.c_int, _dp, C.POINTER(abi.LmOptions),
                           C.c_int], C.c_void_p),
        "mko_lm_cycle": ([C.c_void_p], C.c_int),
        "mko_lm_results": ([C.c_void_p, _dp, C.POINTER(abi.LmRun), C.POINTER(C.c_int)], C.c_int),
        "mko_lm_destroy": ([C.c_void_p], None),
        "mko_lm_graph_cycles": ([C.c_void_p], C.c_int),
        "mko_lm_cycle_begin": ([C.c_void_p], C.c_int),
        "mko_lm_cycle_end": ([C.c_void_p], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise HipError(library().mko_last_error().decode())


_PRECISION = {"fp64": abi.PRECISION_FP64, "mixed": abi.PRECISION_MIXED,
              "fp64_mfma": abi.PRECISION_FP64_MFMA}
_METHOD = {"fd": abi.JACOBIAN_FD, "adjoint": abi.JACOBIAN_ADJOINT}


class Prefilter:
    """PrefilterResult: per-variant stats of the float trace (merit, optical_length, failed,
    uncertain_ok, uncertain_failed), FP64 merit bounds, reject bits (0 = keep), the reference
    merit / variant of criterion (c) and the number kept."""

    def __init__(self, V):
        self.merit = np.zeros(V)
        self.optical_length = np.zeros(V)
        self.failed = np.zeros(V, dtype=np.int64)
        self.uncertain_ok = np.zeros(V, dtype=np.int64)
        self.uncertain_failed = np.zeros(V, dtype=np.int64)
        self.merit_lower = np.zeros(V)
        self.merit_upper = np.zeros(V)
        self.reject = np.zeros(V, dtype=np.uint8)
        self.reference_merit = float("inf")
        self.reference_variant = -1
        self.kept = 0

    def c_struct(self):
        return abi.PrefilterResult(
            abi.ptr(self.merit), abi.ptr(self.optical_length),
            abi.ptr(self.failed, C.c_int64), abi.ptr(self.uncertain_ok, C.c_int64),
            abi.ptr(self.uncertain_failed, C.c_int64), abi.ptr(self.merit_lower),
            abi.ptr(self.merit_upper), abi.ptr(self.reject, C.c_uint8), float("inf"), -1, 0)

    def take(self, r):
        self.reference_merit = r.reference_merit
        self.reference_variant = r.reference_variant
        self.kept = r.kept
        return self


def prefilter_decide(batch, settings, stats, options=None, extra=None):
    """prefilter_decide on given stats (a Prefilter whose stats are filled): host only."""
    options = options or abi.prefilter_options()
    ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.float64)
    r = stats.c_struct()
    b, s = batch.c_struct(), settings.c_struct()
    _check(library().mko_prefilter_decide(C.byref(b), C.byref(s), C.byref(options), abi.ptr(ex),
                                          C.byref(r)))
    return stats.take(r)


class LmResult:
    """Per-run results (LmRunResult) and the cycle count of the whole call."""

    def __init__(self, x, runs, cycles):
        self.x = x
        self.runs = runs
        self.cycles = cycles
        self.merit = np.array([r.merit for r in runs])
        self.status = [abi.LM_STATUS_NAMES[r.status] for r in runs]
        self.iterations = np.array([r.iterations for r in runs])


class LmSession:
    """One batched LM, advanced one cycle per call of cycle()."""

    def __init__(self, engine, base, settings, variables, x0, options, precision):
        self._keep = (base.c_struct(), settings.c_struct(), variables.c_struct(),
                      np.ascontiguousarray(np.atleast_2d(x0), dtype=np.float64), options)
        b, s, v, x, o = self._keep
        self.runs, self.P = x.shape
        self._h = library().mko_lm_create(engine._h, C.byref(b), C.byref(s), C.byref(v), self.runs,
                                          abi.ptr(x), C.byref(o), _PRECISION[precision])
        if not self._h:
            raise HipError(library().mko_last_error().decode())

    def cycle(self):
        rc = library().mko_lm_cycle(self._h)
        if rc < 0:
            _check(rc)
        return rc == 1

    def cycle_begin(self):
        """Launches the next cycle without waiting (1), or runs it eagerly (2); 0 when done."""
        rc = library().mko_lm_cycle_begin(self._h)
        if rc < 0:
            _check(rc)
        return rc

    def cycle_end(self):
        rc = library().mko_lm_cycle_end(self._h)
        if rc < 0:
            _check(rc)
        return rc == 1

    def graph_cycles(self):
        """Cycles that ran as a replay of the captured hipGraph."""
        return library().mko_lm_graph_cycles(self._h)

    def results(self):
        x = np.zeros((self.runs, self.P))
        res = (abi.LmRun * self.runs)()
        cyc = C.c_int(0)
        _check(library().mko_lm_results(self._h, abi.ptr(x), res, C.byref(cyc)))
        return LmResult(x, list(res), cyc.value)

    def close(self):
        if getattr(self, "_h", None):
            library().mko_lm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class TraceEngine:
    """TraceEngine(device, memory_limit_bytes) of include/mikroptim/cuda/trace_engine.hpp."""

    def __init__(self, device=0, memory_limit_bytes=0):
        self._h = library().mko_engine_create(device, memory_limit_bytes)
        if not self._h:
            raise HipError(library().mko_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            library().mko_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    KERNEL_FAMILIES = ("trace", "adjoint", "gram", "fd_difference", "point_batch", "lm_control")

    def tuning(self, pack_rays=None, fused_waves=None, full_row_traces=None):
        """Launch shapes of this engine's kernels (mko_tuning): returns the current (pack_rays,
        fused_waves) and sets those given (0 = automatic; pack_rays 32 / 64 / 128, 256 = no
        tuple packs; fused_waves 1 / 2 / 4). full_row_traces (True: the constraint rows'
        differences trace every ray instead of the checked candidate rays) is set when given
        and not part of the returned pair."""
        t = abi.Tuning()
        _check(library().mko_engine_get_tuning(self._h, C.byref(t)))
        old = (t.pack_rays, t.fused_waves)
        if pack_rays is not None:
            t.pack_rays = int(pack_rays)
        if fused_waves is not None:
            t.fused_waves = int(fused_waves)
        if full_row_traces is not None:
            t.full_row_traces = int(bool(full_row_traces))
        if pack_rays is not None or fused_waves is not None or full_row_traces is not None:
            _check(library().mko_engine_set_tuning(self._h, C.byref(t)))
        return old

    def row_trace_stats(self):
        """(chunks whose constraint-row differences traced only candidate rays, chunks traced
        whole again after the check)."""
        a, b = C.c_int64(0), C.c_int64(0)
        _check(library().mko_engine_row_trace_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, base, settings, variables, x):
        """Times the candidate launch shapes on the first min(n, 64) points x [n][P] (base: one
        variant, or one per point), sets the fastest as this engine's tuning (cached per shape)
        and returns it as (pack_rays, fused_waves). LM sessions keep the tuning of the engine at
        their creation."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        t = abi.Tuning()
        b, s, v = base.c_struct(), settings.c_struct(), variables.c_struct()
        _check(library().mko_engine_autotune(self._h, C.byref(b), C.byref(s), C.byref(v),
                                             x.shape[0], abi.ptr(x), C.byref(t)))
        return t.pack_rays, t.fused_waves

    def kernel_timing(self, enable=True):
        """Device ms and launch counts per kernel family since the last call (HIP events on the
        launch stream); then resets and enables/disables timing."""
        ms = np.zeros(len(self.KERNEL_FAMILIES))
        n = np.zeros(len(self.KERNEL_FAMILIES), dtype=np.int64)
        _check(library().mko_engine_kernel_timing(self._h, int(enable), abi.ptr(ms),
                                                  abi.ptr(n, C.c_int64)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.KERNEL_FAMILIES)}

    def residuals(self, batch, settings):
        """(residuals [V][R], OPD [V][rays], status [V][rays]); R = trace rows + the
        settings' constraint rows."""
        V = batch.variants
        R = residual_rows(batch, settings)
        rays = rays_per_variant(batch, settings)
        res = np.zeros((V, R))
        opd = np.zeros((V, rays))
        st = np.zeros((V, rays), dtype=np.uint8)
        b, s = batch.c_struct(), settings.c_struct()
        _check(library().mko_engine_residuals(self._h, C.byref(b), C.byref(s), abi.ptr(res),
                                              abi.ptr(opd), abi.ptr(st, C.c_uint8)))
        return res, opd, st

    def surface_extrema(self, batch, settings):
        """[V][S][2]: largest ray height and smallest cos I per surface over the rays with
        status Ok (SurfaceExtremaResult)."""
        out = np.zeros((batch.variants, batch.surfaces, 2))
        b, s = batch.c_struct(), settings.c_struct()
        _check(library().mko_engine_surface_extrema(self._h, C.byref(b), C.byref(s),
                                                    abi.ptr(out)))
        return out

    def prefilter(self, batch, settings, options=None, extra=None):
        """TraceEngine::prefilter_variants: the FP32 prefilter of every variant (float
        trace with boundary margins) and its conservative decision."""
        options = options or abi.prefilter_options()
        out = Prefilter(batch.variants)
        ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.float64)
        r = out.c_struct()
        b, s = batch.c_struct(), settings.c_struct()
        _check(library().mko_engine_prefilter(self._h, C.byref(b), C.byref(s), C.byref(options),
                                              abi.ptr(ex), C.byref(r)))
        return out.take(r)

    def merit(self, batch, settings, extra=None):
        V = batch.variants
        m = np.zeros(V)
        f = np.zeros(V, dtype=np.int64)
        h = np.zeros(V, dtype=np.int64)
        ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.float64)
        b, s = batch.c_struct(), settings.c_struct()
        _check(library().mko_engine_merit(self._h, C.byref(b), C.byref(s), abi.ptr(ex),
                                          abi.ptr(m), abi.ptr(f, C.c_int64),
                                          abi.ptr(h, C.c_int64)))
        return m, f, h

    def jacobian(self, base, settings, variables, x0, method="adjoint", precision="fp64",
                 want_jacobian=True):
        """(J [runs][P][R] or None, r0 [runs][R] or None, A [runs][P][P], g [runs][P]).
        want_jacobian=False asks for the normal equations only (the LM's use): J and r0 are
        then not returned, and the mixed-precision adjoint forms A and g without storing J."""
        x0 = np.ascontiguousarray(np.atleast_2d(x0), dtype=np.float64)
        runs, P = x0.shape
        R = residual_rows(base, settings)
        J = np.zeros((runs, P, R)) if want_jacobian else None
        r0 = np.zeros((runs, R)) if want_jacobian else None
        a = np.zeros((runs, P, P))
        g = np.zeros((runs, P))
        b, s, v = base.c_struct(), settings.c_struct(), variables.c_struct()
        _check(library().mko_engine_jacobian(self._h, C.byref(b), C.byref(s), C.byref(v), runs,
                                             abi.ptr(x0), _METHOD[method], _PRECISION[precision],
                                             abi.ptr(J), abi.ptr(r0), abi.ptr(a), abi.ptr(g)))
        return J, r0, a, g

    def normal_equations(self, base, settings, variables, x, runs=None, method="adjoint",
                         precision="fp64"):
        """LmBackend::normal_equations: (A [n][P][P], g [n][P]) at the points x [n][P] of the
        runs `runs` (indices into a per-run base; None = 0..n-1)."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        n, P = x.shape
        r = None if runs is None else np.ascontiguousarray(runs, dtype=np.int32)
        a = np.zeros((n, P, P))
        g = np.zeros((n, P))
        b, s, v = base.c_struct(), settings.c_struct(), variables.c_struct()
        _check(library().mko_engine_normal_equations(
            self._h, C.byref(b), C.byref(s), C.byref(v), n, abi.ptr(r, C.c_int32), abi.ptr(x),
            _METHOD[method], _PRECISION[precision], abi.ptr(a), abi.ptr(g)))
        return a, g

    def point_merit(self, base, settings, variables, x, runs=None):
        """LmBackend::merit: (merit, failed, hard_failed) [n] at the points x [n][P]."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        n, P = x.shape
        r = None if runs is None else np.ascontiguousarray(runs, dtype=np.int32)
        m = np.zeros(n)
        f = np.zeros(n, dtype=np.int64)
        h = np.zeros(n, dtype=np.int64)
        b, s, v = base.c_struct(), settings.c_struct(), variables.c_struct()
        _check(library().mko_engine_point_merit(
            self._h, C.byref(b), C.byref(s), C.byref(v), n, abi.ptr(r, C.c_int32), abi.ptr(x),
            abi.ptr(m), abi.ptr(f, C.c_int64), abi.ptr(h, C.c_int64)))
        return m, f, h

    def optimize(self, base, settings, variables, x0, options=None, precision="fp64"):
        options = options or abi.lm_options()
        x0 = np.ascontiguousarray(np.atleast_2d(x0), dtype=np.float64)
        runs, P = x0.shape
        x = np.zeros((runs, P))
        res = (abi.LmRun * runs)()
        cyc = C.c_int(0)
        b, s, v = base.c_struct(), settings.c_struct(), variables.c_struct()
        _check(library().mko_optimize_lm(self._h, C.byref(b), C.byref(s), C.byref(v), runs,
                                         abi.ptr(x0), C.byref(options), _PRECISION[precision],
                                         abi.ptr(x), res, C.byref(cyc)))
        return LmResult(x, list(res), cyc.value)

    def optimize_groups(self, groups, options=None, precision="fp64", autotune=False):
        """Uneven batches: runs of different topologies (surface, field, wavelength or variable
        counts), one batched LM per topology group; groups = [(base, settings, variables, x0)].
        Returns the LmResult of every group, in order. The groups advance cycle by cycle
        together; groups whose cycle is a captured hipGraph (device control, fused mixed adjoint)
        run concurrently on their own streams. Each group's result is that of optimize() on it,
        bit for bit. With autotune, each group's launch shapes are first chosen by autotune() on
        its own points (the fused kernel's rounding may then differ from the untuned run).
        (The reference's batches hold a single topology,
        include/mikroptim/batch/system_batch.hpp:26.)"""
        sessions = []
        previous = self.tuning()
        try:
            for b, s, v, x0 in groups:
                if autotune:
                    self.autotune(b, s, v, x0)
                sessions.append(self.lm(b, s, v, x0, options, precision))
        except BaseException:
            for sess in sessions:
                sess.close()
            raise
        finally:
            if autotune:
                self.tuning(*previous)
        try:
            # every group's cycle is begun before any is waited for: the groups whose cycles are
            # captured graphs run concurrently, each on its own stream
            active = list(sessions)
            while active:
                begun = [(sess, sess.cycle_begin()) for sess in active]
                active = [sess for sess, rc in begun if rc != 0 and sess.cycle_end()]
            return [sess.results() for sess in sessions]
        finally:
            for sess in sessions:
                sess.close()

    def lm(self, base, settings, variables, x0, options=None, precision="fp64"):
        return LmSession(self, base, settings, variables, x0, options or abi.lm_options(),
                         precision)

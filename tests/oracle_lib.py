"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the reference driver
(oracle/_ref/libmkoref.so). Test infrastructure: the checker side of every parity test."""

import ctypes as C
import os
import subprocess

import numpy as np

from mikroptim_amd import abi
from mikroptim_amd.batch import rays_per_variant, residual_rows

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libmkoref.so")

_dp = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "-j8"], check=True)


def _declare(lib, prefix):
    p = prefix
    bp, sp, vp = C.POINTER(abi.Batch), C.POINTER(abi.ResidualSettings), C.POINTER(abi.Variables)
    sig = {
        "last_error": ([], C.c_char_p),
        "gauss_legendre": ([C.c_int, _dp, _dp], C.c_int),
        "pupil_samples": ([C.c_int, C.c_int, C.c_int, _dp, _dp, _dp], C.c_int),
        "normal_equations": ([_dp, _dp, C.c_int, C.c_int, C.c_int64, _dp, _dp], C.c_int),
        "splitmix64": ([C.c_uint64], C.c_uint64),
        "variant_seed": ([C.c_uint64, C.c_uint64], C.c_uint64),
        "unit_uniform": ([C.c_uint64], C.c_double),
        "df64_ops": ([C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
        "paraxial": ([C.POINTER(abi.Batch), C.c_int, _dp], C.c_int),
    }
    if prefix == "orc_":
        sig.update({
            "residuals": ([bp, sp, _dp, _dp, C.POINTER(C.c_uint8), C.c_int], C.c_int),
            "merit": ([bp, sp, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int], C.c_int),
            "fd_jacobian": ([bp, sp, vp, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int64),
                             C.POINTER(C.c_uint8), C.c_int], C.c_int),
            "adjoint_jacobian": ([bp, sp, vp, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_uint8),
                                  C.c_int], C.c_int),
            "fd_pattern": ([vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_uint8)],
                           C.c_int),
            "perturbed_vector": ([vp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_uint64, C.c_int,
                                  C.c_uint64, _dp], C.c_int),
            "optimize_lm": ([bp, sp, vp, C.c_int, _dp, C.POINTER(abi.LmOptions), _dp,
                             C.POINTER(abi.LmRun), C.POINTER(C.c_int)], C.c_int),
            "lm_linear": ([_dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _dp,
                           C.POINTER(abi.LmOptions), _dp, C.POINTER(abi.LmRun)], C.c_int),
            "trace_ray": ([bp, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_int,
                           _dp, _dp, _dp, _dp], C.c_int),
            "mt19937_64": ([C.c_uint64, C.c_uint64], C.c_uint64),
            "surface_extrema": ([bp, sp, _dp, C.c_int], C.c_int),
            "constraint_rows": ([bp, sp, _dp, _dp, C.c_int], C.c_int),
        })
    else:
        sig.update({
            "residuals": ([bp, sp, _dp, _dp, C.POINTER(C.c_uint8)], C.c_int),
            "merit": ([bp, sp, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
            "fd_jacobian": ([bp, sp, vp, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int64)], C.c_int),
            "adjoint_jacobian": ([bp, sp, vp, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_uint8),
                                  C.c_int], C.c_int),
            "fd_pattern": ([bp, vp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_uint8)], C.c_int),
            "perturbed_vector": ([bp, vp, _dp, _dp, _dp, C.c_int, C.c_uint64, C.c_int,
                                  C.c_uint64, _dp], C.c_int),
            "optimize_lm": ([bp, sp, vp, C.c_int, _dp, C.POINTER(abi.LmOptions), _dp,
                             C.POINTER(abi.LmRun), C.POINTER(C.c_int)], C.c_int),
            "surface_extrema": ([bp, sp, _dp], C.c_int),
            "constraint_rows": ([bp, sp, _dp, _dp], C.c_int),
        })
    for name, (args, res) in sig.items():
        f = getattr(lib, p + name)
        f.argtypes = args
        f.restype = res
    return lib


class Lib:
    """Uniform Python API over liboracle.so ("orc_") or libmkoref.so ("ref_")."""

    def __init__(self, path, prefix):
        self.lib = _declare(C.CDLL(path), prefix)
        self.prefix = prefix

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def check(self, rc):
        if rc != 0:
            raise RuntimeError(self.fn("last_error")().decode())

    def residuals(self, batch, settings, threads=0):
        V = batch.variants
        R = residual_rows(batch, settings)
        rays = rays_per_variant(batch, settings)
        res = np.zeros((V, R))
        opd = np.zeros((V, rays))
        st = np.zeros((V, rays), dtype=np.uint8)
        args = [C.byref(batch.c_struct()), C.byref(settings.c_struct()), abi.ptr(res),
                abi.ptr(opd), abi.ptr(st, C.c_uint8)]
        if self.prefix == "orc_":
            args.append(threads)
        self.check(self.fn("residuals")(*args))
        return res, opd, st

    def merit(self, batch, settings, threads=0):
        V = batch.variants
        m = np.zeros(V)
        f = np.zeros(V, dtype=np.int64)
        h = np.zeros(V, dtype=np.int64)
        args = [C.byref(batch.c_struct()), C.byref(settings.c_struct()), abi.ptr(m),
                abi.ptr(f, C.c_int64), abi.ptr(h, C.c_int64)]
        if self.prefix == "orc_":
            args.append(threads)
        self.check(self.fn("merit")(*args))
        return m, f, h

    def fd_jacobian(self, base, settings, variables, x0, threads=0):
        x0 = np.ascontiguousarray(np.atleast_2d(x0), dtype=np.float64)
        runs, P = x0.shape
        R = residual_rows(base, settings)
        J = np.zeros((runs, P, R))
        r0 = np.zeros((runs, R))
        ch = np.zeros((runs, P), dtype=np.int64)
        args = [C.byref(base.c_struct()), C.byref(settings.c_struct()),
                C.byref(variables.c_struct()), runs, abi.ptr(x0), abi.ptr(J), abi.ptr(r0),
                abi.ptr(ch, C.c_int64)]
        if self.prefix == "orc_":
            args += [None, threads]
        self.check(self.fn("fd_jacobian")(*args))
        return J, r0, ch

    def adjoint_jacobian(self, base, settings, variables, x0, threads=0):
        x0 = np.ascontiguousarray(np.atleast_2d(x0), dtype=np.float64)
        runs, P = x0.shape
        R = residual_rows(base, settings)
        rays = rays_per_variant(base, settings)
        J = np.zeros((runs, P, R))
        r0 = np.zeros((runs, R))
        st = np.zeros((runs, rays), dtype=np.uint8)
        self.check(self.fn("adjoint_jacobian")(
            C.byref(base.c_struct()), C.byref(settings.c_struct()),
            C.byref(variables.c_struct()), runs, abi.ptr(x0), abi.ptr(J), abi.ptr(r0),
            abi.ptr(st, C.c_uint8), threads))
        return J, r0, st

    def surface_extrema(self, batch, settings, threads=0):
        """[V][S][2]: largest ray height and smallest cos I per surface over the valid rays."""
        out = np.zeros((batch.variants, batch.surfaces, 2))
        args = [C.byref(batch.c_struct()), C.byref(settings.c_struct()), abi.ptr(out)]
        if self.prefix == "orc_":
            args.append(threads)
        self.check(self.fn("surface_extrema")(*args))
        return out

    def paraxial(self, batch, v):
        """(A, B, C, D, power, petzval_curvature) of variant v at lambda_ref."""
        out = np.zeros(6)
        self.check(self.fn("paraxial")(C.byref(batch.c_struct()), v, abi.ptr(out)))
        return out

    def constraint_rows(self, batch, settings, extrema=None, threads=0):
        """[V][E + E_t] constraint rows (row order: without extrema, then with); extrema
        [V][S][2] as given (the oracle computes them when None)."""
        if extrema is None and self.prefix == "ref_":
            extrema = self.surface_extrema(batch, settings)
        ext = None if extrema is None else np.ascontiguousarray(extrema, dtype=np.float64)
        out = np.zeros((batch.variants, settings.extra_rows))
        args = [C.byref(batch.c_struct()), C.byref(settings.c_struct()), abi.ptr(ext),
                abi.ptr(out)]
        if self.prefix == "orc_":
            args.append(threads)
        self.check(self.fn("constraint_rows")(*args))
        return out

    def normal_equations(self, J, r):
        J = np.ascontiguousarray(J, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        runs, P, R = J.shape
        a = np.zeros((runs, P, P))
        g = np.zeros((runs, P))
        self.check(self.fn("normal_equations")(abi.ptr(J), abi.ptr(r), runs, P, R, abi.ptr(a),
                                               abi.ptr(g)))
        return a, g

    def fd_pattern(self, batch, variables, x0):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        P = len(x0)
        out = [np.zeros(P) for _ in range(5)]
        st = np.zeros(P, dtype=np.uint8)
        first = (C.byref(variables.c_struct()), batch.surfaces) if self.prefix == "orc_" else (
            C.byref(batch.c_struct()), C.byref(variables.c_struct()))
        self.check(self.fn("fd_pattern")(*first, abi.ptr(x0), *[abi.ptr(o) for o in out],
                                         abi.ptr(st, C.c_uint8)))
        return out + [st]

    def perturbed_vector(self, batch, variables, x0, relative, absolute, distribution, seed,
                         first_is_base, j):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        rel = np.ascontiguousarray(relative, dtype=np.float64)
        ab = np.ascontiguousarray(absolute, dtype=np.float64)
        y = np.zeros(len(x0))
        first = (C.byref(variables.c_struct()), batch.surfaces) if self.prefix == "orc_" else (
            C.byref(batch.c_struct()), C.byref(variables.c_struct()))
        self.check(self.fn("perturbed_vector")(*first, abi.ptr(x0), abi.ptr(rel), abi.ptr(ab),
                                               distribution, seed, int(first_is_base), j,
                                               abi.ptr(y)))
        return y

    def optimize_lm(self, base, settings, variables, x0, options):
        x0 = np.ascontiguousarray(np.atleast_2d(x0), dtype=np.float64)
        runs, P = x0.shape
        x = np.zeros((runs, P))
        res = (abi.LmRun * runs)()
        cycles = C.c_int(0)
        self.check(self.fn("optimize_lm")(
            C.byref(base.c_struct()), C.byref(settings.c_struct()),
            C.byref(variables.c_struct()), runs, abi.ptr(x0), C.byref(options), abi.ptr(x),
            res, C.byref(cycles)))
        return x, list(res), cycles.value

    def optimize_lm_hip_backend(self, engine, base, settings, variables, x0, options,
                                precision):
        """The reference's optimize_lm_rounds driving HipLmBackend, its LmBackend plug-in over the
        HIP library's mko_engine_normal_equations / mko_engine_point_merit (reference build
        only); `engine` is a mikroptim_amd TraceEngine."""
        from mikroptim_amd import engine as E

        lib = E.library()
        calls = (C.c_void_p * 3)(*[C.cast(getattr(lib, n), C.c_void_p).value for n in (
            "mko_engine_normal_equations", "mko_engine_point_merit", "mko_last_error")])
        x0 = np.ascontiguousarray(np.atleast_2d(x0), dtype=np.float64)
        runs, P = x0.shape
        x = np.zeros((runs, P))
        res = (abi.LmRun * runs)()
        cycles = C.c_int(0)
        f = self.fn("optimize_lm_hip_backend")
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                      C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        b, s, v = base.c_struct(), settings.c_struct(), variables.c_struct()
        self.check(f(calls, engine._h, C.byref(b), C.byref(s), C.byref(v), runs, abi.ptr(x0),
                     C.byref(options), E._PRECISION[precision], abi.ptr(x), res,
                     C.byref(cycles)))
        return x, list(res), cycles.value

    def variables_view(self, base, variables):
        """The reference-side adapter's variable description (mko::hip::VariablesView of the
        reference's VariableSet built from this description; reference build only)."""
        P = variables.size
        out = {k: np.zeros(P, dtype=np.int32) for k in ("type", "surface", "component")}
        out.update({k: np.zeros(P) for k in ("lower", "upper", "typical")})
        count = C.c_int32(0)
        self.check(self.fn("variables_view")(
            C.byref(base.c_struct()), C.byref(variables.c_struct()), C.byref(count),
            *[abi.ptr(out[k], C.c_int32) for k in ("type", "surface", "component")],
            *[abi.ptr(out[k]) for k in ("lower", "upper", "typical")]))
        out["count"] = count.value
        return out

    def fit_buchdahl(self, glass, dispersion_terms=3, lambda_min=400.0, lambda_max=700.0,
                     samples=301, nodes="chebyshev", check_samples=3001):
        """The reference's fit_buchdahl of a catalogue glass (reference build only): n0, nu1,
        nu2, nu3, max_abs_error, rms_error."""
        poly = glass.model == "polynomial"
        b = np.array(glass.coefficients if poly else glass.B, dtype=np.float64)
        c = np.array(glass.exponents if poly else glass.C, dtype=np.float64)
        out = np.zeros(6)
        f = self.fn("fit_buchdahl")
        f.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int,
                      C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp]
        f.restype = C.c_int
        self.check(f(1 if poly else 0, glass.A, abi.ptr(b), abi.ptr(c), len(b),
                     glass.range_nm[0], glass.range_nm[1], dispersion_terms, lambda_min,
                     lambda_max, samples, 0 if nodes == "chebyshev" else 1, check_samples,
                     abi.ptr(out)))
        return out

    def glass_hull_rows(self, coords, weight, margin, x):
        """The reference's glass hull of the (n0, nu1, nu2) rows of `coords` and its residuals
        glass_hull_residual at the points x [m][3] (reference build only): (rows [m][3], info:
        dimension, center, scale, vertex and triangle counts)."""
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        out = np.zeros((len(x), 3))
        info = np.zeros(9)
        f = self.fn("glass_hull_rows")
        f.argtypes = [_dp, C.c_int, C.c_double, C.c_double, _dp, C.c_int, _dp, _dp]
        f.restype = C.c_int
        self.check(f(abi.ptr(coords), len(coords), weight, margin, abi.ptr(x), len(x),
                     abi.ptr(out), abi.ptr(info)))
        return out, info

    def gauss_legendre(self, n):
        u = np.zeros(n)
        w = np.zeros(n)
        self.check(self.fn("gauss_legendre")(n, abi.ptr(u), abi.ptr(w)))
        return u, w

    def pupil_samples(self, rings, spokes, mode):
        n = rings * (1 if mode == abi.PUPIL_AXIS else spokes)
        rho, phi, w = np.zeros(n), np.zeros(n), np.zeros(n)
        self.check(self.fn("pupil_samples")(rings, spokes, mode, abi.ptr(rho), abi.ptr(phi),
                                            abi.ptr(w)))
        return rho, phi, w


_cache = {}


def oracle():
    if "orc" not in _cache:
        if not os.path.exists(ORACLE_SO):
            build()
        _cache["orc"] = Lib(ORACLE_SO, "orc_")
    return _cache["orc"]


def reference():
    """The reference build, or None where it is absent (e.g. no sources to build it from)."""
    if "ref" not in _cache:
        if not os.path.exists(REF_SO) and os.path.isdir("/root/reference/include"):
            build()
        _cache["ref"] = Lib(REF_SO, "ref_") if os.path.exists(REF_SO) else None
    return _cache["ref"]


DF64_OPS = {"add": 0, "sub": 1, "mul": 2, "mul_fp": 3, "div": 4, "sqrt": 5}


def df64_split(x):
    """float pairs (hi, lo) of doubles x: hi = float(x), lo = float(x - hi) -> [n][2] float32."""
    x = np.asarray(x, dtype=np.float64)
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    return np.ascontiguousarray(np.stack([hi, lo], axis=-1))


def df64_ops(lib, op, a, b=None):
    """Element-wise df64 operation of the oracle or reference driver over [n][2] float pairs."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.zeros_like(a) if b is None else np.ascontiguousarray(b, dtype=np.float32)
    out = np.zeros_like(a)
    fn = getattr(lib.lib, lib.prefix + "df64_ops")
    rc = fn(DF64_OPS[op], a.shape[0], a.ctypes.data, b.ctypes.data, out.ctypes.data)
    assert rc == 0
    return out

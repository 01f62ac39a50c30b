"""Least-squares Buchdahl model of a glass's index (include/mikroptim/glass/buchdahl_fit.hpp,
src/glass/buchdahl_fit.cpp): the dispersion model glass variables optimize over. A catalogue
glass under a glass variable is replaced by the fit of its catalogue index, as
convert_glass_variables does before the variables are resolved (src/optim/glass_variables.cpp:
26-62).

    n(lambda) = n0 + nu1 w + nu2 w^2 + nu3 w^3,  w = (lambda - lambda0) / (1 + 2.5 (lambda - lambda0))

with lambda in um and lambda0 = 0.546074 um. The fit samples the index at `samples` wavelengths
in [lambda_min, lambda_max] (Chebyshev nodes of the first kind, or equidistant), solves the
overdetermined Vandermonde system in w by Householder QR in float64 (not by normal equations)
and reports the largest and the RMS deviation from the index on an equidistant check grid. The
arithmetic follows the reference's operation order step for step, so a fit equals the reference
build's bit for bit (the tests compare them). Host code: the fit runs once per system file, before
any batch reaches the device."""

import math
from dataclasses import dataclass, field

LAMBDA0_UM = 0.546074
ALPHA = 2.5


class FitError(ValueError):
    """An invalid fit request (the reference's GlassError from the fit)."""


@dataclass
class FitOptions:
    """BuchdahlFitOptions (buchdahl_fit.hpp:32-40)."""
    lambda_min_nm: float = 400.0
    lambda_max_nm: float = 700.0
    samples: int = 301
    nodes: str = "chebyshev"  # or "equidistant"
    check_samples: int = 3001


@dataclass
class BuchdahlFit:
    """BuchdahlFit (buchdahl_fit.hpp:42-52): n0, (nu1, nu2, nu3) (nu3 = 0 for 2 terms) and the
    errors on the check grid."""
    dispersion_terms: int
    n0: float
    nu: tuple
    max_abs_error: float
    rms_error: float
    options: FitOptions = field(default_factory=FitOptions)

    def index(self, lambda_nm):
        """buchdahl_index at lambda_nm (the same evaluation as a Buchdahl medium's index)."""
        return buchdahl_index(self.n0, self.nu[0], self.nu[1], self.nu[2], lambda_nm / 1000.0)

    def material(self):
        """to_buchdahl_material: the medium ("buchdahl", n0, nu1, nu2, nu3) of a system file."""
        return ("buchdahl", self.n0, self.nu[0], self.nu[1], self.nu[2])


def buchdahl_denominator(lambda_um):
    return 1.0 + ALPHA * (lambda_um - LAMBDA0_UM)


def buchdahl_omega(lambda_um):
    return (lambda_um - LAMBDA0_UM) / buchdahl_denominator(lambda_um)


def buchdahl_index(n0, nu1, nu2, nu3, lambda_um):
    w = buchdahl_omega(lambda_um)
    return n0 + w * (nu1 + w * (nu2 + w * nu3))


def _fmt(x):
    return "%.10g" % x


def _validate(o):
    """validate_options (buchdahl_fit.cpp:17-35)."""
    lo, hi = o.lambda_min_nm, o.lambda_max_nm
    if not (math.isfinite(lo) and math.isfinite(hi) and lo < hi):
        raise FitError(f"Buchdahl fit: invalid range [{_fmt(lo)}, {_fmt(hi)}] nm (finite, "
                       "min < max expected)")
    if not buchdahl_denominator(lo / 1000.0) > 0.0:
        raise FitError(f"Buchdahl fit: the Buchdahl coordinate is undefined at {_fmt(lo)} nm "
                       "(needs lambda > 146.074 nm)")
    if o.samples < 2:
        raise FitError(f"Buchdahl fit: at least 2 fit nodes expected, not {o.samples}")
    if o.check_samples < 2:
        raise FitError(f"Buchdahl fit: the check grid needs at least 2 points, not "
                       f"{o.check_samples}")
    if o.nodes not in ("chebyshev", "equidistant"):
        raise FitError(f"Buchdahl fit: unknown node distribution '{o.nodes}'")


def fit_nodes_nm(options=None):
    """buchdahl_fit_nodes_nm (buchdahl_fit.cpp:105-125): Chebyshev nodes
    m - h cos(pi (k + 1/2) / n), ascending, without the end points; or equidistant with both."""
    o = options or FitOptions()
    _validate(o)
    n, lo, hi = o.samples, o.lambda_min_nm, o.lambda_max_nm
    out = []
    for k in range(n):
        if o.nodes == "chebyshev":
            mid = 0.5 * (lo + hi)
            half = 0.5 * (hi - lo)
            out.append(mid - half * math.cos(math.pi * (k + 0.5) / n))
        else:
            out.append(hi if k == n - 1 else lo + (hi - lo) * k / (n - 1))
    return out


def check_grid_nm(options=None):
    """buchdahl_check_grid_nm (buchdahl_fit.cpp:127-137): equidistant, both end points."""
    o = options or FitOptions()
    _validate(o)
    n, lo, hi = o.check_samples, o.lambda_min_nm, o.lambda_max_nm
    return [hi if i == n - 1 else lo + (hi - lo) * i / (n - 1) for i in range(n)]


def _householder_least_squares(a, b, m, p):
    """min ||A x - b|| by Householder QR (buchdahl_fit.cpp:48-101): A column-major (a[j][i]),
    each reflection's sign chosen against cancellation in v_k, applied to the remaining columns
    and to b, then back substitution. a and b are overwritten."""
    for k in range(p):
        col = a[k]
        norm2 = 0.0
        for i in range(k, m):
            norm2 += col[i] * col[i]
        norm = math.sqrt(norm2)
        if not norm > 0.0:
            raise FitError("Buchdahl fit: the least-squares problem is rank deficient")
        alpha = -norm if col[k] > 0.0 else norm
        col[k] -= alpha
        vtv = 0.0
        for i in range(k, m):
            vtv += col[i] * col[i]
        for j in range(k + 1, p):
            cj = a[j]
            s = 0.0
            for i in range(k, m):
                s += col[i] * cj[i]
            f = 2.0 * s / vtv
            for i in range(k, m):
                cj[i] -= f * col[i]
        s = 0.0
        for i in range(k, m):
            s += col[i] * b[i]
        f = 2.0 * s / vtv
        for i in range(k, m):
            b[i] -= f * col[i]
        col[k] = alpha
    x = [0.0] * p
    for k in range(p - 1, -1, -1):
        s = b[k]
        for j in range(k + 1, p):
            s -= a[j][k] * x[j]
        x[k] = s / a[k][k]
    return x


def _checked(index_nm, lambda_nm):
    n = index_nm(lambda_nm)
    if not math.isfinite(n):
        raise FitError(f"Buchdahl fit: the index at {_fmt(lambda_nm)} nm is not finite")
    return n


def fit(index_nm, dispersion_terms=3, options=None):
    """fit_buchdahl (buchdahl_fit.cpp:139-186) to an index function of the wavelength in nm:
    design matrix A[i][j] = w_i^j, right-hand side n(lambda_i); errors on the check grid with the
    Buchdahl medium's own evaluation."""
    o = options or FitOptions()
    if dispersion_terms not in (2, 3):
        raise FitError(f"Buchdahl fit: 2 or 3 dispersion terms expected, not {dispersion_terms}")
    p = dispersion_terms + 1
    nodes = fit_nodes_nm(o)
    m = o.samples
    if m < p:
        raise FitError(f"Buchdahl fit: {m} fit nodes are not enough for {p} unknowns")
    a = [[0.0] * m for _ in range(p)]
    b = [0.0] * m
    for i, lam in enumerate(nodes):
        w = buchdahl_omega(lam / 1000.0)
        wj = 1.0
        for j in range(p):
            a[j][i] = wj
            wj *= w
        b[i] = _checked(index_nm, lam)
    x = _householder_least_squares(a, b, m, p) + [0.0] * (4 - p)
    out = BuchdahlFit(dispersion_terms, x[0], (x[1], x[2], x[3]), 0.0, 0.0, o)
    max_err, sum_sq = 0.0, 0.0
    grid = check_grid_nm(o)
    for lam in grid:
        e = out.index(lam) - _checked(index_nm, lam)
        max_err = max(max_err, abs(e))
        sum_sq += e * e
    out.max_abs_error = max_err
    out.rms_error = math.sqrt(sum_sq / len(grid))
    return out


def fit_glass(glass, dispersion_terms=3, options=None):
    """fit_buchdahl of a catalogue glass (buchdahl_fit.cpp:188-203): the fit range must lie in
    the glass's valid range; errors name the glass."""
    o = options or FitOptions()
    lo, hi = glass.range_nm

    def inside(x):
        return lo <= x <= hi
    if not (inside(o.lambda_min_nm) and inside(o.lambda_max_nm)):
        raise FitError(f"glass {glass.name}: fit range [{_fmt(o.lambda_min_nm)}, "
                       f"{_fmt(o.lambda_max_nm)}] nm is not within the valid range "
                       f"[{_fmt(lo)}, {_fmt(hi)}] nm")
    try:
        return fit(glass.index, dispersion_terms, o)
    except FitError as e:
        raise FitError(f"glass {glass.name}: {e}") from None

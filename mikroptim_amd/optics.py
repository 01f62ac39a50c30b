"""Host-side optics helpers used to build batches: Sellmeier dispersion and the paraxial y-nu
trace.

sellmeier_index follows include/mikroptim/glass/dispersion.hpp:27 (n^2 = 1 + sum B_i l^2 /
(l^2 - C_i), lambda in um, C in um^2). paraxial_trace follows
include/mikroptim/paraxial/ynu_trace.hpp (n'u' = nu - y c (n' - n); y += t u); focal length
1/phi and back focal distance -y/u as in src/paraxial/paraxial.cpp:151-170.
"""

import math

# N-BK7 (data/glass/basis.json of the reference; Schott datasheet coefficients)
NBK7_B = (1.03961212, 0.231792344, 1.01046945)
NBK7_C = (0.00600069867, 0.0200179144, 103.560653)


def sellmeier_index(B, C, lambda_um):
    l2 = lambda_um * lambda_um
    n2 = 1.0
    for b, c in zip(B, C):
        n2 += b * l2 / (l2 - c)
    return math.sqrt(n2)


def synthetic_glass(nd, vd):
    """Sellmeier coefficients of a synthetic glass with index nd at 587.56 nm and Abbe number
    close to vd: the N-BK7 form with the short-wavelength resonances scaled (index) and the UV
    resonance wavelength shifted (dispersion)."""
    # dispersion: move the first UV pole; solve for the pole position that gives vd, then the
    # overall strength that gives nd
    def make(k, c1):
        return (NBK7_B[0] * k, NBK7_B[1] * k, NBK7_B[2]), (c1, NBK7_C[1], NBK7_C[2])

    def abbe(B, C):
        nd_ = sellmeier_index(B, C, 0.58756)
        nf = sellmeier_index(B, C, 0.48613)
        nc = sellmeier_index(B, C, 0.65627)
        return nd_, (nd_ - 1.0) / (nf - nc)

    def solve_k(c1):
        lo, hi = 0.01, 20.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if abbe(*make(mid, c1))[0] < nd:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    lo, hi = 1e-4, 0.12  # larger pole -> more dispersion -> smaller Abbe number
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        k = solve_k(mid)
        if abbe(*make(k, mid))[1] > vd:
            lo = mid
        else:
            hi = mid
    c1 = 0.5 * (lo + hi)
    return make(solve_k(c1), c1)


def paraxial_trace(c, t, n, y=1.0, u=0.0):
    """Marginal ray through surfaces 0 ... S-1 (n = medium after each surface, air before 0).
    Returns (y, u) after the last surface."""
    n_before = 1.0
    for s in range(len(c)):
        if s > 0:
            y += t[s - 1] * u
        phi = c[s] * (n[s] - n_before)
        u = (n_before * u - y * phi) / n[s]
        n_before = n[s]
    return y, u


def focal_and_bfd(c, t, n):
    """Effective focal length and back focal distance (from the last vertex) for an axial
    collimated ray (paraxial.cpp:151-170)."""
    y, u = paraxial_trace(c, t, n, 1.0, 0.0)
    yl, ul = y, u
    # system power: phi = -u' / y_1 with y_1 = 1 (ray enters at unit height)
    phi = -ul
    return 1.0 / phi, -yl / ul

"""Deterministic, platform-independent variant perturbations (host side).

Follows src/optim/variant_generator.cpp:67-127: the generator of variant j is a 64-bit Mersenne
twister (MT19937-64, Matsumoto & Nishimura 2000, as std::mt19937_64) seeded with
splitmix64(seed + j * 0x9E3779B97F4A7C15); uniform u = (bits >> 11) * 2^-53; Uniform xi = 2u - 1,
IrwinHall xi = sum of 12 u - 6; y = fma(s, xi, x0) with s = fma(rel, |x0|, abs), then clamped.
The MT state is advanced for all variants at once (numpy over variants); fma is evaluated exactly
with rational arithmetic and rounded once.
"""

import math
from fractions import Fraction

import numpy as np

from . import abi

MASK64 = (1 << 64) - 1
_NN, _MM = 312, 156
_MATRIX_A = np.uint64(0xB5026F5AA96619E9)
_UM = np.uint64(0xFFFFFFFF80000000)
_LM = np.uint64(0x7FFFFFFF)


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


def variant_seed(seed, j):
    return splitmix64((seed + j * 0x9E3779B97F4A7C15) & MASK64)


def unit_uniform(bits):
    return float(bits >> 11) * 2.0 ** -53


class MT19937_64:
    """Vectorised std::mt19937_64: one independent generator per entry of `seeds`."""

    def __init__(self, seeds):
        seeds = [int(s) & MASK64 for s in seeds]
        n = len(seeds)
        mt = np.zeros((_NN, n), dtype=np.uint64)
        prev = [s for s in seeds]
        mt[0] = np.array(prev, dtype=np.uint64)
        with np.errstate(over="ignore"):
            cur = mt[0].copy()
            for i in range(1, _NN):
                cur = np.uint64(6364136223846793005) * (cur ^ (cur >> np.uint64(62))) + np.uint64(i)
                mt[i] = cur
        self.mt = mt
        self.index = _NN

    def _twist(self):
        mt = self.mt
        one = np.uint64(1)
        zero = np.uint64(0)
        for i in range(_NN):
            x = (mt[i] & _UM) | (mt[(i + 1) % _NN] & _LM)
            xa = (x >> one) ^ np.where((x & one) != zero, _MATRIX_A, zero)
            mt[i] = mt[(i + _MM) % _NN] ^ xa
        self.index = 0

    def __call__(self):
        if self.index >= _NN:
            self._twist()
        y = self.mt[self.index].copy()
        self.index += 1
        y ^= (y >> np.uint64(29)) & np.uint64(0x5555555555555555)
        y ^= (y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000)
        y ^= (y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000)
        y ^= y >> np.uint64(43)
        return y


def _uniform(bits):
    return (bits >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


_SPLIT = 134217729.0  # 2^27 + 1 (Veltkamp)


def _two_product(a, b):
    """p + e = a * b exactly (Dekker 1971), elementwise; valid without over/underflow."""
    p = a * b
    t = _SPLIT * a
    ah = t - (t - a)
    al = a - ah
    t = _SPLIT * b
    bh = t - (t - b)
    bl = b - bh
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
    return p, e


def fma_exact(a, b, c):
    """fma(a, b, c) rounded once: exact product (Dekker) summed with c by math.fsum, which
    rounds the exact sum of its arguments once."""
    a, b, c = np.broadcast_arrays(np.asarray(a, np.float64), np.asarray(b, np.float64),
                                  np.asarray(c, np.float64))
    p, e = _two_product(a, b)
    out = np.empty(a.shape)
    flat_p, flat_e, flat_c, flat_o = p.ravel(), e.ravel(), c.ravel(), out.ravel()
    for i in range(flat_o.size):
        flat_o[i] = math.fsum((flat_p[i], flat_e[i], flat_c[i]))
    return out


def perturbed_vectors(variables, x0, count, seed=0, relative=(0.01, 0.01, 0.01),
                      absolute=(0.0, 0.0, 0.0), distribution="uniform", first_is_base=False):
    """Vectors of variants j = 0 ... count-1 (perturbed_vector, variant_generator.cpp:95).
    relative/absolute: per MKO_VAR_* type (curvature, thickness, working distance[, glass]);
    the glass scale defaults to PerturbationOptions::glass (0.01 relative)."""
    x0 = np.asarray(x0, dtype=np.float64)
    P = len(x0)
    relative = tuple(relative) + (0.01,) * (4 - len(relative))
    absolute = tuple(absolute) + (0.0,) * (4 - len(absolute))
    gen = MT19937_64([variant_seed(seed, j) for j in range(count)])
    out = np.empty((count, P))
    s = [_fma(relative[t], abs(float(x)), absolute[t]) for t, x in zip(variables.type, x0)]
    for k in range(P):
        if distribution == "uniform":
            xi = 2.0 * _uniform(gen()) - 1.0
        else:
            acc = np.zeros(count)
            for _ in range(12):
                acc = acc + _uniform(gen())
            xi = acc - 6.0
        out[:, k] = fma_exact(s[k], xi, x0[k])
    out = variables.clamp(out)
    if first_is_base and count > 0:
        out[0] = x0
    return out

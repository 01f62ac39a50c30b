// Per-ray device math of the trace and of its reverse (adjoint) sweep, templated on the number
// type R (double or df64 for the forward trace; double for the reverse sweep).
//
// Formulas and branch rules follow the reference's shared math core:
//   surface steps        include/mikroptim/trace/surface_ops.hpp:122-203
//   surface propagation  include/mikroptim/trace/trace.hpp:175-229 (stop/plane/sphere only)
//   reference spheres    include/mikroptim/trace/reference_sphere.hpp:56-143
//   reverse steps        include/mikroptim/trace/adjoint_backward.hpp:115-241, 373-464
// The FP64 surface step spells its fmas out (no contraction left to the compiler, so every kernel
// that inlines it computes the same bits); the rest is compiled with FMA contraction (hipcc
// default). Results differ from the FP64 CPU oracle by rounding only, which the parity tests
// bound per ray.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "df64.hpp"
#include "mikroptim_types.h"

namespace mkod {

// Number-type helpers of the forward trace: R = double (every mode) or df64 (the RESIDUALS /
// MERIT traces with MKO_PRECISION_DF64). For double they are the identity.
template <class R>
__device__ __forceinline__ R num(double x) {
    return x;
}
template <>
__device__ __forceinline__ df64 num<df64>(double x) {
    return df64_from(x);
}
__device__ __forceinline__ double to_double(double x) { return x; }
__device__ __forceinline__ double to_double(df64 x) { return df64_to(x); }

template <class R>
struct V3 {
    R x, y, z;
};
template <class R>
__device__ __forceinline__ V3<R> operator+(V3<R> a, V3<R> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class R>
__device__ __forceinline__ V3<R> operator-(V3<R> a, V3<R> b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class R>
__device__ __forceinline__ V3<R> operator*(R s, V3<R> a) {
    return {s * a.x, s * a.y, s * a.z};
}
template <class R>
__device__ __forceinline__ V3<R>& operator+=(V3<R>& a, V3<R> b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
template <class R>
__device__ __forceinline__ R dot(V3<R> a, V3<R> b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
__device__ __forceinline__ V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ df64 dsqrt(df64 x) { return sqrt(x); }
template <class R>
__device__ __forceinline__ R rnorm(V3<R> a) {
    return num<R>(1.0) / dsqrt(dot(a, a));
}
__device__ __forceinline__ float rnorm(V3<float> a) { return 1.0f / dsqrt(dot(a, a)); }

constexpr uint8_t kBranchRegular = 0, kBranchSoftMissed = 1, kBranchSoftTIR = 2;

__device__ __forceinline__ bool in_piston(uint8_t s) {
    return s == MKO_STATUS_OK || s == MKO_STATUS_SOFT_MISSED || s == MKO_STATUS_SOFT_TIR;
}
__device__ __forceinline__ bool is_soft(uint8_t s) {
    return s == MKO_STATUS_SOFT_MISSED || s == MKO_STATUS_SOFT_TIR;
}

// Division, reciprocal and square root of the reverse sweep, of the state reconstruction and of
// the fused kernel's forward trace: hardware reciprocal / reciprocal square root refined by
// Newton steps (a few ulp) instead of the correctly rounded IEEE sequences with their scaling and
// fix-up steps. These paths need ~1e-15 relative accuracy (Jacobian columns are compared to
// 1e-10, OPD to 1e-12 mm per ray); the df64 trace keeps its own operations.
#ifndef MKO_REVERSE_FAST_MATH
#define MKO_REVERSE_FAST_MATH 1
#endif
// Measured on MI355X (2^20 arguments over 2^+-30): v_rcp_f64 and v_rsq_f64 are accurate to
// ~5e-8; one Newton step gives 2.1e-15, two give <= 0.5 ulp. The quotient a * r with one residual
// correction and the square root after one Goldschmidt step plus one correction are within
// 0.5 ulp (1.1e-16) as well.
__device__ __forceinline__ double rcp_nr1(double b) {
#if MKO_REVERSE_FAST_MATH
    const double r = __builtin_amdgcn_rcp(b);
    return __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
#else
    return 1.0 / b;
#endif
}
__device__ __forceinline__ double rev_rcp(double b) {
#if MKO_REVERSE_FAST_MATH
    const double r = rcp_nr1(b);
    return __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
#else
    return 1.0 / b;
#endif
}
__device__ __forceinline__ double rev_div(double a, double b) {
#if MKO_REVERSE_FAST_MATH
    const double r = rcp_nr1(b);
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
#else
    return a / b;
#endif
}
__device__ __forceinline__ double rev_sqrt(double x) {
#if MKO_REVERSE_FAST_MATH
    if (!(x > 0.0)) return __builtin_sqrt(x);
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
#else
    return __builtin_sqrt(x);
#endif
}

// rev_sqrt that also returns 1 / (2 sqrt(x)): the Goldschmidt iteration's refined half reciprocal
// (~1e-15 relative), for callers that divide by twice the root.
__device__ __forceinline__ double rev_sqrt_half_rcp(double x, double& half_rcp) {
#if MKO_REVERSE_FAST_MATH
    if (!(x > 0.0)) {
        const double g = __builtin_sqrt(x);
        half_rcp = 0.5 / g;
        return g;
    }
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    half_rcp = h;
    return __builtin_fma(d, h, g);
#else
    const double g = __builtin_sqrt(x);
    half_rcp = 0.5 / g;
    return g;
#endif
}

template <class R>
struct RayAdj {
    V3<R> p, d;
};

// Forward-trace division / square root: IEEE, or the Newton-refined forms above (kFast, double).
template <bool kFast, class R>
__device__ __forceinline__ R fdiv(R a, R b) {
    if constexpr (kFast) return rev_div(a, b);
    else return a / b;
}
template <bool kFast, class R>
__device__ __forceinline__ R fsqrt(R x) {
    if constexpr (kFast) return rev_sqrt(x);
    else return dsqrt(x);
}

// Accumulator of soft ray failure (SoftFailure, surface_ops.hpp:94).
struct Soft {
    double penalty;
    uint8_t first;
    __device__ __forceinline__ void add(double amount, uint8_t kind) {
        penalty += amount;
        if (first == MKO_STATUS_OK) first = kind;
    }
};

template <class R>
struct RayT {
    V3<R> p, d;
    R opl;
};
using Ray = RayT<double>;

// The FP64 surface step with its arithmetic spelled out: each product that fuses into a sum is
// one fma and every other operation is rounded on its own (no contraction is left to the
// compiler), so every kernel that inlines the step computes the same bits, in hard and in soft
// failure mode.
struct StepQuadratic {
    double s0, px, py, H, B, D;
};
// transfer_to_tangent_plane and the coefficients of intersect_sphere's quadratic
template <bool kFast>
__device__ __forceinline__ StepQuadratic step_quadratic(const RayT<double>& r, double c,
                                                        double t) {
#pragma clang fp contract(off)
    StepQuadratic q;
    q.s0 = fdiv<kFast>(-(r.p.z - t), r.d.z);
    q.px = __builtin_fma(q.s0, r.d.x, r.p.x);
    q.py = __builtin_fma(q.s0, r.d.y, r.p.y);
    q.H = c * __builtin_fma(q.py, q.py, q.px * q.px);
    q.B = __builtin_fma(-c, __builtin_fma(r.d.y, q.py, r.d.x * q.px), r.d.z);
    q.D = __builtin_fma(q.B, q.B, -(c * q.H));
    return q;
}
__device__ __forceinline__ double dot_fma(V3<double> a, V3<double> b) {
#pragma clang fp contract(off)
    return __builtin_fma(a.z, b.z, __builtin_fma(a.y, b.y, a.x * b.x));
}
// the intersection p + s d (p.z = 0 on the tangent plane) and the surface normal there
__device__ __forceinline__ V3<double> step_point(const StepQuadratic& q, V3<double> d, double s) {
    return {__builtin_fma(s, d.x, q.px), __builtin_fma(s, d.y, q.py), __builtin_fma(s, d.z, 0.0)};
}
__device__ __forceinline__ V3<double> step_normal(V3<double> p, double c) {
#pragma clang fp contract(off)
    return {-c * p.x, -c * p.y, __builtin_fma(-c, p.z, 1.0)};
}
// refract_sphere: k = 1 - mu^2 (1 - cos^2 I) and, for k >= 0, mu d + (sqrt(k) - mu cos I) N
__device__ __forceinline__ double refract_k(double mu, double cosI) {
#pragma clang fp contract(off)
    return __builtin_fma(-(mu * mu), __builtin_fma(-cosI, cosI, 1.0), 1.0);
}
__device__ __forceinline__ V3<double> refract_regular(V3<double> d, V3<double> N, double mu,
                                                      double cosI, double sqrt_k) {
#pragma clang fp contract(off)
    const double g = __builtin_fma(-mu, cosI, sqrt_k);
    return {__builtin_fma(N.x, g, d.x * mu), __builtin_fma(N.y, g, d.y * mu),
            __builtin_fma(N.z, g, d.z * mu)};
}

// Hard failure mode: the step's operations run for every ray and one test of the four failure
// conditions (backward, missed, no positive root, TIR; the precedence of the early returns of
// the soft form) ends it, so a surface costs one exit of the surface loop instead of four. The
// ray is updated in every case: the callers drop a ray whose step failed.
template <bool kFast>
__device__ __forceinline__ uint8_t surface_step_hard(RayT<double>& r, double c, double t,
                                                     double nb, double mu) {
#pragma clang fp contract(off)
    const V3<double> d = r.d;
    const StepQuadratic q = step_quadratic<kFast>(r, c, t);
    const double denom = q.B + fsqrt<kFast>(q.D);
    const double s = fdiv<kFast>(q.H, denom);
    const V3<double> p = step_point(q, d, s);
    const V3<double> N = step_normal(p, c);
    const double cosI = dot_fma(d, N);
    const double k = refract_k(mu, cosI);
    r.opl = __builtin_fma(nb, q.s0 + s, r.opl);
    r.p = p;
    r.d = refract_regular(d, N, mu, cosI, fsqrt<kFast>(k));
    const bool backward = !(d.z > 0.0), missed = q.D < 0.0 || !(denom > 0.0), tir = k < 0.0;
    if (backward || missed || tir) [[unlikely]]
        return backward ? uint8_t(MKO_STATUS_BACKWARD)
               : missed ? uint8_t(MKO_STATUS_MISSED)
                        : uint8_t(MKO_STATUS_TIR);
    return MKO_STATUS_OK;
}

// Soft failure mode: a miss on a curved surface continues from the quadratic's vertex (s = B / c,
// unit normal) and TIR continues along the surface tangent, each adding its penalty.
template <bool kFast>
__device__ __forceinline__ uint8_t surface_step_soft(RayT<double>& r, double c, double t,
                                                     double nb, double mu, Soft& soft,
                                                     uint8_t& branch) {
#pragma clang fp contract(off)
    const V3<double> d = r.d;
    if (!(d.z > 0.0)) return MKO_STATUS_BACKWARD;
    const StepQuadratic q = step_quadratic<kFast>(r, c, t);
    double s;
    bool missed = false;
    if (q.D < 0.0) {
        if (!(c > 0.0) && !(c < 0.0)) return MKO_STATUS_MISSED;
        soft.add(-q.D, MKO_STATUS_SOFT_MISSED);
        s = q.B / c;
        missed = true;
    } else {
        const double denom = q.B + fsqrt<kFast>(q.D);
        if (!(denom > 0.0)) return MKO_STATUS_MISSED;
        s = fdiv<kFast>(q.H, denom);
    }
    const V3<double> p = step_point(q, d, s);
    r.p = p;
    r.opl = __builtin_fma(nb, q.s0 + s, r.opl);
    V3<double> N = step_normal(p, c);
    if (missed) {
        const double w = 1.0 / __builtin_sqrt(dot_fma(N, N));
        N = V3<double>{w * N.x, w * N.y, w * N.z};
    }
    const double cosI = dot_fma(d, N);
    const double k = refract_k(mu, cosI);
    if (k < 0.0) {
        soft.add(-k, MKO_STATUS_SOFT_TIR);
        const V3<double> tau{__builtin_fma(-cosI, N.x, d.x), __builtin_fma(-cosI, N.y, d.y),
                             __builtin_fma(-cosI, N.z, d.z)};
        const double w = 1.0 / __builtin_sqrt(dot_fma(tau, tau));
        r.d = V3<double>{w * tau.x, w * tau.y, w * tau.z};
        branch = uint8_t((missed ? kBranchSoftMissed : 0) | kBranchSoftTIR);
        return MKO_STATUS_OK;
    }
    r.d = refract_regular(d, N, mu, cosI, fsqrt<kFast>(k));
    branch = missed ? kBranchSoftMissed : kBranchRegular;
    return MKO_STATUS_OK;
}

// One surface step (propagate_surface_local with transfer, sphere intersection and refraction).
// c = effective curvature (0 for stop and plane), t = vertex distance from the previous surface,
// nb = medium before, mu = nb / n_after. Returns a TraceStatus (Ok also for soft continuations,
// whose kind goes to `soft` and whose branch bits go to `branch`). FP64 takes the spelled-out
// forms above; FP32 (the prefilter) and df64 the general form below.
template <bool kSoft, bool kFast = false, class R>
__device__ __forceinline__ uint8_t surface_step(RayT<R>& r, R c, R t, R nb, R mu, Soft& soft,
                                                uint8_t& branch) {
    if constexpr (std::is_same_v<R, double>) {
        if constexpr (kSoft) {
            return surface_step_soft<kFast>(r, c, t, nb, mu, soft, branch);
        } else {
            branch = kBranchRegular;
            return surface_step_hard<kFast>(r, c, t, nb, mu);
        }
    }
    const R zero = num<R>(0.0), one = num<R>(1.0);
    if (!(r.d.z > zero)) return MKO_STATUS_BACKWARD;
    // transfer_to_tangent_plane
    r.p.z -= t;
    const R s0 = fdiv<kFast>(-r.p.z, r.d.z);
    r.p.x += s0 * r.d.x;
    r.p.y += s0 * r.d.y;
    r.p.z = zero;
    // intersect_sphere
    const R H = c * (r.p.x * r.p.x + r.p.y * r.p.y);
    const R B = r.d.z - c * (r.p.x * r.d.x + r.p.y * r.d.y);
    const R D = B * B - c * H;
    R s;
    bool missed = false;
    if (D < zero) {
        if (!kSoft || (!(c > zero) && !(c < zero))) return MKO_STATUS_MISSED;
        soft.add(-to_double(D), MKO_STATUS_SOFT_MISSED);
        s = B / c;
        missed = true;
    } else {
        const R denom = B + fsqrt<kFast>(D);
        if (!(denom > zero)) return MKO_STATUS_MISSED;
        s = fdiv<kFast>(H, denom);
    }
    r.p = r.p + s * r.d;
    r.opl += nb * (s0 + s);
    // refract_sphere
    V3<R> N{-c * r.p.x, -c * r.p.y, one - c * r.p.z};
    if (kSoft && missed) N = rnorm(N) * N;
    const R cosI = dot(r.d, N);
    const R k = one - mu * mu * (one - cosI * cosI);
    if (k < zero) {
        if (!kSoft) return MKO_STATUS_TIR;
        soft.add(-to_double(k), MKO_STATUS_SOFT_TIR);
        const V3<R> tau = r.d - cosI * N;
        r.d = rnorm(tau) * tau;
        branch = uint8_t((missed ? kBranchSoftMissed : 0) | kBranchSoftTIR);
        return MKO_STATUS_OK;
    }
    r.d = mu * r.d + (fsqrt<kFast>(k) - mu * cosI) * N;
    branch = missed ? kBranchSoftMissed : kBranchRegular;
    return MKO_STATUS_OK;
}

// reference_sphere_step (kind last_surface) / image_sphere_step (image kinds).
template <bool kSoft, bool kFast = false, class R>
__device__ __forceinline__ uint8_t reference_step(RayT<R>& r, V3<R> M, R kappa, int kind,
                                                  R n_image, Soft& soft, uint8_t& branch) {
    const R zero = num<R>(0.0), one = num<R>(1.0);
    const V3<R> q = r.p - M;
    const R b = -dot(q, r.d);
    branch = kBranchRegular;
    if (kind == MKO_REF_LAST_SURFACE) {
        const R g = dot(r.p, r.p - num<R>(2.0) * M);
        const R disc = b * b - g;
        if (disc < zero) {
            const R radius_sq = dot(M, M);
            if (!kSoft || !(radius_sq > zero)) return MKO_STATUS_MISSED;
            soft.add(-to_double(disc) / to_double(radius_sq), MKO_STATUS_SOFT_MISSED);
            r.p = r.p + b * r.d;
            r.opl += n_image * b;
            branch = kBranchSoftMissed;
            return MKO_STATUS_SOFT_MISSED;
        }
        const R root = fsqrt<kFast>(disc);
        const bool real_image = M.z >= zero;
        const R denom = real_image ? b + root : b - root;
        if (!(real_image ? denom > zero : denom < zero)) return MKO_STATUS_MISSED;
        const R s = fdiv<kFast>(g, denom);
        r.p = r.p + s * r.d;
        r.opl += n_image * s;
        return MKO_STATUS_OK;
    }
    const V3<R> cq = cross(q, r.d);
    const R t = -dot(cq, cq);
    const R disc = one + kappa * kappa * t;
    R step;
    uint8_t st = MKO_STATUS_OK;
    if (disc >= zero) {
        step = b - fdiv<kFast>(kappa * t, fsqrt<kFast>(disc) + one);
    } else if (kSoft && disc < zero) {
        step = b - kappa * t;
        st = MKO_STATUS_SOFT_MISSED;
    } else {
        return MKO_STATUS_MISSED;
    }
    if (!(step >= zero) && !(step < zero)) return MKO_STATUS_MISSED;
    if (st == MKO_STATUS_SOFT_MISSED) {
        soft.add(-to_double(disc), MKO_STATUS_SOFT_MISSED);
        branch = kBranchSoftMissed;
    }
    r.p = r.p + b * r.d;
    r.opl += n_image * step;
    return st;
}

// ---- reverse sweep ------------------------------------------------------------------------------

// Reverse of one stop/plane/sphere step (refracting_step_backward). (p, d): state before the
// surface; adj: adjoint of the state after it on entry, of the state before it on return.
// Adds the adjoints of the vertex distance t, the curvature c and the indices to tb/cb/nbb/nab.
template <class R>
__device__ __forceinline__ void refracting_step_backward(V3<R> p, V3<R> d, R t, R c, R n_before,
                                                         R n_after, uint8_t branch, R seed_opl,
                                                         R seed_pen, RayAdj<R>& adj, R& tb,
                                                         R& cb_out, R& nbb, R& nab) {
    const R z = p.z - t;
    const R rdz = rev_rcp(d.z);
    const R s0 = -z * rdz;
    const R x = p.x + s0 * d.x;
    const R y = p.y + s0 * d.y;
    const R rr = x * x + y * y;
    const R H = c * rr;
    const R xd = x * d.x + y * d.y;
    const R B = d.z - c * xd;
    const R D = B * B - c * H;
    const bool missed = (branch & kBranchSoftMissed) != 0;
    const bool tir = (branch & kBranchSoftTIR) != 0;
    R root_d = R(0);
    R s;
    if (missed) {
        s = B / c;
    } else {
        root_d = rev_sqrt(D);
        s = rev_div(H, B + root_d);
    }
    const V3<R> q{x + s * d.x, y + s * d.y, s * d.z};
    const V3<R> n_raw{-c * q.x, -c * q.y, R(1) - c * q.z};
    R inv_norm = R(1);
    V3<R> N = n_raw;
    if (missed) {
        inv_norm = rnorm(n_raw);
        N = inv_norm * n_raw;
    }
    const R mu = rev_div(n_before, n_after);
    const R cos_i = dot(d, N);
    const R sin2 = R(1) - cos_i * cos_i;
    const R k = R(1) - mu * mu * sin2;

    const V3<R> qb_out = adj.p;
    const V3<R> d1b = adj.d;
    RayAdj<R> in{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
    V3<R> nb{R(0), R(0), R(0)};
    R cosb = R(0), kb = R(0), mub = R(0);
    if (tir) {
        const V3<R> tau = d - cos_i * N;
        const R inv_tau = rnorm(tau);
        const V3<R> d1 = inv_tau * tau;
        const V3<R> taub = inv_tau * (d1b - dot(d1, d1b) * d1);
        in.d += taub;
        cosb += -dot(taub, N);
        nb += (-cos_i) * taub;
        kb += -seed_pen;
    } else {
        const R root_k = rev_sqrt(k);
        const R a = root_k - mu * cos_i;
        in.d += mu * d1b;
        nb += a * d1b;
        const R ab = dot(d1b, N);
        mub += dot(d1b, d);
        kb += rev_div(ab, R(2) * root_k);
        mub += -ab * cos_i;
        cosb += -ab * mu;
    }
    mub += -kb * R(2) * mu * sin2;
    cosb += kb * R(2) * mu * mu * cos_i;
    in.d += cosb * N;
    nb += cosb * d;
    nbb += mub / n_after;
    nab += -mub * mu / n_after;
    V3<R> nrb = nb;
    if (missed) nrb = inv_norm * (nb - dot(N, nb) * N);
    const V3<R> qb = qb_out + (-c) * nrb;
    R cb = -dot(q, nrb);
    R xb = qb.x;
    R yb = qb.y;
    R sb = dot(qb, d);
    in.d += s * qb;
    sb += seed_opl * n_before;
    R s0b = seed_opl * n_before;
    nbb += seed_opl * (s0 + s);
    R hb = R(0), bb = R(0);
    if (missed) {
        bb += sb / c;
        cb += -sb * s / c;
        const R db = -seed_pen;
        bb += db * R(2) * B;
        cb += -db * H;
        hb += -db * c;
    } else {
        const R inv = rev_rcp(R(2) * root_d);
        hb += sb * inv;
        bb += -sb * s * R(2) * inv;
        cb += sb * s * s * inv;
    }
    cb += hb * rr;
    xb += hb * R(2) * c * x;
    yb += hb * R(2) * c * y;
    in.d.z += bb;
    cb += -bb * xd;
    xb += -bb * c * d.x;
    yb += -bb * c * d.y;
    in.d.x += -bb * c * x;
    in.d.y += -bb * c * y;
    // transfer
    in.p.x += xb;
    in.p.y += yb;
    s0b += xb * d.x + yb * d.y;
    in.d.x += xb * s0;
    in.d.y += yb * s0;
    const R qq = s0b * rdz;
    tb += qq;
    in.p.z += -qq;
    in.d.z += -qq * s0;
    cb_out += cb;
    adj = in;
}

// Reverse of the reference-sphere step; returns the adjoint of M_z (the working distance) and
// adds the adjoint of the image-space index to nib.
template <class R>
__device__ __forceinline__ R reference_step_backward(V3<R> p, V3<R> d, V3<R> M, R kappa, int kind,
                                                     R n_image, uint8_t branch, R seed_opl,
                                                     R seed_pen, RayAdj<R>& adj, R& nib) {
    const V3<R> q = p - M;
    const R b = -dot(q, d);
    const bool missed = (branch & kBranchSoftMissed) != 0;
    V3<R> center_adj{R(0), R(0), R(0)};
    if (kind == MKO_REF_LAST_SURFACE) {
        const R g = dot(p, p - R(2) * M);
        const R disc = b * b - g;
        const R sigma = M.z >= R(0) ? R(1) : R(-1);
        R s, root = R(0);
        if (missed) {
            s = b;
        } else {
            root = dsqrt(disc);
            s = g / (b + sigma * root);
        }
        const R sb = seed_opl * n_image;
        nib += seed_opl * s;
        R bb, gb;
        if (missed) {
            bb = sb;
            gb = R(0);
            const R radius_sq = dot(M, M);
            const R discb = -seed_pen / radius_sq;
            bb += discb * R(2) * b;
            gb += -discb;
            center_adj += (seed_pen * R(2) * disc / (radius_sq * radius_sq)) * M;
        } else {
            gb = sb * sigma / (R(2) * root);
            bb = -sb * sigma * s / root;
        }
        RayAdj<R> in{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
        in.p += (-bb) * d;
        in.d += (-bb) * q;
        center_adj += bb * d;
        in.p += (R(2) * gb) * (p - M);
        center_adj += (R(-2) * gb) * p;
        adj = in;
        return center_adj.z;
    }
    const V3<R> c = cross(q, d);
    const R t = -dot(c, c);
    R step, dstep_dt, tb = R(0);
    if (missed) {
        step = b - kappa * t;
        dstep_dt = -kappa;
        tb += seed_pen * (-(kappa * kappa));
    } else {
        const R u = dsqrt(R(1) + kappa * kappa * t);
        step = b - kappa * t / (u + R(1));
        dstep_dt = -kappa / (R(2) * u);
    }
    const R sb = seed_opl * n_image;
    nib += seed_opl * step;
    const R bb = sb;
    tb += sb * dstep_dt;
    const V3<R> qb = (-bb) * d + (R(-2) * tb) * cross(d, c);
    const V3<R> db = (-bb) * q + (R(-2) * tb) * cross(c, q);
    adj.p = qb;
    adj.d = db;
    center_adj += (R(-1)) * qb;
    return center_adj.z;
}

// ---- state reconstruction for the reverse sweep (no per-ray record) -----------------------------
//
// A regular (hard-mode) surface step maps the ray state on surface s (position in the frame of s,
// direction in medium n_s) to the state on surface j = s + 1 and is invertible: refraction is
// reversible (the tangential direction component scales by n_j / n_s, the normal component stays
// positive), and the segment in gap s is recovered by intersecting the reversed ray with surface
// s. inverse_step returns the state on s from the state (q, d1) on j; it is the forward transfer
// and intersection of surface_step in the frame mirrored at z = 0 (z -> -z, curvature -> -c),
// applied to the reversed ray, so it uses the same Spencer-Murty root.

template <class R>
struct RayState {
    V3<R> p, d;
};

// State on surface s (curvature c_s, vertex t_s before surface j) from the state on surface j
// (curvature c_j); n_s, n_j: media before and after surface j.
__device__ __forceinline__ RayState<double> inverse_step(V3<double> q, V3<double> d1, double c_s,
                                                         double t_s, double c_j, double n_s,
                                                         double n_j) {
    // inverse refraction at j: d = mu' d1 + (sqrt(k') - mu' cos') N, mu' = n_j / n_s
    const V3<double> N{-c_j * q.x, -c_j * q.y, 1.0 - c_j * q.z};
    const double mu = rev_div(n_j, n_s);
    const double cosI = dot(d1, N);
    const double k = 1.0 - mu * mu * (1.0 - cosI * cosI);
    const V3<double> d = mu * d1 + (rev_sqrt(k) - mu * cosI) * N;
    // reversed ray in the mirrored frame of s: origin (q.x, q.y, -(q.z + t_s)), direction
    // (-d.x, -d.y, d.z); transfer to the tangent plane z = 0 and intersect curvature -c_s
    const double s0 = rev_div(q.z + t_s, d.z);
    const double x = q.x - s0 * d.x;
    const double y = q.y - s0 * d.y;
    const double cm = -c_s;
    const double H = cm * (x * x + y * y);
    const double B = d.z + cm * (x * d.x + y * d.y);
    const double D = B * B - cm * H;
    const double s = rev_div(H, B + rev_sqrt(D));
    RayState<double> out;
    out.p = V3<double>{x - s * d.x, y - s * d.y, -(s * d.z)};
    out.d = d;
    return out;
}

// One reverse step of a regular (hard-mode) ray for the fused kernel: reconstructs the state on
// surface s from the state (q, d1) on surface j = s + 1 and propagates the adjoint (seed: OPL)
// through the surface-j step, sharing the quantities both need. With N the unit normal at q:
// d1 . N = sqrt(k) of the forward refraction, and the incidence cosine d . N equals the
// intersection's sqrt(D) (B - c s = sqrt(D) on the sphere), so the reverse step needs two square
// roots instead of four. mu = n_s / n_j, imu = n_j / n_s. Formulas: inverse_step above and
// refracting_step_backward (adjoint_backward.hpp:119-241) specialised to the regular branch.
// kIndex: also the adjoints of the indices before (n_s) and after the surface, added to nbb and
// nab (glass variables), with 1 / n_after formed as mu / n_s.
template <bool kIndex = false>
__device__ __forceinline__ RayState<double> fused_step_backward(
    V3<double> q, V3<double> d1, double c, double cs, double t, double n_s, double mu, double imu,
    RayAdj<double>& adj, double& tb, double& cb_out, double* nbb = nullptr,
    double* nab = nullptr) {
    // reconstruction
    const V3<double> N{-c * q.x, -c * q.y, 1.0 - c * q.z};
    const double root_k = dot(d1, N);
    const double kp = 1.0 - imu * imu * (1.0 - root_k * root_k);
    double inv; // 1 / (2 cos_i)
    const double cos_i = rev_sqrt_half_rcp(kp, inv); // = sqrt(D) of the forward intersection
    const V3<double> d = imu * d1 + (cos_i - imu * root_k) * N;
    const double rdz = rev_rcp(d.z);
    const double s = q.z * rdz;
    const double x = q.x - s * d.x;
    const double y = q.y - s * d.y;
    const double s0p = (q.z + t) * rdz;
    const double xp = q.x - s0p * d.x;
    const double yp = q.y - s0p * d.y;
    const double cm = -cs;
    const double Hp = cm * (xp * xp + yp * yp);
    const double Bp = d.z + cm * (xp * d.x + yp * d.y);
    const double Dp = Bp * Bp - cm * Hp;
    const double sp = rev_div(Hp, Bp + rev_sqrt(Dp));
    RayState<double> prev;
    prev.p = V3<double>{xp - sp * d.x, yp - sp * d.y, -(sp * d.z)};
    prev.d = d;
    const double s0 = t * rdz + sp;
    // adjoint of the step from s to j
    const double rr = x * x + y * y;
    const double xd = x * d.x + y * d.y;
    const V3<double> qb_out = adj.p;
    const V3<double> d1b = adj.d;
    RayAdj<double> in{{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    const double a = root_k - mu * cos_i;
    in.d += mu * d1b;
    V3<double> nb = a * d1b;
    const double ab = dot(d1b, N);
    const double kb = rev_div(ab, 2.0 * root_k);
    const double cosb = -ab * mu + kb * 2.0 * mu * mu * cos_i;
    in.d += cosb * N;
    nb += cosb * d;
    const V3<double> qb = qb_out + (-c) * nb;
    double cb = -dot(q, nb);
    double xb = qb.x;
    double yb = qb.y;
    const double sb = dot(qb, d) + n_s;
    in.d += s * qb;
    double s0b = n_s;
    const double hb = sb * inv;
    const double bb = -sb * s * 2.0 * inv;
    cb += sb * s * s * inv;
    cb += hb * rr;
    xb += hb * 2.0 * c * x;
    yb += hb * 2.0 * c * y;
    in.d.z += bb;
    cb += -bb * xd;
    xb += -bb * c * d.x;
    yb += -bb * c * d.y;
    in.d.x += -bb * c * x;
    in.d.y += -bb * c * y;
    in.p.x += xb;
    in.p.y += yb;
    s0b += xb * d.x + yb * d.y;
    in.d.x += xb * s0;
    in.d.y += yb * s0;
    const double qq = s0b * rdz;
    tb += qq;
    in.p.z += -qq;
    in.d.z += -qq * s0;
    cb_out += cb;
    if constexpr (kIndex) {
        // mu = n_s / n_after: d(OPL)/d mu through the refraction, then the chain to both indices;
        // the segment in medium n_s has length s0 + s
        const double sin2 = 1.0 - cos_i * cos_i;
        const double mub = dot(d1b, d) - ab * cos_i - kb * 2.0 * mu * sin2;
        const double rna = rev_div(mu, n_s);
        *nbb += mub * rna + (s0 + s);
        *nab += -mub * mu * rna;
    }
    adj = in;
    return prev;
}

} // namespace mkod

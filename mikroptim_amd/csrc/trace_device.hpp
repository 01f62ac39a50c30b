// Per-ray device math of the trace and of its reverse (adjoint) sweep, templated on the number
// type R (double for every forward trace; double or float for the reverse sweep).
//
// Formulas and branch rules follow the reference's shared math core:
//   surface steps        include/mikroptim/trace/surface_ops.hpp:122-203
//   surface propagation  include/mikroptim/trace/trace.hpp:175-229 (stop/plane/sphere only)
//   reference spheres    include/mikroptim/trace/reference_sphere.hpp:56-143
//   reverse steps        include/mikroptim/trace/adjoint_backward.hpp:115-241, 373-464
// The device code is compiled with FMA contraction (hipcc default); results therefore differ
// from the FP64 CPU oracle by rounding only, which the parity tests bound per ray.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mikroptim_types.h"

namespace mkod {

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
template <class R>
__device__ __forceinline__ R rnorm(V3<R> a) {
    return R(1) / dsqrt(dot(a, a));
}

constexpr uint8_t kBranchRegular = 0, kBranchSoftMissed = 1, kBranchSoftTIR = 2;

__device__ __forceinline__ bool in_piston(uint8_t s) {
    return s == MKO_STATUS_OK || s == MKO_STATUS_SOFT_MISSED || s == MKO_STATUS_SOFT_TIR;
}
__device__ __forceinline__ bool is_soft(uint8_t s) {
    return s == MKO_STATUS_SOFT_MISSED || s == MKO_STATUS_SOFT_TIR;
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

struct Ray {
    V3<double> p, d;
    double opl;
};

// One surface step (propagate_surface_local with transfer, sphere intersection and refraction).
// c = effective curvature (0 for stop and plane), t = vertex distance from the previous surface,
// nb = medium before, mu = nb / n_after. Returns a TraceStatus (Ok also for soft continuations,
// whose kind goes to `soft` and whose branch bits go to `branch`).
template <bool kSoft>
__device__ __forceinline__ uint8_t surface_step(Ray& r, double c, double t, double nb, double mu,
                                                Soft& soft, uint8_t& branch) {
    if (!(r.d.z > 0.0)) return MKO_STATUS_BACKWARD;
    // transfer_to_tangent_plane
    r.p.z -= t;
    const double s0 = -r.p.z / r.d.z;
    r.p.x += s0 * r.d.x;
    r.p.y += s0 * r.d.y;
    r.p.z = 0.0;
    // intersect_sphere
    const double H = c * (r.p.x * r.p.x + r.p.y * r.p.y);
    const double B = r.d.z - c * (r.p.x * r.d.x + r.p.y * r.d.y);
    const double D = B * B - c * H;
    double s;
    bool missed = false;
    if (D < 0.0) {
        if (!kSoft || (!(c > 0.0) && !(c < 0.0))) return MKO_STATUS_MISSED;
        soft.add(-D, MKO_STATUS_SOFT_MISSED);
        s = B / c;
        missed = true;
    } else {
        const double denom = B + dsqrt(D);
        if (!(denom > 0.0)) return MKO_STATUS_MISSED;
        s = H / denom;
    }
    r.p = r.p + s * r.d;
    r.opl += nb * (s0 + s);
    // refract_sphere
    V3<double> N{-c * r.p.x, -c * r.p.y, 1.0 - c * r.p.z};
    if (kSoft && missed) N = rnorm(N) * N;
    const double cosI = dot(r.d, N);
    const double k = 1.0 - mu * mu * (1.0 - cosI * cosI);
    if (k < 0.0) {
        if (!kSoft) return MKO_STATUS_TIR;
        soft.add(-k, MKO_STATUS_SOFT_TIR);
        const V3<double> tau = r.d - cosI * N;
        r.d = rnorm(tau) * tau;
        branch = uint8_t((missed ? kBranchSoftMissed : 0) | kBranchSoftTIR);
        return MKO_STATUS_OK;
    }
    r.d = mu * r.d + (dsqrt(k) - mu * cosI) * N;
    branch = missed ? kBranchSoftMissed : kBranchRegular;
    return MKO_STATUS_OK;
}

// reference_sphere_step (kind last_surface) / image_sphere_step (image kinds).
template <bool kSoft>
__device__ __forceinline__ uint8_t reference_step(Ray& r, V3<double> M, double kappa, int kind,
                                                  double n_image, Soft& soft, uint8_t& branch) {
    const V3<double> q = r.p - M;
    const double b = -dot(q, r.d);
    branch = kBranchRegular;
    if (kind == MKO_REF_LAST_SURFACE) {
        const double g = dot(r.p, r.p - 2.0 * M);
        const double disc = b * b - g;
        if (disc < 0.0) {
            const double radius_sq = dot(M, M);
            if (!kSoft || !(radius_sq > 0.0)) return MKO_STATUS_MISSED;
            soft.add(-disc / radius_sq, MKO_STATUS_SOFT_MISSED);
            r.p = r.p + b * r.d;
            r.opl += n_image * b;
            branch = kBranchSoftMissed;
            return MKO_STATUS_SOFT_MISSED;
        }
        const double root = dsqrt(disc);
        const bool real_image = M.z >= 0.0;
        const double denom = real_image ? b + root : b - root;
        if (!(real_image ? denom > 0.0 : denom < 0.0)) return MKO_STATUS_MISSED;
        const double s = g / denom;
        r.p = r.p + s * r.d;
        r.opl += n_image * s;
        return MKO_STATUS_OK;
    }
    const V3<double> cq = cross(q, r.d);
    const double t = -dot(cq, cq);
    const double disc = 1.0 + kappa * kappa * t;
    double step;
    uint8_t st = MKO_STATUS_OK;
    if (disc >= 0.0) {
        step = b - kappa * t / (dsqrt(disc) + 1.0);
    } else if (kSoft && disc < 0.0) {
        step = b - kappa * t;
        st = MKO_STATUS_SOFT_MISSED;
    } else {
        return MKO_STATUS_MISSED;
    }
    if (!(step >= 0.0) && !(step < 0.0)) return MKO_STATUS_MISSED;
    if (st == MKO_STATUS_SOFT_MISSED) {
        soft.add(-disc, MKO_STATUS_SOFT_MISSED);
        branch = kBranchSoftMissed;
    }
    r.p = r.p + b * r.d;
    r.opl += n_image * step;
    return st;
}

// ---- reverse sweep ------------------------------------------------------------------------------

template <class R>
struct RayAdj {
    V3<R> p, d;
};

// Reverse of one stop/plane/sphere step (refracting_step_backward). (p, d): state before the
// surface; adj: adjoint of the state after it on entry, of the state before it on return.
// Adds the adjoints of the vertex distance t, the curvature c and the indices to tb/cb/nbb/nab.
template <class R>
__device__ __forceinline__ void refracting_step_backward(V3<R> p, V3<R> d, R t, R c, R n_before,
                                                         R n_after, uint8_t branch, R seed_opl,
                                                         R seed_pen, RayAdj<R>& adj, R& tb,
                                                         R& cb_out, R& nbb, R& nab) {
    const R z = p.z - t;
    const R s0 = -z / d.z;
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
        root_d = dsqrt(D);
        s = H / (B + root_d);
    }
    const V3<R> q{x + s * d.x, y + s * d.y, s * d.z};
    const V3<R> n_raw{-c * q.x, -c * q.y, R(1) - c * q.z};
    R inv_norm = R(1);
    V3<R> N = n_raw;
    if (missed) {
        inv_norm = rnorm(n_raw);
        N = inv_norm * n_raw;
    }
    const R mu = n_before / n_after;
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
        const R root_k = dsqrt(k);
        const R a = root_k - mu * cos_i;
        in.d += mu * d1b;
        nb += a * d1b;
        const R ab = dot(d1b, N);
        mub += dot(d1b, d);
        kb += ab / (R(2) * root_k);
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
        const R inv = R(1) / (R(2) * root_d);
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
    const R qq = s0b / d.z;
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

} // namespace mkod

// Constraint rows of one variant on the device: the paraxial y-nu trace at lambda_ref and the
// paraxial / mechanical row formulas (ExtraResiduals / TraceExtraResiduals of the reference).
//   paraxial trace   compute_paraxial, src/paraxial/paraxial.cpp:98-175;
//                    paraxial_trace / petzval_sum, include/mikroptim/paraxial/ynu_trace.hpp:53-93
//   paraxial rows    paraxial_deviation / paraxial_residual, src/optim/paraxial_residuals.cpp:257-339
//   real rays        batch_sine_condition, src/optim/paraxial_residuals.cpp:68-91, conrady_osc,
//                    include/mikroptim/paraxial/sine_condition.hpp:25-40; batch_chief_ray_height,
//                    src/batch/system_batch.cpp:999-1030 (collimated_ray / trace_ray at lambda_ref,
//                    include/mikroptim/trace/trace.hpp:110, surface_ops.hpp:122-203)
//   mechanical rows  sag / mechanical_value / mechanical_residual,
//                    src/optim/mechanical_constraints.cpp:58-281
// Every product, sum, quotient and root is rounded on its own (contraction off, correctly rounded
// division and square root), in the reference's operation order, so a row is bitwise the value of
// the reference's host code for the same variant and extrema.
#pragma once

#include <hip/hip_runtime.h>

#include "mikroptim_types.h"

namespace mkoc {

struct Paraxial {
    double A, C, power, petzval_curvature;
    // compute_paraxial, src/paraxial/paraxial.cpp:145-190 (the sine condition's quantities)
    bool afocal, exit_pupil_at_infinity;
    double focal_length, back_focal_distance, exit_pupil_position;
};

__device__ inline bool needs_extrema(const mko_constraint& c) {
    return c.type == MKO_CON_MAX_SEMI_DIAMETER || c.type == MKO_CON_MAX_INCIDENCE_ANGLE ||
           (c.type == MKO_CON_MIN_EDGE_THICKNESS && c.edge_height == MKO_EDGE_RAYS);
}
__device__ inline bool needs_paraxial(const mko_constraint& c) {
    return c.type == MKO_CON_FOCAL_LENGTH || c.type == MKO_CON_PARAXIAL_FOCUS ||
           c.type == MKO_CON_PETZVAL || c.type == MKO_CON_SINE_CONDITION;
}
__device__ inline bool is_paraxial(int type) {
    return type <= MKO_CON_PETZVAL || type == MKO_CON_SINE_CONDITION ||
           type == MKO_CON_DISTORTION;
}

// Marginal (y = 1, u = 0) and chief (y = 0, u = 1) rays through the S surfaces; A = y_L and
// C = n'_L u'_L of the marginal ray, power = -C (0 when afocal), Petzval curvature -n'_L P;
// focal length 1 / phi, back focal distance -y_L / u'_L, exit pupil at -y_chief / u'_chief
// (at infinity when D = n'_L u'_chief = 0).
__device__ inline Paraxial paraxial(const double* c, const double* t, const double* n, int S) {
#pragma clang fp contract(off)
    double y = 1.0, u = 0.0, cy = 0.0, cu = 1.0, nb = 1.0, ps = 0.0;
    for (int j = 0; j < S; ++j) {
        if (j > 0) {
            y = y + t[j - 1] * u;
            cy = cy + t[j - 1] * cu;
        }
        const double phi = c[j] * (n[j] - nb);
        u = __ddiv_rn(nb * u - y * phi, n[j]);
        cu = __ddiv_rn(nb * cu - cy * phi, n[j]);
        ps = ps + __ddiv_rn(phi, nb * n[j]);
        nb = n[j];
    }
    const double n_image = n[S - 1];
    Paraxial p;
    p.A = y;
    p.C = n_image * u;
    p.power = p.C == 0.0 ? 0.0 : -p.C;
    p.petzval_curvature = -n_image * ps;
    p.afocal = p.C == 0.0;
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    p.focal_length = p.afocal ? __longlong_as_double(0x7ff0000000000000LL) : __ddiv_rn(1.0, p.power);
    p.back_focal_distance = p.afocal ? nan : __ddiv_rn(-y, u);
    const double D = n_image * cu;
    p.exit_pupil_at_infinity = D == 0.0;
    p.exit_pupil_position = p.exit_pupil_at_infinity ? nan : __ddiv_rn(-cy, cu);
    return p;
}

// A collimated ray (pupil radius a, field angle theta, pupil point (rho, phi = 0)) traced at
// lambda_ref through surfaces 1 .. S-1 without soft failures; false where it fails (backward,
// missed, total internal reflection). p, d: after the last surface, in its local coordinates.
__device__ inline bool real_ray(const double* c, const double* t, const double* n,
                                uint64_t sphere_mask, int S, double a, double theta, double rho,
                                double p[3], double d[3]) {
#pragma clang fp contract(off)
    const double r = a * rho;
    p[0] = r * 0.0;
    p[1] = r * 1.0;
    p[2] = 0.0;
    d[0] = 0.0;
    if (theta == 0.0) { // sin(+-0) = +-0, cos(0) = 1 exactly
        d[1] = theta;
        d[2] = 1.0;
    } else {
        sincos(theta, &d[1], &d[2]);
    }
    for (int s = 1; s < S; ++s) {
        const double curv = ((sphere_mask >> s) & 1ull) ? c[s] : 0.0;
        if (!(d[2] > 0.0)) return false;
        // transfer to the tangent plane
        p[2] = p[2] - t[s - 1];
        const double s0 = __ddiv_rn(-p[2], d[2]);
        p[0] = p[0] + s0 * d[0];
        p[1] = p[1] + s0 * d[1];
        p[2] = p[2] + s0 * d[2];
        p[2] = 0.0;
        // intersection with the sphere
        const double H = curv * (p[0] * p[0] + p[1] * p[1]);
        const double B = d[2] - curv * (p[0] * d[0] + p[1] * d[1]);
        const double Dk = B * B - curv * H;
        if (Dk < 0.0) return false;
        const double denom = B + __dsqrt_rn(Dk);
        if (!(denom > 0.0)) return false;
        const double si = __ddiv_rn(H, denom);
        p[0] = p[0] + si * d[0];
        p[1] = p[1] + si * d[1];
        p[2] = p[2] + si * d[2];
        // refraction
        const double mu = __ddiv_rn(n[s - 1], n[s]);
        const double N0 = -curv * p[0], N1 = -curv * p[1], N2 = 1.0 - curv * p[2];
        const double cosI = d[0] * N0 + d[1] * N1 + d[2] * N2;
        const double k = 1.0 - mu * mu * (1.0 - cosI * cosI);
        if (k < 0.0) return false;
        const double g = __dsqrt_rn(k) - mu * cosI;
        d[0] = mu * d[0] + g * N0;
        d[1] = mu * d[1] + g * N1;
        d[2] = mu * d[2] + g * N2;
    }
    return true;
}

__device__ inline double sag(double curvature, double height) {
#pragma clang fp contract(off)
    const double h2 = height * height;
    const double a = fabs(curvature * height);
    if (!(a < 1.0)) return curvature * h2;
    return __ddiv_rn(curvature * h2, 1.0 + __dsqrt_rn((1.0 - a) * (1.0 + a)));
}

// Variant view for the rows: curvature / thickness / paraxial index [S], sphere mask (effective
// curvature), pupil radius; ext = the variant's extrema [S][2] (traced rows only).
struct VariantRows {
    const double* c;
    const double* t;
    const double* n;
    uint64_t sphere_mask;
    double pupil_radius;
    int S;
    const double* field_angle;  // [F]
    const double* field_target; // [F]
    int F;
};

// OSC of the zone's meridional ray (conrady_osc): NaN where undefined.
__device__ inline double sine_condition(const VariantRows& v, const Paraxial& pd, double rho) {
#pragma clang fp contract(off)
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    double p[3], d[3];
    if (!real_ray(v.c, v.t, v.n, v.sphere_mask, v.S, v.pupil_radius, 0.0, rho, p, d)) return nan;
    const double height = rho * v.pupil_radius;
    if (d[1] == 0.0 || pd.afocal || !isfinite(pd.focal_length)) return nan;
    const double f_image = v.n[v.S - 1] * pd.focal_length;
    const double f_real = __ddiv_rn(-height, d[1]);
    double factor = 1.0;
    if (!pd.exit_pupil_at_infinity) {
        const double L = p[2] - __ddiv_rn(p[1] * d[2], d[1]);
        factor = __ddiv_rn(pd.back_focal_distance - pd.exit_pupil_position,
                           L - pd.exit_pupil_position);
    }
    return __ddiv_rn(f_real, f_image) * factor - 1.0;
}

// Height of field f's real chief ray in the image plane (batch_chief_ray_height): NaN where the
// ray fails or does not travel forward.
__device__ inline double chief_ray_height(const VariantRows& v, int f) {
#pragma clang fp contract(off)
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    double p[3], d[3];
    if (!real_ray(v.c, v.t, v.n, v.sphere_mask, v.S, v.pupil_radius, v.field_angle[f], 0.0, p,
                  d) ||
        !(d[2] > 0.0))
        return nan;
    const double w = v.t[v.S - 1];
    return p[1] + d[1] * __ddiv_rn(w - p[2], d[2]);
}

__device__ inline double row(const mko_constraint& k, const VariantRows& v, const Paraxial& p,
                             const double* ext) {
#pragma clang fp contract(off)
    const int S = v.S;
    if (is_paraxial(k.type)) {
        double q = 0.0;
        switch (k.type) {
        case MKO_CON_SINE_CONDITION: {
            const double osc = sine_condition(v, p, k.zone);
            q = isfinite(osc) ? osc : 0.0;
            break;
        }
        case MKO_CON_DISTORTION: {
            q = -k.target;
            if (k.field < v.F) {
                const double h = v.field_target[k.field];
                if (h != 0.0) {
                    const double y = chief_ray_height(v, k.field);
                    if (isfinite(y)) q = fabs(__ddiv_rn(y - h, h)) - k.target;
                }
            }
            break;
        }
        case MKO_CON_FOCAL_LENGTH: q = k.target * (1.0 - k.target * p.power); break;
        case MKO_CON_WORKING_DISTANCE: q = v.t[S - 1] - k.target; break;
        case MKO_CON_PARFOCAL_LENGTH: {
            double length = k.offset;
            for (int j = k.surface; j < S; ++j) length = length + v.t[j];
            q = length - k.target;
            break;
        }
        case MKO_CON_PARAXIAL_FOCUS:
            q = v.pupil_radius * (p.A + __ddiv_rn(v.t[S - 1] * p.C, v.n[S - 1]));
            break;
        case MKO_CON_PETZVAL: q = fabs(p.petzval_curvature) - k.target; break;
        default: break;
        }
        if (k.shift != 0.0) {
            if (k.mode == MKO_MODE_MAX) q = q + k.shift;
            else if (k.mode == MKO_MODE_MIN) q = q - k.shift;
        }
        const double m = k.mode == MKO_MODE_MIN ? (0.0 < q ? 0.0 : q)
                         : k.mode == MKO_MODE_MAX ? (q < 0.0 ? 0.0 : q)
                                                  : q;
        return __dsqrt_rn(k.weight) * m;
    }
    auto center = [&] {
        double sum = 0.0;
        for (int j = k.surface; j < k.next_surface; ++j) sum = sum + v.t[j];
        return sum;
    };
    auto curv = [&](int j) { return ((v.sphere_mask >> j) & 1ull) ? v.c[j] : 0.0; };
    auto larger = [](double a, double b) { return a < b ? b : a; }; // std::max
    double value = 0.0;
    switch (k.type) {
    case MKO_CON_MAX_SEMI_DIAMETER: value = ext[2 * k.surface] + k.offset; break;
    case MKO_CON_MIN_CENTER_THICKNESS:
    case MKO_CON_MAX_CENTER_THICKNESS: value = center(); break;
    case MKO_CON_MIN_EDGE_THICKNESS: {
        const double h = (k.edge_height == MKO_EDGE_RAYS
                              ? larger(ext[2 * k.surface], ext[2 * k.next_surface])
                              : larger(k.semi_diameter, k.next_semi_diameter)) +
                         k.offset;
        value = center() - sag(curv(k.surface), h) + sag(curv(k.next_surface), h);
        break;
    }
    default: value = ext[2 * k.surface + 1]; break;
    }
    double g = (k.type == MKO_CON_MAX_SEMI_DIAMETER || k.type == MKO_CON_MAX_CENTER_THICKNESS)
                   ? value - k.target
                   : k.target - value;
    if (k.shift != 0.0) g = g + k.shift;
    return __dsqrt_rn(k.weight) * (g < 0.0 ? 0.0 : g);
}

} // namespace mkoc

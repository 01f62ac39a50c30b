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
//   glass-hull rows  glass_hull_residual, src/optim/glass_variables.cpp:135-155;
//                    project_onto_hull, src/glass/convex_hull.cpp:397-452 (the closest point of a
//                    triangle by Voronoi regions: Ericson, Real-Time Collision Detection, 5.1.5)
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
    const double* buchdahl;     // [S][4] (n0, nu1, nu2, nu3), or NULL
    const double* glass_hull;   // mko_extra_rows.glass_hull (device copy), or NULL
};

// ---- glass hull: the closest point of the hull to u and the penalty row ------------------------

__device__ inline double hull_dot(const double* a, const double* b) {
#pragma clang fp contract(off)
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ inline double hull_length(const double* p, const double* q) {
#pragma clang fp contract(off)
    const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    return __dsqrt_rn(hull_dot(d, d));
}
// closest point of the segment a-b to p (a for a degenerate segment)
__device__ inline void hull_on_segment(const double* p, const double* a, const double* b,
                                       double* q) {
#pragma clang fp contract(off)
    const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double len2 = hull_dot(ab, ab);
    if (!(len2 > 0.0)) {
        q[0] = a[0]; q[1] = a[1]; q[2] = a[2];
        return;
    }
    const double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double t = __ddiv_rn(hull_dot(ap, ab), len2);
    t = t < 0.0 ? 0.0 : (1.0 < t ? 1.0 : t);
    for (int j = 0; j < 3; ++j) q[j] = a[j] + ab[j] * t;
}
// closest point of the triangle abc to p: the Voronoi regions of the corners, the edges and the
// face; a triangle of zero area falls back to the nearest of its three edges
__device__ inline void hull_on_triangle(const double* p, const double* a, const double* b,
                                        const double* c, double* q) {
#pragma clang fp contract(off)
    const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    const double d1 = hull_dot(ab, ap), d2 = hull_dot(ac, ap);
    auto set = [q](const double* x) { q[0] = x[0]; q[1] = x[1]; q[2] = x[2]; };
    if (d1 <= 0.0 && d2 <= 0.0) return set(a);
    const double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    const double d3 = hull_dot(ab, bp), d4 = hull_dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return set(b);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double t = __ddiv_rn(d1, d1 - d3);
        for (int j = 0; j < 3; ++j) q[j] = a[j] + ab[j] * t;
        return;
    }
    const double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    const double d5 = hull_dot(ab, cp), d6 = hull_dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return set(c);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double t = __ddiv_rn(d2, d2 - d6);
        for (int j = 0; j < 3; ++j) q[j] = a[j] + ac[j] * t;
        return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double t = __ddiv_rn(d4 - d3, (d4 - d3) + (d5 - d6));
        for (int j = 0; j < 3; ++j) q[j] = b[j] + (c[j] - b[j]) * t;
        return;
    }
    const double sum = va + vb + vc;
    if (!(sum > 0.0)) {
        double best[3], r[3];
        hull_on_segment(p, a, b, best);
        double bd = hull_length(p, best);
        hull_on_segment(p, b, c, r);
        double d = hull_length(p, r);
        if (d < bd) {
            bd = d;
            best[0] = r[0]; best[1] = r[1]; best[2] = r[2];
        }
        hull_on_segment(p, a, c, r);
        d = hull_length(p, r);
        if (d < bd) {
            best[0] = r[0]; best[1] = r[1]; best[2] = r[2];
        }
        return set(best);
    }
    const double v = __ddiv_rn(vb, sum), w = __ddiv_rn(vc, sum);
    for (int j = 0; j < 3; ++j) q[j] = a[j] + (ab[j] * v + ac[j] * w);
}

// Component `comp` of sqrt(w) e grad for the Buchdahl coordinates x = (n0, nu1, nu2): u = (x - m)
// / s; a 3-d hull with u behind every face plane gives phi = max_f (n_f . u - o_f) + margin and
// grad = that face's normal; otherwise the nearest of the records' closest points gives the
// distance d (phi = d, + margin for the 3-d hull) and grad = (u - closest) / d.
__device__ inline double glass_hull_row(const double* hull, const double* x, double weight,
                                        double margin, int comp) {
#pragma clang fp contract(off)
    const int dim = int(hull[0]), n = int(hull[1]);
    const double* rec = hull + 8;
    double u[3];
    for (int j = 0; j < 3; ++j) u[j] = __ddiv_rn(x[j] - hull[2 + j], hull[5 + j]);
    double grad[3] = {0.0, 0.0, 0.0};
    double phi;
    bool inside_3d = false;
    if (dim == 3) {
        double worst = -__builtin_inf();
        int wf = 0;
        for (int f = 0; f < n; ++f) {
            const double sd = hull_dot(rec + 13 * f + 9, u) - rec[13 * f + 12];
            if (sd > worst) {
                worst = sd;
                wf = f;
            }
        }
        if (worst <= 0.0) {
            inside_3d = true;
            phi = worst + margin;
            for (int j = 0; j < 3; ++j) grad[j] = rec[13 * wf + 9 + j];
        }
    }
    if (!inside_3d) {
        double best = __builtin_inf(), closest[3] = {0.0, 0.0, 0.0}, q[3];
        const int m = dim >= 2 ? n : 1;
        for (int f = 0; f < m; ++f) {
            const double* r = rec + 13 * f;
            if (dim >= 2) hull_on_triangle(u, r, r + 3, r + 6, q);
            else if (dim == 1) hull_on_segment(u, r, r + 3, q);
            else {
                q[0] = r[0]; q[1] = r[1]; q[2] = r[2];
            }
            const double d = hull_length(u, q);
            if (d < best) {
                best = d;
                closest[0] = q[0]; closest[1] = q[1]; closest[2] = q[2];
            }
        }
        if (best > 0.0) {
            const double inv = __ddiv_rn(1.0, best);
            for (int j = 0; j < 3; ++j) grad[j] = (u[j] - closest[j]) * inv;
        }
        phi = dim == 3 ? best + margin : best;
    }
    const double e = 0.0 < phi ? phi : 0.0;
    if (!(e > 0.0)) return 0.0;
    return (__dsqrt_rn(weight) * e) * grad[comp];
}

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
    if (k.type == MKO_CON_GLASS_HULL)
        return glass_hull_row(v.glass_hull, v.buchdahl + 4 * k.surface, k.weight, k.target,
                              k.field);
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

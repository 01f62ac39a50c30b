// Constraint rows of one variant on the device: the paraxial y-nu trace at lambda_ref and the
// paraxial / mechanical row formulas (ExtraResiduals / TraceExtraResiduals of the reference).
//   paraxial trace   compute_paraxial, src/paraxial/paraxial.cpp:98-175;
//                    paraxial_trace / petzval_sum, include/mikroptim/paraxial/ynu_trace.hpp:53-93
//   paraxial rows    paraxial_deviation / paraxial_residual, src/optim/paraxial_residuals.cpp:257-339
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
};

__device__ inline bool needs_extrema(const mko_constraint& c) {
    return c.type == MKO_CON_MAX_SEMI_DIAMETER || c.type == MKO_CON_MAX_INCIDENCE_ANGLE ||
           (c.type == MKO_CON_MIN_EDGE_THICKNESS && c.edge_height == MKO_EDGE_RAYS);
}
__device__ inline bool needs_paraxial(const mko_constraint& c) {
    return c.type == MKO_CON_FOCAL_LENGTH || c.type == MKO_CON_PARAXIAL_FOCUS ||
           c.type == MKO_CON_PETZVAL;
}

// Marginal (y = 1, u = 0) and chief (y = 0, u = 1) rays through the S surfaces; A = y_L and
// C = n'_L u'_L of the marginal ray, power = -C (0 when afocal), Petzval curvature -n'_L P.
__device__ inline Paraxial paraxial(const double* c, const double* t, const double* n, int S) {
#pragma clang fp contract(off)
    double y = 1.0, u = 0.0, nb = 1.0, ps = 0.0;
    for (int j = 0; j < S; ++j) {
        if (j > 0) y = y + t[j - 1] * u;
        const double phi = c[j] * (n[j] - nb);
        u = __ddiv_rn(nb * u - y * phi, n[j]);
        ps = ps + __ddiv_rn(phi, nb * n[j]);
        nb = n[j];
    }
    const double n_image = n[S - 1];
    Paraxial p;
    p.A = y;
    p.C = n_image * u;
    p.power = p.C == 0.0 ? 0.0 : -p.C;
    p.petzval_curvature = -n_image * ps;
    return p;
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
};

__device__ inline double row(const mko_constraint& k, const VariantRows& v, const Paraxial& p,
                             const double* ext) {
#pragma clang fp contract(off)
    const int S = v.S;
    if (k.type <= MKO_CON_PETZVAL) {
        double q = 0.0;
        switch (k.type) {
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
        default: q = fabs(p.petzval_curvature) - k.target; break;
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

// CPU FP64 oracle of the batched forward-and-Jacobian path. TEST INFRASTRUCTURE ONLY: this
// library is the checker used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg; the product (mikroptim_amd, libmikroptim_hip.so) never links, loads or calls it.
//
// A from-scratch restatement of the reference's CPU algorithms in plain C++20, compiled with
// -ffp-contract=off like the reference (CMakeLists.txt:27-28) so that operation order fixes every
// rounding. Each function names the reference file:line whose formulas and summation order it
// follows. Scope: coaxial systems of stop/plane/sphere surfaces (no ghost plates, no
// decenter/tilt), collimated and point sources, the three reference-sphere kinds, soft and hard
// ray failure, curvature/thickness/working-distance variables.
//
// How each part is pinned (tests/test_oracle_*.py): bitwise against the reference's own CPU
// code built from its sources into oracle/_ref (residuals_cpu, merit_cpu, fd_jacobian_cpu,
// adjoint_jacobian_cpu, normal_equations_host, levenberg_marquardt, perturbed_vector, the
// Gauss-Legendre table); against the reference's golden vectors (mt19937_64/splitmix64 values and
// perturbed vectors of tests/test_variant_generator.cpp:72-101, kFdStepFactor); and against closed
// forms (plane-parallel plate, Snell's law, aplanatic points, exact spherical waves, GL moments,
// merit = weighted sum of wavefront RMS^2, adjoint vs Richardson-extrapolated differences, piston
// invariant, LM on a linear problem).

#include "mikroptim_types.h"
#include "gauss_legendre_table.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace orc {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

struct Vec3 {
    double x, y, z;
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(double s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline Vec3& operator+=(Vec3& a, Vec3 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(Vec3 a) { return std::sqrt(dot(a, a)); }

struct Ray {
    Vec3 p, d;
    double opl;
};

enum Status : uint8_t {
    Ok = MKO_STATUS_OK,
    Missed = MKO_STATUS_MISSED,
    TIR = MKO_STATUS_TIR,
    Backward = MKO_STATUS_BACKWARD,
    Unsupported = MKO_STATUS_UNSUPPORTED,
    SoftMissed = MKO_STATUS_SOFT_MISSED,
    SoftTIR = MKO_STATUS_SOFT_TIR,
};
inline bool is_soft(Status s) { return s == SoftMissed || s == SoftTIR; }
inline bool in_piston(Status s) { return s == Ok || is_soft(s); }
inline bool is_hard(Status s) { return !in_piston(s); }

constexpr uint8_t kBranchRegular = 0, kBranchSoftMissed = 1, kBranchSoftTIR = 2;

// SoftFailure, include/mikroptim/trace/surface_ops.hpp:94
struct Soft {
    double penalty = 0.0;
    Status first = Ok;
    void add(double amount, Status kind) {
        penalty += amount;
        if (first == Ok) first = kind;
    }
};

// ---- surface operations: include/mikroptim/trace/surface_ops.hpp:122-203 --------------------

inline double transfer_to_tangent_plane(Ray& r, double t) {
    r.p.z -= t;
    const double s = -r.p.z / r.d.z;
    r.p += s * r.d;
    r.p.z = 0.0;
    return s;
}

inline Status intersect_sphere(Ray& r, double c, double& s, Soft* soft) {
    const double H = c * (r.p.x * r.p.x + r.p.y * r.p.y);
    const double B = r.d.z - c * (r.p.x * r.d.x + r.p.y * r.d.y);
    const double D = B * B - c * H;
    if (D < 0.0) {
        if (soft == nullptr || (!(c > 0.0) && !(c < 0.0))) return Missed;
        soft->add(-D, SoftMissed);
        s = B / c;
        r.p += s * r.d;
        return SoftMissed;
    }
    const double denom = B + std::sqrt(D);
    if (!(denom > 0.0)) return Missed;
    s = H / denom;
    r.p += s * r.d;
    return Ok;
}

inline Status refract_sphere(Ray& r, double c, double mu, Soft* soft, bool off_surface) {
    Vec3 N{-c * r.p.x, -c * r.p.y, 1.0 - c * r.p.z};
    if (off_surface) N = (1.0 / norm(N)) * N;
    const double cosI = dot(r.d, N);
    const double k = 1.0 - mu * mu * (1.0 - cosI * cosI);
    if (k < 0.0) {
        if (soft == nullptr) return TIR;
        soft->add(-k, SoftTIR);
        const Vec3 t = r.d - cosI * N;
        r.d = (1.0 / norm(t)) * t;
        return SoftTIR;
    }
    r.d = mu * r.d + (std::sqrt(k) - mu * cosI) * N;
    return Ok;
}

// propagate_surface_local, include/mikroptim/trace/trace.hpp:175-229 (no ghost plates)
inline Status propagate_surface(Ray& ray, double t, uint8_t kind, double c, double n_before,
                                double n_after, Soft* soft, uint8_t* branch) {
    double curv;
    switch (kind) {
    case MKO_SURFACE_STOP:
    case MKO_SURFACE_PLANE: curv = 0.0; break;
    case MKO_SURFACE_SPHERE: curv = c; break;
    default: return Unsupported;
    }
    if (!(ray.d.z > 0.0)) return Backward;
    const double s0 = transfer_to_tangent_plane(ray, t);
    double s = 0.0;
    const Status hit = intersect_sphere(ray, curv, s, soft);
    if (hit != Ok && hit != SoftMissed) return hit;
    ray.opl += n_before * (s0 + s);
    const Status refr = refract_sphere(ray, curv, n_before / n_after, soft, hit == SoftMissed);
    if (branch) {
        *branch = static_cast<uint8_t>((hit == SoftMissed ? kBranchSoftMissed : kBranchRegular) |
                                       (refr == SoftTIR ? kBranchSoftTIR : kBranchRegular));
    }
    return refr == SoftTIR ? Ok : refr;
}

// ---- one variant's surface table -------------------------------------------------------------

struct Table {
    int S, L;
    const uint8_t* type;
    const double* c;
    const double* t;
    const double* n; // [S][L]
};

// AdjointRecord, include/mikroptim/trace/adjoint_record.hpp:66
struct Record {
    int count = 0;
    Vec3 p[MKO_MAX_SURFACES];
    Vec3 d[MKO_MAX_SURFACES];
    uint8_t branch[MKO_MAX_SURFACES];
    double image_s = 0.0;
    uint8_t image_branch = 0;
};

// trace_segment / trace_ray_recorded, include/mikroptim/trace/trace.hpp:277-358
inline Status trace_ray(const Table& tb, int l, Ray& ray, Soft* soft, Record* rec, int* failed) {
    if (rec) {
        rec->count = 1;
        rec->p[0] = ray.p;
        rec->d[0] = ray.d;
        rec->branch[0] = kBranchRegular;
        rec->image_s = 0.0;
        rec->image_branch = kBranchRegular;
    }
    int soft_surface = -1;
    for (int s = 1; s < tb.S; ++s) {
        const double nb = tb.n[(s - 1) * tb.L + l];
        const double na = tb.n[s * tb.L + l];
        uint8_t branch = kBranchRegular;
        const Status st =
            propagate_surface(ray, tb.t[s - 1], tb.type[s], tb.c[s], nb, na, soft, &branch);
        if (st != Ok) {
            if (failed) *failed = s;
            return st;
        }
        if (soft && soft_surface < 0 && soft->first != Ok) soft_surface = s;
        if (rec) {
            rec->count = s + 1;
            rec->p[s] = ray.p;
            rec->d[s] = ray.d;
            rec->branch[s] = branch;
        }
    }
    if (failed) *failed = soft_surface;
    return soft ? soft->first : Ok;
}

struct RefSphere {
    Vec3 center;
    double kappa;
    int kind;
};

// reference_sphere_step, include/mikroptim/trace/reference_sphere.hpp:56-90
inline Status reference_sphere_step(Ray& ray, Vec3 center, double n_image, double* s_out,
                                    Soft* soft) {
    const Vec3 q = ray.p - center;
    const double b = -dot(q, ray.d);
    const double g = dot(ray.p, ray.p - 2.0 * center);
    const double disc = b * b - g;
    if (disc < 0.0) {
        const double radius_sq = dot(center, center);
        if (soft == nullptr || !(radius_sq > 0.0)) return Missed;
        soft->add(-disc / radius_sq, SoftMissed);
        ray.p += b * ray.d;
        ray.opl += n_image * b;
        if (s_out) *s_out = b;
        return SoftMissed;
    }
    const double root = std::sqrt(disc);
    const bool real_image = center.z >= 0.0;
    const double denom = real_image ? b + root : b - root;
    if (!(real_image ? denom > 0.0 : denom < 0.0)) return Missed;
    const double s = g / denom;
    ray.p += s * ray.d;
    ray.opl += n_image * s;
    if (s_out) *s_out = s;
    return Ok;
}

// image_sphere_step, include/mikroptim/trace/reference_sphere.hpp:110-143
inline Status image_sphere_step(Ray& ray, Vec3 center, double kappa, double n_image,
                                double* s_out, Soft* soft) {
    const Vec3 q = ray.p - center;
    const double b = -dot(q, ray.d);
    const Vec3 c{q.y * ray.d.z - q.z * ray.d.y, q.z * ray.d.x - q.x * ray.d.z,
                 q.x * ray.d.y - q.y * ray.d.x};
    const double t = -dot(c, c);
    const double disc = 1.0 + kappa * kappa * t;
    double step;
    Status st = Ok;
    if (disc >= 0.0) {
        step = b - kappa * t / (std::sqrt(disc) + 1.0);
    } else if (soft != nullptr && disc < 0.0) {
        step = b - kappa * t;
        st = SoftMissed;
    } else {
        return Missed;
    }
    if (!(step >= 0.0) && !(step < 0.0)) return Missed;
    if (st == SoftMissed) soft->add(-disc, SoftMissed);
    ray.p += b * ray.d;
    ray.opl += n_image * step;
    if (s_out) *s_out = step;
    return st;
}

// trace_to_reference_sphere, include/mikroptim/trace/reference_sphere.hpp:183-222
inline Status trace_to_reference(const Table& tb, int l, Ray& ray, const RefSphere& ref,
                                 Soft* soft, Record* rec) {
    const Status st = trace_ray(tb, l, ray, soft, rec, nullptr);
    if (is_hard(st)) return st;
    if (!(ray.d.z > 0.0)) return Backward;
    const double n_image = tb.n[(tb.S - 1) * tb.L + l];
    double s_ref = 0.0;
    const Status rst = ref.kind == MKO_REF_LAST_SURFACE
                           ? reference_sphere_step(ray, ref.center, n_image, &s_ref, soft)
                           : image_sphere_step(ray, ref.center, ref.kappa, n_image, &s_ref, soft);
    if (is_hard(rst)) return rst;
    if (rec) {
        rec->image_s = s_ref;
        rec->image_branch = rst == SoftMissed ? kBranchSoftMissed : kBranchRegular;
    }
    if (st != Ok) return st;
    return rst;
}

// ---- pupil quadrature: include/mikroptim/pupil/quadrature.hpp:40-132 ------------------------

struct Quadrature {
    int rings, spokes, mode;
    double u[MKO_MAX_RINGS];
    double w[MKO_MAX_RINGS];
};

inline Quadrature make_quadrature(int rings, int spokes, int mode) {
    if (rings < 1 || rings > MKO_MAX_RINGS) throw std::invalid_argument("rings outside [1, 16]");
    if (spokes < 1) throw std::invalid_argument("spokes < 1");
    if (mode != MKO_PUPIL_HALF && mode != MKO_PUPIL_FULL && mode != MKO_PUPIL_AXIS)
        throw std::invalid_argument("unknown pupil mode");
    Quadrature q{};
    q.rings = rings;
    q.spokes = mode == MKO_PUPIL_AXIS ? 1 : spokes;
    q.mode = mode;
    const int off = rings * (rings - 1) / 2;
    for (int i = 0; i < rings; ++i) {
        q.u[i] = mko_gl_u[off + i];
        q.w[i] = mko_gl_w[off + i];
    }
    return q;
}

struct Sample {
    double rho, phi, weight;
};

inline Sample sample(const Quadrature& q, int index) {
    const int i = index / q.spokes;
    const int j = index - i * q.spokes;
    Sample s;
    s.rho = std::sqrt(q.u[i]);
    s.weight = q.w[i] / double(q.spokes);
    switch (q.mode) {
    case MKO_PUPIL_HALF: s.phi = double(2 * j + 1) * kPi / double(2 * q.spokes); break;
    case MKO_PUPIL_FULL: s.phi = double(2 * j + 1) * kPi / double(q.spokes); break;
    default: s.phi = 0.0; break;
    }
    return s;
}

inline double residual_weight(const Sample& s) { return std::sqrt(s.weight); }

// ---- batch access ------------------------------------------------------------------------------

struct Batch {
    mko_batch b;
    size_t V() const { return size_t(b.variants); }
    int S() const { return b.surfaces; }
    int L() const { return b.wavelengths; }
    int F() const { return b.fields; }
    Table table(int v) const {
        const size_t S_ = size_t(S());
        return Table{S(), L(), b.type, b.curvature + size_t(v) * S_, b.thickness + size_t(v) * S_,
                     b.index + size_t(v) * S_ * size_t(L())};
    }
};

// collimated_ray / point_source_ray, include/mikroptim/trace/trace.hpp:109-131, selected as in
// batch_source_ray, include/mikroptim/batch/system_batch_view.hpp:190-201
inline Ray source_ray(const Batch& B, int v, int f, int l, const Sample& s) {
    Ray ray;
    if (B.b.source == MKO_SOURCE_POINT) {
        const double sin_max =
            B.b.source_na / B.b.index[(size_t(v) * size_t(B.S()) + 0) * size_t(B.L()) + size_t(l)];
        const double sr = sin_max * s.rho;
        ray.p = Vec3{B.b.source_point[0], B.b.source_point[1], B.b.source_point[2]};
        ray.d = Vec3{sr * std::sin(s.phi), sr * std::cos(s.phi), std::sqrt((1.0 - sr) * (1.0 + sr))};
        ray.opl = 0.0;
        return ray;
    }
    const double a = B.b.pupil_radius[v];
    const double theta = B.b.field_angle[size_t(v) * size_t(B.F()) + size_t(f)];
    const double r = a * s.rho;
    ray.p = Vec3{r * std::sin(s.phi), r * std::cos(s.phi), 0.0};
    ray.d = Vec3{0.0, std::sin(theta), std::cos(theta)};
    ray.opl = dot(ray.p, ray.d);
    return ray;
}

// batch_reference, include/mikroptim/merit/ray_residual.hpp:230-252 (ideal image point)
inline RefSphere reference_of(const Batch& B, const Table& tb, int v, int f) {
    const size_t i = size_t(v) * size_t(B.F()) + size_t(f);
    RefSphere r{Vec3{0.0, B.b.field_target[i], tb.t[tb.S - 1]}, 0.0, B.b.reference_sphere};
    if (B.b.reference_sphere != MKO_REF_LAST_SURFACE) r.kappa = B.b.reference_curvature[i];
    return r;
}

// tuple_weight / residual_scale / penalty_scale, include/mikroptim/merit/ray_residual.hpp:301-385
inline double tuple_weight(const Batch& B, int f, int l) {
    const double wf = B.b.field_weight[f];
    const double wl = B.b.wavelength_weight[l];
    if (!(wf > 0.0) || !(wl > 0.0)) return 0.0;
    double sum_f = 0.0;
    for (int i = 0; i < B.F(); ++i) sum_f += B.b.field_weight[i];
    double sum_l = 0.0;
    for (int i = 0; i < B.L(); ++i) sum_l += B.b.wavelength_weight[i];
    return (wf / sum_f) * (wl / sum_l);
}
inline double residual_scale(const Batch& B, int f, int l) {
    const double w = tuple_weight(B, f, l);
    if (!(w > 0.0)) return 0.0;
    const double lambda_ref_mm = B.b.reference_wavelength_nm * 1e-6;
    return std::sqrt(w) / lambda_ref_mm;
}
inline double penalty_scale(const Batch& B, int f, int l, double fail_weight) {
    const double w = tuple_weight(B, f, l);
    if (!(w > 0.0)) return 0.0;
    return std::sqrt(fail_weight * w);
}
inline double failure_penalty(Status st, double soft_penalty) {
    if (st == Ok) return 0.0;
    if (is_soft(st)) return soft_penalty;
    return soft_penalty + 1.0;
}

// ResidualLayout, src/batch/batch_residuals.cpp:25-57
struct Layout {
    int F = 0, L = 0;
    mko_residual_settings settings{};
    std::vector<Quadrature> quad;
    std::vector<int64_t> field_offset;
    std::vector<int> samples;
    int64_t rays = 0;
    // constraint rows in row order: the items without extrema (E), then those with (E_t)
    std::vector<mko_constraint> items;
    int E = 0, Et = 0;
    const double* hull = nullptr; // mko_extra_rows.glass_hull of glass-hull rows
    bool soft() const { return settings.failure == MKO_FAILURE_SOFT; }
    int64_t penalty_rows() const { return soft() ? rays : 0; }
    int64_t trace_rows() const { return rays + penalty_rows(); }
    int64_t rows() const { return trace_rows() + E + Et; }
    int64_t offset(int f, int l) const { return field_offset[f] + int64_t(l) * samples[f]; }
};

void check_constraint(const mko_constraint& c, int k, const Batch& B);
bool con_needs_extrema(const mko_constraint& c);

inline Layout make_layout(const Batch& B, const mko_residual_settings& s) {
    if (s.failure != MKO_FAILURE_SOFT && s.failure != MKO_FAILURE_HARD)
        throw std::invalid_argument("unknown failure mode");
    if (!(s.fail_weight >= 0.0) || !std::isfinite(s.fail_weight))
        throw std::invalid_argument("fail_weight must be finite and >= 0");
    Layout lay;
    lay.F = B.F();
    lay.L = B.L();
    lay.settings = s;
    int64_t offset = 0;
    for (int f = 0; f < lay.F; ++f) {
        const Quadrature q = make_quadrature(s.rings, s.spokes, B.b.pupil_mode[f]);
        const int n = q.rings * q.spokes;
        lay.quad.push_back(q);
        lay.field_offset.push_back(offset);
        lay.samples.push_back(n);
        offset += int64_t(lay.L) * n;
    }
    lay.rays = offset;
    if (s.extra && s.extra->count > 0) {
        if (!s.extra->items) throw std::invalid_argument("constraint rows without items");
        for (int k = 0; k < s.extra->count; ++k) check_constraint(s.extra->items[k], k, B);
        for (int pass = 0; pass < 2; ++pass)
            for (int k = 0; k < s.extra->count; ++k)
                if (con_needs_extrema(s.extra->items[k]) == (pass == 1))
                    lay.items.push_back(s.extra->items[k]);
        for (const mko_constraint& c : lay.items) {
            (con_needs_extrema(c) ? lay.Et : lay.E) += 1;
            if (c.type == MKO_CON_GLASS_HULL) {
                const double* h = s.extra->glass_hull;
                const int n = s.extra->glass_hull_size;
                if (!h || n < 21 || !(h[0] >= 0.0 && h[0] <= 3.0) || 8 + 13 * int(h[1]) != n)
                    throw std::invalid_argument("glass-hull rows need a valid glass hull");
                if (c.target > 0.0 && h[0] != 3.0)
                    throw std::invalid_argument("glass hull: a margin needs a 3-d hull");
                lay.hull = h;
            }
        }
    }
    return lay;
}

// opd_statistics, include/mikroptim/merit/opd.hpp:449-487
inline void opd_statistics(const double* opl, const double* weight, const Status* st, int n,
                           double* opd) {
    double valid = 0.0;
    double sum = 0.0;
    for (int k = 0; k < n; ++k) {
        if (in_piston(st[k])) {
            valid += weight[k];
            sum += weight[k] * opl[k];
        }
    }
    if (!(valid > 0.0)) {
        for (int k = 0; k < n; ++k) opd[k] = 0.0;
        return;
    }
    const double piston = sum / valid;
    for (int k = 0; k < n; ++k) {
        double e = 0.0;
        if (in_piston(st[k])) e = opl[k] - piston;
        opd[k] = e;
    }
}

// variant_residuals, src/batch/batch_residuals.cpp:181-231
void variant_residuals(const Batch& B, int v, const Layout& lay, double* residuals, double* opd,
                       uint8_t* status) {
    const Table tb = B.table(v);
    const size_t rays = size_t(lay.rays);
    std::vector<double> opl, weight, penalty;
    std::vector<Status> st;
    std::vector<Sample> pupil;
    for (int f = 0; f < lay.F; ++f) {
        const Quadrature& q = lay.quad[f];
        const int n = lay.samples[f];
        pupil.resize(n);
        weight.resize(n);
        opl.resize(n);
        penalty.resize(n);
        st.resize(n);
        for (int k = 0; k < n; ++k) {
            pupil[k] = sample(q, k);
            weight[k] = pupil[k].weight;
        }
        const RefSphere ref = reference_of(B, tb, v, f);
        for (int l = 0; l < lay.L; ++l) {
            const size_t o = size_t(lay.offset(f, l));
            for (int k = 0; k < n; ++k) {
                Soft soft;
                Ray ray = source_ray(B, v, f, l, pupil[k]);
                st[k] = trace_to_reference(tb, l, ray, ref, lay.soft() ? &soft : nullptr, nullptr);
                opl[k] = ray.opl;
                penalty[k] = soft.penalty;
            }
            opd_statistics(opl.data(), weight.data(), st.data(), n, opd + o);
            const double scale = residual_scale(B, f, l);
            for (int k = 0; k < n; ++k) {
                residuals[o + k] = scale * residual_weight(pupil[k]) * opd[o + k];
                status[o + k] = st[k];
            }
            if (lay.soft()) {
                const double pscale = penalty_scale(B, f, l, lay.settings.fail_weight);
                for (int k = 0; k < n; ++k) {
                    residuals[rays + o + k] =
                        pscale * residual_weight(pupil[k]) * failure_penalty(st[k], penalty[k]);
                }
            }
        }
    }
}

// variant_merit, src/batch/batch_merit.cpp:62-89
void variant_merit(const Batch& B, int v, const Layout& lay, double& merit, int64_t& failed,
                   int64_t& hard_failed) {
    const size_t R = size_t(lay.trace_rows());
    std::vector<double> r(R, 0.0), opd(size_t(lay.rays));
    std::vector<uint8_t> st(size_t(lay.rays));
    variant_residuals(B, v, lay, r.data(), opd.data(), st.data());
    double sum = 0.0;
    for (int f = 0; f < lay.F; ++f) {
        for (int l = 0; l < lay.L; ++l) {
            const int64_t o = lay.offset(f, l);
            double tuple = 0.0;
            for (int k = 0; k < lay.samples[f]; ++k) {
                const double x = r[size_t(o + k)];
                tuple += x * x;
            }
            if (lay.soft()) {
                double pen = 0.0;
                for (int k = 0; k < lay.samples[f]; ++k) {
                    const double p = r[size_t(lay.rays + o + k)];
                    pen += p * p;
                }
                tuple += pen;
            }
            sum += tuple;
        }
    }
    merit = sum;
    failed = 0;
    hard_failed = 0;
    for (uint8_t s : st) {
        failed += s != Ok ? 1 : 0;
        hard_failed += is_hard(Status(s)) ? 1 : 0;
    }
}

// ---- surface extrema: src/batch/batch_extrema.cpp:17-66, include/mikroptim/trace/
// surface_extrema.hpp:47-110 -----------------------------------------------------------------
// Per table surface s of variant v, over the rays with status Ok (hard trace) of all fields,
// wavelengths and samples: the largest height^2 x^2 + y^2 at the intersection (the start point
// for s = 0) and the smallest cos I = d_before . N, N = (-c x, -c y, 1 - c z), c the effective
// curvature (0 for stop and plane); neutral 0 and 1 without a valid ray; out[2s] = sqrt(max
// height^2), out[2s+1] = min cos I.
inline double extremum_max(double acc, double x) { return x > acc ? x : acc; }
inline double extremum_min(double acc, double x) { return x < acc ? x : acc; }

void variant_extrema(const Batch& B, int v, const Layout& lay, double* out) {
    const Table tb = B.table(v);
    const int S = B.S();
    std::vector<double> h2(size_t(S), 0.0), ci(size_t(S), 1.0);
    Record rec;
    for (int f = 0; f < lay.F; ++f) {
        const Quadrature& q = lay.quad[f];
        const RefSphere ref = reference_of(B, tb, v, f);
        for (int l = 0; l < lay.L; ++l) {
            for (int k = 0; k < lay.samples[f]; ++k) {
                Ray ray = source_ray(B, v, f, l, sample(q, k));
                if (trace_to_reference(tb, l, ray, ref, nullptr, &rec) != Ok) continue;
                for (int j = 0; j < S; ++j) {
                    const Vec3 p = rec.p[j];
                    const double c =
                        j == 0 ? 0.0 : (tb.type[j] == MKO_SURFACE_SPHERE ? tb.c[j] : 0.0);
                    const Vec3 d = j == 0 ? rec.d[0] : rec.d[j - 1];
                    const Vec3 n{-c * p.x, -c * p.y, 1.0 - c * p.z};
                    h2[size_t(j)] = extremum_max(h2[size_t(j)], p.x * p.x + p.y * p.y);
                    ci[size_t(j)] = extremum_min(ci[size_t(j)], dot(d, n));
                }
            }
        }
    }
    for (int j = 0; j < S; ++j) {
        out[2 * j] = std::sqrt(h2[size_t(j)]);
        out[2 * j + 1] = ci[size_t(j)];
    }
}

// ---- paraxial y-nu trace at lambda_ref: compute_paraxial, src/paraxial/paraxial.cpp:98-175;
// paraxial_trace / petzval_sum, include/mikroptim/paraxial/ynu_trace.hpp:53-93 -------------------
struct Paraxial {
    double A = 1.0, B = 0.0, C = 0.0, D = 1.0, power = 0.0, petzval_curvature = 0.0;
    // compute_paraxial (paraxial.cpp:145-190): the quantities of the sine condition
    bool afocal = false, exit_pupil_at_infinity = false;
    double focal_length = 0.0, back_focal_distance = 0.0, exit_pupil_position = 0.0;
};

Paraxial batch_paraxial(const Batch& B, int v) {
    const int S = B.S();
    if (!B.b.paraxial_index) throw std::invalid_argument("paraxial constraint without paraxial_index");
    const double* c = B.b.curvature + size_t(v) * S;
    const double* t = B.b.thickness + size_t(v) * S;
    const double* n = B.b.paraxial_index + size_t(v) * S;
    for (int j = 0; j < S; ++j) {
        if (!std::isfinite(c[j]) || !std::isfinite(t[j]) || !std::isfinite(n[j]))
            throw std::invalid_argument("paraxial trace: value not finite");
        if (!(n[j] > 0.0)) throw std::invalid_argument("paraxial trace: index must be > 0");
    }
    if (!(std::isfinite(B.b.pupil_radius[v]) && B.b.pupil_radius[v] > 0.0))
        throw std::invalid_argument("paraxial trace: stop radius must be > 0");
    const auto trace = [&](double y, double u) {
        double nb = 1.0;
        for (int j = 0; j < S; ++j) {
            if (j > 0) y += t[j - 1] * u;
            const double phi = c[j] * (n[j] - nb);
            u = (nb * u - y * phi) / n[j];
            nb = n[j];
        }
        return std::pair<double, double>{y, u};
    };
    const double n_image = n[S - 1];
    const auto marginal = trace(1.0, 0.0);
    const auto chief = trace(0.0, 1.0);
    Paraxial p;
    p.A = marginal.first;
    p.C = n_image * marginal.second;
    p.B = chief.first;
    p.D = n_image * chief.second;
    p.power = -p.C;
    p.afocal = p.C == 0.0;
    if (p.afocal) {
        p.power = 0.0;
        p.focal_length = std::numeric_limits<double>::infinity();
        p.back_focal_distance = std::numeric_limits<double>::quiet_NaN();
    } else {
        p.focal_length = 1.0 / p.power;
        p.back_focal_distance = -marginal.first / marginal.second;
    }
    p.exit_pupil_at_infinity = p.D == 0.0;
    p.exit_pupil_position = p.exit_pupil_at_infinity ? std::numeric_limits<double>::quiet_NaN()
                                                     : -chief.first / chief.second;
    double ps = 0.0, nb = 1.0;
    for (int j = 0; j < S; ++j) {
        ps += c[j] * (n[j] - nb) / (nb * n[j]);
        nb = n[j];
    }
    p.petzval_curvature = -n_image * ps;
    return p;
}

// ---- constraint rows: paraxial_residual, src/optim/paraxial_residuals.cpp:257-371;
// mechanical_residual, src/optim/mechanical_constraints.cpp:58-281 -----------------------------
inline bool con_paraxial(int type) {
    return (type >= MKO_CON_FOCAL_LENGTH && type <= MKO_CON_PETZVAL) ||
           type == MKO_CON_SINE_CONDITION || type == MKO_CON_DISTORTION;
}
inline bool con_mechanical(int type) {
    return type >= MKO_CON_MAX_SEMI_DIAMETER && type <= MKO_CON_MAX_INCIDENCE_ANGLE;
}
inline bool con_gap(int type) {
    return type == MKO_CON_MIN_CENTER_THICKNESS || type == MKO_CON_MAX_CENTER_THICKNESS ||
           type == MKO_CON_MIN_EDGE_THICKNESS;
}
bool con_needs_extrema(const mko_constraint& c) {
    return c.type == MKO_CON_MAX_SEMI_DIAMETER || c.type == MKO_CON_MAX_INCIDENCE_ANGLE ||
           (c.type == MKO_CON_MIN_EDGE_THICKNESS && c.edge_height == MKO_EDGE_RAYS);
}
inline bool con_needs_paraxial(const mko_constraint& c) {
    return c.type == MKO_CON_FOCAL_LENGTH || c.type == MKO_CON_PARAXIAL_FOCUS ||
           c.type == MKO_CON_PETZVAL || c.type == MKO_CON_SINE_CONDITION ||
           c.type == MKO_CON_DISTORTION;
}

// check_paraxial_constraints (paraxial_residuals.cpp:152-255), check_mechanical_constraints
// (mechanical_constraints.cpp:148-203)
void check_constraint(const mko_constraint& c, int k, const Batch& B) {
    const auto fail = [k](const std::string& what) {
        throw std::invalid_argument("constraint " + std::to_string(k) + ": " + what);
    };
    const int S = B.S();
    if (!(std::isfinite(c.weight) && c.weight >= 0.0)) fail("weight must be finite and >= 0");
    if (!(std::isfinite(c.shift) && c.shift >= 0.0)) fail("shift must be finite and >= 0");
    if (c.type == MKO_CON_GLASS_HULL) {
        // check_glass_hull_penalty, src/optim/glass_variables.cpp:103-133
        if (c.surface < 0 || c.surface >= S) fail("surface out of range");
        if (c.field < 0 || c.field > 2) fail("component must be 0, 1 or 2");
        if (!(std::isfinite(c.target) && c.target >= 0.0)) fail("margin must be finite and >= 0");
        if (c.shift != 0.0) fail("no shift for the glass hull");
        if (!B.b.buchdahl) fail("batch has no Buchdahl coefficients");
        return;
    }
    if (con_paraxial(c.type)) {
        if (c.mode < MKO_MODE_EQUAL || c.mode > MKO_MODE_MAX) fail("unknown mode");
        if (!std::isfinite(c.target)) fail("target must be finite");
        if (c.shift != 0.0 && c.mode == MKO_MODE_EQUAL) fail("shift only for one-sided modes");
        switch (c.type) {
        case MKO_CON_FOCAL_LENGTH:
            if (c.target == 0.0) fail("target focal length must not be 0");
            if (c.mode != MKO_MODE_EQUAL) fail("defined as an equation only");
            break;
        case MKO_CON_PARAXIAL_FOCUS:
            if (c.mode != MKO_MODE_EQUAL) fail("defined as an equation only");
            if (B.b.source != MKO_SOURCE_COLLIMATED) fail("needs a collimated source");
            break;
        case MKO_CON_PARFOCAL_LENGTH:
            if (c.surface < 0 || c.surface >= S) fail("reference surface out of range");
            if (!std::isfinite(c.offset)) fail("offset must be finite");
            break;
        case MKO_CON_PETZVAL:
            if (c.mode != MKO_MODE_MAX) fail("defined as an upper bound only");
            if (c.target < 0.0) fail("kappa_max must be >= 0");
            break;
        case MKO_CON_SINE_CONDITION:
            if (c.mode != MKO_MODE_EQUAL) fail("defined as an equation only");
            if (c.target != 0.0) fail("target must be 0");
            if (!(std::isfinite(c.zone) && c.zone > 0.0 && c.zone <= 1.0))
                fail("pupil zone must lie in (0, 1]");
            if (B.b.source != MKO_SOURCE_COLLIMATED) fail("needs a collimated source");
            break;
        case MKO_CON_DISTORTION:
            if (c.mode != MKO_MODE_MAX) fail("defined as an upper bound only");
            if (c.target < 0.0) fail("D_max must be >= 0");
            if (c.field < 0) fail("field must be >= 0");
            if (B.b.source != MKO_SOURCE_COLLIMATED) fail("needs a collimated source");
            break;
        default:
            break;
        }
        if (con_needs_paraxial(c) && !B.b.paraxial_index) fail("batch has no paraxial_index");
        return;
    }
    if (!con_mechanical(c.type)) fail("unsupported constraint type");
    if (c.surface < 0 || c.surface >= S) fail("surface out of range");
    if (con_gap(c.type)) {
        if (c.next_surface < 0 || c.next_surface >= S) fail("next surface out of range");
        if (c.next_surface <= c.surface) fail("next surface not behind the surface");
    }
    if (!std::isfinite(c.target)) fail("limit must be finite");
    if (c.type == MKO_CON_MAX_SEMI_DIAMETER && !(c.target > 0.0)) fail("semi-diameter must be > 0");
    if (c.type == MKO_CON_MAX_INCIDENCE_ANGLE && !(c.target > 0.0 && c.target < 1.0))
        fail("cos theta_max must be in (0, 1)");
    if (!(std::isfinite(c.offset) && c.offset >= 0.0)) fail("clearance must be finite and >= 0");
    if (c.type == MKO_CON_MIN_EDGE_THICKNESS) {
        if (c.edge_height != MKO_EDGE_SEMI_DIAMETER && c.edge_height != MKO_EDGE_RAYS)
            fail("unknown edge height");
        if (c.edge_height == MKO_EDGE_SEMI_DIAMETER &&
            !(std::isfinite(c.semi_diameter) && c.semi_diameter > 0.0 &&
              std::isfinite(c.next_semi_diameter) && c.next_semi_diameter > 0.0))
            fail("semi-diameters must be finite and > 0");
    }
}

// sag / edge_thickness, mechanical_constraints.cpp:58-69
inline double sag(double curvature, double height) {
    const double h2 = height * height;
    const double a = std::abs(curvature * height);
    if (!(a < 1.0)) return curvature * h2;
    return curvature * h2 / (1.0 + std::sqrt((1.0 - a) * (1.0 + a)));
}

// A collimated ray at lambda_ref (the variant's table with one wavelength, the paraxial index):
// collimated_ray(a, theta, rho, 0) traced through surfaces 1 .. S-1 in hard mode
// (batch_sine_condition, paraxial_residuals.cpp:68-91; batch_chief_ray_height,
// system_batch.cpp:999-1030). False where the ray fails.
bool lambda_ref_ray(const Batch& B, int v, double theta, double rho, Ray& ray) {
    const int S = B.S();
    const Table tb{S, 1, B.b.type, B.b.curvature + size_t(v) * S, B.b.thickness + size_t(v) * S,
                   B.b.paraxial_index + size_t(v) * S};
    const double r = B.b.pupil_radius[v] * rho;
    ray.p = Vec3{r * std::sin(0.0), r * std::cos(0.0), 0.0};
    ray.d = Vec3{0.0, std::sin(theta), std::cos(theta)};
    ray.opl = dot(ray.p, ray.d);
    return trace_ray(tb, 0, ray, nullptr, nullptr, nullptr) == Ok;
}

// conrady_osc (include/mikroptim/paraxial/sine_condition.hpp:25-40) of the zone's ray; NaN where
// undefined
double sine_condition(const Batch& B, int v, const Paraxial& pd, double rho) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    Ray ray;
    if (!lambda_ref_ray(B, v, 0.0, rho, ray)) return nan;
    const int S = B.S();
    const double height = rho * B.b.pupil_radius[v];
    const Vec3& p = ray.p;
    const Vec3& d = ray.d;
    if (d.y == 0.0 || pd.afocal || !std::isfinite(pd.focal_length)) return nan;
    const double f_image = B.b.paraxial_index[size_t(v) * S + size_t(S - 1)] * pd.focal_length;
    const double f_real = -height / d.y;
    double factor = 1.0;
    if (!pd.exit_pupil_at_infinity) {
        const double L = p.z - p.y * d.z / d.y;
        factor = (pd.back_focal_distance - pd.exit_pupil_position) / (L - pd.exit_pupil_position);
    }
    return f_real / f_image * factor - 1.0;
}

// the real chief ray's height in the image plane, NaN where it fails or runs backward
double chief_ray_height(const Batch& B, int v, int f) {
    Ray ray;
    const double theta = B.b.field_angle[size_t(v) * size_t(B.F()) + size_t(f)];
    if (!lambda_ref_ray(B, v, theta, 0.0, ray) || !(ray.d.z > 0.0))
        return std::numeric_limits<double>::quiet_NaN();
    const double w = B.b.thickness[size_t(v) * B.S() + size_t(B.S() - 1)];
    return ray.p.y + ray.d.y * ((w - ray.p.z) / ray.d.z);
}

// ---- glass hull (project_onto_hull, src/glass/convex_hull.cpp:152-215, 397-452; the closest
// point of a triangle by Voronoi regions, Ericson, Real-Time Collision Detection, 5.1.5) --------

using P3 = std::array<double, 3>;
inline double dot3(const P3& a, const P3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline P3 sub3(const P3& a, const P3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline P3 axpy3(const P3& a, const P3& d, double t) {
    return {a[0] + d[0] * t, a[1] + d[1] * t, a[2] + d[2] * t};
}
inline double len3(const P3& a) { return std::sqrt(dot3(a, a)); }
inline P3 at3(const double* r) { return {r[0], r[1], r[2]}; }

P3 closest_on_segment(const P3& p, const P3& a, const P3& b) {
    const P3 ab = sub3(b, a);
    const double len2 = dot3(ab, ab);
    if (!(len2 > 0.0)) return a;
    const double t = std::clamp(dot3(sub3(p, a), ab) / len2, 0.0, 1.0);
    return axpy3(a, ab, t);
}

P3 closest_on_triangle(const P3& p, const P3& a, const P3& b, const P3& c) {
    const P3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
    const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const P3 bp = sub3(p, b);
    const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return axpy3(a, ab, d1 / (d1 - d3));
    const P3 cp = sub3(p, c);
    const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return axpy3(a, ac, d2 / (d2 - d6));
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
        return axpy3(b, sub3(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6)));
    const double sum = va + vb + vc;
    if (!(sum > 0.0)) {
        P3 best = closest_on_segment(p, a, b);
        double bd = len3(sub3(p, best));
        for (const P3& q : {closest_on_segment(p, b, c), closest_on_segment(p, a, c)}) {
            const double d = len3(sub3(p, q));
            if (d < bd) {
                bd = d;
                best = q;
            }
        }
        return best;
    }
    const double v = vb / sum, w = vc / sum;
    return {a[0] + (ab[0] * v + ac[0] * w), a[1] + (ab[1] * v + ac[1] * w),
            a[2] + (ab[2] * v + ac[2] * w)};
}

// glass_hull_residual (src/optim/glass_variables.cpp:135-155) of the medium's (n0, nu1, nu2):
// component `comp` of sqrt(w) max(0, phi) grad
double glass_hull_row(const double* hull, const P3& x, double weight, double margin, int comp) {
    const int dim = int(hull[0]), n = int(hull[1]);
    const double* rec = hull + 8;
    const P3 u{(x[0] - hull[2]) / hull[5], (x[1] - hull[3]) / hull[6], (x[2] - hull[4]) / hull[7]};
    P3 grad{0.0, 0.0, 0.0};
    double phi = 0.0;
    bool inside = false;
    if (dim == 3) {
        double worst = -std::numeric_limits<double>::infinity();
        int wf = 0;
        for (int f = 0; f < n; ++f) {
            const double sd = dot3(at3(rec + 13 * f + 9), u) - rec[13 * f + 12];
            if (sd > worst) {
                worst = sd;
                wf = f;
            }
        }
        if (worst <= 0.0) {
            inside = true;
            phi = worst + margin;
            grad = at3(rec + 13 * wf + 9);
        }
    }
    if (!inside) {
        double best = std::numeric_limits<double>::infinity();
        P3 closest{0.0, 0.0, 0.0};
        const int m = dim >= 2 ? n : 1;
        for (int f = 0; f < m; ++f) {
            const double* r = rec + 13 * f;
            const P3 q = dim >= 2   ? closest_on_triangle(u, at3(r), at3(r + 3), at3(r + 6))
                         : dim == 1 ? closest_on_segment(u, at3(r), at3(r + 3))
                                    : at3(r);
            const double d = len3(sub3(u, q));
            if (d < best) {
                best = d;
                closest = q;
            }
        }
        if (best > 0.0) {
            const double inv = 1.0 / best;
            grad = {(u[0] - closest[0]) * inv, (u[1] - closest[1]) * inv,
                    (u[2] - closest[2]) * inv};
        }
        phi = dim == 3 ? best + margin : best;
    }
    const double e = std::max(0.0, phi);
    if (!(e > 0.0)) return 0.0;
    return (std::sqrt(weight) * e) * grad[size_t(comp)];
}

double constraint_row(const mko_constraint& c, const Batch& B, int v, const Paraxial& p,
                      const double* ext, const double* hull = nullptr) {
    const int S = B.S();
    const double* t = B.b.thickness + size_t(v) * S;
    if (c.type == MKO_CON_GLASS_HULL) {
        const double* bd = B.b.buchdahl + (size_t(v) * S + size_t(c.surface)) * 4;
        return glass_hull_row(hull, P3{bd[0], bd[1], bd[2]}, c.weight, c.target, c.field);
    }
    if (con_paraxial(c.type)) {
        double q = 0.0;
        switch (c.type) {
        case MKO_CON_FOCAL_LENGTH: q = c.target * (1.0 - c.target * p.power); break;
        case MKO_CON_WORKING_DISTANCE: q = t[S - 1] - c.target; break;
        case MKO_CON_PARFOCAL_LENGTH: {
            double length = c.offset;
            for (int j = c.surface; j < S; ++j) length += t[j];
            q = length - c.target;
            break;
        }
        case MKO_CON_PARAXIAL_FOCUS:
            q = B.b.pupil_radius[v] *
                (p.A + t[S - 1] * p.C / B.b.paraxial_index[size_t(v) * S + size_t(S - 1)]);
            break;
        case MKO_CON_SINE_CONDITION: {
            const double osc = sine_condition(B, v, p, c.zone);
            q = std::isfinite(osc) ? osc : 0.0;
            break;
        }
        case MKO_CON_DISTORTION: {
            q = -c.target;
            if (c.field < B.F()) {
                const double h = B.b.field_target[size_t(v) * size_t(B.F()) + size_t(c.field)];
                if (h != 0.0) {
                    const double y = chief_ray_height(B, v, c.field);
                    if (std::isfinite(y)) q = std::abs((y - h) / h) - c.target;
                }
            }
            break;
        }
        default: q = std::abs(p.petzval_curvature) - c.target; break;
        }
        if (c.shift != 0.0) {
            if (c.mode == MKO_MODE_MAX) q += c.shift;
            else if (c.mode == MKO_MODE_MIN) q -= c.shift;
        }
        const double m = c.mode == MKO_MODE_MIN ? std::min(q, 0.0)
                         : c.mode == MKO_MODE_MAX ? std::max(q, 0.0)
                                                  : q;
        return std::sqrt(c.weight) * m;
    }
    const auto center = [&] {
        double sum = 0.0;
        for (int j = c.surface; j < c.next_surface; ++j) sum += t[j];
        return sum;
    };
    const auto curv = [&](int j) {
        return B.b.type[j] == MKO_SURFACE_SPHERE ? B.b.curvature[size_t(v) * S + size_t(j)] : 0.0;
    };
    double value = 0.0;
    switch (c.type) {
    case MKO_CON_MAX_SEMI_DIAMETER: value = ext[2 * c.surface] + c.offset; break;
    case MKO_CON_MIN_CENTER_THICKNESS:
    case MKO_CON_MAX_CENTER_THICKNESS: value = center(); break;
    case MKO_CON_MIN_EDGE_THICKNESS: {
        const double h = (c.edge_height == MKO_EDGE_RAYS
                              ? std::max(ext[2 * c.surface], ext[2 * c.next_surface])
                              : std::max(c.semi_diameter, c.next_semi_diameter)) +
                         c.offset;
        value = center() - sag(curv(c.surface), h) + sag(curv(c.next_surface), h);
        break;
    }
    default: value = ext[2 * c.surface + 1]; break;
    }
    double g = (c.type == MKO_CON_MAX_SEMI_DIAMETER || c.type == MKO_CON_MAX_CENTER_THICKNESS)
                   ? value - c.target
                   : c.target - value;
    if (c.shift != 0.0) g += c.shift;
    return std::sqrt(c.weight) * std::max(g, 0.0);
}

// The E + E_t constraint rows of variant v (ext: its extrema [S][2], needed when E_t > 0);
// untraced_only: the first E rows only.
void variant_rows(const Batch& B, int v, const Layout& lay, const double* ext, double* out,
                  bool untraced_only = false) {
    Paraxial p;
    bool need_p = false;
    const int n = untraced_only ? lay.E : lay.E + lay.Et;
    for (int k = 0; k < n; ++k) need_p = need_p || con_needs_paraxial(lay.items[size_t(k)]);
    if (need_p) p = batch_paraxial(B, v);
    for (int k = 0; k < n; ++k)
        out[k] = constraint_row(lay.items[size_t(k)], B, v, p, ext, lay.hull);
}

// rows of variant v into out (extrema from its own trace when E_t > 0)
void variant_rows_traced(const Batch& B, int v, const Layout& lay, double* out) {
    std::vector<double> ext;
    if (lay.Et > 0) {
        ext.resize(size_t(B.S()) * 2);
        variant_extrema(B, v, lay, ext.data());
    }
    variant_rows(B, v, lay, ext.data(), out);
}

// merit with constraint rows, CpuLmBackend::merit (src/optim/levenberg_marquardt.cpp:866-888):
// (ray merit + sum e^2) + sum t^2, each sum ascending from 0
void variant_merit_rows(const Batch& B, int v, const Layout& lay, double& merit, int64_t& failed,
                        int64_t& hard_failed) {
    variant_merit(B, v, lay, merit, failed, hard_failed);
    if (lay.E + lay.Et == 0) return;
    std::vector<double> rows(size_t(lay.E + lay.Et));
    variant_rows_traced(B, v, lay, rows.data());
    if (lay.E > 0) {
        double e = 0.0;
        for (int k = 0; k < lay.E; ++k) e += rows[size_t(k)] * rows[size_t(k)];
        merit = merit + e;
    }
    if (lay.Et > 0) {
        double tsum = 0.0;
        for (int k = lay.E; k < lay.E + lay.Et; ++k) tsum += rows[size_t(k)] * rows[size_t(k)];
        merit += tsum;
    }
}

// contiguous split over host threads, result independent of the thread count
// (src/optim/parallel_variants.hpp)
void parallel_for(int first, int last, int threads, const std::function<void(int)>& fn) {
    const int n = last - first;
    if (n <= 0) return;
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    threads = std::min(threads, n);
    if (threads <= 1) {
        for (int i = first; i < last; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    std::exception_ptr error;
    std::mutex m;
    for (int t = 0; t < threads; ++t) {
        const int a = first + int(int64_t(n) * t / threads);
        const int b = first + int(int64_t(n) * (t + 1) / threads);
        pool.emplace_back([&, a, b] {
            try {
                for (int i = a; i < b; ++i) fn(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(m);
                if (!error) error = std::current_exception();
            }
        });
    }
    for (auto& th : pool) th.join();
    if (error) std::rethrow_exception(error);
}

// ---- variables, FD pattern and perturbations ---------------------------------------------------

struct Vars {
    int P = 0;
    std::vector<int> type, surface, component;
    std::vector<double> lower, upper, typical;
    std::vector<int> glass_surfaces; // surfaces with glass variables, ascending
    bool glass() const { return !glass_surfaces.empty(); }
};

inline Vars make_vars(const mko_variables* v, int S) {
    Vars o;
    o.P = v->count;
    for (int k = 0; k < o.P; ++k) {
        if (v->type[k] < MKO_VAR_CURVATURE || v->type[k] > MKO_VAR_GLASS)
            throw std::invalid_argument("unsupported variable type");
        if (v->batch_surface[k] < 0 || v->batch_surface[k] >= S)
            throw std::invalid_argument("variable surface out of range");
        if (!(v->typical[k] > 0.0)) throw std::invalid_argument("typical must be > 0");
        int comp = 0;
        if (v->type[k] == MKO_VAR_GLASS) {
            if (!v->component) throw std::invalid_argument("glass variables need components");
            comp = v->component[k];
            if (comp < 0 || comp > 2) throw std::invalid_argument("glass component outside 0 ... 2");
            if (std::find(o.glass_surfaces.begin(), o.glass_surfaces.end(), v->batch_surface[k]) ==
                o.glass_surfaces.end())
                o.glass_surfaces.push_back(v->batch_surface[k]);
        }
        o.type.push_back(v->type[k]);
        o.surface.push_back(v->batch_surface[k]);
        o.component.push_back(comp);
        o.lower.push_back(v->lower[k]);
        o.upper.push_back(v->upper[k]);
        o.typical.push_back(v->typical[k]);
    }
    std::sort(o.glass_surfaces.begin(), o.glass_surfaces.end());
    return o;
}

// nm_to_um, buchdahl_omega, buchdahl_index (include/mikroptim/glass/dispersion.hpp:18-95)
constexpr double kBuchdahlLambda0Um = 0.546074, kBuchdahlAlpha = 2.5;
inline double buchdahl_omega(double lambda_nm) {
    const double um = lambda_nm / 1000.0;
    const double den = 1.0 + kBuchdahlAlpha * (um - kBuchdahlLambda0Um);
    if (!(den > 0.0)) throw std::invalid_argument("Buchdahl coordinate undefined (lambda <= 146.074 nm)");
    return (um - kBuchdahlLambda0Um) / den;
}
inline double buchdahl_index(const double* c, double lambda_nm) {
    const double w = buchdahl_omega(lambda_nm);
    return c[0] + w * (c[1] + w * (c[2] + w * c[3]));
}

// The per-variant arrays VariableSet::apply writes.
struct Target {
    double* curvature;
    double* thickness;
    double* index;
    double* buchdahl; // [V][S][4] or NULL
    double* pindex;   // [V][S] or NULL
};

// VariableSet::apply, src/optim/variable_set.cpp:237-285: curvature / thickness / working
// distance; glass components into the Buchdahl coefficients, then the glass surfaces' index at
// every wavelength and at lambda_ref.
inline void apply(const Vars& vars, const mko_batch& b, const Target& t, int v, const double* x) {
    const size_t S = size_t(b.surfaces), L = size_t(b.wavelengths);
    if (vars.glass()) {
        if (!t.buchdahl || !b.wavelength_nm)
            throw std::invalid_argument("glass variables need Buchdahl coefficients and wavelengths");
        for (size_t l = 0; l < L; ++l) (void)buchdahl_omega(b.wavelength_nm[l]);
        (void)buchdahl_omega(b.reference_wavelength_nm);
    }
    for (int k = 0; k < vars.P; ++k) {
        if (!std::isfinite(x[k])) throw std::invalid_argument("variable value not finite");
        const size_t s = size_t(vars.surface[k]);
        if (vars.type[k] == MKO_VAR_CURVATURE)
            t.curvature[size_t(v) * S + s] = x[k];
        else if (vars.type[k] == MKO_VAR_GLASS)
            t.buchdahl[(size_t(v) * S + s) * 4 + size_t(vars.component[k])] = x[k];
        else
            t.thickness[size_t(v) * S + s] = x[k];
    }
    for (const int si : vars.glass_surfaces) {
        const size_t s = size_t(si);
        const double* c = t.buchdahl + (size_t(v) * S + s) * 4;
        for (size_t l = 0; l < L; ++l)
            t.index[(size_t(v) * S + s) * L + l] = buchdahl_index(c, b.wavelength_nm[l]);
        if (t.pindex) t.pindex[size_t(v) * S + s] = buchdahl_index(c, b.reference_wavelength_nm);
    }
}

inline void clamp_to_bounds(const Vars& vars, double* x) {
    for (int k = 0; k < vars.P; ++k) x[k] = std::min(std::max(x[k], vars.lower[k]), vars.upper[k]);
}

constexpr double kFdStepFactor = 6.0554544523933395e-06; // 2^(-52/3), variant_generator.hpp:81

struct FdPattern {
    std::vector<double> nominal, plus, minus, step_plus, step_minus;
    std::vector<uint8_t> stencil;
};

// fd_pattern (central scheme), src/optim/variant_generator.cpp:177-237
inline FdPattern fd_pattern(const Vars& vars, const double* x0) {
    FdPattern pat;
    const int P = vars.P;
    pat.nominal.resize(P);
    pat.plus.resize(P);
    pat.minus.resize(P);
    pat.step_plus.resize(P);
    pat.step_minus.resize(P);
    pat.stencil.resize(P);
    for (int k = 0; k < P; ++k) {
        const double x = x0[k];
        if (!std::isfinite(x) || !(x >= vars.lower[k] && x <= vars.upper[k]))
            throw std::invalid_argument("x0 not finite or outside bounds");
        const double h = kFdStepFactor * std::max(std::abs(x), vars.typical[k]);
        pat.nominal[k] = h;
        const double up = x + h;
        const double down = x - h;
        double xa, xb;
        uint8_t stencil = MKO_STENCIL_CENTRAL;
        if (up <= vars.upper[k] && down >= vars.lower[k]) {
            xa = up;
            xb = down;
        } else {
            const bool towards_upper = vars.upper[k] - x >= x - vars.lower[k];
            const double far = towards_upper ? x + 2.0 * h : x - 2.0 * h;
            if (towards_upper ? far <= vars.upper[k] : far >= vars.lower[k]) {
                stencil = MKO_STENCIL_ONE_SIDED;
                xa = towards_upper ? up : down;
                xb = far;
            } else {
                stencil = MKO_STENCIL_TRUNCATED;
                xa = up <= vars.upper[k] ? up : vars.upper[k];
                xb = down >= vars.lower[k] ? down : vars.lower[k];
            }
        }
        pat.stencil[k] = stencil;
        pat.plus[k] = xa;
        pat.minus[k] = xb;
        pat.step_plus[k] = xa - x;
        pat.step_minus[k] = xb - x;
    }
    return pat;
}

// fd_derivative, include/mikroptim/optim/fd_formula.hpp:39-53
inline double fd_derivative(uint8_t stencil, double r0, double ra, double rb, double da,
                            double db) {
    if (stencil == MKO_STENCIL_ONE_SIDED) {
        const double ga = ra - r0;
        const double gb = rb - r0;
        return (db * db * ga - da * da * gb) / (da * db * (db - da));
    }
    return (ra - rb) / (da - db);
}

// splitmix64 / variant_seed / unit_uniform, src/optim/variant_generator.cpp:67-82
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
inline uint64_t variant_seed(uint64_t seed, uint64_t j) {
    return splitmix64(seed + j * 0x9E3779B97F4A7C15ULL);
}
inline double unit_uniform(uint64_t bits) { return double(bits >> 11) * 0x1.0p-53; }

// perturbed_vector, src/optim/variant_generator.cpp:95-127; rel/abs per variable type
inline void perturbed_vector(const Vars& vars, const double* x0, const double* relative,
                             const double* absolute, int distribution, uint64_t seed,
                             int first_is_base, uint64_t j, double* y) {
    for (int k = 0; k < vars.P; ++k) y[k] = x0[k];
    if (first_is_base && j == 0) return;
    std::mt19937_64 engine(variant_seed(seed, j));
    for (int k = 0; k < vars.P; ++k) {
        double xi;
        if (distribution == 0) {
            xi = 2.0 * unit_uniform(engine()) - 1.0;
        } else {
            double sum = 0.0;
            for (int i = 0; i < 12; ++i) sum += unit_uniform(engine());
            xi = sum - 6.0;
        }
        const int t = vars.type[k];
        const double s = std::fma(relative[t], std::abs(x0[k]), absolute[t]);
        y[k] = std::fma(s, xi, x0[k]);
    }
    clamp_to_bounds(vars, y);
}

// ---- residual batches ----------------------------------------------------------------------------

// A batch whose per-variant arrays this oracle owns (FD batches, point batches).
struct OwnedBatch {
    mko_batch b{};
    std::vector<double> curvature, thickness, index, pupil, fangle, ftarget, refcurv, pindex, bd;
    // copies variant `from` of `src` into variant `to`
    void init(const mko_batch& src, int V) {
        b = src;
        b.variants = V;
        const size_t S = size_t(src.surfaces), L = size_t(src.wavelengths), F = size_t(src.fields);
        curvature.assign(size_t(V) * S, 0.0);
        thickness.assign(size_t(V) * S, 0.0);
        index.assign(size_t(V) * S * L, 0.0);
        pupil.assign(size_t(V), 0.0);
        fangle.assign(size_t(V) * F, 0.0);
        ftarget.assign(size_t(V) * F, 0.0);
        if (src.reference_curvature) refcurv.assign(size_t(V) * F, 0.0);
        if (src.paraxial_index) pindex.assign(size_t(V) * S, 0.0);
        b.paraxial_index = src.paraxial_index ? pindex.data() : nullptr;
        if (src.buchdahl) bd.assign(size_t(V) * S * 4, 0.0);
        b.buchdahl = src.buchdahl ? bd.data() : nullptr;
        b.curvature = curvature.data();
        b.thickness = thickness.data();
        b.index = index.data();
        b.pupil_radius = pupil.data();
        b.field_angle = fangle.data();
        b.field_target = ftarget.data();
        b.reference_curvature = src.reference_curvature ? refcurv.data() : nullptr;
    }
    void copy_variant(const mko_batch& src, int from, int to) {
        const size_t S = size_t(src.surfaces), L = size_t(src.wavelengths), F = size_t(src.fields);
        std::copy_n(src.curvature + from * S, S, curvature.data() + to * S);
        std::copy_n(src.thickness + from * S, S, thickness.data() + to * S);
        std::copy_n(src.index + from * S * L, S * L, index.data() + to * S * L);
        pupil[to] = src.pupil_radius[from];
        std::copy_n(src.field_angle + from * F, F, fangle.data() + to * F);
        std::copy_n(src.field_target + from * F, F, ftarget.data() + to * F);
        if (src.reference_curvature)
            std::copy_n(src.reference_curvature + from * F, F, refcurv.data() + to * F);
        if (src.paraxial_index)
            std::copy_n(src.paraxial_index + from * S, S, pindex.data() + to * S);
        if (src.buchdahl) std::copy_n(src.buchdahl + from * S * 4, S * 4, bd.data() + to * S * 4);
    }
    Target target() {
        return Target{curvature.data(), thickness.data(), index.data(),
                      bd.empty() ? nullptr : bd.data(), pindex.empty() ? nullptr : pindex.data()};
    }
};

// RESIDUALS of every variant, rows [V][trace rows + E + E_t] (CpuLmBackend::host_residuals,
// src/optim/levenberg_marquardt.cpp:812-864: constraint rows from the variant's own extrema)
void residuals_all(const Batch& B, const Layout& lay, double* residuals, double* opd,
                   uint8_t* status, int threads) {
    const size_t R = size_t(lay.rows());
    const size_t rays = size_t(lay.rays);
    parallel_for(0, B.b.variants, threads, [&](int v) {
        variant_residuals(B, v, lay, residuals + size_t(v) * R, opd + size_t(v) * rays,
                          status + size_t(v) * rays);
        if (lay.E + lay.Et > 0)
            variant_rows_traced(B, v, lay, residuals + size_t(v) * R + size_t(lay.trace_rows()));
    });
}

// FD Jacobian (central scheme), src/optim/fd_jacobian.cpp:46-182: batch of 2P+1 variants per
// run (variant 0 = x0, 2k+1 = x0 + d+ e_k, 2k+2 = x0 + d- e_k), RESIDUALS of all, then
// fd_derivative per row; J[run][p][R], r0[run][R], status_changes[run][p].
void fd_jacobian(const Batch& base, const Layout& lay, const Vars& vars, int runs, const double* x0,
                 double* J, double* r0, int64_t* changes, uint8_t* status0, int threads) {
    const int P = vars.P;
    const int n = 2 * P + 1;
    const int S = base.S();
    OwnedBatch fd;
    fd.init(base.b, runs * n);
    std::vector<FdPattern> pats;
    std::vector<double> y(P);
    for (int i = 0; i < runs; ++i) {
        const double* xi = x0 + size_t(i) * P;
        const int tmpl = base.b.variants == 1 ? 0 : i;
        pats.push_back(fd_pattern(vars, xi));
        for (int j = 0; j < n; ++j) fd.copy_variant(base.b, tmpl, i * n + j);
        std::copy_n(xi, P, y.data());
        apply(vars, fd.b, fd.target(), i * n, y.data());
        for (int k = 0; k < P; ++k) {
            y[k] = pats[i].plus[k];
            apply(vars, fd.b, fd.target(), i * n + 2 * k + 1, y.data());
            y[k] = pats[i].minus[k];
            apply(vars, fd.b, fd.target(), i * n + 2 * k + 2, y.data());
            y[k] = xi[k];
        }
    }
    const Batch FB{fd.b};
    const size_t R = size_t(lay.rows());
    const size_t rays = size_t(lay.rays);
    std::vector<double> res(size_t(runs) * n * R), opd(size_t(runs) * n * rays);
    std::vector<uint8_t> st(size_t(runs) * n * rays);
    residuals_all(FB, lay, res.data(), opd.data(), st.data(), threads);
    for (int i = 0; i < runs; ++i) {
        const FdPattern& pat = pats[i];
        const size_t v0 = size_t(i) * n;
        double* r0i = r0 + size_t(i) * R;
        for (size_t r = 0; r < R; ++r) r0i[r] = res[v0 * R + r];
        if (status0) std::copy_n(st.data() + v0 * rays, rays, status0 + size_t(i) * rays);
        for (int k = 0; k < P; ++k) {
            const size_t va = v0 + size_t(2 * k + 1), vb = v0 + size_t(2 * k + 2);
            double* col = J + (size_t(i) * P + size_t(k)) * R;
            for (size_t r = 0; r < R; ++r)
                col[r] = fd_derivative(pat.stencil[k], r0i[r], res[va * R + r], res[vb * R + r],
                                       pat.step_plus[k], pat.step_minus[k]);
            int64_t c = 0;
            for (size_t r = 0; r < rays; ++r) {
                const bool ok0 = st[v0 * rays + r] == Ok;
                if ((st[va * rays + r] == Ok) != ok0 || (st[vb * rays + r] == Ok) != ok0) ++c;
            }
            if (changes) changes[size_t(i) * P + k] = c;
        }
    }
}

// ---- adjoint: include/mikroptim/trace/adjoint_backward.hpp ------------------------------------

struct Seed {
    double opl, penalty;
};
struct RayAdj {
    Vec3 p, d;
};
struct SurfAdj {
    double thickness = 0, curvature = 0, n_before = 0, n_after = 0;
};

// refracting_step_backward, include/mikroptim/trace/adjoint_backward.hpp:115-241
inline void refracting_step_backward(const Vec3& p, const Vec3& d, double t, double c,
                                     double n_before, double n_after, uint8_t branch, Seed seed,
                                     RayAdj& adj, SurfAdj& out) {
    // local forward step (surface_ops.hpp formulas)
    const double z = p.z - t;
    const double s0 = -z / d.z;
    const double x = p.x + s0 * d.x;
    const double y = p.y + s0 * d.y;
    const double rr = x * x + y * y;
    const double H = c * rr;
    const double xd = x * d.x + y * d.y;
    const double B = d.z - c * xd;
    const double D = B * B - c * H;
    const bool missed = (branch & kBranchSoftMissed) != 0;
    const bool tir = (branch & kBranchSoftTIR) != 0;
    double root_d = 0.0;
    double s;
    if (missed) {
        s = B / c;
    } else {
        root_d = std::sqrt(D);
        s = H / (B + root_d);
    }
    const Vec3 q{x + s * d.x, y + s * d.y, s * d.z};
    const Vec3 n_raw{-c * q.x, -c * q.y, 1.0 - c * q.z};
    double inv_norm = 1.0;
    Vec3 N = n_raw;
    if (missed) {
        inv_norm = 1.0 / norm(n_raw);
        N = inv_norm * n_raw;
    }
    const double mu = n_before / n_after;
    const double cos_i = dot(d, N);
    const double sin2 = 1.0 - cos_i * cos_i;
    const double k = 1.0 - mu * mu * sin2;

    // reverse
    const Vec3 qb_out = adj.p;
    const Vec3 d1b = adj.d;
    RayAdj in{{0, 0, 0}, {0, 0, 0}};
    Vec3 nb{0, 0, 0};
    double cosb = 0.0, kb = 0.0, mub = 0.0;
    if (tir) {
        const Vec3 tau = d - cos_i * N;
        const double inv_tau = 1.0 / norm(tau);
        const Vec3 d1 = inv_tau * tau;
        const Vec3 taub = inv_tau * (d1b - dot(d1, d1b) * d1);
        in.d += taub;
        cosb += -dot(taub, N);
        nb += (-cos_i) * taub;
        kb += -seed.penalty;
    } else {
        const double root_k = std::sqrt(k);
        const double a = root_k - mu * cos_i;
        in.d += mu * d1b;
        nb += a * d1b;
        const double ab = dot(d1b, N);
        mub += dot(d1b, d);
        kb += ab / (2.0 * root_k);
        mub += -ab * cos_i;
        cosb += -ab * mu;
    }
    mub += -kb * 2.0 * mu * sin2;
    cosb += kb * 2.0 * mu * mu * cos_i;
    in.d += cosb * N;
    nb += cosb * d;
    out.n_before += mub / n_after;
    out.n_after += -mub * mu / n_after;
    Vec3 nrb = nb;
    if (missed) nrb = inv_norm * (nb - dot(N, nb) * N);
    const Vec3 qb = qb_out + (-c) * nrb;
    double cb = -dot(q, nrb);
    double xb = qb.x;
    double yb = qb.y;
    double sb = dot(qb, d);
    in.d += s * qb;
    sb += seed.opl * n_before;
    double s0b = seed.opl * n_before;
    out.n_before += seed.opl * (s0 + s);
    double hb = 0.0, bb = 0.0;
    if (missed) {
        bb += sb / c;
        cb += -sb * s / c;
        const double db = -seed.penalty;
        bb += db * 2.0 * B;
        cb += -db * H;
        hb += -db * c;
    } else {
        const double inv = 1.0 / (2.0 * root_d);
        hb += sb * inv;
        bb += -sb * s * 2.0 * inv;
        cb += sb * s * s * inv;
    }
    cb += hb * rr;
    xb += hb * 2.0 * c * x;
    yb += hb * 2.0 * c * y;
    in.d.z += bb;
    cb += -bb * xd;
    xb += -bb * c * d.x;
    yb += -bb * c * d.y;
    in.d.x += -bb * c * x;
    in.d.y += -bb * c * y;
    // transfer_backward, adjoint_backward.hpp:93-108
    in.p.x += xb;
    in.p.y += yb;
    s0b += xb * d.x + yb * d.y;
    in.d.x += xb * s0;
    in.d.y += yb * s0;
    const double qq = s0b / d.z;
    out.thickness += qq;
    in.p.z += -qq;
    in.d.z += -qq * s0;
    out.curvature += cb;
    adj = in;
}

// reference_step_backward, adjoint_backward.hpp:373-419
inline void reference_step_backward(const Vec3& p, const Vec3& d, const Vec3& center,
                                    double n_image, uint8_t branch, Seed seed, RayAdj& adj,
                                    Vec3& center_adj, double& n_image_adj) {
    const Vec3 q = p - center;
    const double b = -dot(q, d);
    const double g = dot(p, p - 2.0 * center);
    const double disc = b * b - g;
    const bool missed = (branch & kBranchSoftMissed) != 0;
    const double sigma = center.z >= 0.0 ? 1.0 : -1.0;
    double s, root = 0.0;
    if (missed) {
        s = b;
    } else {
        root = std::sqrt(disc);
        s = g / (b + sigma * root);
    }
    const double sb = seed.opl * n_image;
    n_image_adj += seed.opl * s;
    double bb, gb;
    if (missed) {
        bb = sb;
        gb = 0.0;
        const double radius_sq = dot(center, center);
        const double discb = -seed.penalty / radius_sq;
        bb += discb * 2.0 * b;
        gb += -discb;
        center_adj += (seed.penalty * 2.0 * disc / (radius_sq * radius_sq)) * center;
    } else {
        gb = sb * sigma / (2.0 * root);
        bb = -sb * sigma * s / root;
    }
    RayAdj in{{0, 0, 0}, {0, 0, 0}};
    in.p += (-bb) * d;
    in.d += (-bb) * q;
    center_adj += bb * d;
    in.p += (2.0 * gb) * (p - center);
    center_adj += (-2.0 * gb) * p;
    adj = in;
}

// image_step_backward, adjoint_backward.hpp:429-464
inline void image_step_backward(const Vec3& p, const Vec3& d, const Vec3& center, double kappa,
                                double n_image, uint8_t branch, Seed seed, RayAdj& adj,
                                Vec3& center_adj, double& n_image_adj) {
    const Vec3 q = p - center;
    const double b = -dot(q, d);
    const Vec3 c = cross(q, d);
    const double t = -dot(c, c);
    const bool missed = (branch & kBranchSoftMissed) != 0;
    double step, dstep_dt, tb = 0.0;
    if (missed) {
        step = b - kappa * t;
        dstep_dt = -kappa;
        tb += seed.penalty * (-(kappa * kappa));
    } else {
        const double u = std::sqrt(1.0 + kappa * kappa * t);
        step = b - kappa * t / (u + 1.0);
        dstep_dt = -kappa / (2.0 * u);
    }
    const double sb = seed.opl * n_image;
    n_image_adj += seed.opl * step;
    const double bb = sb;
    tb += sb * dstep_dt;
    const Vec3 qb = (-bb) * d + (-2.0 * tb) * cross(d, c);
    const Vec3 db = (-bb) * q + (-2.0 * tb) * cross(c, q);
    adj = RayAdj{qb, db};
    center_adj += (-1.0) * qb;
}

// Writes the per-ray derivative of one seed into J rows (JacobianRowWriter,
// include/mikroptim/merit/ray_jacobian.hpp:50-78; curvature, thickness and glass parameters).
struct RowWriter {
    const int* curv_col; // [S] variable index or -1
    const int* thick_col;
    double* column0;
    int64_t rows;
    int64_t row;
    double factor;
    const int* glass_col = nullptr; // [S][3] variable of (n0, nu1, nu2) of the medium after s
    const double* slope = nullptr;  // [3] dn/d(n0, nu1, nu2) at this wavelength: 1, w, w^2
    void write(int p, double value) const {
        if (p >= 0) column0[int64_t(p) * rows + row] = factor * value;
    }
    void index(int s, double value) const {
        if (!glass_col) return;
        for (int c = 0; c < 3; ++c) {
            const int p = glass_col[s * 3 + c];
            if (p >= 0) write(p, value * slope[c]);
        }
    }
};

// backward_ray, include/mikroptim/trace/adjoint_backward.hpp:496-580 (no ghosts, coaxial)
inline void backward_ray(const Table& tb, int l, const Record& rec, bool reached_image,
                         const RefSphere& ref, bool point_source, Seed seed, const RowWriter& w) {
    const int S = tb.S, L = tb.L, count = rec.count;
    RayAdj adj{{0, 0, 0}, {0, 0, 0}};
    double n_carry = 0.0; // adjoint of index[j] from step j + 1 (or the reference sphere)
    uint8_t next_branch = kBranchRegular;
    Vec3 d0{0, 0, 0};
    for (int s = count - 1; s >= 0; --s) {
        const Vec3& p = rec.p[s];
        const Vec3& d = rec.d[s];
        if (s == count - 1) {
            if (reached_image && count == S) {
                Vec3 center_adj{0, 0, 0};
                double n_image_adj = 0.0;
                const double n_image = tb.n[(S - 1) * L + l];
                if (ref.kind == MKO_REF_LAST_SURFACE)
                    reference_step_backward(p, d, ref.center, n_image, rec.image_branch, seed, adj,
                                            center_adj, n_image_adj);
                else
                    image_step_backward(p, d, ref.center, ref.kappa, n_image, rec.image_branch,
                                        seed, adj, center_adj, n_image_adj);
                w.write(w.thick_col[S - 1], center_adj.z);
                n_carry = n_image_adj;
            }
        } else {
            const int j = s + 1;
            const uint8_t kind = tb.type[j];
            SurfAdj par;
            const double c = kind == MKO_SURFACE_SPHERE ? tb.c[j] : 0.0;
            refracting_step_backward(p, d, tb.t[s], c, tb.n[s * L + l], tb.n[j * L + l],
                                     next_branch, seed, adj, par);
            if (kind == MKO_SURFACE_SPHERE) w.write(w.curv_col[j], par.curvature);
            w.write(w.thick_col[s], par.thickness);
            w.index(j, par.n_after + n_carry);
            n_carry = par.n_before;
        }
        next_branch = rec.branch[s];
        if (s == 0) d0 = d;
    }
    // the point source's direction derivative belongs to no supported variable (no decenter)
    (void)point_source;
    (void)d0;
}

// adjoint_jacobian_cpu, src/optim/adjoint_jacobian.cpp:370-481 (no extra rows)
// Constraint rows of the adjoint Jacobian by central differences, extra_rows_untraced /
// extra_rows_cpu (src/optim/adjoint_jacobian.cpp:245-337): per run, x0 and the 2P points of
// fd_pattern applied to a one-variant copy, rows of each (rows with extrema from the copy's own
// trace), fd_derivative per row into J[run][p][trace rows + e] and r0.
void extra_rows(const Batch& base, const Layout& lay, const Vars& vars, int runs, const double* x0,
                double* J, double* r0, int threads) {
    const int P = vars.P;
    const int S = base.S();
    const int64_t R = lay.rows(), T = lay.trace_rows();
    const int E = lay.E + lay.Et;
    parallel_for(0, runs, threads, [&](int i) {
        OwnedBatch one;
        one.init(base.b, 1);
        one.copy_variant(base.b, base.b.variants == 1 ? 0 : i, 0);
        const Batch OB{one.b};
        const double* xi = x0 + size_t(i) * P;
        const FdPattern pat = fd_pattern(vars, xi);
        std::vector<double> y(xi, xi + P);
        std::vector<double> e0(static_cast<size_t>(E)), ea(static_cast<size_t>(E)),
            eb(static_cast<size_t>(E));
        const auto rows_at = [&](std::vector<double>& out) {
            apply(vars, one.b, one.target(), 0, y.data());
            variant_rows_traced(OB, 0, lay, out.data());
        };
        rows_at(e0);
        for (int e = 0; e < E; ++e) r0[size_t(i) * R + size_t(T + e)] = e0[size_t(e)];
        for (int k = 0; k < P; ++k) {
            y[k] = pat.plus[k];
            rows_at(ea);
            y[k] = pat.minus[k];
            rows_at(eb);
            y[k] = xi[k];
            double* col = J + (size_t(i) * P + size_t(k)) * R + size_t(T);
            for (int e = 0; e < E; ++e)
                col[e] = fd_derivative(pat.stencil[k], e0[size_t(e)], ea[size_t(e)], eb[size_t(e)],
                                       pat.step_plus[k], pat.step_minus[k]);
        }
    });
}

void adjoint_jacobian(const Batch& base, const Layout& lay, const Vars& vars, int runs,
                      const double* x0, double* J, double* r0, uint8_t* status_out, int threads) {
    const int P = vars.P;
    const int S = base.S();
    OwnedBatch ab;
    ab.init(base.b, runs);
    for (int i = 0; i < runs; ++i) {
        ab.copy_variant(base.b, base.b.variants == 1 ? 0 : i, i);
        apply(vars, ab.b, ab.target(), i, x0 + size_t(i) * P);
    }
    const Batch B{ab.b};
    std::vector<int> curv_col(S, -1), thick_col(S, -1), glass_col(size_t(S) * 3, -1);
    for (int k = 0; k < P; ++k) {
        if (vars.type[k] == MKO_VAR_CURVATURE) {
            if (curv_col[vars.surface[k]] >= 0) throw std::invalid_argument("duplicate parameter");
            curv_col[vars.surface[k]] = k;
        } else if (vars.type[k] == MKO_VAR_GLASS) {
            int& c = glass_col[size_t(vars.surface[k]) * 3 + size_t(vars.component[k])];
            if (c >= 0) throw std::invalid_argument("duplicate parameter");
            c = k;
        } else {
            if (thick_col[vars.surface[k]] >= 0) throw std::invalid_argument("duplicate parameter");
            thick_col[vars.surface[k]] = k;
        }
    }
    const int64_t rays = lay.rays;
    const int64_t R = lay.rows();
    std::fill(J, J + size_t(runs) * P * R, 0.0);
    std::vector<uint8_t> status(size_t(runs) * rays);
    parallel_for(0, runs, threads, [&](int v) {
        std::vector<double> opd(static_cast<size_t>(rays));
        variant_residuals(B, v, lay, r0 + size_t(v) * R, opd.data(), status.data() + size_t(v) * rays);
    });
    const bool soft_mode = lay.soft();
    const bool point_source = B.b.source == MKO_SOURCE_POINT;
    const int F = lay.F;
    // dn(lambda_l)/d(n0, nu1, nu2) = (1, w_l, w_l^2) (AdjointParameters, adjoint_jacobian.cpp:
    // 107-127)
    std::vector<double> slope(size_t(lay.L) * 3, 0.0);
    for (int l = 0; l < lay.L; ++l) {
        const double w = vars.glass() ? buchdahl_omega(B.b.wavelength_nm[l]) : 0.0;
        slope[size_t(l) * 3 + 0] = 1.0;
        slope[size_t(l) * 3 + 1] = w;
        slope[size_t(l) * 3 + 2] = w * w;
    }
    const int* gcol = vars.glass() ? glass_col.data() : nullptr;
    parallel_for(0, runs * F, threads, [&](int pair) {
        const int v = pair / F;
        const int f = pair % F;
        Record rec;
        const Table tb = B.table(v);
        double* col0 = J + size_t(v) * P * R;
        const Quadrature& q = lay.quad[f];
        const RefSphere ref = reference_of(B, tb, v, f);
        const int n = lay.samples[f];
        std::vector<double> weight(n), root_weight(n);
        std::vector<uint8_t> piston(n);
        for (int l = 0; l < lay.L; ++l) {
            const int64_t offset = lay.offset(f, l);
            const double scale = residual_scale(B, f, l);
            const double pscale =
                soft_mode ? penalty_scale(B, f, l, lay.settings.fail_weight) : 0.0;
            for (int k = 0; k < n; ++k) {
                const Sample s = sample(q, k);
                Soft soft;
                Ray ray = source_ray(B, v, f, l, s);
                const Status st =
                    trace_to_reference(tb, l, ray, ref, soft_mode ? &soft : nullptr, &rec);
                if (st != status[size_t(v) * rays + size_t(offset + k)])
                    throw std::logic_error("recorded trace gives another status");
                weight[k] = s.weight;
                root_weight[k] = residual_weight(s);
                piston[k] = in_piston(st) ? 1 : 0;
                const double* sl = slope.data() + size_t(l) * 3;
                const RowWriter opl_row{curv_col.data(), thick_col.data(), col0, R, offset + k,
                                        1.0, gcol, sl};
                const RowWriter pen_row{curv_col.data(), thick_col.data(), col0, R,
                                        rays + offset + k, pscale * root_weight[k], gcol, sl};
                // ray_gradient_rows, include/mikroptim/merit/ray_jacobian.hpp:93-107
                const bool reached = in_piston(st);
                if (reached) backward_ray(tb, l, rec, true, ref, point_source, Seed{1.0, 0.0}, opl_row);
                if (soft_mode && soft.first != Ok)
                    backward_ray(tb, l, rec, reached, ref, point_source, Seed{0.0, 1.0}, pen_row);
            }
            double w_ok = 0.0;
            for (int k = 0; k < n; ++k)
                if (piston[k]) w_ok += weight[k];
            for (int p = 0; p < P; ++p) {
                double* col = col0 + int64_t(p) * R + offset;
                if (!(w_ok > 0.0)) {
                    for (int k = 0; k < n; ++k) col[k] = 0.0;
                    continue;
                }
                double sum = 0.0;
                for (int k = 0; k < n; ++k)
                    if (piston[k]) sum += weight[k] * col[k];
                const double mean = sum / w_ok;
                for (int k = 0; k < n; ++k)
                    col[k] = piston[k] ? scale * root_weight[k] * (col[k] - mean) : 0.0;
            }
        }
    });
    if (status_out) std::copy(status.begin(), status.end(), status_out);
    if (lay.E + lay.Et > 0) extra_rows(base, lay, vars, runs, x0, J, r0, threads);
}

// normal_equations_host, include/mikroptim/optim/normal_equations.hpp:19-56
inline double gram_dot(const double* u, const double* v, int64_t n) {
    double acc = 0.0;
    for (int64_t r = 0; r < n; ++r) acc = std::fma(u[r], v[r], acc);
    return acc;
}
void normal_equations(const double* J, const double* r, int P, int64_t R, double* a, double* g) {
    for (int q = 0; q < P; ++q) {
        const double* cq = J + size_t(q) * R;
        for (int p = 0; p <= q; ++p) {
            const double v = gram_dot(J + size_t(p) * R, cq, R);
            a[size_t(q) * P + p] = v;
            a[size_t(p) * P + q] = v;
        }
        g[q] = gram_dot(cq, r, R);
    }
}

// ---- Levenberg-Marquardt control logic: src/optim/levenberg_marquardt.cpp:37-572 ------------

struct Backend {
    // a[j*P*P + q*P + p], g[j*P + p]
    std::function<void(const std::vector<int>&, const std::vector<double>&, std::vector<double>&,
                       std::vector<double>&)>
        normal_equations;
    std::function<void(const std::vector<int>&, const std::vector<double>&, std::vector<double>&,
                       std::vector<int64_t>&, std::vector<int64_t>&)>
        merit;
};

struct Bounds {
    const double *lower, *upper, *typical;
};

struct RunState {
    std::vector<double> x, trial, a, g, d;
    double merit = 0.0;
    int64_t failed = 0, hard_failed = 0;
    double mu = 0.0, nu = 2.0;
    int status = MKO_LM_RUNNING;
    bool need_jacobian = true, rejected = false, solved = false;
    double step_norm = 0.0, predicted = 0.0, step_damping = 0.0, gradient_norm = 0.0;
    double model_decrease = kNaN;
    int iterations = 0, accepted = 0, jacobians = 0;
    std::vector<double> history;
};

constexpr double kModelDamping = 1e-12;

inline double scale_of(double x, double typical) { return std::max(std::abs(x), typical); }

bool cholesky(std::vector<double>& m, int n) {
    const size_t N = size_t(n);
    for (size_t j = 0; j < N; ++j) {
        double diag = m[j * N + j];
        for (size_t k = 0; k < j; ++k) diag -= m[k * N + j] * m[k * N + j];
        if (!(diag > 0.0) || !std::isfinite(diag)) return false;
        const double ljj = std::sqrt(diag);
        m[j * N + j] = ljj;
        for (size_t i = j + 1; i < N; ++i) {
            double s = m[j * N + i];
            for (size_t k = 0; k < j; ++k) s -= m[k * N + i] * m[k * N + j];
            m[j * N + i] = s / ljj;
        }
    }
    return true;
}

void cholesky_solve(const std::vector<double>& l, int n, std::vector<double>& b) {
    const size_t N = size_t(n);
    for (size_t i = 0; i < N; ++i) {
        double s = b[i];
        for (size_t k = 0; k < i; ++k) s -= l[k * N + i] * b[k];
        b[i] = s / l[i * N + i];
    }
    for (size_t i = N; i-- > 0;) {
        double s = b[i];
        for (size_t k = i + 1; k < N; ++k) s -= l[i * N + k] * b[k];
        b[i] = s / l[i * N + i];
    }
}

inline bool bound_blocks(const Bounds& b, size_t k, double x, double dir) {
    if (b.lower[k] == b.upper[k]) return true;
    return (x <= b.lower[k] && dir < 0.0) || (x >= b.upper[k] && dir > 0.0);
}

void prepare_step(RunState& s, const Bounds& b, int P, const mko_lm_options& o) {
    const size_t n = size_t(P);
    double smax = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const double sk = scale_of(s.x[k], b.typical[k]);
        smax = std::max(smax, s.a[k * n + k] * sk * sk);
    }
    double gmax = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const double sk = scale_of(s.x[k], b.typical[k]);
        const double akk = s.a[k * n + k];
        const double floor = smax > 0.0 ? o.diag_floor * smax / (sk * sk) : 1.0 / (sk * sk);
        s.d[k] = std::max(akk, floor);
        if (!bound_blocks(b, k, s.x[k], -s.g[k])) {
            const double denom = std::sqrt(akk * s.merit);
            if (denom > 0.0) gmax = std::max(gmax, std::abs(s.g[k]) / denom);
        }
    }
    s.gradient_norm = gmax;
}

void solve_step(RunState& s, const Bounds& b, int P, double mu) {
    const size_t n = size_t(P);
    std::vector<char> fixed(n);
    for (size_t k = 0; k < n; ++k) fixed[k] = bound_blocks(b, k, s.x[k], -s.g[k]) ? 1 : 0;
    std::vector<double> delta(n, 0.0), m, rhs;
    std::vector<size_t> idx;
    s.solved = true;
    s.step_damping = mu;
    for (size_t round = 0; round <= n; ++round) {
        idx.clear();
        for (size_t k = 0; k < n; ++k)
            if (!fixed[k]) idx.push_back(k);
        const size_t f = idx.size();
        std::fill(delta.begin(), delta.end(), 0.0);
        if (f == 0) break;
        m.assign(f * f, 0.0);
        rhs.assign(f, 0.0);
        for (size_t j = 0; j < f; ++j) {
            for (size_t i = j; i < f; ++i) m[j * f + i] = s.a[idx[j] * n + idx[i]];
            m[j * f + j] += mu * s.d[idx[j]];
            rhs[j] = -s.g[idx[j]];
        }
        if (!cholesky(m, int(f))) {
            s.solved = false;
            return;
        }
        cholesky_solve(m, int(f), rhs);
        for (size_t j = 0; j < f; ++j) delta[idx[j]] = rhs[j];
        bool changed = false;
        for (const size_t k : idx) {
            if (bound_blocks(b, k, s.x[k], delta[k])) {
                fixed[k] = 1;
                changed = true;
            }
        }
        if (!changed) break;
    }
    double nrm = 0.0;
    std::vector<double> step(n);
    for (size_t k = 0; k < n; ++k) {
        const double t = std::min(std::max(s.x[k] + delta[k], b.lower[k]), b.upper[k]);
        s.trial[k] = t;
        step[k] = t - s.x[k];
        nrm = std::max(nrm, std::abs(step[k]) / scale_of(s.x[k], b.typical[k]));
    }
    s.step_norm = nrm;
    double lin = 0.0, quad = 0.0;
    for (size_t q = 0; q < n; ++q) {
        lin += step[q] * s.g[q];
        double aq = 0.0;
        for (size_t p = 0; p < n; ++p) aq += s.a[q * n + p] * step[p];
        quad += step[q] * aq;
    }
    s.predicted = -(2.0 * lin + quad);
}

inline bool all_finite(const double* v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

// levenberg_marquardt, src/optim/levenberg_marquardt.cpp:332-572
std::vector<RunState> levenberg_marquardt(Backend& backend, int P, const Bounds& bounds,
                                          const double* x0, int runs, const mko_lm_options& o,
                                          int* cycles_out) {
    const size_t Pn = size_t(P);
    std::vector<RunState> state(static_cast<size_t>(runs));
    std::vector<int> list;
    std::vector<double> xs, buf_a, buf_g, buf_f;
    std::vector<int64_t> buf_failed, buf_hard;
    auto gather = [&](bool trial) {
        xs.resize(list.size() * Pn);
        for (size_t j = 0; j < list.size(); ++j) {
            const RunState& s = state[size_t(list[j])];
            const std::vector<double>& src = trial ? s.trial : s.x;
            std::copy(src.begin(), src.end(), xs.begin() + std::ptrdiff_t(j * Pn));
        }
    };
    auto evaluate = [&](bool trial) {
        gather(trial);
        buf_f.assign(list.size(), 0.0);
        buf_failed.assign(list.size(), 0);
        buf_hard.assign(list.size(), 0);
        backend.merit(list, xs, buf_f, buf_failed, buf_hard);
    };
    for (int i = 0; i < runs; ++i) {
        RunState& s = state[size_t(i)];
        s.x.assign(x0 + size_t(i) * Pn, x0 + size_t(i + 1) * Pn);
        s.trial = s.x;
        s.a.assign(Pn * Pn, 0.0);
        s.g.assign(Pn, 0.0);
        s.d.assign(Pn, 0.0);
        s.mu = o.initial_damping;
        list.push_back(i);
    }
    evaluate(false);
    for (size_t j = 0; j < list.size(); ++j) {
        RunState& s = state[size_t(list[j])];
        s.merit = buf_f[j];
        s.failed = buf_failed[j];
        s.hard_failed = buf_hard[j];
        s.history.push_back(s.merit);
        if (!std::isfinite(s.merit))
            s.status = MKO_LM_NUMERICAL_ERROR;
        else if (s.merit == 0.0)
            s.status = MKO_LM_CONVERGED_MERIT;
    }
    int cycles = 0;
    for (;;) {
        list.clear();
        for (int i = 0; i < runs; ++i) {
            const RunState& s = state[size_t(i)];
            if (s.status == MKO_LM_RUNNING && s.need_jacobian && s.iterations < o.max_iterations)
                list.push_back(i);
        }
        if (!list.empty()) {
            gather(false);
            buf_a.assign(list.size() * Pn * Pn, 0.0);
            buf_g.assign(list.size() * Pn, 0.0);
            backend.normal_equations(list, xs, buf_a, buf_g);
            for (size_t j = 0; j < list.size(); ++j) {
                RunState& s = state[size_t(list[j])];
                std::copy_n(buf_a.data() + j * Pn * Pn, Pn * Pn, s.a.begin());
                std::copy_n(buf_g.data() + j * Pn, Pn, s.g.begin());
                s.need_jacobian = false;
                ++s.jacobians;
                if (!all_finite(buf_a.data() + j * Pn * Pn, Pn * Pn) ||
                    !all_finite(buf_g.data() + j * Pn, Pn)) {
                    s.status = MKO_LM_NUMERICAL_ERROR;
                    continue;
                }
                prepare_step(s, bounds, P, o);
            }
            parallel_for(0, int(list.size()), o.threads, [&](int j) {
                RunState& s = state[size_t(list[size_t(j)])];
                if (s.status == MKO_LM_RUNNING) solve_step(s, bounds, P, kModelDamping);
            });
            for (const int i : list) {
                RunState& s = state[size_t(i)];
                if (s.status != MKO_LM_RUNNING) continue;
                s.model_decrease = s.solved ? s.predicted / s.merit : kNaN;
                if (s.gradient_norm <= o.gtol)
                    s.status = MKO_LM_CONVERGED_GRADIENT;
                else if (s.solved && s.predicted >= 0.0 && s.predicted <= o.ftol * s.merit)
                    s.status = MKO_LM_CONVERGED_MERIT;
                else if (s.solved && s.step_norm <= o.xtol)
                    s.status = MKO_LM_CONVERGED_STEP;
            }
        }
        list.clear();
        for (int i = 0; i < runs; ++i) {
            RunState& s = state[size_t(i)];
            if (s.status != MKO_LM_RUNNING) continue;
            if (s.iterations >= o.max_iterations) {
                s.status = MKO_LM_MAX_ITERATIONS;
                continue;
            }
            list.push_back(i);
        }
        if (list.empty()) break;
        ++cycles;
        parallel_for(0, int(list.size()), o.threads, [&](int j) {
            RunState& s = state[size_t(list[size_t(j)])];
            solve_step(s, bounds, P, s.mu);
        });
        std::vector<int> cycle_runs = list, eval;
        for (const int i : cycle_runs) {
            RunState& s = state[size_t(i)];
            if (s.solved && s.rejected && s.step_norm <= o.xtol)
                s.status = MKO_LM_CONVERGED_STEP;
            else if (s.solved)
                eval.push_back(i);
        }
        list = eval;
        if (!list.empty()) evaluate(true);
        size_t e = 0;
        for (const int i : cycle_runs) {
            RunState& s = state[size_t(i)];
            if (s.status == MKO_LM_CONVERGED_STEP) continue;
            ++s.iterations;
            bool accept = false;
            if (s.solved) {
                const double f_new = buf_f[e];
                const int64_t failed_new = buf_failed[e];
                const int64_t hard_new = buf_hard[e];
                ++e;
                const double rho = (s.merit - f_new) / s.predicted;
                accept = std::isfinite(f_new) && s.predicted > 0.0 && f_new < s.merit &&
                         !(o.reject_new_failures && hard_new > s.hard_failed);
                if (accept) {
                    s.x = s.trial;
                    s.merit = f_new;
                    s.failed = failed_new;
                    s.hard_failed = hard_new;
                    const double t = 2.0 * rho - 1.0;
                    s.mu *= std::max(1.0 / 3.0, 1.0 - t * t * t);
                    s.nu = 2.0;
                    s.need_jacobian = true;
                    s.rejected = false;
                    ++s.accepted;
                    if (f_new == 0.0) s.status = MKO_LM_CONVERGED_MERIT;
                }
            }
            if (!accept) {
                s.mu *= s.nu;
                s.nu *= 2.0;
                s.rejected = true;
                if (s.mu > o.max_damping) s.status = MKO_LM_DAMPING_LIMIT;
            }
            s.history.push_back(s.merit);
            if (o.stall_window > 0 && s.status == MKO_LM_RUNNING) {
                const size_t w = size_t(o.stall_window);
                if (s.history.size() > w) {
                    const double old = s.history[s.history.size() - 1 - w];
                    if (old - s.merit <= o.stall_tolerance * old) s.status = MKO_LM_CONVERGED_STALL;
                }
            }
        }
    }
    if (cycles_out) *cycles_out = cycles;
    return state;
}

// CpuLmBackend, src/optim/levenberg_marquardt.cpp:678-888 (constraint rows from the settings)
Backend cpu_backend(const Batch& base, const Layout& lay, const Vars& vars, int method,
                    int threads) {
    Backend be;
    const int P = vars.P;
    be.normal_equations = [=, &base, &lay, &vars](const std::vector<int>& runs,
                                                 const std::vector<double>& x,
                                                 std::vector<double>& a, std::vector<double>& g) {
        const int count = int(runs.size());
        OwnedBatch tmpl;
        const mko_batch* tb = &base.b;
        if (base.b.variants != 1) {
            tmpl.init(base.b, count);
            for (int j = 0; j < count; ++j) tmpl.copy_variant(base.b, runs[size_t(j)], j);
            tb = &tmpl.b;
        }
        const Batch TB{*tb};
        const int64_t R = lay.rows();
        std::vector<double> J(size_t(count) * P * R), r0(size_t(count) * R);
        if (method == MKO_JACOBIAN_ADJOINT)
            adjoint_jacobian(TB, lay, vars, count, x.data(), J.data(), r0.data(), nullptr, threads);
        else
            fd_jacobian(TB, lay, vars, count, x.data(), J.data(), r0.data(), nullptr, nullptr,
                        threads);
        parallel_for(0, count, threads, [&](int j) {
            normal_equations(J.data() + size_t(j) * P * R, r0.data() + size_t(j) * R, P, R,
                             a.data() + size_t(j) * P * P, g.data() + size_t(j) * P);
        });
    };
    be.merit = [=, &base, &lay, &vars](const std::vector<int>& runs, const std::vector<double>& x,
                                      std::vector<double>& merit, std::vector<int64_t>& failed,
                                      std::vector<int64_t>& hard) {
        const int count = int(runs.size());
        OwnedBatch pb;
        pb.init(base.b, count);
        for (int j = 0; j < count; ++j) {
            pb.copy_variant(base.b, base.b.variants == 1 ? 0 : runs[size_t(j)], j);
            apply(vars, pb.b, pb.target(), j, x.data() + size_t(j) * P);
        }
        const Batch PB{pb.b};
        parallel_for(0, count, threads, [&](int v) {
            variant_merit_rows(PB, v, lay, merit[size_t(v)], failed[size_t(v)], hard[size_t(v)]);
        });
    };
    return be;
}

// ---- FP32 prefilter: include/mikroptim/merit/prefilter_ray.hpp, src/batch/batch_prefilter.cpp --
// The float ray of the prefilter (hard failure) with its boundary margin (the distance of the
// tracer's checks from their failure boundary, as a cosine), per-variant stats in float summed
// per tuple in double, and the conservative decision in double.
namespace pf {

struct V3f {
    float x, y, z;
};
inline V3f operator+(V3f a, V3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3f operator-(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3f operator*(float s, V3f a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3f cross(V3f a, V3f b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
struct RayF {
    V3f p, d;
    float opl;
};

// prefilter_min_margin / prefilter_failure_linear / prefilter_failure_root (prefilter_ray.hpp:36-56)
inline float min_margin(float acc, float x) {
    if (x < acc) return x;
    return x >= acc ? acc : 0.0f;
}
inline float failure_linear(float x) { return -x > 0.0f ? -x : 0.0f; }
inline float failure_root(float q) { return -q > 0.0f ? std::sqrt(-q) : 0.0f; }

struct TableF {
    int S;
    const uint8_t* type;
    float c[MKO_MAX_SURFACES], t[MKO_MAX_SURFACES], n[MKO_MAX_SURFACES]; // one wavelength
};

inline float curv_of(const TableF& tb, int s) {
    return tb.type[s] == MKO_SURFACE_SPHERE ? tb.c[s] : 0.0f;
}

struct RayResult {
    Status status;
    float opl, margin;
};

// prefilter_ray (prefilter_ray.hpp:171-249) over the float surface table of one wavelength; the
// trace steps are those of surface_ops.hpp:122-203 in float, hard failure.
RayResult prefilter_ray(const TableF& tb, RayF ray, V3f center, int kind, float kappa) {
    RayResult out{Ok, 0.0f, 0.0f};
    float margin = 1.0f;
    for (int s = 1; s < tb.S; ++s) {
        const float c = curv_of(tb, s);
        const float nb = tb.n[s - 1], na = tb.n[s];
        if (!(ray.d.z > 0.0f)) {
            out.status = Backward;
            out.margin = failure_linear(ray.d.z);
            return out;
        }
        const V3f d_before = ray.d;
        // transfer_to_tangent_plane
        ray.p.z -= tb.t[s - 1];
        const float s0 = -ray.p.z / ray.d.z;
        ray.p = ray.p + s0 * ray.d;
        ray.p.z = 0.0f;
        // intersect_sphere
        const float H = c * (ray.p.x * ray.p.x + ray.p.y * ray.p.y);
        const float B = ray.d.z - c * (ray.p.x * ray.d.x + ray.p.y * ray.d.y);
        const float D = B * B - c * H;
        const float denom = D < 0.0f ? 0.0f : B + std::sqrt(D);
        if (D < 0.0f || !(denom > 0.0f)) {
            const float a = failure_root(D), b = failure_linear(B);
            out.status = Missed;
            out.margin = a > b ? a : b;
            return out;
        }
        const float sh = H / denom;
        ray.p = ray.p + sh * ray.d;
        ray.opl += nb * (s0 + sh);
        // refract_sphere
        const float mu = nb / na;
        const V3f N{-c * ray.p.x, -c * ray.p.y, 1.0f - c * ray.p.z};
        const float cosI = dot(ray.d, N);
        const float k = 1.0f - mu * mu * (1.0f - cosI * cosI);
        if (k < 0.0f) {
            out.status = TIR;
            out.margin = failure_root(k);
            return out;
        }
        ray.d = mu * ray.d + (std::sqrt(k) - mu * cosI) * N;
        // PrefilterMarginObserver (prefilter_ray.hpp:68-81)
        margin = min_margin(margin, d_before.z);
        margin = min_margin(margin, dot(d_before, N));
        margin = min_margin(margin, dot(ray.d, N));
    }
    if (!(ray.d.z > 0.0f)) {
        out.status = Backward;
        out.margin = failure_linear(ray.d.z);
        return out;
    }
    margin = min_margin(margin, ray.d.z);
    const float n_image = tb.n[tb.S - 1];
    if (kind != MKO_REF_LAST_SURFACE) {
        // image_sphere_step (reference_sphere.hpp:110-143), hard
        const V3f q = ray.p - center;
        const V3f ci = cross(q, ray.d);
        const float disc_i = 1.0f + kappa * kappa * (-dot(ci, ci));
        const float b = -dot(q, ray.d);
        const float t = -dot(ci, ci);
        const float disc = 1.0f + kappa * kappa * t;
        if (!(disc >= 0.0f)) {
            out.status = Missed;
            out.margin = disc_i < 0.0f ? failure_root(disc_i) : 0.0f;
            return out;
        }
        const float step = b - kappa * t / (std::sqrt(disc) + 1.0f);
        if (!(step >= 0.0f) && !(step < 0.0f)) {
            out.status = Missed;
            out.margin = disc_i < 0.0f ? failure_root(disc_i) : 0.0f;
            return out;
        }
        ray.opl += n_image * step;
        out.opl = ray.opl;
        out.margin = min_margin(margin, std::sqrt(disc_i));
        return out;
    }
    // the reference sphere's checks (reference_sphere_step, reference_sphere.hpp:56-90)
    const float r2 = dot(center, center);
    const float inv_r = r2 > 0.0f ? 1.0f / std::sqrt(r2) : 1.0f;
    const V3f q = ray.p - center;
    const float bq = -dot(q, ray.d);
    const float g = dot(ray.p, ray.p - 2.0f * center);
    const float disc = bq * bq - g;
    const float sigma = center.z >= 0.0f ? 1.0f : -1.0f;
    const float root = disc > 0.0f ? std::sqrt(disc) : 0.0f;
    const float den = sigma * (bq + sigma * root) * inv_r;
    bool ok = !(disc < 0.0f);
    float s_ref = 0.0f;
    if (ok) {
        const float rt = std::sqrt(disc);
        const bool real_image = center.z >= 0.0f;
        const float dn = real_image ? bq + rt : bq - rt;
        ok = real_image ? dn > 0.0f : dn < 0.0f;
        if (ok) s_ref = g / dn;
    }
    if (!ok) {
        const float a = failure_root(disc * inv_r * inv_r);
        const float c = disc >= 0.0f ? failure_linear(den) : 0.0f;
        out.status = Missed;
        out.margin = a > c ? a : c;
        return out;
    }
    ray.opl += n_image * s_ref;
    margin = min_margin(margin, root * inv_r);
    margin = min_margin(margin, den);
    out.opl = ray.opl;
    out.margin = margin;
    return out;
}

struct Stats {
    double merit = 0.0, optical = 0.0;
    int64_t failed = 0, uncertain_ok = 0, uncertain_failed = 0;
};

// variant_stats (src/batch/batch_prefilter.cpp:151-222): float quadrature, float source ray,
// classes 0 certain valid / 1 certain failed / 2, 3 uncertain, piston over the certain-valid
// rays, tuple sums in float added in double.
Stats variant_stats(const Batch& B, int v, const Layout& lay, double boundary) {
    const int S = B.S(), L = B.L();
    const float eta = float(boundary);
    Stats out;
    double merit = 0.0;
    float optical = 0.0f;
    TableF tb;
    tb.S = S;
    tb.type = B.b.type;
    for (int f = 0; f < lay.F; ++f) {
        const Quadrature& qd = lay.quad[f];
        const int n = lay.samples[f];
        const size_t nn = static_cast<size_t>(n);
        std::vector<float> opl(nn), wts(nn), rho(nn), phi(nn);
        std::vector<uint8_t> cls(nn);
        for (int k = 0; k < n; ++k) {
            // sample() of PupilQuadrature<float> (quadrature.hpp:74-95)
            const int i = k / qd.spokes, j = k - i * qd.spokes;
            const float pi = float(3.14159265358979323846264338327950288);
            rho[size_t(k)] = std::sqrt(float(qd.u[i]));
            wts[size_t(k)] = float(qd.w[i]) / float(qd.spokes);
            phi[size_t(k)] = qd.mode == MKO_PUPIL_HALF   ? float(2 * j + 1) * pi / float(2 * qd.spokes)
                             : qd.mode == MKO_PUPIL_FULL ? float(2 * j + 1) * pi / float(qd.spokes)
                                                         : 0.0f;
        }
        for (int l = 0; l < L; ++l) {
            for (int s = 0; s < S; ++s) {
                tb.c[s] = float(B.b.curvature[size_t(v) * S + s]);
                tb.t[s] = float(B.b.thickness[size_t(v) * S + s]);
                tb.n[s] = float(B.b.index[(size_t(v) * S + s) * L + l]);
            }
            // prefilter_opl_offset (prefilter_ray.hpp:264-272)
            float offset = 0.0f;
            for (int s = 0; s + 1 < S; ++s) offset += tb.n[s] * tb.t[s];
            const size_t vf = size_t(v) * B.F() + f;
            const V3f center{0.0f, float(B.b.field_target[vf]), tb.t[S - 1]};
            const float kappa =
                B.b.reference_sphere != MKO_REF_LAST_SURFACE ? float(B.b.reference_curvature[vf]) : 0.0f;
            float weight_ok = 0.0f, weighted_opl = 0.0f;
            for (int k = 0; k < n; ++k) {
                RayF ray;
                if (B.b.source == MKO_SOURCE_POINT) {
                    const float sin_max = float(B.b.source_na) / tb.n[0];
                    const float sr = sin_max * rho[size_t(k)];
                    ray.p = V3f{float(B.b.source_point[0]), float(B.b.source_point[1]),
                                float(B.b.source_point[2])};
                    ray.d = V3f{sr * std::sin(phi[size_t(k)]), sr * std::cos(phi[size_t(k)]),
                                std::sqrt((1.0f - sr) * (1.0f + sr))};
                    ray.opl = 0.0f;
                } else {
                    const float a = float(B.b.pupil_radius[v]);
                    const float th = float(B.b.field_angle[vf]);
                    const float r = a * rho[size_t(k)];
                    ray.p = V3f{r * std::sin(phi[size_t(k)]), r * std::cos(phi[size_t(k)]), 0.0f};
                    ray.d = V3f{0.0f, std::sin(th), std::cos(th)};
                    ray.opl = dot(ray.p, ray.d);
                }
                ray.opl -= offset;
                const RayResult rr = prefilter_ray(tb, ray, center, B.b.reference_sphere, kappa);
                const bool ok = rr.status == Ok;
                const bool certain = rr.margin > eta;
                const uint8_t c = uint8_t(ok ? (certain ? 0 : 2) : (certain ? 1 : 3));
                cls[size_t(k)] = c;
                opl[size_t(k)] = rr.opl;
                if (c == 0) {
                    weight_ok += wts[size_t(k)];
                    weighted_opl += wts[size_t(k)] * rr.opl;
                    const float len = std::abs(offset) + std::abs(rr.opl);
                    optical = len > optical ? len : optical;
                }
                out.failed += c == 1;
                out.uncertain_ok += c == 2;
                out.uncertain_failed += c == 3;
            }
            const bool valid = weight_ok > 0.0f;
            const float piston = valid ? weighted_opl / weight_ok : 0.0f;
            const float scale = float(residual_scale(B, f, l));
            float sum = 0.0f;
            for (int k = 0; k < n; ++k) {
                float e = 0.0f;
                if (valid && cls[size_t(k)] == 0) e = opl[size_t(k)] - piston;
                const float r = scale * std::sqrt(wts[size_t(k)]) * e;
                sum += r * r;
            }
            merit += double(sum);
        }
    }
    out.merit = merit;
    out.optical = double(optical);
    return out;
}

} // namespace pf

thread_local std::string g_error;

} // namespace orc

using namespace orc;

#define ORC_TRY(...)                                                                              \
    try {                                                                                          \
        __VA_ARGS__;                                                                               \
        return 0;                                                                                  \
    } catch (const std::exception& e) {                                                            \
        orc::g_error = e.what();                                                                   \
        return -1;                                                                                 \
    }

// ---- df64 (double-float) arithmetic: include/mikroptim/math/df64.hpp:91-271 -------------------
// Every float operation rounded on its own (this file is built with -ffp-contract=off); the
// products that the algorithms fuse use fmaf.
namespace orc_df64 {

struct D {
    float hi, lo;
};
inline void two_sum(float a, float b, float& s, float& e) { // df64.hpp:91-95
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
inline void fast_two_sum(float a, float b, float& s, float& e) { // df64.hpp:97-100
    s = a + b;
    e = b - (s - a);
}
inline void two_prod(float a, float b, float& p, float& e) { // df64.hpp:101-104
    p = a * b;
    e = std::fmaf(a, b, -p);
}
inline D add(D a, D b) { // AccurateDWPlusDW, df64.hpp:160-170
    float sh, sl, th, tl, vh, vl, zh, zl;
    two_sum(a.hi, b.hi, sh, sl);
    two_sum(a.lo, b.lo, th, tl);
    const float c = sl + th;
    fast_two_sum(sh, c, vh, vl);
    const float w = tl + vl;
    fast_two_sum(vh, w, zh, zl);
    return {zh, zl};
}
inline D neg(D a) { return {-a.hi, -a.lo}; }
inline D mul(D a, D b) { // DWTimesDW3, df64.hpp:177-187
    float ch, cl1, zh, zl;
    two_prod(a.hi, b.hi, ch, cl1);
    const float tl0 = a.lo * b.lo;
    const float tl1 = std::fmaf(a.hi, b.lo, tl0);
    const float cl2 = std::fmaf(a.lo, b.hi, tl1);
    const float cl3 = cl1 + cl2;
    fast_two_sum(ch, cl3, zh, zl);
    return {zh, zl};
}
inline D mul_fp(D a, float b) { // DWTimesFP3, df64.hpp:190-196
    float ch, cl1, zh, zl;
    two_prod(a.hi, b, ch, cl1);
    const float cl3 = std::fmaf(a.lo, b, cl1);
    fast_two_sum(ch, cl3, zh, zl);
    return {zh, zl};
}
inline D div(D a, D b) { // DWDivDW2, df64.hpp:204-215
    const float th = a.hi / b.hi;
    const D r = mul_fp(b, th);
    const float pi_h = a.hi - r.hi;
    const float d_l = a.lo - r.lo;
    const float d = pi_h + d_l;
    const float tl = d / b.hi;
    float zh, zl;
    fast_two_sum(th, tl, zh, zl);
    return {zh, zl};
}
inline D sqrt(D a) { // SQRTDWtoDW, df64.hpp:259-271
    if (!(a.hi > 0.0f) || a.hi > FLT_MAX) return {std::sqrt(a.hi), 0.0f};
    const float sh = std::sqrt(a.hi);
    const float rho1 = std::fmaf(-sh, sh, a.hi);
    const float rho2 = a.lo + rho1;
    const float sl = rho2 / (sh + sh);
    float zh, zl;
    fast_two_sum(sh, sl, zh, zl);
    return {zh, zl};
}

} // namespace orc_df64

extern "C" {

const char* orc_last_error(void) { return orc::g_error.c_str(); }

// Element-wise df64 operation over n pairs (hi, lo): op 0 a + b, 1 a - b, 2 a * b, 3 a * b.hi
// (DWTimesFP3), 4 a / b, 5 sqrt(a). a, b, out: [n][2] floats.
int orc_df64_ops(int op, int64_t n, const float* a, const float* b, float* out) {
    if (op < 0 || op > 5 || n < 0) return -1;
    using namespace orc_df64;
    for (int64_t i = 0; i < n; ++i) {
        const D x{a[2 * i], a[2 * i + 1]};
        const D y = op == 5 ? D{0.0f, 0.0f} : D{b[2 * i], b[2 * i + 1]};
        D z{};
        switch (op) {
        case 0: z = add(x, y); break;
        case 1: z = add(x, neg(y)); break;
        case 2: z = mul(x, y); break;
        case 3: z = mul_fp(x, y.hi); break;
        case 4: z = div(x, y); break;
        default: z = sqrt(x); break;
        }
        out[2 * i] = z.hi;
        out[2 * i + 1] = z.lo;
    }
    return 0;
}

int orc_gauss_legendre(int n, double* u, double* w) {
    ORC_TRY({
        const Quadrature q = make_quadrature(n, 1, MKO_PUPIL_HALF);
        for (int i = 0; i < n; ++i) {
            u[i] = q.u[i];
            w[i] = q.w[i];
        }
    })
}

// pupil samples of one field mode: rho, phi, weight per index k = i*M + j
int orc_pupil_samples(int rings, int spokes, int mode, double* rho, double* phi, double* weight) {
    ORC_TRY({
        const Quadrature q = make_quadrature(rings, spokes, mode);
        for (int k = 0; k < q.rings * q.spokes; ++k) {
            const Sample s = sample(q, k);
            rho[k] = s.rho;
            phi[k] = s.phi;
            weight[k] = s.weight;
        }
    })
}

int64_t orc_rays_per_variant(const mko_batch* b, const mko_residual_settings* s) {
    try {
        return make_layout(Batch{*b}, *s).rays;
    } catch (const std::exception& e) {
        orc::g_error = e.what();
        return -1;
    }
}

// One ray through one variant's table, start ray given; returns the status, fills final state.
int orc_trace_ray(const mko_batch* b, int v, int l, const double* p0, const double* d0,
                  double opl0, int to_reference, int f, int soft_mode, double* p_out,
                  double* d_out, double* opl_out, double* penalty_out) {
    const Batch B{*b};
    const Table tb = B.table(v);
    Ray ray{{p0[0], p0[1], p0[2]}, {d0[0], d0[1], d0[2]}, opl0};
    Soft soft;
    Status st;
    if (to_reference)
        st = trace_to_reference(tb, l, ray, reference_of(B, tb, v, f), soft_mode ? &soft : nullptr,
                                nullptr);
    else
        st = trace_ray(tb, l, ray, soft_mode ? &soft : nullptr, nullptr, nullptr);
    p_out[0] = ray.p.x;
    p_out[1] = ray.p.y;
    p_out[2] = ray.p.z;
    d_out[0] = ray.d.x;
    d_out[1] = ray.d.y;
    d_out[2] = ray.d.z;
    *opl_out = ray.opl;
    if (penalty_out) *penalty_out = soft.penalty;
    return int(st);
}

int orc_residuals(const mko_batch* b, const mko_residual_settings* s, double* residuals,
                  double* opd, uint8_t* status, int threads) {
    ORC_TRY({
        const Batch B{*b};
        const Layout lay = make_layout(B, *s);
        residuals_all(B, lay, residuals, opd, status, threads);
    })
}

int orc_merit(const mko_batch* b, const mko_residual_settings* s, double* merit, int64_t* failed,
              int64_t* hard_failed, int threads) {
    ORC_TRY({
        const Batch B{*b};
        const Layout lay = make_layout(B, *s);
        parallel_for(0, b->variants, threads, [&](int v) {
            variant_merit_rows(B, v, lay, merit[v], failed[v], hard_failed[v]);
        });
    })
}

// Surface extrema [V][S][2] (surface_extrema_cpu, src/batch/batch_extrema.cpp:75-92).
int orc_surface_extrema(const mko_batch* b, const mko_residual_settings* s, double* extrema,
                        int threads) {
    ORC_TRY({
        const Batch B{*b};
        const Layout lay = make_layout(B, *s);
        const size_t S = size_t(b->surfaces);
        parallel_for(0, b->variants, threads,
                     [&](int v) { variant_extrema(B, v, lay, extrema + size_t(v) * S * 2); });
    })
}

// Paraxial data of variant v at lambda_ref: out = {A, B, C, D, power, petzval_curvature}
// (batch_paraxial, src/batch/system_batch.cpp:965; compute_paraxial, src/paraxial/paraxial.cpp:98).
int orc_paraxial(const mko_batch* b, int v, double* out) {
    ORC_TRY({
        const Paraxial p = batch_paraxial(Batch{*b}, v);
        out[0] = p.A;
        out[1] = p.B;
        out[2] = p.C;
        out[3] = p.D;
        out[4] = p.power;
        out[5] = p.petzval_curvature;
    })
}

// Rows of the constraints in s->extra for every variant [V][E + E_t] (row order: untraced,
// then traced); extrema [V][S][2] as given (NULL: from each variant's own trace).
int orc_constraint_rows(const mko_batch* b, const mko_residual_settings* s, const double* extrema,
                        double* rows, int threads) {
    ORC_TRY({
        const Batch B{*b};
        const Layout lay = make_layout(B, *s);
        const size_t E = size_t(lay.E + lay.Et), S = size_t(b->surfaces);
        parallel_for(0, b->variants, threads, [&](int v) {
            if (extrema)
                variant_rows(B, v, lay, extrema + size_t(v) * S * 2, rows + size_t(v) * E);
            else
                variant_rows_traced(B, v, lay, rows + size_t(v) * E);
        });
    })
}

// Rows per variant of the residual vector: trace rows + E + E_t.
int64_t orc_residual_rows(const mko_batch* b, const mko_residual_settings* s) {
    try {
        return make_layout(Batch{*b}, *s).rows();
    } catch (const std::exception& e) {
        orc::g_error = e.what();
        return -1;
    }
}

// Prefilter stats per variant (prefilter_stats_cpu, src/batch/batch_prefilter.cpp:384):
// merit F32, optical length, certain-failed, uncertain-valid, uncertain-failed counts.
int orc_prefilter_stats(const mko_batch* b, const mko_residual_settings* s, double boundary,
                        double* merit, double* optical, int64_t* failed, int64_t* uncertain_ok,
                        int64_t* uncertain_failed, int threads) {
    ORC_TRY({
        if (!(boundary >= 0.0) || !std::isfinite(boundary))
            throw std::invalid_argument("boundary margin must be finite and >= 0");
        const Batch B{*b};
        const Layout lay = make_layout(B, *s);
        parallel_for(0, b->variants, threads, [&](int v) {
            const pf::Stats st = pf::variant_stats(B, v, lay, boundary);
            merit[v] = st.merit;
            optical[v] = st.optical;
            failed[v] = st.failed;
            uncertain_ok[v] = st.uncertain_ok;
            uncertain_failed[v] = st.uncertain_failed;
        });
    })
}

int orc_fd_pattern(const mko_variables* vars, int surfaces, const double* x0, double* nominal,
                   double* plus, double* minus, double* step_plus, double* step_minus,
                   uint8_t* stencil) {
    ORC_TRY({
        const Vars V = make_vars(vars, surfaces);
        const FdPattern p = fd_pattern(V, x0);
        for (int k = 0; k < V.P; ++k) {
            nominal[k] = p.nominal[k];
            plus[k] = p.plus[k];
            minus[k] = p.minus[k];
            step_plus[k] = p.step_plus[k];
            step_minus[k] = p.step_minus[k];
            stencil[k] = p.stencil[k];
        }
    })
}

int orc_fd_jacobian(const mko_batch* base, const mko_residual_settings* s,
                    const mko_variables* vars, int runs, const double* x0, double* J, double* r0,
                    int64_t* status_changes, uint8_t* status0, int threads) {
    ORC_TRY({
        const Batch B{*base};
        const Layout lay = make_layout(B, *s);
        fd_jacobian(B, lay, make_vars(vars, base->surfaces), runs, x0, J, r0, status_changes,
                    status0, threads);
    })
}

int orc_adjoint_jacobian(const mko_batch* base, const mko_residual_settings* s,
                         const mko_variables* vars, int runs, const double* x0, double* J,
                         double* r0, uint8_t* status, int threads) {
    ORC_TRY({
        const Batch B{*base};
        const Layout lay = make_layout(B, *s);
        adjoint_jacobian(B, lay, make_vars(vars, base->surfaces), runs, x0, J, r0, status, threads);
    })
}

int orc_normal_equations(const double* J, const double* r, int runs, int P, int64_t R, double* a,
                         double* g) {
    ORC_TRY({
        for (int j = 0; j < runs; ++j)
            normal_equations(J + size_t(j) * P * R, r + size_t(j) * R, P, R,
                             a + size_t(j) * P * P, g + size_t(j) * P);
    })
}

// Writes x0 of each run into a copy of the base variant (VariableSet::apply).
int orc_apply(mko_batch* b, const mko_variables* vars, int v, const double* x) {
    ORC_TRY({
        apply(make_vars(vars, b->surfaces), *b,
              Target{const_cast<double*>(b->curvature), const_cast<double*>(b->thickness),
                     const_cast<double*>(b->index), const_cast<double*>(b->buchdahl),
                     const_cast<double*>(b->paraxial_index)},
              v, x);
    })
}

uint64_t orc_splitmix64(uint64_t x) { return splitmix64(x); }
uint64_t orc_variant_seed(uint64_t seed, uint64_t j) { return variant_seed(seed, j); }
double orc_unit_uniform(uint64_t bits) { return unit_uniform(bits); }
uint64_t orc_mt19937_64(uint64_t seed, uint64_t discard) {
    std::mt19937_64 e(seed);
    e.discard(discard);
    return e();
}

// relative/absolute: [3] per MKO_VAR_* type; distribution 0 uniform, 1 Irwin-Hall
int orc_perturbed_vector(const mko_variables* vars, int surfaces, const double* x0,
                         const double* relative, const double* absolute, int distribution,
                         uint64_t seed, int first_is_base, uint64_t j, double* y) {
    ORC_TRY({
        perturbed_vector(make_vars(vars, surfaces), x0, relative, absolute, distribution, seed,
                         first_is_base, j, y);
    })
}

// optimize_lm_cpu (src/optim/levenberg_marquardt.cpp:891-983, 985-991): LM with the CPU backend,
// outer penalty rounds in soft mode. x_out [runs][P], run results [runs], history optional
// ([runs][max_iterations+1] merits, NaN-padded; first round only).
int orc_optimize_lm(const mko_batch* base, const mko_residual_settings* s,
                    const mko_variables* vars, int runs, const double* x0,
                    const mko_lm_options* options, double* x_out, mko_lm_run* run_out,
                    int* cycles_out) {
    ORC_TRY({
        const Batch B{*base};
        const Vars V = make_vars(vars, base->surfaces);
        const int P = V.P;
        const Bounds bounds{V.lower.data(), V.upper.data(), V.typical.data()};
        mko_residual_settings rs = *s;
        int total_cycles = 0;
        std::vector<double> weight_of(size_t(runs), rs.fail_weight);
        std::vector<int> rounds(size_t(runs), 1);
        std::vector<RunState> result;
        {
            const Layout lay = make_layout(B, rs);
            Backend be = cpu_backend(B, lay, V, options->jacobian_method, options->threads);
            int cyc = 0;
            result = levenberg_marquardt(be, P, bounds, x0, runs, *options, &cyc);
            total_cycles += cyc;
        }
        double weight = rs.fail_weight;
        if (rs.failure == MKO_FAILURE_SOFT && options->fail_weight_factor > 1.0 && weight > 0.0) {
            for (;;) {
                weight *= options->fail_weight_factor;
                if (!(weight <= options->max_fail_weight)) break;
                std::vector<int> again;
                std::vector<double> xs;
                for (int i = 0; i < runs; ++i) {
                    if (result[size_t(i)].failed > 0 &&
                        result[size_t(i)].status != MKO_LM_NUMERICAL_ERROR) {
                        again.push_back(i);
                        xs.insert(xs.end(), result[size_t(i)].x.begin(), result[size_t(i)].x.end());
                    }
                }
                if (again.empty()) break;
                rs.fail_weight = weight;
                OwnedBatch tmpl;
                const mko_batch* tb = &B.b;
                if (B.b.variants != 1) {
                    tmpl.init(B.b, int(again.size()));
                    for (size_t j = 0; j < again.size(); ++j) tmpl.copy_variant(B.b, again[j], int(j));
                    tb = &tmpl.b;
                }
                const Batch TB{*tb};
                const Layout lay = make_layout(TB, rs);
                Backend be = cpu_backend(TB, lay, V, options->jacobian_method, options->threads);
                int cyc = 0;
                std::vector<RunState> part = levenberg_marquardt(be, P, bounds, xs.data(),
                                                                 int(again.size()), *options, &cyc);
                total_cycles += cyc;
                for (size_t j = 0; j < again.size(); ++j) {
                    RunState& r = result[size_t(again[j])];
                    RunState& q = part[j];
                    r.x = q.x;
                    r.merit = q.merit;
                    r.failed = q.failed;
                    r.hard_failed = q.hard_failed;
                    r.status = q.status;
                    r.iterations += q.iterations;
                    r.accepted += q.accepted;
                    r.jacobians += q.jacobians;
                    r.mu = q.mu;
                    r.gradient_norm = q.gradient_norm;
                    r.model_decrease = q.model_decrease;
                    weight_of[size_t(again[j])] = weight;
                    ++rounds[size_t(again[j])];
                }
            }
        }
        for (int i = 0; i < runs; ++i) {
            const RunState& r = result[size_t(i)];
            std::copy(r.x.begin(), r.x.end(), x_out + size_t(i) * P);
            mko_lm_run& o = run_out[i];
            o.merit = r.merit;
            o.failed_rays = r.failed;
            o.hard_failed_rays = r.hard_failed;
            o.status = r.status;
            o.iterations = r.iterations;
            o.accepted = r.accepted;
            o.jacobians = r.jacobians;
            o.damping = r.mu;
            o.gradient_norm = r.gradient_norm;
            o.model_decrease = r.model_decrease;
            o.fail_weight = weight_of[size_t(i)];
            o.penalty_rounds = rounds[size_t(i)];
        }
        if (cycles_out) *cycles_out = total_cycles;
    })
}

// LM over an analytic residual function r(x) = M x - b (linear least squares; DenseBackend of
// tests/lm_test_systems.hpp:38): A = MtM and g = Mtr exactly, merit = |r|^2.
int orc_lm_linear(const double* M, const double* bvec, int rows, int P, const double* lower,
                  const double* upper, const double* typical, int runs, const double* x0,
                  const mko_lm_options* options, double* x_out, mko_lm_run* run_out) {
    ORC_TRY({
        Backend be;
        auto residual = [&](const double* x, std::vector<double>& r) {
            r.assign(size_t(rows), 0.0);
            for (int i = 0; i < rows; ++i) {
                double acc = -bvec[i];
                for (int k = 0; k < P; ++k) acc += M[size_t(i) * P + k] * x[k];
                r[size_t(i)] = acc;
            }
        };
        be.normal_equations = [&](const std::vector<int>& list, const std::vector<double>& x,
                                  std::vector<double>& a, std::vector<double>& g) {
            std::vector<double> J(size_t(P) * rows), r;
            for (int p = 0; p < P; ++p)
                for (int i = 0; i < rows; ++i) J[size_t(p) * rows + i] = M[size_t(i) * P + p];
            for (size_t j = 0; j < list.size(); ++j) {
                residual(x.data() + j * P, r);
                normal_equations(J.data(), r.data(), P, rows, a.data() + j * P * P,
                                 g.data() + j * P);
            }
        };
        be.merit = [&](const std::vector<int>& list, const std::vector<double>& x,
                       std::vector<double>& f, std::vector<int64_t>& fl, std::vector<int64_t>& hf) {
            std::vector<double> r;
            for (size_t j = 0; j < list.size(); ++j) {
                residual(x.data() + j * P, r);
                double s = 0.0;
                for (double v : r) s += v * v;
                f[j] = s;
                fl[j] = 0;
                hf[j] = 0;
            }
        };
        const Bounds bounds{lower, upper, typical};
        std::vector<RunState> res = levenberg_marquardt(be, P, bounds, x0, runs, *options, nullptr);
        for (int i = 0; i < runs; ++i) {
            std::copy(res[size_t(i)].x.begin(), res[size_t(i)].x.end(), x_out + size_t(i) * P);
            run_out[i].merit = res[size_t(i)].merit;
            run_out[i].status = res[size_t(i)].status;
            run_out[i].iterations = res[size_t(i)].iterations;
            run_out[i].accepted = res[size_t(i)].accepted;
            run_out[i].jacobians = res[size_t(i)].jacobians;
        }
    })
}

} // extern "C"

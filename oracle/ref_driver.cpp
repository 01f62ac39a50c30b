// C-ABI driver over the reference's own CPU hot-path code (TEST INFRASTRUCTURE ONLY). Built by
// oracle/Makefile together with the reference sources where they lie (never copied) into
// oracle/_ref/libmkoref.so. The entry points take the same plain-C batch/settings/variables
// structs as the oracle (mikroptim_types.h) and call residuals_cpu, merit_cpu, fd_jacobian_cpu,
// adjoint_jacobian_cpu, normal_equations_host, perturbed_vector, fd_pattern and optimize_lm_cpu
// of the reference, so tests can compare the oracle's restatement with the original bit for bit.

#include "mikroptim_hip.h"
#include "mikroptim_hip_reference.hpp"
#include "mikroptim_types.h"

#include "mikroptim/batch/batch_merit.hpp"
#include "mikroptim/math/df64.hpp"
#include "mikroptim/batch/batch_residuals.hpp"
#include "mikroptim/batch/system_batch.hpp"
#include "mikroptim/glass/buchdahl_fit.hpp"
#include "mikroptim/glass/glass_hull.hpp"
#include "mikroptim/optim/glass_variables.hpp"
#include "mikroptim/glass/dispersion.hpp"
#include "mikroptim/model/system.hpp"
#include "mikroptim/optim/adjoint_jacobian.hpp"
#include "mikroptim/optim/fd_jacobian.hpp"
#include "mikroptim/batch/batch_extrema.hpp"
#include "mikroptim/batch/batch_prefilter.hpp"
#include "mikroptim/optim/levenberg_marquardt.hpp"
#include "mikroptim/optim/mechanical_constraints.hpp"
#include "mikroptim/optim/normal_equations.hpp"
#include "mikroptim/optim/paraxial_residuals.hpp"
#include "mikroptim/optim/trace_extra_residuals.hpp"
#include "mikroptim/optim/variable_set.hpp"
#include "mikroptim/optim/variant_generator.hpp"
#include "mikroptim/pupil/quadrature.hpp"

#include <algorithm>
#include <memory>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

mko::BatchTopology topology_of(const mko_batch& b) {
    mko::BatchTopology t;
    t.surfaces = b.surfaces;
    t.wavelengths = b.wavelengths;
    t.fields = b.fields;
    t.source = b.source == MKO_SOURCE_POINT ? mko::SourceKind::Point : mko::SourceKind::Collimated;
    for (int s = 0; s < b.surfaces; ++s) t.type.push_back(static_cast<mko::SurfaceKind>(b.type[s]));
    for (int f = 0; f < b.fields; ++f) t.pupil_mode.push_back(static_cast<mko::PupilMode>(b.pupil_mode[f]));
    t.reference_sphere = static_cast<mko::ReferenceSphereKind>(b.reference_sphere);
    return t;
}

mko::SystemBatchHost batch_of(const mko_batch& b) {
    mko::SystemBatchHost h(topology_of(b), b.variants);
    const size_t V = size_t(b.variants), S = size_t(b.surfaces), L = size_t(b.wavelengths),
                 F = size_t(b.fields);
    h.reference_wavelength_nm = b.reference_wavelength_nm;
    h.source_point = {b.source_point[0], b.source_point[1], b.source_point[2]};
    h.source_na = b.source_na;
    for (size_t l = 0; l < L; ++l) {
        h.wavelength_weight(int(l)) = b.wavelength_weight[l];
        // used only by glass variables (Buchdahl index of each wavelength)
        h.wavelength_nm(int(l)) = b.wavelength_nm ? b.wavelength_nm[l] : 500.0;
    }
    for (size_t f = 0; f < F; ++f) h.field_weight(int(f)) = b.field_weight[f];
    std::copy_n(b.curvature, V * S, h.curvature_data().begin());
    std::copy_n(b.thickness, V * S, h.thickness_data().begin());
    std::copy_n(b.index, V * S * L, h.index_data().begin());
    std::copy_n(b.pupil_radius, V, h.pupil_radius_data().begin());
    std::copy_n(b.field_angle, V * F, h.field_angle_data().begin());
    std::copy_n(b.field_target, V * F, h.field_target_data().begin());
    if (b.reference_sphere != MKO_REF_LAST_SURFACE)
        std::copy_n(b.reference_curvature, V * F, h.reference_curvature_data().begin());
    if (b.paraxial_index) std::copy_n(b.paraxial_index, V * S, h.paraxial_index_data().begin());
    if (b.buchdahl) std::copy_n(b.buchdahl, V * S * 4, h.buchdahl_coeff_data().begin());
    return h;
}

bool con_gap(int type) {
    return type == MKO_CON_MIN_CENTER_THICKNESS || type == MKO_CON_MAX_CENTER_THICKNESS ||
           type == MKO_CON_MIN_EDGE_THICKNESS;
}

bool needs_extrema(const mko_constraint& c) {
    return c.type == MKO_CON_MAX_SEMI_DIAMETER || c.type == MKO_CON_MAX_INCIDENCE_ANGLE ||
           (c.type == MKO_CON_MIN_EDGE_THICKNESS && c.edge_height == MKO_EDGE_RAYS);
}

// The constraint items as the reference's own row sources: one ExtraResiduals per untraced item
// (make_paraxial_residuals / make_mechanical_residuals(...).geometric), concatenated in item
// order, and the traced mechanical items in one TraceExtraResiduals.
struct Extras {
    mko::ExtraResiduals extra;
    mko::TraceExtraResiduals trace_extra;
};
Extras extras_of(const mko_residual_settings& s) {
    Extras out;
    if (!s.extra || s.extra->count == 0) return out;
    std::vector<mko::ExtraResiduals> parts;
    mko::MechanicalConstraints traced;
    for (int k = 0; k < s.extra->count; ++k) {
        const mko_constraint& c = s.extra->items[k];
        if (c.type <= MKO_CON_PETZVAL || c.type == MKO_CON_SINE_CONDITION ||
            c.type == MKO_CON_DISTORTION) {
            mko::ParaxialConstraint pc;
            pc.type = static_cast<mko::ConstraintType>(c.type);
            pc.mode = static_cast<mko::ConstraintMode>(c.mode);
            pc.weight = c.weight;
            pc.target = c.target;
            pc.reference_surface = c.surface;
            pc.offset = c.offset;
            pc.shift = c.shift;
            pc.zone = c.zone;
            pc.field = c.field;
            mko::ParaxialConstraints one;
            one.items.push_back(pc);
            parts.push_back(mko::make_paraxial_residuals(one));
            continue;
        }
        mko::MechanicalConstraint mc;
        mc.type = static_cast<mko::ConstraintType>(c.type);
        mc.file_surface = c.surface;
        mc.surface = c.surface;
        mc.next_surface = con_gap(c.type) ? c.next_surface : c.surface;
        mc.weight = c.weight;
        mc.limit = c.target;
        mc.clearance = c.offset;
        mc.height = static_cast<mko::EdgeHeight>(c.edge_height);
        mc.semi_diameter = c.semi_diameter;
        mc.next_semi_diameter = c.next_semi_diameter;
        mc.shift = c.shift;
        if (needs_extrema(c)) {
            traced.items.push_back(mc);
        } else {
            mko::MechanicalConstraints one;
            one.items.push_back(mc);
            parts.push_back(mko::make_mechanical_residuals(one).geometric);
        }
    }
    out.extra = mko::concat_extra_residuals(std::move(parts));
    if (!traced.empty()) out.trace_extra = mko::make_mechanical_residuals(traced).traced;
    return out;
}

mko::ResidualSettings settings_of(const mko_residual_settings& s) {
    mko::ResidualSettings r;
    r.rings = s.rings;
    r.spokes = s.spokes;
    r.failure = s.failure == MKO_FAILURE_HARD ? mko::FailureMode::Hard : mko::FailureMode::Soft;
    r.fail_weight = s.fail_weight;
    return r;
}

// A system description whose surfaces match the batch (types, values of variant 0) and whose
// `variables` section lists one item per variable in order.
mko::SystemDescription description_of(const mko_batch& b, const mko_variables& v) {
    mko::SystemDescription d;
    d.name = "driver";
    for (int s = 0; s < b.surfaces; ++s) {
        mko::Surface sf;
        sf.type = b.type[s] == MKO_SURFACE_STOP    ? mko::SurfaceType::Stop
                  : b.type[s] == MKO_SURFACE_PLANE ? mko::SurfaceType::Plane
                                                   : mko::SurfaceType::Sphere;
        sf.curvature = b.curvature[s];
        sf.thickness = b.thickness[s];
        // glass-variable media as Buchdahl materials (variant 0's coefficients), as
        // convert_glass_variables leaves them
        if (b.buchdahl) {
            for (int k = 0; k < v.count; ++k) {
                if (v.type[k] == MKO_VAR_GLASS && v.batch_surface[k] == s) {
                    const double* c = b.buchdahl + size_t(s) * 4;
                    sf.material.kind = mko::MaterialKind::Buchdahl;
                    sf.material.n0 = c[0];
                    sf.material.nu = {c[1], c[2], c[3]};
                }
            }
        }
        d.surfaces.push_back(sf);
    }
    mko::Variables vars;
    for (int k = 0; k < v.count; ++k) {
        // a glass entry stands for its three components (resolve_variables expands it)
        if (v.type[k] == MKO_VAR_GLASS && v.component[k] != 0) continue;
        mko::VariableSpec spec;
        spec.type = v.type[k] == MKO_VAR_CURVATURE   ? mko::VariableType::Curvature
                    : v.type[k] == MKO_VAR_THICKNESS ? mko::VariableType::Thickness
                    : v.type[k] == MKO_VAR_GLASS     ? mko::VariableType::Glass
                                                     : mko::VariableType::WorkingDistance;
        if (v.type[k] == MKO_VAR_GLASS) {
            spec.selection = mko::SurfaceSelection::One;
            spec.surfaces = {v.batch_surface[k]};
            vars.items.push_back(spec);
            continue;
        }
        if (v.type[k] == MKO_VAR_WORKING_DISTANCE) {
            spec.selection = mko::SurfaceSelection::None;
        } else {
            spec.selection = mko::SurfaceSelection::One;
            spec.surfaces = {v.batch_surface[k]};
        }
        if (std::isfinite(v.lower[k])) spec.min = v.lower[k];
        if (std::isfinite(v.upper[k])) spec.max = v.upper[k];
        spec.typical = v.typical[k];
        vars.items.push_back(spec);
    }
    d.variables = vars;
    return d;
}

#define REF_TRY(...)                                                                               \
    try {                                                                                          \
        __VA_ARGS__;                                                                               \
        return 0;                                                                                  \
    } catch (const std::exception& e) {                                                            \
        g_error = e.what();                                                                        \
        return -1;                                                                                 \
    }

mko::LmOptions lm_options_of(const mko_lm_options& o, const mko_residual_settings& s) {
    mko::LmOptions r;
    r.max_iterations = o.max_iterations;
    r.ftol = o.ftol;
    r.xtol = o.xtol;
    r.gtol = o.gtol;
    r.stall_window = o.stall_window;
    r.stall_tolerance = o.stall_tolerance;
    r.initial_damping = o.initial_damping;
    r.max_damping = o.max_damping;
    r.diag_floor = o.diag_floor;
    r.reject_new_failures = o.reject_new_failures != 0;
    r.threads = o.threads;
    r.jacobian.quadrature = settings_of(s);
    const Extras x = extras_of(s);
    r.jacobian.extra = x.extra;
    r.jacobian.trace_extra = x.trace_extra;
    r.jacobian_method = o.jacobian_method == MKO_JACOBIAN_ADJOINT
                            ? mko::JacobianMethod::Adjoint
                            : mko::JacobianMethod::FiniteDifferences;
    r.fail_weight_factor = o.fail_weight_factor;
    r.max_fail_weight = o.max_fail_weight;
    return r;
}

// ---- the reference's LM plug-in over libmikroptim_hip.so -------------------------------------
// mko::hip::HipLmBackend (include/mikroptim_hip_reference.hpp) is the adapter a maintainer adds on
// the reference side; the reference's own optimize_lm_rounds drives it. The library's entry
// points come in as function pointers (HipApi), so this driver does not link against it.

void copy_results(const mko::LmResult& r, int runs, size_t P, double* x_out, mko_lm_run* run_out,
                  int* cycles_out) {
    for (int i = 0; i < runs; ++i) {
        const mko::LmRunResult& q = r.runs[size_t(i)];
        std::copy(q.x.begin(), q.x.end(), x_out + size_t(i) * P);
        mko_lm_run& out = run_out[i];
        out.merit = q.merit;
        out.failed_rays = q.failed_rays;
        out.hard_failed_rays = q.hard_failed_rays;
        out.status = static_cast<int32_t>(q.status);
        out.iterations = q.iterations;
        out.accepted = q.accepted;
        out.jacobians = q.jacobians;
        out.damping = q.damping;
        out.gradient_norm = q.gradient_norm;
        out.model_decrease = q.model_decrease;
        out.fail_weight = q.fail_weight;
        out.penalty_rounds = q.penalty_rounds;
    }
    if (cycles_out) *cycles_out = r.cycles;
}

} // namespace

extern "C" {

const char* ref_last_error(void) { return g_error.c_str(); }

int ref_gauss_legendre(int n, double* u, double* w) {
    REF_TRY({
        const auto q = mko::make_pupil_quadrature<double>(n, 1, mko::PupilMode::Half);
        for (int i = 0; i < n; ++i) {
            u[i] = q.u[i];
            w[i] = q.w[i];
        }
    })
}

int ref_pupil_samples(int rings, int spokes, int mode, double* rho, double* phi, double* weight) {
    REF_TRY({
        const auto q =
            mko::make_pupil_quadrature<double>(rings, spokes, static_cast<mko::PupilMode>(mode));
        for (int k = 0; k < mko::sample_count(q); ++k) {
            const auto s = mko::sample(q, k);
            rho[k] = s.rho;
            phi[k] = s.phi;
            weight[k] = s.weight;
        }
    })
}

int ref_residuals(const mko_batch* b, const mko_residual_settings* s, double* residuals,
                  double* opd, uint8_t* status) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const mko::ResidualLayout layout(h.topology(), settings_of(*s));
        const mko::ResidualResult r = mko::residuals_cpu(h, layout);
        std::copy(r.opd.begin(), r.opd.end(), opd);
        std::copy(r.status.begin(), r.status.end(), status);
        // constraint rows after the ray rows, as CpuLmBackend::host_residuals
        // (src/optim/levenberg_marquardt.cpp:812-864)
        const Extras x = extras_of(*s);
        mko::FdJacobianOptions fo;
        fo.quadrature = settings_of(*s);
        fo.extra = x.extra;
        fo.trace_extra = x.trace_extra;
        const int E = fo.extra_count();
        const size_t tr = size_t(r.stride), R = tr + size_t(E);
        std::vector<double> ev;
        if (E > 0) {
            const mko::SurfaceExtremaResult ext = x.trace_extra.count > 0
                                                      ? mko::surface_extrema_cpu(h, layout)
                                                      : mko::SurfaceExtremaResult{};
            ev = mko::fd_extra_residuals(fo, h, ext);
        }
        for (size_t v = 0; v < size_t(h.variants()); ++v) {
            std::copy_n(r.residuals.begin() + std::ptrdiff_t(v * tr), tr, residuals + v * R);
            for (int e = 0; e < E; ++e) residuals[v * R + tr + size_t(e)] = ev[v * size_t(E) + size_t(e)];
        }
    })
}

int ref_surface_extrema(const mko_batch* b, const mko_residual_settings* s, double* extrema) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const mko::ResidualLayout layout(h.topology(), settings_of(*s));
        const mko::SurfaceExtremaResult r = mko::surface_extrema_cpu(h, layout);
        std::copy(r.values.begin(), r.values.end(), extrema);
    })
}

int ref_paraxial(const mko_batch* b, int v, double* out) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const mko::ParaxialData p = mko::batch_paraxial(h, v);
        out[0] = p.abcd[0];
        out[1] = p.abcd[1];
        out[2] = p.abcd[2];
        out[3] = p.abcd[3];
        out[4] = p.power;
        out[5] = p.petzval_curvature;
    })
}

// Rows [V][E + E_t] with the given extrema [V][S][2] (fd_extra_residuals,
// src/optim/fd_jacobian.cpp:164-180).
int ref_constraint_rows(const mko_batch* b, const mko_residual_settings* s, const double* extrema,
                        double* rows) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const Extras x = extras_of(*s);
        mko::FdJacobianOptions fo;
        fo.quadrature = settings_of(*s);
        fo.extra = x.extra;
        fo.trace_extra = x.trace_extra;
        mko::SurfaceExtremaResult ext;
        if (x.trace_extra.count > 0) {
            ext.variants = h.variants();
            ext.surfaces = h.surfaces();
            ext.values.assign(extrema, extrema + size_t(ext.variants) * size_t(ext.surfaces) * 2);
        }
        const std::vector<double> ev = mko::fd_extra_residuals(fo, h, ext);
        std::copy(ev.begin(), ev.end(), rows);
    })
}

int ref_merit(const mko_batch* b, const mko_residual_settings* s, double* merit, int64_t* failed,
              int64_t* hard_failed) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const mko::ResidualLayout layout(h.topology(), settings_of(*s));
        // (ray merit + sum e^2) + sum t^2 as CpuLmBackend::merit
        // (src/optim/levenberg_marquardt.cpp:866-888)
        const Extras x = extras_of(*s);
        const std::vector<double> extra = mko::extra_merit_values(x.extra, h, 1);
        for (int v = 0; v < h.variants(); ++v) {
            const mko::MeritResult m = mko::merit_cpu(h, v, layout, x.trace_extra.count > 0);
            merit[v] = extra.empty() ? m.merit[0] : m.merit[0] + extra[size_t(v)];
            if (x.trace_extra.count > 0)
                merit[v] += mko::trace_extra_merit(x.trace_extra, h, v,
                                                   mko::variant_extrema(m.extrema, 0));
            failed[v] = m.failed[0];
            hard_failed[v] = m.hard_failed[0];
        }
    })
}

int ref_fd_pattern(const mko_batch* b, const mko_variables* vars, const double* x0,
                   double* nominal, double* plus, double* minus, double* step_plus,
                   double* step_minus, uint8_t* stencil) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const mko::VariableSet set(description_of(*b, *vars), h.topology());
        const mko::FdPattern p = mko::fd_pattern(
            set, std::span<const double>(x0, size_t(vars->count)), mko::FdScheme::Central);
        for (int k = 0; k < vars->count; ++k) {
            nominal[k] = p.nominal[size_t(k)];
            plus[k] = p.point_plus[size_t(k)];
            minus[k] = p.point_minus[size_t(k)];
            step_plus[k] = p.step_plus[size_t(k)];
            step_minus[k] = p.step_minus[size_t(k)];
            stencil[k] = static_cast<uint8_t>(p.stencil[size_t(k)]);
        }
    })
}

int ref_fd_jacobian(const mko_batch* base, const mko_residual_settings* s,
                    const mko_variables* vars, int runs, const double* x0, double* J, double* r0,
                    int64_t* status_changes) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*base);
        const mko::VariableSet set(description_of(*base, *vars), h.topology());
        mko::FdJacobianOptions o;
        o.quadrature = settings_of(*s);
        const Extras x = extras_of(*s);
        o.extra = x.extra;
        o.trace_extra = x.trace_extra;
        const mko::FdJacobian jac = mko::fd_jacobian_cpu(
            set, h, std::span<const double>(x0, size_t(runs) * size_t(vars->count)), o);
        std::copy(jac.jacobian.begin(), jac.jacobian.end(), J);
        std::copy(jac.residuals.begin(), jac.residuals.end(), r0);
        std::copy(jac.status_changes.begin(), jac.status_changes.end(), status_changes);
    })
}

int ref_adjoint_jacobian(const mko_batch* base, const mko_residual_settings* s,
                         const mko_variables* vars, int runs, const double* x0, double* J,
                         double* r0, uint8_t* status, int threads) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*base);
        const mko::VariableSet set(description_of(*base, *vars), h.topology());
        mko::AdjointJacobianOptions o;
        o.quadrature = settings_of(*s);
        o.threads = threads;
        const Extras x = extras_of(*s);
        o.extra = x.extra;
        o.trace_extra = x.trace_extra;
        const mko::AdjointJacobian jac = mko::adjoint_jacobian_cpu(
            set, h, std::span<const double>(x0, size_t(runs) * size_t(vars->count)), o);
        std::copy(jac.jacobian.begin(), jac.jacobian.end(), J);
        std::copy(jac.residuals.begin(), jac.residuals.end(), r0);
        std::copy(jac.status.begin(), jac.status.end(), status);
    })
}

int ref_normal_equations(const double* J, const double* r, int runs, int P, int64_t R, double* a,
                         double* g) {
    REF_TRY({
        const size_t p = size_t(P), rr = size_t(R);
        for (int j = 0; j < runs; ++j) {
            mko::normal_equations_host(std::span<const double>(J + size_t(j) * p * rr, p * rr),
                                       std::span<const double>(r + size_t(j) * rr, rr), P,
                                       std::span<double>(a + size_t(j) * p * p, p * p),
                                       std::span<double>(g + size_t(j) * p, p));
        }
    })
}

uint64_t ref_splitmix64(uint64_t x) { return mko::splitmix64(x); }
uint64_t ref_variant_seed(uint64_t seed, uint64_t j) { return mko::variant_seed(seed, j); }
double ref_unit_uniform(uint64_t bits) { return mko::unit_uniform(bits); }

int ref_perturbed_vector(const mko_batch* b, const mko_variables* vars, const double* x0,
                         const double* relative, const double* absolute, int distribution,
                         uint64_t seed, int first_is_base, uint64_t j, double* y) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const mko::VariableSet set(description_of(*b, *vars), h.topology());
        mko::PerturbationOptions o;
        o.curvature = {relative[MKO_VAR_CURVATURE], absolute[MKO_VAR_CURVATURE]};
        o.thickness = {relative[MKO_VAR_THICKNESS], absolute[MKO_VAR_THICKNESS]};
        o.working_distance = {relative[MKO_VAR_WORKING_DISTANCE],
                              absolute[MKO_VAR_WORKING_DISTANCE]};
        o.distribution =
            distribution == 0 ? mko::Distribution::Uniform : mko::Distribution::IrwinHall;
        o.seed = seed;
        o.first_is_base = first_is_base != 0;
        const std::vector<double> r = mko::perturbed_vector(
            set, std::span<const double>(x0, size_t(vars->count)), o, j);
        std::copy(r.begin(), r.end(), y);
    })
}

int ref_optimize_lm(const mko_batch* base, const mko_residual_settings* s,
                    const mko_variables* vars, int runs, const double* x0,
                    const mko_lm_options* options, double* x_out, mko_lm_run* run_out,
                    int* cycles_out) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*base);
        const mko::VariableSet set(description_of(*base, *vars), h.topology());
        const mko::LmOptions o = lm_options_of(*options, *s);
        const mko::LmResult r = mko::optimize_lm_cpu(
            set, h, std::span<const double>(x0, size_t(runs) * size_t(vars->count)), o);
        copy_results(r, runs, size_t(vars->count), x_out, run_out, cycles_out);
    })
}

// The reference's optimize_lm_rounds (src/optim/levenberg_marquardt.cpp:891) with HipLmBackend:
// the reference's LM control over the library's normal equations and merit. calls[3] =
// {mko_engine_normal_equations, mko_engine_point_merit, mko_last_error} of the loaded library;
// precision MKO_PRECISION_FP64 or MKO_PRECISION_MIXED (the adjoint's argument).
int ref_optimize_lm_hip_backend(void* const* calls, void* engine, const mko_batch* base,
                                const mko_residual_settings* s, const mko_variables* vars,
                                int runs, const double* x0, const mko_lm_options* options,
                                int precision, double* x_out, mko_lm_run* run_out,
                                int* cycles_out) {
    REF_TRY({
        mko::hip::HipApi api{};
        api.normal_equations = reinterpret_cast<decltype(api.normal_equations)>(calls[0]);
        api.point_merit = reinterpret_cast<decltype(api.point_merit)>(calls[1]);
        api.last_error = reinterpret_cast<decltype(api.last_error)>(calls[2]);
        const mko::SystemBatchHost h = batch_of(*base);
        const mko::VariableSet set(description_of(*base, *vars), h.topology());
        const mko::LmOptions o = lm_options_of(*options, *s);
        // the caller's variable description (its bounds drive the FD patterns exactly as in the
        // library's own LM); the LM control takes its bounds from `set`
        const mko::LmResult r = mko::hip::optimize_lm_hip(
            static_cast<mko_engine*>(engine), set, h,
            std::span<const double>(x0, size_t(runs) * size_t(vars->count)), o, precision, api,
            vars);
        copy_results(r, runs, size_t(vars->count), x_out, run_out, cycles_out);
    })
}

// The variable description the reference-side adapter derives from the reference's VariableSet
// (mko::hip::VariablesView over description_of(base, vars)): type / batch_surface / component
// [P], lower / upper / typical [P]; count = its P.
int ref_variables_view(const mko_batch* base, const mko_variables* vars, int32_t* count,
                       int32_t* type, int32_t* surface, int32_t* component, double* lower,
                       double* upper, double* typical) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*base);
        const mko::VariableSet set(description_of(*base, *vars), h.topology());
        const mko::hip::VariablesView view(set);
        const mko_variables& v = view.c();
        *count = v.count;
        for (int k = 0; k < v.count && k < vars->count; ++k) {
            type[k] = v.type[k];
            surface[k] = v.batch_surface[k];
            component[k] = v.component ? v.component[k] : 0;
            lower[k] = v.lower[k];
            upper[k] = v.upper[k];
            typical[k] = v.typical[k];
        }
    })
}

// prefilter_stats_cpu (src/batch/batch_prefilter.cpp:384).
int ref_prefilter_stats(const mko_batch* b, const mko_residual_settings* s, double boundary,
                        double* merit, double* optical, int64_t* failed, int64_t* uncertain_ok,
                        int64_t* uncertain_failed) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const mko::ResidualLayout layout(h.topology(), settings_of(*s));
        const mko::PrefilterStats st = mko::prefilter_stats_cpu(h, layout, boundary);
        std::copy(st.merit.begin(), st.merit.end(), merit);
        std::copy(st.optical_length.begin(), st.optical_length.end(), optical);
        std::copy(st.failed.begin(), st.failed.end(), failed);
        std::copy(st.uncertain_ok.begin(), st.uncertain_ok.end(), uncertain_ok);
        std::copy(st.uncertain_failed.begin(), st.uncertain_failed.end(), uncertain_failed);
    })
}

// prefilter_decide (src/batch/batch_prefilter.cpp:236) on given stats; options[6] =
// {max_failed_fraction, max_merit, max_merit_factor, boundary, opl_error_factor,
// merit_relative}; extra [V] or NULL.
int ref_prefilter_decide(const mko_batch* b, const mko_residual_settings* s, const double* options,
                         const double* merit, const double* optical, const int64_t* failed,
                         const int64_t* uncertain_ok, const int64_t* uncertain_failed,
                         const double* extra, double* lower, double* upper, uint8_t* reject,
                         double* reference_merit, int* reference_variant) {
    REF_TRY({
        const mko::SystemBatchHost h = batch_of(*b);
        const mko::ResidualLayout layout(h.topology(), settings_of(*s));
        const size_t V = size_t(b->variants);
        mko::PrefilterStats st;
        st.variants = b->variants;
        st.rays_per_variant = layout.rays_per_variant();
        st.merit.assign(merit, merit + V);
        st.optical_length.assign(optical, optical + V);
        st.failed.assign(failed, failed + V);
        st.uncertain_ok.assign(uncertain_ok, uncertain_ok + V);
        st.uncertain_failed.assign(uncertain_failed, uncertain_failed + V);
        mko::PrefilterOptions o;
        o.criteria.max_failed_fraction = options[0];
        o.criteria.max_merit = options[1];
        o.criteria.max_merit_factor = options[2];
        o.margins.boundary = options[3];
        o.margins.opl_error_factor = options[4];
        o.margins.merit_relative = options[5];
        const std::vector<double> ex = extra ? std::vector<double>(extra, extra + V)
                                             : std::vector<double>();
        const mko::PrefilterResult r = mko::prefilter_decide(std::move(st), h, layout, o, ex);
        std::copy(r.merit_lower.begin(), r.merit_lower.end(), lower);
        std::copy(r.merit_upper.begin(), r.merit_upper.end(), upper);
        std::copy(r.reject.begin(), r.reject.end(), reject);
        *reference_merit = r.reference_merit;
        *reference_variant = r.reference_variant;
    })
}

// The reference's fit_buchdahl (src/glass/buchdahl_fit.cpp) of a glass's index: model 0
// Sellmeier (A, B[terms], C[terms]), 1 polynomial (coefficients in b, exponents in c). The index
// function evaluates the reference's inline formulas (dispersion.hpp: sellmeier_index_n,
// polynomial_index) with the range check of Glass::index; Glass::index itself lives in
// src/glass/catalog.cpp with the JSON reader, which needs a library this image lacks. nodes: 0
// Chebyshev, 1 equidistant. out: n0, nu1, nu2, nu3, max_abs_error, rms_error.
int ref_fit_buchdahl(int model, double A, const double* b, const double* c, int terms,
                     double range_lo, double range_hi, int dispersion_terms, double lambda_min,
                     double lambda_max, int samples, int nodes, int check_samples, double* out) {
    REF_TRY({
        mko::BuchdahlFitOptions o;
        o.lambda_min_nm = lambda_min;
        o.lambda_max_nm = lambda_max;
        o.samples = samples;
        o.nodes = nodes == 0 ? mko::BuchdahlFitNodes::Chebyshev : mko::BuchdahlFitNodes::Equidistant;
        o.check_samples = check_samples;
        if (!(lambda_min >= range_lo && lambda_min <= range_hi && lambda_max >= range_lo &&
              lambda_max <= range_hi))
            throw std::invalid_argument("fit range outside the glass range");
        const std::vector<double> B(b, b + terms), C(c, c + terms);
        const auto index = [&](double lambda_nm) {
            if (!(lambda_nm >= range_lo && lambda_nm <= range_hi))
                throw std::invalid_argument("wavelength outside the glass range");
            const double l = mko::nm_to_um(lambda_nm);
            return model == 1 ? mko::polynomial_index(B.data(), C.data(), terms, l)
                              : mko::sellmeier_index_n(A, B.data(), C.data(), terms, l);
        };
        const mko::BuchdahlFit f = mko::fit_buchdahl(index, dispersion_terms, o);
        out[0] = f.n0;
        out[1] = f.nu[0];
        out[2] = f.nu[1];
        out[3] = f.nu[2];
        out[4] = f.max_abs_error;
        out[5] = f.rms_error;
    })
}

// The reference's glass hull (GlassHull, src/glass/glass_hull.cpp; convex_hull,
// src/glass/convex_hull.cpp) of n glasses' coordinates (n0, nu1, nu2) and its rows
// glass_hull_residual (src/optim/glass_variables.cpp:147-155) at m points x: out [m][3]; info:
// dimension, center (3), scale (3), then the hull's vertex count and triangle count.
int ref_glass_hull_rows(const double* coords, int n, double weight, double margin,
                        const double* x, int m, double* out, double* info) {
    REF_TRY({
        std::vector<mko::GlassHullPoint> pts(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            pts[size_t(i)].name = "G" + std::to_string(i);
            pts[size_t(i)].coordinates = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
        }
        mko::GlassHullPenalty pen;
        pen.hull = mko::GlassHull(std::move(pts));
        pen.weight = weight;
        pen.margin = margin;
        for (int k = 0; k < m; ++k) {
            const mko::Point3 r =
                mko::glass_hull_residual(pen, {x[3 * k], x[3 * k + 1], x[3 * k + 2]});
            for (int j = 0; j < 3; ++j) out[3 * k + j] = r[size_t(j)];
        }
        info[0] = pen.hull.dimension();
        for (int j = 0; j < 3; ++j) {
            info[1 + j] = pen.hull.center()[size_t(j)];
            info[4 + j] = pen.hull.scale()[size_t(j)];
        }
        info[7] = double(pen.hull.hull().vertices.size());
        info[8] = double(pen.hull.hull().triangles.size());
    })
}

// The reference's own df64 operations (include/mikroptim/math/df64.hpp), same layout and op
// codes as orc_df64_ops.
int ref_df64_ops(int op, int64_t n, const float* a, const float* b, float* out) {
    if (op < 0 || op > 5 || n < 0) return -1;
    for (int64_t i = 0; i < n; ++i) {
        const mko::df64 x = mko::df64::from_parts(a[2 * i], a[2 * i + 1]);
        const mko::df64 y = op == 5 ? mko::df64::from_parts(0.0f, 0.0f)
                                    : mko::df64::from_parts(b[2 * i], b[2 * i + 1]);
        mko::df64 z{};
        switch (op) {
        case 0: z = x + y; break;
        case 1: z = x - y; break;
        case 2: z = x * y; break;
        case 3: z = x * y.hi; break;
        case 4: z = x / y; break;
        default: z = mko::msqrt(x); break;
        }
        out[2 * i] = z.hi;
        out[2 * i + 1] = z.lo;
    }
    return 0;
}

} // extern "C"

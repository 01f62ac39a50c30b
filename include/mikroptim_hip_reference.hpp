// The reference's side of the boundary: its own types over libmikroptim_hip.so. A build of the
// reference that includes this header (and its include/ directory) drives the HIP path with
// its LM control and its batches:
//   BatchView        SystemBatchHost (include/mikroptim/batch/system_batch.hpp) -> mko_batch,
//                    a view of the batch's own arrays (no copy of per-variant data)
//   VariablesView    VariableSet (include/mikroptim/optim/variable_set.hpp) -> mko_variables
//   settings_of      ResidualSettings (include/mikroptim/batch/batch_residuals.hpp) ->
//                    mko_residual_settings
//   HipLmBackend     mko::LmBackend (include/mikroptim/optim/levenberg_marquardt.hpp:171-182):
//                    normal_equations / merit answered by mko_engine_normal_equations /
//                    mko_engine_point_merit, the counterpart of mko::cuda::GpuLmBackend
//                    (include/mikroptim/cuda/levenberg_marquardt.hpp:48-88)
//   optimize_lm_hip  mko::cuda::optimize_lm_gpu (levenberg_marquardt.hpp:91): the reference's
//                    optimize_lm_rounds (outer penalty rounds included) with HipLmBackend
// Errors: argument errors throw std::invalid_argument, device errors mko_hip::DeviceError, as
// the reference's GPU backend throws std::invalid_argument / CudaError.
// The library's entry points are reached through HipApi: linked_api() takes the linked
// symbols; a caller that loads the library at run time fills HipApi with its own pointers.
// Constraint rows (LmOptions::jacobian.extra / trace_extra) are not forwarded by this adapter:
// the C-ABI takes them as mko_constraint items, which the reference's ExtraResiduals (opaque
// closures) do not expose; an LM with constraint rows is refused.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "mikroptim/batch/system_batch.hpp"
#include "mikroptim/optim/levenberg_marquardt.hpp"
#include "mikroptim/optim/variable_set.hpp"
#include "mikroptim_hip.h"
#include "mikroptim_hip.hpp"

namespace mko::hip {

// The entry points the adapter calls.
struct HipApi {
    decltype(&mko_engine_normal_equations) normal_equations = nullptr;
    decltype(&mko_engine_point_merit) point_merit = nullptr;
    decltype(&mko_last_error) last_error = nullptr;
};

// The symbols of a build linked against libmikroptim_hip.so.
inline HipApi linked_api() {
    return HipApi{&mko_engine_normal_equations, &mko_engine_point_merit, &mko_last_error};
}

// mko_batch over the arrays of a SystemBatchHost; valid while the batch lives and keeps its
// sizes. Ghost plates and decentered surfaces are not on the HIP path and are refused.
class BatchView {
public:
    explicit BatchView(const SystemBatchHost& b) {
        const BatchTopology& t = b.topology();
        if (t.ghost_plates != 0)
            throw std::invalid_argument("mikroptim_hip: ghost plates are not on the HIP path");
        for (SurfaceKind k : b.type_data()) type_.push_back(uint8_t(k));
        for (PupilMode m : t.pupil_mode) mode_.push_back(uint8_t(m));
        if (int(mode_.size()) != b.fields())
            throw std::invalid_argument("mikroptim_hip: the batch needs one pupil mode per field");
        c_.variants = b.variants();
        c_.surfaces = b.surfaces();
        c_.wavelengths = b.wavelengths();
        c_.fields = b.fields();
        c_.source = t.source == SourceKind::Point ? MKO_SOURCE_POINT : MKO_SOURCE_COLLIMATED;
        c_.reference_sphere = int32_t(t.reference_sphere);
        c_.reference_wavelength_nm = b.reference_wavelength_nm;
        for (int i = 0; i < 3; ++i) c_.source_point[i] = b.source_point[std::size_t(i)];
        c_.source_na = b.source_na;
        c_.type = type_.data();
        c_.pupil_mode = mode_.data();
        c_.wavelength_weight = b.wavelength_weight_data().data();
        c_.field_weight = b.field_weight_data().data();
        c_.curvature = b.curvature_data().data();
        c_.thickness = b.thickness_data().data();
        c_.index = b.index_data().data();
        c_.pupil_radius = b.pupil_radius_data().data();
        c_.field_angle = b.field_angle_data().data();
        c_.field_target = b.field_target_data().data();
        c_.reference_curvature = c_.reference_sphere != MKO_REF_LAST_SURFACE
                                     ? b.reference_curvature_data().data()
                                     : nullptr;
        c_.paraxial_index = or_null(b.paraxial_index_data());
        c_.wavelength_nm = or_null(b.wavelength_nm_data());
        c_.buchdahl = or_null(b.buchdahl_coeff_data());
    }
    BatchView(const BatchView&) = delete; // c_ points into the view's own vectors
    BatchView& operator=(const BatchView&) = delete;
    const mko_batch& c() const { return c_; }

private:
    static const double* or_null(std::span<const double> s) { return s.empty() ? nullptr : s.data(); }
    std::vector<uint8_t> type_, mode_;
    mko_batch c_{};
};

// mko_variables of a VariableSet (types, batch surfaces, glass components, bounds, typical
// sizes); valid while the set lives.
class VariablesView {
public:
    explicit VariablesView(const VariableSet& set) {
        const std::span<const Variable> vars = set.variables();
        bool glass = false;
        for (const Variable& v : vars) {
            int32_t t = 0;
            switch (v.type) {
            case VariableType::Curvature: t = MKO_VAR_CURVATURE; break;
            case VariableType::Thickness: t = MKO_VAR_THICKNESS; break;
            case VariableType::WorkingDistance: t = MKO_VAR_WORKING_DISTANCE; break;
            case VariableType::Glass: t = MKO_VAR_GLASS; glass = true; break;
            default:
                throw std::invalid_argument("mikroptim_hip: ghost-plate variables are not on the HIP path");
            }
            if (v.split)
                throw std::invalid_argument("mikroptim_hip: split variables are not on the HIP path");
            type_.push_back(t);
            surface_.push_back(v.batch_surface);
            component_.push_back(v.type == VariableType::Glass ? v.glass_component : 0);
        }
        c_.count = int32_t(vars.size());
        c_.type = type_.data();
        c_.batch_surface = surface_.data();
        c_.lower = set.lower().data();
        c_.upper = set.upper().data();
        c_.typical = set.typical().data();
        c_.component = glass ? component_.data() : nullptr;
    }
    VariablesView(const VariablesView&) = delete;
    VariablesView& operator=(const VariablesView&) = delete;
    const mko_variables& c() const { return c_; }

private:
    std::vector<int32_t> type_, surface_, component_;
    mko_variables c_{};
};

inline mko_residual_settings settings_of(const ResidualSettings& q) {
    mko_residual_settings s{};
    s.rings = q.rings;
    s.spokes = q.spokes;
    s.failure = q.failure == FailureMode::Hard ? MKO_FAILURE_HARD : MKO_FAILURE_SOFT;
    s.fail_weight = q.fail_weight;
    s.precision = MKO_PRECISION_FP64;
    s.extra = nullptr;
    return s;
}

// LmBackend over the library: the reference's LM control, the HIP normal equations and merit.
// precision: MKO_PRECISION_FP64 (the reference's GpuLmBackend), MKO_PRECISION_MIXED (fused
// adjoint normal equations) or MKO_PRECISION_FP64_MFMA. `base` and `vars` must outlive it.
class HipLmBackend final : public LmBackend {
public:
    HipLmBackend(const HipApi& api, mko_engine* engine, const SystemBatchHost& base,
                 const mko_variables& vars, const LmOptions& o,
                 int precision = MKO_PRECISION_FP64)
        : api_(api), engine_(engine), batch_(base), vars_(vars), precision_(precision) {
        if (!api_.normal_equations || !api_.point_merit || !api_.last_error)
            throw std::invalid_argument("HipLmBackend: library entry points missing");
        if (o.jacobian.extra.count != 0 || o.jacobian.trace_extra.count != 0)
            throw std::invalid_argument("HipLmBackend: constraint rows are not forwarded");
        s_ = settings_of(o.jacobian.quadrature);
        method_ = o.jacobian_method == JacobianMethod::Adjoint ? MKO_JACOBIAN_ADJOINT
                                                               : MKO_JACOBIAN_FD;
    }
    void normal_equations(std::span<const int> runs, std::span<const double> x,
                          std::span<double> a, std::span<double> g) override {
        check(api_.normal_equations(engine_, &batch_.c(), &s_, &vars_, int(runs.size()),
                                    runs.data(), x.data(), method_, precision_, a.data(),
                                    g.data()));
    }
    void merit(std::span<const int> runs, std::span<const double> x, std::span<double> merit,
               std::span<std::int64_t> failed, std::span<std::int64_t> hard_failed) override {
        check(api_.point_merit(engine_, &batch_.c(), &s_, &vars_, int(runs.size()), runs.data(),
                               x.data(), merit.data(), failed.data(), hard_failed.data()));
    }

private:
    void check(int rc) const {
        if (rc >= 0) return;
        const char* m = api_.last_error();
        const std::string msg = m ? m : "mikroptim_hip: unknown error";
        if (rc == MKO_ERR_INVALID) throw std::invalid_argument(msg);
        throw mko_hip::DeviceError(msg);
    }
    HipApi api_;
    mko_engine* engine_;
    BatchView batch_;
    mko_variables vars_;
    int precision_;
    int method_ = MKO_JACOBIAN_ADJOINT;
    mko_residual_settings s_{};
};

// optimize_lm_gpu on the HIP path: optimize_lm_rounds with one HipLmBackend per round. `vars`
// (NULL: the view of `set`) is the variable description the library gets; the LM control takes
// its bounds from `set`, as the reference's does.
inline LmResult optimize_lm_hip(mko_engine* engine, const VariableSet& set,
                                const SystemBatchHost& base, std::span<const double> x0,
                                const LmOptions& options, int precision = MKO_PRECISION_FP64,
                                const HipApi& api = linked_api(),
                                const mko_variables* vars = nullptr) {
    std::unique_ptr<VariablesView> own;
    if (!vars) {
        own = std::make_unique<VariablesView>(set);
        vars = &own->c();
    }
    const mko_variables v = *vars;
    const LmBackendFactory make = [&](const LmOptions& ro, const SystemBatchHost& rb) {
        return std::unique_ptr<LmBackend>(new HipLmBackend(api, engine, rb, v, ro, precision));
    };
    return optimize_lm_rounds(make, set, base, x0, options);
}

} // namespace mko::hip

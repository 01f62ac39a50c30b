// Host layer of libmikroptim_hip.so (C++20): the TraceEngine / GPU LM backend / batched
// Levenberg-Marquardt of the reference, over the device entry points of kernels.hip.
//
//   Engine            TraceEngine (include/mikroptim/cuda/trace_engine.hpp; tranches
//                     src/cuda/trace_engine.cu:199, run :275): device buffers that only grow,
//                     per-call memory budget, variants split into tranches that fit it.
//   Engine::jacobian  DeviceFdJacobian::compute (src/cuda/fd_jacobian.cu:203) and
//                     DeviceAdjointJacobian::compute (src/cuda/adjoint_jacobian.cu:401) followed
//                     by DeviceNormalEquations (src/cuda/normal_equations.cu:155); runs chunked so
//                     the Jacobian fits the budget (GpuLmBackend::max_jacobian_runs,
//                     src/cuda/levenberg_marquardt.cpp:35-67).
//   Lm                levenberg_marquardt (src/optim/levenberg_marquardt.cpp:332-572) with the
//                     GPU backend (GpuLmBackend, src/cuda/levenberg_marquardt.cpp:24-248) and
//                     the outer penalty rounds of optimize_lm_rounds (:891-983), cut into cycles.
// All calls are synchronous; arguments are checked before device work.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "error.hpp"
#include "internal.hpp"
#include "lm_device.hpp"
#include "mikroptim_hip.h"

namespace mko_host {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}
inline void dev_check(int rc) {
    if (rc == MKO_ERR_INVALID) throw std::invalid_argument(mko_detail::last_error);
    if (rc != 0) throw DeviceError(mko_detail::last_error);
}

// Device allocation that only grows (reused across calls).
class Buffer {
public:
    Buffer() = default;
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    ~Buffer() {
        if (ptr_) (void)hipFree(ptr_);
    }
    void* reserve(size_t bytes) {
        if (bytes > size_) {
            if (ptr_) hip_check(hipFree(ptr_), "hipFree");
            ptr_ = nullptr;
            size_ = 0;
            hip_check(hipMalloc(&ptr_, std::max<size_t>(bytes, 256)), "hipMalloc");
            size_ = bytes;
        }
        return ptr_;
    }
    template <class T>
    T* as(size_t count) {
        return static_cast<T*>(reserve(count * sizeof(T)));
    }
    size_t size() const { return size_; }

private:
    void* ptr_ = nullptr;
    size_t size_ = 0;
};

// Restores the caller's current device (DeviceGuard of the reference).
struct DeviceGuard {
    int old = 0;
    explicit DeviceGuard(int dev) {
        hip_check(hipGetDevice(&old), "hipGetDevice");
        hip_check(hipSetDevice(dev), "hipSetDevice");
    }
    ~DeviceGuard() { (void)hipSetDevice(old); }
};

inline int64_t rays_of(const mko_batch& b, const mko_residual_settings& s) {
    int64_t rays = 0;
    for (int f = 0; f < b.fields; ++f)
        rays += int64_t(b.wavelengths) * s.rings * (b.pupil_mode[f] == MKO_PUPIL_AXIS ? 1 : s.spokes);
    return rays;
}
inline int64_t trace_rows_of(const mko_batch& b, const mko_residual_settings& s) {
    const int64_t r = rays_of(b, s);
    return s.failure == MKO_FAILURE_SOFT ? 2 * r : r;
}

// Per-variant arrays of a batch: element counts per variant.
struct Strides {
    size_t S, SL, F;
    bool refc, pidx, bd;
    explicit Strides(const mko_batch& b)
        : S(size_t(b.surfaces)), SL(size_t(b.surfaces) * size_t(b.wavelengths)),
          F(size_t(b.fields)), refc(b.reference_sphere != MKO_REF_LAST_SURFACE),
          pidx(b.paraxial_index != nullptr), bd(b.buchdahl != nullptr) {}
    size_t doubles() const {
        return 2 * S + SL + 1 + 2 * F + (refc ? F : 0) + (pidx ? S : 0) + (bd ? 4 * S : 0);
    }
};

void check_batch(const mko_batch& b) {
    if (b.variants < 0) throw std::invalid_argument("negative variant count");
    if (b.surfaces < 2 || b.surfaces > MKO_MAX_SURFACES)
        throw std::invalid_argument("surfaces outside [2, 64]");
    if (b.wavelengths < 1 || b.fields < 1) throw std::invalid_argument("no wavelengths or fields");
    if (!b.type || !b.pupil_mode || !b.wavelength_weight || !b.field_weight)
        throw std::invalid_argument("batch-wide arrays missing");
    if (b.reference_sphere != MKO_REF_LAST_SURFACE && !b.reference_curvature)
        throw std::invalid_argument("image reference sphere needs reference_curvature");
}

// Glass variables (VariableSet::apply, src/optim/variable_set.cpp:244-258): components 0 ... 2,
// the batch's Buchdahl coefficients and wavelengths, a defined Buchdahl coordinate at every
// wavelength and at lambda_ref.
void check_glass(const mko_batch& b, const mko_variables& v) {
    bool glass = false;
    for (int k = 0; k < v.count; ++k) {
        if (v.type[k] != MKO_VAR_GLASS) continue;
        glass = true;
        if (!v.component || v.component[k] < 0 || v.component[k] > 2)
            throw std::invalid_argument("glass variables need components 0 ... 2");
    }
    if (!glass) return;
    if (!b.buchdahl || !b.wavelength_nm)
        throw std::invalid_argument("glass variables need the batch's Buchdahl coefficients and "
                                    "wavelengths");
    for (int l = 0; l <= b.wavelengths; ++l) {
        const double nm = l < b.wavelengths ? b.wavelength_nm[l] : b.reference_wavelength_nm;
        if (!(1.0 + 2.5 * (nm / 1000.0 - 0.546074) > 0.0))
            throw std::invalid_argument("glass variables: Buchdahl coordinate undefined at a "
                                        "wavelength (needs lambda > 146.074 nm)");
    }
}

// validate_values (src/batch/system_batch.cpp:558): finite values, pupil radius > 0, indices > 0.
void validate_values(const mko_batch& b, int first, int count) {
    const Strides st(b);
    for (int v = first; v < first + count; ++v) {
        auto bad = [&](const char* what) {
            throw std::invalid_argument("variant " + std::to_string(v) + ": " + what);
        };
        for (size_t i = 0; i < st.S; ++i) {
            if (!std::isfinite(b.curvature[v * st.S + i]) || !std::isfinite(b.thickness[v * st.S + i]))
                bad("curvature or thickness not finite");
        }
        for (size_t i = 0; i < st.SL; ++i) {
            const double n = b.index[v * st.SL + i];
            if (!(n > 0.0) || !std::isfinite(n)) bad("refractive index not finite and > 0");
        }
        if (!(b.pupil_radius[v] > 0.0) || !std::isfinite(b.pupil_radius[v])) bad("pupil radius");
        for (size_t f = 0; f < st.F; ++f) {
            if (!std::isfinite(b.field_angle[v * st.F + f]) || !std::isfinite(b.field_target[v * st.F + f]))
                bad("field angle or target not finite");
        }
    }
}

// ---- constraint rows (ExtraResiduals / TraceExtraResiduals) -----------------------------------

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
inline bool con_needs_extrema(const mko_constraint& c) {
    return c.type == MKO_CON_MAX_SEMI_DIAMETER || c.type == MKO_CON_MAX_INCIDENCE_ANGLE ||
           (c.type == MKO_CON_MIN_EDGE_THICKNESS && c.edge_height == MKO_EDGE_RAYS);
}
// rows that read the batch's paraxial_index: the paraxial trace (ParaxialConstraints::
// needs_trace) and the real rays at lambda_ref of the sine condition and distortion
inline bool con_needs_paraxial(const mko_constraint& c) {
    return c.type == MKO_CON_FOCAL_LENGTH || c.type == MKO_CON_PARAXIAL_FOCUS ||
           c.type == MKO_CON_PETZVAL || c.type == MKO_CON_SINE_CONDITION ||
           c.type == MKO_CON_DISTORTION;
}

// check_paraxial_constraints (src/optim/paraxial_residuals.cpp:152-255) and
// check_mechanical_constraints (src/optim/mechanical_constraints.cpp:148-203) for one item.
void check_constraint(const mko_constraint& c, int k, const mko_batch& b) {
    const auto fail = [k](const std::string& what) {
        throw std::invalid_argument("constraint " + std::to_string(k) + ": " + what);
    };
    const int S = b.surfaces;
    if (!(std::isfinite(c.weight) && c.weight >= 0.0)) fail("weight must be finite and >= 0");
    if (!(std::isfinite(c.shift) && c.shift >= 0.0)) fail("shift must be finite and >= 0");
    if (c.type == MKO_CON_GLASS_HULL) {
        if (c.surface < 0 || c.surface >= S) fail("surface out of range");
        if (c.field < 0 || c.field > 2) fail("component must be 0, 1 or 2");
        if (!(std::isfinite(c.target) && c.target >= 0.0)) fail("margin must be finite and >= 0");
        if (c.shift != 0.0) fail("no shift for the glass hull");
        if (!b.buchdahl) fail("batch has no Buchdahl coefficients");
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
            if (b.source != MKO_SOURCE_COLLIMATED) fail("needs a collimated source");
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
            if (b.source != MKO_SOURCE_COLLIMATED) fail("needs a collimated source");
            break;
        case MKO_CON_DISTORTION:
            if (c.mode != MKO_MODE_MAX) fail("defined as an upper bound only");
            if (c.target < 0.0) fail("D_max must be >= 0");
            if (c.field < 0) fail("field must be >= 0");
            if (b.source != MKO_SOURCE_COLLIMATED) fail("needs a collimated source");
            break;
        default:
            break;
        }
        if (con_needs_paraxial(c) && !b.paraxial_index) fail("batch has no paraxial_index");
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

// The constraint rows of a call: items in row order (without extrema, then with), on the host
// and (after upload) on the device.
struct RowPlan {
    std::vector<mko_constraint> items;
    int E = 0, Et = 0;
    bool paraxial = false;
    uint64_t surfaces = 0; // surfaces whose extrema the rows read
    uint64_t hsurf = 0, csurf = 0; // of those, the ones read for heights / for cos I
    const mko_constraint* dev = nullptr;
    // the glass hull of MKO_CON_GLASS_HULL rows: host array and its device copy
    const double* hull_host = nullptr;
    int hull_size = 0;
    const double* hull = nullptr;
    int count() const { return E + Et; }
};

// The layout of mko_extra_rows.glass_hull (header, records of 13 doubles, finite, scale > 0).
void check_glass_hull(const double* h, int size) {
    const auto fail = [](const std::string& what) {
        throw std::invalid_argument("glass hull: " + what);
    };
    if (!h || size < 8 + 13) fail("missing or shorter than one record");
    for (int i = 0; i < size; ++i)
        if (!std::isfinite(h[i])) fail("non-finite entry " + std::to_string(i));
    const double dim = h[0], n = h[1];
    if (!(dim == 0.0 || dim == 1.0 || dim == 2.0 || dim == 3.0)) fail("dimension not 0 ... 3");
    if (!(n >= 1.0 && n == std::floor(n) && 8.0 + 13.0 * n == double(size)))
        fail("size is not 8 + 13 records");
    if (dim < 2.0 && n != 1.0) fail("a point or segment hull has one record");
    for (int j = 0; j < 3; ++j)
        if (!(h[5 + j] > 0.0)) fail("scale must be > 0");
}

RowPlan plan_rows(const mko_batch& b, const mko_residual_settings& s) {
    RowPlan r;
    if (!s.extra || s.extra->count <= 0) return r;
    if (!s.extra->items) throw std::invalid_argument("constraint rows without items");
    for (int k = 0; k < s.extra->count; ++k) check_constraint(s.extra->items[k], k, b);
    for (int pass = 0; pass < 2; ++pass)
        for (int k = 0; k < s.extra->count; ++k)
            if (con_needs_extrema(s.extra->items[k]) == (pass == 1))
                r.items.push_back(s.extra->items[k]);
    for (const mko_constraint& c : r.items) {
        if (c.type == MKO_CON_GLASS_HULL && !r.hull_host) {
            check_glass_hull(s.extra->glass_hull, s.extra->glass_hull_size);
            r.hull_host = s.extra->glass_hull;
            r.hull_size = s.extra->glass_hull_size;
            if (c.target > 0.0 && r.hull_host[0] != 3.0)
                throw std::invalid_argument("glass hull: a margin needs a 3-d hull");
        }
        (con_needs_extrema(c) ? r.Et : r.E) += 1;
        r.paraxial = r.paraxial || con_needs_paraxial(c);
        if (con_needs_extrema(c)) {
            r.surfaces |= 1ull << c.surface;
            if (c.type == MKO_CON_MIN_EDGE_THICKNESS) r.surfaces |= 1ull << c.next_surface;
            if (c.type == MKO_CON_MAX_INCIDENCE_ANGLE) {
                r.csurf |= 1ull << c.surface;
            } else {
                r.hsurf |= 1ull << c.surface;
                if (c.type == MKO_CON_MIN_EDGE_THICKNESS) r.hsurf |= 1ull << c.next_surface;
            }
        }
    }
    return r;
}

// compute_paraxial's checks (src/paraxial/paraxial.cpp:104-129) on the paraxial index of a
// host batch that has paraxial rows.
void validate_paraxial(const mko_batch& b, int first, int count) {
    for (int v = first; v < first + count; ++v)
        for (int s = 0; s < b.surfaces; ++s) {
            const double n = b.paraxial_index[size_t(v) * b.surfaces + s];
            if (!(n > 0.0) || !std::isfinite(n))
                throw std::invalid_argument("variant " + std::to_string(v) +
                                            ": paraxial index not finite and > 0");
        }
}

// Device copy of (a range of) a host batch in one allocation.
struct DeviceBatch {
    Buffer mem;
    mko_batch view{};
    void upload(const mko_batch& h, int first, int count, hipStream_t stream) {
        const Strides st(h);
        const size_t V = size_t(count);
        double* base = mem.as<double>(std::max<size_t>(1, V * st.doubles()));
        view = h;
        view.variants = count;
        double* p = base;
        auto put = [&](const double* src, size_t per, const double** dst) {
            if (src == nullptr) {
                *dst = nullptr;
                return;
            }
            hip_check(hipMemcpyAsync(p, src + size_t(first) * per, V * per * sizeof(double),
                                     hipMemcpyHostToDevice, stream),
                      "upload batch");
            *dst = p;
            p += V * per;
        };
        put(h.curvature, st.S, &view.curvature);
        put(h.thickness, st.S, &view.thickness);
        put(h.index, st.SL, &view.index);
        put(h.pupil_radius, 1, &view.pupil_radius);
        put(h.field_angle, st.F, &view.field_angle);
        put(h.field_target, st.F, &view.field_target);
        put(st.refc ? h.reference_curvature : nullptr, st.F, &view.reference_curvature);
        put(h.paraxial_index, st.S, &view.paraxial_index);
        put(h.buchdahl, 4 * st.S, &view.buchdahl);
    }
    // Uninitialised device batch of `count` variants shaped like h (filled by a kernel).
    void allocate(const mko_batch& h, int count) {
        const Strides st(h);
        const size_t V = size_t(count);
        double* p = mem.as<double>(std::max<size_t>(1, V * st.doubles()));
        view = h;
        view.variants = count;
        auto take = [&](size_t per, const double** dst, bool present = true) {
            if (!present) {
                *dst = nullptr;
                return;
            }
            *dst = p;
            p += V * per;
        };
        take(st.S, &view.curvature);
        take(st.S, &view.thickness);
        take(st.SL, &view.index);
        take(1, &view.pupil_radius);
        take(st.F, &view.field_angle);
        take(st.F, &view.field_target);
        take(st.F, &view.reference_curvature, st.refc);
        take(st.S, &view.paraxial_index, st.pidx);
        take(4 * st.S, &view.buchdahl, st.bd);
    }
    double* mut(const double* p) { return const_cast<double*>(p); }
};

// Host batch made of selected variants of another batch (run templates of an LM subset).
struct GatheredBatch {
    std::vector<double> curv, thick, index, pupil, fa, ft, rc, pi, bd;
    mko_batch view{};
    const mko_batch& gather(const mko_batch& base, const int* ids, size_t n) {
        const Strides st(base);
        curv.resize(n * st.S);
        thick.resize(n * st.S);
        index.resize(n * st.SL);
        pupil.resize(n);
        fa.resize(n * st.F);
        ft.resize(n * st.F);
        rc.resize(st.refc ? n * st.F : 0);
        pi.resize(st.pidx ? n * st.S : 0);
        bd.resize(st.bd ? n * 4 * st.S : 0);
        for (size_t j = 0; j < n; ++j) {
            const int id = ids[j];
            if (id < 0 || id >= base.variants) throw std::out_of_range("run index outside the base batch");
            const size_t v = size_t(id);
            std::copy_n(base.curvature + v * st.S, st.S, curv.data() + j * st.S);
            std::copy_n(base.thickness + v * st.S, st.S, thick.data() + j * st.S);
            std::copy_n(base.index + v * st.SL, st.SL, index.data() + j * st.SL);
            pupil[j] = base.pupil_radius[v];
            std::copy_n(base.field_angle + v * st.F, st.F, fa.data() + j * st.F);
            std::copy_n(base.field_target + v * st.F, st.F, ft.data() + j * st.F);
            if (st.refc) std::copy_n(base.reference_curvature + v * st.F, st.F, rc.data() + j * st.F);
            if (st.pidx) std::copy_n(base.paraxial_index + v * st.S, st.S, pi.data() + j * st.S);
            if (st.bd) std::copy_n(base.buchdahl + v * 4 * st.S, 4 * st.S, bd.data() + j * 4 * st.S);
        }
        view = base;
        view.variants = int(n);
        view.curvature = curv.data();
        view.thickness = thick.data();
        view.index = index.data();
        view.pupil_radius = pupil.data();
        view.field_angle = fa.data();
        view.field_target = ft.data();
        view.reference_curvature = st.refc ? rc.data() : nullptr;
        view.paraxial_index = st.pidx ? pi.data() : nullptr;
        view.buchdahl = st.bd ? bd.data() : nullptr;
        return view;
    }
};

// ---- FD pattern (central scheme): fd_pattern, src/optim/variant_generator.cpp:177-237 -------

constexpr double kFdStepFactor = 6.0554544523933395e-06; // 2^(-52/3)

struct Pattern {
    std::vector<double> plus, minus, step_plus, step_minus;
    std::vector<uint8_t> stencil;
};

void fd_pattern(const mko_variables& v, const double* x0, Pattern& out, size_t at) {
    for (int k = 0; k < v.count; ++k) {
        const double x = x0[k];
        const double lo = v.lower[k], hi = v.upper[k];
        if (!std::isfinite(x) || !(x >= lo && x <= hi))
            throw std::invalid_argument("x0 not finite or outside the bounds");
        const double h = kFdStepFactor * std::max(std::abs(x), v.typical[k]);
        const double up = x + h, down = x - h;
        double xa, xb;
        uint8_t sten = MKO_STENCIL_CENTRAL;
        if (up <= hi && down >= lo) {
            xa = up;
            xb = down;
        } else {
            const bool towards_upper = hi - x >= x - lo;
            const double far = towards_upper ? x + 2.0 * h : x - 2.0 * h;
            if (towards_upper ? far <= hi : far >= lo) {
                sten = MKO_STENCIL_ONE_SIDED;
                xa = towards_upper ? up : down;
                xb = far;
            } else {
                sten = MKO_STENCIL_TRUNCATED;
                xa = up <= hi ? up : hi;
                xb = down >= lo ? down : lo;
            }
        }
        out.plus[at + k] = xa;
        out.minus[at + k] = xb;
        out.step_plus[at + k] = xa - x;
        out.step_minus[at + k] = xb - x;
        out.stencil[at + k] = sten;
    }
}

// ---- kernel timing (AdjointJacobianTiming-like, include/mikroptim/cuda/adjoint_jacobian.hpp:39) --

enum KernelFamily { kTrace = 0, kAdjoint = 1, kGram = 2, kFdDifference = 3, kPointBatch = 4,
                    kLmControl = 5, kFamilies = 6 };

// Accumulates the device time of each kernel family with HIP events recorded on the launch
// stream (the null stream) around each launch, when enabled. Recording does not block the host;
// the events are resolved when the totals are read (collect()).
class KernelTimer {
public:
    bool enabled = false;
    double ms[kFamilies] = {};
    int64_t launches[kFamilies] = {};
    ~KernelTimer() { drop(); }
    template <class F>
    void run(KernelFamily fam, F&& launch) {
        if (!enabled) {
            launch();
            return;
        }
        hipEvent_t a = event(), b = event();
        hip_check(hipEventRecord(a, nullptr), "hipEventRecord");
        launch();
        hip_check(hipEventRecord(b, nullptr), "hipEventRecord");
        pending_.push_back(Pending{fam, a, b});
    }
    // time of a launch measured elsewhere (events recorded inside a captured graph)
    void add(KernelFamily fam, float t) {
        ms[fam] += t;
        ++launches[fam];
    }
    void collect() {
        for (const Pending& p : pending_) {
            hip_check(hipEventSynchronize(p.b), "hipEventSynchronize");
            float t = 0.0f;
            hip_check(hipEventElapsedTime(&t, p.a, p.b), "hipEventElapsedTime");
            ms[p.fam] += t;
            ++launches[p.fam];
            pool_.push_back(p.a);
            pool_.push_back(p.b);
        }
        pending_.clear();
    }
    void reset() {
        collect();
        for (int i = 0; i < kFamilies; ++i) {
            ms[i] = 0.0;
            launches[i] = 0;
        }
    }

private:
    struct Pending {
        KernelFamily fam;
        hipEvent_t a, b;
    };
    hipEvent_t event() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    void drop() {
        for (const Pending& p : pending_) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        pending_.clear();
        pool_.clear();
    }
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
};

// ---- Engine --------------------------------------------------------------------------------

// ---- FP32 prefilter: the conservative decision (prefilter_decide,
// src/batch/batch_prefilter.cpp:236-329) --------------------------------------------------------

// check_prefilter_options (src/batch/batch_prefilter.cpp:40-52)
void check_prefilter_options(const mko_prefilter_options& o) {
    if (!(o.max_failed_fraction >= 0.0)) throw std::invalid_argument("max_failed_fraction must be >= 0");
    if (!(o.max_merit >= 0.0)) throw std::invalid_argument("max_merit must be >= 0");
    if (!(o.max_merit_factor >= 0.0)) throw std::invalid_argument("max_merit_factor must be >= 0");
    if (!(o.boundary >= 0.0) || !std::isfinite(o.boundary))
        throw std::invalid_argument("boundary margin must be finite and >= 0");
    if (!(o.opl_error_factor >= 0.0) || !std::isfinite(o.opl_error_factor))
        throw std::invalid_argument("opl_error_factor must be finite and >= 0");
    if (!(o.merit_relative >= 0.0 && o.merit_relative < 1.0))
        throw std::invalid_argument("merit_relative must be in [0, 1)");
}

// residual_scale (include/mikroptim/merit/ray_residual.hpp:132): sqrt(w_f w_l) / lambda_ref[mm]
double residual_scale_of(const mko_batch& b, int f, int l) {
    const double wf = b.field_weight[f], wl = b.wavelength_weight[l];
    if (!(wf > 0.0) || !(wl > 0.0)) return 0.0;
    double sf = 0.0, sl = 0.0;
    for (int i = 0; i < b.fields; ++i) sf += b.field_weight[i];
    for (int i = 0; i < b.wavelengths; ++i) sl += b.wavelength_weight[i];
    const double w = (wf / sf) * (wl / sl);
    if (!(w > 0.0)) return 0.0;
    return std::sqrt(w) / (b.reference_wavelength_nm * 1e-6);
}

void prefilter_decide(const mko_batch& b, const mko_residual_settings& s,
                      const mko_prefilter_options& o, const double* extra,
                      mko_prefilter_result& r) {
    check_prefilter_options(o);
    constexpr double kU = 0x1p-24; // float unit roundoff
    double scale_sq = 0.0;
    for (int f = 0; f < b.fields; ++f)
        for (int l = 0; l < b.wavelengths; ++l) {
            const double sc = residual_scale_of(b, f, l);
            scale_sq += sc * sc;
        }
    const double scale_norm = std::sqrt(scale_sq);
    const int64_t R = rays_of(b, s);
    const bool soft = s.failure == MKO_FAILURE_SOFT;
    const int V = b.variants;
    const double inf = std::numeric_limits<double>::infinity();
    auto too_many = [&](int64_t failed) { return double(failed) > o.max_failed_fraction * double(R); };
    for (int v = 0; v < V; ++v) {
        const double m = r.merit[v];
        const double root = std::sqrt(m < 0.0 ? 0.0 : m);
        const double delta = o.opl_error_factor * kU * r.optical_length[v] * scale_norm;
        const double rd = root - delta;
        const double lower = (0.0 < rd ? rd : 0.0) / (1.0 + o.merit_relative);
        const double upper = (root + delta) / (1.0 - o.merit_relative);
        const double e = extra ? extra[v] : 0.0;
        r.merit_lower[v] = lower * lower + e;
        const bool finite = std::isfinite(root) && std::isfinite(delta);
        const bool penalties = soft && r.failed[v] > 0;
        const int64_t uncertain = r.uncertain_ok[v] + r.uncertain_failed[v];
        r.merit_upper[v] = uncertain == 0 && finite && !penalties ? upper * upper + e : inf;
        if (!finite) r.merit_lower[v] = e;
        uint8_t rej = MKO_PREFILTER_KEEP;
        if (too_many(r.failed[v])) rej |= MKO_PREFILTER_FAILED_RAYS;
        if (r.merit_lower[v] > o.max_merit) rej |= MKO_PREFILTER_MERIT_ABSOLUTE;
        r.reject[v] = rej;
    }
    r.reference_merit = inf;
    r.reference_variant = -1;
    if (std::isfinite(o.max_merit_factor)) {
        for (int v = 0; v < V; ++v) {
            const int64_t uncertain = r.uncertain_ok[v] + r.uncertain_failed[v];
            if (uncertain != 0 || (r.reject[v] & MKO_PREFILTER_FAILED_RAYS) != 0 || r.failed[v] >= R)
                continue;
            if (r.merit_upper[v] < r.reference_merit) {
                r.reference_merit = r.merit_upper[v];
                r.reference_variant = v;
            }
        }
        if (r.reference_variant >= 0) {
            const double limit = o.max_merit_factor * r.reference_merit;
            for (int v = 0; v < V; ++v)
                if (r.merit_lower[v] > limit) r.reject[v] |= MKO_PREFILTER_MERIT_RELATIVE;
        }
    }
    int kept = 0;
    for (int v = 0; v < V; ++v) kept += r.reject[v] == MKO_PREFILTER_KEEP;
    r.kept = kept;
}

class Engine {
public:
    Engine(int device, size_t memory_limit) : device_(device), limit_(memory_limit) {
        int n = 0;
        hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (n == 0) throw DeviceError("no HIP device");
        if (device < 0 || device >= n) throw std::invalid_argument("device index out of range");
    }

    // Usable bytes for one call: free device memory minus a reserve of max(256 MiB, 5 %)
    // (TraceEngine::plan_tranches, src/cuda/trace_engine.cu:199), capped by the option.
    size_t budget() const {
        size_t free_b = 0, total = 0;
        hip_check(hipMemGetInfo(&free_b, &total), "hipMemGetInfo");
        const size_t reserve = std::max<size_t>(size_t(256) << 20, total / 20);
        size_t b = free_b > reserve ? free_b - reserve : 0;
        if (limit_ > 0) b = std::min(b, limit_);
        return b;
    }

    void residuals(const mko_batch& b, const mko_residual_settings& s, double* res, double* opd,
                   uint8_t* status) {
        DeviceGuard g(device_);
        check_batch(b);
        validate_values(b, 0, b.variants);
        RowPlan rp = plan_rows(b, s);
        if (rp.paraxial) validate_paraxial(b, 0, b.variants);
        upload_rows(rp);
        const int64_t T = trace_rows_of(b, s), R = T + rp.count(), rays = rays_of(b, s);
        const size_t tuples = size_t(b.fields) * b.wavelengths, S = size_t(b.surfaces);
        const size_t per = Strides(b).doubles() * 8 + size_t(R) * 8 + size_t(rays) * 9 +
                           (rp.Et ? (tuples + 1) * S * 16 : 0);
        const int tranche = tranche_size(b.variants, per);
        for (int first = 0; first < b.variants; first += tranche) {
            const int count = std::min(tranche, b.variants - first);
            batch_.upload(b, first, count, nullptr);
            double* dres = out_res_.as<double>(size_t(count) * R);
            double* dopd = opd ? out_opd_.as<double>(size_t(count) * rays) : nullptr;
            uint8_t* dst = status ? out_st_.as<uint8_t>(size_t(count) * rays) : nullptr;
            double* ep = rp.Et ? ext_part_.as<double>(size_t(count) * tuples * S * 2) : nullptr;
            double* ex = rp.Et ? ext_.as<double>(size_t(count) * S * 2) : nullptr;
            timer.run(kTrace, [&] {
                dev_check(mko_internal::trace_extrema(&batch_.view, &s, 0, dres, R, dopd, dst,
                                                      nullptr, nullptr, nullptr, ep, ex,
                                                      rp.surfaces, nullptr));
            });
            if (rp.count())
                dev_check(mko_internal::constraint_rows(&batch_.view, rp.dev, rp.count(),
                                                        rp.paraxial, ex, dres, R, T, nullptr,
                                                        rp.hull));
            if (res) copy_out(res + size_t(first) * R, dres, size_t(count) * R * 8);
            if (opd) copy_out(opd + size_t(first) * rays, dopd, size_t(count) * rays * 8);
            if (status) copy_out(status + size_t(first) * rays, dst, size_t(count) * rays);
        }
        hip_check(hipDeviceSynchronize(), "residuals");
    }

    // FP32 prefilter of a host batch: stats on the device (tranches), then the decision.
    void prefilter(const mko_batch& b, const mko_residual_settings& s,
                   const mko_prefilter_options& o, const double* extra, mko_prefilter_result& r) {
        DeviceGuard g(device_);
        check_batch(b);
        check_prefilter_options(o);
        validate_values(b, 0, b.variants);
        if (!r.merit || !r.optical_length || !r.failed || !r.uncertain_ok || !r.uncertain_failed ||
            !r.merit_lower || !r.merit_upper || !r.reject)
            throw std::invalid_argument("prefilter result arrays missing");
        const size_t T = size_t(b.fields) * b.wavelengths;
        const size_t per = Strides(b).doubles() * 8 + T * 24 + 40;
        const int tranche = tranche_size(b.variants, per);
        for (int first = 0; first < b.variants; first += tranche) {
            const int count = std::min(tranche, b.variants - first);
            batch_.upload(b, first, count, nullptr);
            void* ws = part_.reserve(size_t(count) * T * 24);
            double* dm = out_merit_.as<double>(size_t(count) * 2);
            int64_t* dc = out_failed_.as<int64_t>(size_t(count) * 3);
            timer.run(kTrace, [&] {
                dev_check(mko_internal::prefilter_stats(&batch_.view, &s, o.boundary, ws, dm,
                                                        dm + count, dc, dc + count,
                                                        dc + 2 * size_t(count), nullptr));
            });
            copy_out(r.merit + first, dm, size_t(count) * 8);
            copy_out(r.optical_length + first, dm + count, size_t(count) * 8);
            copy_out(r.failed + first, dc, size_t(count) * 8);
            copy_out(r.uncertain_ok + first, dc + count, size_t(count) * 8);
            copy_out(r.uncertain_failed + first, dc + 2 * size_t(count), size_t(count) * 8);
        }
        hip_check(hipDeviceSynchronize(), "prefilter");
        prefilter_decide(b, s, o, extra, r);
    }

    // Surface extrema [V][S][2] of a host batch (the MERIT trace's max reductions).
    void surface_extrema(const mko_batch& b, const mko_residual_settings& s, double* out) {
        DeviceGuard g(device_);
        check_batch(b);
        validate_values(b, 0, b.variants);
        const size_t T = size_t(b.fields) * b.wavelengths, S = size_t(b.surfaces);
        const size_t per = Strides(b).doubles() * 8 + T * 16 + (T + 1) * S * 16;
        const int tranche = tranche_size(b.variants, per);
        for (int first = 0; first < b.variants; first += tranche) {
            const int count = std::min(tranche, b.variants - first);
            batch_.upload(b, first, count, nullptr);
            double* part = part_.as<double>(size_t(count) * T);
            int* pc = partc_.as<int>(size_t(count) * T * 2);
            double* ep = ext_part_.as<double>(size_t(count) * T * S * 2);
            double* ex = ext_.as<double>(size_t(count) * S * 2);
            timer.run(kTrace, [&] {
                dev_check(mko_internal::trace_extrema(&batch_.view, &s, 1, nullptr, 0, nullptr,
                                                      nullptr, part, pc, nullptr, ep, ex, ~0ull,
                                                      nullptr));
            });
            copy_out(out + size_t(first) * S * 2, ex, size_t(count) * S * 2 * 8);
        }
        hip_check(hipDeviceSynchronize(), "surface extrema");
    }

    void merit(const mko_batch& b, const mko_residual_settings& s, const double* extra,
               double* merit, int64_t* failed, int64_t* hard) {
        DeviceGuard g(device_);
        check_batch(b);
        validate_values(b, 0, b.variants);
        RowPlan rp = plan_rows(b, s);
        if (rp.paraxial) validate_paraxial(b, 0, b.variants);
        upload_rows(rp);
        const size_t T = size_t(b.fields) * b.wavelengths;
        const size_t per = Strides(b).doubles() * 8 + T * 16 + 32 + rows_bytes(rp, b);
        const int tranche = tranche_size(b.variants, per);
        for (int first = 0; first < b.variants; first += tranche) {
            const int count = std::min(tranche, b.variants - first);
            batch_.upload(b, first, count, nullptr);
            merit_device(batch_.view, s, rp, extra ? extra + first : nullptr, count);
            copy_out(merit + first, d_merit_, size_t(count) * 8);
            copy_out(failed + first, d_failed_, size_t(count) * 8);
            copy_out(hard + first, d_hard_, size_t(count) * 8);
        }
        hip_check(hipDeviceSynchronize(), "merit");
    }

    // MERIT of a device-resident batch: (ray merit + sum e^2 [+ extra]) + sum t^2 into md / fd /
    // hd (NULL: d_merit_ / d_failed_ / d_hard_); states: the LM's final ray states.
    void merit_device(const mko_batch& dev, const mko_residual_settings& s, const RowPlan& rp,
                      const double* extra_host, int count,
                      const mko_internal::RayStates* states = nullptr, double* md = nullptr,
                      int64_t* fd = nullptr, int64_t* hd = nullptr) {
        const size_t T = size_t(dev.fields) * dev.wavelengths, S = size_t(dev.surfaces);
        double* part = part_.as<double>(size_t(count) * T);
        int* pc = partc_.as<int>(size_t(count) * T * 2);
        if (!md) {
            md = d_merit_ = out_merit_.as<double>(size_t(count));
            fd = d_failed_ = out_failed_.as<int64_t>(size_t(count));
            hd = d_hard_ = out_hard_.as<int64_t>(size_t(count));
        }
        const double* dextra = nullptr;
        if (extra_host) {
            double* e = extra_.as<double>(size_t(count));
            hip_check(hipMemcpy(e, extra_host, size_t(count) * 8, hipMemcpyHostToDevice), "extra");
            dextra = e;
        }
        double* ep = rp.Et ? ext_part_.as<double>(size_t(count) * T * S * 2) : nullptr;
        double* ex = rp.Et ? ext_.as<double>(size_t(count) * S * 2) : nullptr;
        timer.run(kTrace, [&] {
            dev_check(mko_internal::trace_extrema(&dev, &s, 1, nullptr, 0, nullptr, nullptr, part,
                                                  pc, states, ep, ex, rp.surfaces, nullptr));
            const double* esum = dextra;
            const double* tsum = nullptr;
            if (rp.count()) {
                double* rows = rows_.as<double>(size_t(count) * rp.count());
                dev_check(mko_internal::constraint_rows(&dev, rp.dev, rp.count(), rp.paraxial, ex,
                                                        rows, rp.count(), 0, nullptr, rp.hull));
                double* e = (rp.E || dextra) ? esum_.as<double>(size_t(count)) : nullptr;
                double* t = rp.Et ? tsum_.as<double>(size_t(count)) : nullptr;
                dev_check(mko_internal::rows_merit(count, rows, rp.count(), 0, rp.E, rp.Et, dextra,
                                                   e, t, nullptr));
                esum = e;
                tsum = t;
            }
            dev_check(mko_internal::merit_finalize(count, int(T), part, pc, esum, tsum, md, fd, hd,
                                                   nullptr));
        });
    }

    // J / r0 / A / g for `runs` points; base has 1 variant or one per run. Rows per run:
    // trace rows + constraint rows.
    // failed (NULL: unknown): the failed-ray count of each run's point, which the soft split
    // then takes instead of tracing the points; rows_settings (NULL: s): the settings of the
    // constraint rows' central differences.
    void jacobian(const mko_batch& base, const mko_residual_settings& s, const mko_variables& vars,
                  int runs, const double* x0, int method, int precision, double* J, double* r0,
                  double* a, double* g, const int64_t* failed = nullptr,
                  const mko_residual_settings* rows_settings = nullptr) {
        DeviceGuard guard(device_);
        check_batch(base);
        if (base.variants != 1 && base.variants != runs)
            throw std::invalid_argument("base must have one variant or one per run");
        if (vars.count < 1) throw std::invalid_argument("no variables");
        if (method != MKO_JACOBIAN_FD && method != MKO_JACOBIAN_ADJOINT)
            throw std::invalid_argument("unknown Jacobian method");
        if (precision != MKO_PRECISION_FP64 && precision != MKO_PRECISION_MIXED &&
            precision != MKO_PRECISION_FP64_MFMA)
            throw std::invalid_argument("unknown precision");
        validate_values(base, 0, base.variants);
        if (soft_split_applies(s, vars, method, precision) && !J && !r0 && (a || g)) {
            soft_split(base, s, vars, runs, x0, a, g, failed);
            return;
        }
        const mko_residual_settings& rs = rows_settings ? *rows_settings : s;
        RowPlan rp = plan_rows(base, s);
        if (rp.paraxial) validate_paraxial(base, 0, base.variants);
        upload_rows(rp);
        const int P = vars.count;
        const int64_t T = trace_rows_of(base, s), R = T + rp.count(), rays = rays_of(base, s);
        const bool fd = method == MKO_JACOBIAN_FD;
        // The LM consumes only J^T J and J^T r0: for the mixed-precision adjoint those come from
        // the fused kernel and J is never formed (hard failure mode, <= 96 variables).
        // With glass variables the fused kernel needs stored ray states (its index adjoints are
        // in the stored-state kernels): a MERIT trace of the same points keeps them first.
        const bool glass_fused = !fd && precision == MKO_PRECISION_MIXED && !J && !r0 &&
                                 (a || g) && !fused_ok(s, vars) && fused_ok(s, vars, true);
        const bool fused = glass_fused || (!fd && precision == MKO_PRECISION_MIXED && !J && !r0 &&
                                           (a || g) && fused_ok(s, vars));
        const bool pattern = fd || rp.count() > 0;
        const int per_run = fd ? 2 * P + 1 : 1;
        const size_t Tt = size_t(base.fields) * base.wavelengths, S = size_t(base.surfaces);
        const size_t per = size_t(per_run) * (Strides(base).doubles() * 8 +
                                              (fd ? size_t(R) * 8 + size_t(rays) : 0) +
                                              (fd && rp.Et ? (Tt + 1) * S * 16 : 0)) +
                           (fused ? Tt * (size_t(P) * P * 4 + size_t(P) * 64 + 24) + 1024 +
                                        (glass_fused ? size_t(rays) * 56 + Tt * 16 + 32 : 0)
                                  : size_t(P) * R * 8 + size_t(R) * 8 + size_t(rays)) +
                           size_t(P) * (P + 1) * 8 + 64 * size_t(P) +
                           (!fd ? rows_fd_bytes(rp, base, P) : 0);
        const int chunk = tranche_size(runs, per);
        check_glass(base, vars);
        upload_vars(vars);
        Pattern pat;
        if (pattern) {
            pat.plus.resize(size_t(runs) * P);
            pat.minus.resize(size_t(runs) * P);
            pat.step_plus.resize(size_t(runs) * P);
            pat.step_minus.resize(size_t(runs) * P);
            pat.stencil.resize(size_t(runs) * P);
            for (int i = 0; i < runs; ++i) fd_pattern(vars, x0 + size_t(i) * P, pat, size_t(i) * P);
        } else {
            for (size_t i = 0; i < size_t(runs) * P; ++i)
                if (!std::isfinite(x0[i])) throw std::invalid_argument("x0 not finite");
        }
        if (base.variants == 1) tmpl_.upload(base, 0, 1, nullptr);
        for (int first = 0; first < runs; first += chunk) {
            const int count = std::min(chunk, runs - first);
            if (base.variants != 1) tmpl_.upload(base, first, count, nullptr);
            double* dx = x_.as<double>(size_t(count) * P);
            hip_check(hipMemcpy(dx, x0 + size_t(first) * P, size_t(count) * P * 8,
                                hipMemcpyHostToDevice), "x0");
            double *dplus = nullptr, *dminus = nullptr, *dsp = nullptr, *dsm = nullptr;
            uint8_t* dsten = nullptr;
            if (pattern) {
                const size_t o = size_t(first) * P, n = size_t(count) * P;
                dplus = plus_.as<double>(n);
                dminus = minus_.as<double>(n);
                dsp = sp_.as<double>(n);
                dsm = sm_.as<double>(n);
                dsten = sten_.as<uint8_t>(n);
                hip_check(hipMemcpy(dplus, pat.plus.data() + o, n * 8, hipMemcpyHostToDevice), "pattern");
                hip_check(hipMemcpy(dminus, pat.minus.data() + o, n * 8, hipMemcpyHostToDevice), "pattern");
                hip_check(hipMemcpy(dsp, pat.step_plus.data() + o, n * 8, hipMemcpyHostToDevice), "steps");
                hip_check(hipMemcpy(dsm, pat.step_minus.data() + o, n * 8, hipMemcpyHostToDevice), "steps");
                hip_check(hipMemcpy(dsten, pat.stencil.data() + o, n, hipMemcpyHostToDevice), "stencil");
            }
            work_.allocate(base, count * per_run);
            dev_check(mko_internal::point_batch(
                &tmpl_.view, nullptr, base.variants != 1, count, per_run, &dvars_, dx,
                fd ? dplus : nullptr, fd ? dminus : nullptr, work_.mut(work_.view.curvature),
                work_.mut(work_.view.thickness), work_.mut(work_.view.index),
                work_.mut(work_.view.pupil_radius), work_.mut(work_.view.field_angle),
                work_.mut(work_.view.field_target), work_.mut(work_.view.reference_curvature),
                work_.mut(work_.view.paraxial_index),
                    work_.mut(work_.view.buchdahl), nullptr));
            if (fused) {
                double* da = a_.as<double>(size_t(count) * P * P);
                double* dg = g_.as<double>(size_t(count) * P);
                const size_t wbytes =
                    mko_adjoint_normal_workspace(count, base.fields, base.wavelengths, P);
                void* ws = jac_.reserve(wbytes);
                mko_internal::RayStates st{};
                if (glass_fused) {
                    const int64_t nrays = rays_of(base, s);
                    st.data = states_.as<double>(size_t(count) * size_t(nrays) * 7);
                    st.slot = nullptr;
                    st.slot_offset = 0;
                    st.slots = count;
                    st.rays = nrays;
                    mko_residual_settings sp = s;
                    sp.precision = MKO_PRECISION_FP64; // the states come from the FP64 trace
                    sp.extra = nullptr;
                    timer.run(kTrace, [&] {
                        dev_check(mko_internal::merit_states(
                            &work_.view, &sp, smerit_.as<double>(size_t(count)),
                            sfailed_.as<int64_t>(size_t(count)), shard_.as<int64_t>(size_t(count)),
                            spart_.as<double>(size_t(count) * Tt),
                            spc_.as<int>(size_t(count) * Tt * 2), st, nullptr));
                    });
                }
                timer.run(kAdjoint, [&] {
                    dev_check(mko_internal::adjoint_normal_states(&work_.view, &s, &vars_host_dev_,
                                                                  da, dg, ws, wbytes, st,
                                                                  nullptr));
                });
                if (rp.count()) {
                    const int E = rp.count();
                    double* Je = je_.as<double>(size_t(count) * P * E);
                    double* re = re_.as<double>(size_t(count) * E);
                    rows_fd(tmpl_.view, nullptr, base.variants != 1, count, dx, dplus, dminus, dsp,
                            dsm, dsten, rs, rp, P, Je, re, E, 0);
                    dev_check(mko_internal::extra_normal(count, P, E, Je, re, da, dg, nullptr));
                }
                if (a) copy_out(a + size_t(first) * P * P, da, size_t(count) * P * P * 8);
                if (g) copy_out(g + size_t(first) * P, dg, size_t(count) * P * 8);
                hip_check(hipDeviceSynchronize(), "jacobian");
                continue;
            }
            double* dJ = jac_.as<double>(size_t(count) * P * R);
            double* dr0 = r0_.as<double>(size_t(count) * R);
            if (fd) {
                fd_rows(work_.view, s, rp, count, P, dsp, dsm, dsten, T, R, rays, dJ, dr0);
            } else {
                timer.run(kAdjoint, [&] {
                    dev_check(mko_internal::adjoint_jacobian(&work_.view, &s, &vars_host_dev_, dJ,
                                                             dr0, nullptr, R, nullptr));
                });
                if (rp.count())
                    rows_fd(tmpl_.view, nullptr, base.variants != 1, count, dx, dplus, dminus, dsp,
                            dsm, dsten, rs, rp, P, dJ, dr0, R, T);
            }
            if (a || g) {
                double* da = a_.as<double>(size_t(count) * P * P);
                double* dg = g_.as<double>(size_t(count) * P);
                timer.run(kGram, [&] {
                    dev_check(gram(precision, dJ, dr0, count, P, R, da, dg));
                });
                if (a) copy_out(a + size_t(first) * P * P, da, size_t(count) * P * P * 8);
                if (g) copy_out(g + size_t(first) * P, dg, size_t(count) * P * 8);
            }
            if (J) copy_out(J + size_t(first) * P * R, dJ, size_t(count) * P * R * 8);
            if (r0) copy_out(r0 + size_t(first) * R, dr0, size_t(count) * R * 8);
            hip_check(hipDeviceSynchronize(), "jacobian");
        }
    }

    // Merit of `runs` points: copies of the template with x applied (lm_point_batch).
    void point_merit(const mko_batch& base, const mko_residual_settings& s, const mko_variables& vars,
                     int runs, const double* x, double* merit, int64_t* failed, int64_t* hard) {
        DeviceGuard guard(device_);
        const int P = vars.count;
        for (size_t i = 0; i < size_t(runs) * P; ++i)
            if (!std::isfinite(x[i])) throw std::invalid_argument("x not finite");
        RowPlan rp = plan_rows(base, s);
        if (rp.paraxial) validate_paraxial(base, 0, base.variants);
        upload_rows(rp);
        const size_t T = size_t(base.fields) * base.wavelengths;
        const size_t per =
            Strides(base).doubles() * 8 + T * 16 + 64 + size_t(P) * 8 + rows_bytes(rp, base);
        const int chunk = tranche_size(runs, per);
        check_glass(base, vars);
        upload_vars(vars);
        if (base.variants == 1) tmpl_.upload(base, 0, 1, nullptr);
        for (int first = 0; first < runs; first += chunk) {
            const int count = std::min(chunk, runs - first);
            if (base.variants != 1) tmpl_.upload(base, first, count, nullptr);
            double* dx = x_.as<double>(size_t(count) * P);
            hip_check(hipMemcpy(dx, x + size_t(first) * P, size_t(count) * P * 8,
                                hipMemcpyHostToDevice), "x");
            work_.allocate(base, count);
            dev_check(mko_internal::point_batch(
                &tmpl_.view, nullptr, base.variants != 1, count, 1, &dvars_, dx, nullptr, nullptr,
                work_.mut(work_.view.curvature), work_.mut(work_.view.thickness),
                work_.mut(work_.view.index), work_.mut(work_.view.pupil_radius),
                work_.mut(work_.view.field_angle), work_.mut(work_.view.field_target),
                work_.mut(work_.view.reference_curvature), work_.mut(work_.view.paraxial_index),
                work_.mut(work_.view.buchdahl), nullptr));
            merit_device(work_.view, s, rp, nullptr, count);
            copy_out(merit + first, d_merit_, size_t(count) * 8);
            copy_out(failed + first, d_failed_, size_t(count) * 8);
            copy_out(hard + first, d_hard_, size_t(count) * 8);
        }
        hip_check(hipDeviceSynchronize(), "point merit");
    }

    // ---- device-resident variants for the LM control on the device -------------------------

    // Variable types / surfaces for the point batches (host copy kept for the adjoint map).
    // `owner` (an LM session's id, never reused) skips the upload when the engine holds that
    // session's variables already; sessions that share the engine call this before each of their
    // device cycles.
    void set_variables(const mko_variables& v, uint64_t owner = 0) {
        if (owner && owner == vars_owner_) return;
        DeviceGuard guard(device_);
        upload_vars(v);
        vars_owner_ = owner;
    }

    // Normal equations a [count][P][P], g [count][P] (device) of `count` points xd [count][P]
    // (device) on the device template tmpl (one variant, or one per run selected by tidx).
    // Variables must have been set with set_variables; bounds lower / upper / typical [P] on the
    // device (FD pattern).
    void jacobian_dev(const mko_batch& tmpl, const int32_t* tidx, int count, const double* xd,
                      const mko_residual_settings& s, int P, const double* lower,
                      const double* upper, const double* typical, int method, int precision,
                      double* ad, double* gd, const mko_internal::RayStates* states = nullptr,
                      const mko_residual_settings* rows_settings = nullptr) {
        DeviceGuard guard(device_);
        if (count <= 0) return;
        const mko_residual_settings& rs = rows_settings ? *rows_settings : s;
        RowPlan rp = plan_rows(tmpl, s);
        upload_rows(rp);
        const int64_t T = trace_rows_of(tmpl, s), R = T + rp.count(), rays = rays_of(tmpl, s);
        const bool fd = method == MKO_JACOBIAN_FD;
        const bool fused = !fd && precision == MKO_PRECISION_MIXED &&
                           fused_ok(s, vars_host_dev_, states != nullptr);
        const bool pattern = fd || rp.count() > 0;
        const int per_run = fd ? 2 * P + 1 : 1;
        const size_t Tt = size_t(tmpl.fields) * tmpl.wavelengths, S = size_t(tmpl.surfaces);
        const size_t per = size_t(per_run) * (Strides(tmpl).doubles() * 8 +
                                              (fd ? size_t(R) * 8 + size_t(rays) : 0) +
                                              (fd && rp.Et ? (Tt + 1) * S * 16 : 0)) +
                           (fused ? Tt * (size_t(P) * P * 4 + size_t(P) * 64 + 24) + 1024
                                  : size_t(P) * R * 8 + size_t(R) * 8) +
                           size_t(P) * 48 + 64 + (!fd ? rows_fd_bytes(rp, tmpl, P) : 0);
        const int chunk = tranche_size(count, per);
        for (int first = 0; first < count; first += chunk) {
            const int c = std::min(chunk, count - first);
            const double* x = xd + size_t(first) * P;
            double *dplus = nullptr, *dminus = nullptr, *dsp = nullptr, *dsm = nullptr;
            uint8_t* dsten = nullptr;
            if (pattern) {
                dplus = plus_.as<double>(size_t(c) * P);
                dminus = minus_.as<double>(size_t(c) * P);
                dsp = sp_.as<double>(size_t(c) * P);
                dsm = sm_.as<double>(size_t(c) * P);
                dsten = sten_.as<uint8_t>(size_t(c) * P);
                timer.run(kLmControl, [&] {
                    mko_lmdev::launch_fd_pattern(c, P, x, lower, upper, typical, dplus, dminus, dsp,
                                                 dsm, dsten, nullptr);
                    hip_check(hipGetLastError(), "fd pattern launch");
                });
            }
            const int32_t* ti = tidx ? tidx + first : nullptr;
            work_.allocate(tmpl, c * per_run);
            timer.run(kPointBatch, [&] {
                dev_check(mko_internal::point_batch(
                    &tmpl, ti, 0, c, per_run, &dvars_, x, fd ? dplus : nullptr,
                    fd ? dminus : nullptr, work_.mut(work_.view.curvature),
                    work_.mut(work_.view.thickness), work_.mut(work_.view.index),
                    work_.mut(work_.view.pupil_radius), work_.mut(work_.view.field_angle),
                    work_.mut(work_.view.field_target), work_.mut(work_.view.reference_curvature),
                    work_.mut(work_.view.paraxial_index),
                    work_.mut(work_.view.buchdahl), nullptr));
            });
            double* a_out = ad + size_t(first) * P * P;
            double* g_out = gd + size_t(first) * P;
            if (fused) {
                const size_t wbytes =
                    mko_adjoint_normal_workspace(c, tmpl.fields, tmpl.wavelengths, P);
                void* ws = jac_.reserve(wbytes);
                mko_internal::RayStates st{};
                if (states) {
                    st = *states;
                    if (st.slot) st.slot += first;
                    else st.slot_offset += first;
                }
                timer.run(kAdjoint, [&] {
                    dev_check(mko_internal::adjoint_normal_states(&work_.view, &s, &vars_host_dev_,
                                                                  a_out, g_out, ws, wbytes, st,
                                                                  nullptr));
                });
                if (rp.count()) {
                    const int E = rp.count();
                    double* Je = je_.as<double>(size_t(c) * P * E);
                    double* re = re_.as<double>(size_t(c) * E);
                    rows_fd(tmpl, ti, 0, c, x, dplus, dminus, dsp, dsm, dsten, rs, rp, P, Je, re, E,
                            0);
                    dev_check(mko_internal::extra_normal(c, P, E, Je, re, a_out, g_out, nullptr));
                }
                continue;
            }
            double* dJ = jac_.as<double>(size_t(c) * P * R);
            double* dr0 = r0_.as<double>(size_t(c) * R);
            if (fd) {
                fd_rows(work_.view, s, rp, c, P, dsp, dsm, dsten, T, R, rays, dJ, dr0);
            } else {
                timer.run(kAdjoint, [&] {
                    dev_check(mko_internal::adjoint_jacobian(&work_.view, &s, &vars_host_dev_, dJ,
                                                             dr0, nullptr, R, nullptr));
                });
                if (rp.count())
                    rows_fd(tmpl, ti, 0, c, x, dplus, dminus, dsp, dsm, dsten, rs, rp, P, dJ, dr0, R,
                            T);
            }
            timer.run(kGram, [&] {
                dev_check(gram(precision, dJ, dr0, c, P, R, a_out, g_out));
            });
        }
    }

    // MERIT of `count` points xd [count][P] (device) on the device template: merit, failed,
    // hard-failed counts into device arrays.
    void merit_dev(const mko_batch& tmpl, const int32_t* tidx, int count, const double* xd,
                   const mko_residual_settings& s, int P, double* md, int64_t* fd,
                   int64_t* hd, const mko_internal::RayStates* states = nullptr) {
        DeviceGuard guard(device_);
        if (count <= 0) return;
        RowPlan rp = plan_rows(tmpl, s);
        upload_rows(rp);
        const size_t T = size_t(tmpl.fields) * tmpl.wavelengths;
        const size_t per =
            Strides(tmpl).doubles() * 8 + T * 16 + 64 + size_t(P) * 8 + rows_bytes(rp, tmpl);
        const int chunk = tranche_size(count, per);
        for (int first = 0; first < count; first += chunk) {
            const int c = std::min(chunk, count - first);
            work_.allocate(tmpl, c);
            timer.run(kPointBatch, [&] {
                dev_check(mko_internal::point_batch(
                    &tmpl, tidx ? tidx + first : nullptr, 0, c, 1, &dvars_,
                    xd + size_t(first) * P, nullptr, nullptr, work_.mut(work_.view.curvature),
                    work_.mut(work_.view.thickness), work_.mut(work_.view.index),
                    work_.mut(work_.view.pupil_radius), work_.mut(work_.view.field_angle),
                    work_.mut(work_.view.field_target), work_.mut(work_.view.reference_curvature),
                    work_.mut(work_.view.paraxial_index),
                    work_.mut(work_.view.buchdahl), nullptr));
            });
            mko_internal::RayStates st{};
            if (states) {
                st = *states;
                if (st.slot) st.slot += first;
                else st.slot_offset += first;
            }
            merit_device(work_.view, s, rp, nullptr, c, states ? &st : nullptr, md + first,
                         fd + first, hd + first);
        }
    }

    // The mixed-precision normal equations come from the fused kernel under these conditions.
    // Soft failure mode with the mixed-precision adjoint: a run whose point has no failed ray has
    // no penalty residual and no penalty derivative, so its normal equations are those of hard
    // mode, which the fused kernel forms; runs with failed rays take the FP64 path. The choice
    // depends only on the run's own point (its failed-ray count).
    static bool soft_split_applies(const mko_residual_settings& s, const mko_variables& v,
                                   int method, int precision) {
        if (s.failure != MKO_FAILURE_SOFT || method != MKO_JACOBIAN_ADJOINT ||
            precision != MKO_PRECISION_MIXED)
            return false;
        mko_residual_settings hs = s;
        hs.failure = MKO_FAILURE_HARD;
        return fused_ok(hs, v);
    }
    // failed_in (NULL: traced here): the failed-ray count of each point. The trace rows of the
    // clean runs are hard mode's; their constraint rows keep the caller's settings (the
    // differences' perturbed points may fail rays).
    void soft_split(const mko_batch& base, const mko_residual_settings& s, const mko_variables& vars,
                    int runs, const double* x0, double* a, double* g,
                    const int64_t* failed_in = nullptr) {
        const size_t P = size_t(vars.count);
        const size_t nr = size_t(runs);
        std::vector<int64_t> failed(nr);
        if (failed_in) {
            std::copy_n(failed_in, nr, failed.begin());
        } else {
            std::vector<double> merit(nr);
            std::vector<int64_t> hard(nr);
            point_merit(base, s, vars, runs, x0, merit.data(), failed.data(), hard.data());
        }
        std::vector<int> clean, other;
        for (int i = 0; i < runs; ++i) (failed[size_t(i)] == 0 ? clean : other).push_back(i);
        mko_residual_settings hs = s;
        hs.failure = MKO_FAILURE_HARD;
        auto part = [&](const std::vector<int>& ids, const mko_residual_settings& ps, int prec) {
            if (ids.empty()) return;
            const size_t n = ids.size();
            std::vector<double> xs(n * P), as(a ? n * P * P : 0), gs(g ? n * P : 0);
            for (size_t j = 0; j < n; ++j) std::copy_n(x0 + size_t(ids[j]) * P, P, xs.data() + j * P);
            GatheredBatch gb;
            const mko_batch& b = base.variants == 1 ? base : gb.gather(base, ids.data(), n);
            jacobian(b, ps, vars, int(n), xs.data(), MKO_JACOBIAN_ADJOINT, prec, nullptr, nullptr,
                     a ? as.data() : nullptr, g ? gs.data() : nullptr, nullptr, &s);
            for (size_t j = 0; j < n; ++j) {
                const size_t i = size_t(ids[j]);
                if (a) std::copy_n(as.data() + j * P * P, P * P, a + i * P * P);
                if (g) std::copy_n(gs.data() + j * P, P, g + i * P);
            }
        };
        part(clean, hs, MKO_PRECISION_MIXED);
        part(other, s, MKO_PRECISION_FP64);
    }

    // Autotuner (no counterpart in the reference, whose tuple-pack sizes and blocks per SM are
    // fixed: include/mikroptim/cuda/tuple_packs.hpp:61, src/cuda/wavefront_block.hpp:504-526).
    // For the batch's shape the launch shapes of the LM cycle's two dominant kernels are timed on
    // up to 64 of the caller's points and the fastest kept: the tuple-pack segment of the MERIT
    // trace (packing off, or each segment that holds the largest tuple) and the waves per
    // workgroup of the fused normal equations (where they apply). Each candidate runs twice and
    // the second run's kernel time decides. The choice is cached per shape (a later call on the
    // same shape times nothing) and becomes the engine's tuning. Results do not depend on the
    // choice except for the fused kernel's rounding.
    mko_internal::Tuning autotune(const mko_batch& base, const mko_residual_settings& s,
                                  const mko_variables& v, int runs, const double* x) {
        check_batch(base);
        if (runs < 1) throw std::invalid_argument("autotune needs at least one point");
        if (v.count < 1) throw std::invalid_argument("no variables");
        int nmax = 0;
        for (int f = 0; f < base.fields; ++f)
            nmax = std::max(nmax, s.rings * (base.pupil_mode[f] == MKO_PUPIL_AXIS ? 1 : s.spokes));
        std::string key = std::to_string(base.surfaces) + "/" + std::to_string(base.fields) + "/" +
                          std::to_string(base.wavelengths) + "/" + std::to_string(s.rings) + "/" +
                          std::to_string(s.spokes) + "/" + std::to_string(v.count) + "/" +
                          std::to_string(s.failure) + "/" + std::to_string(base.source) + "/" +
                          std::to_string(s.extra ? s.extra->count : 0);
        for (int f = 0; f < base.fields; ++f) key += char('0' + base.pupil_mode[f]);
        const int full_rows = tuning_.full_row_traces; // not a launch shape: kept
        auto hit = tuned_.find(key);
        if (hit != tuned_.end()) {
            tuning_ = hit->second;
            tuning_.full_row_traces = full_rows;
            return tuning_;
        }
        const int n = std::min(runs, 64);
        // the first n points: a batch of one variant per point is cut to its first n variants
        mko_batch b = base;
        if (b.variants > 1) {
            if (b.variants != runs)
                throw std::invalid_argument("base must have one variant or one per point");
            b.variants = n;
        }
        const size_t P = size_t(v.count);
        std::vector<double> merit(n), a(size_t(n) * P * P), g(size_t(n) * P);
        std::vector<int64_t> failed(n), hard(n);
        // the caller's timing totals are kept aside while the trials use the timer
        timer.collect();
        double saved_ms[kFamilies];
        int64_t saved_n[kFamilies];
        std::copy(timer.ms, timer.ms + kFamilies, saved_ms);
        std::copy(timer.launches, timer.launches + kFamilies, saved_n);
        const bool was_enabled = timer.enabled;
        auto trial = [&](const mko_internal::Tuning& t, KernelFamily fam, auto&& work) {
            mko_internal::TuningScope ts(t);
            double best = 0.0;
            for (int rep = 0; rep < 2; ++rep) {
                timer.reset();
                timer.enabled = true;
                work();
                timer.collect();
                best = timer.ms[fam];
            }
            return best;
        };
        mko_internal::Tuning chosen{};
        chosen.full_row_traces = full_rows;
        try {
            std::vector<int> packs{0};
            if (nmax <= 128 && base.fields * base.wavelengths >= 2 &&
                s.precision == MKO_PRECISION_FP64) {
                packs = {256};
                for (int seg : {32, 64, 128})
                    if (seg >= nmax) packs.push_back(seg);
            }
            double best = -1.0;
            for (int pk : packs) {
                const double ms = trial(mko_internal::Tuning{pk, 0, full_rows}, kTrace, [&] {
                    point_merit(b, s, v, n, x, merit.data(), failed.data(), hard.data());
                });
                if (best < 0.0 || ms < best) {
                    best = ms;
                    chosen.pack_rays = pk;
                }
            }
            if (s.failure == MKO_FAILURE_HARD && fused_ok(s, v, true)) {
                std::vector<int> waves{2, 4};
                if (nmax <= 64) waves.insert(waves.begin(), 1);
                best = -1.0;
                for (int w : waves) {
                    const double ms = trial(mko_internal::Tuning{chosen.pack_rays, w, full_rows}, kAdjoint, [&] {
                        jacobian(b, s, v, n, x, MKO_JACOBIAN_ADJOINT, MKO_PRECISION_MIXED,
                                 nullptr, nullptr, a.data(), g.data());
                    });
                    if (best < 0.0 || ms < best) {
                        best = ms;
                        chosen.fused_waves = w;
                    }
                }
            }
        } catch (...) {
            timer.reset();
            std::copy(saved_ms, saved_ms + kFamilies, timer.ms);
            std::copy(saved_n, saved_n + kFamilies, timer.launches);
            timer.enabled = was_enabled;
            throw;
        }
        timer.reset();
        std::copy(saved_ms, saved_ms + kFamilies, timer.ms);
        std::copy(saved_n, saved_n + kFamilies, timer.launches);
        timer.enabled = was_enabled;
        tuned_[key] = chosen;
        tuning_ = chosen;
        return chosen;
    }

    // Normal equations of stored Jacobians: the host's FMA chain (bitwise), or FP64 MFMA for
    // MKO_PRECISION_FP64_MFMA.
    static int gram(int precision, const double* J, const double* r0, int runs, int P, int64_t R,
                    double* a, double* g) {
        return precision == MKO_PRECISION_FP64_MFMA
                   ? mko_gram_mfma_device(J, r0, runs, P, R, a, g, nullptr)
                   : mko_gram_device(J, r0, runs, P, R, a, g, nullptr);
    }

    // The fused mixed-precision normal equations apply: hard failure mode, at most 96 variables;
    // glass variables only where the reverse sweeps start from an LM MERIT trace's stored ray
    // states (`states`), whose kernels carry the index adjoints.
    static bool fused_ok(const mko_residual_settings& s, const mko_variables& v,
                         bool states = false) {
        bool glass = false;
        for (int k = 0; k < v.count; ++k) glass = glass || v.type[k] == MKO_VAR_GLASS;
        return s.failure == MKO_FAILURE_HARD && v.count <= 96 && (!glass || states);
    }

    int device() const { return device_; }
    const mko_internal::Tuning& tuning() const { return tuning_; }
    void set_tuning(const mko_internal::Tuning& t) { tuning_ = t; }
    KernelTimer timer;

private:
    // Device copy of the plan's items (kept while the items do not change).
    void upload_rows(RowPlan& rp) {
        if (rp.count() == 0) return;
        const size_t bytes = rp.items.size() * sizeof(mko_constraint);
        const char* src = reinterpret_cast<const char*>(rp.items.data());
        void* d = items_.reserve(bytes);
        if (items_host_.size() != bytes || std::memcmp(items_host_.data(), src, bytes) != 0) {
            hip_check(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice), "constraint items");
            items_host_.assign(src, src + bytes);
        }
        rp.dev = static_cast<const mko_constraint*>(d);
        if (rp.hull_host) {
            const size_t hb = size_t(rp.hull_size) * sizeof(double);
            const char* hs = reinterpret_cast<const char*>(rp.hull_host);
            void* hd = hull_.reserve(hb);
            if (hull_host_.size() != hb || std::memcmp(hull_host_.data(), hs, hb) != 0) {
                hip_check(hipMemcpy(hd, hs, hb, hipMemcpyHostToDevice), "glass hull");
                hull_host_.assign(hs, hs + hb);
            }
            rp.hull = static_cast<const double*>(hd);
        }
    }
    // device bytes per variant of a MERIT / RESIDUALS call for its constraint rows
    static size_t rows_bytes(const RowPlan& rp, const mko_batch& b) {
        if (rp.count() == 0) return 0;
        const size_t T = size_t(b.fields) * b.wavelengths, S = size_t(b.surfaces);
        return size_t(rp.count()) * 8 + 16 + (rp.Et ? (T + 1) * S * 16 : 0);
    }
    // device bytes per run of rows_fd
    static size_t rows_fd_bytes(const RowPlan& rp, const mko_batch& b, int P) {
        if (rp.count() == 0) return 0;
        const size_t T = size_t(b.fields) * b.wavelengths, S = size_t(b.surfaces);
        // the candidate rays' buffers of the run's point: its batch, partials and extrema, ray
        // flags (at most 2048 rays per tuple), tuple counts and the candidate list
        const size_t cand = rp.Et ? Strides(b).doubles() * 8 + T * (16 + S * 16 + 2048 + 12) +
                                        size_t(kCandidateCap) * 4 + S * 16 + 64
                                  : 0;
        return size_t(2 * P + 1) * (Strides(b).doubles() * 8 + size_t(rp.count()) * 8 +
                                    (rp.Et ? T * 16 + (T + 1) * S * 16 : 0)) +
               size_t(P + 1) * rp.count() * 8 + cand;
    }

    // FD Jacobian of the central FD batch fdb (count runs x (2P + 1) variants): RESIDUALS of
    // every variant with its constraint rows (traced rows from the variant's own extrema) at
    // stride R, then fd_derivative of all R rows (fd_jacobian_cpu, src/optim/fd_jacobian.cpp:182).
    void fd_rows(const mko_batch& fdb, const mko_residual_settings& s, const RowPlan& rp,
                 int count, int P, const double* dsp, const double* dsm, const uint8_t* dsten,
                 int64_t T, int64_t R, int64_t rays, double* dJ, double* dr0) {
        const size_t n = size_t(count) * (2 * P + 1);
        const size_t Tt = size_t(fdb.fields) * fdb.wavelengths, S = size_t(fdb.surfaces);
        double* fres = fdres_.as<double>(n * R);
        uint8_t* fst = fdst_.as<uint8_t>(n * rays);
        double* ep = rp.Et ? ext_part_.as<double>(n * Tt * S * 2) : nullptr;
        double* ex = rp.Et ? ext_.as<double>(n * S * 2) : nullptr;
        timer.run(kTrace, [&] {
            dev_check(mko_internal::trace_extrema(&fdb, &s, 0, fres, R, nullptr, fst, nullptr,
                                                  nullptr, nullptr, ep, ex, rp.surfaces,
                                                  nullptr));
            if (rp.count())
                dev_check(mko_internal::constraint_rows(&fdb, rp.dev, rp.count(), rp.paraxial, ex,
                                                        fres, R, T, nullptr, rp.hull));
        });
        int64_t* dch = ch_.as<int64_t>(size_t(count) * P);
        timer.run(kFdDifference, [&] {
            dev_check(mko_fd_difference_device(fres, R, fst, rays, count, P, R, dsp, dsm, dsten,
                                               dJ, dr0, dch, nullptr));
        });
    }

    // Jacobian of the constraint rows by central differences for the adjoint paths
    // (extra_rows_untraced / extra_rows_cpu, src/optim/adjoint_jacobian.cpp:245-337): the FD
    // batch of the runs' points (2P + 1 variants per run), the rows of every variant (traced rows
    // from each variant's own MERIT extrema), fd_derivative per row into
    // J[(i * P + p) * jstride + joffset + e] and r0[i * jstride + joffset + e].
    void rows_fd(const mko_batch& tmpl, const int32_t* tidx, int template_per_run, int count,
                 const double* x, const double* plus, const double* minus, const double* sp,
                 const double* sm, const uint8_t* sten, const mko_residual_settings& s,
                 const RowPlan& rp, int P, double* J, double* r0, int64_t jstride,
                 int64_t joffset) {
        const int per_run = 2 * P + 1;
        const size_t n = size_t(count) * per_run;
        const int E = rp.count();
        rowfd_.allocate(tmpl, int(n));
        timer.run(kPointBatch, [&] {
            dev_check(mko_internal::point_batch(
                &tmpl, tidx, template_per_run, count, per_run, &dvars_, x, plus, minus,
                rowfd_.mut(rowfd_.view.curvature), rowfd_.mut(rowfd_.view.thickness),
                rowfd_.mut(rowfd_.view.index), rowfd_.mut(rowfd_.view.pupil_radius),
                rowfd_.mut(rowfd_.view.field_angle), rowfd_.mut(rowfd_.view.field_target),
                rowfd_.mut(rowfd_.view.reference_curvature),
                rowfd_.mut(rowfd_.view.paraxial_index),
                rowfd_.mut(rowfd_.view.buchdahl), nullptr));
        });
        double* ex = nullptr;
        if (rp.Et) ex = rows_extrema(tmpl, tidx, template_per_run, count, x, s, rp, P);
        double* rows = rows_.as<double>(n * E);
        dev_check(mko_internal::constraint_rows(&rowfd_.view, rp.dev, E, rp.paraxial, ex, rows, E,
                                                0, nullptr, rp.hull));
        timer.run(kFdDifference, [&] {
            dev_check(mko_internal::fd_difference(rows, E, nullptr, 0, count, P, E, sp, sm, sten, J,
                                                  r0, nullptr, jstride, joffset, nullptr));
        });
    }

    // Extrema [count * (2P + 1)][S][2] of the rows' FD batch (rowfd_). The rows need the largest
    // height / smallest cos I over all rays at each difference point; a central difference moves
    // a ray height by ~1e-5 of itself, far less than the rays' heights differ. So each run's
    // point is traced whole first, and only the candidate rays -- within a relative 1e-3 of the
    // run's extrema at its point on the surfaces the rows read, or not ending with status Ok --
    // are traced at the 2P difference points. A check then asks that the other rays would have
    // to move 10 times as far as the candidates' extrema did to reach them; a chunk where a run
    // misses that margin (or has more than kCandidateCap candidates) is traced whole. The
    // extrema, and so the rows and their differences, are those of tracing every ray: the same
    // ray traces, the same maximum.
    static constexpr int kCandidateCap = 8192;
    double* rows_extrema(const mko_batch& tmpl, const int32_t* tidx, int template_per_run,
                         int count, const double* x, const mko_residual_settings& s,
                         const RowPlan& rp, int P) {
        const int per_run = 2 * P + 1;
        const size_t n = size_t(count) * per_run;
        const size_t Tt = size_t(tmpl.fields) * tmpl.wavelengths, S = size_t(tmpl.surfaces);
        double* part = part_.as<double>(n * Tt);
        int* pc = partc_.as<int>(n * Tt * 2);
        double* ep = ext_part_.as<double>(n * Tt * S * 2);
        double* ex = ext_.as<double>(n * S * 2);
        auto full = [&] {
            timer.run(kTrace, [&] {
                dev_check(mko_internal::trace_extrema(&rowfd_.view, &s, 1, nullptr, 0, nullptr,
                                                      nullptr, part, pc, nullptr, ep, ex,
                                                      rp.surfaces, nullptr));
            });
        };
        if (mko_internal::thread_tuning().full_row_traces || s.precision != MKO_PRECISION_FP64) {
            full();
            return ex;
        }
        // the runs' points, every ray
        rowbase_.allocate(tmpl, count);
        timer.run(kPointBatch, [&] {
            dev_check(mko_internal::point_batch(
                &tmpl, tidx, template_per_run, count, 1, &dvars_, x, nullptr, nullptr,
                rowbase_.mut(rowbase_.view.curvature), rowbase_.mut(rowbase_.view.thickness),
                rowbase_.mut(rowbase_.view.index), rowbase_.mut(rowbase_.view.pupil_radius),
                rowbase_.mut(rowbase_.view.field_angle), rowbase_.mut(rowbase_.view.field_target),
                rowbase_.mut(rowbase_.view.reference_curvature),
                rowbase_.mut(rowbase_.view.paraxial_index), rowbase_.mut(rowbase_.view.buchdahl),
                nullptr));
        });
        const size_t c = size_t(count);
        const int64_t rays = rays_of(tmpl, s);
        double* part0 = cand_part_.as<double>(c * Tt);
        int* pc0 = cand_pc_.as<int>(c * Tt * 2);
        double* ep0 = cand_ep_.as<double>(c * Tt * S * 2);
        double* ex0 = cand_ex_.as<double>(c * S * 2);
        timer.run(kTrace, [&] {
            dev_check(mko_internal::trace_extrema(&rowbase_.view, &s, 1, nullptr, 0, nullptr,
                                                  nullptr, part0, pc0, nullptr, ep0, ex0,
                                                  rp.surfaces, nullptr));
        });
        uint8_t* flags = cand_flags_.as<uint8_t>(c * size_t(rays));
        int* tcount = cand_tcount_.as<int>(c * Tt);
        int* tstart = cand_tstart_.as<int>(c * (Tt + 1));
        int* list = cand_list_.as<int>(c * kCandidateCap);
        int* others = cand_count_.as<int>(c);
        int* overflow = cand_over_.as<int>(c);
        int* flag = cand_flag_.as<int>(c);
        constexpr double kDelta = 1e-3, kMargin = 10.0;
        timer.run(kTrace, [&] {
            dev_check(mko_internal::candidate_extrema(
                &rowbase_.view, &rowfd_.view, &s, per_run, rp.hsurf, rp.csurf, kDelta, ex0, flags,
                tcount, others, tstart, list, kCandidateCap, overflow, ex, nullptr));
        });
        dev_check(mko_internal::extrema_check(count, per_run, int(S), rp.hsurf, rp.csurf, kDelta,
                                              ex, ex0, others, overflow, kMargin, flag, nullptr));
        std::vector<int> flags_h(c);
        hip_check(hipMemcpy(flags_h.data(), flag, c * sizeof(int), hipMemcpyDeviceToHost),
                  "extrema check");
        ++candidate_chunks_;
        if (std::any_of(flags_h.begin(), flags_h.end(), [](int f) { return f != 0; })) {
            ++candidate_fallbacks_;
            full();
        }
        return ex;
    }

public:
    // chunks whose rows' differences used candidate rays, and those traced whole after the check
    std::pair<int64_t, int64_t> candidate_stats() const {
        return {candidate_chunks_, candidate_fallbacks_};
    }

private:
    int tranche_size(int items, size_t per_item) const {
        if (items <= 0) return 1;
        const size_t b = budget();
        const size_t n = per_item ? b / per_item : size_t(items);
        if (n == 0) throw DeviceError("too little device memory for one variant");
        return int(std::min<size_t>(n, size_t(items)));
    }
    void copy_out(void* host, const void* dev, size_t bytes) {
        hip_check(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost), "download");
    }
    void upload_vars(const mko_variables& v) {
        int32_t* t = vtype_.as<int32_t>(size_t(v.count));
        int32_t* s = vsurf_.as<int32_t>(size_t(v.count));
        hip_check(hipMemcpy(t, v.type, size_t(v.count) * 4, hipMemcpyHostToDevice), "vars");
        hip_check(hipMemcpy(s, v.batch_surface, size_t(v.count) * 4, hipMemcpyHostToDevice), "vars");
        dvars_ = v;
        dvars_.type = t;
        dvars_.batch_surface = s;
        dvars_.component = nullptr;
        if (v.component) {
            int32_t* c = vcomp_.as<int32_t>(size_t(v.count));
            hip_check(hipMemcpy(c, v.component, size_t(v.count) * 4, hipMemcpyHostToDevice), "vars");
            dvars_.component = c;
        }
        vars_host_dev_ = v; // the adjoint entry point reads type/surface on the host
        vars_owner_ = 0;
    }

    int device_;
    size_t limit_;
    mko_internal::Tuning tuning_;
    std::map<std::string, mko_internal::Tuning> tuned_; // autotuner's choices per shape
    DeviceBatch batch_, tmpl_, work_, rowfd_, rowbase_;
    Buffer cand_part_, cand_pc_, cand_ep_, cand_ex_, cand_flags_, cand_tcount_, cand_tstart_,
        cand_list_, cand_count_, cand_over_, cand_flag_;
    int64_t candidate_chunks_ = 0, candidate_fallbacks_ = 0;
    Buffer out_res_, out_opd_, out_st_, part_, partc_, out_merit_, out_failed_, out_hard_, extra_;
    Buffer ext_part_, ext_, rows_, esum_, tsum_, je_, re_, items_, hull_;
    std::vector<char> items_host_, hull_host_;
    Buffer states_, smerit_, sfailed_, shard_, spart_, spc_; // ray states of the glass mixed path
    Buffer x_, plus_, minus_, jac_, r0_, fdres_, fdst_, sp_, sm_, sten_, ch_, a_, g_, vtype_, vsurf_,
        vcomp_;
    double* d_merit_ = nullptr;
    int64_t* d_failed_ = nullptr;
    int64_t* d_hard_ = nullptr;
    mko_variables dvars_{};
    mko_variables vars_host_dev_{};
    uint64_t vars_owner_ = 0; // id of the LM session whose variables dvars_ holds
};

// ---- batched Levenberg-Marquardt (src/optim/levenberg_marquardt.cpp:37-572) --------------------

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
constexpr double kModelDamping = 1e-12;

void parallel_for(int n, int threads, const std::function<void(int)>& fn) {
    if (n <= 0) return;
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    threads = std::min(threads, n);
    if (threads <= 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    std::exception_ptr err;
    std::mutex m;
    for (int t = 0; t < threads; ++t) {
        const int a = int(int64_t(n) * t / threads), b = int(int64_t(n) * (t + 1) / threads);
        pool.emplace_back([&, a, b] {
            try {
                for (int i = a; i < b; ++i) fn(i);
            } catch (...) {
                std::lock_guard<std::mutex> l(m);
                if (!err) err = std::current_exception();
            }
        });
    }
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

struct Run {
    std::vector<double> x, trial, a, g, d;
    double merit = 0.0;
    int64_t failed = 0, hard = 0;
    double mu = 0.0, nu = 2.0;
    int status = MKO_LM_RUNNING;
    bool need_jacobian = true, rejected = false, solved = false;
    double step_norm = 0.0, predicted = 0.0, gradient_norm = 0.0, model_decrease = kNaN;
    int iterations = 0, accepted = 0, jacobians = 0;
    std::vector<double> history;
};

inline double scale_of(double x, double typical) { return std::max(std::abs(x), typical); }

// In-place Cholesky of the lower triangle of an n x n column-major matrix.
bool cholesky(std::vector<double>& m, size_t n) {
    for (size_t j = 0; j < n; ++j) {
        double diag = m[j * n + j];
        for (size_t k = 0; k < j; ++k) diag -= m[k * n + j] * m[k * n + j];
        if (!(diag > 0.0) || !std::isfinite(diag)) return false;
        const double l = std::sqrt(diag);
        m[j * n + j] = l;
        for (size_t i = j + 1; i < n; ++i) {
            double s = m[j * n + i];
            for (size_t k = 0; k < j; ++k) s -= m[k * n + i] * m[k * n + j];
            m[j * n + i] = s / l;
        }
    }
    return true;
}

void cholesky_solve(const std::vector<double>& l, size_t n, std::vector<double>& b) {
    for (size_t i = 0; i < n; ++i) {
        double s = b[i];
        for (size_t k = 0; k < i; ++k) s -= l[k * n + i] * b[k];
        b[i] = s / l[i * n + i];
    }
    // back substitution, each sum in ascending k (cholesky_solve,
    // src/optim/levenberg_marquardt.cpp:94-112)
    for (size_t i = n; i-- > 0;) {
        double s = b[i];
        for (size_t k = i + 1; k < n; ++k) s -= l[i * n + k] * b[k];
        b[i] = s / l[i * n + i];
    }
}

struct Bounds {
    const double *lower, *upper, *typical;
    bool blocks(size_t k, double x, double dir) const {
        if (lower[k] == upper[k]) return true;
        return (x <= lower[k] && dir < 0.0) || (x >= upper[k] && dir > 0.0);
    }
};

void prepare_step(Run& s, const Bounds& b, size_t n, const mko_lm_options& o) {
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
        if (!b.blocks(k, s.x[k], -s.g[k])) {
            const double den = std::sqrt(akk * s.merit);
            if (den > 0.0) gmax = std::max(gmax, std::abs(s.g[k]) / den);
        }
    }
    s.gradient_norm = gmax;
}

// (A_FF + mu D_FF) delta_F = -g_F on the free variables with the active set (solve_step).
void solve_step(Run& s, const Bounds& b, size_t n, double mu) {
    std::vector<char> fixed(n);
    for (size_t k = 0; k < n; ++k) fixed[k] = b.blocks(k, s.x[k], -s.g[k]);
    std::vector<double> delta(n, 0.0), m, rhs;
    std::vector<size_t> idx;
    s.solved = true;
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
        if (!cholesky(m, f)) {
            s.solved = false;
            return;
        }
        cholesky_solve(m, f, rhs);
        for (size_t j = 0; j < f; ++j) delta[idx[j]] = rhs[j];
        bool changed = false;
        for (size_t k : idx)
            if (b.blocks(k, s.x[k], delta[k])) {
                fixed[k] = 1;
                changed = true;
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

class Lm {
public:
    ~Lm() { graph_release(); }
    // launch shapes of this session: the engine's tuning when the session was created
    const mko_internal::Tuning& tuning() const { return tuning_; }
    Lm(Engine& eng, const mko_batch& base, const mko_residual_settings& s, const mko_variables& v,
       int runs, const double* x0, const mko_lm_options& o, int precision)
        : eng_(eng), base_(base), settings_(s), vars_(v), runs_(runs), o_(o), precision_(precision),
          tuning_(eng.tuning()) {
        check_batch(base);
        if (runs < 1) throw std::invalid_argument("no runs");
        if (base.variants != 1 && base.variants != runs)
            throw std::invalid_argument("base must have one variant or one per run");
        P_ = size_t(v.count);
        if (P_ < 1) throw std::invalid_argument("no variables");
        if (o.max_iterations < 0 || !(o.initial_damping > 0.0) || !(o.max_damping > o.initial_damping))
            throw std::invalid_argument("invalid LM options");
        if (o.jacobian_method != MKO_JACOBIAN_FD && o.jacobian_method != MKO_JACOBIAN_ADJOINT)
            throw std::invalid_argument("unknown Jacobian method");
        for (size_t k = 0; k < P_; ++k) {
            if (std::isnan(v.lower[k]) || std::isnan(v.upper[k]) || v.lower[k] > v.upper[k])
                throw std::invalid_argument("invalid bounds");
            if (!(v.typical[k] > 0.0) || !std::isfinite(v.typical[k]))
                throw std::invalid_argument("typical must be finite and > 0");
        }
        for (size_t i = 0; i < size_t(runs) * P_; ++i) {
            const size_t k = i % P_;
            if (!std::isfinite(x0[i]) || x0[i] < v.lower[k] || x0[i] > v.upper[k])
                throw std::invalid_argument("start point not finite or outside the bounds");
        }
        // constraint rows: checked here, items owned by the session
        const RowPlan rp = plan_rows(base, s);
        if (rp.paraxial) validate_paraxial(base, 0, base.variants);
        if (s.extra && s.extra->count > 0) {
            items_own_.assign(s.extra->items, s.extra->items + s.extra->count);
            if (s.extra->glass_hull && s.extra->glass_hull_size > 0)
                hull_own_.assign(s.extra->glass_hull,
                                 s.extra->glass_hull + s.extra->glass_hull_size);
            extra_own_ = mko_extra_rows{int32_t(items_own_.size()), items_own_.data(),
                                        hull_own_.empty() ? nullptr : hull_own_.data(),
                                        int32_t(hull_own_.size())};
            settings_.extra = &extra_own_;
        } else {
            settings_.extra = nullptr;
        }
        bounds_ = Bounds{v.lower, v.upper, v.typical};
        x0_.assign(x0, x0 + size_t(runs) * P_);
        weight_.assign(size_t(runs), s.fail_weight);
        rounds_.assign(size_t(runs), 1);
        device_ = o.host_control != 1 && P_ <= size_t(mko_lmdev::kMaxVariables);
        graph_wanted_ = o.host_control == 0;
        if (device_) dev_setup();
        start_round(std::vector<int>(), x0_);
    }

    // One LM cycle (plus, when a round ends, the next penalty round). false when done.
    bool cycle() {
        if (done_) return false;
        if (!round_cycle()) {
            finish_round();
            if (!next_round()) {
                done_ = true;
                return false;
            }
        }
        return true;
    }

    int graph_cycles() const { return graph_cycles_; }

    // Split form of cycle() for running several LMs at once: cycle_begin launches a captured
    // cycle on this LM's own stream without waiting (1), or, when this LM's cycle is not a
    // graph, runs cycle() to the end (2; 0: already done); cycle_end waits for a launched cycle
    // and returns what cycle() would (true while runs are active). Between the two calls the
    // caller may begin other LMs' cycles, which then run concurrently on the device.
    int cycle_begin() {
        if (done_) return 0;
        if (device_ && graph_eligible() && graph_ready() && graph_fits()) {
            graph_launch();
            pending_ = 1;
            return 1;
        }
        sync_result_ = cycle();
        pending_ = 2;
        return 2;
    }
    bool cycle_end() {
        const int p = pending_;
        pending_ = 0;
        if (p == 2) return sync_result_;
        if (p != 1) return !done_;
        if (!graph_finish()) {
            finish_round();
            if (!next_round()) {
                done_ = true;
                return false;
            }
        }
        return true;
    }

    void results(double* x, mko_lm_run* out, int* cycles) {
        if (device_ && !done_) dev_download_state();
        const std::vector<Run>& src = (done_ || !ids_.empty()) ? final_ : state_;
        for (int i = 0; i < runs_; ++i) {
            const Run& r = src[size_t(i)];
            std::copy(r.x.begin(), r.x.end(), x + size_t(i) * P_);
            mko_lm_run& o = out[i];
            o.merit = r.merit;
            o.failed_rays = r.failed;
            o.hard_failed_rays = r.hard;
            o.status = r.status;
            o.iterations = r.iterations;
            o.accepted = r.accepted;
            o.jacobians = r.jacobians;
            o.damping = r.mu;
            o.gradient_norm = r.gradient_norm;
            o.model_decrease = r.model_decrease;
            o.fail_weight = weight_[size_t(i)];
            o.penalty_rounds = rounds_[size_t(i)];
        }
        if (cycles) *cycles = cycles_;
    }

private:
    // ---- one pass of levenberg_marquardt over the runs `ids` (empty = all) ----
    void start_round(const std::vector<int>& ids, const std::vector<double>& xs) {
        if (device_) {
            dev_start_round(ids, xs);
            return;
        }
        ids_ = ids;
        const int n = ids.empty() ? runs_ : int(ids.size());
        state_.assign(size_t(n), Run());
        std::vector<int> list;
        for (int i = 0; i < n; ++i) {
            Run& s = state_[size_t(i)];
            s.x.assign(xs.begin() + std::ptrdiff_t(size_t(i) * P_), xs.begin() + std::ptrdiff_t(size_t(i + 1) * P_));
            s.trial = s.x;
            s.a.assign(P_ * P_, 0.0);
            s.g.assign(P_, 0.0);
            s.d.assign(P_, 0.0);
            s.mu = o_.initial_damping;
            list.push_back(i);
        }
        evaluate(list, false);
        for (size_t j = 0; j < list.size(); ++j) {
            Run& s = state_[size_t(list[j])];
            s.merit = f_[j];
            s.failed = fl_[j];
            s.hard = hf_[j];
            s.history.push_back(s.merit);
            if (!std::isfinite(s.merit)) s.status = MKO_LM_NUMERICAL_ERROR;
            else if (s.merit == 0.0) s.status = MKO_LM_CONVERGED_MERIT;
        }
    }

    // template batch for the runs of this round: base (one variant) or the listed runs' variants
    const mko_batch& template_for(const std::vector<int>& list, std::vector<int>& global) {
        global.clear();
        for (int i : list) global.push_back(ids_.empty() ? i : ids_[size_t(i)]);
        if (base_.variants == 1) return base_;
        gather(global);
        return *gathered_;
    }

    void gather(const std::vector<int>& global) {
        gathered_ = &gathered_batch_.gather(base_, global.data(), global.size());
    }

    void gather_x(const std::vector<int>& list, bool trial) {
        xs_.resize(list.size() * P_);
        for (size_t j = 0; j < list.size(); ++j) {
            const Run& s = state_[size_t(list[j])];
            const std::vector<double>& src = trial ? s.trial : s.x;
            std::copy(src.begin(), src.end(), xs_.begin() + std::ptrdiff_t(j * P_));
        }
    }

    void evaluate(const std::vector<int>& list, bool trial) {
        gather_x(list, trial);
        f_.assign(list.size(), 0.0);
        fl_.assign(list.size(), 0);
        hf_.assign(list.size(), 0);
        if (list.empty()) return;
        std::vector<int> global;
        const mko_batch& tmpl = template_for(list, global);
        eng_.point_merit(tmpl, settings_, vars_, int(list.size()), xs_.data(), f_.data(),
                         fl_.data(), hf_.data());
    }

    bool round_cycle() {
        if (!device_) return host_round_cycle();
        return graph_eligible() ? graph_round_cycle() : dev_round_cycle();
    }

    bool host_round_cycle() {
        const int n = int(state_.size());
        std::vector<int> list;
        for (int i = 0; i < n; ++i) {
            const Run& s = state_[size_t(i)];
            if (s.status == MKO_LM_RUNNING && s.need_jacobian && s.iterations < o_.max_iterations)
                list.push_back(i);
        }
        if (!list.empty()) {
            gather_x(list, false);
            a_.assign(list.size() * P_ * P_, 0.0);
            g_.assign(list.size() * P_, 0.0);
            std::vector<int> global;
            const mko_batch& tmpl = template_for(list, global);
            std::vector<int64_t> failed(list.size()); // the points' failed rays, for the soft split
            for (size_t j = 0; j < list.size(); ++j) failed[j] = state_[size_t(list[j])].failed;
            eng_.jacobian(tmpl, settings_, vars_, int(list.size()), xs_.data(), o_.jacobian_method,
                          precision_, nullptr, nullptr, a_.data(), g_.data(), failed.data());
            for (size_t j = 0; j < list.size(); ++j) {
                Run& s = state_[size_t(list[j])];
                std::copy_n(a_.data() + j * P_ * P_, P_ * P_, s.a.begin());
                std::copy_n(g_.data() + j * P_, P_, s.g.begin());
                s.need_jacobian = false;
                ++s.jacobians;
                bool finite = true;
                for (double v : s.a) finite = finite && std::isfinite(v);
                for (double v : s.g) finite = finite && std::isfinite(v);
                if (!finite) {
                    s.status = MKO_LM_NUMERICAL_ERROR;
                    continue;
                }
                prepare_step(s, bounds_, P_, o_);
            }
            parallel_for(int(list.size()), o_.threads, [&](int j) {
                Run& s = state_[size_t(list[size_t(j)])];
                if (s.status == MKO_LM_RUNNING) solve_step(s, bounds_, P_, kModelDamping);
            });
            for (int i : list) {
                Run& s = state_[size_t(i)];
                if (s.status != MKO_LM_RUNNING) continue;
                s.model_decrease = s.solved ? s.predicted / s.merit : kNaN;
                if (s.gradient_norm <= o_.gtol) s.status = MKO_LM_CONVERGED_GRADIENT;
                else if (s.solved && s.predicted >= 0.0 && s.predicted <= o_.ftol * s.merit)
                    s.status = MKO_LM_CONVERGED_MERIT;
                else if (s.solved && s.step_norm <= o_.xtol) s.status = MKO_LM_CONVERGED_STEP;
            }
        }
        list.clear();
        for (int i = 0; i < n; ++i) {
            Run& s = state_[size_t(i)];
            if (s.status != MKO_LM_RUNNING) continue;
            if (s.iterations >= o_.max_iterations) {
                s.status = MKO_LM_MAX_ITERATIONS;
                continue;
            }
            list.push_back(i);
        }
        if (list.empty()) return false;
        ++cycles_;
        parallel_for(int(list.size()), o_.threads, [&](int j) {
            Run& s = state_[size_t(list[size_t(j)])];
            solve_step(s, bounds_, P_, s.mu);
        });
        std::vector<int> cycle_runs = list, eval;
        for (int i : cycle_runs) {
            Run& s = state_[size_t(i)];
            if (s.solved && s.rejected && s.step_norm <= o_.xtol) s.status = MKO_LM_CONVERGED_STEP;
            else if (s.solved) eval.push_back(i);
        }
        evaluate(eval, true);
        size_t e = 0;
        for (int i : cycle_runs) {
            Run& s = state_[size_t(i)];
            if (s.status == MKO_LM_CONVERGED_STEP) continue;
            ++s.iterations;
            bool accept = false;
            if (s.solved) {
                const double f_new = f_[e];
                const int64_t failed_new = fl_[e], hard_new = hf_[e];
                ++e;
                const double rho = (s.merit - f_new) / s.predicted;
                accept = std::isfinite(f_new) && s.predicted > 0.0 && f_new < s.merit &&
                         !(o_.reject_new_failures && hard_new > s.hard);
                if (accept) {
                    s.x = s.trial;
                    s.merit = f_new;
                    s.failed = failed_new;
                    s.hard = hard_new;
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
                if (s.mu > o_.max_damping) s.status = MKO_LM_DAMPING_LIMIT;
            }
            s.history.push_back(s.merit);
            if (o_.stall_window > 0 && s.status == MKO_LM_RUNNING) {
                const size_t w = size_t(o_.stall_window);
                if (s.history.size() > w) {
                    const double old = s.history[s.history.size() - 1 - w];
                    if (old - s.merit <= o_.stall_tolerance * old) s.status = MKO_LM_CONVERGED_STALL;
                }
            }
        }
        return true;
    }

    void finish_round() {
        if (device_) dev_download_state();
        if (ids_.empty()) {
            final_ = state_;
            return;
        }
        for (size_t j = 0; j < ids_.size(); ++j) {
            Run& r = final_[size_t(ids_[j])];
            const Run& q = state_[j];
            r.x = q.x;
            r.merit = q.merit;
            r.failed = q.failed;
            r.hard = q.hard;
            r.status = q.status;
            r.iterations += q.iterations;
            r.accepted += q.accepted;
            r.jacobians += q.jacobians;
            r.mu = q.mu;
            r.gradient_norm = q.gradient_norm;
            r.model_decrease = q.model_decrease;
            weight_[size_t(ids_[j])] = settings_.fail_weight;
            ++rounds_[size_t(ids_[j])];
        }
    }

    // outer penalty rounds (optimize_lm_rounds): restart runs with failed rays with a larger
    // fail weight while it stays below max_fail_weight
    bool next_round() {
        if (settings_.failure != MKO_FAILURE_SOFT || !(o_.fail_weight_factor > 1.0)) return false;
        double w = settings_.fail_weight;
        if (!(w > 0.0)) return false;
        w *= o_.fail_weight_factor;
        if (!(w <= o_.max_fail_weight)) return false;
        std::vector<int> again;
        std::vector<double> xs;
        for (int i = 0; i < runs_; ++i) {
            const Run& r = final_[size_t(i)];
            if (r.failed > 0 && r.status != MKO_LM_NUMERICAL_ERROR) {
                again.push_back(i);
                xs.insert(xs.end(), r.x.begin(), r.x.end());
            }
        }
        if (again.empty()) return false;
        settings_.fail_weight = w;
        start_round(again, xs);
        return true;
    }

    // ---- LM control on the device (lm_device.hip): per-run state in device memory -----------

    void dev_setup() {
        validate_values(base_, 0, base_.variants);
        check_glass(base_, vars_);
        eng_.set_variables(vars_, id_);
        dev_tmpl_.upload(base_, 0, base_.variants, nullptr);
        double* bnd = bounds_d_.as<double>(3 * P_);
        hip_check(hipMemcpy(bnd, vars_.lower, P_ * 8, hipMemcpyHostToDevice), "bounds");
        hip_check(hipMemcpy(bnd + P_, vars_.upper, P_ * 8, hipMemcpyHostToDevice), "bounds");
        hip_check(hipMemcpy(bnd + 2 * P_, vars_.typical, P_ * 8, hipMemcpyHostToDevice), "bounds");
        dopt_.max_iterations = o_.max_iterations;
        dopt_.stall_window = o_.stall_window;
        dopt_.reject_new_failures = o_.reject_new_failures;
        dopt_.ftol = o_.ftol;
        dopt_.xtol = o_.xtol;
        dopt_.gtol = o_.gtol;
        dopt_.stall_tolerance = o_.stall_tolerance;
        dopt_.initial_damping = o_.initial_damping;
        dopt_.max_damping = o_.max_damping;
        dopt_.diag_floor = o_.diag_floor;
        soft_split_ = Engine::soft_split_applies(settings_, vars_, o_.jacobian_method, precision_);
    }

    // The LM's MERIT traces keep the final ray states for the fused mixed-precision adjoint at
    // the same points (FP64 trace, fused normal equations: hard failure, or soft failure for the
    // runs without failed rays), if they fit in half of the engine's memory budget.
    void dev_allocate_states(int n) {
        use_states_ = false;
        if (o_.jacobian_method != MKO_JACOBIAN_ADJOINT || precision_ != MKO_PRECISION_MIXED ||
            !(Engine::fused_ok(settings_, vars_, true) || soft_split_) ||
            settings_.precision != MKO_PRECISION_FP64)
            return;
        const int64_t rays = rays_of(base_, settings_);
        const size_t bytes = size_t(n) * size_t(rays) * 7 * sizeof(double);
        if (bytes > eng_.budget() / 2) return;
        states_.data = states_mem_.as<double>(size_t(n) * size_t(rays) * 7);
        states_.slot = nullptr;
        states_.slot_offset = 0;
        states_.slots = n;
        states_.rays = rays;
        use_states_ = true;
    }

    // State arrays of n runs in one allocation.
    void dev_allocate(int n) {
        const size_t N = size_t(n), P = P_;
        const int H = o_.stall_window > 0 ? o_.stall_window + 1 : 1;
        const size_t doubles = N * (4 * P + P * P + 7 + size_t(H));
        const size_t bytes = doubles * 8 + N * 2 * 8 + N * 5 * 4 + N * 4 + 1024;
        char* p = static_cast<char*>(state_mem_.reserve(bytes));
        auto take = [&](size_t b) {
            char* q = p;
            p += (b + 15) & ~size_t(15);
            return q;
        };
        mko_lmdev::State& S = ds_;
        S.n = n;
        S.P = int(P);
        S.H = H;
        S.x = reinterpret_cast<double*>(take(N * P * 8));
        S.trial = reinterpret_cast<double*>(take(N * P * 8));
        S.g = reinterpret_cast<double*>(take(N * P * 8));
        S.d = reinterpret_cast<double*>(take(N * P * 8));
        S.a = reinterpret_cast<double*>(take(N * P * P * 8));
        S.merit = reinterpret_cast<double*>(take(N * 8));
        S.mu = reinterpret_cast<double*>(take(N * 8));
        S.nu = reinterpret_cast<double*>(take(N * 8));
        S.step_norm = reinterpret_cast<double*>(take(N * 8));
        S.predicted = reinterpret_cast<double*>(take(N * 8));
        S.gradient_norm = reinterpret_cast<double*>(take(N * 8));
        S.model_decrease = reinterpret_cast<double*>(take(N * 8));
        S.failed = reinterpret_cast<int64_t*>(take(N * 8));
        S.hard = reinterpret_cast<int64_t*>(take(N * 8));
        S.status = reinterpret_cast<int32_t*>(take(N * 4));
        S.iterations = reinterpret_cast<int32_t*>(take(N * 4));
        S.accepted = reinterpret_cast<int32_t*>(take(N * 4));
        S.jacobians = reinterpret_cast<int32_t*>(take(N * 4));
        S.history_len = reinterpret_cast<int32_t*>(take(N * 4));
        S.need_jacobian = reinterpret_cast<uint8_t*>(take(N));
        S.rejected = reinterpret_cast<uint8_t*>(take(N));
        S.solved = reinterpret_cast<uint8_t*>(take(N));
        S.evaluate = reinterpret_cast<uint8_t*>(take(N));
        S.history = reinterpret_cast<double*>(take(N * size_t(H) * 8));
        const double* bnd = static_cast<const double*>(bounds_d_.as<double>(3 * P));
        S.lower = bnd;
        S.upper = bnd + P;
        S.typical = bnd + 2 * P;
        lists_ = lists_mem_.as<int32_t>(6 * N + 8);
        // [0] Jacobian, [1] cycle, [2] evaluate, [3] / [4] the Jacobian list's runs without /
        // with failed rays (soft split), [5] runs of the next Jacobian list with failed rays
        counts_ = lists_ + 6 * N;
    }

    // template index per run of the round (global run ids) when runs have their own template
    const int32_t* dev_round_index() const { return base_.variants == 1 ? nullptr : round_index_; }

    int dev_count(int which) {
        int32_t c = 0;
        hip_check(hipMemcpy(&c, counts_ + which, 4, hipMemcpyDeviceToHost), "LM count");
        return c;
    }

    void dev_start_round(const std::vector<int>& ids, const std::vector<double>& xs) {
        graph_reset(); // the round's state arrays may move
        eng_.set_variables(vars_, id_); // another session may have used the engine
        ids_ = ids;
        const int n = ids.empty() ? runs_ : int(ids.size());
        state_.assign(size_t(n), Run());
        dev_allocate(n);
        dev_allocate_states(n);
        hip_check(hipMemcpy(ds_.x, xs.data(), size_t(n) * P_ * 8, hipMemcpyHostToDevice), "x");
        std::vector<int32_t> index(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) index[size_t(i)] = ids.empty() ? i : ids[size_t(i)];
        round_index_ = round_mem_.as<int32_t>(size_t(n));
        hip_check(hipMemcpy(round_index_, index.data(), size_t(n) * 4, hipMemcpyHostToDevice),
                  "round index");
        eng_.timer.run(kLmControl, [&] { mko_lmdev::launch_init(ds_, dopt_, nullptr); });
        double* fm = fm_.as<double>(size_t(n));
        int64_t* ff = ff_.as<int64_t>(size_t(n));
        int64_t* fh = fh_.as<int64_t>(size_t(n));
        eng_.merit_dev(dev_tmpl_.view, dev_round_index(), n, ds_.x, settings_, int(P_), fm, ff, fh,
                       use_states_ ? &states_ : nullptr);
        eng_.timer.run(kLmControl, [&] { mko_lmdev::launch_start(ds_, fm, ff, fh, nullptr); });
        hip_check(hipGetLastError(), "LM start");
        count_failing();
    }

    // soft split: how many runs of the next cycle's Jacobian list have failed rays (the FP64
    // share of its normal equations), read at the end of each cycle to choose its form
    void count_failing() {
        if (!soft_split_) return;
        eng_.timer.run(kLmControl, [&] {
            mko_lmdev::launch_count_failing(ds_, o_.max_iterations, counts_ + 5, nullptr);
        });
        failing_next_ = dev_count(5);
    }

    bool dev_round_cycle() {
        eng_.set_variables(vars_, id_); // another session may have used the engine
        const int n = ds_.n;
        int32_t* list_j = lists_;
        int32_t* list_c = lists_ + n;
        int32_t* list_e = lists_ + 2 * n;
        int32_t* pos = lists_ + 3 * n;
        const int32_t* rix = dev_round_index();
        double* xs = xs_d_.as<double>(size_t(n) * P_);
        int32_t* tidx = rix ? tidx_d_.as<int32_t>(size_t(n)) : nullptr;
        eng_.timer.run(kLmControl, [&] {
            mko_lmdev::launch_select(ds_, mko_lmdev::kListJacobian, o_.max_iterations, nullptr,
                                     nullptr, list_j, counts_ + 0, nullptr, nullptr);
        });
        const int cj = dev_count(0);
        if (cj > 0) {
            double* ac = ac_.as<double>(size_t(cj) * P_ * P_);
            double* gc = gc_.as<double>(size_t(cj) * P_);
            mko_internal::RayStates st = states_;
            st.slot = list_j;
            if (soft_split_) {
                // soft mode, mixed adjoint (Engine::soft_split): the runs whose point has no
                // failed ray take hard mode's normal equations (fused kernel), the others FP64;
                // the device splits the list, each part is gathered, solved and prepared alone
                int32_t* list_cl = lists_ + 4 * n;
                int32_t* list_ft = lists_ + 5 * n;
                eng_.timer.run(kLmControl, [&] {
                    mko_lmdev::launch_split(ds_, list_j, counts_ + 0, list_cl, counts_ + 3,
                                            list_ft, counts_ + 4, nullptr);
                });
                int32_t split[2] = {0, 0};
                hip_check(hipMemcpy(split, counts_ + 3, sizeof(split), hipMemcpyDeviceToHost),
                          "LM split counts");
                mko_residual_settings hs = settings_;
                hs.failure = MKO_FAILURE_HARD;
                for (int part = 0; part < 2; ++part) {
                    const int c = split[part];
                    if (c == 0) continue;
                    const int32_t* list = part == 0 ? list_cl : list_ft;
                    eng_.timer.run(kLmControl, [&] {
                        mko_lmdev::launch_gather(ds_, list, c, false, rix, xs, tidx, nullptr);
                    });
                    st.slot = list;
                    if (part == 0)
                        eng_.jacobian_dev(dev_tmpl_.view, tidx, c, xs, hs, int(P_), ds_.lower,
                                          ds_.upper, ds_.typical, o_.jacobian_method,
                                          MKO_PRECISION_MIXED, ac, gc,
                                          use_states_ ? &st : nullptr, &settings_);
                    else
                        eng_.jacobian_dev(dev_tmpl_.view, tidx, c, xs, settings_, int(P_),
                                          ds_.lower, ds_.upper, ds_.typical, o_.jacobian_method,
                                          MKO_PRECISION_FP64, ac, gc, nullptr);
                    eng_.timer.run(kLmControl, [&] {
                        mko_lmdev::launch_prepare(ds_, dopt_, list, c, ac, gc, nullptr);
                    });
                }
            } else {
                eng_.timer.run(kLmControl, [&] {
                    mko_lmdev::launch_gather(ds_, list_j, cj, false, rix, xs, tidx, nullptr);
                });
                eng_.jacobian_dev(dev_tmpl_.view, tidx, cj, xs, settings_, int(P_), ds_.lower,
                                  ds_.upper, ds_.typical, o_.jacobian_method, precision_, ac, gc,
                                  use_states_ ? &st : nullptr);
                eng_.timer.run(kLmControl, [&] {
                    mko_lmdev::launch_prepare(ds_, dopt_, list_j, cj, ac, gc, nullptr);
                });
            }
        }
        eng_.timer.run(kLmControl, [&] {
            mko_lmdev::launch_select(ds_, mko_lmdev::kListCycle, o_.max_iterations, nullptr,
                                     nullptr, list_c, counts_ + 1, nullptr, nullptr);
        });
        const int cc = dev_count(1);
        if (cc == 0) return false;
        ++cycles_;
        eager_pending_ = true; // the cycle's last launches are still running on the null stream
        eng_.timer.run(kLmControl, [&] {
            mko_lmdev::launch_step(ds_, dopt_, list_c, cc, nullptr);
            mko_lmdev::launch_select(ds_, mko_lmdev::kListEvaluate, o_.max_iterations, list_c,
                                     counts_ + 1, list_e, counts_ + 2, pos, nullptr);
        });
        const int ce = dev_count(2);
        double* fm = fm_.as<double>(size_t(std::max(ce, 1)));
        int64_t* ff = ff_.as<int64_t>(size_t(std::max(ce, 1)));
        int64_t* fh = fh_.as<int64_t>(size_t(std::max(ce, 1)));
        if (ce > 0) {
            eng_.timer.run(kLmControl, [&] {
                mko_lmdev::launch_gather(ds_, list_e, ce, true, rix, xs, tidx, nullptr);
            });
            mko_internal::RayStates st = states_;
            st.slot = list_e;
            eng_.merit_dev(dev_tmpl_.view, tidx, ce, xs, settings_, int(P_), fm, ff, fh,
                           use_states_ ? &st : nullptr);
        }
        eng_.timer.run(kLmControl, [&] {
            mko_lmdev::launch_update(ds_, dopt_, list_c, cc, pos, fm, ff, fh, nullptr);
        });
        hip_check(hipGetLastError(), "LM cycle");
        count_failing();
        return true;
    }

    // ---- one LM cycle as a captured hipGraph ------------------------------------------------
    struct GraphPart {
        KernelFamily fam;
        int which; // list whose count > 0 makes the part a timed launch (-1: always)
        std::function<void()> launch;
    };
    struct GraphTimed {
        KernelFamily fam;
        int which;
        hipGraphExec_t exec;
        hipEvent_t a, b;
    };
    // Device control with the fused mixed-precision adjoint on stored ray states and no
    // constraint rows (the batched LM of the headline workload): the cycle of dev_round_cycle
    // without its three host reads of list lengths. Every launch is sized for the round's n runs
    // and its kernels read the list length from device memory (`live` / `dcount`), so one
    // capture per round serves every cycle; the host waits once per cycle for the three counts.
    // The graph owns its work buffers (nothing the engine reallocates between cycles) and is
    // rebuilt when a round starts. Same kernels, same inputs per run: the results are those of
    // the eager cycle bit for bit.
    // Soft failure mode (soft split): the cycle splits its Jacobian list on the device, the runs
    // without failed rays take the fused kernel, the others the FP64 stored-J path on buffers
    // for up to gdirty_cap_ runs; a cycle whose list holds more such runs (known from the
    // previous cycle's count) runs eagerly.
    bool graph_eligible() const {
        return graph_wanted_ && !graph_failed_ && use_states_ && settings_.extra == nullptr &&
               precision_ == MKO_PRECISION_MIXED && o_.jacobian_method == MKO_JACOBIAN_ADJOINT &&
               (Engine::fused_ok(settings_, vars_, true) || soft_split_);
    }
    bool graph_fits() const { return !soft_split_ || failing_next_ <= gdirty_cap_; }

    void graph_reset() {
        if (gexec_) (void)hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
        for (GraphTimed& t : gparts_) {
            if (t.exec) (void)hipGraphExecDestroy(t.exec);
            if (t.a) (void)hipEventDestroy(t.a);
            if (t.b) (void)hipEventDestroy(t.b);
        }
        gparts_.clear();
    }

    void graph_release() {
        graph_reset();
        if (gs_) (void)hipStreamDestroy(gs_);
        if (counts_host_) (void)hipHostFree(counts_host_);
        gs_ = nullptr;
        counts_host_ = nullptr;
    }

    // Captures `parts` of the cycle, in order, into one executable graph on gs_ (NULL on failure:
    // nothing captured, nothing run).
    hipGraphExec_t graph_capture_parts(const std::vector<const GraphPart*>& parts) {
        hipGraph_t g = nullptr;
        bool ok = true;
        try {
            hip_check(hipStreamBeginCapture(gs_, hipStreamCaptureModeRelaxed), "begin capture");
            for (const GraphPart* p : parts) p->launch();
            hip_check(hipGetLastError(), "graph cycle launch");
        } catch (const std::exception&) {
            ok = false;
        }
        const hipError_t end = hipStreamEndCapture(gs_, &g);
        ok = ok && end == hipSuccess && g != nullptr;
        hipGraphExec_t exec = nullptr;
        if (ok) ok = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
        if (!ok) {
            if (exec) (void)hipGraphExecDestroy(exec);
            (void)hipGetLastError(); // clear the capture's error state
            return nullptr;
        }
        return exec;
    }

    // The cycle as one graph (gexec_) and, for the engine's kernel timing, as one graph per part
    // (HIP events recorded between their launches on gs_); false on failure.
    bool graph_capture() {
        const int n = ds_.n;
        const mko_batch& tmpl = dev_tmpl_.view;
        if (!gs_) hip_check(hipStreamCreateWithFlags(&gs_, hipStreamNonBlocking), "graph stream");
        if (!counts_host_)
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&counts_host_), 32, hipHostMallocDefault),
                      "pinned counts");
        // buffers of n runs, allocated before the capture (within half of the memory budget;
        // else the eager cycle, which tranches)
        const size_t P = P_, N = size_t(n);
        const size_t T = size_t(tmpl.fields) * tmpl.wavelengths;
        const size_t wbytes = mko_adjoint_normal_workspace(n, tmpl.fields, tmpl.wavelengths, int(P));
        const size_t need =
            wbytes + N * (Strides(tmpl).doubles() * 8 + P * P * 8 + P * 16 + T * 16 + 64);
        const size_t half = eng_.budget() / 2;
        if (need > half) return false;
        // soft split: FP64 stored-J buffers for as many runs with failed rays as fit
        const int64_t R = trace_rows_of(tmpl, settings_);
        const size_t per_dirty =
            Strides(tmpl).doubles() * 8 + (P + 1) * size_t(R) * 8 + P * P * 8 + P * 16 + 64;
        gdirty_cap_ = soft_split_ ? int(std::min<size_t>(N, (half - need) / per_dirty)) : 0;
        const size_t D = size_t(gdirty_cap_);
        gwork_.allocate(tmpl, n);
        void* ws = gws_.reserve(wbytes);
        double* ac = gac_.as<double>(N * P * P);
        double* gc = ggc_.as<double>(N * P);
        double* xs = gxs_.as<double>(N * P);
        const int32_t* rix = dev_round_index();
        int32_t* tidx = rix ? gtidx_.as<int32_t>(N) : nullptr;
        double* fm = gfm_.as<double>(N);
        int64_t* ff = gff_.as<int64_t>(N);
        int64_t* fh = gfh_.as<int64_t>(N);
        double* part = gpart_.as<double>(N * T);
        int* pc = gpc_.as<int>(N * T * 2);
        // soft split buffers (D runs with failed rays)
        if (D > 0) gwork2_.allocate(tmpl, int(D));
        double* xs2 = D ? gxs2_.as<double>(D * P) : nullptr;
        int32_t* tidx2 = rix && D ? gtidx2_.as<int32_t>(D) : nullptr;
        double* J2 = D ? gj2_.as<double>(D * P * size_t(R)) : nullptr;
        double* r02 = D ? gr02_.as<double>(D * size_t(R)) : nullptr;
        double* ac2 = D ? gac2_.as<double>(D * P * P) : nullptr;
        double* gc2 = D ? ggc2_.as<double>(D * P) : nullptr;
        int32_t* vt = gvt_.as<int32_t>(P);
        int32_t* vs = gvs_.as<int32_t>(P);
        hip_check(hipMemcpy(vt, vars_.type, P * 4, hipMemcpyHostToDevice), "graph vars");
        hip_check(hipMemcpy(vs, vars_.batch_surface, P * 4, hipMemcpyHostToDevice), "graph vars");
        mko_variables dv = vars_;
        dv.type = vt;
        dv.batch_surface = vs;
        dv.component = nullptr;
        if (vars_.component) { // glass variables: their components, on the graph's own buffer
            int32_t* vc = gvc_.as<int32_t>(P);
            hip_check(hipMemcpy(vc, vars_.component, P * 4, hipMemcpyHostToDevice), "graph vars");
            dv.component = vc;
        }
        // the round's start (null stream) is complete before the first replay on gs_
        hip_check(hipStreamSynchronize(nullptr), "graph start");
        int32_t* list_j = lists_;
        int32_t* list_c = lists_ + n;
        int32_t* list_e = lists_ + 2 * n;
        int32_t* pos = lists_ + 3 * n;
        const int32_t *cj = counts_ + 0, *cc = counts_ + 1, *ce = counts_ + 2;
        // soft split: the fused part works on the runs without failed rays (list_cl, count [3])
        // under hard mode's settings, the FP64 part on the others (list_ft, count [4])
        int32_t* list_cl = lists_ + 4 * n;
        int32_t* list_ft = lists_ + 5 * n;
        const int32_t* cf = soft_split_ ? counts_ + 3 : cj;
        const int32_t* cd = counts_ + 4;
        int32_t* list_f = soft_split_ ? list_cl : list_j;
        mko_residual_settings hs = settings_;
        if (soft_split_) hs.failure = MKO_FAILURE_HARD;
        mko_internal::RayStates st_j = states_, st_e = states_;
        st_j.slot = list_f;
        st_e.slot = list_e;
        auto point_batch = [this, &tmpl, &dv](DeviceBatch& w, const int32_t* ti, int count,
                                              const double* x, const int32_t* live) {
            dev_check(mko_internal::point_batch(
                &tmpl, ti, 0, count, 1, &dv, x, nullptr, nullptr, w.mut(w.view.curvature),
                w.mut(w.view.thickness), w.mut(w.view.index), w.mut(w.view.pupil_radius),
                w.mut(w.view.field_angle), w.mut(w.view.field_target),
                w.mut(w.view.reference_curvature), w.mut(w.view.paraxial_index),
                w.mut(w.view.buchdahl), gs_, live));
        };
        const int Di = int(D);
        // the parts of the cycle in order; `which`: the count whose value > 0 makes the part a
        // timed launch, as in the eager cycle (-1: always)
        std::vector<GraphPart> parts = {
            {kLmControl, -1, [&] {
                 mko_lmdev::launch_select(ds_, mko_lmdev::kListJacobian, o_.max_iterations,
                                          nullptr, nullptr, list_j, counts_ + 0, nullptr, gs_);
                 if (soft_split_)
                     mko_lmdev::launch_split(ds_, list_j, cj, list_cl, counts_ + 3, list_ft,
                                             counts_ + 4, gs_);
                 mko_lmdev::launch_gather(ds_, list_f, n, false, rix, xs, tidx, gs_, cf);
             }},
            {kPointBatch, soft_split_ ? 3 : 0, [&] { point_batch(gwork_, tidx, n, xs, cf); }},
            {kAdjoint, soft_split_ ? 3 : 0, [&] {
                 dev_check(mko_internal::adjoint_normal_states(&gwork_.view, &hs, &vars_, ac, gc,
                                                               ws, wbytes, st_j, gs_, cf));
             }},
            {kLmControl, -1, [&] {
                 mko_lmdev::launch_prepare(ds_, dopt_, list_f, n, ac, gc, gs_, cf);
                 mko_lmdev::launch_gather(ds_, list_ft, Di, false, rix, xs2, tidx2, gs_, cd);
             }},
            {kPointBatch, 4, [&] { point_batch(gwork2_, tidx2, Di, xs2, cd); }},
            {kAdjoint, 4, [&] {
                 dev_check(mko_internal::adjoint_jacobian(&gwork2_.view, &settings_, &vars_, J2,
                                                          r02, nullptr, R, gs_, cd));
             }},
            {kGram, 4, [&] {
                 dev_check(mko_internal::gram_chain(J2, r02, Di, int(P), R, ac2, gc2, gs_, cd));
             }},
            {kLmControl, -1, [&] {
                 mko_lmdev::launch_prepare(ds_, dopt_, list_ft, Di, ac2, gc2, gs_, cd);
                 mko_lmdev::launch_select(ds_, mko_lmdev::kListCycle, o_.max_iterations, nullptr,
                                          nullptr, list_c, counts_ + 1, nullptr, gs_);
                 mko_lmdev::launch_step(ds_, dopt_, list_c, n, gs_, cc);
                 mko_lmdev::launch_select(ds_, mko_lmdev::kListEvaluate, o_.max_iterations,
                                          list_c, counts_ + 1, list_e, counts_ + 2, pos, gs_);
                 mko_lmdev::launch_gather(ds_, list_e, n, true, rix, xs, tidx, gs_, ce);
             }},
            {kPointBatch, 2, [&] { point_batch(gwork_, tidx, n, xs, ce); }},
            {kTrace, 2, [&] {
                 dev_check(mko_internal::merit_states(&gwork_.view, &settings_, fm, ff, fh, part,
                                                      pc, st_e, gs_, ce));
             }},
            {kLmControl, -1, [&] {
                 mko_lmdev::launch_update(ds_, dopt_, list_c, n, pos, fm, ff, fh, gs_, cc);
                 if (soft_split_)
                     mko_lmdev::launch_count_failing(ds_, o_.max_iterations, counts_ + 5, gs_);
             }},
        };
        if (Di == 0) // no FP64 part: its point batch, adjoint and Gram are left out
            parts.erase(parts.begin() + 4, parts.begin() + 7);
        std::vector<const GraphPart*> all;
        for (const GraphPart& p : parts) all.push_back(&p);
        gexec_ = graph_capture_parts(all);
        if (!gexec_) return false;
        for (const GraphPart& p : parts) {
            GraphTimed t{p.fam, p.which, graph_capture_parts({&p}), nullptr, nullptr};
            if (!t.exec) {
                graph_reset();
                return false;
            }
            hip_check(hipEventCreate(&t.a), "graph event");
            hip_check(hipEventCreate(&t.b), "graph event");
            gparts_.push_back(t);
        }
        return true;
    }

    // The captured cycle is ready (captured now if needed); false: eager cycles from here on.
    bool graph_ready() {
        if (gexec_) return true;
        if (graph_capture()) return true;
        graph_reset();
        graph_failed_ = true; // eager cycles for the rest of this LM
        return false;
    }

    // Launches the cycle (and the read of its counts) on gs_ without waiting.
    void graph_launch() {
        if (eager_pending_) { // an eager cycle's launches on the null stream come first
            hip_check(hipStreamSynchronize(nullptr), "LM cycle");
            eager_pending_ = false;
        }
        if (eng_.timer.enabled) {
            for (const GraphTimed& t : gparts_) {
                hip_check(hipEventRecord(t.a, gs_), "hipEventRecord");
                hip_check(hipGraphLaunch(t.exec, gs_), "graph launch");
                hip_check(hipEventRecord(t.b, gs_), "hipEventRecord");
            }
            timed_launch_ = true;
        } else {
            hip_check(hipGraphLaunch(gexec_, gs_), "graph launch");
            timed_launch_ = false;
        }
        hip_check(hipMemcpyAsync(counts_host_, counts_, 6 * sizeof(int32_t), hipMemcpyDeviceToHost,
                                 gs_),
                  "LM counts");
    }

    // Waits for the launched cycle; false when the round has no active run (as dev_round_cycle).
    bool graph_finish() {
        hip_check(hipStreamSynchronize(gs_), "LM cycle");
        ++graph_cycles_;
        // per family the launches that had work, as the eager cycle times them
        if (timed_launch_) {
            for (const GraphTimed& t : gparts_) {
                if (t.which >= 0 && counts_host_[t.which] == 0) continue;
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, t.a, t.b), "graph timing");
                eng_.timer.add(t.fam, ms);
            }
        }
        if (soft_split_) {
            if (counts_host_[4] > gdirty_cap_)
                throw std::runtime_error("LM cycle: more runs with failed rays than its buffers");
            failing_next_ = counts_host_[5];
        }
        if (counts_host_[1] == 0) return false;
        ++cycles_;
        return true;
    }

    bool graph_round_cycle() {
        if (!graph_ready() || !graph_fits()) return dev_round_cycle();
        graph_launch();
        return graph_finish();
    }

    // Device state -> state_ (the fields finish_round and results read).
    void dev_download_state() {
        const size_t n = size_t(ds_.n);
        if (n == 0) return;
        std::vector<double> x(n * P_), merit(n), mu(n), gn(n), md(n);
        std::vector<int64_t> failed(n), hard(n);
        std::vector<int32_t> status(n), it(n), acc(n), jac(n);
        auto get = [](void* h, const void* d, size_t b) {
            hip_check(hipMemcpy(h, d, b, hipMemcpyDeviceToHost), "LM state");
        };
        get(x.data(), ds_.x, n * P_ * 8);
        get(merit.data(), ds_.merit, n * 8);
        get(mu.data(), ds_.mu, n * 8);
        get(gn.data(), ds_.gradient_norm, n * 8);
        get(md.data(), ds_.model_decrease, n * 8);
        get(failed.data(), ds_.failed, n * 8);
        get(hard.data(), ds_.hard, n * 8);
        get(status.data(), ds_.status, n * 4);
        get(it.data(), ds_.iterations, n * 4);
        get(acc.data(), ds_.accepted, n * 4);
        get(jac.data(), ds_.jacobians, n * 4);
        for (size_t i = 0; i < n; ++i) {
            Run& r = state_[i];
            r.x.assign(x.begin() + std::ptrdiff_t(i * P_), x.begin() + std::ptrdiff_t((i + 1) * P_));
            r.merit = merit[i];
            r.mu = mu[i];
            r.gradient_norm = gn[i];
            r.model_decrease = md[i];
            r.failed = failed[i];
            r.hard = hard[i];
            r.status = status[i];
            r.iterations = it[i];
            r.accepted = acc[i];
            r.jacobians = jac[i];
        }
    }

    Engine& eng_;
    // Session id for the engine's variable upload: unique per process, unlike the address of a
    // session, which a later session may reuse after this one is closed.
    const uint64_t id_ = next_id();
    static uint64_t next_id() {
        static std::atomic<uint64_t> n{0};
        return ++n;
    }
    mko_batch base_;
    mko_residual_settings settings_;
    std::vector<mko_constraint> items_own_;
    std::vector<double> hull_own_;
    mko_extra_rows extra_own_{};
    mko_variables vars_;
    int runs_;
    mko_lm_options o_;
    int precision_;
    mko_internal::Tuning tuning_;
    size_t P_ = 0;
    Bounds bounds_{};
    std::vector<double> x0_;
    std::vector<int> ids_;
    std::vector<Run> state_, final_;
    std::vector<double> weight_;
    std::vector<int> rounds_;
    int cycles_ = 0;
    bool done_ = false;
    bool device_ = false;
    mko_lmdev::State ds_{};
    mko_lmdev::Options dopt_{};
    DeviceBatch dev_tmpl_;
    Buffer state_mem_, lists_mem_, round_mem_, bounds_d_, xs_d_, tidx_d_, ac_, gc_, fm_, ff_, fh_;
    bool use_states_ = false;
    bool soft_split_ = false; // soft mode, mixed adjoint: failure-free runs on the fused kernel
    mko_internal::RayStates states_{};
    Buffer states_mem_;
    int32_t* lists_ = nullptr;
    int32_t* counts_ = nullptr;
    int32_t* round_index_ = nullptr;
    std::vector<double> xs_, a_, g_, f_;
    std::vector<int64_t> fl_, hf_;
    GatheredBatch gathered_batch_;
    const mko_batch* gathered_ = nullptr;
    // captured cycle
    bool graph_wanted_ = false;
    bool graph_failed_ = false;
    bool timed_launch_ = false;
    int graph_cycles_ = 0;
    int pending_ = 0;          // cycle_begin: 1 launched, 2 ran synchronously
    bool sync_result_ = false;
    hipStream_t gs_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    int32_t* counts_host_ = nullptr;
    std::vector<GraphTimed> gparts_;
    DeviceBatch gwork_, gwork2_;
    Buffer gws_, gac_, ggc_, gxs_, gtidx_, gfm_, gff_, gfh_, gpart_, gpc_, gvt_, gvs_, gvc_;
    Buffer gxs2_, gtidx2_, gj2_, gr02_, gac2_, ggc2_; // soft split: the FP64 part
    int gdirty_cap_ = 0;     // runs with failed rays the captured cycle's FP64 part holds
    int failing_next_ = 0;   // runs with failed rays in the next cycle's Jacobian list
    bool eager_pending_ = false;
};

} // namespace mko_host

using namespace mko_host;

struct mko_engine {
    Engine engine;
    mko_engine(int dev, size_t lim) : engine(dev, lim) {}
};
struct mko_lm {
    Lm lm;
    // the LM keeps pointers into the caller's batch and variable arrays; they must stay valid
};

#define MKO_HOST_TRY(...)                                                                          \
    try {                                                                                          \
        __VA_ARGS__;                                                                               \
    } catch (const std::invalid_argument& e) {                                                     \
        mko_detail::last_error = e.what();                                                         \
        return MKO_ERR_INVALID;                                                                    \
    } catch (const std::exception& e) {                                                            \
        mko_detail::last_error = e.what();                                                         \
        return MKO_ERR_DEVICE;                                                                     \
    }

extern "C" {

mko_engine* mko_engine_create(int device, size_t memory_limit_bytes) {
    try {
        return new mko_engine(device, memory_limit_bytes);
    } catch (const std::exception& e) {
        mko_detail::last_error = e.what();
        return nullptr;
    }
}

void mko_engine_destroy(mko_engine* e) { delete e; }

int mko_engine_kernel_timing(mko_engine* e, int enable, double* ms, int64_t* launches) {
    MKO_HOST_TRY({
        KernelTimer& t = e->engine.timer;
        t.collect();
        if (ms)
            for (int i = 0; i < kFamilies; ++i) ms[i] = t.ms[i];
        if (launches)
            for (int i = 0; i < kFamilies; ++i) launches[i] = t.launches[i];
        t.reset();
        t.enabled = enable != 0;
        return 0;
    })
}

int mko_engine_set_tuning(mko_engine* e, const mko_tuning* t) {
    MKO_HOST_TRY({
        if (!e || !t) throw std::invalid_argument("engine and tuning needed");
        if (t->pack_rays != 0 && t->pack_rays != 32 && t->pack_rays != 64 &&
            t->pack_rays != 128 && t->pack_rays != 256)
            throw std::invalid_argument("pack_rays must be 0, 32, 64, 128 or 256");
        if (t->fused_waves != 0 && t->fused_waves != 1 && t->fused_waves != 2 &&
            t->fused_waves != 4)
            throw std::invalid_argument("fused_waves must be 0, 1, 2 or 4");
        if (t->full_row_traces != 0 && t->full_row_traces != 1)
            throw std::invalid_argument("full_row_traces must be 0 or 1");
        e->engine.set_tuning(
            mko_internal::Tuning{t->pack_rays, t->fused_waves, t->full_row_traces});
        return 0;
    })
}

int mko_engine_row_trace_stats(const mko_engine* e, int64_t* candidate_chunks,
                               int64_t* full_chunks) {
    MKO_HOST_TRY({
        if (!e) throw std::invalid_argument("engine needed");
        const auto st = e->engine.candidate_stats();
        if (candidate_chunks) *candidate_chunks = st.first;
        if (full_chunks) *full_chunks = st.second;
        return 0;
    })
}

int mko_engine_get_tuning(const mko_engine* e, mko_tuning* t) {
    MKO_HOST_TRY({
        if (!e || !t) throw std::invalid_argument("engine and tuning needed");
        t->pack_rays = e->engine.tuning().pack_rays;
        t->fused_waves = e->engine.tuning().fused_waves;
        t->full_row_traces = e->engine.tuning().full_row_traces;
        return 0;
    })
}

int mko_engine_autotune(mko_engine* e, const mko_batch* base, const mko_residual_settings* s,
                        const mko_variables* v, int runs, const double* x, mko_tuning* chosen) {
    MKO_HOST_TRY({
        if (!e || !base || !s || !v || !x) throw std::invalid_argument("autotune arguments missing");
        validate_values(*base, 0, base->variants);
        const mko_internal::Tuning t = e->engine.autotune(*base, *s, *v, runs, x);
        if (chosen) {
            chosen->pack_rays = t.pack_rays;
            chosen->fused_waves = t.fused_waves;
        }
        return 0;
    })
}

int mko_engine_residuals(mko_engine* e, const mko_batch* b, const mko_residual_settings* s,
                         double* res, double* opd, uint8_t* st) {
    MKO_HOST_TRY(mko_internal::TuningScope ts(e->engine.tuning());
                 e->engine.residuals(*b, *s, res, opd, st); return 0;)
}

int mko_engine_prefilter(mko_engine* e, const mko_batch* b, const mko_residual_settings* s,
                         const mko_prefilter_options* o, const double* extra,
                         mko_prefilter_result* r) {
    MKO_HOST_TRY(mko_internal::TuningScope ts(e->engine.tuning());
                 e->engine.prefilter(*b, *s, *o, extra, *r); return 0;)
}

int mko_prefilter_decide(const mko_batch* b, const mko_residual_settings* s,
                         const mko_prefilter_options* o, const double* extra,
                         mko_prefilter_result* r) {
    MKO_HOST_TRY({
        check_batch(*b);
        if (!r->merit || !r->optical_length || !r->failed || !r->uncertain_ok ||
            !r->uncertain_failed || !r->merit_lower || !r->merit_upper || !r->reject)
            throw std::invalid_argument("prefilter result arrays missing");
        prefilter_decide(*b, *s, *o, extra, *r);
        return 0;
    })
}

int mko_engine_surface_extrema(mko_engine* e, const mko_batch* b, const mko_residual_settings* s,
                               double* extrema) {
    MKO_HOST_TRY(mko_internal::TuningScope ts(e->engine.tuning());
                 e->engine.surface_extrema(*b, *s, extrema); return 0;)
}

int mko_engine_merit(mko_engine* e, const mko_batch* b, const mko_residual_settings* s,
                     const double* extra, double* merit, int64_t* failed, int64_t* hard) {
    MKO_HOST_TRY(mko_internal::TuningScope ts(e->engine.tuning());
                 e->engine.merit(*b, *s, extra, merit, failed, hard); return 0;)
}

int mko_engine_jacobian(mko_engine* e, const mko_batch* base, const mko_residual_settings* s,
                        const mko_variables* v, int runs, const double* x0, int method,
                        int precision, double* J, double* r0, double* a, double* g) {
    MKO_HOST_TRY(mko_internal::TuningScope ts(e->engine.tuning());
                 e->engine.jacobian(*base, *s, *v, runs, x0, method, precision, J, r0, a, g);
                 return 0;)
}

mko_lm* mko_lm_create(mko_engine* e, const mko_batch* base, const mko_residual_settings* s,
                      const mko_variables* v, int runs, const double* x0,
                      const mko_lm_options* o, int precision) {
    try {
        mko_internal::TuningScope ts(e->engine.tuning());
        return new mko_lm{Lm(e->engine, *base, *s, *v, runs, x0, *o, precision)};
    } catch (const std::exception& ex) {
        mko_detail::last_error = ex.what();
        return nullptr;
    }
}

int mko_lm_cycle(mko_lm* lm) {
    MKO_HOST_TRY(mko_internal::TuningScope ts(lm->lm.tuning());
                 return lm->lm.cycle() ? 1 : 0;)
}

int mko_lm_results(mko_lm* lm, double* x, mko_lm_run* out, int* cycles) {
    MKO_HOST_TRY(lm->lm.results(x, out, cycles); return 0;)
}

void mko_lm_destroy(mko_lm* lm) { delete lm; }

int mko_lm_graph_cycles(const mko_lm* lm) { return lm ? lm->lm.graph_cycles() : 0; }

int mko_lm_cycle_begin(mko_lm* lm) {
    MKO_HOST_TRY(mko_internal::TuningScope ts(lm->lm.tuning());
                 return lm->lm.cycle_begin();)
}

int mko_lm_cycle_end(mko_lm* lm) {
    MKO_HOST_TRY(mko_internal::TuningScope ts(lm->lm.tuning());
                 return lm->lm.cycle_end() ? 1 : 0;)
}

int mko_engine_normal_equations(mko_engine* e, const mko_batch* base,
                                const mko_residual_settings* s, const mko_variables* v, int count,
                                const int32_t* runs, const double* x, int method, int precision,
                                double* a, double* g) {
    MKO_HOST_TRY({
        if (count < 0) throw std::invalid_argument("negative run count");
        if (count == 0) return 0;
        check_batch(*base);
        mko_internal::TuningScope ts(e->engine.tuning());
        GatheredBatch gb;
        const mko_batch* b = base;
        if (base->variants != 1) {
            if (runs) b = &gb.gather(*base, runs, size_t(count));
            else if (base->variants != count)
                throw std::invalid_argument("base must have one variant, one per run, or runs given");
        }
        e->engine.jacobian(*b, *s, *v, count, x, method, precision, nullptr, nullptr, a, g);
        return 0;
    })
}

int mko_engine_point_merit(mko_engine* e, const mko_batch* base, const mko_residual_settings* s,
                           const mko_variables* v, int count, const int32_t* runs,
                           const double* x, double* merit, int64_t* failed, int64_t* hard) {
    MKO_HOST_TRY({
        if (count < 0) throw std::invalid_argument("negative run count");
        if (count == 0) return 0;
        check_batch(*base);
        if (v->count < 1) throw std::invalid_argument("no variables");
        GatheredBatch gb;
        const mko_batch* b = base;
        if (base->variants != 1) {
            if (runs) b = &gb.gather(*base, runs, size_t(count));
            else if (base->variants != count)
                throw std::invalid_argument("base must have one variant, one per run, or runs given");
        }
        validate_values(*b, 0, b->variants);
        mko_internal::TuningScope ts(e->engine.tuning());
        e->engine.point_merit(*b, *s, *v, count, x, merit, failed, hard);
        return 0;
    })
}

int mko_optimize_lm(mko_engine* e, const mko_batch* base, const mko_residual_settings* s,
                    const mko_variables* v, int runs, const double* x0, const mko_lm_options* o,
                    int precision, double* x_out, mko_lm_run* run_out, int* cycles_out) {
    MKO_HOST_TRY({
        mko_internal::TuningScope ts(e->engine.tuning());
        Lm lm(e->engine, *base, *s, *v, runs, x0, *o, precision);
        while (lm.cycle()) {
        }
        lm.results(x_out, run_out, cycles_out);
        return 0;
    })
}

} // extern "C"

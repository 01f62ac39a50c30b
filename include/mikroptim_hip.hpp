// C++ layer over the C-ABI of libmikroptim_hip.so (header only, no HIP or torch types): owning
// handles for the engine and an LM session, spans for the arrays, and the reference's error
// behaviour -- bad arguments throw std::invalid_argument, HIP runtime and device failures throw
// mko_hip::DeviceError (the reference's mko::cuda::CudaError, include/mikroptim/cuda/
// trace_engine.hpp:136-139). It mirrors the host API of the reference's GPU path:
//   TraceEngine    <- mko::cuda::TraceEngine (include/mikroptim/cuda/trace_engine.hpp:141-176)
//                     run(RESIDUALS / MERIT), surface extrema, prefilter_variants;
//                     DeviceFdJacobian / DeviceAdjointJacobian::compute + download
//                     (include/mikroptim/cuda/fd_jacobian.hpp:58, adjoint_jacobian.hpp:70) and
//                     DeviceNormalEquations (include/mikroptim/cuda/normal_equations.hpp:41-60)
//                     as jacobian(); GpuLmBackend::normal_equations / merit
//                     (include/mikroptim/cuda/levenberg_marquardt.hpp:48-93) as
//                     normal_equations() / point_merit()
//   optimize_lm    <- mko::cuda::optimize_lm_gpu (include/mikroptim/cuda/levenberg_marquardt.hpp:91)
//   LmSession      <- the same LM one cycle at a time (no counterpart in the reference)
// Like the reference's engine, an engine is move-only, owns its device buffers, and must not be
// used from several threads at once; after an exception it stays usable.
// The adapter for the reference's own types (SystemBatchHost, VariableSet, LmBackend) is
// include/mikroptim_hip_reference.hpp.
#pragma once

#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mikroptim_hip.h"

namespace mko_hip {

// HIP runtime or device failure (MKO_ERR_DEVICE).
class DeviceError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// Throws for a negative return code of the C-ABI, with the library's message.
inline void check(int rc) {
    if (rc >= 0) return;
    const char* m = mko_last_error();
    const std::string msg = m ? m : "mikroptim_hip: unknown error";
    if (rc == MKO_ERR_INVALID) throw std::invalid_argument(msg);
    throw DeviceError(msg);
}

namespace detail {
template <class T>
T* data_or_null(std::span<T> s) {
    return s.empty() ? nullptr : s.data();
}
inline void need(bool ok, const char* what) {
    if (!ok) throw std::invalid_argument(what);
}
} // namespace detail

// Normal equations' and Jacobians' arithmetic (MKO_PRECISION_*).
enum class Precision : int {
    Fp64 = MKO_PRECISION_FP64,
    Df64 = MKO_PRECISION_DF64,
    Mixed = MKO_PRECISION_MIXED,
    Fp64Mfma = MKO_PRECISION_FP64_MFMA
};
enum class JacobianMethod : int { FiniteDifferences = MKO_JACOBIAN_FD, Adjoint = MKO_JACOBIAN_ADJOINT };

// Device time per kernel family (mko_engine_kernel_timing).
struct KernelTiming {
    double ms[6] = {};
    int64_t launches[6] = {};
    static constexpr const char* names[6] = {"trace", "adjoint", "gram", "fd_difference",
                                             "point_batch", "lm_control"};
};

class TraceEngine {
public:
    // Engine on HIP device `device`; memory_limit_bytes 0: the device's free memory minus a
    // reserve (EngineOptions::memory_limit_bytes).
    explicit TraceEngine(int device = 0, std::size_t memory_limit_bytes = 0)
        : h_(mko_engine_create(device, memory_limit_bytes)) {
        if (!h_) check(MKO_ERR_DEVICE);
    }
    ~TraceEngine() {
        if (h_) mko_engine_destroy(h_);
    }
    TraceEngine(const TraceEngine&) = delete;
    TraceEngine& operator=(const TraceEngine&) = delete;
    TraceEngine(TraceEngine&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    TraceEngine& operator=(TraceEngine&& o) noexcept {
        if (this != &o) {
            if (h_) mko_engine_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }
    mko_engine* handle() const { return h_; }

    // run(batch, TraceMode::Residuals): residuals [V][R]; opd / status [V][rays] (may be empty).
    void residuals(const mko_batch& batch, const mko_residual_settings& s,
                   std::span<double> residuals, std::span<double> opd = {},
                   std::span<uint8_t> status = {}) {
        check(mko_engine_residuals(h_, &batch, &s, detail::data_or_null(residuals),
                                   detail::data_or_null(opd), detail::data_or_null(status)));
    }
    // run(batch, TraceMode::Merit) with extra_merit [V] (may be empty).
    void merit(const mko_batch& batch, const mko_residual_settings& s, std::span<double> merit,
               std::span<int64_t> failed, std::span<int64_t> hard_failed,
               std::span<const double> extra_merit = {}) {
        detail::need(merit.size() >= std::size_t(batch.variants) &&
                         failed.size() >= std::size_t(batch.variants) &&
                         hard_failed.size() >= std::size_t(batch.variants),
                     "merit outputs need one entry per variant");
        check(mko_engine_merit(h_, &batch, &s, detail::data_or_null(extra_merit), merit.data(),
                               failed.data(), hard_failed.data()));
    }
    // TraceSettings::extrema: [V][S][2] largest ray height, smallest cos I.
    void surface_extrema(const mko_batch& batch, const mko_residual_settings& s,
                         std::span<double> extrema) {
        detail::need(extrema.size() >= std::size_t(batch.variants) * std::size_t(batch.surfaces) * 2,
                     "extrema needs [V][S][2]");
        check(mko_engine_surface_extrema(h_, &batch, &s, extrema.data()));
    }
    // prefilter_variants: the result's arrays are the caller's (mko_prefilter_result).
    void prefilter(const mko_batch& batch, const mko_residual_settings& s,
                   const mko_prefilter_options& o, mko_prefilter_result& result,
                   std::span<const double> extra_merit = {}) {
        check(mko_engine_prefilter(h_, &batch, &s, &o, detail::data_or_null(extra_merit),
                                   &result));
    }
    // Jacobians and / or normal equations of `runs` points x0 [runs][P] (any output may be
    // empty): J [runs][P][R], r0 [runs][R], a [runs][P][P], g [runs][P].
    void jacobian(const mko_batch& base, const mko_residual_settings& s, const mko_variables& v,
                  std::span<const double> x0, JacobianMethod method, Precision precision,
                  std::span<double> J, std::span<double> r0, std::span<double> a,
                  std::span<double> g) {
        const int runs = runs_of(x0, v);
        check(mko_engine_jacobian(h_, &base, &s, &v, runs, x0.data(), int(method),
                                  int(precision), detail::data_or_null(J),
                                  detail::data_or_null(r0), detail::data_or_null(a),
                                  detail::data_or_null(g)));
    }
    // LmBackend::normal_equations of the runs `runs` (ascending; empty: 0 .. count-1) at x.
    void normal_equations(const mko_batch& base, const mko_residual_settings& s,
                          const mko_variables& v, std::span<const int32_t> runs,
                          std::span<const double> x, JacobianMethod method, Precision precision,
                          std::span<double> a, std::span<double> g) {
        const int count = runs_of(x, v);
        detail::need(runs.empty() || runs.size() == std::size_t(count), "one run index per point");
        const std::size_t P = std::size_t(v.count);
        detail::need(a.size() >= std::size_t(count) * P * P && g.size() >= std::size_t(count) * P,
                     "a needs [count][P][P], g [count][P]");
        check(mko_engine_normal_equations(h_, &base, &s, &v, count, detail::data_or_null(runs),
                                          x.data(), int(method), int(precision), a.data(),
                                          g.data()));
    }
    // LmBackend::merit of the runs `runs` at x.
    void point_merit(const mko_batch& base, const mko_residual_settings& s,
                     const mko_variables& v, std::span<const int32_t> runs,
                     std::span<const double> x, std::span<double> merit,
                     std::span<int64_t> failed, std::span<int64_t> hard_failed) {
        const int count = runs_of(x, v);
        detail::need(runs.empty() || runs.size() == std::size_t(count), "one run index per point");
        detail::need(merit.size() >= std::size_t(count) && failed.size() >= std::size_t(count) &&
                         hard_failed.size() >= std::size_t(count),
                     "merit outputs need one entry per point");
        check(mko_engine_point_merit(h_, &base, &s, &v, count, detail::data_or_null(runs),
                                     x.data(), merit.data(), failed.data(), hard_failed.data()));
    }

    // Launch shapes (mko_tuning) and the autotuner; no counterpart in the reference.
    mko_tuning tuning() const {
        mko_tuning t{};
        check(mko_engine_get_tuning(h_, &t));
        return t;
    }
    void set_tuning(const mko_tuning& t) { check(mko_engine_set_tuning(h_, &t)); }
    mko_tuning autotune(const mko_batch& base, const mko_residual_settings& s,
                        const mko_variables& v, std::span<const double> x) {
        mko_tuning t{};
        check(mko_engine_autotune(h_, &base, &s, &v, runs_of(x, v), x.data(), &t));
        return t;
    }
    // Kernel times since the previous call; `enable` switches the timing on or off.
    KernelTiming kernel_timing(bool enable) {
        KernelTiming t;
        check(mko_engine_kernel_timing(h_, enable ? 1 : 0, t.ms, t.launches));
        return t;
    }

    static int runs_of(std::span<const double> x, const mko_variables& v) {
        detail::need(v.count > 0, "no variables");
        detail::need(x.size() % std::size_t(v.count) == 0, "x must hold whole points [runs][P]");
        return int(x.size() / std::size_t(v.count));
    }

private:
    mko_engine* h_ = nullptr;
};

// Result of an LM: x [runs][P] and one mko_lm_run per run.
struct LmResult {
    std::vector<double> x;
    std::vector<mko_lm_run> runs;
    int cycles = 0;
};

// optimize_lm_gpu: the whole LM (outer penalty rounds included) on the engine's device.
inline LmResult optimize_lm(TraceEngine& engine, const mko_batch& base,
                            const mko_residual_settings& s, const mko_variables& v,
                            std::span<const double> x0, const mko_lm_options& o,
                            Precision precision = Precision::Fp64) {
    const int runs = TraceEngine::runs_of(x0, v);
    LmResult r;
    r.x.resize(x0.size());
    r.runs.resize(std::size_t(runs));
    check(mko_optimize_lm(engine.handle(), &base, &s, &v, runs, x0.data(), &o, int(precision),
                          r.x.data(), r.runs.data(), &r.cycles));
    return r;
}

// The same LM one cycle at a time. The session keeps pointers into the caller's batch and
// variable arrays: they must outlive it.
class LmSession {
public:
    LmSession(TraceEngine& engine, const mko_batch& base, const mko_residual_settings& s,
              const mko_variables& v, std::span<const double> x0, const mko_lm_options& o,
              Precision precision = Precision::Fp64)
        : runs_(TraceEngine::runs_of(x0, v)), P_(v.count),
          h_(mko_lm_create(engine.handle(), &base, &s, &v, runs_, x0.data(), &o, int(precision))) {
        if (!h_) {
            const char* m = mko_last_error();
            // mko_lm_create reports argument errors and device errors alike through NULL
            throw std::invalid_argument(m ? m : "mko_lm_create failed");
        }
    }
    ~LmSession() {
        if (h_) mko_lm_destroy(h_);
    }
    LmSession(const LmSession&) = delete;
    LmSession& operator=(const LmSession&) = delete;
    LmSession(LmSession&& o) noexcept
        : runs_(o.runs_), P_(o.P_), h_(std::exchange(o.h_, nullptr)) {}

    // One cycle; false when every run is done.
    bool cycle() {
        const int rc = mko_lm_cycle(h_);
        check(rc);
        return rc == 1;
    }
    // Split form for sessions that run at once: begin every session's cycle, then end them.
    int cycle_begin() {
        const int rc = mko_lm_cycle_begin(h_);
        check(rc);
        return rc;
    }
    bool cycle_end() {
        const int rc = mko_lm_cycle_end(h_);
        check(rc);
        return rc == 1;
    }
    int graph_cycles() const { return mko_lm_graph_cycles(h_); }
    LmResult results() {
        LmResult r;
        r.x.resize(std::size_t(runs_) * std::size_t(P_));
        r.runs.resize(std::size_t(runs_));
        check(mko_lm_results(h_, r.x.data(), r.runs.data(), &r.cycles));
        return r;
    }
    mko_lm* handle() const { return h_; }

private:
    int runs_, P_;
    mko_lm* h_ = nullptr;
};

} // namespace mko_hip

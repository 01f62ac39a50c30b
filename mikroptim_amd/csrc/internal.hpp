// Library-internal entry points shared by kernels.hip and the host layer (not exported in the
// public C-ABI).
#pragma once

#include <cstdint>

#include "mikroptim_types.h"

namespace mko_internal {

// Final ray states of a MERIT trace, kept for the fused adjoint at the same point: per ray the
// position and direction on the last surface and the OPL at the reference sphere (NaN for a ray
// that failed), structure of arrays: data[(c * slots + slot) * rays + row], c = p.x, p.y, p.z,
// d.x, d.y, d.z, opl. Variant v of a launch uses slot slot[v] (slot == NULL: slot_offset + v).
// Launch shapes the engine's autotuner chooses: pack_rays (0: automatic; 32, 64, 128: the
// segment of a tuple pack, used where the tuples fit it; 256: one tuple per workgroup) and
// fused_waves (0: automatic; 1, 2, 4: waves per workgroup of the fused normal equations). The
// kernels of the calling thread use them; the host layer sets them for each of its calls.
struct Tuning {
    int pack_rays = 0;
    int fused_waves = 0;
    int full_row_traces = 0; // 1: the constraint rows' differences trace every tuple
};
void set_thread_tuning(const Tuning& t);
Tuning thread_tuning();
struct TuningScope {
    Tuning saved;
    explicit TuningScope(const Tuning& t) : saved(thread_tuning()) { set_thread_tuning(t); }
    ~TuningScope() { set_thread_tuning(saved); }
    TuningScope(const TuningScope&) = delete;
    TuningScope& operator=(const TuningScope&) = delete;
};

struct RayStates {
    double* data;
    const int32_t* slot;
    int32_t slot_offset;
    int64_t slots;
    int64_t rays;
};

// MERIT (mko_merit_device) that also writes the final ray states (hard failure mode, FP64).
// `live` (device, NULL: all variants): only variants v < *live are traced and finalized; the
// launch is sized for the batch, so a captured graph serves any count the device sets.
int merit_states(const mko_batch* batch, const mko_residual_settings* settings, double* merit,
                 int64_t* failed, int64_t* hard_failed, double* partial, int* partial_counts,
                 const RayStates& states, void* stream, const int32_t* live = nullptr);

// mko_adjoint_normal_device starting each ray's reverse sweep from the stored final state
// instead of tracing it forward (the states must come from a MERIT trace of the same points);
// with states, glass variables are taken as well. waves: workgroup size of the fused kernel in
// wavefronts (1, 2 or 4; 0 chooses from the tuple size).
int adjoint_normal_states(const mko_batch* batch, const mko_residual_settings* settings,
                          const mko_variables* vars, double* a, double* g, void* workspace,
                          size_t workspace_bytes, const RayStates& states, void* stream,
                          const int32_t* live = nullptr, int waves = 0);

// mko_point_batch_device with an optional device array template_index [runs]: run i copies
// template variant template_index[i] (NULL: as mko_point_batch_device).
int point_batch(const mko_batch* templ, const int32_t* template_index, int template_per_run,
                int runs, int per_run, const mko_variables* vars_device, const double* x0,
                const double* plus, const double* minus, double* curvature, double* thickness,
                double* index, double* pupil_radius, double* field_angle, double* field_target,
                double* reference_curvature, double* paraxial_index, double* buchdahl,
                void* stream, const int32_t* live = nullptr);

// RESIDUALS (mode 0: rows at `stride` >= trace rows) or MERIT partials (mode 1: partial [V*T],
// partial_counts [2*V*T], finalized by merit_finalize) of a device batch, optionally with the
// LM's final ray states (mode 1) and the surface extrema: ext_partial [V*T][S][2] workspace,
// extrema [V][S][2] = {largest ray height, smallest cos I} over the rays with status Ok
// (surface_extrema_cpu, src/batch/batch_extrema.cpp:75). ext_partial == NULL: no extrema.
// `surfaces`: bit s set = surface s is reduced (the others stay at the neutral 0 / 1).
int trace_extrema(const mko_batch* batch, const mko_residual_settings* settings, int mode,
                  double* residuals, int64_t stride, double* opd, uint8_t* status,
                  double* partial, int* partial_counts, const RayStates* states,
                  double* ext_partial, double* extrema, uint64_t surfaces, void* stream);

// Extrema of the constraint rows' differences over candidate rays. base: the runs' points (V
// runs), fd: their difference batch (V * per_run variants); ex0 [V][S][2] the extrema at the
// points. Marks the candidate rays (flags [V * rays]; failed rays, heights within delta of the
// largest on hmask, cos I within delta (1 - cos) of the smallest on cmask), lists them in order
// (tstart [V][T + 1], list [V][cap], overflow [V]; tcount [V][T], others [V] workspace) and
// writes ext [V * per_run][S][2] over the candidates (extrema_check tells whether that is the
// extrema of every ray). FP64 traces only.
int candidate_extrema(const mko_batch* base, const mko_batch* fd,
                      const mko_residual_settings* settings, int per_run, uint64_t hmask,
                      uint64_t cmask, double delta, const double* ex0, uint8_t* flags,
                      int* tcount, int* others, int* tstart, int* list, int cap, int* overflow,
                      double* ext, void* stream);
// flag[r] = 1 where the list overflowed, or the candidates' extrema over the run's per_run points
// moved by D and the other rays (bounded through delta) lie within K D of them.
int extrema_check(int runs, int per_run, int S, uint64_t hmask, uint64_t cmask, double delta,
                  const double* ex, const double* ex0, const int* others, const int* overflow,
                  double K, int* flag, void* stream);

// FP32 prefilter stats of a device batch [V] (prefilter_kernel + second stage); workspace of
// 24 bytes per (variant, tuple).
int prefilter_stats(const mko_batch* batch, const mko_residual_settings* settings, double boundary,
                    void* workspace, double* merit, double* optical, int64_t* failed,
                    int64_t* uncertain_ok, int64_t* uncertain_failed, void* stream);

// merit[v] = (sum of the tuple partials + extra[v]) + traced[v] (extra / traced may be NULL).
int merit_finalize(int V, int T, const double* partial, const int* partial_counts,
                   const double* extra, const double* traced, double* merit, int64_t* failed,
                   int64_t* hard_failed, void* stream);

// Constraint rows of a device batch: rows[v * stride + offset + k] for the `count` items
// (device memory, row order) with the extrema [V][S][2] of the traced items.
int constraint_rows(const mko_batch* batch, const mko_constraint* items, int count,
                    int need_paraxial, const double* extrema, double* rows, int64_t stride,
                    int64_t offset, void* stream, const double* glass_hull = nullptr);

// e_sum[v] = (extra ? extra[v] + : ) sum of the first E rows squared, t_sum[v] = sum of the
// next Et rows squared (either output may be NULL).
int rows_merit(int V, const double* rows, int64_t stride, int64_t offset, int E, int Et,
               const double* extra, double* e_sum, double* t_sum, void* stream);

// a [runs][P][P] += Je^T Je, g [runs][P] += Je^T re for the constraint rows' Jacobian
// Je [runs][P][E] and values re [runs][E] (fused path).
int extra_normal(int runs, int P, int E, const double* Je, const double* re, double* a, double* g,
                 void* stream);

// mko_adjoint_jacobian_device (FP64) with J / r0 rows at `stride` (0: trace rows); `live`
// (device, NULL: unused): only variants < *live are computed (launch sized for all).
int adjoint_jacobian(const mko_batch* batch, const mko_residual_settings* settings,
                     const mko_variables* vars, double* jacobian, double* r0, uint8_t* status,
                     int64_t stride, void* stream, const int32_t* live = nullptr);

// mko_gram_device (the bitwise FMA chain) with the runs limited to *live (device, NULL: unused).
int gram_chain(const double* jacobian, const double* r0, int runs, int variables, int64_t rows,
               double* a, double* g, void* stream, const int32_t* live);

// mko_fd_difference_device with J column (i, p) at jacobian + (i * P + p) * jstride + joffset,
// r0 of run i at r0 + i * jstride + joffset; status / status_changes may be NULL.
int fd_difference(const double* rows, int64_t stride, const uint8_t* status, int64_t rays,
                  int runs, int variables, int64_t count, const double* step_plus,
                  const double* step_minus, const uint8_t* stencil, double* jacobian, double* r0,
                  int64_t* status_changes, int64_t jstride, int64_t joffset, void* stream);

} // namespace mko_internal

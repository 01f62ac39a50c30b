/* C-ABI of libmikroptim_hip.so: the MI355X (gfx950) implementation of Mikroptim's batched
 * forward-and-Jacobian path. Plain C types only (pointers, sizes, the structs of
 * mikroptim_types.h); every entry point returns 0 or a negative MKO_ERR_* code, and
 * mko_last_error() gives the message of the last failure on the calling thread. Arguments are
 * checked before any device work (std::invalid_argument in the reference -> MKO_ERR_INVALID);
 * device and runtime failures give MKO_ERR_DEVICE (mko::cuda::CudaError in the reference).
 *
 * Two layers:
 *  - *_device: one kernel family each, on a device-resident batch (per-variant arrays of the
 *    mko_batch in device memory, batch-wide arrays in host memory), asynchronous on `stream`
 *    (a hipStream_t, NULL = default stream).
 *  - mko_engine_* / mko_lm_*: the host layer (TraceEngine / GpuLmBackend / levenberg_marquardt
 *    of the reference) on host-resident batches; synchronous like the reference. */
#ifndef MIKROPTIM_HIP_H
#define MIKROPTIM_HIP_H

#include "mikroptim_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MKO_ERR_INVALID (-1) /* bad arguments, nothing launched */
#define MKO_ERR_DEVICE (-2)  /* HIP runtime or device failure */

const char* mko_last_error(void);

/* ---- device layer ------------------------------------------------------------------------ */

/* The RESIDUALS and MERIT entry points trace in settings->precision: MKO_PRECISION_FP64, or
 * MKO_PRECISION_DF64 (Precision::Df64 of the reference: the ray trace in df64, OPL converted to
 * double for the piston and the residuals); other values are rejected. */

/* RESIDUALS mode: residual rows r_k (and, in soft mode, penalty rows after the ray rows) of
 * every variant at residuals[v * stride ...], OPD [V][rays] and status [V][rays] (both may be
 * NULL). Replaces residuals_kernel + DeviceResiduals::run (src/cuda/residuals.cu:58, :240;
 * include/mikroptim/cuda/residuals.hpp). */
int mko_residuals_device(const mko_batch* batch, const mko_residual_settings* settings,
                         double* residuals, int64_t stride, double* opd, uint8_t* status,
                         void* stream);

/* MERIT mode: merit[v] = sum over tuples of (sum r^2 + sum p^2) + extra_merit[v] (NULL = 0),
 * failed / hard-failed ray counts. Workspace: partial [V * F * L] doubles, partial_counts
 * [2 * V * F * L] ints. Replaces merit_kernel + merit_finalize_kernel + DeviceMerit::run
 * (src/cuda/merit.cu:63, :207; include/mikroptim/cuda/merit.hpp). */
int mko_merit_device(const mko_batch* batch, const mko_residual_settings* settings,
                     const double* extra_merit, double* merit, int64_t* failed,
                     int64_t* hard_failed, double* partial, int* partial_counts, void* stream);

/* mko_merit_device without the workspace: the tuple partials live in a static device array of
 * the library (8 MiB per GPU), so the call allocates no device memory and calls no allocating
 * routine. No counterpart in the reference; for callers that cannot allocate (code under
 * hipGraph capture, serving loops, harnesses that forbid hipMalloc). At most 524288
 * (variant, field, wavelength) tuples per call (variants * fields * wavelengths); larger batches
 * are rejected with MKO_ERR_INVALID. All calls on one GPU share the array: calls that overlap
 * on one GPU (other streams or threads) race, calls ordered on one stream do not. Results equal
 * mko_merit_device's bit for bit. */
int mko_merit_device_noalloc(const mko_batch* batch, const mko_residual_settings* settings,
                             const double* extra_merit, double* merit, int64_t* failed,
                             int64_t* hard_failed, void* stream);

/* Writes runs * per_run variants: variant (i, j) is a copy of template variant i (or 0 if
 * !template_per_run) with the variables x0[i] applied; for j > 0, component (j - 1) / 2 is
 * replaced by plus[i] (j odd) or minus[i] (j even). per_run = 2P + 1 gives the central FD
 * pattern (write_fd_variants, src/optim/variant_generator.cpp:240-262), per_run = 1 the point
 * batch of LM merit / adjoint calls (lm_point_batch, src/optim/levenberg_marquardt.cpp:634).
 * vars_device: variable type / surface arrays in device memory; x0, plus, minus [runs][P] in
 * device memory; outputs are the per-variant arrays of the destination batch (paraxial_index
 * NULL: not copied). */
int mko_point_batch_device(const mko_batch* templ, int template_per_run, int runs, int per_run,
                           const mko_variables* vars_device, const double* x0,
                           const double* plus, const double* minus, double* curvature,
                           double* thickness, double* index, double* pupil_radius,
                           double* field_angle, double* field_target,
                           double* reference_curvature, double* paraxial_index, void* stream);

/* FD Jacobian from the residuals of a central FD batch (runs * (2P + 1) variants, rows per
 * variant at stride): J [runs][P][rows] by fd_derivative, r0 [runs][rows], status changes
 * [runs][P]. Replaces fd_difference_kernel (src/cuda/fd_jacobian.cu:65). */
int mko_fd_difference_device(const double* fd_residuals, int64_t stride,
                             const uint8_t* fd_status, int64_t rays, int runs, int variables,
                             int64_t rows, const double* step_plus, const double* step_minus,
                             const uint8_t* stencil, double* jacobian, double* r0,
                             int64_t* status_changes, void* stream);

/* ADJOINT Jacobian of a batch with one variant per run: J [V][P][trace_rows], r0 [V][trace_rows],
 * status [V][rays] (may be NULL), all FP64. precision MKO_PRECISION_FP64 or MKO_PRECISION_MIXED;
 * both compute the stored Jacobian in FP64 (the mixed precision of this library is the MFMA
 * normal equations of mko_adjoint_normal_device). Replaces adjoint_jacobian_kernel +
 * DeviceAdjointJacobian::compute (src/cuda/adjoint_jacobian.cu:119, :401). */
int mko_adjoint_jacobian_device(const mko_batch* batch, const mko_residual_settings* settings,
                                const mko_variables* vars, int precision, double* jacobian,
                                double* r0, uint8_t* status, void* stream);

/* ADJOINT Jacobian fused with the normal equations, the mixed-precision path of the LM: per run
 * (one variant per run) a [V][P][P] = J^T J and g [V][P] = J^T r0 without storing J. Forward
 * trace and reverse sweep in FP64; per-ray gradients staged in LDS as float differences to a
 * per-wave pilot ray; J^T J from split-bf16 products (hi*hi + hi*lo + lo*hi) on MFMA with float
 * accumulation per (field, wavelength) tuple and FP64 across tuples; J^T r0 in FP64. Hard failure
 * mode, at most 96 variables (curvature and thickness; glass variables need the LM path's stored
 * ray states). Workspace: mko_adjoint_normal_workspace() bytes of device memory. Replaces adjoint_jacobian_kernel + gram_kernel
 * (src/cuda/adjoint_jacobian.cu:119, src/cuda/normal_equations.cu:48) on the LM path, where only
 * J^T J and J^T r0 are consumed (src/optim/levenberg_marquardt.cpp:427-447). */
size_t mko_adjoint_normal_workspace(int variants, int fields, int wavelengths, int variables);
int mko_adjoint_normal_device(const mko_batch* batch, const mko_residual_settings* settings,
                              const mko_variables* vars, double* a, double* g, void* workspace,
                              size_t workspace_bytes, void* stream);

/* Vector-Jacobian product of the ADJOINT Jacobian without forming J: out [V][P] = sum over the
 * trace rows of grad[v][row] * J[v][p][row], J the FP64 adjoint Jacobian of
 * mko_adjoint_jacobian_device (piston-projected ray rows; penalty rows in soft mode) of a batch
 * with one variant per run. grad [V][grad_stride] (grad_stride >= trace rows) and out in device
 * memory. Each ray's reverse sweep is seeded with its cotangents (the sweep of
 * ray_gradient_rows, include/mikroptim/merit/ray_jacobian.hpp:93, is linear in its seeds); sums
 * in fixed order. FP64 trace; settings with constraint rows are rejected. Workspace:
 * mko_vjp_workspace() bytes. The reference's bindings form J for this product
 * (python/src/bind_trace.cpp:623); this entry point serves the torch autograd backward. */
size_t mko_vjp_workspace(int variants, int fields, int wavelengths, int variables);
int mko_vjp_device(const mko_batch* batch, const mko_residual_settings* settings,
                   const mko_variables* vars, const double* grad, int64_t grad_stride,
                   double* out, void* workspace, size_t workspace_bytes, void* stream);

/* Normal equations of each run: a [runs][P][P] = J^T J (a[q*P + p]), g [runs][P] = J^T r0.
 * Every element is one FMA chain over the rows in ascending order (bitwise equal to
 * normal_equations_host). Replaces gram_kernel + DeviceNormalEquations::compute
 * (src/cuda/normal_equations.cu:48, :155). */
int mko_gram_device(const double* jacobian, const double* r0, int runs, int variables,
                    int64_t rows, double* a, double* g, void* stream);

/* The same normal equations on FP64 MFMA (v_mfma_f64_16x16x4_f64, 16 x 16 tiles, FP64
 * accumulation): equal to mko_gram_device up to rounding (another summation order, so not
 * bitwise the host chain); up to 127 variables. Replaces gram_kernel
 * (src/cuda/normal_equations.cu:48) where bitwise parity with normal_equations_host is not
 * required (MKO_PRECISION_FP64_MFMA in the host layer). */
int mko_gram_mfma_device(const double* jacobian, const double* r0, int runs, int variables,
                         int64_t rows, double* a, double* g, void* stream);

/* df64 (double-float) arithmetic check: out[i] = a[i] op b[i] over n pairs (hi, lo) of floats in
 * device memory ([n][2] each); op 0 add, 1 sub, 2 mul, 3 mul by the float b.hi, 4 div, 5 sqrt(a).
 * The operations of the MKO_PRECISION_DF64 trace, for comparison with a host implementation
 * (df64_check_kernel, src/cuda/df64_check.cu:17; CudaDf64.DeviceMatchesHostBitwise). */
int mko_df64_ops_device(int op, int64_t n, const float* a, const float* b, float* out,
                        void* stream);

/* ---- host layer ----------------------------------------------------------------------------- */

typedef struct mko_engine mko_engine;

/* TraceEngine(int device, EngineOptions) (include/mikroptim/cuda/trace_engine.hpp):
 * memory_limit_bytes = 0 uses the device's free memory minus a reserve. */
mko_engine* mko_engine_create(int device, size_t memory_limit_bytes);
void mko_engine_destroy(mko_engine* engine);

/* Device time per kernel family, measured with HIP events on the launch stream around each
 * launch while enabled (AdjointJacobianTiming, include/mikroptim/cuda/adjoint_jacobian.hpp:39):
 * ms / launches [6] = trace (RESIDUALS and MERIT), adjoint (stored-J or fused normal
 * equations), gram, fd difference, point batch, LM control (device-side step solves and
 * updates). Returns the totals since the previous call, resets them and sets the enable flag. */
int mko_engine_kernel_timing(mko_engine* engine, int enable, double* ms, int64_t* launches);

/* Launch shapes of the engine's kernels (mko_tuning): tuple-pack segment of the FP64 RESIDUALS /
 * MERIT traces and waves per workgroup of the fused normal equations, 0 = automatic; every later
 * host-layer call on the engine (and its LM sessions) uses them. Packed traces equal unpacked
 * ones bit for bit. No counterpart in the reference (fixed pack and block shapes,
 * include/mikroptim/cuda/tuple_packs.hpp:61). */
int mko_engine_set_tuning(mko_engine* engine, const mko_tuning* tuning);
int mko_engine_get_tuning(const mko_engine* engine, mko_tuning* tuning);

/* Constraint rows from trace extrema differentiated by central differences (the adjoint paths'
 * rows, extra_rows_cpu, src/optim/adjoint_jacobian.cpp:245-337): number of chunks whose
 * difference points traced only the candidate rays, and of those traced whole again because a
 * run missed the check's margin (mko_tuning.full_row_traces). No counterpart in the reference. */
int mko_engine_row_trace_stats(const mko_engine* engine, int64_t* candidate_chunks,
                               int64_t* full_chunks);

/* Autotuner: times the candidate launch shapes of the MERIT trace (tuple packing off, or each
 * pack segment that holds the batch's largest tuple) and of the fused normal equations (1, 2 or 4
 * waves per workgroup where they apply) on the first min(runs, 64) points x [runs][P] of the base
 * (one variant, or one per point), keeps the fastest, caches the choice per shape and sets it as
 * the engine's tuning; `chosen` (may be NULL) receives it. LM sessions keep the tuning the engine
 * had when they were created. No counterpart in the reference (fixed pack and block shapes,
 * include/mikroptim/cuda/tuple_packs.hpp:61, src/cuda/wavefront_block.hpp:504-526). */
int mko_engine_autotune(mko_engine* engine, const mko_batch* base,
                        const mko_residual_settings* settings, const mko_variables* variables,
                        int runs, const double* x, mko_tuning* chosen);

/* Constraint rows (settings->extra, see mikroptim_types.h): every host-layer entry point below
 * appends them to each variant's residual vector, R = trace_rows + E + E_t rows, as the
 * reference's FdJacobianOptions::extra / trace_extra do: residuals, J and r0 carry R rows, the
 * merit is (ray merit + sum e^2) + sum t^2 (CpuLmBackend::merit, src/optim/levenberg_marquardt.cpp:866),
 * a and g include the rows. Rows that need extrema take them from the same trace (every FD
 * variant from its own); the adjoint paths differentiate the rows by central differences of
 * the FD batch (extra_rows_cpu, src/optim/adjoint_jacobian.cpp:245-337). The device-layer
 * entry points above reject settings with constraint rows. */

/* TraceEngine::run(batch, TraceMode::Residuals, ...): host batch in, host results out
 * (residuals [V][R], opd / status [V][rays], either may be NULL). */
int mko_engine_residuals(mko_engine* engine, const mko_batch* batch,
                         const mko_residual_settings* settings, double* residuals, double* opd,
                         uint8_t* status);

/* TraceEngine::run(batch, TraceMode::Merit, ...) with MeritRunOptions::extra_merit (added to
 * the constraint rows' sum e^2 when both are present). */
int mko_engine_merit(mko_engine* engine, const mko_batch* batch,
                     const mko_residual_settings* settings, const double* extra_merit,
                     double* merit, int64_t* failed, int64_t* hard_failed);

/* Max reductions of the MERIT trace (TraceSettings::extrema, SurfaceExtremaResult,
 * include/mikroptim/batch/batch_extrema.hpp:29): extrema [V][S][2] = {largest ray height
 * sqrt(x^2 + y^2) at surface s, smallest cos I = d . N before the refraction} over the rays with
 * status Ok of all fields, wavelengths and samples; surface 0: the start state; neutral 0 / 1
 * without a valid ray. */
int mko_engine_surface_extrema(mko_engine* engine, const mko_batch* batch,
                               const mko_residual_settings* settings, double* extrema);

/* DeviceFdJacobian::compute / DeviceAdjointJacobian::compute + download
 * (include/mikroptim/cuda/fd_jacobian.hpp:58, adjoint_jacobian.hpp:70): base has one variant
 * (shared template) or one per run; x0 [runs][P]. Outputs (host, any may be NULL):
 * J [runs][P][R] double, r0 [runs][R], a [runs][P][P], g [runs][P].
 * method ADJOINT with precision MKO_PRECISION_MIXED, J and r0 NULL and <= 96 variables takes a
 * and g from the fused kernel of mko_adjoint_normal_device (with glass variables on the ray
 * states of a MERIT trace of the same points): in hard failure mode for every run, in soft mode
 * for the runs whose point has no failed ray (their penalty rows and penalty derivatives are
 * zero, so their normal equations are hard mode's). MKO_PRECISION_FP64_MFMA computes J in FP64
 * and a, g with mko_gram_mfma_device. Every other run or combination computes J in FP64 and
 * a, g with mko_gram_device. */
int mko_engine_jacobian(mko_engine* engine, const mko_batch* base,
                        const mko_residual_settings* settings, const mko_variables* vars,
                        int runs, const double* x0, int method, int precision, double* jacobian,
                        double* r0, double* a, double* g);

/* GpuLmBackend::normal_equations / merit (src/cuda/levenberg_marquardt.cpp:70-248): the two
 * pure virtuals of mko::LmBackend (include/mikroptim/optim/levenberg_marquardt.hpp:171-182) for
 * `count` points x [count][P] of the runs `runs` [count] (ascending run indices into base when
 * base has one variant per run; NULL = 0 .. count-1; ignored for a one-variant base).
 * a [count][P][P] (a[j*P*P + q*P + p]), g [count][P]; merit [count], failed / hard [count].
 * A run's results depend only on its index and point. */
int mko_engine_normal_equations(mko_engine* engine, const mko_batch* base,
                                const mko_residual_settings* settings, const mko_variables* vars,
                                int count, const int32_t* runs, const double* x, int method,
                                int precision, double* a, double* g);
int mko_engine_point_merit(mko_engine* engine, const mko_batch* base,
                           const mko_residual_settings* settings, const mko_variables* vars,
                           int count, const int32_t* runs, const double* x, double* merit,
                           int64_t* failed, int64_t* hard_failed);

/* FP32 prefilter (TraceEngine::prefilter_variants, include/mikroptim/cuda/trace_engine.hpp;
 * prefilter_kernel src/cuda/prefilter.cu:87; prefilter_decide src/batch/batch_prefilter.cpp:236):
 * every ray traced in float with its boundary margin, per-variant stats, then the conservative
 * decision in double. extra_merit [V] (FP64, e.g. constraint rows' sum e^2) or NULL. The
 * values only select variants: they never enter a merit, residual or Jacobian. */
int mko_engine_prefilter(mko_engine* engine, const mko_batch* batch,
                         const mko_residual_settings* settings,
                         const mko_prefilter_options* options, const double* extra_merit,
                         mko_prefilter_result* result);

/* The decision alone (host, no device work) from the stats already in `result` (merit,
 * optical_length, failed, uncertain_ok, uncertain_failed): fills merit_lower / merit_upper /
 * reject / reference_* / kept (prefilter_decide, src/batch/batch_prefilter.cpp:236). */
int mko_prefilter_decide(const mko_batch* batch, const mko_residual_settings* settings,
                         const mko_prefilter_options* options, const double* extra_merit,
                         mko_prefilter_result* result);

/* Batched Levenberg-Marquardt on the GPU backend (optimize_lm_gpu,
 * include/mikroptim/cuda/levenberg_marquardt.hpp:91; control logic of
 * src/optim/levenberg_marquardt.cpp:332-572, outer penalty rounds :891). precision applies to
 * the adjoint Jacobian. x_out [runs][P], run results [runs]. */
int mko_optimize_lm(mko_engine* engine, const mko_batch* base,
                    const mko_residual_settings* settings, const mko_variables* vars, int runs,
                    const double* x0, const mko_lm_options* options, int precision,
                    double* x_out, mko_lm_run* run_out, int* cycles_out);

/* Stepwise form of the same LM (one handle, one cycle per mko_lm_cycle call) for callers that
 * time or interleave cycles. mko_lm_cycle returns 1 while runs are active, 0 when all are done,
 * < 0 on error. */
typedef struct mko_lm mko_lm;
mko_lm* mko_lm_create(mko_engine* engine, const mko_batch* base,
                      const mko_residual_settings* settings, const mko_variables* vars, int runs,
                      const double* x0, const mko_lm_options* options, int precision);
int mko_lm_cycle(mko_lm* lm);
int mko_lm_results(mko_lm* lm, double* x_out, mko_lm_run* run_out, int* cycles_out);
void mko_lm_destroy(mko_lm* lm);
/* Number of the session's cycles that ran as one replay of a captured hipGraph (device control
 * with the fused mixed-precision adjoint, see mko_lm_options.host_control); no counterpart in the
 * reference. */
int mko_lm_graph_cycles(const mko_lm* lm);
/* Split form of mko_lm_cycle for LMs that run at once (for instance one per topology of an
 * uneven batch): mko_lm_cycle_begin launches the session's captured cycle on its own stream and
 * returns 1 without waiting (or runs a whole eager cycle and returns 2; 0: the LM is done, < 0:
 * error); mko_lm_cycle_end waits for it and returns what mko_lm_cycle would (1 while runs are
 * active). Begin every session's cycle, then end them: their graphs run concurrently. Results
 * equal those of mko_lm_cycle bit for bit. No counterpart in the reference. */
int mko_lm_cycle_begin(mko_lm* lm);
int mko_lm_cycle_end(mko_lm* lm);

#ifdef __cplusplus
}
#endif

#endif

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
struct RayStates {
    double* data;
    const int32_t* slot;
    int32_t slot_offset;
    int64_t slots;
    int64_t rays;
};

// MERIT (mko_merit_device) that also writes the final ray states (hard failure mode, FP64).
int merit_states(const mko_batch* batch, const mko_residual_settings* settings, double* merit,
                 int64_t* failed, int64_t* hard_failed, double* partial, int* partial_counts,
                 const RayStates& states, void* stream);

// mko_adjoint_normal_device starting each ray's reverse sweep from the stored final state
// instead of tracing it forward (the states must come from a MERIT trace of the same points).
int adjoint_normal_states(const mko_batch* batch, const mko_residual_settings* settings,
                          const mko_variables* vars, double* a, double* g, void* workspace,
                          size_t workspace_bytes, const RayStates& states, void* stream);

// mko_point_batch_device with an optional device array template_index [runs]: run i copies
// template variant template_index[i] (NULL: as mko_point_batch_device).
int point_batch(const mko_batch* templ, const int32_t* template_index, int template_per_run,
                int runs, int per_run, const mko_variables* vars_device, const double* x0,
                const double* plus, const double* minus, double* curvature, double* thickness,
                double* index, double* pupil_radius, double* field_angle, double* field_target,
                double* reference_curvature, void* stream);

} // namespace mko_internal

// Library-internal entry points shared by kernels.hip and the host layer (not exported in the
// public C-ABI).
#pragma once

#include <cstdint>

#include "mikroptim_types.h"

namespace mko_internal {

// mko_point_batch_device with an optional device array template_index [runs]: run i copies
// template variant template_index[i] (NULL: as mko_point_batch_device).
int point_batch(const mko_batch* templ, const int32_t* template_index, int template_per_run,
                int runs, int per_run, const mko_variables* vars_device, const double* x0,
                const double* plus, const double* minus, double* curvature, double* thickness,
                double* index, double* pupil_radius, double* field_angle, double* field_target,
                double* reference_curvature, void* stream);

} // namespace mko_internal

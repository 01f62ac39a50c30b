/* Plain-C data layout shared by the HIP library (libmikroptim_hip.so), its C++ host layer and
 * the CPU oracle: a system batch in structure-of-arrays form, the residual settings and the
 * variable map. Mirrors the reference's value types field by field:
 *   mko_batch              <- SystemBatchView<double>  include/mikroptim/batch/system_batch_view.hpp:108
 *                             (BatchShape :29, BatchTopology include/mikroptim/batch/system_batch.hpp:26)
 *   mko_residual_settings  <- ResidualSettings         include/mikroptim/batch/batch_residuals.hpp:24
 *   mko_variables          <- VariableSet / Variable   include/mikroptim/optim/variable_set.hpp:20,62
 *   mko_lm_options         <- LmOptions                include/mikroptim/optim/levenberg_marquardt.hpp:85
 * All indices are element (not byte) offsets, variant outermost. */
#ifndef MIKROPTIM_TYPES_H
#define MIKROPTIM_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MKO_MAX_SURFACES 64 /* kMaxSurfaces, include/mikroptim/trace/surface_table.hpp:15 */
#define MKO_MAX_RINGS 16    /* kMaxPupilRings, include/mikroptim/pupil/quadrature.hpp:21 */

/* SurfaceKind (include/mikroptim/trace/surface_table.hpp:26); ghost plates are not supported. */
enum { MKO_SURFACE_STOP = 0, MKO_SURFACE_PLANE = 1, MKO_SURFACE_SPHERE = 2 };
/* PupilMode (include/mikroptim/pupil/quadrature.hpp:23) */
enum { MKO_PUPIL_HALF = 0, MKO_PUPIL_FULL = 1, MKO_PUPIL_AXIS = 2 };
/* SourceKind (include/mikroptim/batch/system_batch_view.hpp:21) */
enum { MKO_SOURCE_COLLIMATED = 0, MKO_SOURCE_POINT = 1 };
/* ReferenceSphereKind (include/mikroptim/trace/reference_kind.hpp:11) */
enum { MKO_REF_LAST_SURFACE = 0, MKO_REF_IMAGE_EXIT_PUPIL = 1, MKO_REF_IMAGE_INFINITE = 2 };
/* FailureMode (include/mikroptim/trace/surface_ops.hpp:79) */
enum { MKO_FAILURE_SOFT = 0, MKO_FAILURE_HARD = 1 };
/* TraceStatus (include/mikroptim/trace/surface_ops.hpp:24), stored as one byte per ray */
enum {
    MKO_STATUS_OK = 0,
    MKO_STATUS_MISSED = 1,
    MKO_STATUS_TIR = 2,
    MKO_STATUS_BACKWARD = 3,
    MKO_STATUS_UNSUPPORTED = 4,
    MKO_STATUS_SOFT_MISSED = 5,
    MKO_STATUS_SOFT_TIR = 6
};
/* VariableType (include/mikroptim/model/variables.hpp:17), the kinds the library supports */
enum { MKO_VAR_CURVATURE = 0, MKO_VAR_THICKNESS = 1, MKO_VAR_WORKING_DISTANCE = 2, MKO_VAR_GLASS = 3 };
/* FdStencil (include/mikroptim/optim/fd_formula.hpp:15) */
enum { MKO_STENCIL_CENTRAL = 0, MKO_STENCIL_ONE_SIDED = 1, MKO_STENCIL_TRUNCATED = 2 };
/* JacobianMethod (include/mikroptim/optim/levenberg_marquardt.hpp:73) */
enum { MKO_JACOBIAN_FD = 0, MKO_JACOBIAN_ADJOINT = 1 };
/* LmStatus (include/mikroptim/optim/levenberg_marquardt.hpp:28) */
enum {
    MKO_LM_RUNNING = 0,
    MKO_LM_CONVERGED_MERIT = 1,
    MKO_LM_CONVERGED_STEP = 2,
    MKO_LM_CONVERGED_GRADIENT = 3,
    MKO_LM_CONVERGED_STALL = 4,
    MKO_LM_MAX_ITERATIONS = 5,
    MKO_LM_DAMPING_LIMIT = 6,
    MKO_LM_NUMERICAL_ERROR = 7
};
/* Arithmetic of the device path. FP64: every kernel in double. MIXED: forward trace (OPL,
 * residuals, merit) in double, reverse sweep of the adjoint in float, JtJ on bf16x3 MFMA with
 * float accumulation, Jtr and the LM step in double. */
enum { MKO_PRECISION_FP64 = 0, MKO_PRECISION_MIXED = 1, MKO_PRECISION_DF64 = 2,
       /* Jacobian precision of the host layer: FP64 Jacobian with its normal equations on FP64
        * MFMA (mko_gram_mfma_device) instead of the host's FMA chain (no counterpart in the
        * reference's Precision) */
       MKO_PRECISION_FP64_MFMA = 3 };

/* One batch of V variants of one topology. `type`, `pupil_mode`, `wavelength_weight` and
 * `field_weight` are batch-wide and always live in host memory; the per-variant arrays are host
 * memory for host entry points and device memory for the *_device entry points. */
typedef struct mko_batch {
    int32_t variants;    /* V */
    int32_t surfaces;    /* S, stop included, 2 ... MKO_MAX_SURFACES */
    int32_t wavelengths; /* L >= 1 */
    int32_t fields;      /* F >= 1 */
    int32_t source;      /* MKO_SOURCE_* */
    int32_t reference_sphere;       /* MKO_REF_* */
    double reference_wavelength_nm; /* lambda_ref */
    double source_point[3];         /* point source only */
    double source_na;               /* point source only */
    const uint8_t* type;               /* [S]  MKO_SURFACE_* (host) */
    const uint8_t* pupil_mode;         /* [F]  MKO_PUPIL_* (host) */
    const double* wavelength_weight;   /* [L]  (host) */
    const double* field_weight;        /* [F]  (host) */
    const double* curvature;           /* [V][S]  1/mm, 0 for stop and plane */
    const double* thickness;           /* [V][S]  mm; [v][S-1] = working distance */
    const double* index;               /* [V][S][L] medium after surface s */
    const double* pupil_radius;        /* [V] */
    const double* field_angle;         /* [V][F]  rad */
    const double* field_target;        /* [V][F]  ideal image height h_f */
    const double* reference_curvature; /* [V][F]  kappa (image references only) or NULL */
    /* [V][S] index of the medium after surface s at lambda_ref (SystemBatchView::paraxial_index),
     * or NULL; needed only by the paraxial constraint rows (focal length, paraxial focus,
     * Petzval). */
    const double* paraxial_index;
    /* [L] wavelengths in nm (host), or NULL; needed with glass variables. */
    const double* wavelength_nm;
    /* [V][S][4] Buchdahl coefficients (n0, nu1, nu2, nu3) of the medium after surface s
     * (SystemBatchHost::buchdahl_coeff, include/mikroptim/batch/system_batch.hpp), or NULL;
     * needed with glass variables. Setting a glass variable rewrites its surface's index[v][s][l]
     * and paraxial_index[v][s] as buchdahl_index (glass/dispersion.hpp:93):
     * n = n0 + w (nu1 + w (nu2 + w nu3)), w = (um - 0.546074) / (1 + 2.5 (um - 0.546074)). */
    const double* buchdahl;
} mko_batch;

/* Constraint rows appended to every variant's residual vector (ExtraResiduals /
 * TraceExtraResiduals, include/mikroptim/optim/extra_residuals.hpp:23,
 * trace_extra_residuals.hpp:38). Types and modes keep the numbering of ConstraintType /
 * ConstraintMode / EdgeHeight (include/mikroptim/model/constraints.hpp:28,76,69); supported
 * are the paraxial types (paraxial_residuals.hpp:67) and the mechanical types
 * (mechanical_constraints.hpp:38). */
enum {
    MKO_CON_FOCAL_LENGTH = 0,
    MKO_CON_PARFOCAL_LENGTH = 1,
    MKO_CON_WORKING_DISTANCE = 2,
    MKO_CON_PARAXIAL_FOCUS = 3,
    MKO_CON_PETZVAL = 4,
    MKO_CON_GLASS_HULL = 6,
    MKO_CON_MAX_SEMI_DIAMETER = 7,
    MKO_CON_MIN_CENTER_THICKNESS = 8,
    MKO_CON_MAX_CENTER_THICKNESS = 9,
    MKO_CON_MIN_EDGE_THICKNESS = 10,
    MKO_CON_MAX_INCIDENCE_ANGLE = 11,
    MKO_CON_SINE_CONDITION = 12,
    MKO_CON_DISTORTION = 13
};
enum { MKO_MODE_EQUAL = 0, MKO_MODE_MIN = 1, MKO_MODE_MAX = 2 };
enum { MKO_EDGE_SEMI_DIAMETER = 0, MKO_EDGE_RAYS = 1 };

/* One resolved constraint (ParaxialConstraint, paraxial_residuals.hpp:67; MechanicalConstraint,
 * mechanical_constraints.hpp:38). Surfaces are table (batch) surfaces.
 *   paraxial:   r = sqrt(w) * mode(q) with q = f_s (1 - f_s phi) | offset + sum_{j>=surface}
 *               t_j - L_s | t_{S-1} - w_s | a (A + t_{S-1} C / n'_L) | |kappa_P| - kappa_max
 *               | OSC(zone) (sine condition, 0 where undefined) | |D_field| - D_max
 *               (distortion, -D_max where undefined)
 *               (target = f_s, L_s, w_s, 0, kappa_max, 0, D_max; mode Min: min(q - shift, 0),
 *               Max: max(q + shift, 0)). The sine condition and distortion trace one real ray
 *               per row at lambda_ref (paraxial_index): the zone's meridional ray of the axis
 *               bundle, the field's chief ray (collimated source only).
 *   mechanical: r = sqrt(w) * max(0, g + shift), g = value - limit (max semi-diameter, max
 *               center thickness) or limit - value (others); target = limit (mm, or
 *               cos(theta_max) for the incidence angle), offset = clearance.
 *   glass hull: r = sqrt(w) * e * grad[field] (glass_hull_residual, src/optim/glass_variables.cpp:
 *               147-155) for the medium after `surface` (its Buchdahl n0, nu1, nu2 normalized
 *               by the hull's center and scale to u): e = max(0, phi), phi = signed distance
 *               of u to the hull + margin (3-d hull) or distance (lower dimension); grad the
 *               unit direction u - closest point (outside) or the nearest face's normal
 *               (inside); target = margin >= 0, field = component 0 / 1 / 2 (three rows per
 *               glass medium); the hull is mko_extra_rows.glass_hull. */
typedef struct mko_constraint {
    int32_t type;         /* MKO_CON_* */
    int32_t mode;         /* MKO_MODE_* (paraxial types) */
    double weight;        /* w >= 0 */
    double target;        /* target value or limit */
    int32_t surface;      /* parfocal reference surface / mechanical surface m(i) */
    int32_t next_surface; /* thickness types: surface m(i + 1) of the gap's far side */
    double offset;        /* parfocal offset / mechanical clearance [mm] */
    double semi_diameter, next_semi_diameter; /* edge thickness at fixed heights [mm] */
    int32_t edge_height;  /* MKO_EDGE_* (min edge thickness) */
    double shift;         /* augmented-Lagrangian shift theta >= 0 (0: plain penalty) */
    double zone;          /* sine condition: pupil zone rho in (0, 1] */
    int32_t field;        /* distortion: field f >= 0 (rows of fields >= F are inactive) */
} mko_constraint;

/* The rows of a residual vector after the ray rows: first the items that need no trace
 * extrema (E rows, in item order), then the items that do (E_t rows: max semi-diameter, max
 * incidence angle, min edge thickness at ray heights; in item order). R = trace rows + E + E_t.
 * (host memory) */
typedef struct mko_extra_rows {
    int32_t count;
    const mko_constraint* items; /* [count] */
    /* The convex hull of the catalogue glasses in normalized (n0, nu1, nu2) coordinates
     * (GlassHull, include/mikroptim/glass/glass_hull.hpp), for MKO_CON_GLASS_HULL rows, or NULL:
     * [0] dimension (0 point, 1 segment, 2 polygon, 3 polyhedron), [1] records n, [2..4] center
     * m, [5..7] scale s (u = (x - m) / s), then n records of 13 doubles: triangle vertices a, b,
     * c and (dimension 3) its plane, outward unit normal and offset (inside: normal . u <=
     * offset). Dimension 2: the polygon's fan triangles; 1: one record, the segment a - b;
     * 0: one record, the point a. glass_hull_size = 8 + 13 n. */
    const double* glass_hull;
    int32_t glass_hull_size;
} mko_extra_rows;

typedef struct mko_residual_settings {
    int32_t rings;  /* N, 1 ... 16 */
    int32_t spokes; /* M (axis fields: 1) */
    int32_t failure; /* MKO_FAILURE_* */
    double fail_weight; /* w_fail [waves^2] */
    /* number type of the RESIDUALS / MERIT traces (TraceSettings::precision): MKO_PRECISION_FP64
     * or MKO_PRECISION_DF64. The adjoint entry points take their precision as an argument. */
    int32_t precision;
    /* constraint rows after the ray rows (FdJacobianOptions::extra / trace_extra), or NULL */
    const mko_extra_rows* extra;
} mko_residual_settings;

/* P variables. Curvature k -> curvature[v][batch_surface[k]]; thickness and working distance
 * -> thickness[v][batch_surface[k]] (working distance: batch_surface = S - 1); glass k ->
 * buchdahl[v][batch_surface[k]][component[k]], the medium after the surface, components
 * n0, nu1, nu2 as three consecutive variables (resolve_variables, src/model/variables.cpp:
 * 420-432). Bounds may be +-inf; typical > 0. (host memory) */
typedef struct mko_variables {
    int32_t count;
    const int32_t* type;          /* [P] MKO_VAR_* */
    const int32_t* batch_surface; /* [P] */
    const double* lower;          /* [P] */
    const double* upper;          /* [P] */
    const double* typical;        /* [P] */
    const int32_t* component;     /* [P] glass variables' component 0 / 1 / 2, or NULL */
} mko_variables;

/* Launch shapes of an engine (mko_engine_set_tuning; mko_engine_autotune measures them): no
 * counterpart in the reference, whose tuple-pack sizes and blocks per SM are fixed
 * (include/mikroptim/cuda/tuple_packs.hpp:61, src/cuda/wavefront_block.hpp:504-526). Results do
 * not depend on them: packed traces equal unpacked ones bit for bit; the fused normal
 * equations' waves change only their rounding. */
typedef struct {
    int32_t pack_rays;   /* 0: automatic; 32, 64, 128: segment of a tuple pack (tuples of at
                          * most that many rays share a 256-lane workgroup); 256: no packing */
    int32_t fused_waves; /* 0: automatic; 1, 2, 4: waves per workgroup of the fused kernel */
    int32_t full_row_traces; /* 0: the central differences of constraint rows from trace
                              * extrema trace the candidate rays at the difference points
                              * (the same extrema, checked); 1: every ray */
} mko_tuning;

typedef struct mko_lm_options {
    int32_t max_iterations; /* 100 */
    double ftol, xtol, gtol; /* 1e-10 each */
    int32_t stall_window;   /* 0 = off */
    double stall_tolerance; /* 1e-3 */
    double initial_damping; /* 1e-3 */
    double max_damping;     /* 1e16 */
    double diag_floor;      /* 1e-12 */
    int32_t reject_new_failures; /* 1 */
    int32_t threads;        /* host threads, <= 0: hardware */
    int32_t jacobian_method; /* MKO_JACOBIAN_* */
    double fail_weight_factor; /* 10 */
    double max_fail_weight;    /* 1e14 */
    /* 0: the per-cycle control (step solves, acceptance, convergence tests) runs on the device
     * for problems of at most 127 variables, and with the fused mixed-precision adjoint (hard
     * failure, at most 96 variables, no constraint rows) each cycle is one replay of a hipGraph
     * captured per round;
     * 1: control always on the host; 2: device control launched kernel by kernel (no graph).
     * All take the same steps. No counterpart in LmOptions (the reference's control always runs
     * on the host). */
    int32_t host_control;
} mko_lm_options;

/* FP32 prefilter of multistart variants (PrefilterCriteria / PrefilterMargins,
 * include/mikroptim/batch/batch_prefilter.hpp:32,46; defaults in brackets). A variant is
 * rejected only when the criteria also hold for its FP64 values. */
typedef struct mko_prefilter_options {
    double max_failed_fraction; /* (a) failed rays > fraction * rays per variant [0.5] */
    double max_merit;           /* (b) merit above this [+inf = off] */
    double max_merit_factor;    /* (c) merit above factor * best merit [+inf = off] */
    double boundary;            /* eta: margin under which a ray's decision is uncertain [0.15] */
    double opl_error_factor;    /* kappa of the OPD error model [32] */
    double merit_relative;      /* m of the OPD error model, in [0, 1) [0.01] */
} mko_prefilter_options;

/* PrefilterReason bits (batch_prefilter.hpp:87) */
enum {
    MKO_PREFILTER_KEEP = 0,
    MKO_PREFILTER_FAILED_RAYS = 1,
    MKO_PREFILTER_MERIT_ABSOLUTE = 2,
    MKO_PREFILTER_MERIT_RELATIVE = 4
};

/* PrefilterStats + PrefilterResult (batch_prefilter.hpp:61,96): caller-owned arrays [V]. */
typedef struct mko_prefilter_result {
    double* merit;            /* F32: sum r^2 over the certain-valid rays (float trace) */
    double* optical_length;   /* largest optical length of a certain-valid ray [mm] */
    int64_t* failed;          /* certainly failed rays */
    int64_t* uncertain_ok;    /* valid in float, near the boundary */
    int64_t* uncertain_failed; /* failed in float, near the boundary */
    double* merit_lower;      /* bounds of the FP64 merit (with extra_merit) */
    double* merit_upper;
    uint8_t* reject;          /* MKO_PREFILTER_* bits, 0 = keep */
    double reference_merit;   /* upper bound of the best merit for (c), +inf if none */
    int32_t reference_variant; /* -1 if none */
    int32_t kept;
} mko_prefilter_result;

/* Per-run LM result (LmRunResult, include/mikroptim/optim/levenberg_marquardt.hpp:130). */
typedef struct mko_lm_run {
    double merit;
    int64_t failed_rays;
    int64_t hard_failed_rays;
    int32_t status; /* MKO_LM_* */
    int32_t iterations;
    int32_t accepted;
    int32_t jacobians;
    double damping;
    double gradient_norm;
    double model_decrease;
    double fail_weight;
    int32_t penalty_rounds;
} mko_lm_run;

#ifdef __cplusplus
}
#endif

#endif

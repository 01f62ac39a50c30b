"""Host-side system batch, residual layout and variable map.

`SystemBatch` mirrors SystemBatchHost (include/mikroptim/batch/system_batch.hpp:65): V variants of
one topology as structure-of-arrays in FP64, variant outermost, with the index formulas of
BatchShape (include/mikroptim/batch/system_batch_view.hpp:29). `VariableSet` mirrors the
curvature / thickness / working-distance part of VariableSet
(include/mikroptim/optim/variable_set.hpp:62, src/optim/variable_set.cpp:237).
"""

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi


class SystemBatch:
    def __init__(self, types, wavelengths_nm, wavelength_weights, field_weights, pupil_modes,
                 variants, reference_wavelength_nm, source=abi.SOURCE_COLLIMATED,
                 reference_sphere=abi.REF_LAST_SURFACE):
        self.type = np.ascontiguousarray(types, dtype=np.uint8)
        S = len(self.type)
        if not 2 <= S <= abi.MAX_SURFACES:
            raise ValueError(f"{S} surfaces, allowed 2 ... {abi.MAX_SURFACES}")
        self.wavelength_nm = np.ascontiguousarray(wavelengths_nm, dtype=np.float64)
        self.wavelength_weight = np.ascontiguousarray(wavelength_weights, dtype=np.float64)
        self.field_weight = np.ascontiguousarray(field_weights, dtype=np.float64)
        self.pupil_mode = np.ascontiguousarray(pupil_modes, dtype=np.uint8)
        L, F = len(self.wavelength_nm), len(self.field_weight)
        if len(self.wavelength_weight) != L or len(self.pupil_mode) != F:
            raise ValueError("wavelength weights / pupil modes do not match")
        self.reference_wavelength_nm = float(reference_wavelength_nm)
        self.source = int(source)
        self.reference_sphere = int(reference_sphere)
        self.source_point = np.zeros(3)
        self.source_na = 0.0
        V = int(variants)
        self.curvature = np.zeros((V, S))
        self.thickness = np.zeros((V, S))
        self.index = np.ones((V, S, L))
        self.pupil_radius = np.zeros(V)
        self.field_angle = np.zeros((V, F))
        self.field_target = np.zeros((V, F))
        self.reference_curvature = (np.zeros((V, F)) if reference_sphere != abi.REF_LAST_SURFACE
                                    else None)
        # [V][S] medium index after each surface at lambda_ref (paraxial constraint rows)
        self.paraxial_index = None
        # [V][S][4] Buchdahl coefficients (n0, nu1, nu2, nu3) of the medium after each surface,
        # needed with glass variables
        self.buchdahl = None

    # ---- shape --------------------------------------------------------------------------
    @property
    def variants(self):
        return self.curvature.shape[0]

    @property
    def surfaces(self):
        return len(self.type)

    @property
    def wavelengths(self):
        return len(self.wavelength_nm)

    @property
    def fields(self):
        return len(self.field_weight)

    PER_VARIANT = ("curvature", "thickness", "index", "pupil_radius", "field_angle",
                   "field_target", "reference_curvature", "paraxial_index", "buchdahl")

    def take(self, rows):
        """New batch with variants `rows` of this one (shared fields copied)."""
        out = SystemBatch.__new__(SystemBatch)
        out.__dict__.update({k: v for k, v in self.__dict__.items()})
        rows = np.asarray(rows, dtype=np.int64)
        for name in self.PER_VARIANT:
            a = getattr(self, name)
            setattr(out, name, None if a is None else np.ascontiguousarray(a[rows]))
        out.source_point = self.source_point.copy()
        return out

    def copy(self):
        return self.take(np.arange(self.variants))

    # ---- C view ---------------------------------------------------------------------------
    def c_struct(self, arrays=None):
        """mko_batch over this batch's host arrays (or over `arrays`, a dict of per-variant
        pointers, e.g. device addresses as ints)."""
        for name in self.PER_VARIANT:
            a = getattr(self, name)
            if a is not None and not (a.flags.c_contiguous and a.dtype == np.float64):
                setattr(self, name, np.ascontiguousarray(a, dtype=np.float64))
        b = abi.Batch()
        b.variants = self.variants
        b.surfaces = self.surfaces
        b.wavelengths = self.wavelengths
        b.fields = self.fields
        b.source = self.source
        b.reference_sphere = self.reference_sphere
        b.reference_wavelength_nm = self.reference_wavelength_nm
        for i in range(3):
            b.source_point[i] = float(self.source_point[i])
        b.source_na = float(self.source_na)
        b.type = abi.ptr(self.type, C.c_uint8)
        b.pupil_mode = abi.ptr(self.pupil_mode, C.c_uint8)
        b.wavelength_weight = abi.ptr(self.wavelength_weight)
        b.field_weight = abi.ptr(self.field_weight)
        b.wavelength_nm = abi.ptr(self.wavelength_nm)
        for name in self.PER_VARIANT:
            if arrays is not None:
                value = arrays.get(name)
                setattr(b, name, None if value is None else C.cast(C.c_void_p(value),
                                                                   C.POINTER(C.c_double)))
            else:
                setattr(b, name, abi.ptr(getattr(self, name)))
        return b


@dataclass
class Constraint:
    """One constraint row (ParaxialConstraint, include/mikroptim/optim/paraxial_residuals.hpp:67;
    MechanicalConstraint, include/mikroptim/optim/mechanical_constraints.hpp:38); surfaces are
    batch (table) surfaces, `target` the target value or limit, `offset` the parfocal offset or
    the mechanical clearance."""
    type: int
    weight: float = 1.0
    target: float = 0.0
    mode: int = abi.MODE_EQUAL
    surface: int = 0
    next_surface: int = 0
    offset: float = 0.0
    semi_diameter: float = 0.0
    next_semi_diameter: float = 0.0
    edge_height: int = abi.EDGE_SEMI_DIAMETER
    shift: float = 0.0
    zone: float = 0.0   # sine condition: pupil zone rho in (0, 1]
    field: int = 0      # distortion: field index; glass hull: component 0 / 1 / 2

    @property
    def needs_extrema(self):
        return self.type in (abi.CON_MAX_SEMI_DIAMETER, abi.CON_MAX_INCIDENCE_ANGLE) or (
            self.type == abi.CON_MIN_EDGE_THICKNESS and self.edge_height == abi.EDGE_RAYS)

    def c_struct(self):
        return abi.Constraint(self.type, self.mode, self.weight, self.target, self.surface,
                              self.next_surface, self.offset, self.semi_diameter,
                              self.next_semi_diameter, self.edge_height, self.shift, self.zone,
                              self.field)


def focal_length(target, weight=1.0):
    return Constraint(abi.CON_FOCAL_LENGTH, weight, target)


def working_distance(target, weight=1.0, mode=abi.MODE_EQUAL, shift=0.0):
    return Constraint(abi.CON_WORKING_DISTANCE, weight, target, mode, shift=shift)


def parfocal_length(target, surface=0, offset=0.0, weight=1.0, mode=abi.MODE_EQUAL):
    return Constraint(abi.CON_PARFOCAL_LENGTH, weight, target, mode, surface=surface,
                      offset=offset)


def paraxial_focus(weight=1.0):
    return Constraint(abi.CON_PARAXIAL_FOCUS, weight, 0.0)


def petzval(kappa_max, weight=1.0, shift=0.0):
    return Constraint(abi.CON_PETZVAL, weight, kappa_max, abi.MODE_MAX, shift=shift)


def sine_condition(zone, weight=1.0):
    """OSC (Conrady) of the meridional ray of the axis bundle at pupil zone rho = zone, target 0
    (one row per zone; make_paraxial_constraints uses the zones 1/sqrt(2) and 1 by default)."""
    return Constraint(abi.CON_SINE_CONDITION, weight, 0.0, zone=zone)


def distortion(max_distortion, field, weight=1.0, shift=0.0):
    """max(0, |D_f| - D_max) of field f, D_f = (y_f - h_f) / h_f with y_f the real chief ray's
    height in the image plane (one row per field of the source)."""
    return Constraint(abi.CON_DISTORTION, weight, max_distortion, abi.MODE_MAX, field=field,
                      shift=shift)


def max_semi_diameter(surface, limit, weight=1.0, clearance=0.0, shift=0.0):
    return Constraint(abi.CON_MAX_SEMI_DIAMETER, weight, limit, surface=surface,
                      next_surface=surface, offset=clearance, shift=shift)


def center_thickness(surface, next_surface, limit, weight=1.0, upper=False, shift=0.0):
    return Constraint(abi.CON_MAX_CENTER_THICKNESS if upper else abi.CON_MIN_CENTER_THICKNESS,
                      weight, limit, surface=surface, next_surface=next_surface, shift=shift)


def min_edge_thickness(surface, next_surface, limit, weight=1.0, clearance=0.0,
                       semi_diameters=None, shift=0.0):
    """Edge thickness at the larger of two fixed semi-diameters, or (None) at the larger of the
    two surfaces' largest ray heights."""
    c = Constraint(abi.CON_MIN_EDGE_THICKNESS, weight, limit, surface=surface,
                   next_surface=next_surface, offset=clearance, shift=shift,
                   edge_height=abi.EDGE_RAYS if semi_diameters is None else abi.EDGE_SEMI_DIAMETER)
    if semi_diameters is not None:
        c.semi_diameter, c.next_semi_diameter = map(float, semi_diameters)
    return c


def glass_hull(surface, weight=1e4, margin=0.0):
    """The three rows of the glass-hull penalty of the medium after `surface` (components n0,
    nu1, nu2 of sqrt(w) e grad; glass_hull_residual, src/optim/glass_variables.cpp:147-155);
    the hull itself is ResidualSettings.glass_hull. Default weight kDefaultGlassHullWeight."""
    return [Constraint(abi.CON_GLASS_HULL, weight, margin, surface=surface, field=k)
            for k in range(3)]


def max_incidence_angle(surface, angle_deg, weight=1.0, shift=0.0):
    """limit = cos(theta_max) (make_mechanical_constraints)."""
    limit = math.cos(angle_deg * 3.14159265358979323846 / 180.0)
    return Constraint(abi.CON_MAX_INCIDENCE_ANGLE, weight, limit, surface=surface,
                      next_surface=surface, shift=shift)


@dataclass
class ResidualSettings:
    """ResidualSettings, include/mikroptim/batch/batch_residuals.hpp:24 (reference defaults),
    plus the constraint rows appended to every residual vector (FdJacobianOptions::extra /
    trace_extra)."""
    rings: int = 10
    spokes: int = 8
    failure: int = abi.FAILURE_SOFT
    fail_weight: float = 1e6
    precision: int = abi.PRECISION_FP64  # number type of the RESIDUALS / MERIT traces
    constraints: list = field(default_factory=list)
    # the hull of the glass-hull rows (mko_extra_rows.glass_hull; glass_hull.GlassHull.packed())
    glass_hull: object = None

    def row_constraints(self):
        """Constraints in row order: those without extrema, then those with."""
        return ([c for c in self.constraints if not c.needs_extrema]
                + [c for c in self.constraints if c.needs_extrema])

    @property
    def extra_rows(self):
        return len(self.constraints)

    def c_struct(self):
        s = abi.ResidualSettings(self.rings, self.spokes, self.failure, self.fail_weight,
                                 self.precision, None)
        if self.constraints:
            items = (abi.Constraint * len(self.constraints))(
                *[c.c_struct() for c in self.constraints])
            hull = None
            if self.glass_hull is not None:
                hull = np.ascontiguousarray(self.glass_hull, dtype=np.float64)
            rows = abi.ExtraRows(len(self.constraints), items,
                                 abi.ptr(hull) if hull is not None else None,
                                 0 if hull is None else hull.size)
            s.extra = C.pointer(rows)
            s._keep = (items, rows, hull)
        return s


def layout(batch, settings):
    """ResidualLayout, src/batch/batch_residuals.cpp:25-57: per field (offset, samples);
    offset(f, l) = field_offset[f] + l * samples[f]."""
    offsets, samples = [], []
    off = 0
    for f in range(batch.fields):
        n = settings.rings * (1 if batch.pupil_mode[f] == abi.PUPIL_AXIS else settings.spokes)
        offsets.append(off)
        samples.append(n)
        off += batch.wavelengths * n
    return offsets, samples, off


def rays_per_variant(batch, settings):
    return layout(batch, settings)[2]


def trace_rows(batch, settings):
    rays = rays_per_variant(batch, settings)
    return rays * (2 if settings.failure == abi.FAILURE_SOFT else 1)


def residual_rows(batch, settings):
    """Rows of a variant's residual vector: trace rows + E + E_t."""
    return trace_rows(batch, settings) + settings.extra_rows


BUCHDAHL_LAMBDA0_UM = 0.546074
BUCHDAHL_ALPHA = 2.5


def buchdahl_omega(lambda_nm):
    """buchdahl_omega of nm_to_um(lambda) (include/mikroptim/glass/dispersion.hpp:18-89)."""
    um = lambda_nm / 1000.0
    den = 1.0 + BUCHDAHL_ALPHA * (um - BUCHDAHL_LAMBDA0_UM)
    if not den > 0.0:
        raise ValueError(f"Buchdahl coordinate undefined at {lambda_nm} nm (needs > 146.074 nm)")
    return (um - BUCHDAHL_LAMBDA0_UM) / den


def buchdahl_index(coeff, lambda_nm):
    """n = n0 + w (nu1 + w (nu2 + w nu3)) (buchdahl_index, dispersion.hpp:93)."""
    w = buchdahl_omega(lambda_nm)
    return coeff[0] + w * (coeff[1] + w * (coeff[2] + w * coeff[3]))


# typical sizes of the glass components (kDefaultTypicalGlass*, model/variables.hpp:77-79)
GLASS_TYPICAL = (1.0, 0.01, 0.001)


@dataclass
class VariableSet:
    """Variables x <-> batch elements (curvature[v][s], thickness[v][s]; glass: the Buchdahl
    coefficient buchdahl[v][s][component], which also sets the surface's index and paraxial
    index, VariableSet::apply, src/optim/variable_set.cpp:237-285)."""
    type: np.ndarray
    batch_surface: np.ndarray
    lower: np.ndarray
    upper: np.ndarray
    typical: np.ndarray
    names: list = field(default_factory=list)
    component: np.ndarray = None  # [P] glass component 0 / 1 / 2 (None without glass)

    @property
    def size(self):
        return len(self.type)

    @property
    def has_glass(self):
        return bool(np.any(np.asarray(self.type) == abi.VAR_GLASS))

    def c_struct(self):
        self.type = np.ascontiguousarray(self.type, dtype=np.int32)
        self.batch_surface = np.ascontiguousarray(self.batch_surface, dtype=np.int32)
        self.lower = np.ascontiguousarray(self.lower, dtype=np.float64)
        self.upper = np.ascontiguousarray(self.upper, dtype=np.float64)
        self.typical = np.ascontiguousarray(self.typical, dtype=np.float64)
        comp = None
        if self.component is not None:
            self.component = np.ascontiguousarray(self.component, dtype=np.int32)
            comp = abi.ptr(self.component, C.c_int32)
        return abi.Variables(self.size, abi.ptr(self.type, C.c_int32),
                             abi.ptr(self.batch_surface, C.c_int32), abi.ptr(self.lower),
                             abi.ptr(self.upper), abi.ptr(self.typical), comp)

    def _column(self, batch, k):
        t, s = int(self.type[k]), int(self.batch_surface[k])
        if t == abi.VAR_GLASS:
            if batch.buchdahl is None:
                raise ValueError("glass variables need the batch's Buchdahl coefficients")
            return batch.buchdahl[:, s, int(self.component[k])]
        return (batch.curvature if t == abi.VAR_CURVATURE else batch.thickness)[:, s]

    def extract(self, batch, v=None):
        """x of variant v (all variants if v is None): VariableSet::extract."""
        cols = [self._column(batch, k) for k in range(self.size)]
        x = np.stack(cols, axis=-1)
        return x if v is None else x[v]

    def glass_surfaces(self):
        return sorted({int(s) for t, s in zip(self.type, self.batch_surface)
                       if t == abi.VAR_GLASS})

    def apply(self, batch, x, variants=None):
        """Writes x ([P] or [n, P]) into the given variants (all if None): VariableSet::apply;
        glass surfaces then get index[v][s][l] and paraxial_index[v][s] from their Buchdahl
        coefficients."""
        x = np.atleast_2d(np.asarray(x, dtype=np.float64))
        rows = np.arange(batch.variants) if variants is None else np.asarray(variants)
        if not np.all(np.isfinite(x)):
            raise ValueError("variable value not finite")
        glass = self.glass_surfaces()
        if glass:
            for nm in list(batch.wavelength_nm) + [batch.reference_wavelength_nm]:
                buchdahl_omega(float(nm))
        for k, (t, s) in enumerate(zip(self.type, self.batch_surface)):
            if t == abi.VAR_GLASS:
                batch.buchdahl[rows, s, int(self.component[k])] = x[:, k]
            else:
                dst = batch.curvature if t == abi.VAR_CURVATURE else batch.thickness
                dst[rows, s] = x[:, k]
        for s in glass:
            for v in np.atleast_1d(rows):
                c = [float(z) for z in batch.buchdahl[v, s]]
                for l, nm in enumerate(batch.wavelength_nm):
                    batch.index[v, s, l] = buchdahl_index(c, float(nm))
                if batch.paraxial_index is not None:
                    batch.paraxial_index[v, s] = buchdahl_index(c, batch.reference_wavelength_nm)

    def clamp(self, x):
        return np.minimum(np.maximum(x, self.lower), self.upper)

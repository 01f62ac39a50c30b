"""Host-side system batch, residual layout and variable map.

`SystemBatch` mirrors SystemBatchHost (include/mikroptim/batch/system_batch.hpp:65): V variants of
one topology as structure-of-arrays in FP64, variant outermost, with the index formulas of
BatchShape (include/mikroptim/batch/system_batch_view.hpp:29). `VariableSet` mirrors the
curvature / thickness / working-distance part of VariableSet
(include/mikroptim/optim/variable_set.hpp:62, src/optim/variable_set.cpp:237).
"""

import ctypes as C
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
                   "field_target", "reference_curvature")

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
        for name in self.PER_VARIANT:
            if arrays is not None:
                value = arrays.get(name)
                setattr(b, name, None if value is None else C.cast(C.c_void_p(value),
                                                                   C.POINTER(C.c_double)))
            else:
                setattr(b, name, abi.ptr(getattr(self, name)))
        return b


@dataclass
class ResidualSettings:
    """ResidualSettings, include/mikroptim/batch/batch_residuals.hpp:24 (reference defaults)."""
    rings: int = 10
    spokes: int = 8
    failure: int = abi.FAILURE_SOFT
    fail_weight: float = 1e6
    precision: int = abi.PRECISION_FP64  # number type of the RESIDUALS / MERIT traces

    def c_struct(self):
        return abi.ResidualSettings(self.rings, self.spokes, self.failure, self.fail_weight,
                                    self.precision)


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


@dataclass
class VariableSet:
    """Variables x <-> batch elements (curvature[v][s], thickness[v][s])."""
    type: np.ndarray
    batch_surface: np.ndarray
    lower: np.ndarray
    upper: np.ndarray
    typical: np.ndarray
    names: list = field(default_factory=list)

    @property
    def size(self):
        return len(self.type)

    def c_struct(self):
        self.type = np.ascontiguousarray(self.type, dtype=np.int32)
        self.batch_surface = np.ascontiguousarray(self.batch_surface, dtype=np.int32)
        self.lower = np.ascontiguousarray(self.lower, dtype=np.float64)
        self.upper = np.ascontiguousarray(self.upper, dtype=np.float64)
        self.typical = np.ascontiguousarray(self.typical, dtype=np.float64)
        return abi.Variables(self.size, abi.ptr(self.type, C.c_int32),
                             abi.ptr(self.batch_surface, C.c_int32), abi.ptr(self.lower),
                             abi.ptr(self.upper), abi.ptr(self.typical))

    def extract(self, batch, v=None):
        """x of variant v (all variants if v is None): VariableSet::extract."""
        rows = slice(None) if v is None else v
        cols = []
        for t, s in zip(self.type, self.batch_surface):
            src = batch.curvature if t == abi.VAR_CURVATURE else batch.thickness
            cols.append(src[rows, s])
        return np.stack(cols, axis=-1)

    def apply(self, batch, x, variants=None):
        """Writes x ([P] or [n, P]) into the given variants (all if None): VariableSet::apply."""
        x = np.atleast_2d(np.asarray(x, dtype=np.float64))
        rows = np.arange(batch.variants) if variants is None else np.asarray(variants)
        if not np.all(np.isfinite(x)):
            raise ValueError("variable value not finite")
        for k, (t, s) in enumerate(zip(self.type, self.batch_surface)):
            dst = batch.curvature if t == abi.VAR_CURVATURE else batch.thickness
            dst[rows, s] = x[:, k]

    def clamp(self, x):
        return np.minimum(np.maximum(x, self.lower), self.upper)

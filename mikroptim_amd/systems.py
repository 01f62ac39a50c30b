"""Synthetic systems and variable sets of the measured workloads (built in code from a seed; the
reference ships none of these systems; reference system files go through system_file.py).

singlet(...)        stop + biconvex N-BK7 singlet (c1 = +0.02, c2 = -0.005 1/mm, t = 5 mm),
                    EPD 10 mm, fields 0 / 3.5 / 5 deg, 486.13 / 587.56 / 656.27 nm, image plane at
                    the paraxial focus, field targets f tan(theta); curvatures perturbed
                    uniformly by +-5 % from a seed (variant 0 unperturbed).
objective(...)      24 refracting spherical surfaces (12 elements) behind the stop, collimated
                    input, image-space NA 0.95 in air, 9 synthetic Sellmeier glasses
                    nd 1.45 ... 1.90, 5 fields, 7 wavelengths 405 ... 700 nm; 48 variables
                    (24 curvatures, 23 inner thicknesses, working distance), parameters jittered
                    by N(0, 1 %) (Irwin-Hall) from a seed.
The layout of the objective: a converging relay of ten thin positive elements (beam focused at
sin u = 0.263), then two aplanatic-concentric menisci of index 1.90 that multiply sin u by n each
(0.263 -> 0.5 -> 0.95): in trace direction, each meniscus' first surface is aplanatic for the point
its exit beam converges to, and its second surface is concentric about that point. The pupil radius
follows from the sine condition (a = NA f = 5.47 mm), the working distance is 0.6 mm, the field
reaches 0.3 deg. With that, none of the 4096 x 17920 rays of the seeded batch fails (hard mode),
checked with the oracle.
"""

import functools
import math

import numpy as np

from . import abi, optics, rng
from . import batch as batch_mod
from .batch import ResidualSettings, SystemBatch, VariableSet

LAMBDA_REF = 587.56
F_C_D = (486.13, 587.56, 656.27)


def _index_table(glasses, wavelengths_nm):
    """index[s][l] for a list of media per surface (None = air, else (B, C))."""
    out = np.ones((len(glasses), len(wavelengths_nm)))
    for s, g in enumerate(glasses):
        if g is None:
            continue
        for l, nm in enumerate(wavelengths_nm):
            out[s, l] = optics.sellmeier_index(g[0], g[1], nm / 1000.0)
    return out


def _paraxial(c, t, glasses):
    n = [1.0 if g is None else optics.sellmeier_index(g[0], g[1], LAMBDA_REF / 1000.0)
         for g in glasses]
    return optics.focal_and_bfd(c, t, n)


def singlet(variants=1024, seed=1, perturb=0.05):
    """The singlet workload (one variant unperturbed, the rest curvatures * (1 + perturb xi))."""
    nbk7 = (optics.NBK7_B, optics.NBK7_C)
    glasses = [None, nbk7, None]
    c = [0.0, 0.02, -0.005]
    t = [5.0, 5.0, 0.0]
    f, bfd = _paraxial(c, t, glasses)
    t[2] = bfd
    fields_deg = (0.0, 3.5, 5.0)
    b = SystemBatch([abi.SURFACE_STOP, abi.SURFACE_SPHERE, abi.SURFACE_SPHERE], F_C_D,
                    [1.0, 1.0, 1.0], [1.0, 1.0, 1.0],
                    [abi.PUPIL_AXIS if a == 0.0 else abi.PUPIL_HALF for a in fields_deg],
                    variants, LAMBDA_REF)
    b.curvature[:] = c
    b.thickness[:] = t
    b.index[:] = _index_table(glasses, F_C_D)
    b.paraxial_index = np.tile(_index_table(glasses, [LAMBDA_REF])[:, 0], (variants, 1))
    b.pupil_radius[:] = 5.0
    theta = np.radians(fields_deg)
    b.field_angle[:] = theta
    b.field_target[:] = f * np.tan(theta)
    variables = singlet_variables()
    x0 = variables.extract(b, 0)
    xs = rng.perturbed_vectors(variables, x0, variants, seed=seed,
                               relative=(perturb, 0.0, 0.0), first_is_base=True)
    variables.apply(b, xs)
    return b, variables


def concentric(a, gaps, n, image_distance, fields_deg=(0.0,), variants=1):
    """A system concentric to its stop (the sine-condition and distortion oracle of the
    reference's tests, tests/test_sine_condition.cpp:58-81): the stop (radius a, in the medium
    n[0]) and spheres at distances z_k = gaps[0] + ... + gaps[k-1] with their centres in the
    stop's centre (c_k = -1/z_k), medium n[k] behind sphere k; the last sphere's thickness is
    image_distance. One wavelength (587.56 nm, constant indices); field targets 0."""
    S = len(n)
    fields_deg = tuple(fields_deg)
    b = SystemBatch([abi.SURFACE_STOP] + [abi.SURFACE_SPHERE] * (S - 1), [587.56], [1.0],
                    [1.0] * len(fields_deg),
                    [abi.PUPIL_AXIS if f == 0.0 else abi.PUPIL_HALF for f in fields_deg],
                    variants, 587.56)
    z = 0.0
    c, t = [0.0], [gaps[0]]
    for k in range(1, S):
        z += gaps[k - 1]
        c.append(-1.0 / z)
        t.append(gaps[k] if k < len(gaps) else image_distance)
    b.curvature[:] = c
    b.thickness[:] = t
    b.index[:, :, 0] = n
    b.paraxial_index = np.tile(np.asarray(n, dtype=np.float64), (variants, 1))
    b.pupil_radius[:] = a
    b.field_angle[:] = np.radians(fields_deg)
    return b


def singlet_multistart(variants=1000, seed=11):
    """Multistart variants of the singlet for the prefilter: curvatures x(1 + 0.5 xi), lens
    thickness and working distance x(1 + 0.3 xi) (xi uniform, perturbed_vector seeding), pupil
    radius x 2^eta (eta uniform in [-1, 1]) and every seventh variant extreme (pupil x 4 ... 12),
    so that the batch holds failed rays, near-boundary rays and merits over many decades."""
    b, v = singlet(variants, seed=seed, perturb=0.0)
    x0 = v.extract(b, 0)
    xs = rng.perturbed_vectors(v, x0, variants, seed=seed, relative=(0.5, 0.3, 0.3),
                               first_is_base=True)
    xs = v.clamp(xs)
    v.apply(b, xs)
    g = np.random.default_rng(seed)
    scale = 2.0 ** g.uniform(-1.0, 1.0, variants)
    scale[::7] = g.uniform(4.0, 12.0, len(scale[::7]))
    scale[0] = 1.0
    b.pupil_radius[:] = b.pupil_radius * scale
    return b, v


def singlet_variables():
    """Both curvatures, the lens thickness and the working distance."""
    inf = math.inf
    return VariableSet(
        type=np.array([abi.VAR_CURVATURE, abi.VAR_CURVATURE, abi.VAR_THICKNESS,
                       abi.VAR_WORKING_DISTANCE], dtype=np.int32),
        batch_surface=np.array([1, 2, 1, 2], dtype=np.int32),
        lower=np.array([-0.1, -0.1, 1.0, 10.0]), upper=np.array([0.1, 0.1, 20.0, 500.0]),
        typical=np.array([0.01, 0.01, 1.0, 1.0]), names=["c[1]", "c[2]", "t[1]", "w"])


def singlet_settings(failure=abi.FAILURE_HARD):
    return ResidualSettings(rings=6, spokes=12, failure=failure)


def singlet_image_constraints(fields=3, max_distortion=0.001):
    """Sine condition at the zones 1/sqrt(2) and 1 (make_paraxial_constraints' default zones)
    and the distortion of every field (one row per field, the axis row inactive)."""
    return ([batch_mod.sine_condition(0.70710678118654752440), batch_mod.sine_condition(1.0)]
            + [batch_mod.distortion(max_distortion, f) for f in range(fields)])


def singlet_constraints():
    """Constraint rows of every kind on the singlet: focal length 100 mm, working distance >= 95,
    parfocal length, paraxial focus, Petzval bound, semi-diameter of the first lens surface,
    center and edge thickness of the lens (fixed heights and ray heights), incidence angle at
    the second lens surface."""
    m = batch_mod
    return [m.focal_length(100.0, 0.03), m.working_distance(95.0, 0.1, abi.MODE_MIN),
            m.parfocal_length(150.0, 1, 2.0, 0.1), m.paraxial_focus(0.5),
            m.petzval(0.001, 100.0),
            m.max_semi_diameter(1, 4.0, 2.0, clearance=0.5),
            m.center_thickness(1, 2, 6.0, 1.0),
            m.min_edge_thickness(1, 2, 2.0, 1.0, semi_diameters=(6.0, 6.0)),
            m.max_incidence_angle(2, 5.0, 3.0),
            m.min_edge_thickness(1, 2, 4.7, 1.0, clearance=0.2)]


# ---- the 24-surface objective ------------------------------------------------------------------

OBJECTIVE_GLASSES = [  # (nd, Vd) of the nine synthetic glasses
    (1.45, 68.0), (1.50, 64.0), (1.55, 58.0), (1.60, 52.0), (1.65, 46.0), (1.70, 42.0),
    (1.75, 36.0), (1.80, 30.0), (1.90, 40.0)]

OBJECTIVE_NA = 0.95
OBJECTIVE_PUPIL_RADIUS = 3.125
OBJECTIVE_WD = 0.6


@functools.lru_cache(maxsize=None)
def _glass(i):
    return optics.synthetic_glass(*OBJECTIVE_GLASSES[i])


def _meniscus_chain(n, stages=2, working_distance=OBJECTIVE_WD, r_front=1.5, gaps=(0.8, 1.0)):
    """Aplanatic-concentric menisci of index n, from the image point backwards (trace-direction
    coordinates). Meniscus k: first surface radius r1 (center on +z); a beam converging to the
    far aplanatic conjugate P (distance r1 (1 + n) from the vertex) converges after it to
    O (distance r1 (1 + 1/n)); the second surface is centered on O. Returns the per-meniscus
    (r1, thickness, r2, gap_after) upstream-first and the distance from the first vertex to the
    point the upstream optics must focus on."""
    out = []
    r2 = working_distance
    gap_after = working_distance
    r1 = r_front
    for k in range(stages):
        if k > 0:
            r1 = max(r1 * 2.2, (r2 + 1.0) / (1.0 + 1.0 / n))
        thick = r1 * (1.0 + 1.0 / n) - r2
        out.append((r1, thick, r2, gap_after))
        p_dist = r1 * (1.0 + n)
        gap_after = gaps[min(k, len(gaps) - 1)]
        r2 = p_dist + gap_after
    out.reverse()
    return out, p_dist


def _relay(scale, glasses, lens_t=1.6, air_t=0.4):
    """Ten thin positive elements of equal power `scale` / 10, alternating bending."""
    curv, thick, media = [], [], []
    for i, g in enumerate(glasses):
        gl = _glass(g)
        n = optics.sellmeier_index(*gl, LAMBDA_REF / 1000.0)
        d = scale / len(glasses) / (n - 1.0)
        shape = 0.8 if i % 2 == 0 else 0.5
        curv += [0.5 * d * (1.0 + shape), -0.5 * d * (1.0 - shape)]
        thick += [lens_t, air_t]
        media += [gl, None]
    return curv, thick, media


RELAY_GLASSES = (0, 1, 2, 3, 4, 5, 6, 7, 0, 1)


def objective_design():
    """Nominal 24-surface system: (types, curvature, thickness, media, pupil radius). The relay
    power is solved paraxially at lambda_ref so that the relay focuses on the far aplanatic
    conjugate of the first meniscus (RELAY_GAP behind its last vertex); the pupil radius is then
    a = NA * f (sine condition), which makes the image-space NA 0.95 at lambda_ref."""
    n_front = optics.sellmeier_index(*_glass(8), LAMBDA_REF / 1000.0)
    menisci, p_dist = _meniscus_chain(n_front)

    def build(scale):
        c_r, t_r, m_r = _relay(scale, RELAY_GLASSES)
        types = [abi.SURFACE_STOP] + [abi.SURFACE_SPHERE] * (len(c_r) + 2 * len(menisci))
        curv = [0.0] + c_r
        thick = [2.0] + t_r
        media = [None] + m_r
        thick[-1] = RELAY_GAP
        for r1, t_m, r2, gap_after in menisci:
            curv += [1.0 / r1, 1.0 / r2]
            thick += [t_m, gap_after]
            media += [_glass(8), None]
        return types, curv, thick, media

    def relay_bfd(scale):
        c_r, t_r, m_r = _relay(scale, RELAY_GLASSES)
        t_r[-1] = 0.0
        return _paraxial([0.0] + c_r, [2.0] + t_r, [None] + m_r)[1]

    lo, hi = 1e-3, 1.0  # more power -> shorter back focal distance
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if relay_bfd(mid) > p_dist + RELAY_GAP:
            lo = mid
        else:
            hi = mid
    types, curv, thick, media = build(0.5 * (lo + hi))
    f, _ = _paraxial(curv, thick, media)
    return types, np.array(curv), np.array(thick), media, OBJECTIVE_NA * f


RELAY_GAP = 1.5


def objective(variants=4096, seed=2, jitter=0.01, fields=5, wavelengths=7, max_field_deg=0.3):
    types, c, t, media, a = objective_design()
    wl = np.linspace(405.0, 700.0, wavelengths)
    fields_deg = np.linspace(0.0, max_field_deg, fields)
    b = SystemBatch(types, wl, np.ones(wavelengths), np.ones(fields),
                    [abi.PUPIL_HALF] * fields, variants, LAMBDA_REF)
    f, _ = _paraxial(list(c), list(t), media)
    b.curvature[:] = c
    b.thickness[:] = t
    b.index[:] = _index_table(media, wl)
    b.paraxial_index = np.tile(_index_table(media, [LAMBDA_REF])[:, 0], (variants, 1))
    b.pupil_radius[:] = a
    theta = np.radians(fields_deg)
    b.field_angle[:] = theta
    b.field_target[:] = f * np.tan(theta)
    variables = objective_variables(b)
    x0 = variables.extract(b, 0)
    xs = rng.perturbed_vectors(variables, x0, variants, seed=seed,
                               relative=(jitter, jitter, jitter), distribution="irwin_hall",
                               first_is_base=True)
    variables.apply(b, xs)
    return b, variables


def objective_variables(batch):
    S = batch.surfaces
    types, surf, lo, hi, typ, names = [], [], [], [], [], []
    for s in range(1, S):
        types.append(abi.VAR_CURVATURE)
        surf.append(s)
        lo.append(-math.inf)
        hi.append(math.inf)
        typ.append(0.01)
        names.append(f"c[{s}]")
    for s in range(1, S - 1):
        types.append(abi.VAR_THICKNESS)
        surf.append(s)
        lo.append(0.05)
        hi.append(50.0)
        typ.append(1.0)
        names.append(f"t[{s}]")
    types.append(abi.VAR_WORKING_DISTANCE)
    surf.append(S - 1)
    lo.append(0.01)
    hi.append(10.0)
    typ.append(1.0)
    names.append("w")
    return VariableSet(np.array(types, np.int32), np.array(surf, np.int32), np.array(lo),
                       np.array(hi), np.array(typ), names)


def objective_settings(failure=abi.FAILURE_HARD):
    return ResidualSettings(rings=16, spokes=32, failure=failure)


def objective_constraints(batch, extrema=None):
    """Constraint rows for the objective, relative to variant 0: focal length 1 % above its
    paraxial value, working distance >= 0.65 mm, the first element's center thickness >= 105 %
    of its value, Petzval curvature bound 0 (active); with `extrema` ([V][S][2], the surface
    extrema of the batch) also the rows from ray heights and angles: semi-diameter of surface
    S-3 at 98 % of its largest ray height (+0.05 mm clearance), incidence angle at surface S-4
    one degree below its largest, edge thickness of the element S-5..S-4 at ray heights."""
    S = batch.surfaces
    f, _ = optics.focal_and_bfd(batch.curvature[0], batch.thickness[0], batch.paraxial_index[0])
    out = [batch_mod.focal_length(1.01 * f, 0.5),
           batch_mod.working_distance(0.65, 4.0, abi.MODE_MIN),
           batch_mod.center_thickness(1, 2, 1.05 * float(batch.thickness[0, 1]), 2.0),
           batch_mod.petzval(0.0, 1.0)]
    if extrema is not None:
        angle = float(np.degrees(np.arccos(extrema[0, S - 4, 1])))
        out += [batch_mod.max_semi_diameter(S - 3, 0.98 * float(extrema[0, S - 3, 0]), 3.0,
                                            clearance=0.05),
                batch_mod.max_incidence_angle(S - 4, angle - 1.0, 50.0),
                batch_mod.min_edge_thickness(S - 5, S - 4, 0.3, 1.0, clearance=0.1)]
    return out


# ---- uneven batches: random topologies (one LM per topology group) ----------------------------

UNEVEN_GLASSES = ((1.5168, 64.2), (1.6204, 60.3), (1.5750, 41.5), (1.7200, 50.3),
                  (1.6477, 33.8), (1.4875, 70.4), (1.8052, 25.4), (1.5891, 61.3))


@functools.lru_cache(maxsize=None)
def _uneven_glass(i):
    return optics.synthetic_glass(*UNEVEN_GLASSES[i])


def _buchdahl_fit(glass, lo_nm=400.0, hi_nm=700.0, nodes=201):
    """Buchdahl coefficients (n0, nu1, nu2, nu3) of a Sellmeier glass: least squares over
    `nodes` wavelengths evenly in [lo_nm, hi_nm] (the fitted glass then stands for the synthetic
    one in a batch whose media carry glass variables)."""
    B, C = glass
    lam = np.linspace(lo_nm, hi_nm, nodes)
    n = np.array([optics.sellmeier_index(B, C, x / 1000.0) for x in lam])
    w = np.array([batch_mod.buchdahl_omega(float(x)) for x in lam])
    A = np.stack([np.ones_like(w), w, w * w, w * w * w], axis=1)
    coeff, *_ = np.linalg.lstsq(A, n, rcond=None)
    return [float(c) for c in coeff]


def uneven_topology(index, seed=5, variants=4, jitter=0.01, surfaces=None, variables=None):
    """One random topology of the uneven workload (seeded by (seed, index)): 8 ... 32 refracting
    spherical surfaces behind the stop forming weak positive elements (a cemented first element
    when the count is odd), 3 ... 5 fields up to 1 deg, 3 ... 7 wavelengths 450 ... 650 nm,
    10 ... 96 variables: the curvatures, the inner thicknesses and the working distance, then
    the Buchdahl coefficients (n0, nu1, nu2) of the glass media in surface order, as many as the
    drawn count asks for (media with glass variables are Buchdahl fits of their glass; the
    others keep the Sellmeier index). Image plane at the paraxial focus, pupil radius 0.08 f.
    `variants` copies with the variables jittered by N(0, jitter) (Irwin-Hall; n0 by jitter / 10),
    the first unperturbed. `surfaces` / `variables` replace the drawn refracting-surface and
    variable counts. Returns (batch, variables)."""
    g = np.random.default_rng([seed, index])
    n_refr = int(g.integers(8, 33))
    F = int(g.integers(3, 6))
    L = int(g.integers(3, 8))
    P_draw = int(g.integers(10, 97))
    if surfaces is not None:
        n_refr = int(surfaces)
    if variables is not None:
        P_draw = int(variables)
    # media after each refracting surface: alternate glass / air, the first element cemented
    # (two glasses) when n_refr is odd, so that the last medium is air
    media = []
    k = 1
    if n_refr % 2 == 1:
        media += [_uneven_glass(int(g.integers(len(UNEVEN_GLASSES)))),
                  _uneven_glass(int(g.integers(len(UNEVEN_GLASSES)))), None]
        k = 4
    while len(media) < n_refr:
        media += [_uneven_glass(int(g.integers(len(UNEVEN_GLASSES)))), None]
    c = [0.0]
    t = [float(g.uniform(1.0, 3.0))]
    glasses = [None] + media
    for s in range(1, n_refr + 1):
        glass_after = glasses[s] is not None
        glass_before = glasses[s - 1] is not None
        if glass_after and not glass_before:
            c.append(float(g.uniform(0.004, 0.012)))      # front of a positive element
        elif glass_before and not glass_after:
            c.append(-float(g.uniform(0.002, 0.010)))     # back
        else:
            c.append(float(g.uniform(-0.01, 0.01)))       # cemented interface
        t.append(float(g.uniform(2.0, 4.0)) if glass_after else float(g.uniform(0.5, 2.0)))
    scale = 1.0
    for _ in range(20):
        cs = [ci * scale for ci in c]
        f, bfd = _paraxial(cs, t, glasses)
        if f > 0 and bfd > 2.0:
            break
        scale *= 0.7
    c = cs
    t[-1] = bfd
    wl = np.linspace(450.0, 650.0, L)
    fields_deg = np.linspace(0.0, 1.0, F)
    types = [abi.SURFACE_STOP] + [abi.SURFACE_SPHERE] * n_refr
    b = SystemBatch(types, wl, np.ones(L), np.ones(F),
                    [abi.PUPIL_AXIS if a == 0.0 else abi.PUPIL_HALF for a in fields_deg],
                    variants, LAMBDA_REF)
    b.curvature[:] = c
    b.thickness[:] = t
    b.index[:] = _index_table(glasses, wl)
    b.paraxial_index = np.tile(_index_table(glasses, [LAMBDA_REF])[:, 0], (variants, 1))
    b.pupil_radius[:] = 0.08 * f
    theta = np.radians(fields_deg)
    b.field_angle[:] = theta
    b.field_target[:] = f * np.tan(theta)
    kinds = ([(abi.VAR_CURVATURE, s, 0) for s in range(1, n_refr + 1)] +
             [(abi.VAR_THICKNESS, s, 0) for s in range(1, n_refr)] +
             [(abi.VAR_WORKING_DISTANCE, n_refr, 0)])
    glass_media = [s for s in range(1, n_refr + 1) if glasses[s] is not None]
    for s in glass_media:
        kinds += [(abi.VAR_GLASS, s, comp) for comp in range(3)]
    kinds = kinds[:min(P_draw, len(kinds))]
    fitted = sorted({s for k, s, _ in kinds if k == abi.VAR_GLASS})
    if fitted:
        b.buchdahl = np.zeros((variants, b.surfaces, 4))
        for s in fitted:
            b.buchdahl[:, s, :] = _buchdahl_fit(glasses[s])
    inf = math.inf
    bounds = {abi.VAR_CURVATURE: (-0.2, 0.2, 0.01), abi.VAR_THICKNESS: (0.05, 1000.0, 1.0),
              abi.VAR_WORKING_DISTANCE: (0.05, 1000.0, 1.0)}
    names = {abi.VAR_CURVATURE: "c[%d]", abi.VAR_THICKNESS: "t[%d]",
             abi.VAR_WORKING_DISTANCE: "t[%d]"}
    variables = VariableSet(
        type=np.array([k for k, _, _ in kinds], dtype=np.int32),
        batch_surface=np.array([s for _, s, _ in kinds], dtype=np.int32),
        lower=np.array([-inf if k == abi.VAR_GLASS else bounds[k][0] for k, _, _ in kinds]),
        upper=np.array([inf if k == abi.VAR_GLASS else bounds[k][1] for k, _, _ in kinds]),
        typical=np.array([batch_mod.GLASS_TYPICAL[c] if k == abi.VAR_GLASS else bounds[k][2]
                          for k, _, c in kinds]),
        names=[("%s[%d]" % (("n0", "nu1", "nu2")[c], s)) if k == abi.VAR_GLASS
               else names[k] % s for k, s, c in kinds],
        component=(np.array([c for _, _, c in kinds], dtype=np.int32) if fitted else None))
    x0 = variables.extract(b, 0)
    if fitted:
        variables.apply(b, np.tile(x0, (variants, 1)))  # the fitted media's index tables
    xs = rng.perturbed_vectors(variables, x0, variants, seed=seed * 1000003 + index,
                               relative=(jitter, jitter, jitter, 0.1 * jitter),
                               distribution="irwin_hall", first_is_base=True)
    variables.apply(b, xs)
    return b, variables


def uneven_settings(failure=abi.FAILURE_HARD):
    return ResidualSettings(rings=6, spokes=8, failure=failure)


# ---- workloads at the limits of the path ---------------------------------------------------

def point_singlet(point=(0.0, 2.0, -300.0), na=0.05, variants=4, seed=7, image_distance=120.0):
    """The singlet lit by a point source (one field; pupil Axis on the axis, Half for a point in
    the y-z plane, Full otherwise, as topology_of picks it), image plane at `image_distance`.
    Mirrors the reference's point-source adjoint test system (tests/adjoint_test_systems.hpp:
    76-95: N-BK7 singlet, source (0, 2, -300) mm, NA 0.05, image plane at 120 mm)."""
    nbk7 = (optics.NBK7_B, optics.NBK7_C)
    glasses = [None, nbk7, None]
    x, y, _ = point
    mode = (abi.PUPIL_AXIS if y == 0.0 else abi.PUPIL_HALF) if x == 0.0 else abi.PUPIL_FULL
    b = SystemBatch([abi.SURFACE_STOP, abi.SURFACE_SPHERE, abi.SURFACE_SPHERE], F_C_D,
                    [1.0, 1.0, 1.0], [1.0], [mode], variants, LAMBDA_REF,
                    source=abi.SOURCE_POINT)
    b.curvature[:] = [0.0, 0.02, -0.005]
    b.thickness[:] = [5.0, 5.0, image_distance]
    b.index[:] = _index_table(glasses, F_C_D)
    b.paraxial_index = np.tile(_index_table(glasses, [LAMBDA_REF])[:, 0], (variants, 1))
    b.pupil_radius[:] = 5.0
    b.source_point[:] = point
    b.source_na = na
    v = singlet_variables()
    v.upper[3] = 500.0
    xs = rng.perturbed_vectors(v, v.extract(b, 0), variants, seed=seed,
                               relative=(0.05, 0.0, 0.0), first_is_base=True)
    v.apply(b, xs)
    return b, v


def long_relay(variants=2, seed=3, fields_deg=(0.0, 0.5), pupil_radius=1.0):
    """64 surfaces (the most a batch holds): the stop and 63 weak spheres, alternately into a
    glass of constant index 1.5 and into air (surfaces 1 ... 62), the last one in air; lens and
    air gaps 1 mm, image plane 50 mm behind. 64 variables: curvatures of surfaces 1 ... 32 and
    thicknesses 0 ... 31.
    Per variant the curvatures are jittered by 5 % (seeded)."""
    S = abi.MAX_SURFACES
    types = [abi.SURFACE_STOP] + [abi.SURFACE_SPHERE] * (S - 1)
    fields_deg = tuple(fields_deg)
    modes = [abi.PUPIL_AXIS if a == 0.0 else abi.PUPIL_HALF for a in fields_deg]
    b = SystemBatch(types, F_C_D, [1.0, 1.0, 1.0], [1.0] * len(fields_deg), modes, variants,
                    LAMBDA_REF)
    c = [0.0] + [0.004 * (1 if (s // 2) % 2 == 0 else -1) * (1 if s % 2 else -1)
                 for s in range(1, S)]
    t = [1.0] * (S - 1) + [50.0]
    n_after = [1.0] + [1.5 if (s % 2 == 1 and s < S - 1) else 1.0 for s in range(1, S)]
    b.curvature[:] = c
    b.thickness[:] = t
    b.index[:] = np.array(n_after)[:, None]
    b.paraxial_index = np.tile(np.array(n_after), (variants, 1))
    b.pupil_radius[:] = pupil_radius
    theta = np.radians(fields_deg)
    b.field_angle[:] = theta
    b.field_target[:] = 0.0
    P = 64
    v = VariableSet(
        type=np.array([abi.VAR_CURVATURE] * 32 + [abi.VAR_THICKNESS] * 32, dtype=np.int32),
        batch_surface=np.array(list(range(1, 33)) + list(range(0, 32)), dtype=np.int32),
        lower=np.array([-0.05] * 32 + [0.2] * 32), upper=np.array([0.05] * 32 + [5.0] * 32),
        typical=np.array([0.01] * 32 + [1.0] * 32),
        names=[f"c[{s}]" for s in range(1, 33)] + [f"t[{s}]" for s in range(32)])
    assert v.size == P
    xs = rng.perturbed_vectors(v, v.extract(b, 0), variants, seed=seed,
                               relative=(0.05, 0.0, 0.0), first_is_base=True)
    v.apply(b, xs)
    return b, v


def many_tuples(variants=3, fields=32, wavelengths=32):
    """The singlet with the most fields and wavelengths a batch holds (32 each): field angles
    0 ... 5 deg, wavelengths 450 ... 700 nm, uneven weights."""
    nbk7 = (optics.NBK7_B, optics.NBK7_C)
    glasses = [None, nbk7, None]
    wl = list(np.linspace(450.0, 700.0, wavelengths))
    ang = list(np.linspace(0.0, 5.0, fields))
    b = SystemBatch([abi.SURFACE_STOP, abi.SURFACE_SPHERE, abi.SURFACE_SPHERE], wl,
                    list(np.linspace(0.5, 1.5, wavelengths)), list(np.linspace(1.0, 0.2, fields)),
                    [abi.PUPIL_AXIS if a == 0.0 else abi.PUPIL_HALF for a in ang], variants,
                    LAMBDA_REF)
    c = [0.0, 0.02, -0.005]
    t = [5.0, 5.0, 0.0]
    f, bfd = _paraxial(c, t, glasses)
    t[2] = bfd
    b.curvature[:] = c
    b.thickness[:] = t
    b.index[:] = _index_table(glasses, wl)
    b.paraxial_index = np.tile(_index_table(glasses, [LAMBDA_REF])[:, 0], (variants, 1))
    b.pupil_radius[:] = 5.0
    theta = np.radians(ang)
    b.field_angle[:] = theta
    b.field_target[:] = f * np.tan(theta)
    v = singlet_variables()
    xs = rng.perturbed_vectors(v, v.extract(b, 0), variants, seed=9,
                               relative=(0.05, 0.0, 0.0), first_is_base=True)
    v.apply(b, xs)
    return b, v


def glass_singlet(variants=4, seed=13, coeff=(1.5187, -0.0812, 0.0198, 0.004)):
    """The singlet with its lens medium a Buchdahl glass (n0, nu1, nu2, nu3 = `coeff`, close to
    N-BK7) and glass variables on it: c[1], c[2], (n0, nu1, nu2) of the lens medium and the
    working distance (P = 6). The index table comes from the coefficients (VariableSet::apply);
    per variant the curvatures and n0 are jittered (seeded)."""
    b, _ = singlet(variants, seed=seed, perturb=0.0)
    b.buchdahl = np.zeros((variants, b.surfaces, 4))
    b.buchdahl[:, 1, :] = coeff
    inf = math.inf
    v = VariableSet(
        type=np.array([abi.VAR_CURVATURE, abi.VAR_CURVATURE, abi.VAR_GLASS, abi.VAR_GLASS,
                       abi.VAR_GLASS, abi.VAR_WORKING_DISTANCE], dtype=np.int32),
        batch_surface=np.array([1, 2, 1, 1, 1, 2], dtype=np.int32),
        lower=np.array([-0.1, -0.1, -inf, -inf, -inf, 10.0]),
        upper=np.array([0.1, 0.1, inf, inf, inf, 500.0]),
        typical=np.array([0.01, 0.01, *batch_mod.GLASS_TYPICAL, 1.0]),
        names=["c[1]", "c[2]", "n0[1]", "nu1[1]", "nu2[1]", "w"],
        component=np.array([0, 0, 0, 1, 2, 0], dtype=np.int32))
    x0 = v.extract(b, 0)
    xs = rng.perturbed_vectors(v, x0, variants, seed=seed, relative=(0.05, 0.0, 0.0, 0.002),
                               first_is_base=True)
    v.apply(b, xs)
    return b, v

"""Front end for the reference's system files: `mikroptim.system/1` (with its glass catalogues
`mikroptim.glass/1`) into the batch, variables and constraint rows of this path, and an optimize
call that reports in the shape of `mikroptim.optimize/1`.

Reference: the file format (docs/dateiformat.md), the reader and validation
(src/io/system_json.cpp, src/model/{system,variables,constraints}.cpp), the glass catalogue
(src/glass/catalog.cpp:130-235, include/mikroptim/glass/dispersion.hpp:27-95), the batch
(make_system_batch / build_batch, src/batch/system_batch.cpp:835-960, reference curvature
:1091-1133), the variables (resolve_variables), the constraints (make_paraxial_constraints,
src/optim/paraxial_residuals.cpp:112-150; make_mechanical_constraints,
src/optim/mechanical_constraints.cpp:87-146) and the optimize output (docs/cli.md:816-900).

Supported is what this path computes: stop / plane / sphere surfaces without decenter or tilt,
catalogue (Sellmeier, polynomial), constant and Buchdahl materials, collimated and point sources,
the three reference spheres with the ideal image point, curvature / thickness / working-distance
variables, glass variables on every glass medium (as convert_glass_variables does: a catalogue
glass becomes the Buchdahl fit of its index, a constant medium Buchdahl with n0 = n), paraxial
(focal length, parfocal length, working distance, paraxial focus, Petzval, sine condition per
zone, distortion per field), glass-hull (three rows per glass medium over the convex hull of
the catalogues' glasses) and mechanical constraints. Ghost plates, ghost-coefficient variables,
the ghost-penalty constraint, smoothed extrema and traced reference points are rejected with a
FormatError naming the field; tolerances and a glass hull without glass variables are ignored
with a warning.
"""

import json
import dataclasses
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi
from . import batch as bt
from . import buchdahl
from . import glass_hull
from . import optics

FORMAT_SYSTEM = "mikroptim.system/1"
FORMAT_GLASS = "mikroptim.glass/1"
BUCHDAHL_LAMBDA0_UM = 0.546074
BUCHDAHL_ALPHA = 2.5
K_PI = 3.14159265358979323846


class FormatError(ValueError):
    """A content error of a file, with the JSON path of the offending field."""

    def __init__(self, path, message, origin=""):
        self.json_path = path
        super().__init__(f"{origin + ': ' if origin else ''}{path}: {message}")


# ---- glass catalogues ----------------------------------------------------------------------------

@dataclass
class Glass:
    name: str
    model: str
    range_nm: tuple
    A: float = 0.0
    B: tuple = ()
    C: tuple = ()
    coefficients: tuple = ()
    exponents: tuple = ()

    def index(self, lambda_nm):
        """Glass::index (src/glass/catalog.cpp:222-235): Sellmeier with constant and any number
        of terms, or the polynomial n^2 = sum a_j lambda^p_j; lambda in um."""
        if not (self.range_nm[0] <= lambda_nm <= self.range_nm[1]):
            raise ValueError(f"glass '{self.name}': {lambda_nm} nm outside "
                             f"[{self.range_nm[0]}, {self.range_nm[1]}] nm")
        lam = lambda_nm / 1000.0
        if self.model == "polynomial":
            n2 = 0.0
            for a, p in zip(self.coefficients, self.exponents):
                ie = int(p)
                if float(ie) == p:
                    base = 1.0 / lam if ie < 0 else lam
                    e = -ie if ie < 0 else ie
                    term = 1.0
                    while e > 0:
                        if e & 1:
                            term *= base
                        base *= base
                        e >>= 1
                else:
                    term = math.pow(lam, p)
                n2 += a * term
            return math.sqrt(n2) if n2 >= 0.0 else math.nan
        l2 = lam * lam
        n2 = 1.0 + self.A
        for b, c in zip(self.B, self.C):
            n2 += b * l2 / (l2 - c)
        return math.sqrt(n2) if n2 >= 0.0 else math.nan


def parse_glass_catalog(d, origin=""):
    """name -> Glass of a parsed `mikroptim.glass/1` document (read_glass,
    src/glass/catalog.cpp:120-200): 1 ... 8 Sellmeier terms with optional constant A, or 1 ... 12
    polynomial terms [coefficient, exponent]; no Sellmeier pole and a valid index (finite, > 0)
    on 201 points of the range; names unique and not AIR."""
    if not isinstance(d, dict) or d.get("format") != FORMAT_GLASS:
        raise FormatError("format", f"expected {FORMAT_GLASS} (not a glass catalogue)", origin)
    if "glasses" not in d:
        raise FormatError("glasses", "required field missing", origin)
    out = {}
    for i, g in enumerate(d["glasses"]):
        p = f"glasses[{i}]"
        name = g.get("name")
        if not isinstance(name, str) or not name or name == "AIR":
            raise FormatError(p + ".name", "invalid glass name (empty or the reserved AIR)",
                              origin)
        if name in out:
            raise FormatError(p + ".name", f"glass '{name}' twice", origin)
        rng = tuple(_num(x, f"{p}.range_nm[{k}]", origin) for k, x in enumerate(g["range_nm"]))
        if len(rng) != 2 or not (rng[0] > 0.0 and rng[0] < rng[1]):
            raise FormatError(p + ".range_nm", "0 < min < max violated", origin)
        model = g.get("model")
        if model == "sellmeier":
            B = tuple(_num(x, f"{p}.B[{k}]", origin) for k, x in enumerate(g["B"]))
            C = tuple(_num(x, f"{p}.C[{k}]", origin) for k, x in enumerate(g["C"]))
            if not 1 <= len(B) <= 8 or not 1 <= len(C) <= 8:
                raise FormatError(p + ".B", "1 to 8 terms", origin)
            if len(B) != len(C):
                raise FormatError(p + ".C", "B and C need the same number of terms", origin)
            A = _num(g["A"], p + ".A", origin) if "A" in g else 0.0
            for k, c in enumerate(C):
                if c > 0.0 and rng[0] <= math.sqrt(c) * 1000.0 <= rng[1]:
                    raise FormatError(f"{p}.C[{k}]", "Sellmeier pole inside the range", origin)
            gl = Glass(name, model, rng, A=A, B=B, C=C)
        elif model == "polynomial":
            terms = g.get("terms")
            if not isinstance(terms, list) or not 1 <= len(terms) <= 12:
                raise FormatError(p + ".terms", "1 to 12 terms [coefficient, exponent]", origin)
            for k, t in enumerate(terms):
                if not isinstance(t, list) or len(t) != 2:
                    raise FormatError(f"{p}.terms[{k}]", "exactly 2 numbers", origin)
            gl = Glass(name, model, rng,
                       coefficients=tuple(_num(t[0], f"{p}.terms[{k}]", origin)
                                          for k, t in enumerate(terms)),
                       exponents=tuple(_num(t[1], f"{p}.terms[{k}]", origin)
                                       for k, t in enumerate(terms)))
        else:
            raise FormatError(p + ".model", f"unknown model '{model}' (sellmeier, polynomial)",
                              origin)
        for k in range(201):
            nm = rng[1] if k == 200 else rng[0] + (rng[1] - rng[0]) * (k / 200.0)
            try:
                n = gl.index(nm)
            except ValueError:
                n = math.nan
            if not (math.isfinite(n) and n > 0.0):
                raise FormatError(p, f"no valid index at {nm} nm", origin)
        out[name] = gl
    return out


def load_glass_catalog(path):
    """name -> Glass of a `mikroptim.glass/1` file."""
    with open(path, encoding="utf-8") as f:
        return parse_glass_catalog(json.load(f), os.path.basename(path))


def resolve_catalogs(entries, base_dir):
    """resolve_glass_catalogs: relative entries from the file's directory upwards."""
    paths = []
    for i, e in enumerate(entries):
        if os.path.isabs(e):
            found = e if os.path.exists(e) else None
        else:
            found, d = None, os.path.abspath(base_dir)
            while True:
                cand = os.path.join(d, e)
                if os.path.exists(cand):
                    found = cand
                    break
                parent = os.path.dirname(d)
                if parent == d:
                    break
                d = parent
        if found is None:
            raise FormatError(f"glass_catalogs[{i}]", f"catalogue '{e}' not found")
        paths.append(found)
    return paths


# ---- the system description ------------------------------------------------------------------

@dataclass
class Surface:
    type: str
    curvature: float
    thickness: float
    material: object  # "AIR", a glass name, or ("constant", n) / ("buchdahl", n0, nu1, nu2, nu3)
    semi_diameter: float


@dataclass
class SystemFile:
    name: str
    wavelengths_nm: list
    weights: list
    reference_nm: float
    source: str
    field_angles_deg: list
    field_weights: list
    source_point: list
    numerical_aperture: float
    target: dict
    reference_sphere: str
    surfaces: list
    variables: list = field(default_factory=list)
    constraints: list = field(default_factory=list)
    constraint_items: list = field(default_factory=list)
    glasses: dict = field(default_factory=dict)
    warnings: list = field(default_factory=list)
    path: str = ""
    # GlassConversion per glass variable's medium: surface, from ("catalog", "constant",
    # "buchdahl"), name, max_fit_error, n0, nu (src/optim/glass_variables.cpp:26-62)
    glass_conversions: list = field(default_factory=list)
    # the glass hull of a glass_hull item with glass variables (glass_hull.GlassHull), or None
    glass_hull: object = None


def _num(v, path, origin):
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise FormatError(path, "number expected", origin)
    return float(v)


def _index_of(v, path, origin):
    x = _num(v, path, origin)
    if x < 0 or x != int(x):
        raise FormatError(path, "non-negative integer expected", origin)
    return int(x)


def _material(m, path, origin):
    if m == "AIR":
        return "AIR"
    if isinstance(m, str) and m:
        return m
    if isinstance(m, dict):
        model = m.get("model")
        if model == "constant":
            n = _num(m.get("n"), path + ".n", origin)
            if not n > 0.0:
                raise FormatError(path + ".n", "n must be > 0", origin)
            return ("constant", n)
        if model == "buchdahl":
            n0 = _num(m.get("n0"), path + ".n0", origin)
            nu = [_num(x, f"{path}.nu[{k}]", origin) for k, x in enumerate(m.get("nu", []))]
            if not n0 > 0.0 or len(nu) not in (2, 3):
                raise FormatError(path, "buchdahl needs n0 > 0 and 2 or 3 nu", origin)
            nu += [0.0] * (3 - len(nu))
            return ("buchdahl", n0, nu[0], nu[1], nu[2])
    raise FormatError(path, "unknown material", origin)


UNSUPPORTED_VARIABLES = ("ghost_coefficient",)
VARIABLE_TYPES = {"curvature": abi.VAR_CURVATURE, "thickness": abi.VAR_THICKNESS,
                  "working_distance": abi.VAR_WORKING_DISTANCE, "glass": abi.VAR_GLASS}
# kDefaultTypical* (include/mikroptim/model/variables.hpp:67-69)
DEFAULT_TYPICAL = {"curvature": 0.01, "thickness": 1.0, "working_distance": 1.0}
UNSUPPORTED_CONSTRAINTS = ("ghost_penalty",)
DEFAULT_GLASS_HULL_WEIGHT = 1e4  # kDefaultGlassHullWeight
PARAXIAL_TYPES = {"focal_length": abi.CON_FOCAL_LENGTH, "parfocal_length": abi.CON_PARFOCAL_LENGTH,
                  "working_distance": abi.CON_WORKING_DISTANCE,
                  "paraxial_focus": abi.CON_PARAXIAL_FOCUS, "petzval": abi.CON_PETZVAL,
                  "sine_condition": abi.CON_SINE_CONDITION, "distortion": abi.CON_DISTORTION}
# kDefaultSineConditionZones (include/mikroptim/model/constraints.hpp:99)
DEFAULT_SINE_CONDITION_ZONES = (0.70710678118654752440, 1.0)
MECHANICAL_TYPES = {"max_semi_diameter": abi.CON_MAX_SEMI_DIAMETER,
                    "min_center_thickness": abi.CON_MIN_CENTER_THICKNESS,
                    "max_center_thickness": abi.CON_MAX_CENTER_THICKNESS,
                    "min_edge_thickness": abi.CON_MIN_EDGE_THICKNESS,
                    "max_incidence_angle": abi.CON_MAX_INCIDENCE_ANGLE}
MODES = {"equal": abi.MODE_EQUAL, "min": abi.MODE_MIN, "max": abi.MODE_MAX}


def read_system(path):
    """read_system_file + validate_system + the glass catalogues: a SystemFile (FormatError with
    the JSON path of the first problem)."""
    origin = os.path.basename(path)
    d = json.load(open(path, encoding="utf-8"))
    if d.get("format") != FORMAT_SYSTEM:
        raise FormatError("format", f"expected {FORMAT_SYSTEM}", origin)
    known = {"format", "name", "description", "units", "glass_catalogs", "spectrum", "source",
             "target", "reference_sphere", "reference_point", "surfaces", "ghost_plates",
             "variables", "constraints", "tolerances"}
    warnings = [f"{k}: unknown field ignored" for k in d if k not in known]
    for k in ("name", "units", "glass_catalogs", "spectrum", "source", "surfaces"):
        if k not in d:
            raise FormatError(k, "required field missing", origin)
    if d["units"] != "mm":
        raise FormatError("units", 'only "mm"', origin)
    if d.get("ghost_plates"):
        raise FormatError("ghost_plates", "ghost plates are not part of this path", origin)
    if d.get("reference_point", "ideal") != "ideal":
        raise FormatError("reference_point", "only the ideal image point", origin)
    sp = d["spectrum"]
    wl = [_num(x, f"spectrum.wavelengths_nm[{i}]", origin) for i, x in enumerate(sp["wavelengths_nm"])]
    ws = [_num(x, f"spectrum.weights[{i}]", origin) for i, x in enumerate(sp["weights"])]
    if not wl or len(ws) != len(wl) or any(x <= 0 for x in wl) or any(x < 0 for x in ws) \
            or not sum(ws) > 0:
        raise FormatError("spectrum", "wavelengths > 0 and weights >= 0 (sum > 0) expected",
                          origin)
    ref_nm = _num(sp["reference_nm"], "spectrum.reference_nm", origin)
    src = d["source"]
    if "type" not in src:
        raise FormatError("source.type", "required field missing", origin)
    stype = src["type"]
    angles, fweights, point, na = [], [], [0.0, 0.0, 0.0], 0.0
    if stype == "collimated":
        angles = [_num(x, f"source.field_angles_deg[{i}]", origin)
                  for i, x in enumerate(src["field_angles_deg"])]
        fweights = [_num(x, f"source.field_weights[{i}]", origin)
                    for i, x in enumerate(src["field_weights"])]
        if not angles or len(fweights) != len(angles) or not sum(fweights) > 0 \
                or any(w_ < 0 for w_ in fweights):
            raise FormatError("source", "field angles and weights do not match", origin)
        for i, a in enumerate(angles):
            if not (math.isfinite(a) and abs(a) < 90.0):
                raise FormatError(f"source.field_angles_deg[{i}]", "|angle| < 90 deg", origin)
    elif stype == "point":
        point = [_num(x, f"source.point[{i}]", origin) for i, x in enumerate(src["point"])]
        na = _num(src["numerical_aperture"], "source.numerical_aperture", origin)
        if len(point) != 3 or not na > 0:
            raise FormatError("source", "point needs three coordinates and NA > 0", origin)
    else:
        raise FormatError("source.type", f"unknown source type '{stype}'", origin)
    target = {k: _num(v, f"target.{k}", origin) for k, v in (d.get("target") or {}).items()
              if k in ("focal_length", "parfocal_length", "working_distance")}
    for k, x in target.items():
        if not math.isfinite(x) or (x == 0.0 if k == "focal_length" else not x > 0.0):
            raise FormatError(f"target.{k}", "invalid target", origin)
    refsph = d.get("reference_sphere", "last_surface")
    if refsph not in ("last_surface", "image_exit_pupil", "image_infinite"):
        raise FormatError("reference_sphere", "unknown reference sphere", origin)
    surfaces = []
    for i, s in enumerate(d["surfaces"]):
        p = f"surfaces[{i}]"
        t = s.get("type")
        if t not in ("stop", "plane", "sphere"):
            raise FormatError(p + ".type", "stop, plane or sphere expected", origin)
        if (t == "stop") != (i == 0):
            raise FormatError(p + ".type", "the stop is surface 0 and only there", origin)
        for k in ("decenter", "tilt_deg"):
            if k in s and any(float(x) != 0.0 for x in s[k]):
                raise FormatError(p + "." + k, "decenter / tilt are not part of this path", origin)
        if (t == "sphere") != ("curvature" in s):
            raise FormatError(p + ".curvature", "curvature only and always for spheres", origin)
        c = _num(s["curvature"], p + ".curvature", origin) if t == "sphere" else 0.0
        th = _num(s["thickness"], p + ".thickness", origin)
        sd = _num(s["semi_diameter"], p + ".semi_diameter", origin)
        if not sd > 0:
            raise FormatError(p + ".semi_diameter", "must be > 0", origin)
        m = _material(s["material"], p + ".material", origin)
        if i == 0 and m != "AIR":
            raise FormatError(p + ".material", "the stop lies in air", origin)
        surfaces.append(Surface(t, c, th, m, sd))
    if len(surfaces) < 2:
        raise FormatError("surfaces", "at least two surfaces", origin)
    base = os.path.dirname(os.path.abspath(path))
    glasses = {}
    for cat in resolve_catalogs(d["glass_catalogs"], base):
        for name, g in load_glass_catalog(cat).items():
            if name in glasses:  # GlassCatalog::merge
                raise FormatError("glass_catalogs", f"glass '{name}' in two catalogues", origin)
            glasses[name] = g
    for i, s in enumerate(surfaces):
        if isinstance(s.material, str) and s.material != "AIR" and s.material not in glasses:
            raise FormatError(f"surfaces[{i}].material", f"glass '{s.material}' not in the "
                              "catalogues", origin)
    sysf = SystemFile(d["name"], wl, ws, ref_nm, stype, angles, fweights, point, na, target,
                      refsph, surfaces, glasses=glasses, warnings=warnings, path=path)
    sysf.variables = _variables(d.get("variables"), sysf, origin)
    cons = _constraints(d.get("constraints"), sysf, origin)
    sysf.constraints = [c for _, c in cons]
    sysf.constraint_items = [k for k, _ in cons]
    if d.get("tolerances"):
        sysf.warnings.append("tolerances: not part of this path, ignored")
    return sysf


def material_index(sysf, m, lambda_nm):
    """material_index (src/glass/index_table.cpp:13-43)."""
    if m == "AIR":
        return 1.0
    if isinstance(m, tuple) and m[0] == "constant":
        return m[1]
    if isinstance(m, tuple) and m[0] == "buchdahl":
        lam = lambda_nm / 1000.0
        den = 1.0 + BUCHDAHL_ALPHA * (lam - BUCHDAHL_LAMBDA0_UM)
        if not den > 0.0:
            raise ValueError("Buchdahl coordinate undefined below 146.074 nm")
        w = (lam - BUCHDAHL_LAMBDA0_UM) / den
        return m[1] + w * (m[2] + w * (m[3] + w * m[4]))
    return sysf.glasses[m].index(lambda_nm)


# ---- variables (resolve_variables) -----------------------------------------------------------------

@dataclass
class ResolvedVariable:
    type: int
    surface: int
    lower: float
    upper: float
    typical: float
    name: str
    item: int
    component: int = 0


def _variables(sec, sysf, origin):
    """resolve_variables (src/model/variables.cpp:333-439): entries in file order, each expanded
    over its surfaces; bounds finite, typical > 0, start within the bounds (error under .min /
    .max), no variable twice."""
    if not sec:
        return []
    S = len(sysf.surfaces)
    out = []
    for k, it in enumerate(sec.get("items", [])):
        p = f"variables.items[{k}]"
        t = it.get("type")
        if t in UNSUPPORTED_VARIABLES:
            raise FormatError(p + ".type", f"'{t}' variables are not part of this path", origin)
        if t not in VARIABLE_TYPES:
            raise FormatError(p + ".type", f"unknown variable type '{t}'", origin)
        if t == "glass":
            # three components with their own scales; the glass hull acts as a penalty, not as
            # bounds (resolve_variables, src/model/variables.cpp:364-383)
            for key in ("min", "max", "typical"):
                if key in it:
                    raise FormatError(f"{p}.{key}", "not allowed for glass", origin)
        lo = _num(it["min"], p + ".min", origin) if "min" in it else -math.inf
        hi = _num(it["max"], p + ".max", origin) if "max" in it else math.inf
        for key, x in (("min", lo), ("max", hi)):
            if key in it and not math.isfinite(x):
                raise FormatError(f"{p}.{key}", "number must be finite (open bound: omit the "
                                  "field)", origin)
        if "min" in it and "max" in it and not lo < hi:
            raise FormatError(p + ".max", f"upper bound {hi} must exceed the lower bound {lo}",
                              origin)
        typ = _num(it["typical"], p + ".typical", origin) if "typical" in it \
            else DEFAULT_TYPICAL.get(t, 1.0)
        if not (math.isfinite(typ) and typ > 0):
            raise FormatError(p + ".typical", "typical size must be finite and > 0", origin)
        if t == "working_distance":
            for key in ("surface", "surfaces"):
                if key in it:
                    raise FormatError(f"{p}.{key}", "working_distance is always the thickness "
                                      "of the last surface", origin)
            targets = [(S - 1, p + ".type")]
        elif "surface" in it and "surfaces" in it or not ("surface" in it or "surfaces" in it):
            raise FormatError(p + ".surface", "give exactly one of surface / surfaces", origin)
        elif "surface" in it:
            targets = [(_index_of(it["surface"], p + ".surface", origin), p + ".surface")]
        elif it["surfaces"] == "all":
            def take(i):
                if t == "curvature":
                    return sysf.surfaces[i].type == "sphere"
                if t == "glass":
                    return sysf.surfaces[i].material != "AIR"
                return i < S - 1
            targets = [(i, p + ".surfaces") for i in range(S) if take(i)]
            if not targets:
                raise FormatError(p + ".surfaces", '"all" finds no surface', origin)
        else:
            if not isinstance(it["surfaces"], list) or not it["surfaces"]:
                raise FormatError(p + ".surfaces", "at least one surface", origin)
            targets = [(_index_of(x, f"{p}.surfaces[{j}]", origin), f"{p}.surfaces[{j}]")
                       for j, x in enumerate(it["surfaces"])]
        vt = VARIABLE_TYPES[t]
        for i, tp in targets:
            if not 0 <= i < S:
                raise FormatError(tp, f"surface {i} does not exist (surfaces 0 to {S - 1})",
                                  origin)
            if t == "curvature" and sysf.surfaces[i].type != "sphere":
                raise FormatError(tp, f"curvature variable only on spheres; surface {i} is "
                                  f"'{sysf.surfaces[i].type}'", origin)
            if t == "thickness" and i == S - 1:
                raise FormatError(tp, "the thickness of the last surface is the working "
                                  "distance", origin)
            if t == "glass" and sysf.surfaces[i].material == "AIR":
                raise FormatError(tp, f"glass variable needs a glass after surface {i}", origin)
            for prev in out:
                if prev.type == vt and prev.surface == i:
                    raise FormatError(tp, f"variable twice: {t} of surface {i} is already in "
                                      f"variables.items[{prev.item}]", origin)
            if t == "glass":
                # convert_glass_variables (src/optim/glass_variables.cpp:26-62): a catalogue glass
                # becomes the Buchdahl fit of its index (3 dispersion terms, 400-700 nm, the
                # fit's defaults), a constant medium Buchdahl with n0 = n and nu = 0; n0, nu1 and
                # nu2 are the variables, nu3 stays fixed
                m = sysf.surfaces[i].material
                conv = {"surface": i, "from": "buchdahl", "name": None, "max_fit_error": 0.0}
                if isinstance(m, tuple) and m[0] == "constant":
                    sysf.surfaces[i].material = ("buchdahl", m[1], 0.0, 0.0, 0.0)
                    conv["from"] = "constant"
                elif isinstance(m, str):
                    try:
                        f = buchdahl.fit_glass(sysf.glasses[m], 3)
                    except buchdahl.FitError as e:
                        raise FormatError(tp, str(e), origin) from None
                    sysf.surfaces[i].material = f.material()
                    conv.update({"from": "catalog", "name": m, "max_fit_error": f.max_abs_error})
                conv["n0"], conv["nu"] = sysf.surfaces[i].material[1], \
                    sysf.surfaces[i].material[2:]
                sysf.glass_conversions.append(conv)
                for c in range(3):
                    name = ("n0", "nu1", "nu2")[c] + f"[{i}]"
                    out.append(ResolvedVariable(vt, i, -math.inf, math.inf,
                                                bt.GLASS_TYPICAL[c], name, k, c))
                continue
            start = sysf.surfaces[i].curvature if t == "curvature" else sysf.surfaces[i].thickness
            if start < lo:
                raise FormatError(p + ".min", f"start value {start} below the lower bound", origin)
            if start > hi:
                raise FormatError(p + ".max", f"start value {start} above the upper bound", origin)
            name = {"curvature": f"c[{i}]", "thickness": f"t[{i}]", "working_distance": "w"}[t]
            out.append(ResolvedVariable(vt, i, lo, hi, typ, name, k))
    return out


def variable_set(sysf):
    v = sysf.variables
    glass = any(x.type == abi.VAR_GLASS for x in v)
    return bt.VariableSet(
        type=np.array([x.type for x in v], dtype=np.int32),
        batch_surface=np.array([x.surface for x in v], dtype=np.int32),
        lower=np.array([x.lower for x in v]), upper=np.array([x.upper for x in v]),
        typical=np.array([x.typical for x in v]), names=[x.name for x in v],
        component=np.array([x.component for x in v], dtype=np.int32) if glass else None)


# ---- constraints (make_paraxial_constraints / make_mechanical_constraints) --------------------------

def _medium_after(sysf, i):
    return "air" if sysf.surfaces[i].material == "AIR" else "glass"


GAP_TYPES = ("min_center_thickness", "max_center_thickness", "min_edge_thickness")
EQUALITY_ONLY = ("focal_length", "paraxial_focus", "sine_condition")
MECHANICAL_FIELDS = ("medium", "semi_diameter", "thickness", "angle", "clearance", "height")


def _resolve_surfaces(sysf, it, t, p, origin):
    """resolve_surfaces (src/model/constraints.cpp:132-182): default or "all" = surfaces
    1 ... S-1 (gaps 0 ... S-2 with the requested medium); listed surfaces must be allowed, match
    the medium and appear once."""
    S = len(sysf.surfaces)
    gap = t in GAP_TYPES
    last = S - 2 if gap else S - 1
    medium = it.get("medium")
    ok = (lambda i: medium is None or _medium_after(sysf, i) == medium)
    sel = it.get("surfaces", it.get("surface"))
    if "surface" in it and "surfaces" in it:
        raise FormatError(p + ".surface", "give one of surface / surfaces", origin)
    if sel is None or sel == "all":
        out = [i for i in range(0 if gap else 1, last + 1) if ok(i)]
        if not out:
            raise FormatError(p + ".type", f"{t}: no matching surface", origin)
        return out
    one = "surface" in it
    lst = [sel] if one else sel
    if not isinstance(lst, list) or not lst:
        raise FormatError(p + ".surfaces", "at least one surface", origin)
    out = []
    for j, x in enumerate(lst):
        sp = p + ".surface" if one else f"{p}.surfaces[{j}]"
        i = _index_of(x, sp, origin)
        if i > last:
            raise FormatError(sp, f"surface {i} not allowed (surfaces 0 to {last})", origin)
        if not ok(i):
            raise FormatError(sp, f"after surface {i} lies {_medium_after(sysf, i)}, the entry "
                              f"asks for medium '{medium}'", origin)
        if i in out:
            raise FormatError(sp, f"surface {i} twice", origin)
        out.append(i)
    return out


def _constraints(sec, sysf, origin):
    """check_constraints (src/model/constraints.cpp:487-715) for the types of this path, then
    make_paraxial_constraints (src/optim/paraxial_residuals.cpp:105-144) and
    make_mechanical_constraints (src/optim/mechanical_constraints.cpp:87-141). Returns
    [(item, Constraint)]: the paraxial items first, then the mechanical ones expanded per
    surface, the order of the extra residual rows of `mikroptim optimize`."""
    if not sec:
        return []
    es = sec.get("extrema_smoothing")
    if es and (es.get("height_beta", 0.0) or es.get("cos_beta", 0.0)):
        raise FormatError("constraints.extrema_smoothing", "smoothed extrema are not part of this "
                          "path", origin)
    S = len(sysf.surfaces)
    items = sec.get("items", [])
    paraxial, hull_rows, mechanical = [], [], []
    for k, it in enumerate(items):
        p = f"constraints.items[{k}]"
        t = it.get("type")
        if t in UNSUPPORTED_CONSTRAINTS:
            raise FormatError(p + ".type", f"'{t}' constraints are not part of this path", origin)
        if t == "glass_hull":
            # no target, no mode, at most one entry; weight (default 1e4) and margin (default
            # 0) finite and >= 0 (check_constraints, src/model/constraints.cpp:626-657)
            if "mode" in it:
                raise FormatError(p + ".mode", "glass_hull has no mode (omit the field)", origin)
            for j in range(k):
                if items[j].get("type") == t:
                    raise FormatError(p + ".type", f"constraint twice: glass_hull is already in "
                                      f"constraints.items[{j}]", origin)
            w = _num(it["weight"], p + ".weight", origin) if "weight" in it \
                else DEFAULT_GLASS_HULL_WEIGHT
            if not (math.isfinite(w) and w >= 0.0):
                raise FormatError(p + ".weight", "weight must be finite and >= 0", origin)
            margin = _num(it["margin"], p + ".margin", origin) if "margin" in it else 0.0
            if not (math.isfinite(margin) and margin >= 0.0):
                raise FormatError(p + ".margin", "margin must be finite and >= 0", origin)
            glass = [x for x in sysf.variables if x.type == abi.VAR_GLASS and x.component == 0]
            if not glass:
                # a penalty on glass variables only (optimize_command.cpp:1528-1539)
                sysf.warnings.append(f"{p}: glass_hull without glass variables adds no residual")
                continue
            # make_glass_hull over every glass of the catalogues, then one element per glass
            # medium (make_glass_hull_penalty, src/optim/glass_variables.cpp:65-101)
            try:
                hull = glass_hull.make_glass_hull(sysf.glasses)
            except (buchdahl.FitError, ValueError) as e:
                raise FormatError(p + ".type", str(e), origin) from None
            if margin > 0.0 and hull.dimension < 3:
                raise FormatError(p + ".margin", "a margin needs a three-dimensional hull; the "
                                  f"catalogues' glasses give dimension {hull.dimension}", origin)
            sysf.glass_hull = hull
            for x in glass:
                hull_rows.extend((k, c) for c in bt.glass_hull(x.surface, w, margin))
            continue
        if "margin" in it:
            raise FormatError(p + ".margin", "only for glass_hull", origin)
        if t not in PARAXIAL_TYPES and t not in MECHANICAL_TYPES:
            raise FormatError(p + ".type", f"unknown constraint type '{t}'", origin)
        mech = t in MECHANICAL_TYPES
        if "mode" in it:
            if it["mode"] not in MODES:
                raise FormatError(p + ".mode", "unknown mode", origin)
            if t in EQUALITY_ONLY and it["mode"] != "equal":
                raise FormatError(p + ".mode", f"{t} is an equation only", origin)
            if t in ("petzval", "distortion") and it["mode"] != "max":
                raise FormatError(p + ".mode", f"{t} is an upper bound only", origin)
            if mech:
                raise FormatError(p + ".mode", f"{t} has no mode", origin)
        w = _num(it["weight"], p + ".weight", origin) if "weight" in it else 1.0
        if not (math.isfinite(w) and w >= 0.0):
            raise FormatError(p + ".weight", "weight must be finite and >= 0", origin)
        mode_name = it.get("mode", "max" if t in ("petzval", "distortion") else "equal")
        shift = it.get("shift", [])
        if shift:
            if not (mech or mode_name != "equal"):
                raise FormatError(p + ".shift", "shifts only for one-sided constraints", origin)
            shift = [_num(x, f"{p}.shift[{j}]", origin) for j, x in enumerate(shift)]
            if any(not (math.isfinite(x) and x >= 0.0) for x in shift):
                raise FormatError(p + ".shift", "shifts must be finite and >= 0", origin)
        if not mech:
            for key in ("surface", "surfaces") + MECHANICAL_FIELDS:
                if key in it:
                    raise FormatError(f"{p}.{key}", "only for mechanical constraints", origin)
        if t != "parfocal_length":
            for key in ("reference_surface", "offset"):
                if key in it:
                    raise FormatError(f"{p}.{key}", "only for parfocal_length", origin)
        if t != "petzval" and "max_curvature" in it:
            raise FormatError(p + ".max_curvature", "only for petzval", origin)
        if t != "distortion" and "max_distortion" in it:
            raise FormatError(p + ".max_distortion", "only for distortion", origin)
        if t != "sine_condition" and it.get("zones"):
            raise FormatError(p + ".zones", "only for sine_condition", origin)
        if mech:
            gap = t in GAP_TYPES
            forbid = {"semi_diameter": t != "max_semi_diameter", "thickness": not gap,
                      "angle": t != "max_incidence_angle", "medium": not gap,
                      "clearance": t not in ("max_semi_diameter", "min_edge_thickness"),
                      "height": t != "min_edge_thickness"}
            for key, bad in forbid.items():
                if bad and key in it:
                    raise FormatError(f"{p}.{key}", f"not allowed for {t}", origin)
            sd = _num(it["semi_diameter"], p + ".semi_diameter", origin) \
                if "semi_diameter" in it else None
            if sd is not None and not (math.isfinite(sd) and sd > 0.0):
                raise FormatError(p + ".semi_diameter", "must be finite and > 0", origin)
            clearance = _num(it.get("clearance", 0.0), p + ".clearance", origin)
            if not (math.isfinite(clearance) and clearance >= 0.0):
                raise FormatError(p + ".clearance", "must be finite and >= 0", origin)
            if gap:
                if "thickness" not in it:
                    raise FormatError(p + ".thickness", f"required for {t} (limit in mm)", origin)
                th = _num(it["thickness"], p + ".thickness", origin)
                strict = t == "max_center_thickness"
                if not (math.isfinite(th) and (th > 0.0 if strict else th >= 0.0)):
                    raise FormatError(p + ".thickness", "invalid limit", origin)
            if t == "max_incidence_angle":
                if "angle" not in it:
                    raise FormatError(p + ".angle", "required (limit in degrees)", origin)
                ang = _num(it["angle"], p + ".angle", origin)
                if not (math.isfinite(ang) and 0.0 < ang < 90.0):
                    raise FormatError(p + ".angle", "must lie between 0 and 90 degrees", origin)
            if "medium" in it and it["medium"] not in ("air", "glass"):
                raise FormatError(p + ".medium", 'only "air" or "glass"', origin)
            if "height" in it and it["height"] not in ("semi_diameter", "rays"):
                raise FormatError(p + ".height", 'only "semi_diameter" or "rays"', origin)
            surfs = _resolve_surfaces(sysf, it, t, p, origin)
            shifted = len(shift) == len(surfs)
            for j, i in enumerate(surfs):
                c = bt.Constraint(MECHANICAL_TYPES[t], w, 0.0, surface=i,
                                  next_surface=i + 1 if gap else i, offset=clearance,
                                  shift=shift[j] if shifted else 0.0)
                if t == "max_semi_diameter":
                    c.target = sd if sd is not None else sysf.surfaces[i].semi_diameter
                elif t == "max_incidence_angle":
                    c.target = math.cos(ang * K_PI / 180.0)
                else:
                    c.target = th
                if t == "min_edge_thickness":
                    c.edge_height = (abi.EDGE_RAYS if it.get("height") == "rays"
                                     else abi.EDGE_SEMI_DIAMETER)
                    c.semi_diameter = sysf.surfaces[i].semi_diameter
                    c.next_semi_diameter = sysf.surfaces[i + 1].semi_diameter
                mechanical.append((k, c))
            continue
        # paraxial
        if t == "petzval":
            if "max_curvature" not in it:
                raise FormatError(p + ".max_curvature", "required for petzval", origin)
            target = _num(it["max_curvature"], p + ".max_curvature", origin)
            if not (math.isfinite(target) and target >= 0.0):
                raise FormatError(p + ".max_curvature", "must be finite and >= 0", origin)
        elif t == "paraxial_focus":
            target = 0.0
            if sysf.source != "collimated":
                raise FormatError(p + ".type", "paraxial_focus needs a collimated source", origin)
        elif t == "distortion":
            if "max_distortion" not in it:
                raise FormatError(p + ".max_distortion", "required for distortion (largest "
                                  "allowed |D| as a fraction)", origin)
            target = _num(it["max_distortion"], p + ".max_distortion", origin)
            if not (math.isfinite(target) and target >= 0.0):
                raise FormatError(p + ".max_distortion", "must be finite and >= 0", origin)
            if sysf.source != "collimated":
                raise FormatError(p + ".type", "distortion needs a collimated source", origin)
        elif t == "sine_condition":
            target = 0.0
            zones = [_num(z, f"{p}.zones[{j}]", origin) for j, z in enumerate(it.get("zones", []))]
            for j, z in enumerate(zones):
                if not (math.isfinite(z) and 0.0 < z <= 1.0):
                    raise FormatError(f"{p}.zones[{j}]", "pupil zone must lie in (0, 1]", origin)
                if z in zones[:j]:
                    raise FormatError(f"{p}.zones[{j}]", f"pupil zone {z} twice", origin)
            if sysf.source != "collimated":
                raise FormatError(p + ".type", "sine_condition needs a collimated source", origin)
        else:
            if t not in sysf.target:
                raise FormatError(p + ".type", f"target.{t} missing", origin)
            target = sysf.target[t]
        if "reference_surface" in it:
            rs = _index_of(it["reference_surface"], p + ".reference_surface", origin)
            if rs >= S:
                raise FormatError(p + ".reference_surface", f"surface {rs} does not exist",
                                  origin)
        if "offset" in it and not math.isfinite(_num(it["offset"], p + ".offset", origin)):
            raise FormatError(p + ".offset", "number must be finite", origin)
        for j in range(k):
            if items[j].get("type") == t:
                raise FormatError(p + ".type", f"constraint twice: {t} is already in "
                                  f"constraints.items[{j}]", origin)
        c = bt.Constraint(PARAXIAL_TYPES[t], w, target, MODES[mode_name])
        if t == "parfocal_length":
            c.surface = int(it.get("reference_surface", 0))
            c.offset = float(it.get("offset", 0.0))
        if len(shift) == 1:
            c.shift = shift[0]
        if t == "sine_condition":
            # one row per zone, in file order (make_paraxial_constraints)
            for z in zones or DEFAULT_SINE_CONDITION_ZONES:
                paraxial.append((k, dataclasses.replace(c, zone=float(z))))
            continue
        if t == "distortion":
            # one row per field of the source; the axis rows stay 0
            for f in range(len(sysf.field_angles_deg)):
                paraxial.append((k, dataclasses.replace(c, field=f)))
            continue
        paraxial.append((k, c))
    return paraxial + hull_rows + mechanical


# ---- the batch (make_system_batch) -------------------------------------------------------------------

def _paraxial(c, t, n):
    """compute_paraxial (src/paraxial/paraxial.cpp:98-190) without thin elements: focal length
    (None when afocal, focal_length_of) and exit pupil position (None at infinity)."""
    ni = n[-1]
    my, mu = optics.paraxial_trace(c, t, n, 1.0, 0.0)
    cy, cu = optics.paraxial_trace(c, t, n, 0.0, 1.0)
    C, D = ni * mu, ni * cu
    f = None if C == 0.0 else 1.0 / -C
    return {"focal_length": f if f is not None and math.isfinite(f) else None,
            "exit_pupil": None if D == 0.0 else -cy / cu}


def make_batch(sysf, variants=1):
    """make_system_batch (src/batch/system_batch.cpp:1173): one variant copied `variants` times."""
    S, L, F = len(sysf.surfaces), len(sysf.wavelengths_nm), max(1, len(sysf.field_angles_deg))
    types = [{"stop": abi.SURFACE_STOP, "plane": abi.SURFACE_PLANE,
              "sphere": abi.SURFACE_SPHERE}[s.type] for s in sysf.surfaces]
    collimated = sysf.source == "collimated"
    kind = {"last_surface": abi.REF_LAST_SURFACE, "image_exit_pupil": abi.REF_IMAGE_EXIT_PUPIL,
            "image_infinite": abi.REF_IMAGE_INFINITE}[sysf.reference_sphere]
    # pupil mode per field (topology_of, src/batch/system_batch.cpp:740-753)
    if collimated:
        modes = [abi.PUPIL_AXIS if a == 0.0 else abi.PUPIL_HALF for a in sysf.field_angles_deg]
    else:
        q = sysf.source_point
        modes = [(abi.PUPIL_AXIS if q[1] == 0.0 else abi.PUPIL_HALF) if q[0] == 0.0
                 else abi.PUPIL_FULL]
    b = bt.SystemBatch(types, sysf.wavelengths_nm, sysf.weights,
                       sysf.field_weights if collimated else [1.0], modes, variants,
                       sysf.reference_nm,
                       source=abi.SOURCE_COLLIMATED if collimated else abi.SOURCE_POINT,
                       reference_sphere=kind)
    c = [s.curvature for s in sysf.surfaces]
    t = [s.thickness for s in sysf.surfaces]
    b.curvature[:] = c
    b.thickness[:] = t
    for s, surf in enumerate(sysf.surfaces):
        for l, nm in enumerate(sysf.wavelengths_nm):
            b.index[:, s, l] = material_index(sysf, surf.material, nm)
    npar = [material_index(sysf, s.material, sysf.reference_nm) for s in sysf.surfaces]
    # Buchdahl coefficients of the Buchdahl media (build_batch, system_batch.cpp:890-903)
    if any(isinstance(s.material, tuple) and s.material[0] == "buchdahl" for s in sysf.surfaces):
        b.buchdahl = np.zeros((variants, S, 4))
        for si, surf in enumerate(sysf.surfaces):
            if isinstance(surf.material, tuple) and surf.material[0] == "buchdahl":
                b.buchdahl[:, si, :] = surf.material[1:5]
    b.paraxial_index = np.tile(np.array(npar), (variants, 1))
    b.pupil_radius[:] = sysf.surfaces[0].semi_diameter
    par = _paraxial(c, t, npar)
    if collimated:
        f = sysf.target.get("focal_length")
        if f is None and any(a != 0.0 for a in sysf.field_angles_deg):
            f = par["focal_length"]
        for fi, deg in enumerate(sysf.field_angles_deg):
            theta = deg * (math.pi / 180.0)
            b.field_angle[:, fi] = theta
            b.field_target[:, fi] = f * math.tan(theta) if (theta != 0.0 and f is not None) \
                else 0.0
    else:
        b.source_point[:] = sysf.source_point
        b.source_na = sysf.numerical_aperture
    if kind != abi.REF_LAST_SURFACE:
        for fi in range(F):
            kappa = 0.0
            if kind == abi.REF_IMAGE_EXIT_PUPIL and par["exit_pupil"] is not None:
                dz = t[-1] - par["exit_pupil"]
                radius = math.hypot(float(b.field_target[0, fi]), dz)
                kappa = (1.0 if dz >= 0.0 else -1.0) / radius
            b.reference_curvature[:, fi] = kappa
    return b


@dataclass
class Loaded:
    system: SystemFile
    batch: object
    variables: object
    settings: object


def load(path, variants=1, rings=10, spokes=8, failure=abi.FAILURE_SOFT, fail_weight=1e6):
    """A system file as (batch, variables, residual settings with the constraint rows); the
    quadrature and failure defaults of `mikroptim optimize` (docs/cli.md:498-520)."""
    sysf = read_system(path)
    s = bt.ResidualSettings(rings=rings, spokes=spokes, failure=failure, fail_weight=fail_weight)
    s.constraints = list(sysf.constraints)
    if sysf.glass_hull is not None:
        s.glass_hull = sysf.glass_hull.packed()
    return Loaded(sysf, make_batch(sysf, variants), variable_set(sysf), s)


# ---- optimize, reported as mikroptim.optimize/1 ------------------------------------------------

def _finite(x):
    x = float(x)
    return x if math.isfinite(x) else None


def _paraxial_value(c, b):
    """The "value" of a paraxial constraint item in mikroptim.optimize/1: focal length 1/phi,
    working distance, parfocal length (offset + thicknesses from the reference surface), Petzval
    curvature; None for paraxial_focus."""
    cv, t, n = list(b.curvature[0]), list(b.thickness[0]), list(b.paraxial_index[0])
    if c.type == abi.CON_FOCAL_LENGTH:
        return _paraxial(cv, t, n)["focal_length"]
    if c.type == abi.CON_WORKING_DISTANCE:
        return t[-1]
    if c.type == abi.CON_PARFOCAL_LENGTH:
        x = c.offset
        for k in range(c.surface, len(t)):
            x += t[k]
        return x
    if c.type == abi.CON_PETZVAL:
        p, nb = 0.0, 1.0
        for k in range(len(cv)):
            p += cv[k] * (n[k] - nb) / (nb * n[k])
            nb = n[k]
        return -n[-1] * p
    return None


CONSTRAINT_NAMES = {v_: k for k, v_ in {**PARAXIAL_TYPES, **MECHANICAL_TYPES}.items()}
MODE_NAMES = {v_: k for k, v_ in MODES.items()}
VARIABLE_NAMES = {v_: k for k, v_ in VARIABLE_TYPES.items()}


def optimize(path, engine, runs=1, max_iterations=100, jacobian="fd", precision="fp64",
             failure=abi.FAILURE_SOFT, fail_weight=1e6, rings=10, spokes=8, seed=0,
             perturbation=0.05):
    """`mikroptim optimize <file> --json` on this library (src/cli/optimize_command.cpp:1517-1600):
    the LM from x0 (run 0) and, for runs > 1, from perturbed starts (perturbed_vector with the
    relative perturbation for every variable type, first_is_base), reported in the layout of
    mikroptim.optimize/1 (docs/cli.md:816-856) with the fields this path computes: quadrature,
    settings, variables, paraxial constraints (value and residual at start and result),
    mechanical constraints (residual, satisfied), runs, best run and result. Iteration logs,
    wavefront tables, ghost and glass reports are not produced (None)."""
    from . import rng

    ld = load(path, 1, rings, spokes, failure, fail_weight)
    sysf, v, s, b = ld.system, ld.variables, ld.settings, ld.batch
    if v.size == 0:
        raise ValueError(f"{path}: no variables")
    x0 = v.extract(b, 0)
    starts = (rng.perturbed_vectors(v, x0, runs, seed=seed, relative=(perturbation,) * 4,
                                    first_is_base=True) if runs > 1 else x0[None, :])
    method = abi.JACOBIAN_ADJOINT if jacobian == "adjoint" else abi.JACOBIAN_FD
    opt = abi.lm_options(max_iterations=max_iterations, jacobian_method=method)
    res = engine.optimize(b, s, v, starts, opt, precision=precision)
    pts = b.take(np.zeros(runs, dtype=np.int64))
    v.apply(pts, starts)
    m0, f0, h0 = engine.merit(pts, s)
    best = int(np.argmin([r.merit for r in res.runs]))
    start, end = b.take([0]), b.take([0])
    v.apply(start, starts[:1])
    v.apply(end, res.x[best:best + 1])
    E = len(s.constraints)
    rows0 = engine.residuals(start, s)[0][0, -E:] if E else []
    rows1 = engine.residuals(end, s)[0][0, -E:] if E else []
    row_of = {id(c): i for i, c in enumerate(s.row_constraints())}
    paraxial, mechanical, hull_elements = [], [], {}
    for c, item in zip(s.constraints, sysf.constraint_items):
        r = row_of[id(c)]
        if c.type == abi.CON_GLASS_HULL:
            # the excess max(0, phi) of each glass medium = |rows| / sqrt(w) (|grad| = 1)
            el = hull_elements.setdefault(c.surface, {"surface": c.surface, "rows": [],
                                                      "weight": c.weight, "margin": c.target,
                                                      "item": item})
            el["rows"].append(r)
            continue
        entry = {"item": item, "type": CONSTRAINT_NAMES[c.type], "weight": c.weight}
        if c.type in PARAXIAL_TYPES.values():
            entry.update({"mode": MODE_NAMES[c.mode], "target": c.target,
                          "value_start": _paraxial_value(c, start),
                          "value_result": _paraxial_value(c, end),
                          "residual_start": _finite(rows0[r]),
                          "residual_result": _finite(rows1[r])})
            paraxial.append(entry)
        else:
            gap = c.type in (abi.CON_MIN_CENTER_THICKNESS, abi.CON_MAX_CENTER_THICKNESS,
                             abi.CON_MIN_EDGE_THICKNESS)
            entry.update({"surface": c.surface, "next_surface": c.next_surface if gap else None,
                          "limit": c.target, "residual_start": _finite(rows0[r]),
                          "residual_result": _finite(rows1[r]), "satisfied": rows1[r] == 0.0})
            mechanical.append(entry)
    hull_report = None
    if hull_elements:
        h = sysf.glass_hull
        first = next(iter(hull_elements.values()))
        sw = math.sqrt(first["weight"])

        def excess(rows, idx):
            return _finite(math.sqrt(sum(rows[i] ** 2 for i in idx)) / sw) if sw > 0 else None
        hull_report = {"glasses": len(h.names), "dimension": h.dimension,
                       "vertices": h.vertex_names(), "weight": first["weight"],
                       "margin": first["margin"],
                       "elements": [{"surface": e["surface"],
                                     "excess_start": excess(rows0, e["rows"]),
                                     "excess_result": excess(rows1, e["rows"])}
                                    for e in hull_elements.values()]}
    status = [abi.LM_STATUS_NAMES[r.status] for r in res.runs]
    soft = failure != abi.FAILURE_HARD
    run_objs = [{"run": i, "status": status[i], "iterations": r.iterations,
                 "accepted": r.accepted, "jacobians": r.jacobians,
                 "start_merit": _finite(m0[i]), "merit": _finite(r.merit),
                 "failed_rays": r.failed_rays, "hard_failed_rays": r.hard_failed_rays,
                 "penalty_rounds": r.penalty_rounds, "fail_weight": r.fail_weight if soft else None,
                 "start": [float(x) for x in starts[i]], "x": [float(x) for x in res.x[i]]}
                for i, r in enumerate(res.runs)]
    rb = res.runs[best]
    return {
        "format": "mikroptim.optimize/1", "system": sysf.name, "file": path, "output": None,
        "backend": "hip", "device": engine.device if hasattr(engine, "device") else 0,
        "precision": precision, "merit_unit": "waves^2",
        "quadrature": {"rings": rings, "spokes": spokes},
        "settings": {"max_iterations": max_iterations, "runs": runs,
                     "perturbation": perturbation if runs > 1 else None,
                     "seed": seed if runs > 1 else None, "failure": "soft" if soft else "hard",
                     "fail_weight": s.fail_weight if soft else None, "jacobian": jacobian,
                     "global": None},
        "variables": [{"name": x.name, "type": VARIABLE_NAMES[x.type], "surface": x.surface,
                       "plate": None, "order": None, "lower": _finite(x.lower),
                       "upper": _finite(x.upper), "typical": x.typical,
                       "start": float(starts[0][k]), "result": float(res.x[best][k])}
                      for k, x in enumerate(sysf.variables)],
        "constraints": paraxial, "glass_hull": hull_report, "mechanical_constraints": mechanical,
        "ghost_diagnosis": None, "iterations": None, "runs": run_objs, "best_run": best,
        "global_search": None,
        "result": {"status": status[best], "converged": status[best].startswith("converged"),
                   "iterations": rb.iterations, "accepted": rb.accepted,
                   "jacobians": rb.jacobians, "start_merit": _finite(m0[0]),
                   "merit": _finite(rb.merit), "start_failed_rays": int(f0[0]),
                   "failed_rays": rb.failed_rays, "start_hard_failed_rays": int(h0[0]),
                   "hard_failed_rays": rb.hard_failed_rays, "penalty_rounds": rb.penalty_rounds,
                   "fail_weight": rb.fail_weight if soft else None, "damping": rb.damping,
                   "gradient_norm": rb.gradient_norm, "model_decrease": rb.model_decrease},
        "reference_nm": sysf.reference_nm, "wavefront": None,
        "wavefront_error": "not computed by this library", "warnings": sysf.warnings,
    }

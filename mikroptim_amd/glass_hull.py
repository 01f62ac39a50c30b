"""The glass hull: the convex hull of a catalogue's glasses in normalized Buchdahl coordinates,
over which the glass-hull rows penalize glass variables (include/mikroptim/glass/glass_hull.hpp,
src/glass/glass_hull.cpp; the rows: src/optim/glass_variables.cpp:135-196).

A glass's coordinates are (n0, nu1, nu2) of its 3-term Buchdahl fit over 400-700 nm (the fit
that converts a start glass under a glass variable). They are normalized per coordinate by the
mean m_j and the span s_j = max - min over the hull's glasses (s_j = 1 for a zero span), so a
unit of u = (x - m) / s is the catalogue's extent in that coordinate. Glasses whose valid range
does not hold the fit range are left out (`excluded`).

The hull is built once per system file on the host; the rows evaluate it per variant on the
device (the closest point of the hull, constraints.hpp) from the packed layout of
mko_extra_rows.glass_hull. Degenerate sets give a hull of lower dimension, decided with the
absolute tolerance eps = 1e-10 max(bounding-box diagonal, largest |coordinate|) as the reference
does: all points within eps of the lexicographically smallest one (a point), of the line to the
point farthest from it (a segment), of the plane through the point farthest from that line (a
polygon, ordered counter-clockwise about the plane's normal, fanned into triangles); otherwise a
polyhedron (Qhull via scipy.spatial.ConvexHull, triangulated facets with outward unit normals)."""

import math
from dataclasses import dataclass, field

import numpy as np

from . import buchdahl

RELATIVE_TOLERANCE = 1e-10


@dataclass
class GlassHull:
    names: list                 # hull glasses in catalogue order
    coordinates: np.ndarray     # [n][3] (n0, nu1, nu2)
    max_fit_error: list
    excluded: list
    center: np.ndarray          # m
    scale: np.ndarray           # s
    dimension: int
    tolerance: float
    triangles: np.ndarray = field(default_factory=lambda: np.zeros((0, 3, 3)))  # [t][3][3] in u
    planes: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))  # [t] normal, offset
    vertices: list = field(default_factory=list)  # indices into names (ascending for 3-d,
    # around the polygon for 2-d, the segment's ends, the point)

    def vertex_names(self):
        return [self.names[i] for i in self.vertices]

    def normalize(self, x):
        x = np.asarray(x, dtype=np.float64)
        return (x - self.center) / self.scale

    def packed(self):
        """mko_extra_rows.glass_hull: [dim, n, m(3), s(3)] + n records (a, b, c, normal,
        offset)."""
        t = self.triangles
        rec = np.zeros((len(t), 13))
        rec[:, 0:9] = t.reshape(len(t), 9)
        if self.dimension == 3:
            rec[:, 9:13] = self.planes
        head = [float(self.dimension), float(len(t)), *self.center, *self.scale]
        return np.concatenate([np.array(head), rec.ravel()])


def _lexmin(points):
    best = 0
    for i in range(1, len(points)):
        if tuple(points[i]) < tuple(points[best]):
            best = i
    return best


def convex_hull(u, relative_tolerance=RELATIVE_TOLERANCE):
    """(dimension, tolerance, triangles [t][3][3], planes [t][4], vertex indices) of the
    points u [n][3]."""
    u = np.asarray(u, dtype=np.float64)
    if not (relative_tolerance > 0.0 and math.isfinite(relative_tolerance)):
        raise ValueError("convex hull: the relative tolerance must be finite and > 0")
    if len(u) == 0:
        return -1, 0.0, np.zeros((0, 3, 3)), np.zeros((0, 4)), []
    if not np.all(np.isfinite(u)):
        raise ValueError("convex hull: coordinate not finite")
    lo, hi = u.min(axis=0), u.max(axis=0)
    eps = relative_tolerance * max(float(np.linalg.norm(hi - lo)), float(np.abs(u).max()))
    i0 = _lexmin(u)
    p0 = u[i0]
    r = np.linalg.norm(u - p0, axis=1)
    i1 = int(np.argmax(r))
    if not r[i1] > eps:                                   # a point
        return 0, eps, np.array([[p0, p0, p0]]), np.zeros((1, 4)), [i0]
    e1 = (u[i1] - p0) / r[i1]
    along = (u - p0) @ e1
    off_line = (u - p0) - along[:, None] * e1[None, :]
    dl = np.linalg.norm(off_line, axis=1)
    i2 = int(np.argmax(dl))
    if not dl[i2] > eps:                                  # a segment
        ia, ib = int(np.argmin(along)), int(np.argmax(along))
        return 1, eps, np.array([[u[ia], u[ib], u[ib]]]), np.zeros((1, 4)), [ia, ib]
    nrm = np.cross(u[i1] - p0, u[i2] - p0)
    nrm /= np.linalg.norm(nrm)
    dp = np.abs((u - p0) @ nrm)
    if not dp.max() > eps:                                # a polygon in the plane
        e2 = np.cross(nrm, e1)
        xy = np.stack([(u - p0) @ e1, (u - p0) @ e2], axis=1)
        ring = _polygon(xy, eps)
        if len(ring) < 3:
            ia, ib = int(np.argmin(along)), int(np.argmax(along))
            return 1, eps, np.array([[u[ia], u[ib], u[ib]]]), np.zeros((1, 4)), [ia, ib]
        tri = np.array([[u[ring[0]], u[ring[k]], u[ring[k + 1]]] for k in range(1, len(ring) - 1)])
        return 2, eps, tri, np.zeros((len(tri), 4)), list(ring)
    from scipy.spatial import ConvexHull                  # a polyhedron
    h = ConvexHull(u)
    tris, planes = [], []
    for simplex, eq in zip(h.simplices, h.equations):
        a, b, c = (tuple(float(x) for x in u[i]) for i in simplex)
        pl = _plane(a, b, c)
        if pl[0] * eq[0] + pl[1] * eq[1] + pl[2] * eq[2] < 0.0:   # outward, as Qhull's facet
            b, c = c, b
            pl = _plane(a, b, c)
        tris.append((a, b, c))
        planes.append(pl)
    return 3, eps, np.array(tris), np.array(planes), sorted(int(i) for i in h.vertices)


def _plane(a, b, c):
    """Unit normal of the triangle abc (b - a) x (c - a) and its offset n . a."""
    e = (b[0] - a[0], b[1] - a[1], b[2] - a[2])
    f = (c[0] - a[0], c[1] - a[1], c[2] - a[2])
    n = (e[1] * f[2] - e[2] * f[1], e[2] * f[0] - e[0] * f[2], e[0] * f[1] - e[1] * f[0])
    length = math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
    if length > 0.0:
        inv = 1.0 / length
        n = (n[0] * inv, n[1] * inv, n[2] * inv)
    return (n[0], n[1], n[2], n[0] * a[0] + n[1] * a[1] + n[2] * a[2])


def _polygon(xy, eps):
    """Indices of the convex polygon of the 2-d points, counter-clockwise (monotone chain;
    points within eps of an edge are left out)."""
    order = sorted(range(len(xy)), key=lambda i: (xy[i][0], xy[i][1]))

    def cross(o, a, b):
        return (xy[a][0] - xy[o][0]) * (xy[b][1] - xy[o][1]) - \
            (xy[a][1] - xy[o][1]) * (xy[b][0] - xy[o][0])

    def chain(idx):
        out = []
        for i in idx:
            while len(out) >= 2:
                a, b = out[-2], out[-1]
                edge = math.hypot(xy[b][0] - xy[a][0], xy[b][1] - xy[a][1])
                if cross(a, b, i) <= eps * max(edge, 1e-300):
                    out.pop()
                else:
                    break
            out.append(i)
        return out
    lower, upper = chain(order), chain(order[::-1])
    return lower[:-1] + upper[:-1]


def make_glass_hull(glasses, dispersion_terms=3, options=None,
                    relative_tolerance=RELATIVE_TOLERANCE):
    """make_glass_hull (src/glass/glass_hull.cpp:85-106) of the glasses name -> Glass in
    catalogue order; the mean is summed in that order."""
    o = options or buchdahl.FitOptions()
    names, coords, errors, excluded = [], [], [], []
    for name, g in glasses.items():
        lo, hi = g.range_nm
        if not (lo <= o.lambda_min_nm <= hi and lo <= o.lambda_max_nm <= hi):
            excluded.append(name)
            continue
        f = buchdahl.fit_glass(g, dispersion_terms, o)
        names.append(name)
        coords.append((f.n0, f.nu[0], f.nu[1]))
        errors.append(f.max_abs_error)
    if not names:
        raise buchdahl.FitError("glass hull: no glass of the catalogues covers the fit range")
    x = np.array(coords, dtype=np.float64)
    center, scale = np.zeros(3), np.ones(3)
    for j in range(3):
        s = 0.0
        for row in coords:
            s += row[j]
        center[j] = s / len(coords)
        lo, hi = min(r[j] for r in coords), max(r[j] for r in coords)
        scale[j] = hi - lo if hi > lo else 1.0
    u = np.array([[(r[j] - center[j]) / scale[j] for j in range(3)] for r in coords])
    dim, eps, tri, planes, verts = convex_hull(u, relative_tolerance)
    return GlassHull(names, x, errors, excluded, center, scale, dim, eps, tri, planes, verts)

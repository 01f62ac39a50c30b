"""Glass-hull rows (mikroptim_amd.glass_hull builds the hull; the rows are evaluated per variant
by constraints.hpp on the device and restated in the oracle). Pinned:
- against the reference's own GlassHull / convex_hull / glass_hull_residual built from its
  sources (oracle/_ref, entry ref_glass_hull_rows): the normalization (center and scale bit for
  bit), the hull's dimension and vertex count, and the oracle's rows at random points in and
  around the hulls of both catalogues, with and without a margin, and at degenerate point sets
  (a point, a segment, collinear points, a plane). The two hulls are triangulated differently
  (Qhull here, the reference's incremental hull there), so rows agree to rounding of the closest
  point, |dr| <= 1e-12 max(sqrt(w), |r|); points within 1e-3 of the hull's face planes are left
  out (the row's direction is discontinuous on the surface and, with a margin, ill-conditioned
  just outside it: (margin + d) (u - closest) / d for a small d);
- by closed forms on a cube: the distance of a point beyond a face, an edge and a corner, and
  inside with a margin the nearest face's normal;
- the system-file items: three rows per glass medium after the paraxial rows and before the
  mechanical ones, defaults, errors."""

import json
import math
import os

import numpy as np
import pytest

from mikroptim_amd import abi, glass_hull as gh, systems
from mikroptim_amd import batch as bt
from mikroptim_amd import system_file as sf
from oracle_lib import oracle, reference

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


def _catalog(name):
    return sf.load_glass_catalog(os.path.join(DATA, "glass", name + ".json"))


def _oracle_rows(hull, x, weight, margin):
    """The oracle's rows of the points x [m][3] as the Buchdahl coefficients of a glass medium."""
    b, v = systems.glass_singlet()
    b = b.take(np.zeros(len(x), dtype=np.int64))
    surf = int(np.asarray(v.batch_surface)[np.asarray(v.type) == abi.VAR_GLASS][0])
    b.buchdahl[:, surf, :3] = x
    s = systems.singlet_settings(abi.FAILURE_HARD)
    s.constraints = bt.glass_hull(surf, weight, margin)
    s.glass_hull = hull.packed()
    return oracle().constraint_rows(b, s)


def _surface_distance(hull, x):
    """|signed distance| of u(x) to the 3-d hull's surface (inside: nearest face plane; outside:
    at least the largest plane excess)."""
    u = hull.normalize(x)
    return np.abs((hull.planes[:, :3] @ u.T - hull.planes[:, 3:4]).max(axis=0))


def _hull_of(coords):
    glasses = {f"G{i}": None for i in range(len(coords))}
    x = np.asarray(coords, dtype=np.float64)
    center, scale = np.zeros(3), np.ones(3)
    for j in range(3):
        s = 0.0
        for r in x:
            s += r[j]
        center[j] = s / len(x)
        scale[j] = x[:, j].max() - x[:, j].min() if x[:, j].max() > x[:, j].min() else 1.0
    u = (x - center) / scale
    dim, eps, tri, planes, verts = gh.convex_hull(u)
    return gh.GlassHull(list(glasses), x, [0.0] * len(x), [], center, scale, dim, eps, tri,
                        planes, verts)


@pytest.mark.parametrize("cat", ["basis", "synthetisch"])
def test_rows_match_reference(cat):
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    h = gh.make_glass_hull(_catalog(cat))
    rng = np.random.default_rng(11)
    lo, hi = h.coordinates.min(0), h.coordinates.max(0)
    x = lo + (hi - lo) * rng.uniform(-0.3, 1.3, size=(1500, 3))
    if h.dimension == 3:
        x = x[_surface_distance(h, x) > 1e-3]
    for margin in ([0.0, 0.05, 0.3] if h.dimension == 3 else [0.0]):
        r, info = ref.glass_hull_rows(h.coordinates, 1e4, margin, x)
        assert info[0] == h.dimension
        assert np.array_equal(info[1:4], h.center) and np.array_equal(info[4:7], h.scale)
        assert info[7] == len(h.vertices)
        mine = _oracle_rows(h, x, 1e4, margin)
        bound = 1e-12 * np.maximum(100.0, np.abs(r).max(axis=1))
        assert np.all(np.abs(mine - r).max(axis=1) <= bound)
        assert np.count_nonzero(np.abs(r).sum(axis=1)) > len(x) // 2


@pytest.mark.parametrize("coords,dim", [
    ([[1.5, -0.05, 0.005]], 0),
    ([[1.5, -0.05, 0.005], [1.7, -0.09, 0.012]], 1),
    ([[1.5, -0.05, 0.005], [1.6, -0.07, 0.0085], [1.7, -0.09, 0.012]], 1),
    ([[1.5, -0.05, 0.005], [1.7, -0.05, 0.005], [1.5, -0.09, 0.005], [1.7, -0.09, 0.005],
      [1.6, -0.07, 0.005]], 2),
])
def test_degenerate_hulls(coords, dim):
    h = _hull_of(coords)
    assert h.dimension == dim
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    rng = np.random.default_rng(3)
    c = np.asarray(coords)
    x = c.mean(0) + rng.normal(size=(300, 3)) * np.array([0.1, 0.02, 0.005])
    r, info = ref.glass_hull_rows(c, 1e4, 0.0, x)
    assert info[0] == dim and np.array_equal(info[1:7], np.concatenate([h.center, h.scale]))
    mine = _oracle_rows(h, x, 1e4, 0.0)
    assert np.all(np.abs(mine - r).max(axis=1) <= 1e-12 * np.maximum(100.0, np.abs(r).max(1)))


def test_cube_closed_forms():
    """Coordinates spanning [0, 1]^3 normalize to u = x - 1/2 (scale 1, mean 1/2): the hull is
    the cube [-1/2, 1/2]^3. Beyond a face the row is sqrt(w) d n, beyond an edge or a corner
    the unit direction from the nearest edge or corner point; inside, with a margin delta, sqrt(w)
    (delta - depth) times the nearest face's normal."""
    corners = np.array([[i, j, k] for i in (0.0, 1.0) for j in (0.0, 1.0) for k in (0.0, 1.0)])
    h = _hull_of(corners)
    assert h.dimension == 3 and np.array_equal(h.scale, np.ones(3))
    assert np.array_equal(h.center, np.full(3, 0.5)) and len(h.vertices) == 8
    w = 4.0
    x = np.array([[1.25, 0.5, 0.5],          # beyond the face x = 1 by 0.25
                  [1.3, 1.4, 0.5],           # beyond the edge (1, 1, z)
                  [1.1, -0.2, 1.3],          # beyond the corner (1, 0, 1)
                  [0.9, 0.5, 0.5]])          # inside, 0.1 below the face x = 1
    r = _oracle_rows(h, x, w, 0.0)
    assert np.allclose(r[0], [2 * 0.25, 0, 0], rtol=0, atol=1e-15)
    d = np.array([0.3, 0.4, 0.0])
    assert np.allclose(r[1], 2 * d, atol=1e-15)
    d = np.array([0.1, -0.2, 0.3])
    assert np.allclose(r[2], 2 * d, atol=1e-15)
    assert np.array_equal(r[3], [0.0, 0.0, 0.0])
    rm = _oracle_rows(h, x[3:], w, 0.25)
    assert np.allclose(rm[0], [2 * 0.15, 0, 0], atol=1e-15)


def _system(tmp_path, edit):
    d = json.load(open(os.path.join(DATA, "systems", "einzellinse_bk7_biegung.json"),
                       encoding="utf-8"))
    d["glass_catalogs"].append("data/glass/synthetisch.json")
    edit(d)
    (tmp_path / "data" / "systems").mkdir(parents=True, exist_ok=True)
    (tmp_path / "data" / "glass").mkdir(exist_ok=True)
    for c in ("basis", "synthetisch"):
        (tmp_path / "data" / "glass" / f"{c}.json").write_text(
            open(os.path.join(DATA, "glass", c + ".json"), encoding="utf-8").read())
    p = tmp_path / "data" / "systems" / "hull.json"
    p.write_text(json.dumps(d))
    return str(p)


def test_system_file_items(tmp_path):
    def edit(d):
        d["variables"]["items"].append({"type": "glass", "surface": 1})
        d["constraints"]["items"].append({"type": "glass_hull", "margin": 0.02})
        d["constraints"]["items"].append({"type": "min_center_thickness", "medium": "glass",
                                          "thickness": 1.0})
    ld = sf.load(_system(tmp_path, edit))
    types = [(c.type, c.surface, c.field) for c in ld.settings.constraints]
    assert types[0][0] == abi.CON_FOCAL_LENGTH
    assert types[1:4] == [(abi.CON_GLASS_HULL, 1, k) for k in range(3)]
    assert types[4][0] == abi.CON_MIN_CENTER_THICKNESS
    c = ld.settings.constraints[1]
    assert c.weight == 1e4 and c.target == 0.02
    h = ld.system.glass_hull
    assert h.dimension == 3 and len(h.names) == 39 and h.excluded == ["SYN-UV", "SYN-IR"]
    assert np.array_equal(ld.settings.glass_hull, h.packed())
    # the rows of the start point: N-BK7 (converted) lies within the margin of the hull's surface
    res, _, _ = oracle().residuals(ld.batch, ld.settings)
    rows = res[0, -5:][1:4]
    assert 0.0 < np.linalg.norm(rows) <= 100.0 * 0.02

    def nohull(d):
        d["constraints"]["items"].append({"type": "glass_hull"})
    ld = sf.load(_system(tmp_path, nohull))
    assert not any(c.type == abi.CON_GLASS_HULL for c in ld.settings.constraints)
    assert any("glass_hull without glass variables" in w for w in ld.system.warnings)


@pytest.mark.parametrize("item,path", [
    ({"type": "glass_hull", "mode": "max"}, "constraints.items[1].mode"),
    ({"type": "glass_hull", "margin": -0.1}, "constraints.items[1].margin"),
    ({"type": "glass_hull", "weight": -1.0}, "constraints.items[1].weight"),
    ({"type": "focal_length", "margin": 0.1}, "constraints.items[1].margin"),
])
def test_system_file_errors(tmp_path, item, path):
    def edit(d):
        d["variables"]["items"].append({"type": "glass", "surface": 1})
        d["constraints"]["items"].append(item)
    with pytest.raises(sf.FormatError) as e:
        sf.read_system(_system(tmp_path, edit))
    assert e.value.json_path == path


def test_margin_needs_a_3d_hull(tmp_path):
    def edit(d):
        d["glass_catalogs"].remove("data/glass/synthetisch.json")   # basis: 3 glasses, a plane
        d["variables"]["items"].append({"type": "glass", "surface": 1})
        d["constraints"]["items"].append({"type": "glass_hull", "margin": 0.1})
    with pytest.raises(sf.FormatError, match="three-dimensional") as e:
        sf.read_system(_system(tmp_path, edit))
    assert e.value.json_path == "constraints.items[1].margin"

    def twice(d):
        d["variables"]["items"].append({"type": "glass", "surface": 1})
        d["constraints"]["items"] += [{"type": "glass_hull"}, {"type": "glass_hull"}]
    with pytest.raises(sf.FormatError, match="twice"):
        sf.read_system(_system(tmp_path, twice))

"""Glass-hull rows on the HIP path (constraints.hpp on the device): the singlet file with a glass
variable on N-BK7 and a glass_hull item over both catalogues (39 glasses, a 3-d hull), at points
whose glass lies inside, within the margin and outside the hull. Rows and merits bit for bit
the oracle's (the same packed hull, the same operations); FD and adjoint Jacobians (the rows by
central differences on the device) against the oracle's; the LM with device control equal to
host control and pulling a glass from outside the hull back to it."""

import json
import os

import numpy as np
import pytest

from mikroptim_amd import abi
from mikroptim_amd import system_file as sf
from oracle_lib import oracle
from test_gpu_limits import _check_jacobians

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


def _load(tmp_path, margin=0.02, failure=abi.FAILURE_HARD):
    d = json.load(open(os.path.join(DATA, "systems", "einzellinse_bk7_biegung.json"),
                       encoding="utf-8"))
    d["glass_catalogs"].append("data/glass/synthetisch.json")
    d["variables"]["items"].append({"type": "glass", "surface": 1})
    d["constraints"]["items"].append({"type": "glass_hull", "margin": margin})
    (tmp_path / "data" / "systems").mkdir(parents=True, exist_ok=True)
    (tmp_path / "data" / "glass").mkdir(exist_ok=True)
    for c in ("basis", "synthetisch"):
        (tmp_path / "data" / "glass" / f"{c}.json").write_text(
            open(os.path.join(DATA, "glass", c + ".json"), encoding="utf-8").read())
    p = tmp_path / "data" / "systems" / "hull.json"
    p.write_text(json.dumps(d))
    return sf.load(str(p), failure=failure)


def _points(ld, n=64, seed=5):
    """Starts with the glass moved uniformly by up to +-0.05 in n0 and +-20 % of nu1 and nu2
    (seeded): inside, near and outside the hull; the other variables at the file's values."""
    v, b = ld.variables, ld.batch
    x0 = v.extract(b, 0)
    g = np.flatnonzero(np.asarray(v.type) == abi.VAR_GLASS)
    xi = np.random.default_rng(seed).uniform(-1.0, 1.0, size=(n, 3))
    x = np.tile(x0, (n, 1))
    x[:, g] += xi * np.array([0.05, 0.2 * abs(x0[g[1]]), 0.2 * abs(x0[g[2]])])
    x[0] = x0
    return x


def test_rows_and_merit_equal_oracle(engine, tmp_path):
    ld = _load(tmp_path)
    x = _points(ld)
    pts = ld.batch.take(np.zeros(len(x), dtype=np.int64))
    ld.variables.apply(pts, x)
    s = ld.settings
    res, _, _ = engine.residuals(pts, s)
    reso, _, _ = oracle().residuals(pts, s, threads=8)
    E = len(s.constraints)
    assert np.array_equal(res[:, -E:], reso[:, -E:])
    hull = res[:, -E:][:, 1:4]
    outside = np.linalg.norm(hull, axis=1)
    assert np.count_nonzero(outside == 0.0) > 0 and np.count_nonzero(outside > 1.0) > 10
    m, f, h = engine.merit(pts, s)
    mo, fo, ho = oracle().merit(pts, s, threads=8)
    assert np.max(np.abs(m / mo - 1.0)) <= 1e-11


def test_jacobians_match_oracle(engine, tmp_path):
    ld = _load(tmp_path)
    x = _points(ld, 4, seed=9)
    b = ld.batch.take(np.zeros(len(x), dtype=np.int64))
    ld.variables.apply(b, x)
    _check_jacobians(engine, b, ld.variables, ld.settings, runs=4)


@pytest.mark.parametrize("method", [abi.JACOBIAN_FD, abi.JACOBIAN_ADJOINT])
def test_lm_device_equals_host_and_reaches_the_hull(engine, tmp_path, method):
    ld = _load(tmp_path)
    x = _points(ld, 6, seed=13)
    base, v, s = ld.batch.take([0]), ld.variables, ld.settings
    kw = dict(max_iterations=15, threads=4, jacobian_method=method)
    rd = engine.optimize(base, s, v, x, abi.lm_options(**kw))
    rh = engine.optimize(base, s, v, x, abi.lm_options(host_control=True, **kw))
    assert np.array_equal(rd.x, rh.x) and np.array_equal(rd.merit, rh.merit)
    assert rd.status == rh.status and rd.cycles == rh.cycles
    E = len(s.constraints)

    def hull_rows(xs):
        pts = base.take(np.zeros(len(xs), dtype=np.int64))
        v.apply(pts, xs)
        return engine.residuals(pts, s)[0][:, -E:][:, 1:4]
    before, after = np.linalg.norm(hull_rows(x), axis=1), np.linalg.norm(hull_rows(rd.x), axis=1)
    assert np.all(after <= np.maximum(before, 1e-3))
    assert before.max() > 1.0 and after.max() < 0.1 * before.max()


def test_optimize_report_glass_hull(engine, tmp_path):
    """`optimize` on the file reports the hull (glasses, dimension, vertices, weight, margin) and
    the glass medium's excess before and after, as the optimize/1 output's glass_hull."""
    ld = _load(tmp_path)
    rep = sf.optimize(ld.system.path, engine, max_iterations=10, jacobian="adjoint")
    g = rep["glass_hull"]
    h = ld.system.glass_hull
    assert g["glasses"] == 39 and g["dimension"] == 3 and g["vertices"] == h.vertex_names()
    assert g["weight"] == 1e4 and g["margin"] == 0.02
    (e,) = g["elements"]
    assert e["surface"] == 1 and e["excess_start"] > 0.0
    assert e["excess_result"] < e["excess_start"]
    assert not any(c["type"] == "glass_hull" for c in rep["constraints"])

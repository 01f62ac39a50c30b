"""The Buchdahl fit of catalogue glasses (mikroptim_amd.buchdahl; reference
src/glass/buchdahl_fit.cpp), which turns a catalogue glass under a glass variable into the
Buchdahl medium the variables act on. Pinned:
- bit for bit against the reference's own fit_buchdahl: the reference build where it is present,
  and always its golden vectors (tests/golden/reference_outputs.npz) for every glass of both
  catalogues, 2 and 3 dispersion terms, Chebyshev and equidistant nodes;
- by the reference tests' analytic oracle (tests/test_buchdahl_fit.cpp:60-146): an index that is
  a Buchdahl polynomial is recovered to 1e-12 (tolerances.hpp kBuchdahlFitExact) for both node
  kinds, the interpolating case (4 nodes, 4 unknowns) and another range; 3 terms fitted to a
  quadratic give nu3 = 0;
- the acceptance and round-trip checks of the same tests (on the basis catalogue: max error < 1e-5
  with 3 terms, 2 terms never better, n0 ~ n(546.074 nm); nu1 < 0; the fitted medium's index
  reaches max_abs_error on the check grid exactly), the node layouts and the error cases."""

import math
import os

import numpy as np
import pytest

from mikroptim_amd import buchdahl as bd
from mikroptim_amd import system_file as sf
from oracle_lib import reference

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "reference_outputs.npz")
EXACT = 1e-12    # kBuchdahlFitExact
FIT3 = 1e-5      # kBuchdahlFit3
LENGTH = 1e-11   # kLengthFp64


def _catalog(name):
    return sf.load_glass_catalog(os.path.join(DATA, "glass", name + ".json"))


def _row(f):
    return np.array([f.n0, *f.nu, f.max_abs_error, f.rms_error])


def _covering(cat):
    return [g for g in cat.values() if g.range_nm[0] <= 400.0 and g.range_nm[1] >= 700.0]


@pytest.mark.parametrize("cat", ["basis", "synthetisch"])
@pytest.mark.parametrize("terms", [2, 3])
@pytest.mark.parametrize("nodes", ["chebyshev", "equidistant"])
def test_fit_equals_reference_bit_for_bit(cat, terms, nodes):
    glasses = _covering(_catalog(cat))
    mine = np.array([_row(bd.fit_glass(g, terms, bd.FitOptions(nodes=nodes))) for g in glasses])
    golden = np.load(GOLDEN)[f"buchdahl_fit_{cat}_{terms}_{nodes}"]
    assert np.array_equal(mine, golden)
    ref = reference()
    if ref is not None:
        assert np.array_equal(mine, np.array([ref.fit_buchdahl(g, terms, nodes=nodes)
                                              for g in glasses]))


def test_other_options_equal_reference():
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent (the golden vectors cover the default options)")
    g = _catalog("basis")["N-BK7"]
    for kw in (dict(samples=4), dict(samples=57, lambda_min=350.0, lambda_max=1000.0),
               dict(samples=31, check_samples=11, nodes="equidistant")):
        o = bd.FitOptions(lambda_min_nm=kw.get("lambda_min", 400.0),
                          lambda_max_nm=kw.get("lambda_max", 700.0),
                          samples=kw["samples"], nodes=kw.get("nodes", "chebyshev"),
                          check_samples=kw.get("check_samples", 3001))
        for terms in (2, 3):
            assert np.array_equal(_row(bd.fit_glass(g, terms, o)),
                                  ref.fit_buchdahl(g, terms, **kw)), (kw, terms)


def _poly(n0, nu1, nu2, nu3):
    return lambda lam: bd.buchdahl_index(n0, nu1, nu2, nu3, lam / 1000.0)


def _expect_exact(f, c):
    assert abs(f.n0 - c[0]) <= EXACT
    for k in range(3):
        assert abs(f.nu[k] - c[k + 1]) <= EXACT
    assert f.max_abs_error <= EXACT and f.rms_error <= f.max_abs_error


def test_recovers_exact_polynomial():
    for c in ((1.518724, -0.05146, 0.00528, -0.000575), (1.434941, -0.02924, 0.00513, 0.00715)):
        p = _poly(*c)
        _expect_exact(bd.fit(p, 3), c)
        _expect_exact(bd.fit(p, 3, bd.FitOptions(nodes="equidistant")), c)
        _expect_exact(bd.fit(p, 3, bd.FitOptions(samples=4)), c)
        _expect_exact(bd.fit(p, 3, bd.FitOptions(lambda_min_nm=300.0, lambda_max_nm=1100.0,
                                                 samples=57)), c)
    quad = (1.6, -0.08, 0.012, 0.0)
    _expect_exact(bd.fit(_poly(*quad), 2), quad)
    _expect_exact(bd.fit(_poly(*quad), 3), quad)


def test_acceptance_and_roundtrip():
    basis = _covering(_catalog("basis"))
    for g in basis + _covering(_catalog("synthetisch")):
        f3, f2 = bd.fit_glass(g, 3), bd.fit_glass(g, 2)
        assert f3.dispersion_terms == 3 and f2.dispersion_terms == 2 and f2.nu[2] == 0.0
        assert f3.nu[0] < 0.0 and f2.nu[0] < 0.0
        if g in basis:
            # the reference's acceptance catalogue (the synthetic glasses' strong dispersion is
            # not held to these bounds there either)
            assert f3.max_abs_error < FIT3, g.name
            assert abs(f3.n0 - g.index(546.074)) < FIT3
            assert f2.rms_error >= f3.rms_error and f2.max_abs_error >= f3.max_abs_error
        # the fitted medium's index (material_index of a Buchdahl medium) is fit.index, and its
        # largest deviation from the catalogue index on the check grid is max_abs_error exactly
        for f in (f2, f3):
            grid = bd.check_grid_nm(f.options)
            err = 0.0
            for lam in grid:
                n = sf.material_index(None, f.material(), lam)
                assert n == f.index(lam)
                err = max(err, abs(n - g.index(lam)))
            assert err == f.max_abs_error


def test_nodes_and_check_grid():
    cheb = bd.fit_nodes_nm()
    assert len(cheb) == 301 and cheb[0] > 400.0 and cheb[-1] < 700.0
    assert abs(cheb[150] - 550.0) <= LENGTH
    for i in range(1, len(cheb)):
        assert cheb[i - 1] < cheb[i]
        assert abs(cheb[i - 1] + cheb[len(cheb) - i] - 1100.0) <= LENGTH
    eq = bd.fit_nodes_nm(bd.FitOptions(nodes="equidistant"))
    assert len(eq) == 301 and eq[0] == 400.0 and eq[-1] == 700.0
    assert abs(eq[1] - eq[0] - 1.0) <= LENGTH
    grid = bd.check_grid_nm()
    assert len(grid) == 3001 and grid[0] == 400.0 and grid[-1] == 700.0
    assert abs(grid[1] - grid[0] - 0.1) <= LENGTH


def test_errors():
    bk7 = _catalog("basis")["N-BK7"]
    narrow = sf.Glass("NARROW", bk7.model, (450.0, 2500.0), bk7.A, bk7.B, bk7.C)
    with pytest.raises(bd.FitError, match=r"NARROW.*\[400, 700\].*\[450, 2500\]"):
        bd.fit_glass(narrow, 3)
    with pytest.raises(bd.FitError, match="2600"):
        bd.fit_glass(narrow, 3, bd.FitOptions(lambda_max_nm=2600.0))
    for terms in (0, 1, 4):
        with pytest.raises(bd.FitError, match="N-BK7.*2 or 3 dispersion terms"):
            bd.fit_glass(bk7, terms)
    with pytest.raises(bd.FitError, match="3 fit nodes are not enough for 4 unknowns"):
        bd.fit_glass(bk7, 3, bd.FitOptions(samples=3))
    bd.fit_glass(bk7, 2, bd.FitOptions(samples=3))
    p = _poly(1.518724, -0.05146, 0.00528, -0.000575)
    with pytest.raises(bd.FitError, match="at least 2 fit nodes"):
        bd.fit(p, 2, bd.FitOptions(samples=1))
    with pytest.raises(bd.FitError, match="check grid"):
        bd.fit(p, 3, bd.FitOptions(check_samples=1))
    with pytest.raises(bd.FitError, match="invalid range"):
        bd.fit(p, 3, bd.FitOptions(lambda_min_nm=700.0, lambda_max_nm=400.0))
    with pytest.raises(bd.FitError, match="invalid range"):
        bd.fit(p, 3, bd.FitOptions(lambda_min_nm=math.nan))
    with pytest.raises(bd.FitError, match="140 nm.*146.074"):
        bd.fit(p, 3, bd.FitOptions(lambda_min_nm=140.0))
    with pytest.raises(bd.FitError, match="not finite"):
        bd.fit(lambda lam: math.inf if lam > 600.0 else 1.5, 3)


def test_other_glass_models_equal_reference():
    """Polynomial glasses (integer powers by repeated squaring, a fractional power by pow) and a
    5-term Sellmeier with a constant, fitted bit for bit as the reference fits them."""
    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    cat = sf.parse_glass_catalog({"format": "mikroptim.glass/1", "glasses": [
        {"name": "S5", "model": "sellmeier", "A": 0.05, "B": [0.9, 0.2, 0.05, 0.01, 0.8],
         "C": [0.005, 0.02, 0.04, 0.06, 90.0], "range_nm": [300, 2500]},
        {"name": "P", "model": "polynomial", "terms": [[2.27, 0], [-0.0094, 2], [0.0110, -2],
                                                       [0.000149, -4], [-1.1e-5, -6],
                                                       [8.0e-7, -8]], "range_nm": [365, 1100]},
        {"name": "Q", "model": "polynomial", "terms": [[2.1, 0], [0.01, -2.5]],
         "range_nm": [400, 800]}]})
    for g in cat.values():
        for terms in (2, 3):
            for nodes in ("chebyshev", "equidistant"):
                mine = _row(bd.fit_glass(g, terms, bd.FitOptions(nodes=nodes)))
                assert np.array_equal(mine, ref.fit_buchdahl(g, terms, nodes=nodes)), g.name

"""Host logic of calibrate_csda_correction (src/Tomography.jl:1043-1179),
run on the oracle's corrected CSDA fluxes (the GPU run is in
tests/test_gpu_calibration.py).  Pinned by closed forms of the fit:

* MC fluxes generated with kappa_hard = 22 (a grid point) and no geometric
  residual: the grid search returns exactly 22, the objective is exactly 0,
  the ridged least squares of an all-zero residual returns zeros;
* with fit_geometric = false the installed residual is held and a
  kappa_hard of 36 is recovered exactly;
* a heavy anchor pins kappa_hard to the anchor;
* Float64 ranges are the correctly rounded lerps Julia's ranges return;
* bins with non-positive MC flux are dropped, none left is an error.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip import calibration as CAL
from diffpumas_hip.tomography import DirectionalPaths, MaterialConfig, SiteConfig


@pytest.fixture(scope="module")
def setup():
    ph = dp.create_physics_csda(("StandardRock", "Water", "Air"))
    rng = np.random.default_rng(3)
    n_cells = 40
    paths = []
    for b in range(14):
        k = int(rng.integers(2, 7))
        cells = rng.choice(n_cells, size=k, replace=False)
        segs = [(int(c), float(rng.uniform(5.0, 40.0))) for c in cells]
        paths.append((float(rng.uniform(5.0, 70.0)), float(rng.uniform(0, 360)), segs,
                      float(rng.uniform(0.0, 300.0)), 500.0))
    P = DirectionalPaths.from_list(paths)
    w = rng.uniform(0.0, 0.4, n_cells)
    sh = np.zeros(n_cells, np.uint8)
    s = dp.csda_energy_quadrature(24, 1.0, 1e4)
    m = MaterialConfig(0, 1, 2, -1, ph.scalars[0, 0], ph.scalars[1, 0])
    site = SiteConfig(800.0, 3e4)
    return ph, P, w, sh, s, m, site


@pytest.fixture(autouse=True)
def _reset():
    yield
    dp.set_csda_correction(enabled=False)
    dp.set_csda_material_correction(enabled=False)


def _oracle_flux(ph, m, site, w, sh, s):
    def f(sub):
        c = dp.get_csda_correction()
        corr = O.corr_array(enabled=c.enabled, kappa_strag=c.kappa_strag,
                            kappa_hard=c.kappa_hard, resid_a=c.resid_a, resid_b=c.resid_b,
                            resid_c=c.resid_c, resid_d=c.resid_d)
        return O.w_flux(ph, O.matcfg_arrays(0, 1, 2, -1, m.rock_density, m.water_density),
                        (site.detector_elevation, site.primary_altitude),
                        dict(row_ptr=sub.row_ptr, cells=sub.cells, lengths=sub.lengths,
                             remaining=sub.remaining_rock, exit_z=sub.exit_z,
                             cos_theta=sub.cos_theta, valid=sub.valid),
                        w, sh, s.energy, s.weight, s.charge, corr=corr)
    return f


def test_float_range():
    assert CAL.float_range(0.0, 80.0, 41) == [2.0 * i for i in range(41)]
    assert CAL.float_range(0.1, 0.3, 3)[1] == 0.2
    assert CAL.float_range(0.0, 1.0, 10)[3] == 1.0 / 3.0
    assert CAL.float_range(20.0, 24.0, 9) == [20.0 + 0.5 * i for i in range(9)]


def test_path_rock_length(setup):
    _, P, *_ = setup
    for b in range(P.n_rays):
        a, e = P.row_ptr[b], P.row_ptr[b + 1]
        assert CAL.path_rock_length(P, b) == pytest.approx(P.remaining_rock[b] +
                                                           P.lengths[a:e].sum(), rel=1e-15)


def test_recovers_kappa_hard(setup):
    ph, P, w, sh, s, m, site = setup
    flux = _oracle_flux(ph, m, site, w, sh, s)
    bins = np.arange(1, P.n_rays)
    dp.set_csda_correction(enabled=True, kappa_strag=1.0, kappa_hard=22.0)
    mc = flux(P.select(bins))
    dp.set_csda_correction(enabled=False)
    ks, kh, st = dp.calibrate_csda_correction(ph, sh, m, site, P, w, s, mc, bins, _flux=flux)
    assert ks == 1.0 and kh == 22.0
    # rel_after compares with exp(log(mc)), as the reference does: zero to rounding
    assert st["mse_log"] == 0.0 and st["rel_after"] < 1e-14 and st["rel_before"] > 1e-3
    assert (st["resid_a"], st["resid_b"], st["resid_c"], st["resid_d"]) == (0.0, 0.0, 0.0, 0.0)
    c = dp.get_csda_correction()
    assert c.enabled and c.kappa_hard == 22.0 and c.kappa_strag == 1.0


def test_held_geometric_residual(setup):
    ph, P, w, sh, s, m, site = setup
    flux = _oracle_flux(ph, m, site, w, sh, s)
    bins = np.arange(P.n_rays)
    geo = dict(resid_a=0.05, resid_b=-0.02, resid_c=0.003, resid_d=0.1)
    dp.set_csda_correction(enabled=True, kappa_strag=1.0, kappa_hard=36.0, **geo)
    mc = flux(P.select(bins))
    ks, kh, st = dp.calibrate_csda_correction(ph, sh, m, site, P, w, s, mc, bins,
                                              fit_geometric=False, _flux=flux)
    assert kh == 36.0 and st["mse_log"] == 0.0 and st["rel_after"] < 1e-14
    assert (st["resid_a"], st["resid_b"], st["resid_c"], st["resid_d"]) == tuple(geo.values())


def test_anchor_and_geometric_fit(setup):
    ph, P, w, sh, s, m, site = setup
    flux = _oracle_flux(ph, m, site, w, sh, s)
    bins = np.arange(P.n_rays)
    dp.set_csda_correction(enabled=True, kappa_strag=1.0, kappa_hard=30.0, resid_a=0.08)
    mc = flux(P.select(bins))
    ks, kh, st = dp.calibrate_csda_correction(ph, sh, m, site, P, w, s, mc, bins,
                                              kh_anchor=10.0, kh_anchor_weight=1e6, _flux=flux)
    assert kh == 10.0
    # the geometric residual then absorbs most of the misfit left by the pinned kappa_hard
    assert st["rel_after"] < st["rel_before"]
    assert math.isfinite(st["resid_a"]) and st["n_bins"] == P.n_rays


def test_bad_inputs(setup):
    ph, P, w, sh, s, m, site = setup
    flux = _oracle_flux(ph, m, site, w, sh, s)
    with pytest.raises(dp.ArgumentError):
        dp.calibrate_csda_correction(ph, sh, m, site, P, w, s, [1.0], [0, 1], _flux=flux)
    with pytest.raises(dp.TransportError):
        dp.calibrate_csda_correction(ph, sh, m, site, P, w, s, [0.0, -1.0, math.nan], [0, 1, 2],
                                     _flux=flux)

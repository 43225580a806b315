"""GPU parity of the w-model CSDA flux and Jacobian with the MC-calibrated
corrections on (CsdaCorrection, CsdaMaterialCorrection,
src/Tomography.jl:112-172, 360-412, 555-637) against the CPU oracle.

Bars: flux and Jacobian within 1e-10 relative of the oracle; switching the
corrections off again restores the plain model bit for bit; the corrected
Jacobian matches central differences of the corrected GPU flux.
"""
import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip.tomography import DirectionalPaths, MaterialConfig, SiteConfig
from test_gpu_wmodel import PATH1, PATH2, PATH3, _oracle_cfg, _oracle_paths, _traced

pytestmark = pytest.mark.gpu

COEF = dict(kappa_strag=1.1, kappa_hard=0.6, resid_a=0.05, resid_b=-0.02, resid_c=0.003,
            resid_d=0.1)
MAT = dict(linear=-0.8, quadratic=1.7)


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def ph4():
    return dp.create_physics(("StandardRock", "Water", "Air"))


@pytest.fixture(autouse=True)
def _reset():
    yield
    dp.set_csda_correction(enabled=False)
    dp.set_csda_material_correction(enabled=False)


def _cfg(ph, porous=False):
    rho_r, rho_w = ph.scalars[0, 0], ph.scalars[1, 0]
    if porous:
        return MaterialConfig(0, 1, 2, 1, rho_r, rho_w, 1500.0, 50.0)
    return MaterialConfig(0, 1, 2, -1, rho_r, rho_w)


def _set(strag=True, mat=True, **over):
    c = dict(COEF, **over)
    dp.set_csda_correction(enabled=strag, **c)
    dp.set_csda_material_correction(enabled=mat, **MAT)
    return O.corr_array(enabled=strag, **c, material=(MAT["linear"], MAT["quadratic"]) if mat else None)


def _compare(ph, m, site, paths, w, sh, s, ctx, corr):
    fg, jg = dp.flux_and_segment_jacobian(ph, sh, m, site, paths, w, s, ctx=ctx)
    fo, jo = O.w_jacobian(ph, _oracle_cfg(m), (site.detector_elevation, site.primary_altitude),
                          _oracle_paths(paths), w, sh, s.energy, s.weight, s.charge, corr=corr)
    ok = np.isfinite(fo)
    assert np.array_equal(np.isnan(fg), np.isnan(fo))
    np.testing.assert_allclose(fg[ok], fo[ok], rtol=1e-10)
    scale = np.abs(jo).max() if len(jo) else 1.0
    np.testing.assert_allclose(jg, jo, rtol=1e-10, atol=1e-13 * scale)
    f0 = dp.compute_directional_flux_csda(ph, sh, m, site, paths, w, s, ctx=ctx)
    np.testing.assert_allclose(f0[ok], fg[ok], rtol=1e-14)
    return fg, jg


@pytest.mark.parametrize("porous", [False, True])
@pytest.mark.parametrize("which", ["both", "var_only", "material_only"])
def test_fixture_paths_corrected(ph4, ctx, porous, which):
    m = _cfg(ph4, porous)
    site = SiteConfig(1000.0, 12000.0)
    paths = DirectionalPaths.from_list([PATH1, PATH2, PATH3])
    w = np.array([0.10, 0.20, 0.05, 0.15])
    sh = np.array([1, 0, 1, 0], np.uint8)
    s = dp.sample_energy_set(6, 1e-2, 1e5, 42)
    f_plain = dp.compute_directional_flux_csda(ph4, sh, m, site, paths, w, s, ctx=ctx)
    corr = _set(strag=which != "material_only", mat=which != "var_only")
    _say(f"w-model corrected fixture porous={porous} {which}")
    fg, _ = _compare(ph4, m, site, paths, w, sh, s, ctx, corr)
    assert np.all(np.abs(fg / f_plain - 1) > 1e-4)


def test_traced_paths_corrected_many_nodes(ph4, ctx):
    """300 samples (> one block of lanes), porous cells, traced paths."""
    paths, n_cells = _traced(ctx, 24, 13)
    rng = np.random.default_rng(21)
    w = rng.uniform(0.0, 0.6, n_cells)
    sh = (rng.random(n_cells) < 0.4).astype(np.uint8)
    s = dp.csda_energy_quadrature(300, 0.5, 1e5)
    corr = _set()
    _say(f"w-model corrected traced: {paths.n_rays} rays, {paths.n_segments} segments")
    _compare(ph4, _cfg(ph4, True), SiteConfig(600.0, 3e4), paths, w, sh, s, ctx, corr)


def test_traced_paths_corrected_two_bins_per_wave(ph4, ctx):
    """32 samples: two bins per 64-lane block, odd ray count."""
    paths, n_cells = _traced(ctx, 37, 19)
    rng = np.random.default_rng(23)
    w = rng.uniform(0.0, 0.6, n_cells)
    sh = (rng.random(n_cells) < 0.4).astype(np.uint8)
    s = dp.csda_energy_quadrature(32, 1.0, 1e4)
    corr = _set()
    _say(f"w-model corrected packed: {paths.n_rays} rays, {paths.n_segments} segments")
    _compare(ph4, _cfg(ph4, True), SiteConfig(600.0, 3e4), paths, w, sh, s, ctx, corr)


def test_switching_off_restores_plain_model(ph4, ctx):
    paths, n_cells = _traced(ctx, 16, 5)
    w = np.random.default_rng(2).uniform(0.0, 0.5, n_cells)
    sh = np.zeros(n_cells, np.uint8)
    s = dp.csda_energy_quadrature(64, 1.0, 1e4)
    site = SiteConfig(500.0, 3e4)
    m = _cfg(ph4)
    f0, j0 = dp.flux_and_segment_jacobian(ph4, sh, m, site, paths, w, s, ctx=ctx)
    _set()
    f1, _ = dp.flux_and_segment_jacobian(ph4, sh, m, site, paths, w, s, ctx=ctx)
    assert not np.allclose(f0, f1)
    dp.set_csda_correction(enabled=False)
    dp.set_csda_material_correction(enabled=False)
    f2, j2 = dp.flux_and_segment_jacobian(ph4, sh, m, site, paths, w, s, ctx=ctx)
    assert np.array_equal(f0, f2) and np.array_equal(j0, j2)


def test_corrected_jacobian_matches_central_difference(ph4, ctx):
    paths, n_cells = _traced(ctx, 4, 9)
    w = np.random.default_rng(4).uniform(0.05, 0.5, n_cells)
    sh = np.zeros(n_cells, np.uint8)
    s = dp.csda_energy_quadrature(32, 1.0, 1e4)
    site = SiteConfig(500.0, 3e4)
    m = _cfg(ph4)
    _set()
    f, J = dp.assemble_forward_and_jacobian(ph4, sh, m, site, paths, w, s, n_cells, ctx=ctx)
    J = J.tocsr()
    h = 1e-5
    checked = 0
    for r in range(paths.n_rays):
        row = J.getrow(r)
        for c, v in zip(row.indices[:3], row.data[:3]):
            wp, wm = w.copy(), w.copy()
            wp[c] += h
            wm[c] -= h
            fp = dp.compute_directional_flux_csda(ph4, sh, m, site, paths, wp, s, ctx=ctx)[r]
            fm = dp.compute_directional_flux_csda(ph4, sh, m, site, paths, wm, s, ctx=ctx)[r]
            assert v == pytest.approx((fp - fm) / (2 * h), rel=1e-4)
            checked += 1
    assert checked >= 6

"""GPU parity of the water-fraction tomography CSDA flux and sparse Jacobian
(dp_csda_w_flux_jacobian) against the CPU oracle.

Bars: flux and Jacobian within 1e-10 relative of the oracle (FP64, the same
discretised program); forward-only and Jacobian launches give identical
fluxes; repeated calls are bitwise identical; the Jacobian matches central
differences of the GPU flux.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip.tomography import DirectionalPaths, MaterialConfig, SiteConfig

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def ph3():
    return dp.create_physics_csda(("StandardRock", "Water", "Air"))


def _cfg(ph, porous=False):
    rho_r, rho_w = ph.scalars[0, 0], ph.scalars[1, 0]
    if porous:
        return MaterialConfig(0, 1, 2, 2, rho_r, rho_w, 1500.0, 50.0)
    return MaterialConfig(0, 1, 2, -1, rho_r, rho_w)


def _oracle_cfg(m):
    return O.matcfg_arrays(m.rock_material, m.water_material, m.air_material, m.porous_material,
                           m.rock_density, m.water_density, m.porous_density,
                           m.porous_top_thickness)


def _oracle_paths(p):
    return dict(row_ptr=p.row_ptr, cells=p.cells, lengths=p.lengths, remaining=p.remaining_rock,
                exit_z=p.exit_z, cos_theta=p.cos_theta, valid=p.valid)


def _compare(ph, m, site, paths, w, sh, s, ctx):
    fg, jg = dp.flux_and_segment_jacobian(ph, sh, m, site, paths, w, s, ctx=ctx)
    fo, jo = O.w_jacobian(ph, _oracle_cfg(m), (site.detector_elevation, site.primary_altitude),
                          _oracle_paths(paths), w, sh, s.energy, s.weight, s.charge)
    ok = np.isfinite(fo)
    assert np.array_equal(np.isnan(fg), np.isnan(fo))
    np.testing.assert_allclose(fg[ok], fo[ok], rtol=1e-10)
    scale = np.abs(jo).max() if len(jo) else 1.0
    np.testing.assert_allclose(jg, jo, rtol=1e-10, atol=1e-13 * scale)
    return fg, jg


PATH1 = (20.0, 0.0, [(0, 30.0), (1, 40.0), (2, 30.0)], 0.0, 60.0)
PATH2 = (35.0, 90.0, [(1, 50.0), (3, 35.0)], 0.0, 55.0)
PATH3 = (25.0, 10.0, [(0, 12.0), (2, 30.0), (0, 8.0), (3, 20.0)], 40.0, 80.0)
BAD = (20.0, 0.0, [(0, 10.0)], 0.0, 60.0, 0)
HORIZ = (95.0, 0.0, [(1, 10.0)], 0.0, 60.0)


@pytest.mark.parametrize("porous", [False, True])
def test_reference_fixture_paths(ph3, ctx, porous):
    """The hand-built paths of test/test_tomography.jl:31-41, plus a path with
    a repeated cell and a dry remainder, an invalid path and a horizontal one."""
    m = _cfg(ph3, porous)
    site = SiteConfig(1000.0, 12000.0)
    paths = DirectionalPaths.from_list([PATH1, PATH2, PATH3, BAD, HORIZ])
    w = np.array([0.10, 0.20, 0.05, 0.15])
    sh = np.array([1, 0, 1, 0], np.uint8)
    s = dp.sample_energy_set(6, 1e-3, 1e6, 42)
    _say(f"w-model fixture porous={porous}")
    fg, jg = _compare(ph3, m, site, paths, w, sh, s, ctx)
    assert math.isnan(fg[3]) and math.isnan(fg[4])
    assert jg[paths.row_ptr[3]] == 0.0 and jg[paths.row_ptr[4]] == 0.0
    assert jg[paths.row_ptr[2] + 2] == 0.0   # repeated cell merged into its first occurrence


def _traced(ctx, n_dirs, seed, nx=16, ny=16, nz=8, size=10.0):
    rng = np.random.default_rng(seed)
    xe = np.linspace(-nx * size / 2, nx * size / 2, nx + 1)
    ye = np.linspace(-ny * size / 2, ny * size / 2, ny + 1)
    ze = np.linspace(0.0, nz * size, nz + 1)
    zen = rng.uniform(5.0, 60.0, n_dirs)
    az = rng.uniform(0.0, 360.0, n_dirs)
    # the surface lies 30-200 m above the top of the grid along the ray
    top = nz * size
    cz = np.cos(np.radians(zen))
    surface = (top + rng.uniform(30.0, 200.0, n_dirs)) / cz
    exit_z = surface * cz
    return dp.trace_voxel_paths(xe, ye, ze, zen, az, surface, exit_z, ctx=ctx), nx * ny * nz


def test_traced_paths_quadrature(ph3, ctx):
    paths, n_cells = _traced(ctx, 48, 3)
    rng = np.random.default_rng(7)
    w = rng.uniform(0.0, 0.6, n_cells)
    sh = np.zeros(n_cells, np.uint8)
    s = dp.csda_energy_quadrature(64, 1.0, 1e4)
    site = SiteConfig(800.0, 3e4)
    _say(f"w-model traced: {paths.n_rays} rays, {paths.n_segments} segments")
    _compare(ph3, _cfg(ph3), site, paths, w, sh, s, ctx)


def test_traced_paths_many_nodes_porous(ph3, ctx):
    """More samples than one block (300 > 256) and the porous recipe."""
    paths, n_cells = _traced(ctx, 12, 11)
    rng = np.random.default_rng(8)
    w = rng.uniform(0.0, 0.9, n_cells)
    sh = (rng.random(n_cells) < 0.5).astype(np.uint8)
    s = dp.csda_energy_quadrature(300, 0.5, 1e5)
    _compare(ph3, _cfg(ph3, True), SiteConfig(0.0, 3e4), paths, w, sh, s, ctx)


@pytest.mark.parametrize("n_nodes", [32, 17])
def test_two_bins_per_wave(ph3, ctx, n_nodes):
    """At most 32 samples put two bins in one 64-lane block; an odd ray count
    leaves the last block with one live bin."""
    paths, n_cells = _traced(ctx, 45, 17)
    rng = np.random.default_rng(31)
    w = rng.uniform(0.0, 0.6, n_cells)
    sh = (rng.random(n_cells) < 0.3).astype(np.uint8)
    s = dp.csda_energy_quadrature(n_nodes, 1.0, 1e4)
    site = SiteConfig(700.0, 3e4)
    m = _cfg(ph3, True)
    _say(f"w-model packed: {paths.n_rays} rays, {n_nodes} nodes")
    _, j = _compare(ph3, m, site, paths, w, sh, s, ctx)
    f0 = dp.compute_directional_flux_csda(ph3, sh, m, site, paths, w, s, ctx=ctx)
    f1, j1 = dp.flux_and_segment_jacobian(ph3, sh, m, site, paths, w, s, ctx=ctx)
    assert np.array_equal(f0, f1) and np.array_equal(j, j1)


def test_forward_only_equals_jacobian_flux_and_deterministic(ph3, ctx):
    paths, n_cells = _traced(ctx, 32, 5)
    w = np.random.default_rng(2).uniform(0.0, 0.5, n_cells)
    sh = np.zeros(n_cells, np.uint8)
    s = dp.csda_energy_quadrature(96, 1.0, 1e4)
    site = SiteConfig(500.0, 3e4)
    m = _cfg(ph3)
    f0 = dp.compute_directional_flux_csda(ph3, sh, m, site, paths, w, s, ctx=ctx)
    f1, j1 = dp.flux_and_segment_jacobian(ph3, sh, m, site, paths, w, s, ctx=ctx)
    f2, j2 = dp.flux_and_segment_jacobian(ph3, sh, m, site, paths, w, s, ctx=ctx)
    assert np.array_equal(f0, f1)
    assert np.array_equal(f1, f2) and np.array_equal(j1, j2)


def test_jacobian_matches_central_difference(ph3, ctx):
    """test/test_tomography.jl:109-121 on the GPU flux (h = 1e-5, rtol 1e-4)."""
    paths, n_cells = _traced(ctx, 4, 9)
    w = np.random.default_rng(4).uniform(0.05, 0.5, n_cells)
    sh = np.zeros(n_cells, np.uint8)
    s = dp.csda_energy_quadrature(32, 1.0, 1e4)
    site = SiteConfig(500.0, 3e4)
    m = _cfg(ph3)
    f, J = dp.assemble_forward_and_jacobian(ph3, sh, m, site, paths, w, s, n_cells, ctx=ctx)
    J = J.tocsr()
    h = 1e-5
    checked = 0
    for r in range(paths.n_rays):
        row = J.getrow(r)
        for c, v in zip(row.indices[:3], row.data[:3]):
            wp, wm = w.copy(), w.copy()
            wp[c] += h
            wm[c] -= h
            fp = dp.compute_directional_flux_csda(ph3, sh, m, site, paths, wp, s, ctx=ctx)[r]
            fm = dp.compute_directional_flux_csda(ph3, sh, m, site, paths, wm, s, ctx=ctx)[r]
            assert v == pytest.approx((fp - fm) / (2 * h), rel=1e-4)
            checked += 1
    assert checked >= 6


def test_bad_input(ph3, ctx):
    paths = DirectionalPaths.from_list([(20.0, 0.0, [(7, 10.0)], 0.0, 60.0)])
    s = dp.csda_energy_quadrature(4, 1.0, 10.0)
    with pytest.raises(dp.ArgumentError):
        dp.compute_directional_flux_csda(ph3, np.zeros(4, np.uint8), _cfg(ph3), SiteConfig(0, 3e4),
                                         paths, np.zeros(4), s, ctx=ctx)
    bad = dp.EnergySamples(np.array([-1.0]), np.array([1.0]), np.array([0.0]))
    ok_paths = DirectionalPaths.from_list([(20.0, 0.0, [(1, 10.0)], 0.0, 60.0)])
    with pytest.raises(dp.ArgumentError):
        dp.compute_directional_flux_csda(ph3, np.zeros(4, np.uint8), _cfg(ph3), SiteConfig(0, 3e4),
                                         ok_paths, np.zeros(4), bad, ctx=ctx)

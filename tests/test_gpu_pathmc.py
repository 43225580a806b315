"""GPU parity of the fixed-path (tomography) MC, dp_mc_fixed_paths, against
the CPU oracle.  The GPU draws the same Philox stream per event as the
oracle: most events agree to rounding (a rounding difference can flip a
rejection test, after which an event legitimately diverges).  Bars: per-event
agreement for the large majority of events, per-path flux within 3 sigma,
CSDA mode within 1e-10, bitwise determinism.
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
    return dp.create_physics(("StandardRock", "Water", "Air"))


def _traced(ctx, n_dirs, seed, nx=12, ny=12, nz=6, size=10.0):
    rng = np.random.default_rng(seed)
    xe = np.linspace(-nx * size / 2, nx * size / 2, nx + 1)
    ye = np.linspace(-ny * size / 2, ny * size / 2, ny + 1)
    ze = np.linspace(0.0, nz * size, nz + 1)
    zen = rng.uniform(5.0, 60.0, n_dirs)
    az = rng.uniform(0.0, 360.0, n_dirs)
    cz = np.cos(np.radians(zen))
    surface = (nz * size + rng.uniform(10.0, 60.0, n_dirs)) / cz
    return dp.trace_voxel_paths(xe, ye, ze, zen, az, surface, surface * cz, ctx=ctx), nx * ny * nz


def _oracle_paths(p):
    return dict(row_ptr=p.row_ptr, cells=p.cells, lengths=p.lengths, remaining=p.remaining_rock,
                exit_z=p.exit_z, cos_theta=p.cos_theta, valid=p.valid)


def _setup(ph, n_cells, seed):
    m = MaterialConfig(0, 1, 2, -1, 2650.0, 1000.0)
    w = np.random.default_rng(seed).uniform(0.0, 0.4, n_cells)
    w[::5] = 0.0
    media, cm, cd = dp.build_cell_properties_for_mc(np.zeros(n_cells, bool), np.round(w, 1), m)
    return m, media, cm, cd


def test_w_field_events_match_oracle(ph3, ctx):
    paths, n_cells = _traced(ctx, 8, 1)
    m, media, cm, cd = _setup(ph3, n_cells, 2)
    s = dp.sample_energy_set(400, 1.0, 1e4, 5)
    site = SiteConfig(600.0, 3e4)
    _say(f"path MC: {paths.n_rays} paths, {paths.n_segments} segments, {len(media)} media")
    f, sg, ev = dp.compute_directional_flux_mc(ph3, m, site, paths, media, cm, cd, s, 11,
                                               return_events=True, ctx=ctx)
    fo, so, evo = O.mc_paths(ph3, media, cm, cd, 0, 2, 2650.0, (600.0, 3e4), _oracle_paths(paths),
                             (s.energy, s.weight, s.charge), 11, per_event=True)
    evo = evo.reshape(ev.shape)
    same = (np.abs(ev - evo) <= 1e-9 * np.maximum(np.abs(evo), 1e-300)) | ((ev == 0) & (evo == 0))
    one_zero = int(((ev == 0) != (evo == 0)).sum())
    _say(f"  per-event agreement {same.mean():.6f} ({same.sum()}/{same.size}), "
         f"zero on one side only: {one_zero}")
    for i in np.argwhere(~same)[:10]:
        _say(f"    path/event {tuple(i)}: gpu {ev[tuple(i)]!r} oracle {evo[tuple(i)]!r}")
    assert same.mean() >= 0.99 and one_zero == 0
    assert np.all(np.abs(f - fo) <= 3 * np.hypot(sg, so) + 1e-12 * np.abs(fo))


def test_density_field_csda_matches_oracle(ph3, ctx):
    """One rock medium with per-cell densities (the density-field workload),
    deterministic CSDA steps: agreement to rounding."""
    paths, n_cells = _traced(ctx, 16, 3)
    rho = np.random.default_rng(4).uniform(2000.0, 3000.0, n_cells)
    m = MaterialConfig(0, 1, 2, -1, 2650.0, 1000.0)
    s = dp.sample_energy_set(64, 1.0, 1e4, 6)
    site = SiteConfig(0.0, 3e4)
    f, sg = dp.compute_directional_flux_mc(ph3, m, site, paths, [((0,), (1.0,))],
                                           np.zeros(n_cells, np.int32), rho, s, 1,
                                           straggling=False, ctx=ctx)
    fo, so = O.mc_paths(ph3, [([0], [1.0])], np.zeros(n_cells, np.int32), rho, 0, 2, 2650.0,
                        (0.0, 3e4), _oracle_paths(paths), (s.energy, s.weight, s.charge), 1,
                        straggling=False)
    np.testing.assert_allclose(f, fo, rtol=1e-10)
    np.testing.assert_allclose(sg, so, rtol=1e-8)


def test_deterministic_and_invalid(ph3, ctx):
    paths, n_cells = _traced(ctx, 6, 7)
    paths.valid[2] = 0
    m, media, cm, cd = _setup(ph3, n_cells, 8)
    s = dp.sample_energy_set(300, 1.0, 1e3, 9)
    site = SiteConfig(0.0, 3e4)
    a = dp.compute_directional_flux_mc(ph3, m, site, paths, media, cm, cd, s, 3, ctx=ctx)
    b = dp.compute_directional_flux_mc(ph3, m, site, paths, media, cm, cd, s, 3, ctx=ctx)
    assert np.array_equal(a[0], b[0], equal_nan=True) and np.array_equal(a[1], b[1], equal_nan=True)
    assert math.isnan(a[0][2]) and math.isnan(a[1][2])
    assert np.all(np.isfinite(np.delete(a[0], 2)))


def test_rejects_quadrature_nodes_and_scattering(ph3, ctx):
    paths = DirectionalPaths.from_list([(20.0, 0.0, [(0, 100.0)], 0.0, 60.0)])
    m = MaterialConfig(0, 1, 2, -1, 2650.0, 1000.0)
    with pytest.raises(dp.ArgumentError):
        dp.compute_directional_flux_mc(ph3, m, SiteConfig(1000.0, 12000.0), paths, [((0,), (1.0,))],
                                       [0], [2650.0], dp.csda_energy_quadrature(4, 1.0, 10.0), 1,
                                       ctx=ctx)
    with pytest.raises(dp.ArgumentError):
        dp.compute_directional_flux_mc(ph3, m, SiteConfig(1000.0, 12000.0), paths, [((0,), (1.0,))],
                                       [0], [2650.0], dp.sample_energy_set(4, 1.0, 10.0, 1), 1,
                                       scattering=True, ctx=ctx)


def test_mc_and_csda_in_one_launch(ph3, ctx):
    """dp_mc_fixed_paths_csda: a path's MC samples and its CSDA quadrature
    nodes share one persistent launch.  The MC part equals a separate
    dp_mc_fixed_paths run bit for bit; the CSDA part (charge-summed nodes,
    straggling off) equals the oracle's deterministic walk to 1e-10."""
    paths, n_cells = _traced(ctx, 24, 21)
    paths.valid[3] = 0
    m, media, cm, cd = _setup(ph3, n_cells, 22)
    s = dp.sample_energy_set(700, 1.0, 1e4, 23)
    nodes = dp.csda_energy_quadrature(96, 1.0, 1e4)
    site = SiteConfig(600.0, 3e4)
    _say(f"combined launch: {paths.n_rays} paths, {len(s.energy)} samples + "
         f"{len(nodes.energy)} nodes per path")
    f, sg, cf = dp.compute_directional_flux_mc_and_csda(ph3, m, site, paths, media, cm, cd, s,
                                                        nodes, 31, ctx=ctx)
    f2, sg2 = dp.compute_directional_flux_mc(ph3, m, site, paths, media, cm, cd, s, 31, ctx=ctx)
    assert np.array_equal(f, f2, equal_nan=True) and np.array_equal(sg, sg2, equal_nan=True)
    fo, _ = O.mc_paths(ph3, media, cm, cd, 0, 2, 2650.0, (600.0, 3e4), _oracle_paths(paths),
                       (nodes.energy, nodes.weight, nodes.charge), 1, straggling=False)
    ok = paths.valid.astype(bool)
    assert np.all(np.isnan(cf[~ok]))
    np.testing.assert_allclose(cf[ok], fo[ok], rtol=1e-10)
    _say(f"  csda flux range {cf[ok].min():.4g}..{cf[ok].max():.4g}; mc/csda mean ratio "
         f"{np.nanmean(f / cf):.4f}")

"""Launch autotuning (dp_set_autotune / dp_set_tuning / dp_tuning_report).

The tuned kernels are the density adjoint (tables in LDS or read through L2,
rays per queue pop, grid size), the straggled backward MC (blocks per
resident slot) and the fixed-path MC (blocks per resident slot).  Each
candidate launch configuration must give the same outputs bit for bit: the
per-ray and per-event work does not depend on which block does it, and the
reductions have a fixed shape.  The test tunes each kernel once, then pins
every candidate in turn and compares its outputs with the untuned default.
"""
import numpy as np
import pytest

import diffpumas_hip as dp
from diffpumas_hip.tomography import MaterialConfig, SiteConfig
from workloads import cube_rays, energy_nodes, uniform_densities

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def ph3():
    return dp.create_physics(("StandardRock", "Water", "Air"))


def _sweep(tctx, prefix, run):
    """Default result, then tuning, then each tried candidate pinned; all
    results must equal the default's bit for bit."""
    tctx.set_autotune(False)
    tctx.clear_tuning()
    ref = run()
    tctx.set_autotune(True)
    tuned = run()
    tctx.set_autotune(False)
    rep = {k: v for k, v in tctx.tuning().items() if k.startswith(prefix)}
    assert len(rep) == 1, rep
    key, ent = next(iter(rep.items()))
    _say(f"{key}: best {ent['best']} {ent['ms']:.3f} ms; tried "
         + ", ".join(f"{k}:{m:.3f}" for k, m in ent["tried"]))
    assert len(ent["tried"]) >= 2
    assert ent["best"] in [k for k, _ in ent["tried"]]
    _eq(ref, tuned, key + " tuned")
    for knobs, _ in ent["tried"]:
        tctx.set_tuning(key, *knobs)
        _eq(ref, run(), f"{key} pinned {knobs}")
    return key


def _eq(a, b, label):
    if a is None or b is None:
        assert a is None and b is None, label
        return
    if isinstance(a, dict):
        assert a.keys() == b.keys(), label
        for k in a:
            _eq(a[k], b[k], f"{label}[{k}]")
        return
    if isinstance(a, (tuple, list)):
        for x, y in zip(a, b):
            _eq(x, y, label)
        return
    assert np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True), label


def test_density_adjoint_candidates_agree(physics):
    tctx = dp.Context(0)
    geo = dp.CubeGeometry(160.0, 32, 32, 16, physics.get_material_index("StandardRock"),
                          physics.get_material_index("Air"), width=320.0)
    sxy, dirs, cos_t = cube_rays(geo, 4, 256, seed=5)
    paths = dp.trace_cube_paths(geo, sxy, dirs, cos_t, ctx=tctx)
    rho = uniform_densities(geo.n_cells, seed=6)
    ne, nw, nq = energy_nodes(256, 1.0, 1e4)
    _say(f"adjoint autotune: {paths.n_rays} rays, {paths.n_segments} ray*voxels")
    _sweep(tctx, "csda_cube",
           lambda: dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=tctx))


def test_straggled_mc_candidates_agree(ph3):
    tctx = dp.Context(0)
    rock, water, air = (ph3.get_material_index(n) for n in ("StandardRock", "Water", "Air"))
    kw = dict(rock_thickness=300.0, rock_density=2400.0, elevation=30.0, energy_min=1.0,
              energy_max=1e4, n_samples=1 << 16, seed=4, straggling=True, scattering=True,
              rock_mixture=[(rock, 0.85), (water, 0.14), (air, 0.01)], return_events=True)
    _say(f"MC autotune: {kw['n_samples']} events")
    _sweep(tctx, "mc_flat_stochastic", lambda: dp.run_backward_mc(ph3, ctx=tctx, **kw))


def test_path_mc_candidates_agree(ph3):
    tctx = dp.Context(0)
    nx, ny, nz, size, n_dirs = 12, 12, 6, 10.0, 48
    rng = np.random.default_rng(3)
    xe = np.linspace(-nx * size / 2, nx * size / 2, nx + 1)
    ye = np.linspace(-ny * size / 2, ny * size / 2, ny + 1)
    ze = np.linspace(0.0, nz * size, nz + 1)
    zen = rng.uniform(5.0, 60.0, n_dirs)
    az = rng.uniform(0.0, 360.0, n_dirs)
    cz = np.cos(np.radians(zen))
    surface = (nz * size + rng.uniform(10.0, 60.0, n_dirs)) / cz
    paths = dp.trace_voxel_paths(xe, ye, ze, zen, az, surface, surface * cz, ctx=tctx)
    n_cells = nx * ny * nz
    m = MaterialConfig(0, 1, 2, -1, 2650.0, 1000.0)
    w = np.random.default_rng(4).uniform(0.0, 0.4, n_cells)
    media, cm, cd = dp.build_cell_properties_for_mc(np.zeros(n_cells, bool), np.round(w, 1), m)
    s = dp.sample_energy_set(600, 1.0, 1e4, 5)
    nodes = dp.csda_energy_quadrature(64, 1.0, 1e4)
    site = SiteConfig(600.0, 3e4)
    _say(f"path MC autotune: {paths.n_rays} paths x ({len(s.energy)} events + "
         f"{len(nodes.energy)} nodes)")
    _sweep(tctx, "path_mc",
           lambda: dp.compute_directional_flux_mc_and_csda(ph3, m, site, paths, media, cm, cd,
                                                           s, nodes, 9, ctx=tctx))


def test_report_and_pinning_are_per_context(physics):
    a, b = dp.Context(0), dp.Context(0)
    a.set_tuning("csda_cube K=1 nodes=1 rays=1 max_seg=1 jac=1", 0, 4, 2)
    assert "csda_cube K=1" in a.tuning_report()
    assert b.tuning_report() == ""
    a.clear_tuning()
    assert a.tuning_report() == ""

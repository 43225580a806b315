"""GPU parity: HIP kernels through the C ABI against the CPU oracle.

Bars: traversal bit-exact; CSDA flux and Jacobian within 1e-10 relative;
deterministic (bitwise) repeats; full-size properties where the oracle is
too slow.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip import _lib as L
from workloads import cube_rays, energy_nodes, lognormal_cavity_densities, uniform_densities

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _say(msg):
    print(msg, flush=True)


def cube(physics, nx=12, ny=10, nz=8, width=120.0, thick=80.0):
    return dp.CubeGeometry(thick, nx, ny, nz, physics.get_material_index("StandardRock"),
                           physics.get_material_index("Air"), width=width, width_y=width * 0.9)


# ------------------------------------------------------------ traversal
@pytest.mark.parametrize("seed", [1, 2])
def test_trace_cube_bit_exact(physics, ctx, seed):
    geo = cube(physics)
    sxy, dirs, cos_t = cube_rays(geo, 6, 200, seed=seed, el_range=(5.0, 89.0))
    _say(f"trace_cube seed={seed} rays={len(dirs)}")
    paths = dp.trace_cube_paths(geo, sxy, dirs, cos_t, ctx=ctx)
    rp, cells, lens, ze = O.trace_cube(geo, sxy, dirs)
    assert np.array_equal(paths.row_ptr, rp)
    assert np.array_equal(paths.cell, cells)
    assert np.array_equal(paths.length.view(np.uint64), lens.view(np.uint64))
    assert np.array_equal(paths.z_exit.view(np.uint64), ze.view(np.uint64))


def test_trace_cube_edge_cases(physics, ctx):
    geo = cube(physics, 4, 4, 4, 40.0, 40.0)
    # vertical, axis-aligned in x, start on a face, grazing, start outside
    sxy = [[0.0, 0.0], [0.0, 0.0], [-20.0, -18.0], [5.0, 5.0], [100.0, 0.0]]
    dirs = [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.6, 0.0, 0.8], [0.0, 0.99999, 0.00447], [0, 0, 1.0]]
    cos_t = [d[2] for d in dirs]
    paths = dp.trace_cube_paths(geo, sxy, dirs, cos_t, ctx=ctx)
    rp, cells, lens, ze = O.trace_cube(geo, sxy, dirs)
    assert np.array_equal(paths.row_ptr, rp)
    assert np.array_equal(paths.cell, cells)
    assert np.array_equal(paths.length, lens)
    assert rp[-1] - rp[-2] == 0  # the ray starting outside has no segment


def test_trace_cube_empty(physics, ctx):
    geo = cube(physics)
    paths = dp.trace_cube_paths(geo, np.zeros((0, 2)), np.zeros((0, 3)), np.zeros(0), ctx=ctx)
    assert paths.n_rays == 0 and paths.n_segments == 0


def test_trace_edges_bit_exact(ctx):
    rng = np.random.default_rng(9)
    xe = np.cumsum(np.r_[-300.0, rng.uniform(10, 40, 20)])
    ye = np.cumsum(np.r_[-250.0, rng.uniform(10, 40, 18)])
    ze = np.cumsum(np.r_[-1.0, rng.uniform(20, 60, 12)])
    n = 3000
    el = np.radians(rng.uniform(10, 89, n))
    az = np.radians(rng.uniform(0, 360, n))
    dirs = np.stack([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az), np.sin(el)], 1)
    surface = rng.uniform(50, 900, n)
    origins = np.zeros((n, 3))
    _say(f"trace_edges rays={n}")
    rp, cells, lens, tr, rem = dp.trace_edge_paths(xe, ye, ze, origins, dirs, surface, ctx=ctx)
    orp, ocells, olens, otr = O.trace_edges(xe, ye, ze, origins, dirs, surface)
    assert np.array_equal(rp, orp)
    assert np.array_equal(cells, ocells)
    assert np.array_equal(lens.view(np.uint64), olens.view(np.uint64))
    assert np.array_equal(tr.view(np.uint64), otr.view(np.uint64))


def test_trace_edges_edge_cases(ctx):
    """Origins outside the volume and on edge planes, axis-aligned and
    vertical directions, surfaces shorter than the start offset, and a ray
    count that leaves part of a wave empty."""
    xe = np.array([-30.0, -10.0, 0.0, 15.0, 40.0])
    ye = np.array([-20.0, 0.0, 5.0, 25.0])
    ze = np.array([0.0, 10.0, 30.0, 45.0])
    origins = np.array([[0.0, 0.0, 0.0], [-50.0, 0.0, 1.0], [0.0, 5.0, 10.0], [15.0, -20.0, 0.0],
                        [1.0, 1.0, 1.0], [2.0, 2.0, 2.0], [0.0, 0.0, 44.0]])
    dirs = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.6, 0.0, 0.8], [0.0, 0.6, 0.8],
                     [0.48, 0.6, 0.64], [0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    surface = np.array([60.0, 100.0, 50.0, 80.0, 1e-4, 0.0, 500.0])
    rp, cells, lens, tr, rem = dp.trace_edge_paths(xe, ye, ze, origins, dirs, surface, ctx=ctx)
    orp, ocells, olens, otr = O.trace_edges(xe, ye, ze, origins, dirs, surface)
    assert np.array_equal(rp, orp)
    assert np.array_equal(cells, ocells)
    assert np.array_equal(lens.view(np.uint64), olens.view(np.uint64))
    assert np.array_equal(tr.view(np.uint64), otr.view(np.uint64))
    assert rp[2] - rp[1] == 0              # starts outside: no segment
    assert rp[5] - rp[4] == 0 and rp[6] - rp[5] == 0   # surface inside the start offset


# ------------------------------------------------------ per-voxel CSDA
def _problem(physics, ctx, n_det=3, n_dir=24, n_nodes=40, seed=4, nx=12, ny=10, nz=8):
    geo = cube(physics, nx, ny, nz)
    sxy, dirs, cos_t = cube_rays(geo, n_det, n_dir, seed=seed)
    paths = dp.trace_cube_paths(geo, sxy, dirs, cos_t, ctx=ctx)
    rho = lognormal_cavity_densities(geo, seed)
    ne, nw, nq = energy_nodes(n_nodes, 1.0, 1e4)
    return geo, paths, rho, ne, nw, nq


@pytest.mark.parametrize("n_nodes", [1, 40, 64, 300])
def test_csda_cube_flux_and_jacobian(physics, ctx, n_nodes):
    geo, paths, rho, ne, nw, nq = _problem(physics, ctx, n_nodes=n_nodes)
    _say(f"csda_cube rays={paths.n_rays} nodes={n_nodes} segments={paths.n_segments}")
    flux, jac, jrem = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, scale=0.5, ctx=ctx)
    oflux, ograds = O.cube_csda(physics, geo, paths.row_ptr, paths.cell, paths.length,
                                paths.cos_theta, paths.z_exit, rho, ne, nw, nq, scale=0.5)
    assert np.all(oflux > 0)
    np.testing.assert_allclose(flux, oflux, rtol=RTOL, atol=0)
    for r in range(paths.n_rays):
        g = {}
        for s in range(paths.row_ptr[r], paths.row_ptr[r + 1]):
            g[int(paths.cell[s])] = g.get(int(paths.cell[s]), 0.0) + jac[s]
        if jrem[r] != 0.0:
            g[0] = g.get(0, 0.0) + jrem[r]
        og = ograds[r]
        assert set(k for k, v in g.items() if v != 0.0) <= set(og)
        scale = max(abs(v) for v in og.values())
        for cell, v in og.items():
            assert g.get(cell, 0.0) == pytest.approx(v, rel=RTOL, abs=1e-13 * scale), (r, cell)


def test_csda_cube_clamped_lookups(physics, ctx):
    """Node energies at and above the top of the tables (1e9 GeV): lookups
    clamp, c = dE_out/dE_in is 0 on those segments, and the FP64 kernel's
    prefix form of the reverse recursion closes and restarts there.  Flux and
    per-cell gradients against the oracle's forward-mode derivative."""
    geo, paths, rho, _, _, _ = _problem(physics, ctx, n_nodes=8)
    ne, nw, nq = energy_nodes(64, 3e8, 5e9)
    top = physics.data[geo.rock_material, L.T_ENERGY][-1]
    assert ne.min() < top < ne.max()
    _say(f"csda_cube clamped: rays={paths.n_rays} nodes 3e8-5e9 GeV (table top {top:g})")
    flux, jac, jrem = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=ctx)
    oflux, ograds = O.cube_csda(physics, geo, paths.row_ptr, paths.cell, paths.length,
                                paths.cos_theta, paths.z_exit, rho, ne, nw, nq)
    np.testing.assert_allclose(flux, oflux, rtol=RTOL, atol=0)
    for r in range(paths.n_rays):
        g = {}
        for s in range(paths.row_ptr[r], paths.row_ptr[r + 1]):
            g[int(paths.cell[s])] = g.get(int(paths.cell[s]), 0.0) + jac[s]
        if jrem[r] != 0.0:
            g[0] = g.get(0, 0.0) + jrem[r]
        og = ograds[r]
        scale = max([abs(v) for v in og.values()] + [1e-300])
        for cell, v in og.items():
            assert g.get(cell, 0.0) == pytest.approx(v, rel=RTOL, abs=1e-13 * scale), (r, cell)


def test_csda_cube_vjp_dense(physics, ctx):
    geo, paths, rho, ne, nw, nq = _problem(physics, ctx, n_nodes=32)
    flux, jac, jrem = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=ctx)
    fbar = np.random.default_rng(1).uniform(0.5, 2.0, paths.n_rays)
    grad = dp.csda_cube_vjp(paths, jac, jrem, fbar, geo.n_cells, ctx=ctx)
    _, ograds = O.cube_csda(physics, geo, paths.row_ptr, paths.cell, paths.length,
                            paths.cos_theta, paths.z_exit, rho, ne, nw, nq)
    dense = np.zeros(geo.n_cells)
    for r, og in enumerate(ograds):
        for cell, v in og.items():
            dense[cell] += fbar[r] * v
    np.testing.assert_allclose(grad, dense, rtol=1e-9, atol=1e-14 * np.abs(dense).max())
    # deterministic: bitwise identical on repeat
    grad2 = dp.csda_cube_vjp(paths, jac, jrem, fbar, geo.n_cells, ctx=ctx)
    assert np.array_equal(grad, grad2)


def test_csda_cube_deterministic(physics, ctx):
    geo, paths, rho, ne, nw, nq = _problem(physics, ctx, n_nodes=256)
    a = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=ctx)
    b = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=ctx)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_csda_cube_gradient_fd_full_path(physics, ctx):
    """Adjoint vs central FD (h = 1e-3 kg/m3) of the GPU forward itself."""
    geo, paths, rho, ne, nw, nq = _problem(physics, ctx, n_det=2, n_dir=6, n_nodes=256)
    flux, jac, jrem = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=ctx)
    grad = dp.csda_cube_vjp(paths, jac, jrem, np.ones(paths.n_rays), geo.n_cells, ctx=ctx)
    rng = np.random.default_rng(3)
    on_path = np.unique(paths.cell)
    h = 1e-3
    for c in rng.choice(on_path, size=min(12, len(on_path)), replace=False):
        p = rho.copy(); p[c] += h
        m = rho.copy(); m[c] -= h
        fp = dp.csda_cube_flux(physics, geo, paths, p, ne, nw, nq, jacobian=False, ctx=ctx).sum()
        fm = dp.csda_cube_flux(physics, geo, paths, m, ne, nw, nq, jacobian=False, ctx=ctx).sum()
        fd = (fp - fm) / (2 * h)
        assert grad[c] == pytest.approx(fd, rel=2e-6, abs=1e-9 * abs(flux.sum()))


def test_csda_cube_rejects_bad_input(physics, ctx):
    geo, paths, rho, ne, nw, nq = _problem(physics, ctx)
    with pytest.raises(dp.ArgumentError):
        dp.csda_cube_flux(physics, geo, paths, rho[:3], ne, nw, nq, ctx=ctx)
    with pytest.raises(dp.ArgumentError):
        dp.csda_cube_flux(physics, geo, paths, rho, -ne, nw, nq, ctx=ctx)


# ------------------------------------------------- flat CSDA Monte Carlo
@pytest.mark.parametrize("thick,elev,emin,emax,n", [(100.0, 45.0, 1.0, 100.0, 4000),
                                                    (10.0, 90.0, 2.0, 20.0, 1000),
                                                    (0.0, 90.0, 10.0, 10.0, 1)])
def test_mc_flat_csda_vs_oracle(physics, ctx, thick, elev, emin, emax, n):
    _say(f"mc_flat_csda thick={thick} elev={elev} n={n}")
    r = dp.run_backward_mc(physics, rock_thickness=thick, rock_density=2650.0, elevation=elev,
                           energy_min=emin, energy_max=emax, n_samples=n, compute_gradient=True,
                           seed=1, straggling=False, scattering=False, ctx=ctx)
    rock, air = physics.get_material_index("StandardRock"), physics.get_material_index("Air")
    c = math.cos((90 - elev) * math.pi / 180)
    o = O.mc_flat_csda(physics, [thick, 2650.0, 0.0, 0.0, 3e4, 0, 0, 0], [rock, air, -1],
                       emin, emax, c, n, 1)
    assert r["flux"] == pytest.approx(o[0], rel=1e-10)
    assert r["sigma"] == pytest.approx(o[1], rel=1e-8, abs=1e-12 * o[0])
    assert r["gradient_flux"] == pytest.approx(o[2], rel=1e-10)
    assert r["gradient"] == pytest.approx(o[3], rel=1e-9, abs=1e-13)
    assert r["flux"] == pytest.approx(r["gradient_flux"], rel=1e-12)
    if thick == 0.0:
        assert r["gradient"] == 0.0 and r["sigma"] == 0.0


def test_mc_flat_csda_deterministic(physics, ctx):
    kw = dict(rock_thickness=100.0, elevation=45.0, energy_min=1.0, energy_max=100.0,
              n_samples=20000, compute_gradient=True, seed=5, straggling=False, scattering=False)
    a = dp.run_backward_mc(physics, ctx=ctx, **kw)
    b = dp.run_backward_mc(physics, ctx=ctx, **kw)
    assert a == b

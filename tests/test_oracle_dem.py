"""Pins of the oracle's topography ray exit: map_elevation (src/Turtle.jl:
523-553) and ray_surface_exit_distance (examples/lvd_tomography.jl:219-252).

The reference holds no numbers for them; pinned by closed forms:
  * bilinear interpolation returns the node values at nodes, the corner mean at
    cell centres, and reproduces any plane exactly; off the map it reports
    outside;
  * over a flat or planar terrain the exit distance is the analytic
    intersection of the ray with the plane (40 bisections of a 40 m bracket:
    4e-11 m);
  * rays that leave the map first, or point downwards, find no exit.
"""
import math

import numpy as np
import pytest

import _oracle as O


def _dem(f, nx=41, ny=31, x0=-2.0, y0=-1.5, dx=0.1, dy=0.1):
    xs = x0 + dx * np.arange(nx)
    ys = y0 + dy * np.arange(ny)
    data = np.array([[f(x, y) for x in xs] for y in ys])   # data[iy, ix]
    return dict(nx=nx, ny=ny, x0=x0, y0=y0, dx=dx, dy=dy, data=data), xs, ys


def _dir(zen, az):
    t, p = math.radians(zen), math.radians(az)
    return [math.sin(t) * math.sin(p), math.sin(t) * math.cos(p), math.cos(t)]


def test_bilinear_nodes_centres_planes():
    rng = np.random.default_rng(1)
    dem, xs, ys = _dem(lambda x, y: 0.0)
    dem["data"] = rng.uniform(900.0, 1400.0, dem["data"].shape)
    pts = [(xs[3], ys[5]), (xs[10], ys[0]), (xs[-1], ys[-1])]
    z, inside = O.dem_elevation(dem, pts)
    assert inside.all()
    # (x - x0) / dx of a node can land one ulp below the integer: rounding-level agreement
    np.testing.assert_allclose(z, [dem["data"][5, 3], dem["data"][0, 10], dem["data"][-1, -1]],
                               rtol=1e-12)
    zc, _ = O.dem_elevation(dem, [(xs[3] + 0.05, ys[5] + 0.05)])
    d = dem["data"]
    assert zc[0] == pytest.approx((d[5, 3] + d[5, 4] + d[6, 3] + d[6, 4]) / 4, rel=1e-14)
    plane, xs, ys = _dem(lambda x, y: 1000.0 + 35.0 * x - 12.5 * y)
    q = rng.uniform([-2.0, -1.5], [2.0, 1.5], (50, 2))
    zp, inside = O.dem_elevation(plane, q)
    assert inside.all()
    np.testing.assert_allclose(zp, 1000.0 + 35.0 * q[:, 0] - 12.5 * q[:, 1], rtol=1e-13)
    zo, io = O.dem_elevation(plane, [(2.5, 0.0), (0.0, -1.6), (float("nan"), 0.0)])
    assert not io.any() and (zo == 0.0).all()


@pytest.mark.parametrize("zen,az", [(0.0, 0.0), (30.0, 45.0), (60.0, 200.0), (75.0, 300.0)])
def test_flat_and_planar_terrain(zen, az):
    d = _dir(zen, az)
    flat, _, _ = _dem(lambda x, y: 1263.0, nx=201, ny=201, x0=-10.0, y0=-10.0)
    dist, found, ez = O.ray_surface_exit(flat, [d], z_ref=963.0)
    assert found[0]
    assert dist[0] == pytest.approx(300.0 / d[2], abs=1e-8)
    assert ez[0] == pytest.approx(300.0, abs=1e-8)
    a, b, c = 1200.0, 40.0, -25.0   # metres, metres per km
    tilt, _, _ = _dem(lambda x, y: a + b * x + c * y, nx=201, ny=201, x0=-10.0, y0=-10.0)
    dist, found, _ = O.ray_surface_exit(tilt, [d], z_ref=963.0)
    s = (a - 963.0) / (d[2] - 1e-3 * (b * d[0] + c * d[1]))
    assert found[0] and dist[0] == pytest.approx(s, abs=1e-8)


def test_no_exit_cases():
    flat, _, _ = _dem(lambda x, y: 2000.0, nx=21, ny=21, x0=-0.5, y0=-0.5, dx=0.05, dy=0.05)
    # leaves the 1 km map (half width 0.5 km) before reaching 1037 m of height
    dist, found, ez = O.ray_surface_exit(flat, [_dir(60.0, 0.0), [0.0, 0.0, -1.0], [0.6, 0.8, 0.0]],
                                         z_ref=963.0)
    assert not found.any() and np.isnan(dist).all() and np.isnan(ez).all()
    # a terrain below the detector: the first step is already above it
    low, _, _ = _dem(lambda x, y: 900.0)
    dist, found, _ = O.ray_surface_exit(low, [_dir(20.0, 10.0)], z_ref=963.0)
    assert found[0] and 0.0 <= dist[0] <= 40.0

"""GPU parity of the topography ray exit (dp_ray_surface_exit,
ray_surface_exit_distance over map_elevation) against the CPU oracle:
distances, flags and exit heights bit-exact on a rough synthetic terrain,
with rays that cross, rays that leave the map, rays pointing down, rays
that run out of march steps, and ray counts that do not fill a wave.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp

pytestmark = pytest.mark.gpu


def _terrain(seed, n=129, half_km=6.0):
    """A deterministic mountain: smooth bumps plus seeded roughness (m)."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(-half_km, half_km, n)
    X, Y = np.meshgrid(xs, xs)            # X[iy, ix]
    z = 1150.0 + 500.0 * np.exp(-((X - 0.8) ** 2 + (Y + 0.4) ** 2) / 3.0)
    z += 180.0 * np.exp(-((X + 2.0) ** 2 + (Y - 1.5) ** 2) / 0.8)
    z += rng.normal(0.0, 6.0, z.shape)
    d = xs[1] - xs[0]
    return dp.ElevationMap(n, n, -half_km, -half_km, d, d, z)


def _dirs(n, seed):
    rng = np.random.default_rng(seed)
    zen = np.radians(rng.uniform(0.0, 88.0, n))
    az = np.radians(rng.uniform(0.0, 360.0, n))
    d = np.stack([np.sin(zen) * np.sin(az), np.sin(zen) * np.cos(az), np.cos(zen)], axis=1)
    d[::17] *= -1.0                       # some rays point down
    return d


def test_ray_exit_bit_exact(ctx):
    emap = _terrain(3)
    dirs = _dirs(4096, 5)
    print(f"dem ray exit: {len(dirs)} rays on a {emap.nx}x{emap.ny} map", flush=True)
    dist, found, ez = dp.ray_surface_exit_distance(emap, dirs, 963.0, ctx=ctx)
    od, of, oz = O.ray_surface_exit(dict(nx=emap.nx, ny=emap.ny, x0=emap.x0, y0=emap.y0,
                                         dx=emap.dx, dy=emap.dy, data=emap.data), dirs,
                                    z_ref=963.0)
    assert np.array_equal(found, of)
    assert np.array_equal(np.isnan(dist), np.isnan(od))
    assert np.array_equal(dist[found], od[of]) and np.array_equal(ez[found], oz[of])
    assert 0.5 < found.mean() < 0.99      # crossings, map exits and downward rays all occur


@pytest.mark.parametrize("n,step,max_d", [(1, 40.0, 25_000.0), (77, 40.0, 25_000.0),
                                           (301, 13.0, 900.0), (130, 40.0, 85.0)])
def test_ray_exit_ragged_and_short_march(ctx, n, step, max_d):
    """Ray counts that do not fill a wave, other step lengths, and march
    limits short enough that rays run out of steps before crossing."""
    emap = _terrain(4)
    dirs = _dirs(n, 6 + n)
    dist, found, ez = dp.ray_surface_exit_distance(emap, dirs, 963.0, step_m=step,
                                                   max_distance_m=max_d, ctx=ctx)
    od, of, oz = O.ray_surface_exit(dict(nx=emap.nx, ny=emap.ny, x0=emap.x0, y0=emap.y0,
                                         dx=emap.dx, dy=emap.dy, data=emap.data), dirs,
                                    z_ref=963.0, step=step, max_distance=max_d)
    assert np.array_equal(found, of)
    assert np.array_equal(np.isnan(dist), np.isnan(od))
    assert np.array_equal(dist[found], od[of]) and np.array_equal(ez[found], oz[of])
    if max_d < 100.0:
        assert found.mean() < 0.5         # most rays stop at the march limit


def test_ray_exit_empty(ctx):
    emap = _terrain(4)
    dist, found, ez = dp.ray_surface_exit_distance(emap, np.zeros((0, 3)), 963.0, ctx=ctx)
    assert dist.shape == (0,) and found.shape == (0,) and ez.shape == (0,)


def test_bad_map(ctx):
    emap = dp.ElevationMap(1, 4, 0.0, 0.0, 0.1, 0.1, np.zeros((4, 1)))
    with pytest.raises(dp.ArgumentError):
        dp.ray_surface_exit_distance(emap, [[0.0, 0.0, 1.0]], 963.0, ctx=ctx)

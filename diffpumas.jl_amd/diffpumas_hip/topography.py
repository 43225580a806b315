"""Topography: where rays from the detector leave a gridded elevation map.

Host mirror of ElevationMap / map_elevation (src/Turtle.jl:469-553) and of
ray_surface_exit_distance (examples/lvd_tomography.jl:219-252).  The search
runs on the GPU (dp_ray_surface_exit); there is no CPU path.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError

SURFACE_SEARCH_STEP_M = 40.0   # examples/lvd_tomography.jl:147


@dataclass
class ElevationMap:
    """ElevationMap (src/Turtle.jl:469-480): nx x ny nodes from (x0, y0) with
    spacings (dx, dy) in map units; data[iy, ix] is the elevation (m) of node
    (ix, iy)."""
    nx: int
    ny: int
    x0: float
    y0: float
    dx: float
    dy: float
    data: np.ndarray

    def c(self):
        d = np.asfortranarray(self.data, dtype=np.float64)   # storage index iy + ny * ix
        if d.shape != (self.ny, self.nx):
            raise ArgumentError("data must have shape (ny, nx)")
        return _lib.ElevationMap(int(self.nx), int(self.ny), float(self.x0), float(self.y0),
                                 float(self.dx), float(self.dy), d.ctypes.data), d


def ray_surface_exit_distance(emap, dirs, detector_elevation, step_m=SURFACE_SEARCH_STEP_M,
                              max_distance_m=25_000.0, xy_scale=1e-3, ctx=None):
    """ray_surface_exit_distance for every row of dirs (n x 3 unit vectors from
    the detector): (distance m, found, exit_z m); NaN / False where the ray
    points down or leaves the map first.  Map coordinates are the local
    metres times xy_scale (the reference's maps are in km)."""
    ctx = ctx or _lib.default_context()
    d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float64).reshape(-1, 3))
    n = d.shape[0]
    m, keep = emap.c()
    dist, ez = np.zeros(n), np.zeros(n)
    found = np.zeros(n, np.uint8)
    _lib.check(ctx.lib.dp_ray_surface_exit(ctx.h, C.byref(m), float(xy_scale),
                                           float(detector_elevation), n, _lib.ptr(d),
                                           float(step_m), float(max_distance_m), _lib.ptr(dist),
                                           _lib.ptr(found, C.c_uint8), _lib.ptr(ez)))
    return dist, found.astype(bool), ez

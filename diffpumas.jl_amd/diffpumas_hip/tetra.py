"""Tetrahedral sensitivity volumes: host mirror of TetraVolume and
trace_path(::TetraVolume) (examples/lvd_tomography.jl:172-179, 415-451,
694-736).  Neighbour tables are built on the host (dp_tet_neighbors, setup);
the ray walks run on the GPU (dp_trace_tetra); there is no CPU path.
Indices are 0-based; a missing neighbour is -1.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError, f64, i32, ptr
from .tomography import DirectionalPaths, direction_vector


def build_tetra_neighbors(tets):
    """build_tetra_neighbors: neighbors[t, f] is the tet across face f of tet
    t (the face without vertex 3 - f), -1 when no other tet shares it."""
    t = i32(np.asarray(tets).reshape(-1, 4))
    nbr = np.zeros_like(t)
    _lib.check(_lib.load().dp_tet_neighbors(len(t), ptr(t, C.c_int32), ptr(nbr, C.c_int32)))
    return nbr


@dataclass
class TetraVolume:
    """TetraVolume: mesh points (n, 3), tets (m, 4) and their neighbour table."""
    points: np.ndarray
    tets: np.ndarray
    neighbors: np.ndarray = None

    def __post_init__(self):
        self.points = f64(np.asarray(self.points).reshape(-1, 3))
        self.tets = i32(np.asarray(self.tets).reshape(-1, 4))
        if self.neighbors is None:
            self.neighbors = build_tetra_neighbors(self.tets)
        self.neighbors = i32(np.asarray(self.neighbors).reshape(-1, 4))
        if self.neighbors.shape != self.tets.shape:
            raise ArgumentError("neighbors must have one row per tet")

    @property
    def n_cells(self):
        return len(self.tets)

    def c_mesh(self):
        return _lib.TetMesh(len(self.points), len(self.tets), ptr(self.points),
                            ptr(self.tets, C.c_int32), ptr(self.neighbors, C.c_int32))


def trace_tetra_rays(volume, origins, directions, surface_distance, ctx=None):
    """The tetrahedral walk for many rays: (row_ptr, cells, lengths, traveled)."""
    ctx = ctx or _lib.default_context()
    o = f64(np.asarray(origins).reshape(-1, 3))
    d = f64(np.asarray(directions).reshape(-1, 3))
    sd = f64(np.asarray(surface_distance).reshape(-1))
    n = len(d)
    if len(o) != n or len(sd) != n:
        raise ArgumentError("origins, directions and surface distances must match")
    m = volume.c_mesh()
    counts = np.zeros(n, np.int32)
    traveled = np.zeros(n)
    _lib.check(ctx.lib.dp_trace_tetra(ctx.h, C.byref(m), n, ptr(o), ptr(d), ptr(sd),
                                      ptr(counts, C.c_int32), None, None, None, ptr(traveled)))
    row_ptr = np.zeros(n + 1, np.int32)
    np.cumsum(counts, out=row_ptr[1:])
    ns = int(row_ptr[-1])
    cells = np.zeros(max(ns, 1), np.int32)
    lens = np.zeros(max(ns, 1))
    _lib.check(ctx.lib.dp_trace_tetra(ctx.h, C.byref(m), n, ptr(o), ptr(d), ptr(sd), None,
                                      ptr(row_ptr, C.c_int32), ptr(cells, C.c_int32), ptr(lens),
                                      ptr(traveled)))
    return row_ptr, cells[:ns], lens[:ns], traveled


def trace_tetra_paths(volume, zeniths, azimuths, surface_distance, exit_z, detectors=None,
                      ctx=None):
    """trace_path(::TetraVolume) for many (zenith, azimuth) bins from the
    detector (the origin of the mesh frame unless detectors[n, 3] is given).
    surface_distance and exit_z come from ray_surface_exit_distance; a bin
    whose surface distance is not finite is an invalid path, as in the
    reference."""
    zen, az = f64(zeniths), f64(azimuths)
    sd, ez = f64(surface_distance), f64(exit_z)
    dirs = np.array([direction_vector(z, a) for z, a in zip(zen, az)]).reshape(-1, 3)
    org = np.zeros_like(dirs) if detectors is None else f64(detectors).reshape(-1, 3)
    valid = np.isfinite(sd)
    rp, cells, lens, traveled = trace_tetra_rays(volume, org, dirs, np.where(valid, sd, 0.0),
                                                 ctx=ctx)
    remaining = np.where(valid, np.maximum(0.0, sd - traveled), np.nan)
    return DirectionalPaths(rp, cells, lens, zen, az, remaining, np.where(valid, ez, np.nan),
                            valid.astype(np.uint8), distance_in_volume=np.where(valid, traveled,
                                                                                np.nan),
                            surface_distance=sd)

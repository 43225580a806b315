"""Per-voxel-density CSDA through a tessellated cube, on the GPU.

Host mirror of examples/diff_flux_cube.jl: CubeGeometry (:85-129),
trace_cube_path (:304-364), compute_flux_from_segments (:374-468) and the
flux/gradient pair of compute_flux_gradient_cube (:489-550).  Cell indices
are 0-based (Julia's are 1-based).
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import f64, i32, ptr


class CubeGeometry:
    """CubeGeometry{Float64} (examples/diff_flux_cube.jl:85-129)."""

    def __init__(self, rock_thickness, nx, ny, nz, rock_material, air_material, width=200.0,
                 width_y=None):
        wy = width if width_y is None else width_y
        self.x_min, self.x_max = -width / 2, width / 2
        self.y_min, self.y_max = -wy / 2, wy / 2
        self.z_min, self.z_max = 0.0, float(rock_thickness)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx = (self.x_max - self.x_min) / self.nx
        self.dy = (self.y_max - self.y_min) / self.ny
        self.dz = (self.z_max - self.z_min) / self.nz
        self.rock_material = int(rock_material)
        self.air_material = int(air_material)

    @property
    def n_cells(self):
        return self.nx * self.ny * self.nz

    def c_grid(self):
        return _lib.CubeGrid(self.x_min, self.x_max, self.y_min, self.y_max, self.z_min,
                             self.z_max, self.nx, self.ny, self.nz, self.dx, self.dy, self.dz)

    def as_array(self):
        return np.array([self.x_min, self.x_max, self.y_min, self.y_max, self.z_min, self.z_max,
                         self.nx, self.ny, self.nz, self.dx, self.dy, self.dz], dtype=np.float64)


def ray_direction(elevation_deg, azimuth_deg):
    """Upward tracing direction of trace_cube_path (:320-328)."""
    el = math.radians(elevation_deg)
    az = math.radians(azimuth_deg)
    cos_t = math.sin(el)
    sin_t = math.cos(el)
    return (sin_t * math.sin(az), sin_t * math.cos(az), cos_t), cos_t


class RayPaths:
    """CSR segment lists of a batch of rays (rows = rays)."""

    def __init__(self, row_ptr, cell, length, cos_theta, z_exit):
        self.row_ptr = i32(row_ptr)
        self.cell = i32(cell)
        self.length = f64(length)
        self.cos_theta = f64(cos_theta)
        self.z_exit = f64(z_exit)

    @property
    def n_rays(self):
        return len(self.row_ptr) - 1

    @property
    def n_segments(self):
        return int(self.row_ptr[-1])

    @property
    def max_segments(self):
        return int(np.max(np.diff(self.row_ptr))) if self.n_rays else 0

    def c_paths(self):
        return _lib.Paths(self.n_rays, self.n_segments, self.max_segments, ptr(self.row_ptr, C.c_int32),
                          ptr(self.cell, C.c_int32), ptr(self.length), ptr(self.cos_theta),
                          ptr(self.z_exit))


def trace_cube_paths(geo, start_xy, directions, cos_theta, ctx=None):
    """trace_cube_path for many rays in two GPU launches (count, fill)."""
    ctx = ctx or _lib.default_context()
    sxy = f64(start_xy).reshape(-1, 2)
    d = f64(directions).reshape(-1, 3)
    n = len(d)
    counts = np.zeros(n, dtype=np.int32)
    z_exit = np.zeros(n)
    grid = geo.c_grid()
    _lib.check(ctx.lib.dp_trace_cube(ctx.h, C.byref(grid), n, ptr(sxy), ptr(d),
                                     ptr(counts, C.c_int32), None, None, None, ptr(z_exit)))
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(counts, out=row_ptr[1:])
    ns = int(row_ptr[-1])
    cells = np.zeros(max(ns, 1), dtype=np.int32)
    lens = np.zeros(max(ns, 1))
    _lib.check(ctx.lib.dp_trace_cube(ctx.h, C.byref(grid), n, ptr(sxy), ptr(d), None,
                                     ptr(row_ptr, C.c_int32), ptr(cells, C.c_int32), ptr(lens),
                                     None))
    return RayPaths(row_ptr, cells[:ns], lens[:ns], cos_theta, z_exit)


def trace_edge_paths(x_edges, y_edges, z_edges, origins, directions, surface_distance, ctx=None):
    """trace_path(::VoxelVolume) (examples/lvd_tomography.jl:738-779) for many
    rays.  Returns (RayPaths-like CSR arrays, traveled, remaining_rock)."""
    ctx = ctx or _lib.default_context()
    xe, ye, ze = f64(x_edges), f64(y_edges), f64(z_edges)
    o = f64(origins).reshape(-1, 3)
    d = f64(directions).reshape(-1, 3)
    sd = f64(surface_distance)
    n = len(d)
    counts = np.zeros(n, dtype=np.int32)
    traveled = np.zeros(n)
    _lib.check(ctx.lib.dp_trace_edges(ctx.h, ptr(xe), len(xe), ptr(ye), len(ye), ptr(ze), len(ze),
                                      n, ptr(o), ptr(d), ptr(sd), ptr(counts, C.c_int32), None,
                                      None, None, ptr(traveled)))
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(counts, out=row_ptr[1:])
    ns = int(row_ptr[-1])
    cells = np.zeros(max(ns, 1), dtype=np.int32)
    lens = np.zeros(max(ns, 1))
    _lib.check(ctx.lib.dp_trace_edges(ctx.h, ptr(xe), len(xe), ptr(ye), len(ye), ptr(ze), len(ze),
                                      n, ptr(o), ptr(d), ptr(sd), None, ptr(row_ptr, C.c_int32),
                                      ptr(cells, C.c_int32), ptr(lens), None))
    remaining = np.maximum(0.0, sd - traveled)
    return row_ptr, cells[:ns], lens[:ns], traveled, remaining


def csda_cube_flux(physics, geo, paths, densities, energies, weights, charges, scale=1.0,
                   primary_altitude=3e4, jacobian=True, ctx=None):
    """Per-ray CSDA flux  flux[r] = scale * sum_n 2 w_n flux(r, E_n, q_n)  and,
    with jacobian=True, d flux[r]/d rho per segment (aligned with the CSR
    arrays) and through the remaining-rock segment (cell 0)."""
    ctx = ctx or _lib.default_context()
    h = physics.device_handle(ctx)
    rho = f64(densities)
    ne, nw, nq = f64(energies), f64(weights), f64(charges)
    nodes = _lib.CsdaNodes(geo.rock_material, geo.air_material, geo.z_max, float(primary_altitude),
                           float(scale), len(ne), ptr(ne), ptr(nw), ptr(nq))
    cp = paths.c_paths()
    flux = np.zeros(paths.n_rays)
    jac = np.zeros(max(paths.n_segments, 1)) if jacobian else None
    jrem = np.zeros(paths.n_rays) if jacobian else None
    _lib.check(ctx.lib.dp_csda_cube(ctx.h, h, C.byref(cp), C.byref(nodes), ptr(rho), len(rho),
                                    ptr(flux), ptr(jac), ptr(jrem)))
    if not jacobian:
        return flux
    return flux, jac[:paths.n_segments], jrem


def csda_cube_vjp(paths, jac, jac_rem, flux_bar, n_cells, ctx=None):
    """grad[c] = sum_r flux_bar[r] d flux[r]/d rho[c] (deterministic order)."""
    ctx = ctx or _lib.default_context()
    cp = paths.c_paths()
    j = f64(jac) if len(jac) else np.zeros(1)
    grad = np.zeros(n_cells)
    _lib.check(ctx.lib.dp_csda_cube_vjp(ctx.h, C.byref(cp), ptr(j), ptr(f64(jac_rem)),
                                        ptr(f64(flux_bar)), int(n_cells), ptr(grad)))
    return grad


def compute_flux_gradient_cube(physics, geo, cell_densities, samples, primary_altitude=3e4,
                               ctx=None):
    """compute_flux_gradient_cube (examples/diff_flux_cube.jl:489-550) given
    the traced samples: `samples` is a list of (elevation, azimuth, start_x,
    start_y, energy, charge, weight).  Returns (flux, dense gradient)."""
    n = len(samples)
    dirs, cos_t, sxy = [], [], []
    for el, az, x0, y0, *_ in samples:
        d, c = ray_direction(el, az)
        dirs.append(d)
        cos_t.append(c)
        sxy.append((x0, y0))
    paths = trace_cube_paths(geo, sxy, dirs, cos_t, ctx=ctx)
    # one node per traced sample: evaluate each ray with its own node by
    # giving rays their own energy (batched as one-node calls per distinct ray)
    flux_tot = 0.0
    grad = np.zeros(geo.n_cells)
    for r, (_, _, _, _, e, q, w) in enumerate(samples):
        sub = RayPaths(np.array([0, paths.row_ptr[r + 1] - paths.row_ptr[r]]),
                       paths.cell[paths.row_ptr[r]:paths.row_ptr[r + 1]],
                       paths.length[paths.row_ptr[r]:paths.row_ptr[r + 1]],
                       paths.cos_theta[r:r + 1], paths.z_exit[r:r + 1])
        f, j, jr = csda_cube_flux(physics, geo, sub, cell_densities, [e], [w], [q],
                                  scale=1.0 / n, primary_altitude=primary_altitude, ctx=ctx)
        flux_tot += f[0]
        grad += csda_cube_vjp(sub, j, jr, np.ones(1), geo.n_cells, ctx=ctx)
    return flux_tot, grad

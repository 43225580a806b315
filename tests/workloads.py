"""Seeded synthetic workloads shared by tests and bench.py.

Shapes follow the reference's benchmark configurations: a 10 m voxel cube
with detectors on its lower face and upward rays between 15 and 75 degrees
elevation, log-spaced energy nodes, lognormal or uniform densities.
"""
import math

import numpy as np


def cube_rays(geo, n_detectors, n_dirs, seed, el_range=(15.0, 75.0), az_range=(0.0, 360.0)):
    """Detectors on z = 0 inside the lower face; directions uniform in
    azimuth and elevation.  Returns (start_xy, dirs, cos_theta)."""
    rng = np.random.default_rng(seed)
    span_x = geo.x_max - geo.x_min
    span_y = geo.y_max - geo.y_min
    det = np.stack([geo.x_min + span_x * rng.uniform(0.25, 0.75, n_detectors),
                    geo.y_min + span_y * rng.uniform(0.25, 0.75, n_detectors)], axis=1)
    el = rng.uniform(*el_range, size=(n_detectors, n_dirs))
    az = rng.uniform(*az_range, size=(n_detectors, n_dirs))
    sxy, dirs, cos_t = [], [], []
    for d in range(n_detectors):
        for k in range(n_dirs):
            e = math.radians(el[d, k])
            a = math.radians(az[d, k])
            c = math.sin(e)
            s = math.cos(e)
            dirs.append((s * math.sin(a), s * math.cos(a), c))
            cos_t.append(c)
            sxy.append(det[d])
    return np.array(sxy), np.array(dirs), np.array(cos_t)


def energy_nodes(n, emin, emax):
    """csda_energy_quadrature (src/Tomography.jl:263-274): log-midpoint
    nodes, charge 0 (charge-summed), weight 0.5 * E * span."""
    lo = math.log(emin)
    span = math.log(emax) - lo
    e = np.array([math.exp(lo + span * (i - 0.5) / n) for i in range(1, n + 1)])
    return e, 0.5 * e * span, np.zeros(n)


def uniform_densities(n_cells, seed, lo=2000.0, hi=3000.0):
    return np.random.default_rng(seed).uniform(lo, hi, n_cells)


def lognormal_cavity_densities(geo, seed, mean=2650.0, sigma=0.1, cavity=1800.0):
    rng = np.random.default_rng(seed)
    n = geo.n_cells
    rho = mean * np.exp(sigma * rng.standard_normal(n) - 0.5 * sigma * sigma)
    i = np.arange(n)
    ix = i % geo.nx
    iy = (i // geo.nx) % geo.ny
    iz = i // (geo.nx * geo.ny)
    cx, cy, cz = 0.5 * geo.nx, 0.45 * geo.ny, 0.6 * geo.nz
    rx, ry, rz = 0.15 * geo.nx, 0.1 * geo.ny, 0.12 * geo.nz
    inside = ((ix - cx) / rx) ** 2 + ((iy - cy) / ry) ** 2 + ((iz - cz) / rz) ** 2 <= 1.0
    rho[inside] = cavity
    return rho


def diamond_square(n_pow, relief, seed, roughness=0.55):
    """Seeded diamond-square height field on (2^n_pow + 1)^2 nodes, shifted and
    scaled to span [0, relief] m.  h[iy, ix]."""
    rng = np.random.default_rng(seed)
    n = 2 ** n_pow + 1
    h = np.zeros((n, n))
    h[0, 0], h[0, -1], h[-1, 0], h[-1, -1] = rng.uniform(-1.0, 1.0, 4)
    step, amp = n - 1, 1.0
    while step > 1:
        half = step // 2
        for i in range(half, n, step):
            for j in range(half, n, step):
                h[i, j] = 0.25 * (h[i - half, j - half] + h[i - half, j + half] +
                                  h[i + half, j - half] + h[i + half, j + half])
                h[i, j] += amp * rng.uniform(-1.0, 1.0)
        for i in range(0, n, half):
            for j in range((i + half) % step, n, step):
                s, k = 0.0, 0
                for a, b in ((i - half, j), (i + half, j), (i, j - half), (i, j + half)):
                    if 0 <= a < n and 0 <= b < n:
                        s += h[a, b]
                        k += 1
                h[i, j] = s / k + amp * rng.uniform(-1.0, 1.0)
        step, amp = half, amp * roughness
    return (h - h.min()) / (h.max() - h.min()) * relief


def kuhn_mesh(nx, ny, nz, h, origin=(0.0, 0.0, 0.0), jitter=0.0, seed=0):
    """Conforming tetrahedral mesh of an nx x ny x nz block of cubes of side h:
    each cube is cut into the 6 Kuhn tetrahedra (corner 000 to corner 111 along
    the 6 axis orders).  Tet 6 c + k lies in cube c = i + nx (j + ny k).
    jitter > 0 moves interior vertices by up to jitter * h (seeded).
    Returns (points[n, 3], tets[n_tets, 4]) with 0-based indices."""
    import itertools
    ox, oy, oz = origin
    gx, gy, gz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")
    vid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    pts = np.stack([ox + h * gx.ravel(), oy + h * gy.ravel(), oz + h * gz.ravel()], axis=1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        inner = ((gx > 0) & (gx < nx) & (gy > 0) & (gy < ny) & (gz > 0) & (gz < nz)).ravel()
        pts[inner] += rng.uniform(-jitter * h, jitter * h, (inner.sum(), 3))
    tets = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for perm in itertools.permutations(range(3)):
                    v = [i, j, k]
                    ids = [vid(*v)]
                    for ax in perm:
                        v[ax] += 1
                        ids.append(vid(*v))
                    tets.append(ids)
    return pts, np.array(tets, np.int32)


def inverse_workload(dp, ctx, n_pow=8, spacing=20.0, nv=(96, 96, 48), vsize=10.0, n_dirs=4096,
                     dets=((-150.0, -150.0), (150.0, -150.0), (-150.0, 150.0), (150.0, 150.0)),
                     z_det=963.0, seed=11, zen_max=60.0):
    """The Gauss-Newton inverse workload: a (2^n_pow + 1)^2-node diamond-square
    DEM (400 m relief on a dome above 1200 m), detectors at z_det under a
    block of nv voxels of vsize m, n_dirs directions per detector (zenith
    U[0, zen_max] deg); DEM exits and voxel traversal on the GPU.  Returns the
    traced paths (bins whose ray leaves through the terrain), the edges and
    the cell count."""
    from diffpumas_hip.tomography import direction_vector
    n = 2 ** n_pow + 1
    xs = (np.arange(n) - (n - 1) / 2) * spacing
    X, Y = np.meshgrid(xs, xs)
    z = 1200.0 + diamond_square(n_pow, 400.0, seed) + 300.0 * np.exp(
        -(X ** 2 + Y ** 2) / (0.35 * xs[-1]) ** 2)
    rng = np.random.default_rng(seed + 1)
    zen_all, az_all, surf_all, ez_all, org_all = [], [], [], [], []
    for xd, yd in dets:
        emap = dp.ElevationMap(n, n, (xs[0] - xd) * 1e-3, (xs[0] - yd) * 1e-3, spacing * 1e-3,
                               spacing * 1e-3, z)
        zen = rng.uniform(0.0, zen_max, n_dirs)
        az = rng.uniform(0.0, 360.0, n_dirs)
        dirs = np.array([direction_vector(a, b) for a, b in zip(zen, az)])
        dist, found, ez = dp.ray_surface_exit_distance(emap, dirs, z_det, ctx=ctx)
        zen_all.append(zen[found]); az_all.append(az[found])
        surf_all.append(dist[found]); ez_all.append(ez[found])
        org_all.append(np.tile([xd, yd, z_det], (int(found.sum()), 1)))
    zen, az = np.concatenate(zen_all), np.concatenate(az_all)
    nx, ny, nz = nv
    xe = (np.arange(nx + 1) - nx / 2) * vsize
    ye = (np.arange(ny + 1) - ny / 2) * vsize
    ze = z_det - vsize + np.arange(nz + 1) * vsize
    paths = dp.trace_voxel_paths(xe, ye, ze, zen, az, np.concatenate(surf_all),
                                 np.concatenate(ez_all), detectors=np.concatenate(org_all),
                                 ctx=ctx)
    return paths, (xe, ye, ze), nx * ny * nz

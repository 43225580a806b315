"""Oracle transport and traversal checks on the CPU (no GPU).

Traversal: brute-force point sampling, conservation of length, closed forms.
CSDA: knot-exact closed form, central finite differences of the oracle's
forward-mode derivatives (per-voxel density and flat-geometry stepping).
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip import _lib as L
from workloads import cube_rays, energy_nodes, uniform_densities


def small_cube(rock=0, air=1):
    return dp.CubeGeometry(60.0, 6, 5, 4, rock, air, width=80.0, width_y=70.0)


def _cell_of(geo, p):
    i = min(max(int(math.floor((p[0] - geo.x_min) / geo.dx)), 0), geo.nx - 1)
    j = min(max(int(math.floor((p[1] - geo.y_min) / geo.dy)), 0), geo.ny - 1)
    k = min(max(int(math.floor((p[2] - geo.z_min) / geo.dz)), 0), geo.nz - 1)
    return i + j * geo.nx + k * geo.nx * geo.ny


def test_trace_cube_bruteforce():
    """Each segment is recorded in the cell that contains its start point
    (floor rule of examples/diff_flux_cube.jl:132-144) and its length is the
    distance from that point to the nearest face of that cell further than
    1e-10 (:194-280).  Recomputed here in numpy, point by point."""
    geo = small_cube()
    sxy, dirs, cos_t = cube_rays(geo, 4, 64, seed=3)
    rp, cells, lens, ze = O.trace_cube(geo, sxy, dirs)
    geometric = total = 0
    for r in range(len(dirs)):
        p = np.array([sxy[r][0], sxy[r][1], 0.0])
        u = np.asarray(dirs[r])
        for s in range(rp[r], rp[r + 1]):
            c = _cell_of(geo, p)
            assert cells[s] == c
            i, j, k = c % geo.nx, (c // geo.nx) % geo.ny, c // (geo.nx * geo.ny)
            lo = np.array([geo.x_min + i * geo.dx, geo.y_min + j * geo.dy, geo.z_min + k * geo.dz])
            hi = lo + np.array([geo.dx, geo.dy, geo.dz])
            ts = [((hi[a] if u[a] > 0 else lo[a]) - p[a]) / u[a] for a in range(3) if abs(u[a]) > 1e-10]
            t = min(x for x in ts if x > 1e-10)
            assert lens[s] == pytest.approx(t, rel=1e-12, abs=1e-12)
            geometric += _cell_of(geo, p + 0.5 * lens[s] * u) == c
            total += 1
            p = p + lens[s] * u
        assert ze[r] == pytest.approx(p[2], rel=1e-12, abs=1e-9)
        assert rp[r + 1] > rp[r]
    # most segments are also geometrically inside their cell
    assert geometric > 0.5 * total


def test_trace_cube_lower_face_quirk():
    """examples/diff_flux_cube.jl:132-144 maps a point on a cell's lower face
    to that cell, so after a crossing in the -x direction the next segment is
    still attributed to the cell just left; the tracer reproduces this."""
    geo = small_cube()
    u = np.array([-0.6, 0.0, 0.8])
    rp, cells, lens, ze = O.trace_cube(geo, [[-6.0, 1.0]], [u])
    # first segment ends on the x = -13.33 face (cell i=2), the second starts
    # there and is recorded in the same cell although it lies in i=1
    i0 = int((-6.0 - geo.x_min) // geo.dx)
    assert cells[0] % geo.nx == i0 and cells[1] % geo.nx == i0
    p = np.array([-6.0, 1.0, 0.0]) + (lens[0] + 0.5 * lens[1]) * u
    assert _cell_of(geo, p) % geo.nx == i0 - 1


def test_trace_cube_vertical_closed_form():
    geo = small_cube()
    rp, cells, lens, ze = O.trace_cube(geo, [[1.0, 2.0]], [[0.0, 0.0, 1.0]])
    assert rp[-1] == geo.nz
    assert np.allclose(lens, geo.dz, rtol=1e-15)
    assert ze[0] == pytest.approx(geo.z_max)
    i, j = int((1.0 - geo.x_min) // geo.dx), int((2.0 - geo.y_min) // geo.dy)
    assert list(cells) == [i + j * geo.nx + k * geo.nx * geo.ny for k in range(geo.nz)]


def test_trace_edges_bruteforce():
    rng = np.random.default_rng(5)
    xe = np.cumsum(np.r_[-50.0, rng.uniform(5, 15, 8)])
    ye = np.cumsum(np.r_[-40.0, rng.uniform(5, 15, 7)])
    ze = np.cumsum(np.r_[-1.0, rng.uniform(5, 20, 6)])
    n = 200
    el = np.radians(rng.uniform(15, 85, n))
    az = np.radians(rng.uniform(0, 360, n))
    dirs = np.stack([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az), np.sin(el)], 1)
    surface = rng.uniform(20, 120, n)
    origins = np.zeros((n, 3))
    rp, cells, lens, tr = O.trace_edges(xe, ye, ze, origins, dirs, surface)
    nx, ny = len(xe) - 1, len(ye) - 1
    for r in range(n):
        assert np.sum(lens[rp[r]:rp[r + 1]]) == pytest.approx(tr[r] - 1e-4, rel=1e-9, abs=1e-9)
        assert tr[r] <= surface[r] + 1e-9
        t = 1e-4
        for s in range(rp[r], rp[r + 1]):
            mid = (t + 0.5 * lens[s]) * dirs[r]
            i = np.searchsorted(xe, mid[0], side="right") - 1
            j = np.searchsorted(ye, mid[1], side="right") - 1
            k = np.searchsorted(ze, mid[2], side="right") - 1
            assert cells[s] == i + j * nx + k * nx * ny
            t += lens[s] + 1e-4


# --------------------------------------------------------------- CSDA
def test_csda_knot_closed_form(physics):
    """Range inversion that lands on a knot reproduces the knot energy and the
    tabulated stopping-power ratio (compute_flux_from_segments step)."""
    rock = physics.get_material_index("StandardRock")
    e = physics.data[rock, L.T_ENERGY]
    R = physics.data[rock, L.T_CSDA_RANGE]
    i, j = 120, 131
    x, _ = O.property_(physics, rock, L.T_CSDA_RANGE, [e[i]])
    assert x[0] == R[i]
    k, _ = O.kinetic(physics, rock, [R[i] + (R[j] - R[i])])
    assert k[0] == pytest.approx(e[j], rel=1e-14)


def _cube_problem(physics, n_det=2, n_dir=8, n_nodes=6, seed=11):
    geo = small_cube(physics.get_material_index("StandardRock"), physics.get_material_index("Air"))
    sxy, dirs, cos_t = cube_rays(geo, n_det, n_dir, seed=seed)
    rp, cells, lens, ze = O.trace_cube(geo, sxy, dirs)
    rho = uniform_densities(geo.n_cells, seed)
    ne, nw, nq = energy_nodes(n_nodes, 1.0, 1e4)
    return geo, rp, cells, lens, cos_t, ze, rho, ne, nw, nq


def test_cube_csda_gradient_vs_fd(physics):
    geo, rp, cells, lens, cos_t, ze, rho, ne, nw, nq = _cube_problem(physics)
    flux, grads = O.cube_csda(physics, geo, rp, cells, lens, cos_t, ze, rho, ne, nw, nq)
    assert np.all(flux > 0)
    h = 1e-3  # kg/m3, as in the reference's FD check
    checked = 0
    for r in range(0, len(flux), 3):
        for cell, g in list(grads[r].items())[:4]:
            rp_ = rho.copy(); rp_[cell] += h
            rm_ = rho.copy(); rm_[cell] -= h
            fp = O.cube_csda(physics, geo, rp, cells, lens, cos_t, ze, rp_, ne, nw, nq, gradient=False)[r]
            fm = O.cube_csda(physics, geo, rp, cells, lens, cos_t, ze, rm_, ne, nw, nq, gradient=False)[r]
            fd = (fp - fm) / (2 * h)
            assert g == pytest.approx(fd, rel=1e-5, abs=1e-12 * abs(flux[r])), (r, cell)
            checked += 1
    assert checked > 5


def test_cube_csda_reverse_matches_forward_seeding(physics):
    """The reverse-mode oracle (CPU baseline) agrees with forward seeding."""
    geo, rp, cells, lens, cos_t, ze, rho, ne, nw, nq = _cube_problem(physics, n_dir=12)
    flux, grads = O.cube_csda(physics, geo, rp, cells, lens, cos_t, ze, rho, ne, nw, nq)
    f2, jac, jrem = O.cube_csda_reverse(physics, geo.rock_material, geo.air_material, rp, cells,
                                        lens, cos_t, ze, geo.z_max, rho, ne, nw, nq)
    np.testing.assert_allclose(f2, flux, rtol=1e-13)
    for r in range(len(flux)):
        g = {}
        for s in range(rp[r], rp[r + 1]):
            g[int(cells[s])] = g.get(int(cells[s]), 0.0) + jac[s]
        if jrem[r]:
            g[0] = g.get(0, 0.0) + jrem[r]
        sc = max(abs(v) for v in grads[r].values())
        for c, v in grads[r].items():
            assert g.get(c, 0.0) == pytest.approx(v, rel=1e-10, abs=1e-14 * sc)


def test_flat_csda_single_step_fd(physics):
    """transport through a thin slab: derivative of the deterministic
    adaptive walk (src/Geometry.jl:534-672) against central differences."""
    rock, air = physics.get_material_index("StandardRock"), physics.get_material_index("Air")
    for thick, E in ((10.0, 10.0), (3.0, 120.0), (100.0, 30.0)):
        geo = [thick, 2650.0, -1, 0.0, 0.0, 3e4, 0.0, 0.0]
        geo = [thick, 2650.0, 0.0, 0.0, 3e4, 0, 0, 0]
        mats = [rock, air, -1]
        c = math.cos((90 - 45.0) * math.pi / 180)
        out = O.flat_csda(physics, geo, mats, E, c, 1.0)
        h = 1e-2
        gp = list(geo); gp[1] += h
        gm = list(geo); gm[1] -= h
        fd = (O.flat_csda(physics, gp, mats, E, c, 1.0)[0] - O.flat_csda(physics, gm, mats, E, c, 1.0)[0]) / (2 * h)
        assert out[0] > 0 and out[1] < 0
        assert out[1] == pytest.approx(fd, rel=5e-4)


def test_flat_csda_zero_rock(physics):
    """Zero rock: the density derivative vanishes (test_geometry_ad.jl:81-89)."""
    rock, air = physics.get_material_index("StandardRock"), physics.get_material_index("Air")
    out = O.flat_csda(physics, [0.0, 2650.0, 0.0, 0.0, 3e4, 0, 0, 0], [rock, air, -1], 10.0, 1.0, 1.0)
    assert out[0] > 0 and out[1] == 0.0


def test_mc_flat_csda_matches_surrogate(physics):
    """With straggling and scattering off the MC estimate equals its CSDA
    gradient surrogate (test_geometry_ad.jl:60-79: result.flux ≈ gradient_flux)."""
    rock, air = physics.get_material_index("StandardRock"), physics.get_material_index("Air")
    c = math.cos((90 - 45.0) * math.pi / 180)
    out = O.mc_flat_csda(physics, [100.0, 2650.0, 0.0, 0.0, 3e4, 0, 0, 0], [rock, air, -1],
                         1.0, 100.0, c, 500, 1)
    assert out[0] > 0 and out[1] > 0
    assert out[2] == pytest.approx(out[0], rel=1e-12)
    assert out[3] < 0

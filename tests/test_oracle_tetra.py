"""Pins the oracle's tetrahedral ray walk (trace_path(::TetraVolume),
examples/lvd_tomography.jl:694-736) and its pieces against what geometry
fixes.  Meshes are Kuhn subdivisions of a block of cubes (6 tets per cube,
conforming), plain or with jittered interior vertices; the reference's
meshes come from TetGen, which is not available here.

* neighbours: symmetric, exactly the boundary triangles of the block are
  open (2 per boundary square);
* point in tet: agrees with a numpy barycentric solve away from faces; every
  centroid lies in its own tet and in no other; on the plain mesh the first
  tet holding a point lies in the point's cube;
* walk: segment k spans [a_k, a_k + len_k] along the ray with
  a_0 = POINT_EPS and a_{k+1} = a_k + len_k + POINT_EPS (the nudge after each
  face crossing), and the midpoint of every span lies in the segment's tet;
  the walk ends at min(surface distance, exit from the block); on the plain
  mesh the per-cube sums of lengths equal the voxel walk's cell lengths to
  within the nudges; a detector outside the mesh gives no segment.
"""
import numpy as np
import pytest

import _oracle as O
from workloads import kuhn_mesh

EPS = 1e-4


def _bary_inside(pts, tet, p, tol):
    a = pts[tet[0]]
    M = np.stack([pts[tet[1]] - a, pts[tet[2]] - a, pts[tet[3]] - a], axis=1)
    b = np.linalg.solve(M, p - a)
    return b.min() >= -tol and b.sum() <= 1 + tol, b


def _box_exit(o, d, lo, hi):
    t = np.inf
    for ax in range(3):
        if d[ax] > 0:
            t = min(t, (hi[ax] - o[ax]) / d[ax])
        elif d[ax] < 0:
            t = min(t, (lo[ax] - o[ax]) / d[ax])
    return t


def _rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = np.column_stack([rng.uniform(lo[0] + 1, hi[0] - 1, n), rng.uniform(lo[1] + 1, hi[1] - 1, n),
                         np.full(n, lo[2] + 0.5)])
    zen = np.radians(rng.uniform(0, 70, n))
    az = np.radians(rng.uniform(0, 360, n))
    d = np.column_stack([np.sin(zen) * np.sin(az), np.sin(zen) * np.cos(az), np.cos(zen)])
    return o, d


def test_neighbours():
    for nx, ny, nz in ((3, 2, 2), (4, 4, 3)):
        pts, tets = kuhn_mesh(nx, ny, nz, 10.0, jitter=0.05, seed=1)
        nbr = O.tet_neighbors(tets)
        assert (nbr < 0).sum() == 4 * (nx * ny + ny * nz + nx * nz)
        for t in range(len(tets)):
            for f in range(4):
                u = nbr[t, f]
                if u >= 0:
                    shared = set(tets[t]) - {tets[t][3 - f]}
                    assert shared <= set(tets[u]) and t in nbr[u]


@pytest.mark.parametrize("jitter", [0.0, 0.08])
def test_point_in_tet(jitter):
    pts, tets = kuhn_mesh(3, 3, 2, 10.0, jitter=jitter, seed=2)
    cen = pts[tets].mean(axis=1)
    inside = O.point_in_tet(pts, tets, cen)
    assert np.array_equal(inside, np.eye(len(tets), dtype=bool))
    rng = np.random.default_rng(3)
    p = rng.uniform([0, 0, 0], [30, 30, 20], (400, 3))
    got = O.point_in_tet(pts, tets, p)
    for t in range(len(tets)):
        for i in range(len(p)):
            ins, b = _bary_inside(pts, tets[t], p[i], 0.0)
            margin = min(b.min(), 1 - b.sum())
            if abs(margin) > 1e-6:
                assert got[t, i] == (margin > 0)
    if jitter == 0.0:
        first = got.argmax(axis=0)
        cube = (np.floor(p[:, 0] / 10) + 3 * (np.floor(p[:, 1] / 10) + 3 * np.floor(p[:, 2] / 10)))
        assert np.array_equal(first // 6, cube.astype(int))


@pytest.mark.parametrize("jitter", [0.0, 0.08])
def test_walk_spans_and_exit(jitter):
    nx, ny, nz, h = 4, 4, 3, 10.0
    pts, tets = kuhn_mesh(nx, ny, nz, h, jitter=jitter, seed=4)
    nbr = O.tet_neighbors(tets)
    lo, hi = np.zeros(3), np.array([nx * h, ny * h, nz * h])
    o, d = _rays(60, 5, lo, hi)
    surf = np.full(len(o), 1e3)
    surf[::4] = 12.0                         # some rays stop at the topography first
    rp, cells, lens, tr, start = O.trace_tetra(pts, tets, nbr, o, d, surf)
    assert (start >= 0).all()
    for r in range(len(o)):
        a = EPS
        seg = range(rp[r], rp[r + 1])
        for k, s in enumerate(seg):
            mid = o[r] + (a + 0.5 * lens[s]) * d[r]
            ins, _ = _bary_inside(pts, tets[cells[s]], mid, 1e-9)
            assert ins, (r, k)
            a += lens[s] + (EPS if k + 1 < len(seg) else 0.0)
        end = min(surf[r], _box_exit(o[r], d[r], lo, hi))
        n_cross = len(seg) - 1
        assert tr[r] == pytest.approx(np.sum(lens[rp[r]:rp[r + 1]]) + EPS, abs=1e-9)
        if surf[r] < _box_exit(o[r], d[r], lo, hi):
            assert tr[r] == pytest.approx(surf[r], abs=1e-9)
        else:
            assert tr[r] + n_cross * EPS == pytest.approx(end, abs=1e-7)


def test_plain_mesh_matches_voxel_walk():
    nx, ny, nz, h = 5, 4, 3, 10.0
    pts, tets = kuhn_mesh(nx, ny, nz, h)
    nbr = O.tet_neighbors(tets)
    lo, hi = np.zeros(3), np.array([nx * h, ny * h, nz * h])
    o, d = _rays(40, 6, lo, hi)
    surf = np.full(len(o), 1e3)
    rp, cells, lens, tr, _ = O.trace_tetra(pts, tets, nbr, o, d, surf)
    xe, ye, ze = (np.linspace(0, n * h, n + 1) for n in (nx, ny, nz))
    vrp, vcells, vlens, vtr = O.trace_edges(xe, ye, ze, o, d, surf)[:4]
    for r in range(len(o)):
        tet_sum = np.bincount(cells[rp[r]:rp[r + 1]] // 6, lens[rp[r]:rp[r + 1]], nx * ny * nz)
        vox_sum = np.bincount(vcells[vrp[r]:vrp[r + 1]], vlens[vrp[r]:vrp[r + 1]], nx * ny * nz)
        crossings = rp[r + 1] - rp[r]
        np.testing.assert_allclose(tet_sum, vox_sum, atol=(crossings + 2) * EPS)


def test_outside_detector():
    pts, tets = kuhn_mesh(2, 2, 2, 10.0)
    nbr = O.tet_neighbors(tets)
    rp, cells, lens, tr, start = O.trace_tetra(pts, tets, nbr, [[50.0, 5.0, 5.0]],
                                               [[0.0, 0.0, 1.0]], [100.0])
    assert start[0] == -1 and rp[-1] == 0 and tr[0] == 0.0

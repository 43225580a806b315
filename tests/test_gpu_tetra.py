"""GPU parity of the tetrahedral ray walk (dp_trace_tetra; trace_path(::TetraVolume),
examples/lvd_tomography.jl:694-736) against the CPU oracle: start tets,
segment cells and lengths, and travelled distances bit for bit, on a plain
Kuhn mesh with detectors on mesh vertices (ties resolved by the first tet in
index order) and on a jittered mesh; the DirectionalPath fields of the
trace_tetra_paths wrapper; a detector outside the mesh; a bad mesh.
"""
import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from workloads import kuhn_mesh

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _rays(dets, n_per, seed):
    rng = np.random.default_rng(seed)
    o, d = [], []
    for det in dets:
        zen = np.radians(rng.uniform(0.0, 75.0, n_per))
        az = np.radians(rng.uniform(0.0, 360.0, n_per))
        d.append(np.column_stack([np.sin(zen) * np.sin(az), np.sin(zen) * np.cos(az), np.cos(zen)]))
        o.append(np.tile(det, (n_per, 1)))
    return np.concatenate(o), np.concatenate(d)


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_walk_bit_exact(ctx, jitter):
    nx, ny, nz, h = 12, 12, 8, 10.0
    pts, tets = kuhn_mesh(nx, ny, nz, h, jitter=jitter, seed=7)
    vol = dp.TetraVolume(pts, tets)
    # detectors on mesh vertices (plain mesh) and inside cubes
    dets = [(40.0, 40.0, 0.0), (60.0, 80.0, 10.0), (85.0, 33.0, 5.0), (20.0, 95.0, 2.5)]
    o, d = _rays(dets, 400, 11)
    surf = np.full(len(o), 500.0)
    surf[::7] = 35.0
    _say(f"tetra walk: {len(tets)} tets, {len(o)} rays, jitter {jitter}")
    rp, cells, lens, tr = dp.trace_tetra_rays(vol, o, d, surf, ctx=ctx)
    orp, ocells, olens, otr, ostart = O.trace_tetra(pts, tets, vol.neighbors, o, d, surf)
    assert (ostart >= 0).all()
    assert np.array_equal(rp, orp) and np.array_equal(cells, ocells)
    assert np.array_equal(lens, olens) and np.array_equal(tr, otr)
    assert rp[-1] > 10 * len(o)


def test_paths_wrapper_and_outside(ctx):
    pts, tets = kuhn_mesh(6, 6, 4, 10.0, origin=(-30.0, -30.0, -5.0), jitter=0.05, seed=2)
    vol = dp.TetraVolume(pts, tets)
    zen = np.array([0.0, 20.0, 45.0, 70.0, 30.0])
    az = np.array([0.0, 90.0, 200.0, 300.0, 10.0])
    surf = np.array([30.0, 60.0, np.nan, 1000.0, 12.0])
    ez = surf * np.cos(np.radians(zen))
    p = dp.trace_tetra_paths(vol, zen, az, surf, ez, ctx=ctx)
    assert list(p.valid) == [1, 1, 0, 1, 1]
    assert p.row_ptr[3] == p.row_ptr[2]                    # the invalid bin has no segment
    ok = p.valid.astype(bool)
    np.testing.assert_allclose(p.remaining_rock[ok], np.maximum(0.0, surf[ok] - p.distance_in_volume[ok]))
    assert np.isnan(p.remaining_rock[2]) and np.isnan(p.exit_z[2])
    # a detector outside the mesh: no start tet, nothing traversed
    rp, cells, lens, tr = dp.trace_tetra_rays(vol, [[500.0, 0.0, 0.0]], [[0.0, 0.0, 1.0]], [100.0],
                                              ctx=ctx)
    assert rp[-1] == 0 and tr[0] == 0.0


def test_bad_mesh(ctx):
    pts, tets = kuhn_mesh(2, 2, 2, 10.0)
    bad = tets.copy()
    bad[3, 2] = len(pts) + 5
    vol = dp.TetraVolume(pts, tets)
    vol.tets = bad
    with pytest.raises(dp.ArgumentError):
        dp.trace_tetra_rays(vol, [[5.0, 5.0, 1.0]], [[0.0, 0.0, 1.0]], [50.0], ctx=ctx)

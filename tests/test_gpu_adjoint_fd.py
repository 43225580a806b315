"""The per-voxel-density adjoint at the differentiation workload's full size
against central finite differences (examples/diff_flux_cube.jl:489-550):
128x128x64 voxels of 10 m, densities U[2000, 3000] kg/m3 (seed 7), 16384 rays
(8 detectors x 2048 directions), 256 log-spaced charge-summed nodes 1-1e4
GeV.  The checked quantity is the gradient of the scalar total flux
(flux_fn of compute_flux_gradient_cube) at 64 seeded voxels among those
crossed by at least 16 rays (over 1024 rays in all; a random on-path voxel
is crossed by 2.5 rays on average).  Only rays crossing voxel v depend on
rho_v, so d F / d rho_v = d F_v / d rho_v with F_v the summed flux of those
rays; the adjoint's value (the rays' Jacobian entries in v summed) is compared
with the central difference (F_v(rho + h e_v) - F_v(rho - h e_v)) / 2h at
h = 1e-3 kg/m3, relative error <= 1e-8.  The difference is formed by the CPU
oracle with the whole program in long double: in FP64 its round-off over 2h
(about 1e-15 F_v / 2h) is itself 5e-9 of the derivative, so FP64 differences
cannot separate an adjoint error of 1e-8 from their own noise.  The FP64
central difference of the GPU forward is printed beside it, per voxel and per
(ray, voxel) pair, for information.

The CSDA flux is piecewise smooth in rho: a node whose energy path crosses a
table knot between rho - h and rho + h has a kink there, and the central
difference then averages the two one-sided slopes.  Such (ray, voxel) pairs
are identified, not skipped silently: the analytic Jacobians at rho +- h
differ (a jump) and the difference lies between them.  They are counted and
printed, and must stay a small minority.
"""
import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip.cube import RayPaths
from workloads import cube_rays, energy_nodes, uniform_densities

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _subpaths(paths, rays):
    rp = [0]
    cells, lens = [], []
    for r in rays:
        a, b = paths.row_ptr[r], paths.row_ptr[r + 1]
        cells.append(paths.cell[a:b])
        lens.append(paths.length[a:b])
        rp.append(rp[-1] + (b - a))
    return RayPaths(rp, np.concatenate(cells), np.concatenate(lens), paths.cos_theta[rays],
                    paths.z_exit[rays])


def _voxel_jac(sub, jac, v):
    """Per ray of `sub`: sum of its segment Jacobians in voxel v."""
    out = np.zeros(sub.n_rays)
    for i in range(sub.n_rays):
        a, b = sub.row_ptr[i], sub.row_ptr[i + 1]
        m = sub.cell[a:b] == v
        out[i] = jac[a:b][m].sum()
    return out


def test_adjoint_against_central_differences_full_size(physics, ctx):
    geo = dp.CubeGeometry(640.0, 128, 128, 64, physics.get_material_index("StandardRock"),
                          physics.get_material_index("Air"), width=1280.0)
    sxy, dirs, cos_t = cube_rays(geo, 8, 2048, seed=2024)
    paths = dp.trace_cube_paths(geo, sxy, dirs, cos_t, ctx=ctx)
    rho = uniform_densities(geo.n_cells, seed=7)
    ne, nw, nq = energy_nodes(256, 1.0, 1e4)
    _say(f"adjoint FD: {paths.n_rays} rays, {paths.n_segments} ray*voxels")
    rng = np.random.default_rng(64)
    ray_of_seg = np.repeat(np.arange(paths.n_rays), np.diff(paths.row_ptr))
    pairs = np.unique(np.stack([paths.cell.astype(np.int64), ray_of_seg]), axis=1)
    crossings = np.bincount(pairs[0], minlength=geo.n_cells)
    cand = np.flatnonzero(crossings >= 16)
    cand = cand[cand != 0]                    # cell 0 also takes the remaining-rock term
    voxels = rng.choice(cand, size=64, replace=False)
    h = 1e-3
    errs, straddle, n_pairs, n_rays_used = [], 0, 0, set()
    verrs, vstraddle, f64errs = [], 0, []
    for v in voxels:
        rays = np.unique(ray_of_seg[paths.cell == v])
        n_rays_used.update(rays.tolist())
        sub = _subpaths(paths, rays)
        _, j0, _ = dp.csda_cube_flux(physics, geo, sub, rho, ne, nw, nq, ctx=ctx)
        p = rho.copy(); p[v] += h
        m = rho.copy(); m[v] -= h
        fp, jp, _ = dp.csda_cube_flux(physics, geo, sub, p, ne, nw, nq, ctx=ctx)
        fm, jm, _ = dp.csda_cube_flux(physics, geo, sub, m, ne, nw, nq, ctx=ctx)
        fd = (fp - fm) / (2 * h)
        J0, Jp, Jm = (_voxel_jac(sub, x, v) for x in (j0, jp, jm))
        G0, Gp, Gm = J0.sum(), Jp.sum(), Jm.sum()
        f64errs.append(abs((fp.sum() - fm.sum()) / (2 * h) - G0) / abs(G0))
        FD, _ = O.cube_fd_ld(physics, geo, sub.row_ptr, sub.cell, sub.length, sub.cos_theta,
                             sub.z_exit, rho, ne, nw, nq, v, h)
        vrel = abs(FD - G0) / abs(G0)
        if vrel > 1e-8 and abs(Gp - Gm) > 1e-6 * abs(G0) and \
                min(Gp, Gm) - 1e-9 * abs(G0) <= FD <= max(Gp, Gm) + 1e-9 * abs(G0):
            vstraddle += 1
            _say(f"  knot-straddling voxel {v}: FD {FD!r} dF/drho {G0!r} at rho-h {Gm!r} "
                 f"at rho+h {Gp!r}")
        else:
            verrs.append(vrel)
        for i in range(sub.n_rays):
            n_pairs += 1
            rel = abs(fd[i] - J0[i]) / abs(J0[i])
            kink = abs(Jp[i] - Jm[i]) > 1e-6 * abs(J0[i]) and \
                min(Jp[i], Jm[i]) - 1e-9 * abs(J0[i]) <= fd[i] <= max(Jp[i], Jm[i]) + 1e-9 * abs(J0[i])
            if rel > 1e-8 and kink:
                straddle += 1
                _say(f"  knot-straddling: voxel {v} ray {rays[i]}: FD {fd[i]!r} J {J0[i]!r} "
                     f"J(rho-h) {Jm[i]!r} J(rho+h) {Jp[i]!r}")
                continue
            errs.append(rel)
    errs, verrs = np.array(errs), np.array(verrs)
    _say(f"adjoint FD per (ray, voxel) pair (information): {n_pairs} pairs, relative error max "
         f"{errs.max():.3g}, median {np.median(errs):.3g}, knot-straddling {straddle}")
    _say(f"adjoint vs FP64 central difference of the GPU forward (information): relative "
         f"error max {max(f64errs):.3g}, median {np.median(f64errs):.3g}")
    _say(f"adjoint vs long-double central difference of the total flux: {len(voxels)} voxels, "
         f"{len(n_rays_used)} rays; relative error max {verrs.max():.3g}, median "
         f"{np.median(verrs):.3g}; knot-straddling voxels {vstraddle}")
    assert len(n_rays_used) >= 1024
    assert vstraddle <= 2
    assert verrs.max() <= 1e-8

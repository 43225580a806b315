"""FP32 stepping variant of the per-voxel-density CSDA flux and adjoint
(dp_set_fp32_stepping) against the FP64 CPU oracle at the differentiation
workload's size: 128x128x64 voxels of 10 m, densities U[2000, 3000] kg/m3,
16384 rays (8 detectors x 2048 directions), 256 charge-summed nodes 1-1e4 GeV
(examples/diff_flux_cube.jl:489-550, which computes in Float64 only).

Bars (relative, 1e-5): every ray's flux; every ray's Jacobian row, measured
against that row's largest entry; the density gradient of the total flux
(the VJP with flux_bar = 1), measured against its largest entry.  Repeated
FP32 calls are bit-identical, and switching FP32 stepping off again gives
the FP64 results bit for bit.
"""
import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from workloads import cube_rays, energy_nodes, uniform_densities

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def test_fp32_stepping_full_size(physics):
    c32 = dp.Context(0)
    c32.set_fp32_stepping(True)
    geo = dp.CubeGeometry(640.0, 128, 128, 64, physics.get_material_index("StandardRock"),
                          physics.get_material_index("Air"), width=1280.0)
    sxy, dirs, cos_t = cube_rays(geo, 8, 2048, seed=2024)
    paths = dp.trace_cube_paths(geo, sxy, dirs, cos_t, ctx=c32)
    rho = uniform_densities(geo.n_cells, seed=7)
    ne, nw, nq = energy_nodes(256, 1.0, 1e4)
    _say(f"fp32 stepping: {paths.n_rays} rays, {paths.n_segments} ray*voxels, 256 nodes")
    f32, j32, r32 = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=c32)
    ms32 = c32.last_kernel_ms()
    g32 = dp.csda_cube_vjp(paths, j32, r32, np.ones(paths.n_rays), geo.n_cells, ctx=c32)
    a32 = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=c32)
    assert all(np.array_equal(x, y) for x, y in zip((f32, j32, r32), a32))
    of, oj, orr = O.cube_csda_reverse(physics, geo.rock_material, geo.air_material,
                                      paths.row_ptr, paths.cell, paths.length, paths.cos_theta,
                                      paths.z_exit, geo.z_max, rho, ne, nw, nq, threads=16)
    og = dp.csda_cube_vjp(paths, oj, orr, np.ones(paths.n_rays), geo.n_cells, ctx=c32)
    ef = np.abs(f32 - of) / np.abs(of)
    rows = np.repeat(np.arange(paths.n_rays), np.diff(paths.row_ptr))
    row_max = np.zeros(paths.n_rays)
    np.maximum.at(row_max, rows, np.abs(oj))
    row_err = np.zeros(paths.n_rays)
    np.maximum.at(row_err, rows, np.abs(j32 - oj))
    ej = row_err / np.where(row_max > 0, row_max, 1.0)
    er = np.abs(r32 - orr) / np.maximum(np.abs(orr), 1e-300)
    eg = np.abs(g32 - og).max() / np.abs(og).max()
    c64 = dp.Context(0)
    f64_, j64, r64 = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=c64)
    ms64 = c64.last_kernel_ms()
    _say(f"  flux rel err max {ef.max():.3g} median {np.median(ef):.3g}; Jacobian row err max "
         f"{ej.max():.3g} median {np.median(ej):.3g}; remaining-rock term max {er.max():.3g}; "
         f"density gradient {eg:.3g}")
    _say(f"  kernel {ms32:.3f} ms (FP32 stepping) vs {ms64:.3f} ms (FP64)")
    assert ef.max() <= 1e-5
    assert ej.max() <= 1e-5
    assert eg <= 1e-5
    c32.set_fp32_stepping(False)
    b64 = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=c32)
    assert all(np.array_equal(x, y) for x, y in zip((f64_, j64, r64), b64))
    c32.close()
    c64.close()

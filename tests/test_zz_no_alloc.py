"""dp_csda_cube_vjp_dev_noalloc (the density VJP whose scratch is a static
device array) against dp_csda_cube_vjp_dev through the C ABI: bit-identical
gradients on two calls in a row, the second with a smaller problem (the
static buffer keeps the first call's contents), and a problem beyond the
static buffer refused with DP_E_CAPACITY before anything runs."""
import ctypes as C

import numpy as np
import pytest

import diffpumas_hip as dp
from diffpumas_hip import _lib as L
from _hip import DeviceArray as D
from workloads import cube_rays, energy_nodes, uniform_densities

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _case(physics, ctx, nx, ny, nz, n_det, n_dir, seed):
    geo = dp.CubeGeometry(nz * 10.0, nx, ny, nz, physics.get_material_index("StandardRock"),
                          physics.get_material_index("Air"), width=nx * 10.0)
    sxy, dirs, cos_t = cube_rays(geo, n_det, n_dir, seed=seed)
    paths = dp.trace_cube_paths(geo, sxy, dirs, cos_t, ctx=ctx)
    rho = uniform_densities(geo.n_cells, seed=seed)
    ne, nw, nq = energy_nodes(64, 1.0, 1e4)
    flux, jac, jrem = dp.csda_cube_flux(physics, geo, paths, rho, ne, nw, nq, ctx=ctx)
    fbar = np.random.default_rng(seed).uniform(0.5, 2.0, paths.n_rays)
    return geo, paths, jac, jrem, fbar


def _vjp(ctx, fn, geo, paths, jac, jrem, fbar):
    keep = [D(paths.row_ptr.astype(np.int32)), D(paths.cell.astype(np.int32)),
            D(paths.length), D(paths.cos_theta), D(paths.z_exit),
            D(jac if len(jac) else np.zeros(1)), D(jrem), D(fbar), D(shape=(geo.n_cells,))]
    rp, cell, ln, cs, ze, j, jr, fb, g = keep
    cast = lambda x, t: C.cast(x.ptr, t)
    cp = L.Paths(paths.n_rays, paths.n_segments, paths.max_segments, cast(rp, L.i32_p),
                 cast(cell, L.i32_p), cast(ln, L.dbl_p), cast(cs, L.dbl_p), cast(ze, L.dbl_p))
    L.check(fn(ctx.h, C.byref(cp), j.ptr, jr.ptr, fb.ptr, geo.n_cells, g.ptr))
    ctx.synchronize()
    return g.numpy()


def test_noalloc_vjp_equals_allocating_vjp(physics, ctx):
    big = _case(physics, ctx, 48, 48, 24, 6, 512, 3)
    small = _case(physics, ctx, 20, 16, 10, 3, 64, 4)
    for name, case in (("first, larger", big), ("second, smaller", small)):
        geo, paths, jac, jrem, fbar = case
        _say(f"no-alloc VJP ({name}): {paths.n_rays} rays, {paths.n_segments} segments, "
             f"{geo.n_cells} cells")
        ref = _vjp(ctx, ctx.lib.dp_csda_cube_vjp_dev, *case)
        got = _vjp(ctx, ctx.lib.dp_csda_cube_vjp_dev_noalloc, *case)
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        assert np.any(got != 0.0)


def test_noalloc_vjp_refuses_oversized_problem(physics, ctx):
    # sizes only: the pointers are never read, the call fails before any launch
    dummy = C.cast(C.c_void_p(16), L.i32_p)
    cp = L.Paths(1000, 3_000_000, 3000, dummy, dummy, C.cast(dummy, L.dbl_p),
                 C.cast(dummy, L.dbl_p), C.cast(dummy, L.dbl_p))
    rc = ctx.lib.dp_csda_cube_vjp_dev_noalloc(ctx.h, C.byref(cp), 16, 16, 16, 100, 16)
    assert rc == L.DP_E_CAPACITY
    assert b"static buffer" in ctx.lib.dp_last_error()

"""HIP-graph capture of the per-iteration device work of a Gauss-Newton loop
(gauss_newton_reconstruct, src/Tomography_inverse.jl:448-603): the w-model
flux + Jacobian (dp_csda_w_flux_jacobian_dev) followed by the normal
equations of that Jacobian (dp_normal_equations_dev), captured once and
replayed for three water-fraction fields.  Every replay, with eager work and
synchronisations between replays, must equal the eager host-API results for
the same field bit for bit.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import diffpumas_hip as dp
from diffpumas_hip import _lib as L
from diffpumas_hip.tomography import MaterialConfig, SiteConfig
from _hip import DeviceArray as D
from test_gpu_wmodel import _traced

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def test_gauss_newton_step_graph(ctx):
    ph = dp.create_physics_csda(("StandardRock", "Water", "Air"))
    paths, n_cells = _traced(ctx, 64, 17)
    s = dp.csda_energy_quadrature(32, 1.0, 1e4)
    m = MaterialConfig(0, 1, 2, -1, ph.scalars[0, 0], ph.scalars[1, 0])
    site = SiteConfig(800.0, 3e4)
    sh = np.zeros(n_cells, np.uint8)
    rng = np.random.default_rng(5)
    fields = [rng.uniform(0.0, 0.4, n_cells) for _ in range(3)]
    nr, ns = paths.n_rays, paths.n_segments
    W = rng.uniform(0.5, 2.0, nr)
    res = rng.normal(0.0, 1e-3, nr)
    active = np.flatnonzero(np.bincount(paths.cells, minlength=n_cells) >= 2).astype(np.int32)
    na = len(active)
    _say(f"graph: {nr} bins, {ns} segments, {na} free cells")
    # device buffers
    d_rp, d_cell = D(paths.row_ptr.astype(np.int32)), D(paths.cells.astype(np.int32))
    d_len, d_cos = D(paths.lengths), D(paths.cos_theta)
    d_rem, d_ez, d_val = D(paths.remaining_rock), D(paths.exit_z), D(paths.valid.astype(np.uint8))
    d_w, d_sh = D(fields[0]), D(sh)
    d_e, d_wt, d_q = D(s.energy), D(s.weight), D(s.charge)
    d_flux, d_jac = D(shape=(nr,)), D(shape=(ns,))
    d_W, d_res, d_act = D(W), D(res), D(active)
    d_N, d_g = D(shape=(na * na,)), D(shape=(na,))
    cast = lambda x, t: C.cast(x.ptr, t)
    cp = L.DirectionalPaths(nr, ns, paths.max_segments(), cast(d_rp, L.i32_p), cast(d_cell, L.i32_p),
                            cast(d_len, L.dbl_p), cast(d_cos, L.dbl_p), cast(d_rem, L.dbl_p),
                            cast(d_ez, L.dbl_p), C.cast(d_val.ptr, C.POINTER(C.c_uint8)))
    es = L.EnergySamples(len(s.energy), cast(d_e, L.dbl_p), cast(d_wt, L.dbl_p), cast(d_q, L.dbl_p))
    mc, sc = m.c(), L.SiteConfig(site.detector_elevation, site.primary_altitude)
    csr = L.Csr(nr, n_cells, ns, d_rp.ptr, d_cell.ptr, d_jac.ptr)
    h = ph.device_handle(ctx)
    dp.set_csda_correction(enabled=False)
    dp.set_csda_material_correction(enabled=False)
    L.check(ctx.lib.dp_set_csda_correction(ctx.h, C.byref(L.CsdaCorrection(0, 0, 0, 0, 0, 0, 0))))
    L.check(ctx.lib.dp_set_csda_material_correction(ctx.h, C.byref(L.CsdaMaterialCorrection(0, 0, 0))))

    def work():
        L.check(ctx.lib.dp_csda_w_flux_jacobian_dev(ctx.h, h, C.byref(mc), C.byref(sc), C.byref(cp),
                                                    n_cells, d_w.ptr, d_sh.ptr, C.byref(es),
                                                    d_flux.ptr, d_jac.ptr))
        L.check(ctx.lib.dp_normal_equations_dev(ctx.h, C.byref(csr), d_W.ptr, d_res.ptr, na,
                                                d_act.ptr, d_N.ptr, d_g.ptr))

    work()                      # eager once: every buffer gets its size
    ctx.synchronize()
    graph = ctx.capture(work)
    rows = np.repeat(np.arange(nr), np.diff(paths.row_ptr))
    for k, w in enumerate(fields):
        f_ref, j_ref = dp.flux_and_segment_jacobian(ph, sh, m, site, paths, w, s, ctx=ctx)
        keep = np.isfinite(j_ref) & (j_ref != 0.0)
        J = sp.csr_matrix((j_ref[keep], (rows[keep], paths.cells[keep])), shape=(nr, n_cells))
        N_ref, g_ref = dp.normal_equations(J, res, W, active, ctx=ctx)
        hip = __import__("_hip").hip()
        assert hip.hipMemcpy(d_w.ptr, np.ascontiguousarray(w).ctypes.data, 8 * n_cells, 1) == 0
        graph.launch()
        ctx.synchronize()
        _say(f"graph: replay {k}")
        f_g = d_flux.numpy()
        ok = np.isfinite(f_ref)
        assert np.array_equal(np.isnan(f_g), ~ok) and np.array_equal(f_g[ok], f_ref[ok])
        assert np.array_equal(d_jac.numpy(), j_ref)
        assert np.array_equal(d_N.numpy().reshape(na, na).T, N_ref)
        assert np.array_equal(d_g.numpy(), g_ref)
    graph.close()

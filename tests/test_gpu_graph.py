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


def _w_dev_setup(ctx, ph, n_dirs, seed, max_seg=None):
    """Device copies of a traced w-model problem and a closure that enqueues
    dp_csda_w_flux_jacobian_dev on them (optionally with a max_segments
    smaller than the longest path)."""
    paths, n_cells = _traced(ctx, n_dirs, seed)
    s = dp.csda_energy_quadrature(32, 1.0, 1e4)
    m = MaterialConfig(0, 1, 2, -1, ph.scalars[0, 0], ph.scalars[1, 0])
    nr, ns = paths.n_rays, paths.n_segments
    keep = [D(paths.row_ptr.astype(np.int32)), D(paths.cells.astype(np.int32)), D(paths.lengths),
            D(paths.cos_theta), D(paths.remaining_rock), D(paths.exit_z),
            D(paths.valid.astype(np.uint8)), D(np.full(n_cells, 0.1)), D(np.zeros(n_cells, np.uint8)),
            D(s.energy), D(s.weight), D(s.charge), D(shape=(nr,)), D(shape=(ns,))]
    rp, cell, ln, cs, rem, ez, val, w, sh, e, wt, q, flux, jac = keep
    cast = lambda x, t: C.cast(x.ptr, t)
    cp = L.DirectionalPaths(nr, ns, max_seg or paths.max_segments(), cast(rp, L.i32_p),
                            cast(cell, L.i32_p), cast(ln, L.dbl_p), cast(cs, L.dbl_p),
                            cast(rem, L.dbl_p), cast(ez, L.dbl_p),
                            C.cast(val.ptr, C.POINTER(C.c_uint8)))
    es = L.EnergySamples(len(s.energy), cast(e, L.dbl_p), cast(wt, L.dbl_p), cast(q, L.dbl_p))
    mc, sc = m.c(), L.SiteConfig(600.0, 3e4)
    h = ph.device_handle(ctx)

    def work():
        return ctx.lib.dp_csda_w_flux_jacobian_dev(ctx.h, h, C.byref(mc), C.byref(sc), C.byref(cp),
                                                   n_cells, w.ptr, sh.ptr, C.byref(es), flux.ptr,
                                                   jac.ptr)
    return work, keep, paths


def test_graph_refused_after_reallocation():
    """A graph points into the context's scratch: once a larger call moves
    that buffer, the graph is refused (not replayed into freed memory)."""
    ph = dp.create_physics_csda(("StandardRock", "Water", "Air"))
    c2 = dp.Context(0)            # a fresh context: its scratch starts small
    work, keep, _ = _w_dev_setup(c2, ph, 4, 3)
    L.check(work())
    c2.synchronize()
    g = c2.capture(lambda: L.check(work()))
    g.launch()
    c2.synchronize()
    _say("graph: replay before reallocation ok")
    big, keep2, paths2 = _w_dev_setup(c2, ph, 256, 4)   # more rays, longer paths: scratch grows
    _say(f"graph: larger call with {paths2.n_rays} bins, {paths2.n_segments} segments")
    L.check(big())
    c2.synchronize()
    with pytest.raises(dp.ArgumentError):
        g.launch()
    g.close()
    g2 = c2.capture(lambda: L.check(work()))   # a new capture uses the new buffers
    g2.launch()
    c2.synchronize()
    g2.close()
    c2.check_errors()
    c2.close()


def test_check_errors_reports_device_capacity():
    """A _dev call whose paths exceed the max_segments it was given marks
    those bins NaN and sets the sticky word; dp_check_errors raises it once
    and clears it."""
    ph = dp.create_physics_csda(("StandardRock", "Water", "Air"))
    c2 = dp.Context(0)
    ok, keep, paths = _w_dev_setup(c2, ph, 16, 5)
    L.check(ok())
    c2.check_errors()                          # nothing to report
    short = int(np.diff(paths.row_ptr).max()) - 1
    assert short >= 1
    bad, keep2, _ = _w_dev_setup(c2, ph, 16, 5, max_seg=short)
    L.check(bad())                             # enqueued without error
    flux = keep2[12].numpy()
    over = np.diff(paths.row_ptr) > short
    _say(f"check_errors: {int(over.sum())} bins over max_segments {short}")
    assert np.all(np.isnan(flux[over]))
    with pytest.raises(dp.CapacityError):
        c2.check_errors()
    c2.check_errors()                          # cleared by the first check
    L.check(ok())                              # a good call after the bad one leaves it clear
    c2.check_errors()
    c2.close()

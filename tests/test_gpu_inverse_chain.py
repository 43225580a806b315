"""The inverse loop of the tomography study on the GPU, end to end: a seeded
diamond-square DEM -> ray exit (dp_ray_surface_exit) -> voxel traversal
(dp_trace_edges) -> w-model flux and Jacobian (dp_csda_w_flux_jacobian_dev)
-> residual and data cost (dp_residual_dev) -> normal equations
(dp_normal_equations_dev), the last three captured once as a HIP graph and
replayed for every model evaluation of a Gauss-Newton / Levenberg-Marquardt
loop.  The host loop is a test-side restatement of gauss_newton_reconstruct
in its legacy mode without a prior (src/Tomography_inverse.jl:448-603): the
damped dense solve, the projected step onto the box, the halving line
search and the acceptance test f_new < f_cur - 1e-10 |f_cur|.

Reduced size (against the oracle chain): 65x65 DEM, 16x16x8 voxels, 2
detectors x 128 directions, 3 iterations.  At every accepted iterate the
graph's prediction, Jacobian, cost, normal matrix and gradient are compared
with the oracle's at the same water fractions (1e-10 / 1e-9 relative), the
ray exits and traversals bit for bit, and a second loop driven by the
oracle's model must take the same accept/reject decisions and reach the
same iterates (1e-6 absolute).

Full size (properties): 257x257 DEM (256x256 cells, 400 m relief), 96x96x48
voxels of 10 m with two buried water-filled anomalies, Poisson counts for
4 detectors x 4096 directions, 20 iterations over the best-covered 3000
cells (the others held at their true values): replays are deterministic,
the objective never increases and falls well below its start, and at the
final iterate the graph's outputs equal the oracle's.
"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip import _lib as L
from diffpumas_hip.tomography import DirectionalPaths, MaterialConfig, SiteConfig, direction_vector
from _hip import DeviceArray as D, hip
from workloads import diamond_square

pytestmark = pytest.mark.gpu

Z_DET = 963.0
BOX = (0.0, 0.9)          # (0, MAX_WATER_FRACTION), src/Tomography.jl:89


def _say(m):
    print(m, flush=True)


# --------------------------------------------------------------- geometry
def _terrain(n_pow, spacing, seed):
    n = 2 ** n_pow + 1
    xs = (np.arange(n) - (n - 1) / 2) * spacing
    X, Y = np.meshgrid(xs, xs)
    half = xs[-1]
    z = 1200.0 + diamond_square(n_pow, 400.0, seed) + 300.0 * np.exp(-(X ** 2 + Y ** 2) /
                                                                     (0.35 * half) ** 2)
    return xs, z


def _chain_geometry(ctx, n_pow, spacing, nv, vsize, dets, n_dirs, seed, zen_max=60.0):
    """DEM exits and voxel traversal for every (detector, direction): GPU
    results, checked against the oracle (exits bit for bit; traversal bit for
    bit).  Returns the traced paths (valid bins only) and the edges."""
    xs, z = _terrain(n_pow, spacing, seed)
    n = len(xs)
    rng = np.random.default_rng(seed + 1)
    zen_all, az_all, surf_all, ez_all, org_all = [], [], [], [], []
    for k, (xd, yd) in enumerate(dets):
        emap = dp.ElevationMap(n, n, (xs[0] - xd) * 1e-3, (xs[0] - yd) * 1e-3, spacing * 1e-3,
                               spacing * 1e-3, z)
        zen = rng.uniform(0.0, zen_max, n_dirs)
        az = rng.uniform(0.0, 360.0, n_dirs)
        dirs = np.array([direction_vector(a, b) for a, b in zip(zen, az)])
        dist, found, ez = dp.ray_surface_exit_distance(emap, dirs, Z_DET, ctx=ctx)
        od, of, oz = O.ray_surface_exit(dict(nx=n, ny=n, x0=emap.x0, y0=emap.y0, dx=emap.dx,
                                             dy=emap.dy, data=z), dirs, z_ref=Z_DET)
        assert np.array_equal(found, of) and np.array_equal(dist[found], od[of])
        assert np.array_equal(ez[found], oz[of])
        zen_all.append(zen[found]); az_all.append(az[found])
        surf_all.append(dist[found]); ez_all.append(ez[found])
        org_all.append(np.tile([xd, yd, Z_DET], (int(found.sum()), 1)))
    zen, az = np.concatenate(zen_all), np.concatenate(az_all)
    surf, ez, org = np.concatenate(surf_all), np.concatenate(ez_all), np.concatenate(org_all)
    nx, ny, nz = nv
    xe = (np.arange(nx + 1) - nx / 2) * vsize
    ye = (np.arange(ny + 1) - ny / 2) * vsize
    ze = Z_DET - vsize + np.arange(nz + 1) * vsize
    paths = dp.trace_voxel_paths(xe, ye, ze, zen, az, surf, ez, detectors=org, ctx=ctx)
    dirs = np.array([direction_vector(a, b) for a, b in zip(zen, az)])
    rp, cells, lens, tr = O.trace_edges(xe, ye, ze, org, dirs, surf)
    assert np.array_equal(paths.row_ptr, rp) and np.array_equal(paths.cells, cells)
    assert np.array_equal(paths.lengths, lens)
    return paths, (xe, ye, ze), nx * ny * nz


# ------------------------------------------------------------------ models
class GraphModel:
    """model(w) of make_csda_operator on the device, with the residual, data
    cost and normal equations of the same iterate, captured once as a graph."""

    def __init__(self, ctx, ph, m, site, paths, n_cells, nodes, obs, W, active):
        self.ctx, self.paths, self.n_cells = ctx, paths, n_cells
        nr, ns = paths.n_rays, paths.n_segments
        self.na = na = len(active)
        self.keep = [D(paths.row_ptr.astype(np.int32)), D(paths.cells.astype(np.int32)),
                     D(paths.lengths), D(paths.cos_theta), D(paths.remaining_rock),
                     D(paths.exit_z), D(paths.valid.astype(np.uint8)),
                     D(np.zeros(n_cells)), D(np.zeros(n_cells, np.uint8)),
                     D(nodes.energy), D(nodes.weight), D(nodes.charge)]
        rp, cell, ln, cs, rem, ezd, val, self.d_w, sh, e, wt, q = self.keep
        self.d_flux, self.d_jac, self.d_res = D(shape=(nr,)), D(shape=(ns,)), D(shape=(nr,))
        self.d_cost, self.d_obs, self.d_W = D(shape=(1,)), D(obs), D(W)
        self.d_act = D(np.asarray(active, np.int32))
        self.d_N, self.d_g = D(shape=(na * na,)), D(shape=(na,))
        cast = lambda x, t: C.cast(x.ptr, t)
        self.cp = L.DirectionalPaths(nr, ns, paths.max_segments(), cast(rp, L.i32_p),
                                     cast(cell, L.i32_p), cast(ln, L.dbl_p), cast(cs, L.dbl_p),
                                     cast(rem, L.dbl_p), cast(ezd, L.dbl_p),
                                     C.cast(val.ptr, C.POINTER(C.c_uint8)))
        self.es = L.EnergySamples(len(nodes.energy), cast(e, L.dbl_p), cast(wt, L.dbl_p),
                                  cast(q, L.dbl_p))
        self.mc, self.sc = m.c(), L.SiteConfig(site.detector_elevation, site.primary_altitude)
        self.csr = L.Csr(nr, n_cells, ns, rp.ptr, cell.ptr, self.d_jac.ptr)
        self.h = ph.device_handle(ctx)
        self.sh = sh
        L.check(ctx.lib.dp_set_csda_correction(ctx.h, C.byref(L.CsdaCorrection(0, 0, 0, 0, 0, 0, 0))))
        L.check(ctx.lib.dp_set_csda_material_correction(
            ctx.h, C.byref(L.CsdaMaterialCorrection(0, 0, 0))))
        self._work()              # eager once: every buffer gets its size
        ctx.synchronize()
        ctx.check_errors()
        self.graph = ctx.capture(self._work)
        self.replays = 0

    def _work(self):
        lib, c = self.ctx.lib, self.ctx.h
        nr = self.paths.n_rays
        L.check(lib.dp_csda_w_flux_jacobian_dev(c, self.h, C.byref(self.mc), C.byref(self.sc),
                                                C.byref(self.cp), self.n_cells, self.d_w.ptr,
                                                self.sh.ptr, C.byref(self.es), self.d_flux.ptr,
                                                self.d_jac.ptr))
        L.check(lib.dp_residual_dev(c, nr, self.d_flux.ptr, self.d_obs.ptr, self.d_W.ptr,
                                    self.d_res.ptr, self.d_cost.ptr))
        L.check(lib.dp_normal_equations_dev(c, C.byref(self.csr), self.d_W.ptr, self.d_res.ptr,
                                            self.na, self.d_act.ptr, self.d_N.ptr, self.d_g.ptr))

    def evaluate(self, w):
        """Replay at w; returns the data cost (the rest stays on the device)."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert hip().hipMemcpy(self.d_w.ptr, w.ctypes.data, 8 * self.n_cells, 1) == 0
        self.graph.launch()
        self.ctx.synchronize()
        self.replays += 1
        return float(self.d_cost.numpy()[0])

    def outputs(self):
        """(pred, jac, residual, normal, grad) of the last replay."""
        na = self.na
        return (self.d_flux.numpy(), self.d_jac.numpy(), self.d_res.numpy(),
                self.d_N.numpy().reshape(na, na).T, self.d_g.numpy())

    def close(self):
        self.graph.close()
        self.ctx.check_errors()


class OracleModel:
    """The same evaluation on the CPU oracle."""

    def __init__(self, ph, m, site, paths, n_cells, nodes, obs, W, active):
        self.ph, self.paths, self.n_cells = ph, paths, n_cells
        self.mc = O.matcfg_arrays(m.rock_material, m.water_material, m.air_material,
                                  m.porous_material, m.rock_density, m.water_density,
                                  m.porous_density, m.porous_top_thickness)
        self.site = (site.detector_elevation, site.primary_altitude)
        self.op = dict(row_ptr=paths.row_ptr, cells=paths.cells, lengths=paths.lengths,
                       remaining=paths.remaining_rock, exit_z=paths.exit_z,
                       cos_theta=paths.cos_theta, valid=paths.valid)
        self.nodes, self.obs, self.W, self.active = nodes, obs, W, np.asarray(active)
        self.rows = np.repeat(np.arange(paths.n_rays), np.diff(paths.row_ptr))
        self.last = None

    def evaluate(self, w):
        n = self.nodes
        pred, jac = O.w_jacobian(self.ph, self.mc, self.site, self.op, w,
                                 np.zeros(self.n_cells, np.uint8), n.energy, n.weight, n.charge,
                                 threads=8)
        r = pred - self.obs
        cost = 0.5 * float(np.sum(self.W * (r * r)))
        keep = np.isfinite(jac) & (jac != 0.0)
        J = sp.csr_matrix((jac[keep], (self.rows[keep], self.paths.cells[keep])),
                          shape=(self.paths.n_rays, self.n_cells))
        N, g = O.normal_equations(J, r, self.W, self.active)
        self.last = (pred, jac, r, N, g)
        return cost

    def outputs(self):
        return self.last


# --------------------------------------------------------- the host loop
def gauss_newton(model, w0, active, n_iter, lm_damping=1e-3, lm_factor=3.0, on_accept=None):
    """gauss_newton_reconstruct, legacy mode, no prior
    (src/Tomography_inverse.jl:448-603).  Returns (w, objective history,
    decisions) with decisions the (try, line-search step) of each update."""
    w = np.clip(np.asarray(w0, np.float64), *BOX)
    f_cur = model.evaluate(w)
    _, _, r, N, g = model.outputs()
    if on_accept:
        on_accept(0, w, model)
    mu = lm_damping
    hist, decisions = [f_cur], []
    for it in range(1, n_iter + 1):
        d = np.diag(N).copy()
        dmax = d.max() if d.max() > 0 else 1.0
        accepted = False
        for tr in range(1, 7):
            A = N + np.diag(mu * np.maximum(d, 1e-12 * dmax))
            try:
                delta = -scipy.linalg.solve(A, g, check_finite=True)
            except (np.linalg.LinAlgError, ValueError):
                delta = np.zeros(len(active))
            if not np.all(np.isfinite(delta)):
                mu = min(mu * lm_factor, 1e12)
                continue
            alpha = 1.0
            for ls in range(1, 15):
                w_new = w.copy()
                w_new[active] = np.clip(w[active] + alpha * delta, *BOX)
                if np.array_equal(w_new, w):
                    break
                f_new = model.evaluate(w_new)
                if math.isfinite(f_new) and f_new < f_cur - 1e-10 * abs(f_cur):
                    w, f_cur = w_new, f_new
                    _, _, r, N, g = model.outputs()
                    accepted = True
                    decisions.append((tr, ls))
                    break
                alpha *= 0.5
            if accepted:
                mu = max(mu / lm_factor, 1e-12)
                break
            mu = min(mu * lm_factor, 1e12)
        hist.append(f_cur)
        if not accepted:
            decisions.append(None)
            break
        if on_accept:
            on_accept(it, w, model)
    return w, hist, decisions


def _problem(ph, paths, n_cells, w_true, nodes, m, site, ctx, exposure, seed):
    f_true, _ = dp.flux_and_segment_jacobian(ph, np.zeros(n_cells, np.uint8), m, site, paths,
                                             w_true, nodes, ctx=ctx)
    assert np.all(np.isfinite(f_true))
    counts = np.random.default_rng(seed).poisson(f_true * exposure)
    obs = counts / exposure
    W = exposure / np.maximum(obs, 1.0 / exposure)   # 1 / Poisson variance of obs
    return obs, W


def _free(paths, n_cells, min_crossings, cap):
    rows = np.repeat(np.arange(paths.n_rays), np.diff(paths.row_ptr))
    pairs = np.unique(np.stack([paths.cells.astype(np.int64), rows]), axis=1)
    cross = np.bincount(pairs[0], minlength=n_cells)
    m = min_crossings
    while (cross >= m).sum() > cap:
        m += 1
    return np.flatnonzero(cross >= m).astype(np.int32), m


def _setup(ctx, n_pow, spacing, nv, vsize, dets, n_dirs, seed, anomalies, exposure,
           min_crossings, cap, n_nodes=32):
    ph = dp.create_physics_csda(("StandardRock", "Water", "Air"))
    m = MaterialConfig(0, 1, 2, -1, ph.scalars[0, 0], ph.scalars[1, 0])
    site = SiteConfig(Z_DET, 3e4)
    nodes = dp.csda_energy_quadrature(n_nodes, 1.0, 1e4)
    paths, (xe, ye, ze), n_cells = _chain_geometry(ctx, n_pow, spacing, nv, vsize, dets, n_dirs,
                                                   seed)
    xc, yc, zc = (0.5 * (e[1:] + e[:-1]) for e in (xe, ye, ze))
    Zc, Yc, Xc = np.meshgrid(zc, yc, xc, indexing="ij")      # cell id = ix + nx (iy + ny iz)
    w_true = np.full(n_cells, 0.03)
    for (ax, ay, az_, rad, wa) in anomalies:
        inside = ((Xc - ax) ** 2 + (Yc - ay) ** 2 + (Zc - az_) ** 2).ravel() <= rad ** 2
        w_true[inside] = wa
    obs, W = _problem(ph, paths, n_cells, w_true, nodes, m, site, ctx, exposure, seed + 2)
    active, mc = _free(paths, n_cells, min_crossings, cap)
    return ph, m, site, nodes, paths, n_cells, w_true, obs, W, active, mc


def _cmp(a, b, rtol, label):
    scale = np.abs(b).max() if b.size else 1.0
    err = np.abs(a - b).max() / scale if b.size else 0.0
    assert np.allclose(a, b, rtol=rtol, atol=rtol * 1e-3 * scale), (label, err)
    return err


def test_gauss_newton_chain_against_oracle(ctx):
    ph, m, site, nodes, paths, n_cells, w_true, obs, W, active, mcr = _setup(
        ctx, 6, 40.0, (16, 16, 8), 20.0, [(-40.0, 30.0), (50.0, -20.0)], 128, 3,
        [(-30.0, 20.0, Z_DET + 70.0, 35.0, 0.4)], 2e8, 3, 1500)
    _say(f"GN chain (reduced): {paths.n_rays} bins, {paths.n_segments} segments, "
         f"{len(active)} free cells (>= {mcr} crossings)")
    gm = GraphModel(ctx, ph, m, site, paths, n_cells, nodes, obs, W, active)
    om = OracleModel(ph, m, site, paths, n_cells, nodes, obs, W, active)
    errs = []

    def check(it, w, model):
        pred, jac, r, N, g = model.outputs()
        cost_g = float(model.d_cost.numpy()[0])
        oc = om.evaluate(w)
        op, oj, orr, oN, og = om.outputs()
        e = [_cmp(pred, op, 1e-10, "pred"), _cmp(jac, oj, 1e-10, "jacobian"),
             _cmp(N, oN, 1e-9, "normal"), _cmp(g, og, 1e-9, "gradient")]
        assert abs(cost_g - oc) <= 1e-9 * abs(oc), (cost_g, oc)
        errs.append(max(e))
        _say(f"  iterate {it}: cost {cost_g:.6g} (oracle {oc:.6g}); max rel err pred/J/N/g "
             + " ".join(f"{x:.2g}" for x in e))

    w0 = np.full(n_cells, 0.1)
    wg, hg, dg = gauss_newton(gm, w0, active, 3, on_accept=check)
    gm.close()
    _say(f"  GPU loop: objective {hg}, decisions {dg}, {gm.replays} graph replays")
    wo, ho, do = gauss_newton(om, w0, active, 3)
    _say(f"  oracle loop: objective {ho}, decisions {do}; max |w_gpu - w_oracle| "
         f"{np.abs(wg - wo).max():.3g}")
    assert dg == do
    assert np.allclose(hg, ho, rtol=1e-8, atol=0)
    assert np.abs(wg - wo).max() <= 1e-6
    assert all(b <= a for a, b in zip(hg, hg[1:])) and hg[-1] < hg[0]
    assert len(errs) == 1 + sum(d is not None for d in dg)


def test_gauss_newton_full_size_properties(ctx):
    dets = [(-150.0, -150.0), (150.0, -150.0), (-150.0, 150.0), (150.0, 150.0)]
    anomalies = [(-150.0, -150.0, Z_DET + 45.0, 30.0, 0.4), (150.0, 150.0, Z_DET + 60.0, 35.0, 0.35)]
    ph, m, site, nodes, paths, n_cells, w_true, obs, W, active, mcr = _setup(
        ctx, 8, 20.0, (96, 96, 48), 10.0, dets, 4096, 11, anomalies, 5e7, 16, 3000)
    _say(f"GN chain (full): {paths.n_rays} bins, {paths.n_segments} segments, {n_cells} cells, "
         f"{len(active)} free cells (>= {mcr} crossings)")
    assert paths.n_rays >= 4 * 4096 * 0.95
    gm = GraphModel(ctx, ph, m, site, paths, n_cells, nodes, obs, W, active)
    w0 = w_true.copy()            # cells outside the free set are known
    w0[active] = 0.1
    c1 = gm.evaluate(w0)
    o1 = gm.outputs()
    c2 = gm.evaluate(w0)
    o2 = gm.outputs()
    assert c1 == c2 and all(np.array_equal(a, b, equal_nan=True) for a, b in zip(o1, o2))
    _say(f"  replays at w0 bit-identical; cost {c1:.6g}")

    def progress(it, w, model):
        _say(f"  iteration {it}: cost {float(model.d_cost.numpy()[0]):.6g}, "
             f"{model.replays} replays")

    wg, hg, dg = gauss_newton(gm, w0, active, 20, on_accept=progress)
    _say(f"  objective {hg[0]:.6g} -> {hg[-1]:.6g} in {len(hg) - 1} updates, decisions {dg}")
    assert all(b <= a for a, b in zip(hg, hg[1:]))
    assert hg[-1] < 0.5 * hg[0]
    # the final iterate's graph outputs against the oracle
    gm.evaluate(wg)
    pred, jac, r, N, g = gm.outputs()
    om = OracleModel(ph, m, site, paths, n_cells, nodes, obs, W, active)
    om.evaluate(wg)
    op, oj, orr, oN, og = om.outputs()
    e = [_cmp(pred, op, 1e-10, "pred"), _cmp(jac, oj, 1e-10, "jacobian"),
         _cmp(N, oN, 1e-9, "normal"), _cmp(g, og, 1e-9, "gradient")]
    _say("  final iterate vs oracle: max rel err pred/J/N/g " + " ".join(f"{x:.2g}" for x in e))
    inside = w_true[active] > 0.1
    _say(f"  recovered w: mean {wg[active][inside].mean():.3f} in anomaly cells "
         f"({inside.sum()}), {wg[active][~inside].mean():.3f} elsewhere (truth 0.03)")
    gm.close()

"""GPU parity of the Gauss-Newton normal equations (dp_normal_equations;
gauss_newton_reconstruct, src/Tomography_inverse.jl:489-494) against the CPU
oracle, bit for bit: both sum every entry over rows in ascending order.

Cases: the w-model Jacobian of traced paths through make_csda_operator with a
free mask and Poisson-like weights; a random Jacobian with several row
tiles per output column and a column crossed by every row; the
device-pointer entry point; zero and non-finite entries; argument errors.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip import _lib as L
from diffpumas_hip.tomography import MaterialConfig, SiteConfig
from test_gpu_wmodel import _traced

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _rand_csr(n_rows, n_cols, per_row, seed, dense_col=None):
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for r in range(n_rows):
        c = set(rng.choice(n_cols, size=int(rng.integers(1, per_row + 1)), replace=False).tolist())
        if dense_col is not None:
            c.add(dense_col)
        c = sorted(c)
        rows += [r] * len(c)
        cols += c
        vals += list(rng.normal(0.0, 1.0, len(c)) * 10.0 ** rng.uniform(-3, 1, len(c)))
    return sp.csr_matrix((vals, (rows, cols)), shape=(n_rows, n_cols))


def test_wmodel_operator_bit_exact(ctx):
    ph = dp.create_physics_csda(("StandardRock", "Water", "Air"))
    paths, n_cells = _traced(ctx, 160, 21)
    rng = np.random.default_rng(4)
    w = rng.uniform(0.0, 0.4, n_cells)
    sh = np.zeros(n_cells, np.uint8)
    s = dp.csda_energy_quadrature(32, 1.0, 1e4)
    m = MaterialConfig(0, 1, 2, -1, ph.scalars[0, 0], ph.scalars[1, 0])
    site = SiteConfig(800.0, 3e4)
    valid = np.arange(0, paths.n_rays, 2)
    model = dp.make_csda_operator(ph, sh, m, site, paths, s, n_cells, valid, ctx=ctx)
    pred, J = model(w)
    full = dp.compute_directional_flux_csda(ph, sh, m, site, paths, w, s, ctx=ctx)
    assert np.array_equal(pred, full[valid])
    counts = np.bincount(J.indices, minlength=n_cells)
    free = counts >= 2
    obs = pred * (1.0 + 0.05 * rng.standard_normal(len(pred)))
    W = 1.0 / np.maximum(obs, 1e-12)
    _say(f"normal eqs (w-model): {J.shape[0]} bins, {J.nnz} entries, {free.sum()} free cells")
    N, g = dp.normal_equations(J, pred - obs, W, free, ctx=ctx)
    No, go = O.normal_equations(J, pred - obs, W, np.flatnonzero(free))
    assert np.array_equal(N, No) and np.array_equal(g, go)
    N0, g0 = dp.normal_equations(J, pred - obs, None, None, ctx=ctx)
    No0, go0 = O.normal_equations(J, pred - obs, None, None)
    assert np.array_equal(N0, No0) and np.array_equal(g0, go0)


def test_multi_tile_and_long_column(ctx):
    J = _rand_csr(4000, 6000, 40, 5, dense_col=17)   # > 16 384 digit counts: a multi-chunk scan
    rng = np.random.default_rng(6)
    W = rng.uniform(0.2, 3.0, 4000)
    r = rng.normal(0.0, 1.0, 4000)
    _say(f"normal eqs (random): {J.shape}, {J.nnz} entries, 6000 active (several row tiles)")
    N, g = dp.normal_equations(J, r, W, None, ctx=ctx)
    No, go = O.normal_equations(J, r, W, None)
    assert np.array_equal(N, No) and np.array_equal(g, go)
    act = np.flatnonzero(rng.random(6000) < 0.8)     # 4800 active: several row tiles
    N, g = dp.normal_equations(J, r, None, act, ctx=ctx)
    No, go = O.normal_equations(J, r, None, act)
    assert np.array_equal(N, No) and np.array_equal(g, go)


def test_long_rows(ctx):
    """Rows longer than one wave (up to 200 entries) and columns with more than
    64 rows: the row walk fetches row bounds in batches of 64 and each row in
    chunks of 64 entries."""
    J = _rand_csr(700, 1500, 200, 12, dense_col=3)
    rng = np.random.default_rng(13)
    W, r = rng.uniform(0.2, 3.0, 700), rng.normal(0.0, 1.0, 700)
    act = np.flatnonzero(rng.random(1500) < 0.7)
    _say(f"normal eqs (long rows): {J.shape}, {J.nnz} entries, {len(act)} active")
    N, g = dp.normal_equations(J, r, W, act, ctx=ctx)
    No, go = O.normal_equations(J, r, W, act)
    assert np.array_equal(N, No) and np.array_equal(g, go)


def test_device_entry_point(ctx):
    from _hip import DeviceArray as D
    J = _rand_csr(400, 300, 20, 8)
    rng = np.random.default_rng(9)
    W, r = rng.uniform(0.5, 2.0, 400), rng.normal(0.0, 1.0, 400)
    act = np.flatnonzero(rng.random(300) < 0.5).astype(np.int32)
    rp, col = D(J.indptr.astype(np.int32)), D(J.indices.astype(np.int32))
    val, dW, dr, da = D(J.data), D(W), D(r), D(act)
    n = len(act)
    csr = L.Csr(400, 300, J.nnz, rp.ptr, col.ptr, val.ptr)
    No, go = O.normal_equations(J, r, W, act)
    for _ in range(2):
        out_n = D(np.full(n * n, np.nan))
        out_g = D(np.full(n, np.nan))
        L.check(ctx.lib.dp_normal_equations_dev(ctx.h, C.byref(csr), dW.ptr, dr.ptr, n, da.ptr,
                                                out_n.ptr, out_g.ptr))
        L.check(ctx.lib.dp_synchronize(ctx.h))
        assert np.array_equal(out_n.numpy().reshape(n, n).T, No)
        assert np.array_equal(out_g.numpy(), go)


def test_skipped_entries_and_gradient_only(ctx):
    J = _rand_csr(60, 50, 8, 10).tocoo()
    d = J.data.copy()
    d[::5] = np.nan
    d[2::9] = 0.0
    d[4::11] = -np.inf
    Jb = sp.csr_matrix((d, (J.row, J.col)), shape=J.shape)
    r = np.linspace(-1.0, 1.0, 60)
    N, g = dp.normal_equations(Jb, r, None, None, ctx=ctx)
    No, go = O.normal_equations(Jb, r, None, None)
    assert np.isfinite(N).all() and np.array_equal(N, No) and np.array_equal(g, go)
    N, g = dp.normal_equations(Jb, r, normal=False, ctx=ctx)
    assert N is None and np.array_equal(g, go)
    N, g = dp.normal_equations(Jb, None, None, np.zeros(50, bool), ctx=ctx)
    assert N.shape == (0, 0) and g is None


def test_argument_errors(ctx):
    rp = np.array([0, 2, 3], np.int32)
    col = np.array([1, 0, 2], np.int32)       # row 0 columns descending
    val = np.ones(3)
    out = np.zeros(9)
    bad = L.Csr(2, 3, 3, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    with pytest.raises(dp.ArgumentError):
        L.check(ctx.lib.dp_normal_equations(ctx.h, C.byref(bad), None, None, 3, None,
                                            out.ctypes.data, None))
    col_ok = np.array([0, 1, 2], np.int32)
    good = L.Csr(2, 3, 3, rp.ctypes.data, col_ok.ctypes.data, val.ctypes.data)
    act = np.array([2, 1], np.int32)          # not ascending
    with pytest.raises(dp.ArgumentError):
        L.check(ctx.lib.dp_normal_equations(ctx.h, C.byref(good), None, None, 2, act.ctypes.data,
                                            out.ctypes.data, None))
    with pytest.raises(dp.ArgumentError):     # the gradient needs a residual
        L.check(ctx.lib.dp_normal_equations(ctx.h, C.byref(good), None, None, 3, None, None,
                                            out.ctypes.data))
    with pytest.raises(dp.ArgumentError):
        dp.normal_equations(sp.csr_matrix((2, 3)), np.zeros(2), np.ones(3), ctx=ctx)

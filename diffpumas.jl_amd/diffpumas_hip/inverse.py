"""Gauss-Newton building blocks of the tomography inversion on the GPU.

Host mirror of make_csda_operator (src/Tomography_inverse.jl:614-626) and of
the normal-equation assembly inside gauss_newton_reconstruct (:489-494, :595)
and laplace_posterior (:922-924).  The products come from the HIP kernels
behind dp_normal_equations; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import ArgumentError, f64, i32
from .tomography import assemble_forward_and_jacobian


def make_csda_operator(physics, shallow_flags, matcfg, site, paths, energy_samples, n_cells,
                       valid_bins, ctx=None):
    """make_csda_operator: a closure model(w) -> (pred, J) over the paths of
    valid_bins (0-based), J a scipy CSR matrix (bins x cells)."""
    sub = paths.select(valid_bins)

    def model(w):
        return assemble_forward_and_jacobian(physics, shallow_flags, matcfg, site, sub, w,
                                             energy_samples, n_cells, ctx=ctx)
    return model


def _csr_parts(J):
    import scipy.sparse as sp
    if not sp.issparse(J):
        raise ArgumentError("J must be a scipy.sparse matrix")
    J = sp.csr_matrix(J, dtype=np.float64, copy=True)
    J.sum_duplicates()          # as sparse(I, J, V) in the reference; also sorts the columns
    J.sort_indices()
    return J, i32(J.indptr), i32(J.indices), f64(J.data)


def normal_equations(J, residual=None, weights=None, active=None, normal=True, ctx=None):
    """(normal, grad) of the weighted Gauss-Newton system restricted to the
    active columns:

        Ja = J[:, active]; normal = Ja' (W .* Ja); grad = Ja' (W .* residual)

    as gauss_newton_reconstruct forms them (weights None: unweighted).  active
    is a boolean mask over the columns (the reference's `free`), ascending
    0-based indices, or None for every column (laplace_posterior's JtWJ).
    normal=False returns (None, grad); residual None returns (normal, None).
    Entries of J that are zero or not finite are skipped."""
    ctx = ctx or _lib.default_context()
    J, rp, col, val = _csr_parts(J)
    n_rows, n_cols = J.shape
    if active is None:
        act = None
        n_act = n_cols
    else:
        a = np.asarray(active)
        if a.dtype == bool:
            if a.shape != (n_cols,):
                raise ArgumentError("free mask must match the number of cells")
            a = np.flatnonzero(a)
        act = i32(a)
        n_act = len(act)
    w = None if weights is None else f64(weights)
    if w is not None and w.shape != (n_rows,):
        raise ArgumentError("weights must have one entry per bin")
    r = None if residual is None else f64(residual)
    if r is not None and r.shape != (n_rows,):
        raise ArgumentError("residual must have one entry per bin")
    out_n = np.zeros(n_act * n_act) if normal else None
    out_g = np.zeros(n_act) if r is not None else None
    csr = _lib.Csr(n_rows, n_cols, len(val), rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    p = lambda x: None if x is None else x.ctypes.data
    _lib.check(ctx.lib.dp_normal_equations(ctx.h, C.byref(csr), p(w), p(r), n_act, p(act),
                                           p(out_n), p(out_g)))
    # the library writes column-major (Julia's Matrix layout)
    return (None if out_n is None else out_n.reshape(n_act, n_act).T), out_g

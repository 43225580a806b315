"""Fit of the MC-calibrated CSDA correction: host mirror of
calibrate_csda_correction (src/Tomography.jl:1043-1179).

The fit is host logic around the forward model: every objective evaluation
installs a trial CsdaCorrection and evaluates the corrected CSDA flux of all
calibration bins in one launch of the w-model kernel
(compute_directional_flux_csda).  The host arithmetic (bin order of the sums,
the coarse-to-fine grid, the ridged 4-parameter least squares) follows the
reference step by step.
"""
import math
from fractions import Fraction

import numpy as np

from . import tomography as T
from ._lib import ArgumentError, TransportError


def path_rock_length(paths, b):
    """path_rock_length (src/Tomography.jl:393-399) of bin b: remaining rock
    plus the segment lengths, summed in path order."""
    total = float(paths.remaining_rock[b])
    for s in range(int(paths.row_ptr[b]), int(paths.row_ptr[b + 1])):
        total += float(paths.lengths[s])
    return total


def float_range(start, stop, length):
    """collect(range(start, stop; length)) for Float64: element i is the
    correctly rounded start + (stop - start) (i - 1) / (length - 1), which is
    what Julia's twice-precision ranges return."""
    n = int(length)
    if n == 1:
        return [float(start)]
    a, b = Fraction(float(start)), Fraction(float(stop))
    return [float(a + (b - a) * i / (n - 1)) for i in range(n)]


def calibrate_csda_correction(physics, shallow_flags, matcfg, site, paths, water_fractions,
                              energy_samples, mc_fluxes, bin_indices, kh_grid=None, refine=4,
                              fit_geometric=True, alpha_strag=1.0, kh_anchor=None,
                              kh_anchor_weight=0.0, ridge_geom=1.0e-3, ctx=None, _flux=None):
    """calibrate_csda_correction: fit kappa_hard (coarse-to-fine 1-D grid,
    kappa_strag pinned at alpha_strag) and, with fit_geometric, the residual
    exp(a + b logL + c logL^2 + d cos(zenith)) so that the corrected CSDA flux
    matches mc_fluxes on the bins bin_indices (0-based).  Installs the result
    with set_csda_correction and returns (alpha_strag, kappa_hard, stats).

    _flux(sub_paths) -> fluxes may replace the GPU evaluation (tests use it to
    run the same fit on the oracle's fluxes)."""
    if len(mc_fluxes) != len(bin_indices):
        raise ArgumentError("mc_fluxes and bin_indices must have the same length")
    valid = [i for i in range(len(bin_indices))
             if math.isfinite(mc_fluxes[i]) and mc_fluxes[i] > 0]
    if not valid:
        raise TransportError("calibrate_csda_correction: no positive MC reference fluxes")
    logmc = [math.log(float(mc_fluxes[i])) for i in valid]
    bins = [int(bin_indices[i]) for i in valid]
    sub = paths.select(bins)

    def fluxes():
        if _flux is not None:
            return _flux(sub)
        return T.compute_directional_flux_csda(physics, shallow_flags, matcfg, site, sub,
                                               water_fractions, energy_samples, ctx=ctx)

    cur = T.get_csda_correction()
    fa, fb, fc, fd = ((0.0, 0.0, 0.0, 0.0) if fit_geometric else
                      (cur.resid_a, cur.resid_b, cur.resid_c, cur.resid_d))

    def set_corr(kh):
        T.set_csda_correction(enabled=True, kappa_strag=alpha_strag, kappa_hard=kh,
                              resid_a=fa, resid_b=fb, resid_c=fc, resid_d=fd)

    def anchor_pen(kh):
        if kh_anchor is None or kh_anchor_weight <= 0.0:
            return 0.0
        return kh_anchor_weight * (kh - kh_anchor) ** 2

    def objective(kh):
        set_corr(kh)
        f = fluxes()
        s, n = 0.0, 0
        for j in range(len(bins)):
            fj = float(f[j])
            if not (math.isfinite(fj) and fj > 0):
                continue
            d = math.log(fj) - logmc[j]
            s += d * d
            n += 1
        return math.inf if n == 0 else s / n + anchor_pen(kh)

    def rel_only():
        f = fluxes()
        acc, n = 0.0, 0
        for j in range(len(bins)):
            fj = float(f[j])
            if not (math.isfinite(fj) and fj > 0):
                continue
            acc += abs(1.0 - fj / math.exp(logmc[j]))
            n += 1
        return math.nan if n == 0 else acc / n

    # "before": pure CSDA (no straggling enhancement, no hard-loss deficit)
    T.set_csda_correction(enabled=True, kappa_strag=0.0, kappa_hard=0.0, resid_a=fa,
                          resid_b=fb, resid_c=fc, resid_d=fd)
    rel_before = rel_only()

    best_kh, best_obj = 0.0, math.inf
    khg = float_range(0.0, 80.0, 41) if kh_grid is None else [float(k) for k in kh_grid]
    for _ in range(int(refine) + 1):
        for kh in khg:
            o = objective(kh)
            if o < best_obj:
                best_obj, best_kh = o, kh
        dkh = (khg[1] - khg[0]) if len(khg) > 1 else 0.5
        khg = float_range(max(0.0, best_kh - dkh), best_kh + dkh, 9)
    best_ks = alpha_strag

    ra, rb, rc, rd = fa, fb, fc, fd
    if fit_geometric:
        T.set_csda_correction(enabled=True, kappa_strag=best_ks, kappa_hard=best_kh)
        f = fluxes()
        feats, ys = [], []
        for j, b in enumerate(bins):
            fj = float(f[j])
            if not (math.isfinite(fj) and fj > 0):
                continue
            logL = math.log(max(path_rock_length(paths, b), 1.0))
            feats.append([1.0, logL, logL * logL, math.cos(math.radians(paths.zenith[b]))])
            ys.append(logmc[j] - math.log(fj))
        nfit = len(ys)
        ra = rb = rc = rd = 0.0
        if nfit >= 4:
            A = np.array(feats)
            R = np.diag([0.0, ridge_geom, ridge_geom, ridge_geom])
            coef = np.linalg.solve(A.T @ A + R, A.T @ np.array(ys))
            ra, rb, rc, rd = (float(x) for x in coef)
        elif nfit > 0:
            ra = sum(ys) / nfit

    T.set_csda_correction(enabled=True, kappa_strag=best_ks, kappa_hard=best_kh, resid_a=ra,
                          resid_b=rb, resid_c=rc, resid_d=rd)
    rel_after = rel_only()
    stats = dict(rel_before=rel_before, rel_after=rel_after, mse_log=best_obj,
                 n_bins=len(bins), kappa_strag=best_ks, kappa_hard=best_kh, resid_a=ra,
                 resid_b=rb, resid_c=rc, resid_d=rd)
    return best_ks, best_kh, stats

"""Pins of the oracle's MC-calibrated CSDA corrections of the w model.

CsdaCorrection and CsdaMaterialCorrection (src/Tomography.jl:112-172, 360-412,
555-580) are off by default; the reference tests hold no numbers for them
(test/test_tomography_material.jl:11-24 only restates the multiplier).  They
are pinned here by what the mathematics fixes:
  * switched off, the flux and Jacobian are bit-identical to the plain model;
  * the spectral index is the logarithmic slope of the surface flux
    (exact tangent of flux_gccly), and runs from ~2 at 10 GeV to ~3.7 at
    10 TeV as the reference states;
  * V and H of one node equal the per-segment sums of straggling / hard loss
    at the mid energy times the segment grammage, each recomputed here from
    table lookups and the segment's own end energy;
  * the geometric residual is L^b, exp(c log^2 L), exp(d cos theta) in closed
    form, with L the rock path length;
  * the material multiplier of a path of equal-w cells is exp(a w + b w^2)
    (column water fraction = w), and 1 on a dry path;
  * the reverse sweep with the correction partials equals forward mode on the
    whole program and central differences.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip._lib import T_CSDA_STOP, T_MIXED_STOP, T_STRAGGLING
from test_oracle_wmodel import _paths, _samples, SITE, WFIELD

SHALLOW4 = np.array([0, 1, 0, 1], np.uint8)


@pytest.fixture(scope="module")
def ph4():
    return dp.create_physics(("StandardRock", "Water", "Air"))


@pytest.fixture(scope="module")
def mc_plain(ph4):
    return O.matcfg_arrays(0, 1, 2, -1, ph4.scalars[0, 0], ph4.scalars[1, 0])


@pytest.fixture(scope="module")
def mc_porous(ph4):
    # porous recipe: half rock, half "porous" (water tables at 1500 kg/m3) above 40 m
    return O.matcfg_arrays(0, 1, 2, 1, ph4.scalars[0, 0], ph4.scalars[1, 0], 1500.0, 40.0)


CORR = O.corr_array(kappa_strag=1.1, kappa_hard=0.6, resid_a=0.05, resid_b=-0.02,
                    resid_c=0.003, resid_d=0.1, material=(-0.8, 1.7))
PATH_R = (25.0, [(0, 12.0), (2, 30.0), (0, 8.0), (3, 20.0)], 40.0, 80.0, 1)


def test_disabled_is_plain_model(ph4, mc_plain):
    ne, nw, nq = _samples(8, 1e-2, 1e5, 3)
    paths = _paths([PATH_R, (35.0, [(1, 50.0), (3, 35.0)], 0.0, 55.0, 1)])
    off = O.corr_array(enabled=False, kappa_strag=3.0, kappa_hard=2.0, resid_a=1.0)
    f0, j0 = O.w_jacobian(ph4, mc_plain, SITE, paths, WFIELD, SHALLOW4, ne, nw, nq)
    f1, j1 = O.w_jacobian(ph4, mc_plain, SITE, paths, WFIELD, SHALLOW4, ne, nw, nq, corr=off)
    assert np.array_equal(f0, f1) and np.array_equal(j0, j1)
    f2 = O.w_flux(ph4, mc_plain, SITE, paths, WFIELD, SHALLOW4, ne, nw, nq, corr=off)
    assert np.array_equal(f0, f2)


@pytest.mark.parametrize("energy", [3.0, 10.0, 100.0, 1e3, 1e4])
@pytest.mark.parametrize("cos_t", [1.0, 0.4])
def test_spectral_index_is_log_slope(energy, cos_t):
    b = O.spectral_index(energy, cos_t)
    exact = -energy * O.dflux_gccly_dk(cos_t, energy, 1.0) / O.flux_gccly(cos_t, energy, 1.0)
    # symmetric difference at eta = 1e-2: error O(eta^2) of the slope
    assert b == pytest.approx(exact, rel=2e-4)


def test_spectral_index_range():
    """src/Tomography.jl:360-366: b runs ~2 at 10 GeV to ~3.7 above 1 TeV."""
    assert 1.7 < O.spectral_index(10.0, 1.0) < 2.4
    assert 3.55 < O.spectral_index(1e4, 1.0) < 3.75
    assert O.spectral_index(1e4, 1.0) > O.spectral_index(1e3, 1.0) > O.spectral_index(10.0, 1.0)


def test_var_correction_terms():
    """exp(+kappa_s b(b+1)/2 V/E^2 - kappa_h max(H,0)/E) with E floored at 1."""
    c = O.corr_array(kappa_strag=0.9, kappa_hard=0.4)
    b = O.spectral_index(50.0, 0.8)
    got = O.var_correction(50.0, 30.0, 2.0, 0.8, c)
    assert got == pytest.approx(math.exp(0.9 * 0.5 * b * (b + 1) * 30.0 / 2500.0 - 0.4 * 2.0 / 50.0),
                                rel=1e-14)
    assert got < math.exp(0.9 * 0.5 * b * (b + 1) * 30.0 / 2500.0)   # hard loss is a deficit
    assert O.var_correction(50.0, 30.0, -5.0, 0.8, c) == pytest.approx(
        math.exp(0.9 * 0.5 * b * (b + 1) * 30.0 / 2500.0), rel=1e-14)
    b1 = O.spectral_index(1.0, 0.8)
    assert O.var_correction(0.3, 0.2, 0.1, 0.8, c) == pytest.approx(
        math.exp(0.9 * 0.5 * b1 * (b1 + 1) * 0.2 - 0.4 * 0.1), rel=1e-14)
    assert O.var_correction(50.0, 30.0, 2.0, 0.8, O.corr_array(enabled=False, kappa_strag=1.0)) == 1.0


def _lookup(ph, m, col, e):
    return O.property_(ph, m, col, [e])[0][0]


def _mix(ph, mc, w, sh, col, e):
    """Linear mass-fraction mixing of a table column (cell_straggling /
    cell_stopping_power recipes)."""
    mi, md = mc
    rock, water, porous = int(mi[0]), int(mi[1]), int(mi[3])
    if sh and porous >= 0 and md[3] > 0:
        return (0.5 * (1 - w) * _lookup(ph, porous, col, e) + 0.5 * (1 - w) * _lookup(ph, rock, col, e)
                + w * _lookup(ph, water, col, e))
    return (1 - w) * _lookup(ph, rock, col, e) + w * _lookup(ph, water, col, e)


def _rho(mc, w, sh):
    mi, md = mc
    if sh and mi[3] >= 0 and md[3] > 0:
        return 1.0 / (0.5 * (1 - w) / md[2] + 0.5 * (1 - w) / md[0] + w / md[1])
    return 1.0 / ((1 - w) / md[0] + w / md[1])


@pytest.mark.parametrize("porous", [False, True])
def test_node_V_H_are_segment_sums(ph4, mc_plain, mc_porous, porous):
    mc = mc_porous if porous else mc_plain
    segs = [(0, 12.0), (1, 30.0), (3, 20.0)]
    rem, ez, ct = 40.0, 80.0, math.cos(math.radians(25.0))
    cells = [c for c, _ in segs]
    lens = [d for _, d in segs]
    e0 = 30.0
    out = O.w_node_walk(ph4, mc, SITE, cells, lens, rem, ez, ct, WFIELD, SHALLOW4, e0)
    assert out[4] == 1.0
    V = H = 0.0
    e = e0
    for c, d in segs:
        w, sh = WFIELD[c], bool(SHALLOW4[c])
        seg = O.w_cell_segment(ph4, mc, w, sh, d, e)
        er = 0.5 * (e + seg[0])
        gx = d * _rho(mc, w, sh)
        V += _mix(ph4, mc, w, sh, T_STRAGGLING, er) * gx
        H += (_mix(ph4, mc, w, sh, T_CSDA_STOP, er) - _mix(ph4, mc, w, sh, T_MIXED_STOP, er)) * gx
        e = seg[0]
    # remainder: a dry rock cell of the same length crosses to the same energy
    seg = O.w_cell_segment(ph4, mc_plain, 0.0, False, rem, e)
    er = 0.5 * (e + seg[0])
    gx = rem * float(mc[1][0])
    V += _lookup(ph4, 0, T_STRAGGLING, er) * gx
    H += (_lookup(ph4, 0, T_CSDA_STOP, er) - _lookup(ph4, 0, T_MIXED_STOP, er)) * gx
    assert out[2] == pytest.approx(V, rel=1e-12)
    assert out[3] == pytest.approx(H, rel=1e-12)
    assert V > 0 and H > 0


def _flux1(ph, mc, spec, w, sh, corr, nodes):
    ne, nw, nq = nodes
    return O.w_flux(ph, mc, SITE, _paths([spec]), w, sh, ne, nw, nq, corr=corr)[0]


@pytest.mark.parametrize("rem", [40.0, 0.3])
def test_geometric_residual_closed_form(ph4, mc_plain, rem):
    nodes = _samples(6, 1e-1, 1e4, 5)
    spec = (25.0, [(0, 12.0), (2, 30.0)], rem, 80.0, 1)
    base = _flux1(ph4, mc_plain, spec, WFIELD, SHALLOW4, None, nodes)
    L = 12.0 + 30.0 + rem
    ct = math.cos(math.radians(25.0))
    rb = _flux1(ph4, mc_plain, spec, WFIELD, SHALLOW4, O.corr_array(resid_b=0.7), nodes)
    assert rb / base == pytest.approx(L ** 0.7, rel=1e-13)
    rc = _flux1(ph4, mc_plain, spec, WFIELD, SHALLOW4, O.corr_array(resid_c=-0.05), nodes)
    assert rc / base == pytest.approx(math.exp(-0.05 * math.log(L) ** 2), rel=1e-13)
    rd = _flux1(ph4, mc_plain, spec, WFIELD, SHALLOW4, O.corr_array(resid_a=0.2, resid_d=0.3), nodes)
    assert rd / base == pytest.approx(math.exp(0.2 + 0.3 * ct), rel=1e-13)
    # a path shorter than 1 m has log(max(L, 1)) = 0
    short = (25.0, [(0, 0.4)], 0.3, 80.0, 1)
    b2 = _flux1(ph4, mc_plain, short, WFIELD, SHALLOW4, None, nodes)
    assert _flux1(ph4, mc_plain, short, WFIELD, SHALLOW4, O.corr_array(resid_b=0.7), nodes) == \
        pytest.approx(b2, rel=1e-14)


def test_material_multiplier_closed_form(ph4, mc_plain):
    nodes = _samples(6, 1e-1, 1e4, 5)
    spec = (30.0, [(0, 20.0), (1, 15.0), (2, 30.0)], 0.0, 70.0, 1)
    w = np.full(4, 0.23)
    base = _flux1(ph4, mc_plain, spec, w, SHALLOW4, None, nodes)
    got = _flux1(ph4, mc_plain, spec, w, SHALLOW4, O.corr_array(enabled=False, material=(-0.8, 1.7)),
                 nodes)
    assert got / base == pytest.approx(math.exp(-0.8 * 0.23 + 1.7 * 0.23 ** 2), rel=1e-13)
    dry = np.zeros(4)
    b0 = _flux1(ph4, mc_plain, spec, dry, SHALLOW4, None, nodes)
    assert _flux1(ph4, mc_plain, spec, dry, SHALLOW4, O.corr_array(enabled=False, material=(-0.8, 1.7)),
                  nodes) == b0
    # with a dry remainder the column fraction is diluted by its mass
    spec2 = (30.0, [(0, 20.0)], 50.0, 70.0, 1)
    rho_r, rho_w = float(mc_plain[1][0]), float(mc_plain[1][1])
    gx = 20.0 / ((1 - 0.23) / rho_r + 0.23 / rho_w)
    q = 0.23 * gx / (gx + 50.0 * rho_r)
    r = _flux1(ph4, mc_plain, spec2, w, SHALLOW4, O.corr_array(enabled=False, material=(-0.8, 1.7)),
               nodes) / _flux1(ph4, mc_plain, spec2, w, SHALLOW4, None, nodes)
    assert r == pytest.approx(math.exp(-0.8 * q + 1.7 * q * q), rel=1e-13)


@pytest.mark.parametrize("spec", [PATH_R, (35.0, [(1, 50.0), (3, 35.0)], 0.0, 55.0, 1),
                                  (15.0, [(3, 45.0), (1, 10.0), (3, 5.0)], 20.0, 30.0, 1)])
@pytest.mark.parametrize("porous", [False, True])
def test_reverse_equals_forward_mode_corrected(ph4, mc_plain, mc_porous, spec, porous):
    mc = mc_porous if porous else mc_plain
    ne, nw, nq = _samples(6, 1e-2, 1e5, 42)
    paths = _paths([spec])
    f, jac = O.w_jacobian(ph4, mc, SITE, paths, WFIELD, SHALLOW4, ne, nw, nq, corr=CORR)
    fp = O.w_flux(ph4, mc, SITE, paths, WFIELD, SHALLOW4, ne, nw, nq, corr=CORR)
    f0 = O.w_flux(ph4, mc, SITE, paths, WFIELD, SHALLOW4, ne, nw, nq)
    assert f[0] == pytest.approx(fp[0], rel=1e-14)
    assert abs(f[0] / f0[0] - 1) > 1e-3    # the corrections do act
    zen, segs, rem, ez, _ = spec
    cells = [c for c, _ in segs]
    lens = [d for _, d in segs]
    ct = math.cos(math.radians(zen))
    seen = set()
    for k, c in enumerate(cells):
        if c in seen:
            assert jac[k] == 0.0
            continue
        seen.add(c)
        ref = O.w_dflux(ph4, mc, SITE, cells, lens, rem, ez, ct, WFIELD, SHALLOW4, c, ne, nw, nq,
                        corr=CORR)
        assert jac[k] == pytest.approx(ref, rel=1e-9, abs=1e-300)
        assert ref != 0.0


def test_reverse_matches_central_difference_corrected(ph4, mc_porous):
    ne, nw, nq = _samples(6, 1e-2, 1e5, 42)
    paths = _paths([PATH_R])
    _, jac = O.w_jacobian(ph4, mc_porous, SITE, paths, WFIELD, SHALLOW4, ne, nw, nq, corr=CORR)
    h = 1e-5
    for k, c in [(0, 0), (1, 2), (3, 3)]:
        wp, wm = WFIELD.copy(), WFIELD.copy()
        wp[c] += h
        wm[c] -= h
        fd = (O.w_flux(ph4, mc_porous, SITE, paths, wp, SHALLOW4, ne, nw, nq, corr=CORR)[0] -
              O.w_flux(ph4, mc_porous, SITE, paths, wm, SHALLOW4, ne, nw, nq, corr=CORR)[0]) / (2 * h)
        assert jac[k] == pytest.approx(fd, rel=1e-5)


def test_setters_mirror_reference():
    """set_csda_correction! / set_csda_material_correction! (src/Tomography.jl:
    149-172): defaults off; non-finite material coefficients are an
    ArgumentError.  The C ABI keeps them on the context."""
    assert not dp.get_csda_correction().enabled
    assert not dp.get_csda_material_correction().enabled
    try:
        c = dp.set_csda_correction(kappa_strag=1.0, resid_b=0.5)
        assert c.enabled and c.kappa_strag == 1.0 and c.resid_b == 0.5 and c.resid_c == 0.0
        with pytest.raises(dp.ArgumentError):
            dp.set_csda_material_correction(linear=float("inf"))
        assert not dp.get_csda_material_correction().enabled
    finally:
        dp.set_csda_correction(enabled=False)
        dp.set_csda_material_correction(enabled=False)

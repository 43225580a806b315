"""Pins of the oracle's water-fraction (w) tomography CSDA model.

The reference holds no numbers for this path, only closed forms and
self-consistency (test/test_shallow_mixture.jl:53-74, test/test_tomography.jl:44-141).
They are restated here against the oracle:
  * constant-stopping-power slab: E_out = E_in + d rho(w) S(w) and the analytic
    dE/dw (mixture density is harmonic, stopping power linear in w);
  * atmosphere with constant stopping power: the mid-step density sum is a
    geometric series in the 60 m step;
  * the charge-summed energy quadrature equals explicit +/- nodes;
  * the reverse sweep over segment partials equals forward mode on the whole
    program, is zero off path, and matches central differences.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip.physics import PhysicsTables, T_CSDA_STOP
from workloads import energy_nodes


@pytest.fixture(scope="module")
def ph3():
    return dp.create_physics_csda(("StandardRock", "Water", "Air"))


def _const_tables(ph, values):
    data = ph.data.copy()
    for m, v in enumerate(values):
        data[m, T_CSDA_STOP, :] = v
    return PhysicsTables(ph.names, data, ph.scalars, ph.mass, ph.ctau)


CSDA = (0.002, 0.004, 0.008)


@pytest.mark.parametrize("w", [0.0, 1e-12, 0.3, 1.0])
def test_slab_closed_form_shallow(ph3, w):
    """test/test_shallow_mixture.jl:53-74: porous recipe, constant tables."""
    ph = _const_tables(ph3, CSDA)
    mc = O.matcfg_arrays(0, 1, 2, 2, 2700.0, 1000.0, 1500.0, 50.0)
    d, e0 = 0.2, 10.0
    rho = 1.0 / (0.5 * (1 - w) / 2700.0 + 0.5 * (1 - w) / 1500.0 + w / 1000.0)
    s = 0.5 * (1 - w) * CSDA[2] + 0.5 * (1 - w) * CSDA[0] + w * CSDA[1]
    out = O.w_cell_segment(ph, mc, w, True, d, e0)
    assert out[2] == 1.0
    assert out[0] == pytest.approx(e0 + d * rho * s, rel=1e-14)
    a = 0.5 / 2700.0 + 0.5 / 1500.0
    b = 1 / 1000.0 - a
    s0 = 0.5 * (CSDA[0] + CSDA[2])
    ds = CSDA[1] - s0
    assert out[3] == pytest.approx(d * (ds * a - s0 * b) / (a + b * w) ** 2, rel=1e-12)
    assert out[4] == pytest.approx(1.0, rel=1e-14)   # constant S: dE_out/dE_in = 1


@pytest.mark.parametrize("w", [0.0, 0.25, 0.9])
def test_slab_closed_form_plain(ph3, w):
    ph = _const_tables(ph3, CSDA)
    mc = O.matcfg_arrays(0, 1, 2, -1, 2650.0, 1000.0)
    d, e0 = 35.0, 50.0
    a, b = 1 / 2650.0, 1 / 1000.0 - 1 / 2650.0
    rho = 1.0 / ((1 - w) * a + w / 1000.0)
    s = (1 - w) * CSDA[0] + w * CSDA[1]
    out = O.w_cell_segment(ph, mc, w, False, d, e0)
    assert out[0] == pytest.approx(e0 + d * rho * s, rel=1e-13)
    ds = CSDA[1] - CSDA[0]
    assert out[3] == pytest.approx(d * (ds * a - CSDA[0] * b) / (a + b * w) ** 2, rel=1e-12)
    # shallow flag without a porous material is the plain recipe
    out2 = O.w_cell_segment(ph, mc, w, True, d, e0)
    assert np.array_equal(out, out2)


@pytest.mark.parametrize("cos_t", [1.0, 0.5])
def test_air_geometric_series(ph3, cos_t):
    """60 m steps at the mid-step altitude of rho0 exp(-z/12 km)."""
    s_air = 2e-4
    ph = _const_tables(ph3, (CSDA[0], CSDA[1], s_air))
    site = (1000.0, 12000.0)
    z0, n, e0 = 60.0, 10, 10.0
    out = O.w_air_segment(ph, 2, site, z0, cos_t, 60.0 * n, e0)
    h0 = site[0] + z0
    r = math.exp(-60.0 * cos_t / 12e3)
    series = math.exp(-(h0 + 30.0 * cos_t) / 12e3) * (1 - r ** n) / (1 - r)
    assert out[2] == 1.0
    assert out[0] == pytest.approx(e0 + s_air * 60.0 * 1.205 * series, rel=1e-12)
    assert out[3] == pytest.approx(1.0, rel=1e-14)


def test_air_minimum_step(ph3):
    """A 0.1 mm remainder is crossed with the 1 mm floor step."""
    s_air = 2e-4
    ph = _const_tables(ph3, (CSDA[0], CSDA[1], s_air))
    out = O.w_air_segment(ph, 2, (0.0, 3e4), 100.0, 1.0, 1e-4, 5.0)
    rho = 1.205 * math.exp(-(100.0 + 0.5e-3) / 12e3)
    assert out[0] == pytest.approx(5.0 + s_air * 1e-3 * rho, rel=1e-14)


# ------------------------------------------------------------ path fixture
# test/test_tomography.jl:18-42 (cells 0-based here).
SITE = (1000.0, 12000.0)
WFIELD = np.array([0.10, 0.20, 0.05, 0.15])
SHALLOW = np.zeros(4, np.uint8)


def _paths(specs):
    """specs: (zenith, [(cell, distance)], remaining, exit_z, valid)."""
    rp, cells, lens, rem, ez, ct, va = [0], [], [], [], [], [], []
    for zen, segs, remaining, exit_z, valid in specs:
        for c, d in segs:
            cells.append(c)
            lens.append(d)
        rp.append(len(cells))
        rem.append(remaining)
        ez.append(exit_z)
        ct.append(math.cos(math.radians(zen)))
        va.append(valid)
    return dict(row_ptr=np.array(rp, np.int32), cells=np.array(cells, np.int32),
                lengths=np.array(lens), remaining=np.array(rem), exit_z=np.array(ez),
                cos_theta=np.array(ct), valid=np.array(va, np.uint8))


PATH1 = (20.0, [(0, 30.0), (1, 40.0), (2, 30.0)], 0.0, 60.0, 1)
PATH2 = (35.0, [(1, 50.0), (3, 35.0)], 0.0, 55.0, 1)


def _samples(n, emin, emax, seed):
    """Log-uniform energies with +/-1 charges and weight E log(emax/emin)."""
    rng = np.random.default_rng(seed)
    rk = math.log(emax / emin)
    e = emin * np.exp(rk * rng.random(n))
    q = np.where(rng.random(n) > 0.5, 1.0, -1.0)
    return e, e * rk, q


@pytest.fixture(scope="module")
def setup(ph3):
    mc = O.matcfg_arrays(0, 1, 2, -1, ph3.scalars[0, 0], ph3.scalars[1, 0])
    return ph3, mc


def test_charge_summed_quadrature(setup):
    """test/test_tomography.jl:64-75: charge-0 nodes equal explicit pairs."""
    ph, mc = setup
    ne, nw, nq = energy_nodes(32, 1.0, 100.0)
    pe = np.repeat(ne, 2)
    pw = np.repeat(2 * nw, 2)
    pq = np.tile([-1.0, 1.0], len(ne))
    paths = _paths([PATH1, PATH2])
    f1, j1 = O.w_jacobian(ph, mc, SITE, paths, WFIELD, SHALLOW, ne, nw, nq)
    f2, j2 = O.w_jacobian(ph, mc, SITE, paths, WFIELD, SHALLOW, pe, pw, pq)
    np.testing.assert_allclose(f1, f2, rtol=1e-13)
    np.testing.assert_allclose(j1, j2, rtol=1e-10)


def test_flux_of_jacobian_pass_equals_forward(setup):
    ph, mc = setup
    ne, nw, nq = _samples(6, 1e-3, 1e6, 42)
    paths = _paths([PATH1, PATH2])
    f = O.w_flux(ph, mc, SITE, paths, WFIELD, SHALLOW, ne, nw, nq)
    fj, _ = O.w_jacobian(ph, mc, SITE, paths, WFIELD, SHALLOW, ne, nw, nq)
    assert np.all(f > 0)
    np.testing.assert_allclose(fj, f, rtol=1e-14)


@pytest.mark.parametrize("spec", [PATH1, PATH2,
                                  (25.0, [(0, 12.0), (2, 30.0), (0, 8.0), (3, 20.0)], 40.0, 80.0, 1)])
def test_reverse_equals_forward_mode(setup, spec):
    """Sparse reverse sweep == full-vector forward mode; off-path cells are 0
    (test/test_tomography.jl:87-107).  The third path repeats a cell and has
    a dry-rock remainder."""
    ph, mc = setup
    ne, nw, nq = _samples(6, 1e-3, 1e6, 42)
    paths = _paths([spec])
    _, jac = O.w_jacobian(ph, mc, SITE, paths, WFIELD, SHALLOW, ne, nw, nq)
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
        ref = O.w_dflux(ph, mc, SITE, cells, lens, rem, ez, ct, WFIELD, SHALLOW, c, ne, nw, nq)
        assert jac[k] == pytest.approx(ref, rel=1e-10, abs=1e-300)
        assert ref != 0.0
    for c in set(range(4)) - set(cells):
        assert O.w_dflux(ph, mc, SITE, cells, lens, rem, ez, ct, WFIELD, SHALLOW, c, ne, nw, nq) == 0.0


def test_reverse_matches_central_difference(setup):
    """test/test_tomography.jl:109-121 (h = 1e-5, rtol 1e-4)."""
    ph, mc = setup
    ne, nw, nq = _samples(6, 1e-3, 1e6, 42)
    paths = _paths([PATH1])
    _, jac = O.w_jacobian(ph, mc, SITE, paths, WFIELD, SHALLOW, ne, nw, nq)
    h = 1e-5
    for k, c in enumerate([0, 1, 2]):
        wp, wm = WFIELD.copy(), WFIELD.copy()
        wp[c] += h
        wm[c] -= h
        fd = (O.w_flux(ph, mc, SITE, paths, wp, SHALLOW, ne, nw, nq)[0] -
              O.w_flux(ph, mc, SITE, paths, wm, SHALLOW, ne, nw, nq)[0]) / (2 * h)
        assert jac[k] == pytest.approx(fd, rel=1e-4)


def test_dry_remainder_equals_pure_rock_cell(setup):
    """The dry-rock remainder is a w = 0 plain cell of the same length."""
    ph, mc = setup
    ne, nw, nq = energy_nodes(16, 1.0, 1e3)
    a = _paths([(30.0, [(0, 20.0)], 45.0, 70.0, 1)])
    b = _paths([(30.0, [(0, 20.0), (3, 45.0)], 0.0, 70.0, 1)])
    w = WFIELD.copy()
    w[3] = 0.0
    fa = O.w_flux(ph, mc, SITE, a, w, SHALLOW, ne, nw, nq)
    fb = O.w_flux(ph, mc, SITE, b, w, SHALLOW, ne, nw, nq)
    np.testing.assert_allclose(fa, fb, rtol=1e-12)


def test_invalid_paths_are_nan(setup):
    ph, mc = setup
    ne, nw, nq = energy_nodes(4, 1.0, 100.0)
    paths = _paths([(20.0, [(0, 10.0)], 0.0, 60.0, 0), (95.0, [(1, 10.0)], 0.0, 60.0, 1), PATH2])
    f, j = O.w_jacobian(ph, mc, SITE, paths, WFIELD, SHALLOW, ne, nw, nq)
    assert math.isnan(f[0]) and math.isnan(f[1]) and f[2] > 0
    assert j[0] == 0.0 and j[1] == 0.0

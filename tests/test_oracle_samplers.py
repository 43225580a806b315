"""Pins of the oracle's transport samplers (oracle/oracle_mc.hpp) by what the
mathematics fixes, not by re-deriving their code.  CPU only.

  rotate_direction (src/Straggling.jl:645-710): unit result, polar cosine
      1 - 2 mu, azimuth mapping phi = pi (1 - 2u) with the basis handedness
      u1 = u0 x d read off the cross product of two rotated vectors, and the
      co-vector of each branch (largest |component|, and the ties).
  truncated_randn (:59-78): Kolmogorov-Smirnov against a normal truncated at
      3 sigma, scaled by 1/1.015387.
  fluctuate_energy_loss (:103-173), backward: in the Gaussian and uniform
      regimes the mean final energy is the CSDA energy k1 and the variance is
      the trapezoid estimate dk12 = dX (Omega0 + Omega1) / 2 of the
      accumulated straggling variance (Gaussian: times the truncated-normal
      variance); in the mixed regime the model keeps the mean loss dk0 and
      the variance dk12; the large-step correction factor; draws consumed per
      regime; the energy never decreases going backward.
  sample_soft_scattering (:539-635): mu is exponential with mean
      half the transport-path integral (trapezoid over the step), truncated
      at 1, and capped at 1 for large steps.
  sample_scattering_angle (:446-517): centre-of-mass draws follow the Wentzel
      law (A + mu)^-2 on [mu0, 1] without spin, and (1 - f mu)/(A + mu)^2
      with it (KS); the frame change equals an explicit Lorentz boost of the
      outgoing muon.
  DEL polar angles (:746-914): ionisation = two-body kinematics on an
      electron at rest; photonuclear = the muon angle whose four-momentum
      transfer is the log-uniform Q^2 drawn; bremsstrahlung / pair =
      rejection sampler reproduces its target density (KS).
"""
import math

import numpy as np
import pytest
import scipy.integrate as SI
import scipy.stats as ST

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip import _lib as L

MUON = 0.10565839
ME = 0.510998910e-3
NEUTRON = 0.939565
PROTON = 0.938272
PION = 0.13957018


@pytest.fixture(scope="module")
def ph():
    return dp.create_physics(("StandardRock", "Water", "Air"))


def _prop(ph, m, a, k):
    return O.property_(ph, m, a, [k])[0][0]


# ---------------------------------------------------------------- rotate
DIRS = [(0.9, 0.3, 0.3), (0.3, -0.9, 0.2), (0.2, 0.3, -0.9), (1.0, 1.0, 0.0), (1.0, 0.0, 1.0),
        (0.0, 1.0, 1.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (-0.5, 0.5, 0.5)]


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("d", DIRS)
@pytest.mark.parametrize("mu", [1e-6, 0.1, 0.5, 0.93, 1.0 - 1e-9])
def test_rotate_polar_cosine_and_norm(d, mu):
    d = _unit(d)
    u = np.linspace(0.0, 0.999, 7)
    out = O.rotate(np.tile(d, (len(u), 1)), mu, u)
    np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, atol=1e-15)
    np.testing.assert_allclose(out @ d, 1.0 - 2.0 * mu, atol=2e-15)


@pytest.mark.parametrize("d", DIRS)
def test_rotate_azimuth_mapping_and_handedness(d):
    """phi = pi (1 - 2u) in the basis (u0, u1 = u0 x d): the transverse parts
    t(u) of two rotations satisfy t1.t2 = cos(2 pi du) and
    (t1 x t2).d = sin(2 pi du) with du = u2 - u1."""
    d = _unit(d)
    mu = 0.3
    ct, st = 1 - 2 * mu, math.sqrt(4 * mu * (1 - mu))
    for u1, du in ((0.1, 0.25), (0.37, 0.1), (0.8, -0.45), (0.5, 0.3)):
        a, b = O.rotate(np.array([d, d]), mu, [u1, u1 + du])
        t1, t2 = (a - ct * d) / st, (b - ct * d) / st
        assert np.dot(t1, t1) == pytest.approx(1.0, abs=1e-14)
        assert np.dot(t1, d) == pytest.approx(0.0, abs=1e-14)
        assert np.dot(t1, t2) == pytest.approx(math.cos(2 * math.pi * du), abs=1e-13)
        assert np.dot(np.cross(t1, t2), d) == pytest.approx(math.sin(2 * math.pi * du), abs=1e-13)


@pytest.mark.parametrize("d,zero", [((0.9, 0.3, 0.3), 1), ((0.3, -0.9, 0.2), 2),
                                    ((0.2, 0.3, -0.9), 0), ((1.0, 1.0, 0.0), 2),
                                    ((1.0, 0.0, 1.0), 0), ((0.0, 1.0, 1.0), 0)])
def test_rotate_covector_branches(d, zero):
    """u = 1/2 gives phi = 0, so the transverse part is the co-vector u0 the
    reference picks: orthogonal to the component of largest magnitude's
    partner, i.e. zero in component `zero` (ties go to the later branch:
    |dx| = |dy| > |dz| picks the y branch, |dx| = |dz| the z branch)."""
    d = _unit(d)
    mu = 0.2
    ct, st = 1 - 2 * mu, math.sqrt(4 * mu * (1 - mu))
    t = (O.rotate(np.array([d]), mu, [0.5])[0] - ct * d) / st
    assert abs(t[zero]) < 1e-15
    expect = {1: np.array([-d[2], 0.0, d[0]]), 2: np.array([d[1], -d[0], 0.0]),
              0: np.array([0.0, d[2], -d[1]])}[zero]
    np.testing.assert_allclose(t, _unit(expect), atol=1e-14)


def test_rotate_zero_sine_is_identity():
    """sin^2 theta = 4 mu (1 - mu) <= 0 returns the direction unchanged
    (src/Straggling.jl:653-655): mu = 0, and also mu = 1 exactly, where the
    reference does not reverse the direction."""
    d = _unit((0.2, -0.4, 0.7))
    np.testing.assert_array_equal(O.rotate(np.array([d]), 0.0, [0.3])[0], d)
    np.testing.assert_array_equal(O.rotate(np.array([d]), 1.0, [0.3])[0], d)


# ------------------------------------------------------- truncated_randn
def test_truncated_randn_distribution():
    x = O.truncated_randn(11, 200_000)
    s = 0.9848416472
    assert np.abs(x).max() <= 3.0 * s
    p = ST.kstest(x / s, ST.truncnorm(-3.0, 3.0).cdf).pvalue
    assert p > 1e-3, p


# ----------------------------------------------------------- fluctuation
def _fluct_setup(ph, ki, dX):
    rock = ph.get_material_index("StandardRock")
    X0 = _prop(ph, rock, L.T_MIXED_RANGE, ki)
    k1 = O.kinetic(ph, rock, [X0 + dX], mixed=True)[0][0]
    O0, O1 = _prop(ph, rock, L.T_STRAGGLING, ki), _prop(ph, rock, L.T_STRAGGLING, k1)
    dk12 = 0.5 * dX * (O0 + O1)
    return rock, X0, k1, k1 - ki, dk12


def _regime(dk0, dk12):
    r = dk0 / math.sqrt(dk12)
    return "gauss" if r >= 3 else ("uniform" if r >= 1.7320508 else "mixed")


N_FL = 200_000


def test_fluctuation_gaussian_regime(ph):
    ki, dX = 0.5, 100.0
    rock, X0, k1, dk0, dk12 = _fluct_setup(ph, ki, dX)
    assert _regime(dk0, dk12) == "gauss" and dX / X0 <= 0.05
    k, r = O.fluctuate(ph, [rock], [1.0], ki, dX, X0, seed=3, n=N_FL)
    dk1 = math.sqrt(dk12)
    z = (k - k1) / dk1
    vt = ST.truncnorm(-3.0, 3.0).var() * 0.9848416472 ** 2
    assert abs(z.mean()) < 5 * math.sqrt(vt / N_FL)
    assert z.var() == pytest.approx(vt, rel=0.015)
    assert ST.kstest(z / 0.9848416472, ST.truncnorm(-3.0, 3.0).cdf).pvalue > 1e-3
    assert np.all(k >= ki)
    np.testing.assert_allclose(r, np.abs(ki - k) / dk0, rtol=1e-12)
    _, _, used = O.fluctuate(ph, [rock], [1.0], ki, dX, X0, script=[0.3, 0.6])
    assert used == 2                       # one Box-Muller pair (|u| <= 3)


def test_fluctuation_uniform_regime(ph):
    ki, dX = 0.5, 25.0
    rock, X0, k1, dk0, dk12 = _fluct_setup(ph, ki, dX)
    assert _regime(dk0, dk12) == "uniform"
    k, _ = O.fluctuate(ph, [rock], [1.0], ki, dX, X0, seed=4, n=N_FL)
    z = (k - k1) / math.sqrt(dk12)
    assert np.abs(z).max() <= 1.7320508 + 1e-12
    assert ST.kstest(z, ST.uniform(-1.7320508, 2 * 1.7320508).cdf).pvalue > 1e-3
    assert z.var() == pytest.approx(1.0, rel=0.015)          # variance dk12
    assert np.all(k >= ki)
    kk, _, used = O.fluctuate(ph, [rock], [1.0], ki, dX, X0, script=[0.25])
    assert used == 1 and kk == pytest.approx(k1 + 1.7320508 * 0.5 * math.sqrt(dk12), rel=1e-14)


@pytest.mark.parametrize("ki,dX", [(0.5, 10.0), (10.0, 10.0), (80.0, 1.0)])
def test_fluctuation_mixed_regime_keeps_mean_and_variance(ph, ki, dX):
    rock, X0, k1, dk0, dk12 = _fluct_setup(ph, ki, dX)
    assert _regime(dk0, dk12) == "mixed"
    k, _ = O.fluctuate(ph, [rock], [1.0], ki, dX, X0, seed=5, n=N_FL)
    g = k - ki
    m2 = dk12 + dk0 ** 2                   # second moment of the gain
    assert np.all(g >= 0)
    assert g.mean() == pytest.approx(dk0, abs=5 * math.sqrt(m2 / N_FL))
    assert (g ** 2).mean() == pytest.approx(m2, rel=0.03)
    # the no-loss branch: energy unchanged, one draw; the loss branch: two
    _, _, used = O.fluctuate(ph, [rock], [1.0], ki, dX, X0, script=[0.999999])
    _, _, used2 = O.fluctuate(ph, [rock], [1.0], ki, dX, X0, script=[0.0, 0.5])
    assert used == 1 and used2 == 2


def test_fluctuation_large_step_correction(ph):
    """dX / Xtot > 0.05: dk12 is scaled by 1 - dX/dk0 (S_mixed(k1) -
    S_mixed(ki)) in backward mode (src/Straggling.jl:136-140); here 0.898,
    which a sign slip (1.102) or a dropped factor (1) fails."""
    ki, dX = 0.5, 1e4
    rock, X0, k1, dk0, dk12 = _fluct_setup(ph, ki, dX)
    assert dX / X0 > 0.05
    S0, S1 = _prop(ph, rock, L.T_MIXED_STOP, ki), _prop(ph, rock, L.T_MIXED_STOP, k1)
    f = 1.0 - dX / dk0 * (S1 - S0)
    assert 0.85 < f < 0.95
    k, _ = O.fluctuate(ph, [rock], [1.0], ki, dX, X0, seed=6, n=N_FL)
    vt = ST.truncnorm(-3.0, 3.0).var() * 0.9848416472 ** 2
    assert ((k - k1) ** 2).mean() / (dk12 * f) == pytest.approx(vt, rel=0.015)


# -------------------------------------------------------- soft scattering
@pytest.mark.parametrize("k0,k1,step,rho", [(3.0, 3.2, 2.0, 2650.0), (0.3, 0.35, 1.0, 2650.0),
                                            (50.0, 51.0, 40.0, 1000.0)])
def test_soft_scattering_is_truncated_exponential(ph, k0, k1, step, rho):
    rock = ph.get_material_index("StandardRock")
    l0, l1 = _prop(ph, rock, L.T_TRANSPORT_PATH, k0), _prop(ph, rock, L.T_TRANSPORT_PATH, k1)
    mean = 0.5 * (0.5 * step * (rho / l0 + rho / l1))   # half the transport-path integral
    mu = O.soft_scattering(ph, [rock], [1.0], k0, k1, step, rho, 7, 100_000)
    assert 1e-12 < mean < 1
    assert mu.min() >= 0 and mu.max() <= 1
    cdf = lambda x: (1 - np.exp(-x / mean)) / (1 - math.exp(-1 / mean))
    assert ST.kstest(mu, cdf).pvalue > 1e-3


def test_soft_scattering_cap_and_floor(ph):
    rock = ph.get_material_index("StandardRock")
    big = O.soft_scattering(ph, [rock], [1.0], 0.05, 0.06, 500.0, 2650.0, 8, 50_000)
    cdf = lambda x: (1 - np.exp(-x)) / (1 - math.exp(-1))   # mean capped at 1
    assert ST.kstest(big, cdf).pvalue > 1e-3
    tiny = O.soft_scattering(ph, [rock], [1.0], 1e4, 1e4, 1e-12, 1.2, 9, 10)
    assert np.all(tiny == 0.0)


# ------------------------------------------------- Wentzel hard scattering
def _wentzel_cdf(A, mu0, f=0.0):
    """CDF on [mu0, 1] of (1 - f mu) / (A + mu)^2."""
    def F(x):
        x = np.asarray(x, float)
        lo = 1 / (A + mu0)
        # integral of (1 - f mu)/(A+mu)^2 = -(1 + f A)/(A+mu) - f log(A+mu)
        G = lambda y: -(1 + f * A) / (A + y) - f * np.log(A + y)
        return (G(x) - G(mu0)) / (G(1.0) - G(mu0)) if f else (lo - 1 / (A + x)) / (lo - 1 / (A + 1))
    return F


@pytest.mark.parametrize("k", [0.2, 3.0, 80.0])
def test_wentzel_centre_of_mass_draws(ph, k):
    rock = ph.get_material_index("StandardRock")
    data = ph.data.copy()
    data[:, L.T_SPIN_FACTOR, :] = 0.0                      # spin rejection off
    mu1, kin = O.scattering_angle(ph, [rock], [1.0], k, mode=1, seed=12, n=80_000, data=data)
    mu0, A, f = kin[0], kin[1], kin[2]
    assert f == 0.0 and mu0 == _prop(ph, rock, L.T_SCREENING, k) and A == mu0 / 4
    assert ST.kstest(mu1, _wentzel_cdf(A, mu0)).pvalue > 1e-3
    mu1s, kin = O.scattering_angle(ph, [rock], [1.0], k, mode=1, seed=13, n=80_000)
    assert kin[2] == _prop(ph, rock, L.T_SPIN_FACTOR, k) and kin[2] > 0
    assert ST.kstest(mu1s, _wentzel_cdf(A, mu0, kin[2])).pvalue > 1e-3


def _boost_mu(k, m, Mt, mu1):
    """Lab polar mu of the muon scattered elastically at centre-of-mass mu1 off
    a target of mass Mt at rest: explicit boost of its four-momentum."""
    E = k + m
    p = math.sqrt(k * (k + 2 * m))
    s = m * m + Mt * Mt + 2 * E * Mt
    beta = p / (E + Mt)
    gam = (E + Mt) / math.sqrt(s)
    ps = p * Mt / math.sqrt(s)
    Es = math.sqrt(ps * ps + m * m)
    out = []
    for x in mu1:
        ct = 1 - 2 * x
        st_ = 2 * math.sqrt(x * (1 - x))
        th = math.atan2(ps * st_, gam * (ps * ct + beta * Es))
        out.append(math.sin(th / 2) ** 2)
    return np.array(out)


@pytest.mark.parametrize("k", [0.05, 1.0, 30.0, 2e3])
def test_wentzel_lab_transform(ph, k):
    """The frame change is tan(theta_lab) = sin(theta*) / (g (cos(theta*) +
    tau)), evaluated stably as mu = sin^2(theta_lab / 2), with the first-order
    form mu1 / (g (1 + tau))^2 below mu1 = 1e-6.  (g, tau) are the
    reference's expressions of src/Straggling.jl:461-474, restated here from
    the same kinematic inputs; they are not the Lorentz factor and velocity
    ratio of the centre-of-mass frame, so these lab angles differ from an
    explicit boost (checked below, so the difference stays a documented
    property of the reference rather than a silent change)."""
    rock = ph.get_material_index("StandardRock")
    mu1 = np.concatenate([np.logspace(-9, -6.05, 6), np.logspace(-5.9, 0, 30)[:-1]])
    lab, kin = O.scattering_angle(ph, [rock], [1.0], k, mode=2, mu1=mu1)
    m = MUON
    Mt = NEUTRON / ph.scalars[rock, 2]          # target mass: A/Z neutron masses
    s = m * m + Mt * Mt + 2 * (k + m) * Mt
    k0 = (s - (m + Mt) ** 2) / (2 * math.sqrt(s))
    pcm, plab = math.sqrt(k0 * (k0 + 2 * m)), math.sqrt(k * (k + 2 * m))
    g = ((k + m) * (k0 + m) + plab * pcm) / (m * math.sqrt(s))
    tau = ((k + m) * pcm - (k0 + m) * plab) / (plab * pcm)
    assert kin[3] == pytest.approx(g, rel=1e-13) and kin[4] == pytest.approx(tau, rel=1e-11)
    big = mu1 > 1e-6
    ct = 1 - 2 * mu1[big]
    stt = 2 * np.sqrt(mu1[big] * (1 - mu1[big]))
    th = np.arctan2(stt, g * (ct + tau))
    # (1 - cos)/2 in the reference loses ~1e-16 absolute for small lab angles
    np.testing.assert_allclose(lab[big], np.sin(th / 2) ** 2, rtol=1e-9, atol=4e-16)
    np.testing.assert_allclose(lab[~big], mu1[~big] / (g * (1 + tau)) ** 2, rtol=1e-14)
    # continuity of the two forms at the switch
    lo = O.scattering_angle(ph, [rock], [1.0], k, mode=2, mu1=[1e-6 * (1 - 1e-12)])[0][0]
    hi = O.scattering_angle(ph, [rock], [1.0], k, mode=2, mu1=[1e-6 * (1 + 1e-12)])[0][0]
    assert hi == pytest.approx(lo, rel=1e-3)        # the forms agree to first order
    # the reference's frame change is not the Lorentz boost (kept as is)
    ref = _boost_mu(k, m, Mt, mu1[big])
    assert np.max(np.abs(lab[big] / ref - 1)) > 0.5


# -------------------------------------------------------- DEL polar angles
@pytest.mark.parametrize("ki", [1.0, 50.0, 3e3])
def test_polar_ionisation_is_two_body_kinematics(ki):
    m = MUON
    p = math.sqrt(ki * (ki + 2 * m))
    wmax = 2 * ME * p * p / (m * m + ME * (ME + 2 * (ki + m)))
    for nu in wmax * np.array([1e-4, 0.01, 0.3, 0.9]):
        kf = ki - nu
        pf = math.sqrt(kf * (kf + 2 * m))
        pe2 = nu * (nu + 2 * ME)                          # recoil electron
        c = (p * p + pf * pf - pe2) / (2 * p * pf)
        assert O.polar(4, m, ki, kf)[0] == pytest.approx(0.5 * (1 - c), rel=1e-9)


@pytest.mark.parametrize("ki,q", [(10.0, 2.0), (300.0, 40.0), (5e3, 3e3)])
def test_polar_photonuclear_momentum_transfer(ki, q):
    """The returned angle is the one whose Q^2 = -(p - p')^2 equals the
    log-uniform draw between the kinematic bounds."""
    m = MUON
    M = 0.5 * (PROTON + NEUTRON)
    E, E1 = ki + m, ki + m - q
    Q2min = m * m * (q * q - 0.5 * m * m) / (E * (E - q))
    Q2max = 2 * M * (q - PION) - PION * PION
    p = math.sqrt(ki * (ki + 2 * m))
    p1 = math.sqrt(E1 * E1 - m * m)
    fwd = 2 * (E * E1 - p * p1 - m * m)         # transfer of forward scattering
    for u in (0.0, 0.3, 0.5, 0.97):
        mu, used = O.polar(3, m, ki, ki - q, script=[u])
        assert used == 1
        want = Q2min * (Q2max / Q2min) ** u
        if want <= fwd:
            # below the forward transfer the angle clamps to 0 (the reference's
            # Q2min is a little below the exact forward value)
            assert mu == 0.0
            continue
        Q2 = 2 * (E * E1 - p * p1 * (1 - 2 * mu) - m * m)
        assert Q2 == pytest.approx(want, rel=1e-6, abs=1e-9 * Q2max)


@pytest.mark.parametrize("ki,kf,Z", [(20.0, 12.0, 11.0), (500.0, 100.0, 8.0), (3.0, 2.9, 26.0)])
def test_polar_brems_rejection_reproduces_its_density(ki, kf, Z):
    """Proposal mu = r mu0 / (mu0 + 1 - r), of density ~ (mu0 + mu)^-2 on
    [0, 1]; acceptance [c1/(mu0+mu)^2 - mu c2/(mu0+mu)^4] (x(mu)/x0) /
    (c1/(mu0+mu)^2).  Accepted samples therefore follow
    (c1/(mu0+mu)^2 - mu c2/(mu0+mu)^4) x(mu), normalised, with x the nuclear
    form-factor term of src/Straggling.jl:760-820 (the target density is the
    reference's; the test checks the sampling machinery reproduces it)."""
    m = MUON
    mu = O.polar(1, m, ki, kf, Z=Z, seed=21, n=60_000)
    assert np.all((mu >= 0) & (mu <= 1))
    E, nu = ki + m, ki - kf
    y = nu / E
    mu0 = (0.5 * m * nu / (E * (E - nu))) ** 2
    c1 = (2 * (1 - y) + y * y) * mu0
    c2 = 4 * (1 - y) * mu0 * mu0
    R = {8: 2.548, 11: 2.993, 26: 3.776}[int(Z)] * 1e-15   # nuclear_radius table entries
    muni = (m * R / 1.973269804e-16) ** 2 / 6

    def x_of(t):
        l2 = (mu0 + t) ** 2 / mu0 ** 2
        qq, r = muni * l2, mu0 * l2
        return (1 + 2 * qq * r) * np.log((1 + qq * r) / (r + qq * r)) - \
            (1 - r) * (1 + 2 * qq) / (1 + qq)

    dens = lambda t: (c1 / (mu0 + t) ** 2 - t * c2 / (mu0 + t) ** 4) * max(x_of(t), 0.0) \
        * (1.0 if muni * mu0 * ((mu0 + t) / mu0) ** 4 <= 1e5 else 0.0)
    # CDF on a log grid (the density spans many decades above mu0)
    grid = np.concatenate([[0.0], np.logspace(math.log10(mu0) - 6, 0, 3000)])
    pieces = [SI.quad(dens, a, b, limit=200)[0] for a, b in zip(grid[:-1], grid[1:])]
    cdfv = np.concatenate([[0.0], np.cumsum(pieces)])
    cdfv /= cdfv[-1]
    F = lambda v: np.interp(v, grid, cdfv)
    assert ST.kstest(mu, F).pvalue > 1e-3

"""CPU pins of the oracle's restatements of sample_mixture_material
(src/Physics.jl:1331-1365) and the legacy sample_del_event
(src/Straggling.jl:217-317), checked against what the mathematics fixes:
the draw count of every branch (scripted uniforms), the inverse CDF of the
1/nu^2 transfer law in closed form (nu(u = 1/2) = 2c / (1 + c)), a KS test of
seeded transfers against F(nu) = (1/c - 1/nu) / (1/c - 1), the event
probability 1 - exp(-sigma dX) of a step with a flat cross-section, process
and material frequencies against their weights, and the fraction-only
fallback when every cross-section is zero."""
import math

import numpy as np
import pytest

import diffpumas_hip as dp
from diffpumas_hip import _lib as L
import _oracle as O


@pytest.fixture(scope="module")
def ph3():
    return dp.create_physics(("StandardRock", "Water", "Air"))


def _xs(ph, m, e):
    return float(O.property_(ph, m, L.T_CROSS_SECTION, [e])[0][0])


def _binom_ok(count, n, p, z=5.0):
    return abs(count - n * p) <= z * math.sqrt(n * p * (1 - p)) + 1


def test_mixture_single_material_draws_nothing(ph3):
    m, used = O.sample_mixture_material(ph3, [2], [1.0], [50.0], script=[0.3])
    assert m == 2 and used == 0


def test_mixture_frequencies(ph3):
    mats, fr = [0, 1, 2], [0.6, 0.3, 0.1]
    e = 100.0
    w = np.array([f * _xs(ph3, m, e) for m, f in zip(mats, fr)])
    p = w / w.sum()
    n = 200_000
    out = O.sample_mixture_material(ph3, mats, [6.0, 3.0, 1.0], np.full(n, e), seed=11)
    counts = np.bincount(out, minlength=3)
    print("mixture frequencies", counts / n, "expected", p)
    assert all(_binom_ok(counts[j], n, p[j]) for j in range(3))
    # one draw per sample, on the cumulative weights, "<=" inclusive
    m, used = O.sample_mixture_material(ph3, mats, fr, [e], script=[p[0] * 0.999])
    assert m == 0 and used == 1
    m, _ = O.sample_mixture_material(ph3, mats, fr, [e], script=[p[0] + 1e-9])
    assert m == 1
    m, _ = O.sample_mixture_material(ph3, mats, fr, [e], script=[0.999999])
    assert m == 2


def test_mixture_fallback_on_fractions(ph3):
    zero = ph3.data.copy()
    zero[:, L.T_CROSS_SECTION] = 0.0

    class Z:
        n_materials, n_knots, scalars = ph3.n_materials, ph3.n_knots, ph3.scalars
        data = zero
    n = 100_000
    out = O.sample_mixture_material(Z, [0, 1, 2], [0.5, 0.3, 0.2], np.full(n, 10.0), seed=3)
    counts = np.bincount(out, minlength=3)
    assert all(_binom_ok(counts[j], n, q) for j, q in enumerate([0.5, 0.3, 0.2]))
    assert O.sample_mixture_material(Z, [0, 1, 2], [0.5, 0.3, 0.2], [10.0], script=[0.5])[0] == 0
    assert O.sample_mixture_material(Z, [0, 1, 2], [0.5, 0.3, 0.2], [10.0], script=[0.8])[0] == 1


def _step(ph, m, ki, gram):
    """A backward step of grammage gram ending at ki: (kf, Xtot)."""
    X0 = float(O.property_(ph, m, L.T_MIXED_RANGE, [ki])[0][0])
    kf = float(O.kinetic(ph, m, [X0 + gram], mixed=True)[0][0])
    return kf, X0 + gram


def test_del_branches_and_transfer_law(ph3):
    m, ki, c = 0, 50.0, 0.05
    kf, Xtot = _step(ph3, m, ki, 2000.0)
    xs = max(_xs(ph3, m, ki), _xs(ph3, m, kf))
    dX = 2000.0
    args = (ph3, m, ki, kf, dX, Xtot, 0.0)
    # X >= dX: one draw, no event
    row, used = O.sample_del_event(*args, cutoff=c, backward=True, script=[math.exp(-2 * xs * dX)])
    assert used == 1 and row[0] == 0
    row, used = O.sample_del_event(*args, cutoff=c, backward=True, script=[1e-13])
    assert used == 1 and row[0] == 0
    # accepted event, transfer drawn at u = 1/2, process draw 0 (bremsstrahlung)
    z = math.exp(-0.01 * xs * dX)
    row, used = O.sample_del_event(*args, cutoff=c, backward=True, script=[z, 0.0, 0.5, 0.0])
    nu = 2 * c / (1 + c)
    assert used == 4 and row[0] == 1 and row[3] == 1
    assert row[1] == pytest.approx(0.01 * dX, rel=1e-12)
    assert row[2] == pytest.approx(ki / (1 - nu), rel=1e-12)      # ratio 0: event at ki
    # forward mode loses the same fraction
    row, _ = O.sample_del_event(*args, cutoff=c, backward=False, script=[z, 0.0, 0.5, 0.0])
    assert row[2] == pytest.approx(ki * (1 - nu), rel=1e-12)
    # a process draw near 1 lands on photonuclear, the last process
    row, _ = O.sample_del_event(*args, cutoff=c, backward=True, script=[z, 0.0, 0.5, 1 - 1e-12])
    assert row[3] == 3
    # rejection on the local cross-section (the end energy given with a larger
    # one, 10 GeV): an acceptance draw above sigma(ki)/sigma(10) rejects after
    # two draws, one below it goes on
    k10 = 10.0
    r = _xs(ph3, m, ki) / _xs(ph3, m, k10)
    assert r < 0.5
    z10 = math.exp(-0.01 * _xs(ph3, m, k10) * dX)
    a10 = (ph3, m, ki, k10, dX, Xtot, 0.0)
    row, used = O.sample_del_event(*a10, cutoff=c, backward=True, script=[z10, 0.5 * (1 + r)])
    assert used == 2 and row[0] == 0
    row, used = O.sample_del_event(*a10, cutoff=c, backward=True, script=[z10, 0.5 * r, 0.5, 0.0])
    assert used == 4 and row[0] == 1


def test_del_statistics(ph3):
    m, ki, c = 0, 200.0, 0.05
    n = 200_000
    # flat cross-section over the step (kf = ki): P(event) = 1 - exp(-sigma dX)
    xs = _xs(ph3, m, ki)
    dX = 0.3 / xs
    X0 = float(O.property_(ph3, m, L.T_MIXED_RANGE, [ki])[0][0])
    rows = O.sample_del_event(ph3, m, np.full(n, ki), ki, dX, X0, 0.0, cutoff=c, backward=True,
                              seed=5)
    occ = rows[:, 0] == 1
    p = 1 - math.exp(-0.3)
    print(f"DEL frequency {occ.mean():.5f} expected {p:.5f}")
    assert _binom_ok(occ.sum(), n, p)
    # transfers: nu = 1 - ki / k_del against F(nu) = (1/c - 1/nu) / (1/c - 1)
    nu = np.sort(1.0 - ki / rows[occ, 2])
    F = (1 / c - 1 / nu) / (1 / c - 1)
    ks = np.max(np.abs(F - np.arange(1, len(nu) + 1) / len(nu)))
    print(f"transfer KS distance {ks:.4f} over {len(nu)} events")
    assert ks < 1.63 / math.sqrt(len(nu))           # 1 % level
    assert nu.min() >= c
    # processes in proportion to brems : pair : photonuclear at ki
    w = np.array([float(O.property_(ph3, m, a, [ki])[0][0])
                  for a in (L.T_BREMS, L.T_PAIR, L.T_PHOTO)])
    pr = w / w.sum()
    counts = np.bincount(rows[occ, 3].astype(int), minlength=4)[1:]
    assert all(_binom_ok(counts[j], occ.sum(), pr[j]) for j in range(3))

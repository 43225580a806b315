"""Closed-form pins of the oracle's mixed kernel (transport_mixed_step,
reverse_del_event) and EHS thinning, restated from the reference's own tests
with scripted draws:
  test/test_mixed_kernel.jl:6-27   limiting cases (no hard events, splitting,
                                   energy limit, DomainError)
  test/test_mixed_kernel.jl:33-64  hard-event weight closed form, null closure,
                                   optical survival exp(-1)
  test/test_mixed_kernel.jl:66-83  several hard events in one segment
  test/test_mixed_kernel.jl:133-149 sub-ULP range increments
  test/test_ehs_thinning.jl:14-63  EHS acceptance 2/3 (400 of 600)
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip import _lib as L
from diffpumas_hip.physics import PhysicsTables

AVOGADRO = 6.02214076e23
NA_ELEMENT = (11.0, 22.0, 149e-9, 1.0)


def test_limiting_cases():
    c = O.Kernel([1.0, 1e5], [2.0, 2.0], [0.0, 0.0], [0.0, (1e5 - 1) / 2], [0.0, 0.0],
                 [0.0, 0.0])
    s = dict(E=10.0)
    rc, a, _ = O.mixed_step(c, s, 1000.0, 0.01, seed=1)
    rc2, b, _ = O.mixed_step(c, s, 1000.0, 0.01, seed=2)
    assert rc == rc2 == 0
    assert a["E"] == pytest.approx(30.0, rel=1e-14)
    assert a["E"] == b["E"]
    assert a["L"] == pytest.approx(0.01, rel=1e-14)
    assert a["X"] == pytest.approx(10.0, rel=1e-14)
    assert a["w"] == 1.0
    _, first, _ = O.mixed_step(c, s, 1000.0, 0.003, seed=3)
    _, split, _ = O.mixed_step(c, first, 1000.0, 0.007, seed=4)
    assert split["E"] == pytest.approx(a["E"], rel=1e-14)
    assert split["L"] == pytest.approx(a["L"], rel=1e-14)
    assert split["X"] == pytest.approx(a["X"], rel=1e-14)
    _, lim, _ = O.mixed_step(c, s, 1000.0, 0.01, energy_limit=20.0, seed=1)
    assert lim["E"] == pytest.approx(20.0, rel=1e-14)
    assert lim["L"] == pytest.approx(0.005, rel=1e-12)
    rc, _, _ = O.mixed_step(c, s, 1000.0, 1000.0, seed=1)
    assert rc == -2   # DomainError: energy tables exhausted before the segment ends


def _kernel_na(onset=0.001):
    return O.Kernel([0.001, 1e9], [2.0, 2.0], [0.2, 0.2], [0.0, 5e8], [0.0, 1e8], [0.0, 0.0],
                    elements=[NA_ELEMENT], onset=onset)


def test_hard_event_weight_closed_form():
    c = _kernel_na()
    anc, w, used = O.reverse_del(c, 1000.0, [0.5])
    assert used == 1
    nu = 1 / (1 / 0.05 - 0.5 * (1 / 0.05 - 1))
    assert anc == pytest.approx(1000 / (1 - nu), rel=1e-14)
    q = anc - 1000
    Z, A, I, _ = NA_ELEMENT
    lib = L.load()
    raw = sum(lib.dp_dcs(p, Z, A, I, O.MUON_MASS, anc, q) for p in range(4))
    pdf = 1 / ((1 / 0.05 - 1) * nu ** 2)
    assert w == pytest.approx(anc * raw * AVOGADRO / (22 * 0.001) / (pdf * (1 - nu) * 0.2),
                              rel=1e-12)
    assert w > 0


def test_null_closure_below_onset():
    low = _kernel_na(onset=100.0)
    k, w, used = O.reverse_del(low, 0.001, [0.999999])
    assert (k, w) == (0.001, 1.0)


def test_optical_survival():
    """No DCS elements: the first proposed hard event absorbs; survivors
    measure exp(-sigma * grammage / S) = exp(-1)."""
    c = O.Kernel([1.0, 1e5], [2.0, 2.0], [0.1, 0.1], [0.0, (1e5 - 1) / 2],
                 [0.0, (1e5 - 1) * 0.05], [0.0, 0.0], onset=1.0)
    survivors = 0
    for seed in range(6000):
        rc, f, _ = O.mixed_step(c, dict(E=10.0), 1000.0, 0.01, energy_limit=1e5, seed=seed)
        assert rc == 0
        survivors += f["w"] > 0
    assert abs(survivors / 6000 - math.exp(-1)) < 0.025


def test_multiple_hard_events_consume_segment():
    c = O.Kernel([1.0, 1e9], [2.0, 2.0], [0.1, 0.1], [0.0, (1e9 - 1) / 2],
                 [0.0, (1e9 - 1) * 0.05], [0.0, 0.0], elements=[NA_ELEMENT], onset=1.0)
    s = dict(E=1000.0, pos=(1.0, 2.0, 3.0), dir=(0.0, 0.6, 0.8))
    rc, f, used = O.mixed_step(c, s, 1000.0, 0.01, energy_limit=1e9,
                               script=[0.1, 0.5, 0.2, 0.5, 10.0])
    assert rc == 0 and used == 5
    r = 1 - 1 / (1 / 0.05 - 0.5 * (1 / 0.05 - 1))
    assert f["E"] == pytest.approx(((1000 + 2) / r + 4) / r + 14, rel=1e-12)
    assert f["L"] == pytest.approx(0.01, rel=1e-12)
    assert f["X"] == pytest.approx(10.0, rel=1e-12)
    np.testing.assert_allclose(f["pos"], np.array(s["pos"]) - 0.01 * np.array(s["dir"]),
                               rtol=1e-12)
    assert math.isfinite(f["w"]) and f["w"] > 0


def test_sub_ulp_range_increment():
    e0 = 9.690953602732087e8
    c = O.Kernel([1.0, 1e9], [1e-6, 1e-6], [0.0, 0.0], [0.0, (1e9 - 1) * 1e6], [0.0, 0.0],
                 [0.0, 1e8])
    distance, density = 7.821108738426119e-8, 0.09128485154324607
    d0 = 29999.999899921786
    rng0 = 0.0 + (e0 - 1.0) / (1e9 - 1.0) * ((1e9 - 1) * 1e6 - 0.0)  # lookup(e0, energies, range)
    assert rng0 + distance * density == rng0
    rc, f, _ = O.mixed_step(c, dict(E=e0, L=d0), density, distance, energy_limit=1e9, seed=42)
    assert rc == 0
    assert abs((f["L"] - d0) - distance) <= 4 * np.spacing(d0)
    assert f["X"] == pytest.approx(distance * density, rel=1e-12)
    assert f["E"] >= e0
    assert f["t"] > 0
    assert f["w"] > 0


@pytest.fixture(scope="module")
def ehs_tables():
    """Two synthetic materials on K = 1, 2, 4, 8 GeV: unit stopping power,
    range(K) = K and lambda_h(K) = i (1 + K), i = 1, 2."""
    ref = dp.create_physics_csda(("StandardRock", "Air"))
    e = np.array([1.0, 2.0, 4.0, 8.0])
    data = np.zeros((2, L.T_NARRAYS, 4))
    for i in range(2):
        data[i, L.T_ENERGY] = e
        data[i, L.T_LOG_ENERGY] = np.log(e)
        data[i, L.T_CSDA_STOP] = data[i, L.T_MIXED_STOP] = 1.0
        data[i, L.T_CSDA_RANGE] = data[i, L.T_MIXED_RANGE] = e
        data[i, L.T_ELASTIC_PATH] = (i + 1) * (1 + e) * e * (e + 2 * O.MUON_MASS)
        data[i, L.T_TRANSPORT_PATH] = 1.0
    return PhysicsTables(("A", "B"), data, ref.scalars[:2].copy())


@pytest.mark.parametrize("mats,fracs", [([0], [1.0]), ([0, 1], [0.3, 0.7])])
def test_ehs_thinning(ehs_tables, mats, fracs):
    ph = ehs_tables
    ki, kf = 1.0, 4.0
    r0 = O.mix_range(ph, mats, fracs, ki)
    cand = abs(O.mix_range(ph, mats, fracs, 2.0) - r0)
    gram = abs(O.mix_range(ph, mats, fracs, kf) - r0)
    lref = O.ehs_path(ph, mats, fracs, 1.0)
    first = math.exp(-cand / lref)

    def sample(u):
        return O.ehs_sample(ph, mats, fracs, ki, kf, gram, r0, 1.0, [first, u])

    acc, x, k, _ = sample(0.6)
    assert acc
    assert x == pytest.approx(cand, rel=1e-12)
    assert k == pytest.approx(2.0, rel=1e-12)
    assert not sample(0.8)[0]
    assert sample(0.0)[0]
    assert not sample(1.0)[0]
    assert sum(sample((j - 0.5) / 600)[0] for j in range(1, 601)) == 400
    acc, _, _, used = O.ehs_sample(ph, mats, fracs, ki, kf, gram, r0, 1.0,
                                   [math.exp(-(gram + 0.01) / lref)])
    assert not acc and used == 1
    exact = -math.log(0.5) * lref
    acc, _, _, used = O.ehs_sample(ph, mats, fracs, ki, kf, exact, r0, 1.0, [0.5])
    assert not acc and used == 1

"""Statistical pin of the oracle's stochastic flat backward MC against the
reference's recorded runs (CPU only).

docs/revision_evidence/corrected_transport/full102_comparison.csv holds, for
102 (thickness, zenith) points of a flat two-layer geometry (StandardRock at
2650 kg/m3, primary altitude 1000 m, E 1e-3..1e9 GeV, straggling and
scattering on, mode switch 100 GeV; driver dev/flat_comparison_julia.jl), the
Julia flux and its sigma at 10^4 histories, and the official C PUMAS value.
"""
import csv
import os

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def phys3():
    return dp.create_physics(("StandardRock", "Water", "Air"))


def test_full102_statistical_agreement(phys3):
    ph = phys3
    rock, air = ph.get_material_index("StandardRock"), ph.get_material_index("Air")
    rows = list(csv.DictReader(open(os.path.join(GOLD, "full102_comparison.csv"))))
    assert len(rows) == 102
    n = 4000
    zs = []
    for r in rows:
        th, zen = float(r["thickness_m"]), float(r["zenith_deg"])
        f, _, ev = O.mc_flat(ph, [rock], [1.0], 2650.0, th, air, 90.0 - zen, 1e-3, 1e9, n, 99,
                             palt=1000.0, per_event=True)
        s = ev.std() / np.sqrt(n)
        zs.append((f - float(r["julia_flux"])) / np.hypot(s, float(r["julia_sigma"])))
    z = np.abs(np.array(zs))
    # 3 sigma combined at every point; a Gaussian distribution of z overall
    assert z.max() < 3.5, z.max()
    assert np.median(z) < 0.9
    assert (z > 2).sum() <= 10


def test_mixture_of_one_equals_material(phys3):
    """A one-component runtime mixture follows the single-material path."""
    ph = phys3
    rock, air = ph.get_material_index("StandardRock"), ph.get_material_index("Air")
    a = O.mc_flat(ph, [rock], [1.0], 2650.0, 100.0, air, 45.0, 1.0, 1e3, 500, 3)
    b = O.mc_flat(ph, [rock, rock], [0.5, 0.5], 2650.0, 100.0, air, 45.0, 1.0, 1e3, 500, 3)
    assert b[0] == pytest.approx(a[0], rel=0.05)


def test_straggled_mixture_runs(phys3):
    ph = phys3
    ids = [ph.get_material_index(n) for n in ("StandardRock", "Water", "Air")]
    f, s = O.mc_flat(ph, ids, [0.85, 0.14, 0.01], 2400.0, 300.0, ids[2], 30.0, 1.0, 1e4, 2000, 1)
    assert f > 0 and 0 < s < f


# ------------------------------------------- porous top layer (TwoLayerGeometry)
def test_porous_layer_over_whole_slab_is_a_slab_of_that_material(phys3):
    """TwoLayerGeometry (src/Geometry.jl:58-78) with a porous layer as thick as
    the rock: every step is in the porous medium with the porous density and
    the same boundaries, so the flux equals that of a plain slab of the porous
    material at the porous density, event by event (CSDA and straggled)."""
    ph = phys3
    rock, water, air = (ph.get_material_index(n) for n in ("StandardRock", "Water", "Air"))
    c = np.cos(np.radians(90 - 50.0))
    a = O.mc_flat_csda(ph, [120.0, 2650.0, 1000.0, 120.0, 3e4, 0, 0, 0], [rock, air, water],
                       1.0, 1e3, c, 300, 4, threads=1, gradient=False)
    b = O.mc_flat_csda(ph, [120.0, 1000.0, 0.0, 0.0, 3e4, 0, 0, 0], [water, air, -1],
                       1.0, 1e3, c, 300, 4, threads=1, gradient=False)
    assert a[0] == b[0] and a[1] == b[1]
    _, _, ea = O.mc_flat(ph, [rock], [1.0], 2650.0, 120.0, air, 50.0, 1.0, 1e3, 300, 4,
                         porous=water, porous_rho=1000.0, porous_thick=120.0, threads=1,
                         per_event=True)
    _, _, eb = O.mc_flat(ph, [water], [1.0], 1000.0, 120.0, air, 50.0, 1.0, 1e3, 300, 4,
                         threads=1, per_event=True)
    assert np.array_equal(ea, eb)


def test_porous_layer_sits_between_its_bounds(phys3):
    """A 40 m water layer on top of 160 m of rock (deterministic CSDA): less
    column than 200 m of rock, more than 160 m of rock alone, so the flux lies
    strictly between the two.  Using the rock density or skipping the layer
    would land on one of the bounds."""
    ph = phys3
    rock, water, air = (ph.get_material_index(n) for n in ("StandardRock", "Water", "Air"))
    c = np.cos(np.radians(90 - 60.0))
    run = lambda geo, mats: O.mc_flat_csda(ph, geo, mats, 1.0, 1e3, c, 400, 6, threads=1,
                                           gradient=False)[0]
    f_por = run([200.0, 2650.0, 1000.0, 40.0, 3e4, 0, 0, 0], [rock, air, water])
    f_200 = run([200.0, 2650.0, 0.0, 0.0, 3e4, 0, 0, 0], [rock, air, -1])
    f_160 = run([160.0, 2650.0, 0.0, 0.0, 3e4, 0, 0, 0], [rock, air, -1])
    assert f_200 < f_por < f_160

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

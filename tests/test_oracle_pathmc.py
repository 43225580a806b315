"""Pins of the oracle's fixed-path (tomography) MC, compute_directional_flux_mc
(src/Tomography.jl:979-1041):
  * segment conservation: a 100 m rock segment is consumed exactly
    (test/test_mixed_kernel.jl:104-110: distance 100 m, grammage 265 000 kg/m2);
  * CSDA mode of a vertical column of cells equals the flat two-layer CSDA
    walk over the same rock (range-based steps compose; only the decay
    quadrature differs between step partitions);
  * straggled mode agrees statistically with the flat straggled MC on the same
    column (scattering off), an independent code path of the reference;
  * sigma uses (E[w2] - E[w]^2)/(n - 1) and is 0 for one sample; invalid paths
    and cos(zenith) <= 0 give NaN.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp


@pytest.fixture(scope="module")
def ph3():
    return dp.create_physics(("StandardRock", "Water", "Air"))


def _column(n_cells, size, zenith=0.0, remaining=0.0, valid=1):
    c = math.cos(math.radians(zenith))
    lens = [size] * n_cells
    return dict(row_ptr=np.array([0, n_cells], np.int32), cells=np.arange(n_cells, dtype=np.int32),
                lengths=np.array(lens, float), remaining=np.array([remaining]),
                exit_z=np.array([(n_cells * size + remaining) * c]),
                cos_theta=np.array([c]), valid=np.array([valid], np.uint8))


def test_segment_conservation(ph3):
    rock = ph3.get_material_index("StandardRock")
    for seed in range(5):
        rc, st = O.mc_uniform_segment(ph3, rock, 2650.0, 100.0, dict(E=99.0), seed=seed)
        assert rc == 0
        assert st["L"] == pytest.approx(100.0, abs=1e-8)
        assert st["X"] == pytest.approx(265000.0, abs=1e-5)
        assert st["E"] > 99.0 and st["w"] > 0


def test_csda_single_sample_sigma_zero(ph3):
    rock, air = ph3.get_material_index("StandardRock"), ph3.get_material_index("Air")
    paths = _column(1, 100.0, zenith=20.0)
    f, s = O.mc_paths(ph3, [([rock], [1.0])], [0], [2650.0], rock, air, 2650.0, (1000.0, 12000.0),
                      paths, (np.array([1000.0]), np.array([1.0]), np.array([1.0])), 42,
                      straggling=False)
    assert f[0] > 0 and math.isfinite(f[0]) and s[0] == 0.0


def test_csda_column_equals_flat_csda(ph3):
    rock, air = ph3.get_material_index("StandardRock"), ph3.get_material_index("Air")
    paths = _column(10, 10.0)
    for e in (5.0, 50.0, 500.0):
        f, _ = O.mc_paths(ph3, [([rock], [1.0])], np.zeros(10, np.int32), np.full(10, 2650.0),
                          rock, air, 2650.0, (0.0, 3e4), paths,
                          (np.array([e]), np.array([0.5]), np.array([1.0])), 1, straggling=False)
        flat = O.flat_csda(ph3, [100.0, 2650.0, 0.0, 0.0, 3e4, 0.0, 0.0, 0.0], [rock, air, -1], e,
                           1.0, 1.0)
        assert f[0] == pytest.approx(2 * 0.5 * flat[0], rel=1e-7)


def test_straggled_column_matches_flat_mc(ph3):
    """Fixed-path straggled MC over 20 cells of 10 m vs the flat straggled MC
    over a 200 m rock layer (both without scattering)."""
    rock, air = ph3.get_material_index("StandardRock"), ph3.get_material_index("Air")
    paths = _column(20, 10.0)
    n = 4000
    rng = np.random.default_rng(3)
    rk = math.log(1e4 / 1.0)
    e = np.exp(rk * rng.random(n))
    w = e * rk
    q = np.where(rng.random(n) > 0.5, 1.0, -1.0)
    f, s = O.mc_paths(ph3, [([rock], [1.0])], np.zeros(20, np.int32), np.full(20, 2650.0), rock,
                      air, 2650.0, (0.0, 3e4), paths, (e, w, q), 7, straggling=True)
    ff, sf = O.mc_flat(ph3, [rock], [1.0], 2650.0, 200.0, air, 90.0, 1.0, 1e4, n, 11,
                       straggling=True, scattering=False, threshold=1e9)
    assert abs(f[0] - ff) <= 3.0 * math.hypot(s[0], sf)


def test_water_cells_lower_the_attenuation(ph3):
    """Lighter cells (w > 0) transmit more flux than dry rock (CSDA mode)."""
    rock, water, air = (ph3.get_material_index(n) for n in ("StandardRock", "Water", "Air"))
    paths = _column(10, 10.0)
    samples = (np.array([20.0, 60.0]), np.array([1.0, 1.0]), np.array([1.0, -1.0]))
    w = 0.3
    rho = 1.0 / ((1 - w) / 2650.0 + w / 1000.0)
    dry, _ = O.mc_paths(ph3, [([rock], [1.0])], np.zeros(10, np.int32), np.full(10, 2650.0), rock,
                        air, 2650.0, (0.0, 3e4), paths, samples, 1, straggling=False)
    wet, _ = O.mc_paths(ph3, [([rock, water], [1 - w, w])], np.zeros(10, np.int32),
                        np.full(10, rho), rock, air, 2650.0, (0.0, 3e4), paths, samples, 1,
                        straggling=False)
    assert wet[0] > dry[0] > 0


def test_invalid_and_horizontal_paths(ph3):
    rock, air = ph3.get_material_index("StandardRock"), ph3.get_material_index("Air")
    a = _column(2, 10.0, valid=0)
    b = _column(2, 10.0, zenith=95.0)
    for p in (a, b):
        f, s = O.mc_paths(ph3, [([rock], [1.0])], [0, 0], [2650.0, 2650.0], rock, air, 2650.0,
                          (0.0, 3e4), p, (np.array([10.0]), np.array([1.0]), np.array([1.0])), 1)
        assert math.isnan(f[0]) and math.isnan(s[0])

"""The empirical table path (create_material_table_empirical,
src/Physics.jl:280-408; electronic stopping power with the Fano density
effect, src/Materials.jl:345-528) on the CPU: the library's host builder
against values fixed outside it.

* The electronic stopping power and the density effect of the MDF standard
  rock (the fictitious Z = 11, A = 22 element, whose shells are C + 3 O + Ca)
  equal the Ionization and delta columns of the PUMAS v1.1 dE/dx table of
  the package data within the table's printed rounding (4-digit mantissa,
  4 decimals): PUMAS computed those columns with the same model.  Water and
  air agree to 5e-4 (their MDF Z/A and I differ slightly from the ones PUMAS
  tabulated with).
* delta(gamma) -> 2 ln(gamma omega_p / I) - 1 at high energy (the
  Sternheimer limit, fixed by the shell normalisation), 0 below the
  threshold where the bisection has no root, and the reference's absolute
  stopping rule at very high gamma.
* The empirical tables reproduce the PUMAS dE/dx files' stopping powers per
  process (2e-3) on the files' own knots.
* The reference's only deterministic logged flux: run_backward_mc on
  create_physics(MUON; n_energies=50, K_min=0.01, K_max=1e6) with no dE/dx
  files (empirical tables), rock_thickness 0, elevation 90, 10 GeV, one
  sample, straggling and scattering off (test/test_geometry_ad.jl:7,81-89).
  The current reference logs 0.4610008899907813
  (docs/revision_evidence/integrated_kernel_tests/final_precision_fix/
  julia-1.11-pkg-tests.txt and 12 more logs); that event drew a positive
  charge, so the flux is 2 flux(+).  Older logs
  (docs/revision_evidence/tests-julia-1.11.txt:270) hold
  0.4612840622755491, computed before radiative DCS moments dropped an
  extra q/(K+m) factor (RADIATIVE_DCS_TABLE_VERSION 2,
  src/Materials.jl:15-17); rebuilding the radiative columns with that
  factor (mode-2 moment / (K + m)) reproduces it as well.
"""
import ctypes as C
import math
import warnings

import numpy as np
import pytest

import diffpumas_hip as dp
from diffpumas_hip import _lib as L
from diffpumas_hip import physics as P
import _oracle as O

GOLDEN_FLUX = 0.4610008899907813      # current reference logs
GOLDEN_FLUX_V1 = 0.4612840622755491   # logs before the radiative-moment fix


def _recipes():
    _, rec, _ = P.load_mdf(P.default_data_dir() + "/materials.xml")
    return rec


def _pumas(name):
    path = P.default_data_dir() + "/pumas-materials/dedx/muon/" + P.dedx_file_name(name)
    rows = [l.split() for l in open(path) if l.strip() and l.split()[0][0].isdigit()]
    return path, rows


@pytest.mark.parametrize("name,exact", [("StandardRock", True), ("Water", False), ("Air", False)])
def test_electronic_loss_and_density_effect_match_pumas(name, exact):
    rec = _recipes()[name]
    path, rows = _pumas(name)
    T = np.array([float(r[0]) for r in rows]) * 1e-3               # GeV
    ion_txt = [r[2] for r in rows]
    ion = np.array([float(t) for t in ion_txt]) * 1e-4              # GeV m2/kg
    delta = np.array([float(r[9]) for r in rows])
    mine = np.array([dp.electronic_stopping_power(rec, k) for k in T])
    dmine = np.array([dp.electronic_density_effect(rec, 1.0 + k / P.MUON_MASS) for k in T])
    half = np.array([0.5 * 10.0 ** (int(t.split("E")[1]) - 3) for t in ion_txt]) * 1e-4
    ratio = np.abs(mine - ion) / half
    print(f"{name}: |ionization - PUMAS| max {ratio.max():.3f} half-units of the last digit; "
          f"|delta - PUMAS| max {np.abs(dmine - delta).max():.2e}")
    if exact:
        assert ratio.max() <= 1.0
        assert np.abs(dmine - delta).max() <= 5e-5 + 1e-12
    else:
        assert np.abs(mine / ion - 1).max() <= 5e-4
        assert np.abs(dmine - delta).max() <= 2e-3


def test_density_effect_limits():
    """Sternheimer's high-energy limit 2 ln(gamma omega_p / I) - 1 (the
    shells' strength-weighted mean log energy is log(I / omega_p)), reached
    as 1/gamma^2; zero below the root's threshold.  The bisection stops once
    |sum f/(E^2 + l2) - 1/gamma^2| <= eps(Float64) in absolute terms, as in
    the reference: for 1/gamma^2 < eps/2 (gamma >= 1e8) that holds at the
    first midpoint l2 = gamma^2/2, so delta = limit + 1/2 - ln 2 exactly."""
    rec = _recipes()["StandardRock"]
    wp = 28.816 * math.sqrt(rec.ZoA * rec.density * 1e-3) * 1e-9      # GeV
    lim = lambda g: 2.0 * math.log(g * wp / rec.I) - 1.0
    for g, tol in ((1e4, 5e-5), (1e5, 5e-7), (1e6, 1e-8)):
        assert abs(dp.electronic_density_effect(rec, g) - lim(g)) <= tol
    for g in (1e8, 1e9):
        first = lim(g) + 0.5 - math.log(2.0)
        assert dp.electronic_density_effect(rec, g) == pytest.approx(first, rel=1e-10)
    assert dp.electronic_density_effect(rec, 1.01) == 0.0


def test_empirical_tables_match_pumas_files():
    lib = L.load()
    st = L.PhysicsSettings(0.05, 0.05, P.MUON_MASS)
    rec = _recipes()
    for name in ("StandardRock", "Air"):
        d = P.load_dedx_file(_pumas(name)[0])
        e = np.ascontiguousarray(d.kinetic_energy)
        r, keep = P._recipe(rec[name])
        out = np.zeros((L.T_NARRAYS, len(e)))
        L.check(lib.dp_build_material_table_empirical(C.byref(r), L.ptr(e), len(e), C.byref(st),
                                                      L.ptr(out)))
        for col, ref in ((L.T_CSDA_STOP, d.stopping_power), (L.T_IONIZATION, d.ionization),
                         (L.T_BREMS, d.bremsstrahlung), (L.T_PAIR, d.pair_production),
                         (L.T_PHOTO, d.photonuclear)):
            m = ref > 1e-9 * ref.max()
            rel = np.abs(out[col][m] / ref[m] - 1)
            assert rel.max() <= 2e-3, (name, col, rel.max())
        assert np.all(np.diff(out[L.T_CSDA_RANGE]) > 0)


def _empirical_physics():
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        ph = dp.create_physics(("StandardRock", "Air"), n_energies=50, K_min=0.01, K_max=1e6,
                               dedx_dir="/nonexistent-dedx-dir")
    assert sum("Using empirical formulas" in str(x.message) for x in w) == 2
    return ph


def _air_flux(ph, charge):
    rock, air = ph.get_material_index("StandardRock"), ph.get_material_index("Air")
    o = O.flat_csda(ph, [0.0, 2650.0, 0.0, 0.0, 3e4, 0.0, 0.0, 0.0], [rock, air, -1], 10.0, 1.0,
                    charge)
    return 2.0 * o[0]


def test_golden_air_only_flux():
    ph = _empirical_physics()
    e = ph.data[0, L.T_ENERGY]
    assert len(e) == 50 and e[0] == 0.01 and e[-1] == pytest.approx(1e6, rel=1e-15)
    f = _air_flux(ph, 1.0)
    print(f"air-only flux {f!r} (reference {GOLDEN_FLUX!r}); negative charge {_air_flux(ph, -1.0)!r}")
    assert f == pytest.approx(GOLDEN_FLUX, rel=1e-14)


def test_golden_air_only_flux_before_the_moment_fix():
    """The older logged value, from radiative stopping powers with the extra
    q/(K+m) factor: the mode-1 moment replaced by the mode-2 moment / (K+m)."""
    ph = _empirical_physics()
    lib = L.load()
    bi = P.builtin_materials()
    m = P.MUON_MASS
    for name in ("StandardRock", "Air"):
        i = ph.get_material_index(name)
        E = ph.data[i, L.T_ENERGY]
        S = ph.data[i, L.T_IONIZATION].copy()
        for k, K in enumerate(E):
            for p in range(3):
                for el, fr in zip(bi[name].elements, bi[name].fractions):
                    S[k] += fr * lib.dp_dcs_integral(p, el.Z, el.A, el.I, m, K, 1e-6, 1.0, 2,
                                                     180) / (K + m)
        R = np.zeros_like(S)
        R[0] = E[0] / S[0]
        for k in range(1, len(E)):
            R[k] = R[k - 1] + (E[k] - E[k - 1]) / ((S[k] + S[k - 1]) / 2)
        ph.data[i, L.T_CSDA_STOP] = S
        ph.data[i, L.T_CSDA_RANGE] = R
    assert _air_flux(ph, 1.0) == pytest.approx(GOLDEN_FLUX_V1, rel=1e-14)

"""GPU parity of the direct CSDA flux and its density tangent (dp_csda_direct,
compute_flux_csda_direct / compute_flux_gradient_csda, src/Geometry.jl:
1126-1262) against the CPU oracle: flux and derivative within 1e-12 relative
(the same FP64 operations), scalar and array calls, clamped far samples.
"""
import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ph():
    return dp.create_physics_csda(("StandardRock", "Air"))


def test_samples_match_oracle(ph, ctx):
    rng = np.random.default_rng(4)
    n = 5000
    rho = rng.uniform(1800.0, 3000.0, n)
    thick = rng.uniform(0.0, 800.0, n)
    elev = rng.uniform(5.0, 90.0, n)
    e = np.exp(rng.uniform(np.log(0.1), np.log(1e5), n))
    q = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    rock, air = ph.get_material_index("StandardRock"), ph.get_material_index("Air")
    print(f"direct CSDA: {n} samples", flush=True)
    f, d = dp.compute_flux_csda_direct(ph, rock, air, rho, thick, elev, e, q, gradient=True, ctx=ctx)
    fo, do = O.csda_direct(ph, rock, air, rho, thick, elev, e, q)
    np.testing.assert_allclose(f, fo, rtol=1e-12, atol=0)
    np.testing.assert_allclose(d, do, rtol=1e-10, atol=1e-14 * np.abs(do).max())
    assert (f > 0).mean() > 0.3 and np.all(f >= 0)


def test_scalar_calls_and_named_materials(ph, ctx):
    f = dp.compute_flux_differentiable_csda(ph, 2650.0, 120.0, 40.0, 30.0, 1.0, ctx=ctx)
    f2, g = dp.compute_flux_gradient_csda(ph, 2650.0, 120.0, 40.0, 30.0, 1.0, ctx=ctx)
    fo, do = O.csda_direct(ph, 0, 1, 2650.0, 120.0, 40.0, 30.0, 1.0)
    assert isinstance(f, float) and f == f2
    assert f == pytest.approx(fo[0], rel=1e-12)
    assert g == pytest.approx(do[0], rel=1e-10)


def test_far_samples_and_bad_material(ph, ctx):
    f, d = dp.compute_flux_csda_direct(ph, 0, 1, [2650.0, 2650.0], [5000.0, 1e6], [5.0, 5.0],
                                       [5e8, 10.0], [1.0, 1.0], gradient=True, ctx=ctx)
    fo, do = O.csda_direct(ph, 0, 1, 2650.0, np.array([5000.0, 1e6]), 5.0,
                           np.array([5e8, 10.0]), 1.0)
    np.testing.assert_allclose(f, fo, rtol=1e-12)
    # 1000 km of rock: the inverse range clamps at the last knot (property_kinetic_energy),
    # where the density tangent is 0
    assert d[1] == do[1] == 0.0 and f[1] > 0
    with pytest.raises(dp.ArgumentError):
        dp.compute_flux_csda_direct(ph, 0, 7, 2650.0, 10.0, 45.0, 10.0, 1.0, ctx=ctx)

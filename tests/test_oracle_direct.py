"""Pins of the oracle's direct CSDA flux (compute_flux_csda_direct and
compute_flux_gradient_csda, src/Geometry.jl:1126-1262).

The reference holds no numbers for it (examples/diff_flux.jl:430-445 only
calls it).  Pinned by:
  * without decay (c tau -> infinity) and with a 1 cm atmosphere, the one-shot
    range inversion equals the step-by-step CSDA walk of compute_flux_single
    (range is additive and the stopping-power ratio telescopes);
  * a thin slab: the mean-energy decay factor equals the walk's proper-time
    decay to second order;
  * the density tangent equals central differences;
  * a rock slab of zero thickness leaves only the atmosphere.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip.physics import PhysicsTables


@pytest.fixture(scope="module")
def ph():
    return dp.create_physics_csda(("StandardRock", "Air"))


def _no_decay(p):
    return PhysicsTables(p.names, p.data, p.scalars, p.mass, 1e300)


def _walk(p, thick, rho, elev, e, q, palt):
    rock, air = p.get_material_index("StandardRock"), p.get_material_index("Air")
    ct = math.cos(math.pi / 2 - elev * (math.pi / 180.0))
    return O.flat_csda(p, [thick, rho, 0.0, 0.0, palt, 0.0, 0.0, 0.0], [rock, air, -1], e, ct, q)


@pytest.mark.parametrize("thick,rho,elev", [(100.0, 2650.0, 45.0), (400.0, 2300.0, 70.0),
                                            (30.0, 2700.0, 20.0)])
@pytest.mark.parametrize("e", [1.0, 30.0, 2000.0])
def test_no_decay_equals_stepwise_walk(ph, thick, rho, elev, e):
    p = _no_decay(ph)
    rock, air = p.get_material_index("StandardRock"), p.get_material_index("Air")
    palt = thick + 0.01
    f, d = O.csda_direct(p, rock, air, rho, thick, elev, e, 1.0, primary_altitude=palt)
    w = _walk(p, thick, rho, elev, e, 1.0, palt)
    assert f[0] > 0
    assert f[0] == pytest.approx(w[0], rel=1e-7)
    assert d[0] == pytest.approx(w[1], rel=1e-6)


@pytest.mark.parametrize("thick", [0.02, 0.002])
def test_thin_slab_decay(ph, thick):
    """The decay factor (flux over its no-decay value) of rock plus air equals
    the walk's proper-time decay; dropping either layer's factor would change
    the exponent by a third or more."""
    rock, air = ph.get_material_index("StandardRock"), ph.get_material_index("Air")
    nd = _no_decay(ph)
    for e in (0.5, 5.0, 50.0):
        args = (2650.0, thick, 60.0, e, -1.0)
        f, _ = O.csda_direct(ph, rock, air, *args, primary_altitude=thick + 0.01)
        f0, _ = O.csda_direct(nd, rock, air, *args, primary_altitude=thick + 0.01)
        w = _walk(ph, thick, 2650.0, 60.0, e, -1.0, thick + 0.01)
        w0 = _walk(nd, thick, 2650.0, 60.0, e, -1.0, thick + 0.01)
        assert math.log(f[0] / f0[0]) == pytest.approx(math.log(w[0] / w0[0]), rel=2e-4)


@pytest.mark.parametrize("e", [2.0, 100.0])
def test_density_tangent_matches_central_difference(ph, e):
    rock, air = ph.get_material_index("StandardRock"), ph.get_material_index("Air")
    f, d = O.csda_direct(ph, rock, air, 2650.0, 150.0, 50.0, e, 1.0)
    h = 1e-3
    fp, _ = O.csda_direct(ph, rock, air, 2650.0 + h, 150.0, 50.0, e, 1.0)
    fm, _ = O.csda_direct(ph, rock, air, 2650.0 - h, 150.0, 50.0, e, 1.0)
    assert d[0] < 0
    assert d[0] == pytest.approx((fp[0] - fm[0]) / (2 * h), rel=1e-6)


def test_zero_thickness_is_atmosphere_only(ph):
    rock, air = ph.get_material_index("StandardRock"), ph.get_material_index("Air")
    f0, d0 = O.csda_direct(ph, rock, air, 2650.0, 0.0, 35.0, 10.0, 1.0)
    f1, _ = O.csda_direct(ph, rock, air, 1000.0, 0.0, 35.0, 10.0, 1.0)
    assert f0[0] == f1[0] and d0[0] == 0.0
    # an energy beyond the table's last knot leaves no flux
    big, _ = O.csda_direct(ph, rock, air, 2650.0, 1e4, 10.0, 5e8, 1.0)
    assert big[0] >= 0.0 and math.isfinite(big[0])

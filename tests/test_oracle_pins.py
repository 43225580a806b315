"""Pins of the CPU oracle against reference-held values and mathematics.

CPU only (no GPU): known-answer vectors, the reference's golden table
values, interpolation/derivative conventions, flux-model identities.
"""
import csv
import math
import os

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip import _lib as L

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# --------------------------------------------------------------- Philox
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32-10
    assert O.philox_block((0, 0, 0, 0), (0, 0)) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox_block((0xffffffff,) * 4, (0xffffffff,) * 2) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox_block((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
                          (0xa4093822, 0x299f31d0)) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_stream_layout():
    # event e, block b -> counter (e lo, e hi, b, 0); uniform = (w0<<32|w1)>>11 * 2^-53
    seed, ev = 0x1234_5678_9abc_def0, 77
    blk = O.philox_block((ev, 0, 0, 0), (seed & 0xffffffff, seed >> 32))
    u = O.philox_uniforms(seed, ev, 2)
    assert u[0] == ((blk[0] << 32 | blk[1]) >> 11) * 2.0 ** -53
    assert u[1] == ((blk[2] << 32 | blk[3]) >> 11) * 2.0 ** -53
    big = O.philox_uniforms(7, 3, 20000)
    assert 0.0 <= big.min() and big.max() < 1.0
    assert abs(big.mean() - 0.5) < 0.01


# ----------------------------------------------------------- dE/dx parser
@pytest.mark.parametrize("separators", [[], [""], ["", "# data follows", "-" * 80]])
def test_dedx_first_record(tmp_path, separators):
    # test/test_dedx_first_record.jl:4-26
    header = ["Incident particle is a Muon with M = 105.65839 MeV", "Index = 1: StandardRock",
              "<Z/A> = 0.5", "T [MeV] p [MeV/c] ion brems pair photo rad total range delta beta"]
    rows = ["1.0 14.5 20.0 0.1 0.2 0.3 0.6 20.6 0.04 0.0 0.136",
            "1.2 16.0 19.0 0.2 0.3 0.4 0.9 19.9 0.05 0.0 0.149"]
    f = tmp_path / "fixture.txt"
    f.write_text("\n".join(header + separators + rows) + "\n")
    d = dp.load_dedx_file(str(f))
    assert d.material_name == "StandardRock"
    assert d.particle_mass == pytest.approx(0.10565839)
    assert list(d.kinetic_energy) == [0.001, 0.0012]
    assert np.allclose(d.momentum, [0.0145, 0.016])
    assert np.allclose(d.ionization, np.array([20.0, 19.0]) * 1e-4)
    assert np.allclose(d.stopping_power, np.array([20.6, 19.9]) * 1e-4)
    assert np.allclose(d.csda_range, [0.4, 0.5])
    assert len(d.beta) == 2


def test_pinned_dedx_tables(physics):
    assert physics.n_knots == 193  # physics-metadata.toml nodes = 193
    e = physics.data[0, L.T_ENERGY]
    assert e[0] == 1e-3 and e[-1] == 1e9


# ------------------------------------------------- golden table values
def _golden():
    rows = {}
    with open(os.path.join(GOLD, "physics_julia_golden.csv")) as f:
        for r in csv.DictReader(f):
            rows[(r["material"], float(r["energy"]), r["property"])] = float(r["value"])
    return rows


@pytest.mark.parametrize("prop,arr", [("stopping_power_0", L.T_CSDA_STOP),
                                      ("range_0", L.T_CSDA_RANGE),
                                      ("proper_time_0", L.T_CSDA_TIME)])
def test_golden_csda_columns(physics, prop, arr):
    # docs/revision_evidence/corrected_transport/physics-julia.csv (Julia values)
    g = _golden()
    for name in ("StandardRock", "Air"):
        m = physics.get_material_index(name)
        for e in (10.0, 100.0, 1000.0, 1e4, 1e5):
            v, _ = O.property_(physics, m, arr, [e])
            assert v[0] == pytest.approx(g[(name, e, prop)], rel=1e-13), (name, e, prop)


# ---------------------------------------- interpolation conventions
def _toy_tables(xs, ys):
    K = len(xs)
    data = np.zeros((1, L.T_NARRAYS, K))
    data[0, L.T_ENERGY] = xs
    data[0, L.T_LOG_ENERGY] = np.log(xs)
    data[0, L.T_CSDA_STOP] = ys
    return dp.PhysicsTables(["toy"], data, np.zeros((1, 4)))


def test_interp_knot_conventions():
    # test/test_segment_tangents.jl:20-25
    ph = _toy_tables([1.0, 2.0, 4.0], [1.0, 3.0, 2.0])
    v, d = O.property_(ph, 0, L.T_CSDA_STOP, [1.0, 4.0, 2.0, 3.0, 0.5, 8.0])
    assert d[0] == 0.0 and d[1] == 0.0
    assert d[2] == pytest.approx(-1 / (2 * math.log(2.0)), rel=1e-15)
    assert list(v[[0, 1, 2, 4, 5]]) == [1.0, 2.0, 3.0, 1.0, 2.0]
    # log-linear midpoint: y(3) = 3 + log(3/2)/log(2) * (2 - 3)
    assert v[3] == pytest.approx(3 - math.log(1.5) / math.log(2.0), rel=1e-14)


def test_inverse_range_roundtrip(physics):
    m = physics.get_material_index("StandardRock")
    e = np.geomspace(2e-3, 5e8, 400)
    x, _ = O.property_(physics, m, L.T_CSDA_RANGE, e)
    k, dk = O.kinetic(physics, m, x)
    assert np.allclose(k, e, rtol=1e-12)
    # clamped outside the table: constant, zero derivative
    k2, dk2 = O.kinetic(physics, m, [0.0, 1e30])
    assert list(k2) == [1e-3, 1e9] and list(dk2) == [0.0, 0.0]
    # d E / d X = 1 / (d X / d E) inside an interval
    _, dx = O.property_(physics, m, L.T_CSDA_RANGE, e)
    assert np.allclose(dk * dx, 1.0, rtol=1e-9)


# -------------------------------------------------------- flux model
def test_flux_model_identities():
    m = dp.MUON_MASS
    for E in (1.0, 10.0, 1e3):
        # charge fractions sum to the charge-summed flux (src/GaisserFlux.jl:36-46)
        tot = O.flux_gccly(0.7, E, 0.0)
        assert O.flux_gccly(0.7, E, 1.0) + O.flux_gccly(0.7, E, -1.0) == pytest.approx(tot, rel=1e-15)
        assert O.flux_gccly(0.7, E, 1.0) / O.flux_gccly(0.7, E, -1.0) == pytest.approx(1.2766, rel=1e-14)
        # cos_theta_star(1) == 1 exactly, so at vertical GCCLY = factor * Gaisser(1)
        fac = (1 + 3.64 / (E + m)) ** -2.7
        assert O.flux_gccly(1.0, E, 0.0) == pytest.approx(fac * O.flux_gaisser(1.0, E, 0.0), rel=1e-14)
    # Gaisser power law: at E << 115 GeV * (1/1.1) the bracket tends to 1.054
    E = 1e-2
    assert O.flux_gaisser(1.0, E, 0.0) == pytest.approx(1.4e3 * (E + m) ** -2.7 *
                                                       (1 / (1 + 1.1 * (E + m) / 115) +
                                                        0.054 / (1 + 1.1 * (E + m) / 850)), rel=1e-15)
    # d flux / dE from the dual path against central differences
    for E in (3.0, 300.0):
        h = E * 1e-6
        fd = (O.flux_gccly(0.6, E + h, 1.0) - O.flux_gccly(0.6, E - h, 1.0)) / (2 * h)
        assert O.lib().orc_dflux_gccly_dk(0.6, E, 1.0) == pytest.approx(fd, rel=1e-8)


def test_long_double_central_difference_equals_forward_mode():
    """orc_cube_fd_ld (the extended-precision central difference the full-size
    adjoint test uses) against the oracle's forward-mode derivative of the
    same program: agreement to 1e-10, while the FP64 difference at the same h
    is only good to ~1e-8."""
    from workloads import cube_rays, energy_nodes, uniform_densities
    ph = dp.create_physics_csda(("StandardRock", "Air"))
    geo = dp.CubeGeometry(80.0, 12, 10, 8, ph.get_material_index("StandardRock"),
                          ph.get_material_index("Air"), width=120.0)
    sxy, dirs, cos_t = cube_rays(geo, 2, 8, seed=4)
    rp, cells, lens, ze = O.trace_cube(geo, sxy, dirs)
    rho = uniform_densities(geo.n_cells, 3)
    ne, nw, nq = energy_nodes(128, 1.0, 1e4)
    flux, grads = O.cube_csda(ph, geo, rp, cells, lens, cos_t, ze, rho, ne, nw, nq)
    for v in sorted(set(int(c) for c in cells[::7]))[:6]:
        if v == 0:
            continue
        G = sum(g.get(v, 0.0) for g in grads)
        d, F = O.cube_fd_ld(ph, geo, rp, cells, lens, cos_t, ze, rho, ne, nw, nq, v, 1e-3)
        assert F == pytest.approx(flux.sum(), rel=1e-14)
        assert d == pytest.approx(G, rel=1e-10)

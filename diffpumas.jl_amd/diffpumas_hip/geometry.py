"""Flat two-layer backward Monte Carlo on the GPU.

Host mirror of run_backward_mc / compute_flux (src/Geometry.jl:899-1338).
Materials are 0-based; the energy and charge of event e come from the
Philox stream (seed, e) instead of Julia's MersenneTwister, so results agree
with the reference statistically, not draw for draw.
"""
import ctypes as C
import math

from . import _lib

PRIMARY_ALTITUDE = 3e4  # src/Geometry.jl:36


def _zenith(elevation):
    theta = (90.0 - elevation) * math.pi / 180.0
    return math.cos(theta), math.sin(theta)


def run_backward_mc(physics, rock_thickness=100.0, rock_density=2650.0, elevation=45.0,
                    energy_min=1.0, energy_max=None, n_samples=10000, compute_gradient=False,
                    seed=42, straggling=True, scattering=True, energy_threshold_low=100.0,
                    source_energy_max=None, porous_material=-1, porous_density=0.0,
                    porous_thickness=0.0, primary_altitude=PRIMARY_ALTITUDE, rock_mixture=None,
                    return_events=False, ctx=None):
    """run_backward_mc (src/Geometry.jl:1286-1338); returns the same fields.

    rock_mixture: optional [(material index, mass fraction), ...] runtime
    mixture for the rock layer (an extension: the reference flat geometry takes
    one material).  return_events adds each event's contribution."""
    ctx = ctx or _lib.default_context()
    if energy_max is None:
        energy_max = energy_min
    rock = physics.get_material_index("StandardRock")
    air = physics.get_material_index("Air")
    if rock == -1:
        rock = 0
    if air == -1:
        air = physics.n_materials - 1
    geo = _lib.FlatGeometry(float(rock_thickness), float(rock_density), rock, air,
                            int(porous_material), float(porous_density), float(porous_thickness),
                            float(primary_altitude))
    if rock_mixture:
        if len(rock_mixture) > 4:
            raise _lib.ArgumentError("at most 4 rock mixture components")
        geo.rock_n_components = len(rock_mixture)
        for j, (m, f) in enumerate(rock_mixture):
            geo.rock_components[j] = int(m)
            geo.rock_fractions[j] = float(f)
    c, s = _zenith(float(elevation))
    cfg = _lib.McConfig(float(elevation), c, s, float(energy_min), float(energy_max),
                        float(source_energy_max) if source_energy_max else 0.0,
                        float(energy_threshold_low), int(n_samples), int(seed) & (2**64 - 1),
                        int(bool(straggling)), int(bool(scattering)), int(bool(compute_gradient)))
    out = _lib.McResult()
    h = physics.device_handle(ctx)
    events = None
    if return_events:
        import numpy as np
        events = np.zeros(int(n_samples))
        _lib.check(ctx.lib.dp_mc_flat_backward_events(ctx.h, h, C.byref(geo), C.byref(cfg),
                                                      C.byref(out), _lib.ptr(events)))
    else:
        _lib.check(ctx.lib.dp_mc_flat_backward(ctx.h, h, C.byref(geo), C.byref(cfg),
                                               C.byref(out)))
    g = compute_gradient
    res = {"flux": out.flux, "sigma": out.sigma,
            "gradient": out.gradient if g else None,
            "gradient_flux": out.gradient_flux if g else None,
            "gradient_method": "deterministic_csda_sample_average" if g else None,
            "source_energy_max": out.source_energy_max}
    if return_events:
        res["events"] = events
    return res


def compute_flux(physics, rock_density, rock_thickness, elevation, energy_min, energy_max,
                 n_samples=1000, seed=42, straggling=True, scattering=True,
                 energy_threshold_low=100.0, **kw):
    """compute_flux (src/Geometry.jl:921-928): (flux, sigma)."""
    r = run_backward_mc(physics, rock_thickness=rock_thickness, rock_density=rock_density,
                        elevation=elevation, energy_min=energy_min, energy_max=energy_max,
                        n_samples=n_samples, seed=seed, straggling=straggling,
                        scattering=scattering, energy_threshold_low=energy_threshold_low, **kw)
    return r["flux"], r["sigma"]


def compute_flux_csda_direct(physics, rock_material, air_material, rock_density, rock_thickness,
                             elevation, energy_initial, charge, primary_altitude=PRIMARY_ALTITUDE,
                             gradient=False, ctx=None):
    """compute_flux_csda_direct (src/Geometry.jl:1126-1209) on the GPU for
    scalars or broadcastable arrays of samples.  With gradient=True returns
    (flux, d flux / d rock_density) as in compute_flux_gradient_csda
    (:1247-1262)."""
    import numpy as np
    ctx = ctx or _lib.default_context()
    h = physics.device_handle(ctx)
    arrs = np.broadcast_arrays(*(np.asarray(x, dtype=np.float64) for x in
                                 (rock_density, rock_thickness, elevation, energy_initial, charge)))
    shape = arrs[0].shape
    a = [np.ascontiguousarray(x.ravel()) for x in arrs]
    n = a[0].size
    flux, d = np.zeros(n), (np.zeros(n) if gradient else None)
    _lib.check(ctx.lib.dp_csda_direct(ctx.h, h, int(rock_material), int(air_material),
                                      float(primary_altitude), n, *[_lib.ptr(x) for x in a],
                                      _lib.ptr(flux), _lib.ptr(d)))
    flux = flux.reshape(shape)
    if not gradient:
        return flux[()] if shape == () else flux
    d = d.reshape(shape)
    return (flux[()], d[()]) if shape == () else (flux, d)


def _rock_air(physics):
    # compute_flux_differentiable_csda (src/Geometry.jl:1220-1238): StandardRock
    # and Air by name, else the first and the last material
    rock = physics.get_material_index("StandardRock")
    air = physics.get_material_index("Air")
    return (0 if rock < 0 else rock), (len(physics.names) - 1 if air < 0 else air)


def compute_flux_differentiable_csda(physics, rock_density, rock_thickness, elevation,
                                     energy_final, charge, ctx=None):
    """compute_flux_differentiable_csda (src/Geometry.jl:1220-1238)."""
    rock, air = _rock_air(physics)
    return compute_flux_csda_direct(physics, rock, air, rock_density, rock_thickness, elevation,
                                    energy_final, charge, ctx=ctx)


def compute_flux_gradient_csda(physics, rock_density, rock_thickness, elevation, energy_final,
                               charge, ctx=None):
    """compute_flux_gradient_csda (src/Geometry.jl:1247-1262): (flux, d flux /
    d rock_density), the exact derivative of the same discretised program."""
    rock, air = _rock_air(physics)
    return compute_flux_csda_direct(physics, rock, air, rock_density, rock_thickness, elevation,
                                    energy_final, charge, gradient=True, ctx=ctx)

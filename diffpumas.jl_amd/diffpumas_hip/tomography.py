"""Water-fraction (w) tomography: CSDA directional flux and its sparse
Jacobian on the GPU.

Host mirror of the Tomography module's forward model (src/Tomography.jl:
179-820).  Cells and materials are 0-based.  Every flux and Jacobian value
comes from the HIP kernel behind dp_csda_w_flux_jacobian; there is no CPU
path.  The MC-calibrated corrections (CsdaCorrection, CsdaMaterialCorrection)
are module state as in the reference (global Refs, disabled by default) and
are handed to the library with every call.
"""
import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError, f64, i32, ptr
from .cube import trace_edge_paths


@dataclass
class MaterialConfig:
    """MaterialConfig (src/Tomography.jl:179-188); porous_material < 0 disables
    the shallow porous recipe."""
    rock_material: int
    water_material: int
    air_material: int
    porous_material: int
    rock_density: float
    water_density: float
    porous_density: float = 0.0
    porous_top_thickness: float = 0.0

    def c(self):
        return _lib.MaterialConfig(int(self.rock_material), int(self.water_material),
                                   int(self.air_material), int(self.porous_material),
                                   float(self.rock_density), float(self.water_density),
                                   float(self.porous_density), float(self.porous_top_thickness))


@dataclass
class SiteConfig:
    """SiteConfig (src/Tomography.jl:198-201)."""
    detector_elevation: float
    primary_altitude: float


@dataclass
class EnergySamples:
    """Vector{EnergySample} (src/Tomography.jl:203-207) as three arrays."""
    energy: np.ndarray
    weight: np.ndarray
    charge: np.ndarray

    def __len__(self):
        return len(self.energy)


@dataclass(frozen=True)
class CsdaCorrection:
    """CsdaCorrection (src/Tomography.jl:112-137)."""
    enabled: bool
    kappa_strag: float
    kappa_hard: float
    resid_a: float
    resid_b: float
    resid_c: float = 0.0
    resid_d: float = 0.0


@dataclass(frozen=True)
class CsdaMaterialCorrection:
    """CsdaMaterialCorrection (src/Tomography.jl:164-166)."""
    enabled: bool
    linear: float
    quadratic: float


_CSDA_CORRECTION = CsdaCorrection(False, 0.0, 0.0, 0.0, 0.0)            # :139
_CSDA_MATERIAL_CORRECTION = CsdaMaterialCorrection(False, 0.0, 0.0)     # :167


def set_csda_correction(enabled=True, kappa_strag=0.0, kappa_hard=0.0, resid_a=0.0,
                        resid_b=0.0, resid_c=0.0, resid_d=0.0):
    """set_csda_correction! (src/Tomography.jl:149-155)."""
    global _CSDA_CORRECTION
    _CSDA_CORRECTION = CsdaCorrection(bool(enabled), float(kappa_strag), float(kappa_hard),
                                      float(resid_a), float(resid_b), float(resid_c),
                                      float(resid_d))
    return _CSDA_CORRECTION


def get_csda_correction():
    return _CSDA_CORRECTION


def set_csda_material_correction(enabled=True, linear=0.0, quadratic=0.0):
    """set_csda_material_correction! (src/Tomography.jl:168-172); non-finite
    coefficients raise ArgumentError."""
    global _CSDA_MATERIAL_CORRECTION
    if not (math.isfinite(linear) and math.isfinite(quadratic)):
        raise ArgumentError("nonfinite material coefficients")
    _CSDA_MATERIAL_CORRECTION = CsdaMaterialCorrection(bool(enabled), float(linear),
                                                       float(quadratic))
    return _CSDA_MATERIAL_CORRECTION


def get_csda_material_correction():
    return _CSDA_MATERIAL_CORRECTION


def _push_corrections(ctx):
    c, m = _CSDA_CORRECTION, _CSDA_MATERIAL_CORRECTION
    cc = _lib.CsdaCorrection(int(c.enabled), c.kappa_strag, c.kappa_hard, c.resid_a, c.resid_b,
                             c.resid_c, c.resid_d)
    mm = _lib.CsdaMaterialCorrection(int(m.enabled), m.linear, m.quadratic)
    _lib.check(ctx.lib.dp_set_csda_correction(ctx.h, C.byref(cc)))
    _lib.check(ctx.lib.dp_set_csda_material_correction(ctx.h, C.byref(mm)))


def csda_energy_quadrature(n, energy_min, energy_max):
    """csda_energy_quadrature (src/Tomography.jl:263-274): log-midpoint nodes,
    charge 0 (charge-summed flux), weight 0.5 E span."""
    lo, hi = float(energy_min), float(energy_max)
    if int(n) <= 0:
        raise ArgumentError("CSDA quadrature requires at least one energy node")
    if not (math.isfinite(lo) and math.isfinite(hi) and 0 < lo < hi):
        raise ArgumentError("CSDA quadrature requires finite 0 < energy_min < energy_max")
    log_lo = math.log(lo)
    span = math.log(hi) - log_lo
    e = np.array([math.exp(log_lo + span * (i - 0.5) / n) for i in range(1, int(n) + 1)])
    return EnergySamples(e, 0.5 * e * span, np.zeros(int(n)))


def sample_energy_set(n, energy_min, energy_max, seed):
    """sample_energy_set (src/Tomography.jl:236-248): log-uniform energies with
    weight E log(emax/emin) and charges +/-1.  The stream is numpy's PCG64
    seeded with `seed`, not Julia's MersenneTwister."""
    rng = np.random.default_rng(int(seed))
    rk = math.log(float(energy_max) / float(energy_min))
    e = np.empty(int(n))
    w = np.empty(int(n))
    q = np.empty(int(n))
    for i in range(int(n)):
        if rk > 0:
            e[i] = energy_min * math.exp(rk * rng.random())
            w[i] = e[i] * rk
        else:
            e[i] = energy_min
            w[i] = 1.0
        q[i] = 1.0 if rng.random() > 0.5 else -1.0
    return EnergySamples(e, w, q)


def direction_vector(zenith_deg, azimuth_deg):
    """direction_vector (src/Tomography.jl:229-234)."""
    t, p = math.radians(zenith_deg), math.radians(azimuth_deg)
    dh = math.sin(t)
    return (dh * math.sin(p), dh * math.cos(p), math.cos(t))


def cell_density(w, shallow, matcfg):
    """cell_density (src/Tomography.jl:295-302): 1/rho = sum f_i/rho_i."""
    if shallow and matcfg.porous_material >= 0 and matcfg.porous_top_thickness > 0.0:
        return 1.0 / (0.5 * (1.0 - w) / matcfg.porous_density +
                      0.5 * (1.0 - w) / matcfg.rock_density + w / matcfg.water_density)
    return 1.0 / ((1.0 - w) / matcfg.rock_density + w / matcfg.water_density)


class DirectionalPaths:
    """Many DirectionalPath (src/Tomography.jl:214-223) in CSR form."""

    def __init__(self, row_ptr, cells, lengths, zenith, azimuth, remaining_rock, exit_z,
                 valid=None, distance_in_volume=None, surface_distance=None):
        self.row_ptr = i32(row_ptr)
        self.cells = i32(cells)
        self.lengths = f64(lengths)
        self.zenith = f64(zenith)
        self.azimuth = f64(azimuth)
        self.remaining_rock = f64(remaining_rock)
        self.exit_z = f64(exit_z)
        n = len(self.row_ptr) - 1
        self.valid = np.ones(n, np.uint8) if valid is None else np.ascontiguousarray(valid, np.uint8)
        self.distance_in_volume = distance_in_volume
        self.surface_distance = surface_distance
        self.cos_theta = np.array([math.cos(math.radians(z)) for z in self.zenith])

    @classmethod
    def from_list(cls, paths):
        """paths: iterable of (zenith, azimuth, [(cell, distance), ...],
        remaining_rock, exit_z[, valid])."""
        rp, cells, lens, zen, az, rem, ez, va = [0], [], [], [], [], [], [], []
        for p in paths:
            z, a, segs, r, e = p[:5]
            for c, d in segs:
                cells.append(int(c))
                lens.append(float(d))
            rp.append(len(cells))
            zen.append(z); az.append(a); rem.append(r); ez.append(e)
            va.append(p[5] if len(p) > 5 else 1)
        return cls(rp, cells, lens, zen, az, rem, ez, va)

    @property
    def n_rays(self):
        return len(self.row_ptr) - 1

    @property
    def n_segments(self):
        return int(self.row_ptr[-1])

    def max_segments(self):
        return int(np.diff(self.row_ptr).max()) if self.n_rays else 0

    def select(self, bins):
        """The paths of the given bins (0-based), in that order (paths[valid_bins])."""
        bins = np.asarray(bins, np.int64).reshape(-1)
        if bins.size and (bins.min() < 0 or bins.max() >= self.n_rays):
            raise ArgumentError("bin index out of range")
        a, b = self.row_ptr[bins], self.row_ptr[bins + 1]
        take = np.concatenate([np.arange(x, y) for x, y in zip(a, b)]) if bins.size else \
            np.zeros(0, np.int64)
        rp = np.concatenate([[0], np.cumsum(b - a)])
        pick = lambda v: None if v is None else np.asarray(v)[bins]
        return DirectionalPaths(rp, self.cells[take], self.lengths[take], self.zenith[bins],
                                self.azimuth[bins], self.remaining_rock[bins], self.exit_z[bins],
                                self.valid[bins], pick(self.distance_in_volume),
                                pick(self.surface_distance))

    def c_paths(self):
        return _lib.DirectionalPaths(self.n_rays, self.n_segments, self.max_segments(),
                                     ptr(self.row_ptr, C.c_int32), ptr(self.cells, C.c_int32),
                                     ptr(self.lengths), ptr(self.cos_theta),
                                     ptr(self.remaining_rock), ptr(self.exit_z),
                                     ptr(self.valid, C.c_uint8))


def trace_voxel_paths(x_edges, y_edges, z_edges, zeniths, azimuths, surface_distance, exit_z,
                      detectors=None, ctx=None):
    """trace_path(::VoxelVolume) (examples/lvd_tomography.jl:738-779) for many
    (zenith, azimuth) bins; surface_distance and exit_z come from the
    topography (ray_surface_exit_distance).  detectors[n, 3] places each bin's
    detector in the frame of the edges (default: the origin, as the reference's
    volume frame is centred on the detector)."""
    zen, az = f64(zeniths), f64(azimuths)
    dirs = np.array([direction_vector(z, a) for z, a in zip(zen, az)]).reshape(-1, 3)
    origins = np.zeros_like(dirs) if detectors is None else f64(detectors).reshape(-1, 3)
    rp, cells, lens, traveled, remaining = trace_edge_paths(x_edges, y_edges, z_edges, origins,
                                                            dirs, surface_distance, ctx=ctx)
    return DirectionalPaths(rp, cells, lens, zen, az, remaining, exit_z,
                            distance_in_volume=traveled, surface_distance=f64(surface_distance))


def _run(physics, shallow_flags, matcfg, site, paths, water_fractions, samples, jacobian, ctx):
    ctx = ctx or _lib.default_context()
    h = physics.device_handle(ctx)
    w = f64(water_fractions)
    sh = np.ascontiguousarray(shallow_flags, dtype=np.uint8)
    if len(sh) != len(w):
        raise ArgumentError("shallow_flags and water_fractions must have one entry per cell")
    ne, nw, nq = f64(samples.energy), f64(samples.weight), f64(samples.charge)
    es = _lib.EnergySamples(len(ne), ptr(ne), ptr(nw), ptr(nq))
    mc = matcfg.c()
    sc = _lib.SiteConfig(float(site.detector_elevation), float(site.primary_altitude))
    cp = paths.c_paths()
    flux = np.zeros(paths.n_rays)
    jac = np.zeros(max(paths.n_segments, 1)) if jacobian else None
    _push_corrections(ctx)
    _lib.check(ctx.lib.dp_csda_w_flux_jacobian(ctx.h, h, C.byref(mc), C.byref(sc), C.byref(cp),
                                               len(w), ptr(w), ptr(sh, C.c_uint8), C.byref(es),
                                               ptr(flux), ptr(jac)))
    return flux, (jac[:paths.n_segments] if jacobian else None)


def compute_directional_flux_csda(physics, shallow_flags, matcfg, site, paths, water_fractions,
                                  samples, ctx=None):
    """compute_directional_flux_csda (src/Tomography.jl:667-726) for every path
    (NaN for an invalid path or cos(zenith) <= 0)."""
    return _run(physics, shallow_flags, matcfg, site, paths, water_fractions, samples, False,
                ctx)[0]


def flux_and_segment_jacobian(physics, shallow_flags, matcfg, site, paths, water_fractions,
                              samples, ctx=None):
    """(flux[n_rays], jac[n_segments]): jac holds d flux / d w of each path's
    cell at its first occurrence on the path, 0 at later occurrences."""
    return _run(physics, shallow_flags, matcfg, site, paths, water_fractions, samples, True, ctx)


def directional_flux_and_grad_csda(physics, shallow_flags, matcfg, site, paths, water_fractions,
                                   samples, ray=0, ctx=None):
    """directional_flux_and_grad_csda (src/Tomography.jl:736-762) for one path:
    (flux, cellids, values) with the on-path cells in first-appearance order."""
    flux, jac = flux_and_segment_jacobian(physics, shallow_flags, matcfg, site, paths,
                                          water_fractions, samples, ctx)
    if not paths.valid[ray] or not paths.cos_theta[ray] > 0.0:
        return math.nan, np.zeros(0, np.int64), np.zeros(0)
    a, b = paths.row_ptr[ray], paths.row_ptr[ray + 1]
    cells, vals, seen = [], [], set()
    for k in range(a, b):
        c = int(paths.cells[k])
        if c not in seen:
            seen.add(c)
            cells.append(c)
            vals.append(jac[k])
    return float(flux[ray]), np.array(cells, np.int64), np.array(vals)


def assemble_forward_and_jacobian(physics, shallow_flags, matcfg, site, paths, water_fractions,
                                  samples, n_cells, ctx=None):
    """assemble_forward_and_jacobian (src/Tomography.jl:773-820): flux per bin
    and J[bin, cell] = d flux / d w as a scipy.sparse CSR matrix, keeping only
    finite non-zero entries."""
    import scipy.sparse as sp
    flux, jac = flux_and_segment_jacobian(physics, shallow_flags, matcfg, site, paths,
                                          water_fractions, samples, ctx)
    rows = np.repeat(np.arange(paths.n_rays), np.diff(paths.row_ptr))
    keep = np.isfinite(jac) & (jac != 0.0)
    J = sp.csr_matrix((jac[keep], (rows[keep], paths.cells[keep])), shape=(paths.n_rays, n_cells))
    return flux, J


# ------------------------------------------------------------ fixed-path MC
def mc_cell_mixture_and_density(w, shallow, matcfg):
    """mc_cell_mixture_and_density (src/Tomography.jl:826-845):
    ((materials, fractions), density) of one cell."""
    if shallow and matcfg.porous_material >= 0 and matcfg.porous_top_thickness > 0.0:
        pf = rf = 0.5 * (1.0 - w)
        mix = ((matcfg.porous_material, matcfg.rock_material, matcfg.water_material), (pf, rf, w))
        return mix, 1.0 / (pf / matcfg.porous_density + rf / matcfg.rock_density +
                           w / matcfg.water_density)
    if w <= 1e-12:
        return ((matcfg.rock_material,), (1.0,)), matcfg.rock_density
    mix = ((matcfg.rock_material, matcfg.water_material), (1.0 - w, w))
    return mix, 1.0 / ((1.0 - w) / matcfg.rock_density + w / matcfg.water_density)


def build_cell_properties_for_mc(shallow_flags, water_fractions, matcfg):
    """build_cell_properties_for_mc (src/Tomography.jl:847-861) in indexed form:
    (media [(materials, fractions)], cell_medium[n_cells], cell_density[n_cells])
    with one medium per distinct (shallow, w)."""
    media, index = [], {}
    cm = np.zeros(len(water_fractions), np.int32)
    cd = np.zeros(len(water_fractions))
    for i, (sh, w) in enumerate(zip(shallow_flags, water_fractions)):
        key = (bool(sh), float(w))
        if key not in index:
            mix, rho = mc_cell_mixture_and_density(float(w), bool(sh), matcfg)
            index[key] = (len(media), rho)
            media.append(mix)
        cm[i], cd[i] = index[key]
    return media, cm, cd


def compute_directional_flux_mc(physics, matcfg, site, paths, media, cell_medium, cell_density,
                                samples, seed, straggling=True, energy_threshold_low=100.0,
                                scattering=False, return_events=False, ctx=None):
    """compute_directional_flux_mc (src/Tomography.jl:979-1041) for every path
    in one launch: (flux[n_rays], sigma[n_rays]).  Event s of path r draws
    from Philox keyed (seed, r * n_samples + s)."""
    if scattering:
        raise ArgumentError("Fixed RayPath MC supports energy-loss fluctuations only; use "
                            "Geometry for scattered boundary retracing")
    ctx = ctx or _lib.default_context()
    h = physics.device_handle(ctx)
    mix = (_lib.Mixture * len(media))()
    for j, (mats, fr) in enumerate(media):
        if len(mats) > 4:
            raise ArgumentError("at most 4 components per medium")
        mix[j].n_components = len(mats)
        for k, (m, f) in enumerate(zip(mats, fr)):
            mix[j].components[k] = int(m)
            mix[j].fractions[k] = float(f)
    cm = i32(cell_medium)
    cd = f64(cell_density)
    ne, nw, nq = f64(samples.energy), f64(samples.weight), f64(samples.charge)
    es = _lib.EnergySamples(len(ne), ptr(ne), ptr(nw), ptr(nq))
    mc = matcfg.c()
    sc = _lib.SiteConfig(float(site.detector_elevation), float(site.primary_altitude))
    cp = paths.c_paths()
    flux = np.zeros(paths.n_rays)
    sigma = np.zeros(paths.n_rays)
    ev = np.zeros(paths.n_rays * len(ne)) if return_events else None
    _lib.check(ctx.lib.dp_mc_fixed_paths(ctx.h, h, C.byref(mc), C.byref(sc), C.byref(cp), len(cm),
                                         len(media), mix, ptr(cm, C.c_int32), ptr(cd),
                                         C.byref(es), int(seed) & (2**64 - 1), int(bool(straggling)),
                                         float(energy_threshold_low), ptr(flux), ptr(sigma),
                                         ptr(ev)))
    if return_events:
        return flux, sigma, ev.reshape(paths.n_rays, len(ne))
    return flux, sigma


def compute_directional_flux_mc_and_csda(physics, matcfg, site, paths, media, cell_medium,
                                         cell_density, samples, nodes, seed,
                                         energy_threshold_low=100.0, ctx=None):
    """In one persistent launch: the straggled fixed-path MC of every path for
    `samples` (compute_directional_flux_mc, src/Tomography.jl:979-1041) and
    the deterministic CSDA flux of every path over the quadrature `nodes`
    (the same stepping with straggling off; charge-summed nodes allowed).
    Returns (mc_flux, mc_sigma, csda_flux), each n_rays long."""
    ctx = ctx or _lib.default_context()
    h = physics.device_handle(ctx)
    mix = (_lib.Mixture * len(media))()
    for j, (mats, fr) in enumerate(media):
        if len(mats) > 4:
            raise ArgumentError("at most 4 components per medium")
        mix[j].n_components = len(mats)
        for k, (m, f) in enumerate(zip(mats, fr)):
            mix[j].components[k] = int(m)
            mix[j].fractions[k] = float(f)
    cm = i32(cell_medium)
    cd = f64(cell_density)
    se, sw, sq = f64(samples.energy), f64(samples.weight), f64(samples.charge)
    ne, nw, nq = f64(nodes.energy), f64(nodes.weight), f64(nodes.charge)
    es = _lib.EnergySamples(len(se), ptr(se), ptr(sw), ptr(sq))
    en = _lib.EnergySamples(len(ne), ptr(ne), ptr(nw), ptr(nq))
    mc = matcfg.c()
    sc = _lib.SiteConfig(float(site.detector_elevation), float(site.primary_altitude))
    cp = paths.c_paths()
    flux = np.zeros(paths.n_rays)
    sigma = np.zeros(paths.n_rays)
    cflux = np.zeros(paths.n_rays)
    _lib.check(ctx.lib.dp_mc_fixed_paths_csda(ctx.h, h, C.byref(mc), C.byref(sc), C.byref(cp),
                                              len(cm), len(media), mix, ptr(cm, C.c_int32),
                                              ptr(cd), C.byref(es), C.byref(en),
                                              int(seed) & (2**64 - 1),
                                              float(energy_threshold_low), ptr(flux),
                                              ptr(sigma), ptr(cflux)))
    return flux, sigma, cflux

"""ctypes access to the CPU oracle (oracle/liboracle_dp.so) for the tests.

The oracle is the checker only; the product path never loads it.
"""
import ctypes as C
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle_dp.so")

d_p = C.POINTER(C.c_double)
i_p = C.POINTER(C.c_int32)
u_p = C.POINTER(C.c_uint32)

_SIGS = {
    "orc_philox_block": (None, [C.c_uint32] * 6 + [u_p]),
    "orc_philox_uniforms": (None, [C.c_uint64, C.c_uint64, C.c_int, d_p]),
    "orc_flux_gccly": (C.c_double, [C.c_double] * 3),
    "orc_flux_gaisser": (C.c_double, [C.c_double] * 3),
    "orc_dflux_gccly_dk": (C.c_double, [C.c_double] * 3),
    "orc_property": (None, [C.c_int, C.c_int, d_p, d_p, C.c_int, C.c_int, C.c_int, d_p, d_p, d_p]),
    "orc_kinetic": (None, [C.c_int, C.c_int, d_p, d_p, C.c_int, C.c_int, C.c_int, d_p, d_p, d_p]),
    "orc_trace_cube": (None, [d_p, C.c_int, d_p, d_p, i_p, i_p, i_p, d_p, d_p]),
    "orc_trace_edges": (None, [d_p, C.c_int, d_p, C.c_int, d_p, C.c_int, C.c_int, d_p, d_p, d_p,
                               i_p, i_p, i_p, d_p, d_p]),
    "orc_cube_csda": (None, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, C.c_int, C.c_int,
                             C.c_int, i_p, i_p, d_p, d_p, d_p, C.c_double, C.c_double, C.c_int,
                             d_p, C.c_int, d_p, d_p, d_p, C.c_double, d_p, i_p, i_p, d_p, i_p,
                             C.c_int]),
    "orc_cube_csda_reverse": (None, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, C.c_int,
                                     C.c_int, C.c_int, i_p, i_p, d_p, d_p, d_p, C.c_double,
                                     C.c_double, d_p, C.c_int, d_p, d_p, d_p, C.c_double, d_p, d_p,
                                     d_p, C.c_int]),
    "orc_flat_csda": (None, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, d_p, i_p,
                             C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                             d_p]),
    "orc_mc_flat_csda": (None, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, d_p, i_p,
                                C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_int, C.c_longlong, C.c_ulonglong, C.c_int, C.c_int, d_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(ORACLE)
        for k, (r, a) in _SIGS.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def P(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def tabargs(ph):
    return (ph.n_materials, ph.n_knots, P(ph.data), P(ph.scalars))


def philox_block(ctr, key):
    out = (C.c_uint32 * 4)()
    lib().orc_philox_block(*[C.c_uint32(x) for x in ctr], *[C.c_uint32(x) for x in key], out)
    return list(out)


def philox_uniforms(seed, event, n):
    out = np.zeros(n)
    lib().orc_philox_uniforms(seed, event, n, P(out))
    return out


def flux_gccly(c, k, q):
    return lib().orc_flux_gccly(c, k, q)


def dflux_gccly_dk(c, k, q):
    return lib().orc_dflux_gccly_dk(c, k, q)


def flux_gaisser(c, k, q):
    return lib().orc_flux_gaisser(c, k, q)


def property_(ph, m, a, energies):
    e = f64(energies)
    out, dout = np.zeros(len(e)), np.zeros(len(e))
    lib().orc_property(*tabargs(ph), m, a, len(e), P(e), P(out), P(dout))
    return out, dout


def kinetic(ph, m, ranges, mixed=False):
    x = f64(ranges)
    out, dout = np.zeros(len(x)), np.zeros(len(x))
    lib().orc_kinetic(*tabargs(ph), m, int(mixed), len(x), P(x), P(out), P(dout))
    return out, dout


def _csr(counts):
    rp = np.zeros(len(counts) + 1, dtype=np.int32)
    np.cumsum(counts, out=rp[1:])
    return rp


def trace_cube(geo, start_xy, dirs):
    g = geo.as_array()
    sxy = f64(start_xy).reshape(-1, 2)
    d = f64(dirs).reshape(-1, 3)
    n = len(d)
    counts = np.zeros(n, dtype=np.int32)
    ze = np.zeros(n)
    lib().orc_trace_cube(P(g), n, P(sxy), P(d), P(counts, C.c_int32), None, None, None, P(ze))
    rp = _csr(counts)
    ns = int(rp[-1])
    cells = np.zeros(max(ns, 1), dtype=np.int32)
    lens = np.zeros(max(ns, 1))
    lib().orc_trace_cube(P(g), n, P(sxy), P(d), None, P(rp, C.c_int32), P(cells, C.c_int32),
                         P(lens), None)
    return rp, cells[:ns], lens[:ns], ze


def trace_edges(xe, ye, ze, origins, dirs, surface):
    xe, ye, ze = f64(xe), f64(ye), f64(ze)
    o = f64(origins).reshape(-1, 3)
    d = f64(dirs).reshape(-1, 3)
    sd = f64(surface)
    n = len(d)
    counts = np.zeros(n, dtype=np.int32)
    tr = np.zeros(n)
    L = lib()
    L.orc_trace_edges(P(xe), len(xe), P(ye), len(ye), P(ze), len(ze), n, P(o), P(d), P(sd),
                      P(counts, C.c_int32), None, None, None, P(tr))
    rp = _csr(counts)
    ns = int(rp[-1])
    cells = np.zeros(max(ns, 1), dtype=np.int32)
    lens = np.zeros(max(ns, 1))
    L.orc_trace_edges(P(xe), len(xe), P(ye), len(ye), P(ze), len(ze), n, P(o), P(d), P(sd), None,
                      P(rp, C.c_int32), P(cells, C.c_int32), P(lens), None)
    return rp, cells[:ns], lens[:ns], tr


def cube_csda(ph, geo, rp, cells, lens, cos_t, z_exit, dens, ne, nw, nq, scale=1.0,
              palt=3e4, gradient=True, threads=0):
    """Per-ray flux and (optionally) per-(ray, distinct cell) derivatives by
    forward-mode seeding.  Returns flux[, dict ray -> {cell: d flux/d rho}]."""
    rp, cells, lens = i32(rp), i32(cells), f64(lens)
    cos_t, z_exit, dens = f64(cos_t), f64(z_exit), f64(dens)
    ne, nw, nq = f64(ne), f64(nw), f64(nq)
    n = len(rp) - 1
    flux = np.zeros(n)
    gp = gc = gv = gn = None
    if gradient:
        gp = _csr(np.diff(rp) + 1)
        cap = int(gp[-1])
        gc = np.zeros(max(cap, 1), dtype=np.int32)
        gv = np.zeros(max(cap, 1))
        gn = np.zeros(n, dtype=np.int32)
    lib().orc_cube_csda(*tabargs(ph), ph.mass, ph.ctau, geo.rock_material, geo.air_material, n,
                        P(rp, C.c_int32), P(cells if len(cells) else np.zeros(1, np.int32), C.c_int32),
                        P(lens if len(lens) else np.zeros(1)), P(cos_t), P(z_exit), geo.z_max, palt,
                        len(dens), P(dens), len(ne), P(ne), P(nw), P(nq), scale, P(flux),
                        P(gp, C.c_int32), P(gc, C.c_int32), P(gv), P(gn, C.c_int32), threads)
    if not gradient:
        return flux
    grads = []
    for r in range(n):
        o = gp[r]
        grads.append({int(gc[o + k]): float(gv[o + k]) for k in range(gn[r])})
    return flux, grads


def cube_csda_reverse(ph, rock, air, rp, cells, lens, cos_t, z_exit, z_max, dens, ne, nw, nq,
                      scale=1.0, palt=3e4, threads=0):
    """Reverse-mode oracle: (flux[n_rays], jac[n_segments], jac_rem[n_rays])."""
    rp, cells, lens = i32(rp), i32(cells), f64(lens)
    cos_t, z_exit, dens = f64(cos_t), f64(z_exit), f64(dens)
    ne, nw, nq = f64(ne), f64(nw), f64(nq)
    n = len(rp) - 1
    flux, jrem = np.zeros(n), np.zeros(n)
    jac = np.zeros(max(int(rp[-1]), 1))
    lib().orc_cube_csda_reverse(*tabargs(ph), ph.mass, ph.ctau, rock, air, n, P(rp, C.c_int32),
                                P(cells if len(cells) else np.zeros(1, np.int32), C.c_int32),
                                P(lens if len(lens) else np.zeros(1)), P(cos_t), P(z_exit), z_max,
                                palt, P(dens), len(ne), P(ne), P(nw), P(nq), scale, P(flux), P(jac),
                                P(jrem), threads)
    return flux, jac[:int(rp[-1])], jrem


def flat_csda(ph, geo8, mats3, energy, cos_t, charge, threshold=1e9, mode_switch=100.0,
              scat_limit=False):
    out = np.zeros(5)
    g = f64(geo8)
    m = i32(mats3)
    lib().orc_flat_csda(*tabargs(ph), ph.mass, ph.ctau, P(g), P(m, C.c_int32), energy, cos_t,
                        charge, threshold, mode_switch, int(scat_limit), P(out))
    return out


def mc_flat_csda(ph, geo8, mats3, emin, emax, cos_t, n, seed, threshold=1e9, mode_switch=100.0,
                 threads=0, gradient=True):
    out = np.zeros(4)
    g = f64(geo8)
    m = i32(mats3)
    lib().orc_mc_flat_csda(*tabargs(ph), ph.mass, ph.ctau, P(g), P(m, C.c_int32), emin, emax,
                           cos_t, threshold, mode_switch, 0, n, seed, threads, int(gradient), P(out))
    return out


def recipes_arrays(ph):
    """Per-material element recipes of a PhysicsTables as the flat arrays
    orc_mc_flat takes."""
    n_el, Z, A, I, f = [], [], [], [], []
    for name in ph.names:
        r = ph.recipes[name]
        n_el.append(len(r.elements))
        for e, fr in zip(r.elements, r.fractions):
            Z.append(e.Z); A.append(e.A); I.append(e.I); f.append(fr)
    return (np.array(n_el, dtype=np.int32), np.array(Z), np.array(A), np.array(I), np.array(f))


_SIGS["orc_mc_flat"] = (C.c_int, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, i_p, d_p,
                                  d_p, d_p, d_p, C.c_int, i_p, d_p, C.c_double, C.c_double,
                                  C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                                  C.c_longlong, C.c_ulonglong, C.c_int, d_p, d_p,
                                  C.POINTER(C.c_longlong)])


def mc_flat(ph, rock_mats, rock_fracs, rock_rho, thick, air, elevation, emin, emax, n, seed,
            threshold=1e9, mode_switch=100.0, straggling=True, scattering=True, palt=3e4,
            porous=-1, porous_rho=0.0, porous_thick=0.0, threads=0, per_event=False,
            cutoff=0.05, elastic_ratio=0.05):
    """Stochastic flat backward MC on the CPU oracle: (flux, sigma[, w_events])."""
    import math
    L = lib()
    n_el, Z, A, I, f = recipes_arrays(ph)
    rm, rf = i32(rock_mats), f64(rock_fracs)
    theta = (90.0 - elevation) * math.pi / 180.0
    out = np.zeros(2)
    ev = np.zeros(n) if per_event else None
    bad = C.c_longlong(-1)
    rc = L.orc_mc_flat(*tabargs(ph), ph.mass, ph.ctau, P(n_el, C.c_int32), P(Z), P(A), P(I), P(f),
                       len(rm), P(rm, C.c_int32), P(rf), rock_rho, thick, air, porous, porous_rho,
                       porous_thick, palt, math.cos(theta), math.sin(theta), emin, emax,
                       threshold, mode_switch, int(straggling), int(scattering), cutoff,
                       elastic_ratio, n, seed, threads, P(ev), P(out), C.byref(bad))
    if rc != 0:
        raise RuntimeError(f"oracle MC error {rc} at event {bad.value}")
    return (out[0], out[1], ev) if per_event else (out[0], out[1])


# ------------------------------------------------ w-parameterised tomography
u8_p = C.POINTER(C.c_uint8)
_W_COMMON = [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, i_p, d_p, d_p]
_W_RAYS = [C.c_int, i_p, i_p, d_p, d_p, d_p, d_p, u8_p, d_p, u8_p, C.c_int, d_p, d_p, d_p]
_SIGS["orc_w_cell_segment"] = (None, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, i_p,
                                      d_p, C.c_double, C.c_int, C.c_double, C.c_double, d_p])
_SIGS["orc_w_air_segment"] = (None, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, C.c_int,
                                     d_p, C.c_double, C.c_double, C.c_double, C.c_double, d_p])
_SIGS["orc_w_flux"] = (None, _W_COMMON + _W_RAYS + [d_p, d_p, C.c_int])
_SIGS["orc_w_jacobian"] = (None, _W_COMMON + _W_RAYS + [d_p, d_p, d_p, C.c_int])
_SIGS["orc_w_dflux"] = (C.c_double, _W_COMMON + [C.c_int, i_p, d_p, C.c_double, C.c_double,
                                                 C.c_double, C.c_int, d_p, u8_p, C.c_int,
                                                 C.c_int, d_p, d_p, d_p, d_p])
_SIGS["orc_w_node_walk"] = (None, _W_COMMON + [C.c_int, i_p, d_p, C.c_double, C.c_double,
                                               C.c_double, d_p, u8_p, C.c_double, d_p])
_SIGS["orc_spectral_index"] = (C.c_double, [C.c_double, C.c_double])
_SIGS["orc_var_correction"] = (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double, d_p])


_SIGS["orc_csda_direct"] = (None, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, C.c_int,
                                   C.c_int, C.c_double, C.c_int, d_p, d_p, d_p, d_p, d_p, d_p,
                                   d_p])


_SIGS["orc_ray_surface_exit"] = (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, d_p, C.c_double, C.c_double, C.c_int, d_p,
                                        C.c_double, C.c_double, d_p, u8_p, d_p])
_SIGS["orc_dem_elevation"] = (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_double, d_p, C.c_int, d_p, d_p, u8_p])


def _dem_args(dem):
    """dem: dict nx, ny, x0, y0, dx, dy, data (ny x nx array, z[iy, ix])."""
    data = np.asfortranarray(dem["data"], dtype=np.float64)   # column-major: iy + ny * ix
    return (dem["nx"], dem["ny"], dem["x0"], dem["y0"], dem["dx"], dem["dy"],
            data.ctypes.data_as(d_p)), data


def ray_surface_exit(dem, dirs, xy_scale=1e-3, z_ref=0.0, step=40.0, max_distance=25000.0):
    """(distance, found, exit_z) of ray_surface_exit_distance for each direction."""
    args, keep = _dem_args(dem)
    d = f64(np.asarray(dirs).reshape(-1))
    n = len(d) // 3
    dist, ez = np.zeros(n), np.zeros(n)
    found = np.zeros(n, np.uint8)
    lib().orc_ray_surface_exit(*args, xy_scale, z_ref, n, P(d), step, max_distance, P(dist),
                               P(found, C.c_uint8), P(ez))
    return dist, found.astype(bool), ez


def dem_elevation(dem, xy):
    args, keep = _dem_args(dem)
    p = f64(np.asarray(xy).reshape(-1))
    n = len(p) // 2
    z, inside = np.zeros(n), np.zeros(n, np.uint8)
    lib().orc_dem_elevation(*args, n, P(p), P(z), P(inside, C.c_uint8))
    return z, inside.astype(bool)


def csda_direct(ph, rock, air, rho, thick, elev, energy, charge, primary_altitude=3e4):
    """(flux, d flux / d rho) of compute_flux_csda_direct for arrays of samples."""
    a = [f64(np.broadcast_to(x, np.broadcast(rho, thick, elev, energy, charge).shape).ravel())
         for x in (rho, thick, elev, energy, charge)]
    n = len(a[0])
    flux, d = np.zeros(n), np.zeros(n)
    lib().orc_csda_direct(*tabargs(ph), ph.mass, ph.ctau, rock, air, primary_altitude, n,
                          *[P(x) for x in a], P(flux), P(d))
    return flux, d


def corr_array(enabled=True, kappa_strag=0.0, kappa_hard=0.0, resid_a=0.0, resid_b=0.0,
               resid_c=0.0, resid_d=0.0, material=None):
    """CsdaCorrection (src/Tomography.jl:112-137) and, if material = (linear,
    quadratic), an enabled CsdaMaterialCorrection (:164-172), packed as the
    oracle's 10 doubles."""
    lin, quad = material if material is not None else (0.0, 0.0)
    return f64([float(enabled), kappa_strag, kappa_hard, resid_a, resid_b, resid_c, resid_d,
                float(material is not None), lin, quad])


def _cp(corr):
    return None if corr is None else P(corr)


def matcfg_arrays(rock, water, air, porous, rho_rock, rho_water, rho_porous=0.0, porous_top=0.0):
    """MaterialConfig (src/Tomography.jl:179-188) with 0-based materials
    (porous < 0: no shallow porous recipe)."""
    return i32([rock, water, air, porous]), f64([rho_rock, rho_water, rho_porous, porous_top])


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def w_cell_segment(ph, mc, w, shallow, distance, energy):
    out = np.zeros(7)
    mi, md = mc
    lib().orc_w_cell_segment(*tabargs(ph), ph.mass, ph.ctau, P(mi, C.c_int32), P(md), w,
                             int(shallow), distance, energy, P(out))
    return out


def w_air_segment(ph, air, site, start_z, cos_t, distance, energy):
    out = np.zeros(5)
    s = f64(site)
    lib().orc_w_air_segment(*tabargs(ph), ph.mass, ph.ctau, air, P(s), start_z, cos_t, distance,
                            energy, P(out))
    return out


def _wrays(paths):
    rp, cells, lens = i32(paths["row_ptr"]), i32(paths["cells"]), f64(paths["lengths"])
    n = len(rp) - 1
    if len(cells) == 0:
        cells, lens = np.zeros(1, np.int32), np.zeros(1)
    return (n, P(rp, C.c_int32), P(cells, C.c_int32), P(lens), P(f64(paths["remaining"])),
            P(f64(paths["exit_z"])), P(f64(paths["cos_theta"])),
            P(_u8(paths["valid"]), C.c_uint8)), (rp, cells, lens)


def w_flux(ph, mc, site, paths, w, shallow, ne, nw, nq, threads=0, corr=None):
    mi, md = mc
    s = f64(site)
    args, keep = _wrays(paths)
    w, sh = f64(w), _u8(shallow)
    ne, nw, nq = f64(ne), f64(nw), f64(nq)
    flux = np.zeros(args[0])
    lib().orc_w_flux(*tabargs(ph), ph.mass, ph.ctau, P(mi, C.c_int32), P(md), P(s), *args, P(w),
                     P(sh, C.c_uint8), len(ne), P(ne), P(nw), P(nq), _cp(corr), P(flux),
                     threads)
    return flux


def w_jacobian(ph, mc, site, paths, w, shallow, ne, nw, nq, threads=0, corr=None):
    """(flux[n_rays], jac[n_segments]) with repeated cells merged into their
    first occurrence."""
    mi, md = mc
    s = f64(site)
    args, keep = _wrays(paths)
    w, sh = f64(w), _u8(shallow)
    ne, nw, nq = f64(ne), f64(nw), f64(nq)
    flux = np.zeros(args[0])
    jac = np.zeros(max(len(paths["cells"]), 1))
    lib().orc_w_jacobian(*tabargs(ph), ph.mass, ph.ctau, P(mi, C.c_int32), P(md), P(s), *args,
                         P(w), P(sh, C.c_uint8), len(ne), P(ne), P(nw), P(nq), _cp(corr),
                         P(flux), P(jac), threads)
    return flux, jac[:len(paths["cells"])]


def w_dflux(ph, mc, site, cells, lens, remaining, exit_z, cos_t, w, shallow, cell, ne, nw, nq,
            corr=None):
    """Forward-mode d flux / d w[cell] of one path."""
    mi, md = mc
    s = f64(site)
    cells, lens = i32(cells), f64(lens)
    w, sh = f64(w), _u8(shallow)
    ne, nw, nq = f64(ne), f64(nw), f64(nq)
    return lib().orc_w_dflux(*tabargs(ph), ph.mass, ph.ctau, P(mi, C.c_int32), P(md), P(s),
                             len(cells), P(cells, C.c_int32), P(lens), remaining, exit_z, cos_t,
                             len(w), P(w), P(sh, C.c_uint8), cell, len(ne), P(ne), P(nw), P(nq),
                             _cp(corr))


def w_node_walk(ph, mc, site, cells, lens, remaining, exit_z, cos_t, w, shallow, energy):
    """One energy node: (surface energy, weight, V, H, ok)."""
    mi, md = mc
    s = f64(site)
    cells, lens = i32(cells), f64(lens)
    ns = len(cells)
    if ns == 0:
        cells, lens = np.zeros(1, np.int32), np.zeros(1)
    w, sh = f64(w), _u8(shallow)
    out = np.zeros(5)
    lib().orc_w_node_walk(*tabargs(ph), ph.mass, ph.ctau, P(mi, C.c_int32), P(md), P(s),
                          ns, P(cells, C.c_int32),
                          P(lens), remaining, exit_z, cos_t, P(w), P(sh, C.c_uint8), energy, P(out))
    return out


def spectral_index(energy, cos_t):
    return lib().orc_spectral_index(energy, cos_t)


def var_correction(energy, V, H, cos_t, corr):
    return lib().orc_var_correction(energy, V, H, cos_t, P(corr))


# ----------------------------------------- mixed kernel / EHS scripted pins
_K_ARGS = [C.c_int, d_p, d_p, d_p, d_p, d_p, d_p, C.c_int, d_p, d_p, d_p, d_p, C.c_double,
           C.c_double, C.c_double, C.c_double]
_SIGS["orc_mixed_step"] = (C.c_int, _K_ARGS + [d_p, C.c_double, C.c_double, C.c_double, d_p,
                                               C.c_int, C.c_uint64, C.POINTER(C.c_int)])
_SIGS["orc_reverse_del"] = (C.c_int, _K_ARGS + [C.c_double, d_p, C.c_int, d_p])
_SIGS["orc_ehs_sample"] = (C.c_int, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, C.c_int,
                                     i_p, d_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, d_p, C.c_int, d_p])
_SIGS["orc_mix_range"] = (C.c_double, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double,
                                       C.c_int, i_p, d_p, C.c_int, C.c_double])
_SIGS["orc_ehs_path"] = (C.c_double, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double,
                                      C.c_int, i_p, d_p, C.c_double, C.c_double])

MUON_MASS = 0.10565839
MUON_C_TAU = 658.654


class Kernel:
    """MixedKernel fields (src/MixedTransport.jl:15-29) as arrays."""

    def __init__(self, energies, stopping, sigma, range_, hazard, proper, elements=(),
                 onset=math.inf, cutoff=0.05, mass=MUON_MASS,
                 ctau=MUON_C_TAU):
        self.arrs = [f64(a) for a in (energies, stopping, sigma, range_, hazard, proper)]
        el = list(elements)
        self.Z = f64([e[0] for e in el] or [0.0])
        self.A = f64([e[1] for e in el] or [1.0])
        self.I = f64([e[2] for e in el] or [0.0])
        self.f = f64([e[3] for e in el] or [0.0])
        self.n_el = len(el)
        self.onset, self.cutoff, self.mass, self.ctau = onset, cutoff, mass, ctau

    def args(self):
        return ([len(self.arrs[0])] + [P(a) for a in self.arrs] +
                [self.n_el, P(self.Z), P(self.A), P(self.I), P(self.f), self.onset, self.cutoff,
                 self.mass, self.ctau])


def mixed_step(kern, state, density, distance, energy_limit=-1.0, script=None, seed=0):
    """state: dict E, w, t, pos, dir, X, L.  Returns (rc, new state, draws used)."""
    st = f64([state["E"], state.get("w", 1.0), state.get("t", 0.0), *state.get("pos", (0, 0, 0)),
              *state.get("dir", (0, 0, 1)), state.get("X", 0.0), state.get("L", 0.0)])
    sc = None if script is None else f64(script)
    used = C.c_int(0)
    rc = lib().orc_mixed_step(*kern.args(), P(st), density, distance, energy_limit, P(sc),
                              0 if sc is None else len(sc), seed, C.byref(used))
    out = dict(E=st[0], w=st[1], t=st[2], pos=st[3:6].copy(), dir=st[6:9].copy(), X=st[9],
               L=st[10])
    return rc, out, used.value


def reverse_del(kern, k, script):
    sc = f64(script)
    out = np.zeros(2)
    used = lib().orc_reverse_del(*kern.args(), k, P(sc), len(sc), P(out))
    return out[0], out[1], used


def _mix(mats, fracs):
    return i32(mats), f64(fracs)


def ehs_sample(ph, mats, fracs, ki, kf, dX, Xtot, ratio, script, er=0.05):
    m, f = _mix(mats, fracs)
    sc = f64(script)
    out = np.zeros(3)
    used = lib().orc_ehs_sample(*tabargs(ph), ph.mass, ph.ctau, len(m), P(m, C.c_int32), P(f), ki,
                                kf, dX, Xtot, ratio, er, P(sc), len(sc), P(out))
    return bool(out[0]), out[1], out[2], used


def mix_range(ph, mats, fracs, k, mixed=True):
    m, f = _mix(mats, fracs)
    return lib().orc_mix_range(*tabargs(ph), ph.mass, ph.ctau, len(m), P(m, C.c_int32), P(f),
                               int(mixed), k)


def ehs_path(ph, mats, fracs, k, er=0.05):
    m, f = _mix(mats, fracs)
    return lib().orc_ehs_path(*tabargs(ph), ph.mass, ph.ctau, len(m), P(m, C.c_int32), P(f), k, er)


# ------------------------------------------------------------ fixed-path MC
_SIGS["orc_mc_paths"] = (C.c_int, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double, i_p, d_p,
                                   d_p, d_p, d_p, C.c_double, C.c_double, C.c_int, i_p, i_p, d_p,
                                   i_p, d_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_int, i_p, i_p, d_p, d_p, d_p, d_p, u8_p,
                                   C.c_int, d_p, d_p, d_p, C.c_ulonglong, C.c_int, C.c_double,
                                   C.c_int, d_p, d_p, d_p, C.POINTER(C.c_longlong)])
_SIGS["orc_mc_uniform_segment"] = (C.c_int, [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double,
                                             i_p, d_p, d_p, d_p, d_p, C.c_double, C.c_int,
                                             C.c_double, C.c_double, C.c_int, C.c_double,
                                             C.c_ulonglong, d_p])


def mc_paths(ph, media, cell_medium, cell_density, rock, air, rho_rock, site, paths, samples,
             seed, straggling=True, thr_low=100.0, threads=0, per_event=False, cutoff=0.05,
             elastic_ratio=0.05):
    """compute_directional_flux_mc for every path: (flux, sigma[, events]).
    media: [(mats, fracs)], cell_medium/cell_density per cell; samples
    (energy, weight, charge) arrays."""
    n_el, Z, A, I, f = recipes_arrays(ph)
    mn = i32([len(m) for m, _ in media])
    mm = np.zeros((len(media), 4), np.int32)
    mf = np.zeros((len(media), 4))
    for j, (m, fr) in enumerate(media):
        mm[j, :len(m)] = m
        mf[j, :len(fr)] = fr
    cmed, cden = i32(cell_medium), f64(cell_density)
    args, keep = _wrays(paths)
    se, sw, sq = (f64(a) for a in samples)
    n = args[0]
    flux, sigma = np.zeros(n), np.zeros(n)
    ev = np.zeros(n * len(se)) if per_event else None
    bad = C.c_longlong(-1)
    rc = lib().orc_mc_paths(*tabargs(ph), ph.mass, ph.ctau, P(n_el, C.c_int32), P(Z), P(A), P(I),
                            P(f), cutoff, elastic_ratio, len(media), P(mn, C.c_int32),
                            P(mm, C.c_int32), P(mf), P(cmed, C.c_int32), P(cden), rock, air,
                            rho_rock, site[0], site[1], *args[:6], args[6], args[7], len(se),
                            P(se), P(sw), P(sq), seed, int(straggling), thr_low, threads,
                            P(flux), P(sigma), P(ev), C.byref(bad))
    if rc != 0:
        raise RuntimeError(f"oracle path MC error {rc} at event {bad.value}")
    return (flux, sigma, ev) if per_event else (flux, sigma)


def mc_uniform_segment(ph, material, density, distance, state, straggling=True, thr_low=100.0,
                       seed=0, cutoff=0.05):
    n_el, Z, A, I, f = recipes_arrays(ph)
    st = f64([state.get("E"), state.get("w", 1.0), state.get("t", 0.0), state.get("L", 0.0),
              state.get("X", 0.0)])
    rc = lib().orc_mc_uniform_segment(*tabargs(ph), ph.mass, ph.ctau, P(n_el, C.c_int32), P(Z),
                                      P(A), P(I), P(f), cutoff, material, density, distance,
                                      int(straggling), thr_low, seed, P(st))
    return rc, dict(E=st[0], w=st[1], t=st[2], L=st[3], X=st[4])


def normal_equations(J, residual=None, weights=None, active=None, normal=True):
    """(normal, grad) of the Gauss-Newton system (gauss_newton_reconstruct,
    src/Tomography_inverse.jl:489-494) in the reference's summation order.
    J: scipy CSR (columns ascending within rows); active: ascending column ids
    or None (every column)."""
    import scipy.sparse as sp
    J = sp.csr_matrix(J, dtype=np.float64, copy=True)
    J.sort_indices()
    rp, col, val = i32(J.indptr), i32(J.indices), f64(J.data)
    n_rows, n_cols = J.shape
    act = None if active is None else i32(active)
    n_act = n_cols if act is None else len(act)
    out_n = np.zeros(n_act * n_act) if normal else None
    out_g = np.zeros(n_act) if residual is not None else None
    w = None if weights is None else f64(weights)
    r = None if residual is None else f64(residual)
    q = lambda x, t=C.c_double: None if x is None else P(x, t)
    lib().orc_normal_equations(n_rows, n_cols, P(rp, C.c_int), P(col, C.c_int), P(val), q(w),
                               q(r), n_act, q(act, C.c_int), q(out_n), q(out_g))
    return (None if out_n is None else out_n.reshape(n_act, n_act).T), out_g


def tet_neighbors(tets):
    """build_tetra_neighbors (examples/lvd_tomography.jl:431-451): 0-based,
    -1 for a boundary face; face f of a tet omits vertex 3 - f."""
    t = i32(np.asarray(tets).reshape(-1))
    nbr = np.zeros(len(t), np.int32)
    lib().orc_tet_neighbors(len(t) // 4, P(t, C.c_int), P(nbr, C.c_int))
    return nbr.reshape(-1, 4)


def point_in_tet(points, tets, p):
    """point_in_tetrahedron for every (tet, point): bool[n_tets, n_points]."""
    pts, t, q = f64(np.asarray(points).reshape(-1)), i32(np.asarray(tets).reshape(-1)), \
        f64(np.asarray(p).reshape(-1))
    out = np.zeros((len(t) // 4) * (len(q) // 3), np.uint8)
    lib().orc_point_in_tet(len(pts) // 3, len(t) // 4, P(pts), P(t, C.c_int), len(q) // 3, P(q),
                           P(out, C.c_uint8))
    return out.reshape(len(t) // 4, len(q) // 3).astype(bool)


def trace_tetra(points, tets, nbr, origins, dirs, surface):
    """trace_path(::TetraVolume) per ray: (row_ptr, cells, lengths, traveled, start)."""
    pts, t, nb = f64(np.asarray(points).reshape(-1)), i32(np.asarray(tets).reshape(-1)), \
        i32(np.asarray(nbr).reshape(-1))
    o, d, s = f64(np.asarray(origins).reshape(-1)), f64(np.asarray(dirs).reshape(-1)), f64(surface)
    n = len(s)
    counts, start = np.zeros(n, np.int32), np.zeros(n, np.int32)
    tr = np.zeros(n)
    args = (len(pts) // 3, len(t) // 4, P(pts), P(t, C.c_int), P(nb, C.c_int), n, P(o), P(d), P(s))
    lib().orc_trace_tetra(*args, P(counts, C.c_int), None, None, None, P(tr), P(start, C.c_int))
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(counts, out=rp[1:])
    cells, lens = np.zeros(max(rp[-1], 1), np.int32), np.zeros(max(rp[-1], 1))
    lib().orc_trace_tetra(*args, P(counts, C.c_int), P(rp, C.c_int), P(cells, C.c_int), P(lens),
                          P(tr), P(start, C.c_int))
    return rp, cells[:rp[-1]], lens[:rp[-1]], tr, start


# ------------------------------------------------ DCS models and samplers
_TAB = [C.c_int, C.c_int, d_p, d_p, C.c_double, C.c_double]
_SIGS["orc_dcs"] = (None, [C.c_int] + [C.c_double] * 5 + [C.c_int, d_p, d_p])
_SIGS["orc_dcs_integral"] = (C.c_double, [C.c_int] + [C.c_double] * 7 + [C.c_int, C.c_int])
_SIGS["orc_dilog"] = (None, [C.c_int, d_p, d_p])
_SIGS["orc_gauss"] = (None, [C.c_int, d_p, d_p])
_SIGS["orc_rotate"] = (None, [C.c_int, d_p, d_p, d_p, d_p])
_SIGS["orc_truncated_randn"] = (None, [C.c_ulonglong, C.c_int, d_p])
_SIGS["orc_fluctuate"] = (C.c_int, _TAB + [C.c_int, i_p, d_p, C.c_double, C.c_double, C.c_double,
                                           C.c_ulonglong, C.c_int, d_p, C.c_int, d_p, d_p])
_SIGS["orc_soft_scattering"] = (None, _TAB + [C.c_int, i_p, d_p, C.c_double, C.c_double,
                                              C.c_double, C.c_double, C.c_ulonglong, C.c_int,
                                              d_p])
_SIGS["orc_scattering_angle"] = (None, _TAB + [C.c_int, i_p, d_p, C.c_double, C.c_int,
                                               C.c_ulonglong, C.c_int, d_p, d_p, d_p])
_SIGS["orc_polar"] = (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_ulonglong, C.c_int, d_p, C.c_int, d_p])
_lib = None


def dcs(p, Z, A, I, m, K, q):
    q = f64(np.atleast_1d(q))
    out = np.zeros(len(q))
    lib().orc_dcs(p, Z, A, I, m, K, len(q), P(q), P(out))
    return out


def dcs_integral(p, Z, A, I, m, K, xlow, xhigh, mode, npoints=180):
    return lib().orc_dcs_integral(p, Z, A, I, m, K, xlow, xhigh, mode, npoints)


def dilog(x):
    x = f64(np.atleast_1d(x))
    out = np.zeros(len(x))
    lib().orc_dilog(len(x), P(x), P(out))
    return out


def gauss_rule(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss(n, P(x), P(w))
    return x, w


def rotate(d, mu, u):
    d = f64(np.atleast_2d(d))
    mu, u = f64(np.broadcast_to(mu, len(d))), f64(np.broadcast_to(u, len(d)))
    out = np.zeros_like(d)
    lib().orc_rotate(len(d), P(d), P(mu), P(u), P(out))
    return out


def truncated_randn(seed, n):
    out = np.zeros(n)
    lib().orc_truncated_randn(seed, n, P(out))
    return out


def fluctuate(ph, mats, fracs, ki, dX, Xtot, seed=0, n=1, script=None):
    """(kout, ratio) arrays from n seeded streams, or (kout, ratio, draws
    consumed) of one scripted call."""
    m, f = _mix(mats, fracs)
    sc = None if script is None else f64(script)
    nn = 1 if sc is not None else n
    k, r = np.zeros(nn), np.zeros(nn)
    used = lib().orc_fluctuate(*tabargs(ph), ph.mass, ph.ctau, len(m), P(m, C.c_int32), P(f), ki,
                               dX, Xtot, seed, nn, P(sc), 0 if sc is None else len(sc), P(k), P(r))
    return (k[0], r[0], used) if sc is not None else (k, r)


def soft_scattering(ph, mats, fracs, k0, k1, step, density, seed, n):
    m, f = _mix(mats, fracs)
    out = np.zeros(n)
    lib().orc_soft_scattering(*tabargs(ph), ph.mass, ph.ctau, len(m), P(m, C.c_int32), P(f), k0,
                              k1, step, density, seed, n, P(out))
    return out


def scattering_angle(ph, mats, fracs, k, mode=0, seed=0, n=1, mu1=None, data=None):
    """mode 0: lab-frame samples; 1: centre-of-mass draws; 2: lab transform
    of the given mu1.  Returns (values, (mu0, A, fspin, gCM, tau)).  `data`
    replaces the table array (e.g. with the spin factor zeroed)."""
    m, f = _mix(mats, fracs)
    mu1 = None if mu1 is None else f64(mu1)
    nn = len(mu1) if mode == 2 else n
    out, kin = np.zeros(nn), np.zeros(5)
    d = f64(ph.data if data is None else data)
    lib().orc_scattering_angle(ph.n_materials, ph.n_knots, P(d), P(ph.scalars), ph.mass, ph.ctau,
                               len(m), P(m, C.c_int32), P(f), k, mode, seed, nn, P(mu1), P(out),
                               P(kin))
    return out, kin


def polar(proc, m, ki, kf, Z=11.0, seed=0, n=1, script=None):
    sc = None if script is None else f64(script)
    nn = 1 if sc is not None else n
    out = np.zeros(nn)
    used = lib().orc_polar(proc, m, ki, kf, Z, seed, nn, P(sc), 0 if sc is None else len(sc),
                           P(out))
    return (out[0], used) if sc is not None else out


_SIGS["orc_cube_fd_ld"] = (None, _TAB + [C.c_int, C.c_int, C.c_int, i_p, i_p, d_p, d_p, d_p,
                                         C.c_double, C.c_double, C.c_int, d_p, C.c_int, d_p, d_p,
                                         d_p, C.c_double, C.c_int, C.c_double, C.c_int, d_p])
_lib = None


def cube_fd_ld(ph, geo, rp, cells, lens, cos_t, z_exit, dens, ne, nw, nq, voxel, h, scale=1.0,
               palt=3e4, threads=0):
    """Central difference of the summed flux of the given rays with respect
    to the density of `voxel`, evaluated in long double.  Returns (d, F)."""
    out = np.zeros(2)
    rp, cells, lens = i32(rp), i32(cells), f64(lens)
    lib().orc_cube_fd_ld(*tabargs(ph), ph.mass, ph.ctau, geo.rock_material, geo.air_material,
                         len(rp) - 1, P(rp, C.c_int32), P(cells, C.c_int32), P(lens),
                         P(f64(cos_t)), P(f64(z_exit)), geo.z_max, palt, len(dens), P(f64(dens)),
                         len(ne), P(f64(ne)), P(f64(nw)), P(f64(nq)), scale, int(voxel), h,
                         threads, P(out))
    return out[0], out[1]


def sample_mixture_material(ph, mats, fracs, energies, seed=0, script=None):
    """sample_mixture_material per energy (seeded), or one scripted call:
    (material, draws consumed)."""
    m, f = i32(mats), f64(fracs)
    e = f64(np.atleast_1d(energies))
    out = np.zeros(len(e), np.int32)
    sc = None if script is None else f64(script)
    used = lib().orc_sample_mixture_material(ph.n_materials, ph.n_knots, P(ph.data),
                                             P(ph.scalars), len(m), P(m, C.c_int32), P(f),
                                             len(e), P(e), C.c_ulonglong(seed),
                                             None if sc is None else P(sc),
                                             0 if sc is None else len(sc), P(out, C.c_int32))
    if used < 0:
        raise RuntimeError("scripted draws exhausted")
    return (int(out[0]), used) if script is not None else out


def sample_del_event(ph, m, ki, kf, dX, Xtot, ratio, cutoff=0.05, backward=False, seed=0,
                     script=None):
    """sample_del_event: rows (occurred, X_del, k_del, process) per sample
    (seeded), or one scripted call: (row, draws consumed)."""
    cols = [f64(np.atleast_1d(np.asarray(x, dtype=np.float64))) for x in
            np.broadcast_arrays(ki, kf, dX, Xtot, ratio)]
    n = len(cols[0])
    out = np.zeros(4 * n)
    sc = None if script is None else f64(script)
    used = lib().orc_sample_del_event(ph.n_materials, ph.n_knots, P(ph.data), P(ph.scalars), m,
                                      n, *[P(c) for c in cols], C.c_double(cutoff),
                                      1 if backward else 0,
                                      C.c_ulonglong(seed), None if sc is None else P(sc),
                                      0 if sc is None else len(sc), P(out))
    if used < 0:
        raise RuntimeError("scripted draws exhausted")
    rows = out.reshape(n, 4)
    return (rows[0], used) if script is not None else rows

// CPU oracle: voxel traversal and deterministic CSDA transport.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#pragma once
#include "oracle_core.hpp"
#include <algorithm>

namespace orc {

// ================================================================ traversal
// Uniform cube grid of examples/diff_flux_cube.jl:85-129 (fields as stored
// by the CubeGeometry constructor; cell sizes are passed, not recomputed).
struct CubeGrid {
  double x_min, x_max, y_min, y_max, z_min, z_max;
  int nx, ny, nz;
  double dx, dy, dz;
};

// get_cell_indices, examples/diff_flux_cube.jl:132-144 (1-based i, j, k)
inline bool cube_cell(const CubeGrid& g, double x, double y, double z, int& i, int& j, int& k) {
  if (x < g.x_min || x > g.x_max || y < g.y_min || y > g.y_max || z < g.z_min || z > g.z_max)
    return false;
  i = std::min(std::max((int)std::floor((x - g.x_min) / g.dx) + 1, 1), g.nx);
  j = std::min(std::max((int)std::floor((y - g.y_min) / g.dy) + 1, 1), g.ny);
  k = std::min(std::max((int)std::floor((z - g.z_min) / g.dz) + 1, 1), g.nz);
  return true;
}

// distance_to_boundary, examples/diff_flux_cube.jl:194-280.  Returns the
// distance; *exits is set when the next cell lies outside the cube.
inline double cube_boundary(const CubeGrid& g, double x, double y, double z,
                           double ux, double uy, double uz, bool* exits, bool* inside) {
  int i, j, k;
  if (!cube_cell(g, x, y, z, i, j, k)) { *inside = false; *exits = true; return 0.0; }
  *inside = true;
  double x_lo = g.x_min + (i - 1) * g.dx, x_hi = g.x_min + i * g.dx;
  double y_lo = g.y_min + (j - 1) * g.dy, y_hi = g.y_min + j * g.dy;
  double z_lo = g.z_min + (k - 1) * g.dz, z_hi = g.z_min + k * g.dz;
  double t_min = INFINITY;
  bool exiting = false;
  const double eps = 1e-10;
  if (std::fabs(ux) > eps) {
    if (ux > 0) { double t = (x_hi - x) / ux; if (t > eps && t < t_min) { t_min = t; exiting = (i + 1 > g.nx); } }
    else        { double t = (x_lo - x) / ux; if (t > eps && t < t_min) { t_min = t; exiting = (i - 1 < 1); } }
  }
  if (std::fabs(uy) > eps) {
    if (uy > 0) { double t = (y_hi - y) / uy; if (t > eps && t < t_min) { t_min = t; exiting = (j + 1 > g.ny); } }
    else        { double t = (y_lo - y) / uy; if (t > eps && t < t_min) { t_min = t; exiting = (j - 1 < 1); } }
  }
  if (std::fabs(uz) > eps) {
    if (uz > 0) { double t = (z_hi - z) / uz; if (t > eps && t < t_min) { t_min = t; exiting = (k + 1 > g.nz); } }
    else        { double t = (z_lo - z) / uz; if (t > eps && t < t_min) { t_min = t; exiting = (k - 1 < 1); } }
  }
  *exits = exiting || t_min == INFINITY;
  return t_min;
}

// trace_cube_path, examples/diff_flux_cube.jl:304-364.  The direction is
// supplied by the caller (computed once on the host from elevation/azimuth).
// Cells are returned 0-based.  Returns the z where the path left the cube.
inline double trace_cube(const CubeGrid& g, double x0, double y0, double ux, double uy, double uz,
                         std::vector<int>& cells, std::vector<double>& lens) {
  double px = x0, py = y0, pz = 0.0, z_exit = 0.0;
  for (int step = 0; step < 1000; ++step) {
    int i, j, k;
    if (!cube_cell(g, px, py, pz, i, j, k)) break;
    int cell = (i - 1) + (j - 1) * g.nx + (k - 1) * g.nx * g.ny;
    bool exits, inside;
    double dist = cube_boundary(g, px, py, pz, ux, uy, uz, &exits, &inside);
    if (dist <= 1e-10 || !std::isfinite(dist)) {
      px += ux * 0.01; py += uy * 0.01; pz += uz * 0.01;
      continue;
    }
    cells.push_back(cell);
    lens.push_back(dist);
    px += ux * dist; py += uy * dist; pz += uz * dist;
    z_exit = pz;
    if (exits) break;
  }
  return z_exit;
}

// Rectilinear voxel volume of examples/lvd_tomography.jl:181-187.
// searchsortedlast on the edge vector (1-based result, 0 when x < edges[1]).
inline int searchsortedlast(const double* e, int n, double x) {
  int lo = 0, hi = n;  // count of edges <= x
  while (lo < hi) { int mid = (lo + hi) >> 1; if (e[mid] <= x) lo = mid + 1; else hi = mid; }
  return lo;
}
struct EdgeGrid {
  const double* xe; int nxe;
  const double* ye; int nye;
  const double* ze; int nze;
};
// voxel_cell_indices, examples/lvd_tomography.jl:583-594
inline bool edge_cell(const EdgeGrid& g, double x, double y, double z, int& i, int& j, int& k) {
  if (x < g.xe[0] || x > g.xe[g.nxe - 1] || y < g.ye[0] || y > g.ye[g.nye - 1] ||
      z < g.ze[0] || z > g.ze[g.nze - 1])
    return false;
  i = std::min(std::max(searchsortedlast(g.xe, g.nxe, x), 1), g.nxe - 1);
  j = std::min(std::max(searchsortedlast(g.ye, g.nye, y), 1), g.nye - 1);
  k = std::min(std::max(searchsortedlast(g.ze, g.nze, z), 1), g.nze - 1);
  return true;
}
// distance_to_voxel_boundary, examples/lvd_tomography.jl:596-627
inline double edge_boundary(const EdgeGrid& g, double x, double y, double z,
                            double ux, double uy, double uz) {
  int i, j, k;
  if (!edge_cell(g, x, y, z, i, j, k)) return 0.0;
  double x_lo = g.xe[i - 1], x_hi = g.xe[i], y_lo = g.ye[j - 1], y_hi = g.ye[j];
  double z_lo = g.ze[k - 1], z_hi = g.ze[k];
  double t_min = INFINITY;
  const double e = 1e-10;
  if (std::fabs(ux) > e) { double t = ux > 0 ? (x_hi - x) / ux : (x_lo - x) / ux; if (t > e) t_min = std::min(t_min, t); }
  if (std::fabs(uy) > e) { double t = uy > 0 ? (y_hi - y) / uy : (y_lo - y) / uy; if (t > e) t_min = std::min(t_min, t); }
  if (std::fabs(uz) > e) { double t = uz > 0 ? (z_hi - z) / uz : (z_lo - z) / uz; if (t > e) t_min = std::min(t_min, t); }
  return std::isfinite(t_min) ? t_min : 0.0;
}
// trace_path(::VoxelVolume), examples/lvd_tomography.jl:738-779.  The ray
// origin is the detector (the reference puts it at the origin); the surface
// distance comes from the caller.  Returns the distance travelled in volume.
inline double trace_edges(const EdgeGrid& g, double ox, double oy, double oz,
                          double ux, double uy, double uz, double surface_distance,
                          std::vector<int>& cells, std::vector<double>& lens) {
  const double POINT_EPS = 1e-4;
  double x = ox + ux * POINT_EPS, y = oy + uy * POINT_EPS, z = oz + uz * POINT_EPS;
  double traveled = POINT_EPS;
  const int nx = g.nxe - 1, ny = g.nye - 1;
  while (traveled < surface_distance - POINT_EPS) {
    int i, j, k;
    if (!edge_cell(g, x, y, z, i, j, k)) break;
    int cell = (i - 1) + (j - 1) * nx + (k - 1) * nx * ny;
    double dist = edge_boundary(g, x, y, z, ux, uy, uz);
    if (!std::isfinite(dist) || dist <= 0.0) break;
    double remaining = surface_distance - traveled;
    double step = std::min(dist, remaining);
    if (step > 1e-8) { cells.push_back(cell); lens.push_back(step); traveled += step; }
    if (dist >= remaining - 1e-6) break;
    x += ux * (step + POINT_EPS);
    y += uy * (step + POINT_EPS);
    z += uz * (step + POINT_EPS);
  }
  return traveled;
}

// ===================================================== per-voxel-density CSDA
// compute_flux_from_segments, examples/diff_flux_cube.jl:374-468.  R is the
// density type (double, or Dual to get one directional derivative).
template <class R>
inline R cube_csda_flux(const Tables& T, int rock, int air, const int* cells, const double* lens,
                        int nseg, double cos_theta, double z_exit, double z_max,
                        const R* dens, double energy0, double charge, double palt) {
  R energy = R(energy0);
  R weight = R(1.0);
  for (int s = 0; s < nseg; ++s) {
    R rho = dens[cells[s]];
    double dist = lens[s];
    R gram = dist * rho;
    R Xi = prop_range(T, false, rock, energy);
    R Xf = Xi + gram;
    R ea = prop_kinetic(T, false, rock, Xf);
    if (val(ea) >= 1e12 || !std::isfinite(val(ea))) return R(0.0);
    R si = prop_stop(T, false, rock, energy);
    R sf = prop_stop(T, false, rock, ea);
    weight = weight * (sf / si);
    R Eavg = (energy + ea) / 2.0;
    weight = weight * decay_weight(T, dist, Eavg);
    energy = ea;
  }
  double cs = std::max(std::fabs(cos_theta), 0.01);
  if (z_exit < z_max - EPS64) {
    double rem = (z_max - z_exit) / cs;
    R rho1 = dens[0];
    R Xi = prop_range(T, false, rock, energy);
    R Xf = Xi + rem * rho1;
    R ea = prop_kinetic(T, false, rock, Xf);
    if (val(ea) >= 1e12 || !std::isfinite(val(ea))) return R(0.0);
    R si = prop_stop(T, false, rock, energy);
    R sf = prop_stop(T, false, rock, ea);
    weight = weight * (sf / si);
    weight = weight * decay_weight(T, rem, (energy + ea) / 2.0);
    energy = ea;
  }
  double Xair = air_grammage(z_max, palt, cos_theta);
  R Xi = prop_range(T, false, air, energy);
  R ef = prop_kinetic(T, false, air, Xi + Xair);
  if (val(ef) >= 1e12 || !std::isfinite(val(ef))) return R(0.0);
  R si = prop_stop(T, false, air, energy);
  R sf = prop_stop(T, false, air, ef);
  weight = weight * (sf / si);
  double path_air = (palt - z_max) / cs;
  weight = weight * decay_weight(T, path_air, (energy + ef) / 2.0);
  return weight * flux_gccly(cos_theta, ef, charge);
}

// ========================================== flat two-layer deterministic CSDA
// TwoLayerGeometry, src/Geometry.jl:58-78 (materials 0-based; porous < 0
// disables the porous layer).
struct Layered {
  double rock_thickness, rock_density;
  int rock, air, porous;
  double porous_density, porous_thickness, primary_altitude;
};

template <class R> struct CsdaState {
  R energy, weight, time, distance, grammage;
  R z;   // only the vertical position matters for the flat geometry
};

// transport_backward_step, src/Geometry.jl:424-457 (rng = nothing branch).
template <class R, class D>
inline void csda_step(const Tables& T, CsdaState<R>& s, int m, D density, R ds, double uz_motion) {
  R ki = s.energy;
  R gs = ds * density;
  R Xi = prop_range(T, false, m, ki);
  R kf = prop_kinetic(T, false, m, Xi + gs);
  R gi = 1.0 + ki / T.mass;
  R gc = 1.0 + kf / T.mass;
  R ga = (gi + gc) / 2.0;
  R b2 = 1.0 - 1.0 / (ga * ga);
  R beta = sqrt(val(b2) >= 1e-12 ? b2 : R(1e-12));
  R dt = ds / (beta * ga);
  // position -= ds * direction; direction_z = -uz_motion
  s.z = s.z - ds * (-uz_motion);
  R si = prop_stop(T, false, m, ki);
  R sc = prop_stop(T, false, m, kf);
  R wf = sc / si;
  R decay = exp(-dt / T.ctau);
  s.weight = s.weight * wf * decay;
  s.energy = kf;
  s.distance = s.distance + ds;
  s.grammage = s.grammage + gs;
  s.time = s.time + dt;
}

// _backward_geometric_step_limit, src/Geometry.jl:463-476
template <class R, class D>
inline R geometric_limit(const Tables& T, R energy, int m, D density, double uz, bool in_air,
                         bool use_scattering) {
  R Xi = prop_range(T, false, m, energy);
  R range_limit = 0.01 * Xi / vmax(density, 1e-6);
  R lim = range_limit;
  if (use_scattering) {
    R lb1 = property(T, m, T_TRANSPORT_PATH, energy);
    if (val(lb1) > 0 && val(density) > 0) lim = vmin(lim, 0.01 * lb1 / density);
  }
  if (in_air) lim = vmin(lim, R(0.05 * 12e3 / std::max(std::fabs(uz), 0.05)));
  return lim;
}

// transport_backward_through_geometry with rng = nothing,
// src/Geometry.jl:534-672: every step is the deterministic CSDA step.  The
// rock density is the differentiated parameter (R = Dual for the tangent).
template <class R>
inline CsdaState<R> flat_csda_walk(const Tables& T, const Layered& g, R rock_density,
                                   double energy0, double cos_theta, double threshold,
                                   double mode_switch, bool use_scattering_limit) {
  CsdaState<R> s{R(energy0), R(1.0), R(0.0), R(0.0), R(0.0), R(0.0)};
  const double palt = g.primary_altitude;
  const double uz = cos_theta;  // uz_motion = -direction_z = cos(theta)
  for (int outer = 0; outer < 64; ++outer) {
    if (!(val(s.energy) < threshold - EPS64)) break;
    double z = val(s.z);
    if (z < 0) break;
    if (z >= palt - EPS64) break;
    if (val(s.weight) <= 0 || !std::isfinite(val(s.weight))) break;
    bool below = val(s.energy) < mode_switch - EPS64;
    bool use_scat = below ? use_scattering_limit : false;
    double limit = below ? std::min(mode_switch, threshold) : threshold;
    for (int inner = 0; inner < 50000; ++inner) {
      z = val(s.z);
      if (z < 0 || z >= palt - EPS64) break;
      if (val(s.energy) >= limit - EPS64) break;
      if (val(s.weight) <= 0 || !std::isfinite(val(s.weight))) break;
      bool has_porous = g.porous >= 0 && g.porous_thickness > 0;
      double porous_z = has_porous ? g.rock_thickness - g.porous_thickness : g.rock_thickness;
      int m;
      R density;
      R stb;
      bool in_air = !(z < g.rock_thickness);
      if (!in_air) {
        if (has_porous && z >= porous_z) { m = g.porous; density = R(g.porous_density); }
        else { m = g.rock; density = rock_density; }
        if (uz > EPS64) {
          double nb = (has_porous && z < porous_z) ? porous_z : g.rock_thickness;
          stb = (nb - s.z) / uz;
        } else if (uz < -EPS64) {
          double pb = (has_porous && z >= porous_z) ? porous_z : 0.0;
          stb = (pb - s.z) / uz;
        } else stb = R(1000.0);
      } else {
        m = g.air;
        // locals_air_at_altitude, src/Geometry.jl:95-106, evaluated at the
        // (differentiable) position
        density = 1.205 * exp(-s.z / 12e3);
        if (uz > EPS64) stb = (palt - s.z) / uz;
        else if (uz < -EPS64) stb = (g.rock_thickness - s.z) / uz;
        else stb = R(1000.0);
      }
      R mg = geometric_limit(T, s.energy, m, density, uz, in_air, use_scat);
      R step = vmin(stb + 1e-7, mg);
      // clamp(step, STEP_MIN, 1000)
      if (val(step) > 1000.0) step = R(1000.0);
      else if (val(step) < STEP_MIN) step = R(STEP_MIN);
      csda_step(T, s, m, density, step, uz);
    }
  }
  return s;
}

// compute_flux_single, src/Geometry.jl:694-757 (deterministic CSDA walk).
template <class R>
inline R flat_csda_flux(const Tables& T, const Layered& g, R rho, double energy, double cos_t,
                        double charge, double threshold, double mode_switch, bool scat_limit) {
  CsdaState<R> s = flat_csda_walk(T, g, rho, energy, cos_t, threshold, mode_switch, scat_limit);
  if (val(s.z) >= g.primary_altitude - 1.0 && val(s.energy) < threshold) {
    double cf = cos_t;  // the CSDA walk never changes the direction
    cf = std::min(std::max(cf, 0.0), 1.0);
    if (cf <= 0) return R(0.0);
    return s.weight * flux_gccly(cf, s.energy, charge);
  }
  return R(0.0);
}


// compute_flux_csda_direct, src/Geometry.jl:1126-1209: one range inversion
// through the rock slab, one through the analytic air column, decay at the
// mean energy of each layer.  R is the rock-density type (double, or a dual
// for compute_flux_gradient_csda, :1247-1262).
template <class R>
inline R flux_csda_direct(const Tables& T, int rock, int air, R rock_density, double thickness,
                          double elevation, double energy, double charge, double primary_altitude) {
  const double el = elevation * (M_PI / 180.0);
  const double cos_t = std::cos(M_PI / 2 - el);
  const double cs = std::max(std::fabs(cos_t), 0.01);
  // compute_rock_grammage, :1064-1068
  R x_rock = (thickness / cs) * rock_density;
  const double x0 = prop_range(T, false, rock, energy);
  R x1 = x0 + x_rock;
  R e1 = prop_kinetic(T, false, rock, x1);
  if (!(val(e1) < 1e12) || !std::isfinite(val(e1))) return R(0.0);
  const double s0 = prop_stop(T, false, rock, energy);
  R s1 = prop_stop(T, false, rock, e1);
  R w_rock = s1 / s0;
  R d_rock = decay_weight(T, thickness / cs, (energy + e1) / 2.0);
  const double x_air = air_grammage(thickness, primary_altitude, cos_t);
  R xa0 = prop_range(T, false, air, e1);
  R xa1 = xa0 + x_air;
  R e2 = prop_kinetic(T, false, air, xa1);
  if (!(val(e2) < 1e12) || !std::isfinite(val(e2))) return R(0.0);
  R sa0 = prop_stop(T, false, air, e1);
  R sa1 = prop_stop(T, false, air, e2);
  R w_air = sa1 / sa0;
  R d_air = decay_weight(T, (primary_altitude - thickness) / cs, (e1 + e2) / 2.0);
  R w = w_rock * d_rock * w_air * d_air;
  return w * flux_gccly(cos_t, e2, charge);
}

// ================================================== topography: DEM ray exit
// ElevationMap (src/Turtle.jl:469-480): nx x ny nodes from (x0, y0) with
// spacings dx, dy; z(ix, iy) = data[iy + ny * ix] (a Julia Matrix{Float64}
// of size (ny, nx) as stored).
struct Dem {
  int nx, ny;
  double x0, y0, dx, dy;
  const double* data;
};
// map_elevation, src/Turtle.jl:523-553: bilinear interpolation; inside = false
// (elevation 0) off the map.
inline double dem_elevation(const Dem& m, double x, double y, bool& inside) {
  inside = false;
  if (std::isnan(x) || std::isnan(y)) return 0.0;
  double hx = (x - m.x0) / m.dx;
  double hy = (y - m.y0) / m.dy;
  if (hx > m.nx - 1 || hx < 0 || hy > m.ny - 1 || hy < 0) return 0.0;
  int ix = (int)std::floor(hx);
  int iy = (int)std::floor(hy);
  if (ix == m.nx - 1) { ix -= 1; hx = 1.0; } else { hx -= ix; }
  if (iy == m.ny - 1) { iy -= 1; hy = 1.0; } else { hy -= iy; }
  const double z00 = m.data[iy + (size_t)m.ny * ix];
  const double z10 = m.data[iy + (size_t)m.ny * (ix + 1)];
  const double z01 = m.data[iy + 1 + (size_t)m.ny * ix];
  const double z11 = m.data[iy + 1 + (size_t)m.ny * (ix + 1)];
  inside = true;
  return z00 * (1.0 - hx) * (1.0 - hy) + z01 * (1.0 - hx) * hy + z10 * hx * (1.0 - hy) +
         z11 * hx * hy;
}
// local_surface_height, examples/lvd_tomography.jl:214-217: map coordinates
// are the local metres times xy_scale (1e-3 for a map in km); heights are
// relative to the detector elevation.
inline double surface_height(const Dem& m, double xy_scale, double z_ref, double x, double y) {
  bool inside;
  const double e = dem_elevation(m, x * xy_scale, y * xy_scale, inside);
  return inside ? e - z_ref : NAN;
}
// ray_surface_exit_distance, examples/lvd_tomography.jl:219-252: march in
// steps of `step` up to max_distance, then 40 bisections.
inline void ray_surface_exit(const Dem& m, double xy_scale, double z_ref, const double* d,
                             double step, double max_distance, double& dist, bool& found,
                             double& exit_z) {
  dist = NAN; found = false; exit_z = NAN;
  if (d[2] <= 0.0) return;
  const long n = (long)std::floor(max_distance / step + 1e-9);
  double prev = 0.0;
  for (long k = 1; k <= n; ++k) {
    const double s = (double)k * step;
    const double x = d[0] * s, y = d[1] * s, z = d[2] * s;
    const double sz = surface_height(m, xy_scale, z_ref, x, y);
    if (!std::isfinite(sz)) return;
    if (z > sz) {
      double lo = prev, hi = s;
      for (int it = 0; it < 40; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double ms = surface_height(m, xy_scale, z_ref, d[0] * mid, d[1] * mid);
        if (d[2] * mid > ms) hi = mid; else lo = mid;
      }
      dist = 0.5 * (lo + hi);
      found = true;
      exit_z = d[2] * dist;
      return;
    }
    prev = s;
  }
}

// ============================================= w-parameterised tomography CSDA
// MaterialConfig (src/Tomography.jl:179-188), materials 0-based (porous < 0
// disables the shallow porous recipe) and SiteConfig (:198-201).
struct MatCfg {
  int rock, water, air, porous;
  double rho_rock, rho_water, rho_porous, porous_top;
};
struct Site { double detector_elevation, primary_altitude; };

// _csda_min / _csda_max, src/Tomography.jl:421-434: primal comparison with
// the left operand kept at ties.
template <class A, class B> inline A cmin(A a, B b) { return val(a) <= val(b) ? a : A(b); }
template <class A, class B> inline A cmax(A a, B b) { return val(a) >= val(b) ? a : A(b); }

inline bool shallow_porous(bool shallow, const MatCfg& m) {
  return shallow && m.porous >= 0 && m.porous_top > 0.0;
}
// cell_density, src/Tomography.jl:295-302
template <class R> inline R cell_density(R w, bool shallow, const MatCfg& m) {
  if (shallow_porous(shallow, m))
    return 1.0 / (0.5 * (1.0 - w) / m.rho_porous + 0.5 * (1.0 - w) / m.rho_rock + w / m.rho_water);
  return 1.0 / ((1.0 - w) / m.rho_rock + w / m.rho_water);
}
// cell_stopping_power, src/Tomography.jl:304-314
template <class R, class E> inline R cell_stop(const Tables& T, R w, bool shallow, const MatCfg& m, E e) {
  if (shallow_porous(shallow, m)) {
    R sp = prop_stop(T, false, m.porous, R(e)), sr = prop_stop(T, false, m.rock, R(e)),
      sw = prop_stop(T, false, m.water, R(e));
    return 0.5 * (1.0 - w) * sp + 0.5 * (1.0 - w) * sr + w * sw;
  }
  R sr = prop_stop(T, false, m.rock, R(e)), sw = prop_stop(T, false, m.water, R(e));
  return (1.0 - w) * sr + w * sw;
}

constexpr double CSDA_MAX_RELATIVE_GAIN = 0.02;  // examples/lvd_tomography.jl:155
constexpr double CSDA_MAX_STEP_M = 60.0;         // examples/lvd_tomography.jl:156

// propagate_csda_cell_segment, src/Tomography.jl:444-473 (Heun sub-steps)
template <class R>
inline void cell_segment(const Tables& T, R w, bool shallow, const MatCfg& m, double distance,
                         R& energy, R& weight, bool& ok) {
  R density = cell_density(w, shallow, m);
  R remaining = R(distance);
  ok = true;
  while (val(remaining) > 1e-8) {
    R dedx0 = cell_stop(T, w, shallow, m, energy);
    if (!std::isfinite(val(dedx0)) || val(dedx0) <= 0.0) { weight = R(0.0); ok = false; return; }
    R grel = CSDA_MAX_RELATIVE_GAIN * cmax(energy, 1e-6) / dedx0;
    R gstep = cmin(cmin(remaining * density, cmax(grel, 1e-6)), CSDA_MAX_STEP_M * density);
    R ds = cmin(remaining, gstep / cmax(density, 1e-12));
    if (val(ds) <= 0.0) { weight = R(0.0); ok = false; return; }
    R gram = ds * density;
    R epred = energy + dedx0 * gram;
    R dedx1 = cell_stop(T, w, shallow, m, epred);
    R dE = 0.5 * (dedx0 + dedx1) * gram;
    R enext = energy + dE;
    if (!std::isfinite(val(enext)) || val(enext) >= 1e12) { energy = enext; weight = R(0.0); ok = false; return; }
    R dnext = cell_stop(T, w, shallow, m, enext);
    weight = weight * (dnext / cmax(dedx0, 1e-30));
    weight = weight * decay_weight(T, ds, 0.5 * (energy + enext));
    energy = enext;
    remaining = remaining - ds;
  }
}

// propagate_csda_uniform_segment, src/Tomography.jl:481-509
template <class R>
inline void uniform_segment(const Tables& T, int mat, double density, double distance, R& energy,
                            R& weight, bool& ok) {
  R remaining = R(distance);
  ok = true;
  while (val(remaining) > 1e-8) {
    R dedx0 = prop_stop(T, false, mat, energy);
    if (!std::isfinite(val(dedx0)) || val(dedx0) <= 0.0) { weight = R(0.0); ok = false; return; }
    R grel = CSDA_MAX_RELATIVE_GAIN * cmax(energy, 1e-6) / dedx0;
    R gstep = cmin(cmin(remaining * density, cmax(grel, 1e-6)), R(CSDA_MAX_STEP_M * density));
    R ds = cmin(remaining, gstep / std::max(density, 1e-12));
    if (val(ds) <= 0.0) { weight = R(0.0); ok = false; return; }
    R gram = ds * density;
    R epred = energy + dedx0 * gram;
    R dedx1 = prop_stop(T, false, mat, epred);
    R enext = energy + 0.5 * (dedx0 + dedx1) * gram;
    if (!std::isfinite(val(enext)) || val(enext) >= 1e12) { energy = enext; weight = R(0.0); ok = false; return; }
    R dnext = prop_stop(T, false, mat, enext);
    weight = weight * (dnext / cmax(dedx0, 1e-30));
    weight = weight * decay_weight(T, ds, 0.5 * (energy + enext));
    energy = enext;
    remaining = remaining - ds;
  }
}

// propagate_csda_air_segment, src/Tomography.jl:516-551
template <class R>
inline void air_segment(const Tables& T, int air, const Site& site, double start_z, double cos_t,
                        double distance, R& energy, R& weight, bool& ok) {
  R remaining = R(distance);
  R traversed = R(0.0);
  ok = true;
  while (val(remaining) > 1e-8) {
    R dedx0 = prop_stop(T, false, air, energy);
    if (!std::isfinite(val(dedx0)) || val(dedx0) <= 0.0) { weight = R(0.0); ok = false; return; }
    R grel = CSDA_MAX_RELATIVE_GAIN * cmax(energy, 1e-6) / dedx0;
    // the first density only feeds the step choice through the same formula
    R alt0 = site.detector_elevation + start_z + cos_t * (traversed + 0.5 * cmin(remaining, CSDA_MAX_STEP_M));
    R dens0 = 1.205 * exp(-alt0 / 12e3);
    R ds = cmin(cmin(remaining, CSDA_MAX_STEP_M), grel / cmax(dens0, 1e-12));
    ds = cmax(ds, 1e-3);
    R alt = site.detector_elevation + start_z + cos_t * (traversed + 0.5 * ds);
    R density = 1.205 * exp(-alt / 12e3);
    R gram = ds * density;
    R epred = energy + dedx0 * gram;
    R dedx1 = prop_stop(T, false, air, epred);
    R enext = energy + 0.5 * (dedx0 + dedx1) * gram;
    if (!std::isfinite(val(enext)) || val(enext) >= 1e12) { energy = enext; weight = R(0.0); ok = false; return; }
    R dnext = prop_stop(T, false, air, enext);
    weight = weight * (dnext / cmax(dedx0, 1e-30));
    weight = weight * decay_weight(T, ds, 0.5 * (energy + enext));
    energy = enext;
    traversed = traversed + ds;
    remaining = remaining - ds;
  }
}

// ------------------------------------------- MC-calibrated CSDA corrections
// CsdaCorrection (src/Tomography.jl:112-137) and CsdaMaterialCorrection
// (:164-172); both are disabled by default (:139, :167).
struct Corr {
  bool enabled = false;
  double kappa_strag = 0, kappa_hard = 0;
  double resid_a = 0, resid_b = 0, resid_c = 0, resid_d = 0;
  bool mat_enabled = false;
  double linear = 0, quadratic = 0;
};

// property_straggling, src/Physics.jl:820-824
template <class R> inline R prop_strag(const Tables& T, int mat, R e) {
  return property(T, mat, T_STRAGGLING, e);
}
// cell_straggling, src/Tomography.jl:329-339
template <class R, class E>
inline R cell_strag(const Tables& T, R w, bool shallow, const MatCfg& m, E e) {
  if (shallow_porous(shallow, m)) {
    R sp = prop_strag(T, m.porous, R(e)), sr = prop_strag(T, m.rock, R(e)),
      sw = prop_strag(T, m.water, R(e));
    return 0.5 * (1.0 - w) * sp + 0.5 * (1.0 - w) * sr + w * sw;
  }
  R sr = prop_strag(T, m.rock, R(e)), sw = prop_strag(T, m.water, R(e));
  return (1.0 - w) * sr + w * sw;
}
// dedx_CSDA - dedx_MIXED of one material (rock_hardloss, :318-321)
template <class R> inline R hardloss(const Tables& T, int mat, R e) {
  return prop_stop(T, false, mat, e) - prop_stop(T, true, mat, e);
}
// cell_hardloss, src/Tomography.jl:341-352
template <class R, class E>
inline R cell_hard(const Tables& T, R w, bool shallow, const MatCfg& m, E e) {
  R hr = hardloss(T, m.rock, R(e)), hw = hardloss(T, m.water, R(e));
  if (shallow_porous(shallow, m)) {
    R hp = hardloss(T, m.porous, R(e));
    return 0.5 * (1.0 - w) * hp + 0.5 * (1.0 - w) * hr + w * hw;
  }
  return (1.0 - w) * hr + w * hw;
}
// local_spectral_index, src/Tomography.jl:367-372: symmetric log-difference
// of the surface flux at charge +1.
template <class R> inline R spectral_index(R E, double cos_t) {
  const double eta = 1.0e-2;
  R fp = flux_gccly(cos_t, E * (1.0 + eta), 1.0);
  R fm = flux_gccly(cos_t, E * (1.0 - eta), 1.0);
  return -(log(fp) - log(fm)) / (2 * eta);
}
// csda_var_correction, src/Tomography.jl:384-391
template <class R>
inline R var_correction(R energy_surface, R V, R H, double cos_t, const Corr& c) {
  if (!c.enabled) return R(1.0);
  R e = cmax(energy_surface, 1.0);
  R b = spectral_index(e, cos_t);
  R cs = 0.5 * b * (b + 1.0);
  return exp(c.kappa_strag * cs * V / (e * e) - c.kappa_hard * cmax(H, 0.0) / e);
}
// path_rock_length + csda_geom_multiplier, src/Tomography.jl:394-412
inline double geom_multiplier(const double* lens, int nseg, double remaining_rock, double cos_t,
                              const Corr& c) {
  if (!(c.enabled && (c.resid_a != 0.0 || c.resid_b != 0.0 || c.resid_c != 0.0 ||
                      c.resid_d != 0.0)))
    return 1.0;
  double L = remaining_rock;
  for (int s = 0; s < nseg; ++s) L += lens[s];
  const double lL = std::log(std::max(L, 1.0));
  return std::exp(c.resid_a + c.resid_b * lL + c.resid_c * lL * lL + c.resid_d * cos_t);
}
// _column_water_fraction, src/Tomography.jl:560-570
template <class R>
inline R column_water_fraction(const MatCfg& m, double remaining_rock, const double* lens,
                               const int* cells, int nseg, const R* w, const uint8_t* shallow) {
  R water = R(0.0);
  R mass = R(remaining_rock * m.rho_rock);
  for (int s = 0; s < nseg; ++s) {
    R wk = w[cells[s]];
    R gx = lens[s] * cell_density(wk, shallow[cells[s]] != 0, m);
    water = water + wk * gx;
    mass = mass + gx;
  }
  return water / cmax(mass, 1e-30);
}
// _material_multiplier, src/Tomography.jl:571-574
template <class R> inline R material_multiplier(R q, const Corr& c) {
  return c.mat_enabled ? exp(c.linear * q + c.quadratic * q * q) : R(1.0);
}

// Per energy node: the surface energy, the weight, the accumulated soft-loss
// variance proxy V and hard loss H of _directional_flux_core
// (src/Tomography.jl:585-637).
template <class R> struct NodeWalk { R energy, weight, V, H; bool ok; };

template <class R>
inline NodeWalk<R> walk_node_w(const Tables& T, const MatCfg& m, const Site& site, const int* cells,
                               const double* lens, int nseg, double remaining_rock, double exit_z,
                               double cos_t, const R* w, const uint8_t* shallow, double e0,
                               bool accumulate) {
  NodeWalk<R> o{R(e0), R(1.0), R(0.0), R(0.0), true};
  for (int s = 0; s < nseg && o.ok; ++s) {
    const bool sh = shallow[cells[s]] != 0;
    R e_in = o.energy;
    cell_segment(T, w[cells[s]], sh, m, lens[s], o.energy, o.weight, o.ok);
    if (!o.ok || !accumulate) continue;
    R e_rep = 0.5 * (e_in + o.energy);
    R wseg = w[cells[s]];
    R gx = lens[s] * cell_density(wseg, sh, m);
    o.V = o.V + cell_strag(T, wseg, sh, m, e_rep) * gx;
    o.H = o.H + cell_hard(T, wseg, sh, m, e_rep) * gx;
  }
  if (o.ok && remaining_rock > 1e-8) {
    R e_in = o.energy;
    uniform_segment(T, m.rock, m.rho_rock, remaining_rock, o.energy, o.weight, o.ok);
    if (o.ok && accumulate) {
      R e_rep = 0.5 * (e_in + o.energy);
      const double gx = remaining_rock * m.rho_rock;
      o.V = o.V + prop_strag(T, m.rock, e_rep) * gx;
      o.H = o.H + hardloss(T, m.rock, e_rep) * gx;
    }
  }
  if (o.ok) {
    const double air_d = std::max(0.0, site.primary_altitude - exit_z) / std::max(cos_t, 1e-3);
    air_segment(T, m.air, site, exit_z, cos_t, air_d, o.energy, o.weight, o.ok);
  }
  return o;
}

// compute_directional_flux_csda, src/Tomography.jl:667-726: sum over energy
// nodes of 2 w_n flux / N, times the geometric and material multipliers.
// R is the water-fraction type (double, or a dual seeded on cells).
template <class R>
inline R directional_flux_w(const Tables& T, const MatCfg& m, const Site& site, const int* cells,
                            const double* lens, int nseg, double remaining_rock, double exit_z,
                            double cos_t, const R* w, const uint8_t* shallow, int n_nodes,
                            const double* ne, const double* nw, const double* nq,
                            const Corr& corr = Corr{}) {
  R sum = R(0.0);
  for (int n = 0; n < n_nodes; ++n) {
    NodeWalk<R> o = walk_node_w(T, m, site, cells, lens, nseg, remaining_rock, exit_z, cos_t, w,
                                shallow, ne[n], corr.enabled);
    R fs = R(0.0);
    if (o.ok) {
      R c = var_correction(o.energy, o.V, o.H, cos_t, corr);
      fs = o.weight * c * flux_gccly(cos_t, o.energy, nq[n]);
    }
    sum = sum + 2.0 * nw[n] * fs;
  }
  R q = column_water_fraction(m, remaining_rock, lens, cells, nseg, w, shallow);
  return geom_multiplier(lens, nseg, remaining_rock, cos_t, corr) * material_multiplier(q, corr) *
         sum / (double)n_nodes;
}

}  // namespace orc

// CPU oracle: extern "C" entry points loaded by tests/ (ctypes) and by the
// cpu_baseline leg of bench.py.  TEST INFRASTRUCTURE ONLY.
#include "oracle_core.hpp"
#include "oracle_geometry.hpp"
#include "oracle_philox.hpp"
#include "oracle_flat_mc.hpp"
#include "oracle_inverse.hpp"
#include "oracle_tetra.hpp"
#include <cstring>
#include <map>
#include <omp.h>

using namespace orc;

static Tables make_tables(int M, int K, const double* data, const double* scal, double mass,
                          double ctau) {
  Tables T;
  T.M = M; T.K = K; T.data = data; T.scalars = scal; T.mass = mass; T.ctau = ctau;
  return T;
}

extern "C" {

void orc_philox_block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                      uint32_t k1, uint32_t* out) {
  uint32_t c[4] = {c0, c1, c2, c3};
  Philox4x32::block10(c, k0, k1);
  std::memcpy(out, c, sizeof(c));
}

void orc_philox_uniforms(uint64_t seed, uint64_t event, int n, double* out) {
  Philox4x32 r(seed, event);
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
}

double orc_flux_gccly(double c, double k, double q) { return flux_gccly(c, k, q); }

// ---------------------------------------------------------------- DCS models
void orc_dcs(int p, double Z, double A, double I, double m, double K, int n, const double* q,
             double* out) {
  for (int i = 0; i < n; ++i) out[i] = dcs::dcs(p, Z, A, I, m, K, q[i]);
}
double orc_dcs_integral(int p, double Z, double A, double I, double m, double K, double xlow,
                        double xhigh, int mode, int npoints) {
  return dcs::integral(p, Z, A, I, m, K, xlow, xhigh, mode, npoints);
}
void orc_dilog(int n, const double* x, double* out) {
  for (int i = 0; i < n; ++i) out[i] = dcs::dilog(x[i]);
}
void orc_gauss(int n, double* x, double* w) {
  const dcs::GaussRule g(n);
  for (int i = 0; i < n; ++i) { x[i] = g.x[i]; w[i] = g.w[i]; }
}

// ----------------------------------------------------------------- samplers
// rotate_direction with the azimuth draw given: n directions, polar mu[i],
// uniform u[i] for the azimuth.
void orc_rotate(int n, const double* d_in, const double* mu, const double* u, double* d_out) {
  for (int i = 0; i < n; ++i) {
    double d[3] = {d_in[3 * i], d_in[3 * i + 1], d_in[3 * i + 2]};
    Philox4x32 rng(0, 0);
    rng.script = u + i; rng.script_n = 1;
    rotate(d, mu[i], rng);
    for (int a = 0; a < 3; ++a) d_out[3 * i + a] = d[a];
  }
}
// truncated_randn from n seeded streams (seed, i)
void orc_truncated_randn(unsigned long long seed, int n, double* out) {
  for (int i = 0; i < n; ++i) {
    Philox4x32 rng(seed, (uint64_t)i);
    out[i] = truncated_randn(rng);
  }
}
double orc_flux_gaisser(double c, double k, double q) { return flux_gaisser(c, k, q); }
double orc_dflux_gccly_dk(double c, double k, double q) {
  return flux_gccly(c, Dual(k, 1.0), q).d;
}

// Property lookup on material m, array a, with its derivative d/dE.
void orc_property(int M, int K, const double* data, const double* scal, int m, int a, int n,
                  const double* e, double* out, double* dout) {
  Tables T = make_tables(M, K, data, scal, MUON_MASS, MUON_C_TAU);
  for (int i = 0; i < n; ++i) {
    Dual y = property(T, m, a, Dual(e[i], 1.0));
    out[i] = y.v;
    if (dout) dout[i] = y.d;
  }
}
void orc_kinetic(int M, int K, const double* data, const double* scal, int m, int mixed, int n,
                 const double* x, double* out, double* dout) {
  Tables T = make_tables(M, K, data, scal, MUON_MASS, MUON_C_TAU);
  for (int i = 0; i < n; ++i) {
    Dual y = prop_kinetic(T, mixed != 0, m, Dual(x[i], 1.0));
    out[i] = y.v;
    if (dout) dout[i] = y.d;
  }
}

// ---------------------------------------------------------------- tracing
// Pass 1 (cells == nullptr): writes counts[r].  Pass 2: fills the CSR arrays
// at row_ptr offsets.  z_exit may be null.
void orc_trace_cube(const double* grid /*12: xmin xmax ymin ymax zmin zmax nx ny nz dx dy dz*/,
                    int n_rays, const double* start_xy, const double* dirs, int32_t* counts,
                    const int32_t* row_ptr, int32_t* cells, double* lens, double* z_exit) {
  CubeGrid g{grid[0], grid[1], grid[2], grid[3], grid[4], grid[5],
             (int)grid[6], (int)grid[7], (int)grid[8], grid[9], grid[10], grid[11]};
#pragma omp parallel for schedule(dynamic, 16)
  for (int r = 0; r < n_rays; ++r) {
    std::vector<int> c;
    std::vector<double> l;
    double ze = trace_cube(g, start_xy[2 * r], start_xy[2 * r + 1], dirs[3 * r], dirs[3 * r + 1],
                           dirs[3 * r + 2], c, l);
    if (counts) counts[r] = (int32_t)c.size();
    if (z_exit) z_exit[r] = ze;
    if (cells) {
      int32_t o = row_ptr[r];
      for (size_t s = 0; s < c.size(); ++s) { cells[o + s] = c[s]; lens[o + s] = l[s]; }
    }
  }
}

void orc_trace_edges(const double* xe, int nxe, const double* ye, int nye, const double* ze,
                     int nze, int n_rays, const double* origins, const double* dirs,
                     const double* surface_distance, int32_t* counts, const int32_t* row_ptr,
                     int32_t* cells, double* lens, double* traveled) {
  EdgeGrid g{xe, nxe, ye, nye, ze, nze};
#pragma omp parallel for schedule(dynamic, 16)
  for (int r = 0; r < n_rays; ++r) {
    std::vector<int> c;
    std::vector<double> l;
    double tr = trace_edges(g, origins[3 * r], origins[3 * r + 1], origins[3 * r + 2],
                            dirs[3 * r], dirs[3 * r + 1], dirs[3 * r + 2],
                            surface_distance[r], c, l);
    if (counts) counts[r] = (int32_t)c.size();
    if (traveled) traveled[r] = tr;
    if (cells) {
      int32_t o = row_ptr[r];
      for (size_t s = 0; s < c.size(); ++s) { cells[o + s] = c[s]; lens[o + s] = l[s]; }
    }
  }
}

// ------------------------------------------------- per-voxel-density CSDA
// flux[r] = scale * sum_n 2 w_n flux(ray r, E_n, q_n)   (the summand of
// compute_flux_gradient_cube, examples/diff_flux_cube.jl:531-541).
// If grad_cells != null, the derivative of flux[r] with respect to each
// distinct cell on ray r (and cell 0 through the remaining-rock segment) is
// computed by forward-mode seeding, one cell at a time, and written as
// (grad_cells, grad_vals) triples at grad_ptr[r] (capacity = nseg + 1).
void orc_cube_csda(int M, int K, const double* data, const double* scal, double mass, double ctau,
                   int rock, int air, int n_rays, const int32_t* row_ptr, const int32_t* cells,
                   const double* lens, const double* cos_theta, const double* z_exit, double z_max,
                   double palt, int n_cells, const double* dens, int n_nodes, const double* ne,
                   const double* nw, const double* nq, double scale, double* flux,
                   const int32_t* grad_ptr, int32_t* grad_cells, double* grad_vals,
                   int32_t* grad_count, int threads) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    std::vector<Dual> ddens;
#pragma omp for schedule(dynamic, 4)
    for (int r = 0; r < n_rays; ++r) {
      const int b = row_ptr[r], ns = row_ptr[r + 1] - b;
      const int* c = cells + b;
      const double* l = lens + b;
      double acc = 0.0;
      for (int n = 0; n < n_nodes; ++n)
        acc += 2.0 * nw[n] * cube_csda_flux(T, rock, air, c, l, ns, cos_theta[r], z_exit[r], z_max,
                                           dens, ne[n], nq[n], palt);
      flux[r] = scale * acc;
      if (!grad_cells) continue;
      // distinct cells of the ray, plus cell 0 for the remaining rock
      std::map<int, int> uniq;
      for (int s = 0; s < ns; ++s) uniq.emplace(c[s], 0);
      if (z_exit[r] < z_max - EPS64) uniq.emplace(0, 0);
      if ((int)ddens.size() != n_cells) {
        ddens.assign(n_cells, Dual());
        for (int i = 0; i < n_cells; ++i) ddens[i] = Dual(dens[i], 0.0);
      }
      int o = grad_ptr[r], cnt = 0;
      for (auto& kv : uniq) {
        int cell = kv.first;
        ddens[cell].d = 1.0;
        double g = 0.0;
        for (int n = 0; n < n_nodes; ++n)
          g += 2.0 * nw[n] * cube_csda_flux(T, rock, air, c, l, ns, cos_theta[r], z_exit[r],
                                            z_max, ddens.data(), ne[n], nq[n], palt).d;
        ddens[cell].d = 0.0;
        grad_cells[o + cnt] = cell;
        grad_vals[o + cnt] = scale * g;
        ++cnt;
      }
      grad_count[r] = cnt;
    }
  }
}

// Central difference (F(rho + h e_v) - F(rho - h e_v)) / 2h of the summed
// flux F of the given rays (compute_flux_from_segments,
// examples/diff_flux_cube.jl:374-468, over the charge-summed nodes), with the
// whole program evaluated in long double (64-bit mantissa): its round-off
// over 2h is about 2000 times below that of the same difference in double.
// out[0] = the difference, out[1] = F(rho) (both rounded to double).
void orc_cube_fd_ld(int M, int K, const double* data, const double* scal, double mass,
                    double ctau, int rock, int air, int n_rays, const int* row_ptr,
                    const int* cells, const double* lens, const double* cos_theta,
                    const double* z_exit, double z_max, double palt, int n_cells,
                    const double* dens, int n_nodes, const double* ne, const double* nw,
                    const double* nq, double scale, int voxel, double h, int threads,
                    double* out) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  std::vector<long double> d(dens, dens + n_cells);
  if (threads > 0) omp_set_num_threads(threads);
  auto total = [&](long double rv) {
    d[voxel] = rv;
    std::vector<long double> per(n_rays, 0.0L);
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < n_rays; ++r) {
      const int b = row_ptr[r], ns = row_ptr[r + 1] - b;
      long double acc = 0.0L;
      for (int n = 0; n < n_nodes; ++n)
        acc += 2.0L * nw[n] * cube_csda_flux<long double>(T, rock, air, cells + b, lens + b, ns,
                                                          cos_theta[r], z_exit[r], z_max,
                                                          d.data(), ne[n], nq[n], palt);
      per[r] = (long double)scale * acc;
    }
    long double F = 0.0L;
    for (int r = 0; r < n_rays; ++r) F += per[r];
    return F;
  };
  const long double r0 = dens[voxel];
  const long double Fp = total(r0 + (long double)h);
  const long double Fm = total(r0 - (long double)h);
  const long double F0 = total(r0);
  out[0] = (double)((Fp - Fm) / (2.0L * (long double)h));
  out[1] = (double)F0;
}

// Reverse-mode form of the same derivative (used as the CPU baseline of the
// adjoint; O(segments) per ray·node).  Each CSDA segment is a map
// (E_in, rho) -> (E_out, log w) differentiated by two dual evaluations; the
// reverse sweep accumulates d log(flux)/dE backwards along the ray.
// jac[s] = d flux[r]/d rho through segment s; jac_rem[r] through the
// remaining rock (density of cell 0).
}  // extern "C"
namespace {
struct SegJac { double c, r, bE, bR; };  // dE/dE_in, dE/drho, dlogw/dE_in, dlogw/drho

template <class R>
inline void seg_map(const Tables& T, int m, R E, R rho, double dist, R& Eo, R& logw, bool& ok) {
  R Xi = prop_range(T, false, m, E);
  R ea = prop_kinetic(T, false, m, Xi + dist * rho);
  ok = !(val(ea) >= 1e12) && std::isfinite(val(ea));
  R si = prop_stop(T, false, m, E);
  R sf = prop_stop(T, false, m, ea);
  logw = log(sf / si) + log(decay_weight(T, dist, (E + ea) / 2.0));
  Eo = ea;
}
inline SegJac seg_jac(const Tables& T, int m, double E, double rho, double dist, double& Eo,
                      double& w, bool& ok) {
  Dual eo, lw;
  seg_map(T, m, Dual(E, 1.0), Dual(rho, 0.0), dist, eo, lw, ok);
  SegJac j;
  j.c = eo.d; j.bE = lw.d;
  seg_map(T, m, Dual(E, 0.0), Dual(rho, 1.0), dist, eo, lw, ok);
  j.r = eo.d; j.bR = lw.d;
  Eo = eo.v;
  w = std::exp(lw.v);
  return j;
}
}  // namespace
extern "C" {

void orc_cube_csda_reverse(int M, int K, const double* data, const double* scal, double mass,
                           double ctau, int rock, int air, int n_rays, const int32_t* row_ptr,
                           const int32_t* cells, const double* lens, const double* cos_theta,
                           const double* z_exit, double z_max, double palt, const double* dens,
                           int n_nodes, const double* ne, const double* nw, const double* nq,
                           double scale, double* flux, double* jac, double* jac_rem, int threads) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    std::vector<SegJac> tape;
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < n_rays; ++r) {
      const int b = row_ptr[r], ns = row_ptr[r + 1] - b;
      const double cs = std::max(std::fabs(cos_theta[r]), 0.01);
      const bool has_rem = z_exit[r] < z_max - EPS64;
      const double rem = (z_max - z_exit[r]) / cs;
      const double Xair = air_grammage(z_max, palt, cos_theta[r]);
      const double path_air = (palt - z_max) / cs;
      double facc = 0.0, racc = 0.0;
      for (int s = 0; s < ns; ++s) jac[b + s] = 0.0;
      tape.resize(ns);
      for (int n = 0; n < n_nodes; ++n) {
        double E = ne[n], W = 1.0, w;
        bool ok = true, okk;
        for (int s = 0; s < ns; ++s) {
          double Eo;
          tape[s] = seg_jac(T, rock, E, dens[cells[b + s]], lens[b + s], Eo, w, okk);
          ok = ok && okk; W *= w; E = Eo;
        }
        SegJac jr{0, 0, 0, 0};
        if (has_rem) {
          double Eo;
          jr = seg_jac(T, rock, E, dens[0], rem, Eo, w, okk);
          ok = ok && okk; W *= w; E = Eo;
        }
        // air + flux as one function of the energy entering the air
        Dual eo, lw;
        bool oka;
        Dual Xi = prop_range(T, false, air, Dual(E, 1.0));
        eo = prop_kinetic(T, false, air, Xi + Xair);
        oka = !(eo.v >= 1e12) && std::isfinite(eo.v);
        lw = log(prop_stop(T, false, air, eo) / prop_stop(T, false, air, Dual(E, 1.0))) +
             log(decay_weight(T, path_air, (Dual(E, 1.0) + eo) / 2.0)) +
             log(flux_gccly(cos_theta[r], eo, nq[n]));
        ok = ok && oka;
        const double f = ok ? 2.0 * nw[n] * W * std::exp(lw.v) : 0.0;
        facc += f;
        if (f == 0.0) continue;
        double Tad = lw.d;  // d log(tail)/dE at the exit of the last cell segment
        if (has_rem) { racc += f * (jr.bR + Tad * jr.r); Tad = jr.bE + Tad * jr.c; }
        for (int s = ns - 1; s >= 0; --s) {
          const SegJac& j = tape[s];
          jac[b + s] += f * (j.bR + Tad * j.r);
          Tad = j.bE + Tad * j.c;
        }
      }
      flux[r] = scale * facc;
      for (int s = 0; s < ns; ++s) jac[b + s] *= scale;
      if (jac_rem) jac_rem[r] = scale * racc;
    }
  }
}

// ------------------------------------------- flat two-layer CSDA (one event)
// Returns the flux of compute_flux_single (deterministic CSDA walk) and its
// density derivative (forward mode through the adaptive stepping).
void orc_flat_csda(int M, int K, const double* data, const double* scal, double mass, double ctau,
                   const double* geo /*8*/, const int32_t* mats /*3: rock air porous*/,
                   double energy, double cos_theta, double charge, double threshold,
                   double mode_switch, int scat_limit, double* out /*flux, dflux, E, z, w*/) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Layered g{geo[0], geo[1], mats[0], mats[1], mats[2], geo[2], geo[3], geo[4]};
  Dual f = flat_csda_flux(T, g, Dual(g.rock_density, 1.0), energy, cos_theta, charge, threshold,
                          mode_switch, scat_limit != 0);
  CsdaState<double> s = flat_csda_walk(T, g, g.rock_density, energy, cos_theta, threshold,
                                       mode_switch, scat_limit != 0);
  out[0] = f.v; out[1] = f.d; out[2] = s.energy; out[3] = s.z; out[4] = s.weight;
}

// _compute_flux_estimate with straggling = scattering = false
// (src/Geometry.jl:818-896) plus _sampled_csda_flux_gradient (:933-945).
// Event e draws its energy and charge uniforms from Philox(seed, e).
// out = flux, sigma, gradient_flux, gradient.
void orc_mc_flat_csda(int M, int K, const double* data, const double* scal, double mass,
                      double ctau, const double* geo, const int32_t* mats, double emin,
                      double emax, double cos_theta, double threshold, double mode_switch,
                      int scat_limit, long long n, unsigned long long seed, int threads,
                      int with_gradient, double* out) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Layered g{geo[0], geo[1], mats[0], mats[1], mats[2], geo[2], geo[3], geo[4]};
  // gradient surrogate geometry: TwoLayerGeometry(thickness, rho, rock, air)
  Layered gg{geo[0], geo[1], mats[0], mats[1], -1, 0.0, 0.0, PRIMARY_ALTITUDE};
  if (threads > 0) omp_set_num_threads(threads);
  double sw = 0, sw2 = 0, sgf = 0, sg = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : sw, sw2, sgf, sg)
  for (long long e = 0; e < n; ++e) {
    Philox4x32 rng(seed, (uint64_t)e);
    double k, w;
    sample_loguniform(emin, emax, rng.uniform(), k, w);
    double q = rng.uniform() > 0.5 ? 1.0 : -1.0;
    w *= 2.0;
    double f = flat_csda_flux(T, g, g.rock_density, k, cos_theta, q, threshold, mode_switch,
                              scat_limit != 0);
    double wi = w * f;
    sw += wi;
    sw2 += wi * wi;
    if (with_gradient) {
      Dual fd = flat_csda_flux(T, gg, Dual(g.rock_density, 1.0), k, cos_theta, q, threshold,
                               mode_switch, false);
      sgf += w * fd.v;
      sg += w * fd.d;
    }
  }
  double flux = sw / (double)n;
  double var = sw2 / (double)n - flux * flux;
  out[0] = flux;
  out[1] = g.rock_thickness > 0 ? std::sqrt(std::max(var, 0.0) / (double)n) : 0.0;
  out[2] = sgf / (double)n;
  out[3] = sg / (double)n;
}


// Stochastic flat backward MC (_compute_flux_estimate, src/Geometry.jl:818-896)
// with a runtime-mixture rock layer.  recipes: n_el[M] elements per material,
// then Z, A, I, f concatenated in material order.  Returns 0, or -2 on a
// DomainError, -3 on another transport error (first failing event in
// err_event).  w_events (optional) receives each event's contribution.
int orc_mc_flat(int M, int K, const double* data, const double* scal, double mass, double ctau,
                const int32_t* n_el, const double* eZ, const double* eA, const double* eI,
                const double* ef, int n_rock, const int32_t* rock_mats, const double* rock_fracs,
                double rock_rho, double thick, int air, int porous, double porous_rho,
                double porous_thick, double palt, double cos_t, double sin_t, double emin,
                double emax, double threshold, double mode_switch, int straggling, int scattering,
                double cutoff, double elastic_ratio, long long n, unsigned long long seed,
                int threads, double* w_events, double* out, long long* err_event) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  std::vector<std::vector<ElementRec>> rec(M);
  int o = 0;
  for (int m = 0; m < M; ++m)
    for (int i = 0; i < n_el[m]; ++i, ++o) rec[m].push_back({eZ[o], eA[o], eI[o], ef[o]});
  Mixture mrock = make_mixture(T, std::vector<int>(rock_mats, rock_mats + n_rock),
                               std::vector<double>(rock_fracs, rock_fracs + n_rock));
  Mixture mair = make_mixture(T, {air}, {1.0});
  Mixture mpor = make_mixture(T, {porous >= 0 ? porous : rock_mats[0]}, {1.0});
  MixedKernel krock = prepare_mixed_kernel(T, mrock, rec, cutoff);
  MixedKernel kair = prepare_mixed_kernel(T, mair, rec, cutoff);
  MixedKernel kpor = prepare_mixed_kernel(T, mpor, rec, cutoff);
  McSetup S;
  S.T = &T;
  S.rock = {&mrock, &krock};
  S.air = {&mair, &kair};
  S.porous = {&mpor, &kpor};
  S.has_porous = porous >= 0 && porous_thick > 0;
  S.thick = thick; S.rock_rho = rock_rho; S.porous_rho = porous_rho; S.porous_thick = porous_thick;
  S.palt = palt; S.cutoff = cutoff; S.elastic_ratio = elastic_ratio;
  if (threads > 0) omp_set_num_threads(threads);
  double sw = 0, sw2 = 0;
  int status = 0;
  long long first_bad = -1;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sw, sw2)
  for (long long e = 0; e < n; ++e) {
    Philox4x32 rng(seed, (uint64_t)e);
    double k, w;
    sample_loguniform(emin, emax, rng.uniform(), k, w);
    const double q = rng.uniform() > 0.5 ? 1.0 : -1.0;
    w *= 2.0;
    double f = 0.0;
    try {
      f = flat_event(S, k, cos_t, sin_t, q, threshold, mode_switch, straggling != 0,
                     scattering != 0, rng);
    } catch (const DomainErr&) {
#pragma omp critical
      { if (status == 0 || e < first_bad) { status = -2; first_bad = e; } }
    } catch (const std::exception&) {
#pragma omp critical
      { if (status == 0 || e < first_bad) { status = -3; first_bad = e; } }
    }
    const double wi = w * f;
    if (w_events) w_events[e] = wi;
    sw += wi;
    sw2 += wi * wi;
  }
  const double flux = sw / (double)n;
  const double var = sw2 / (double)n - flux * flux;
  out[0] = flux;
  out[1] = thick > 0 ? std::sqrt(std::max(var, 0.0) / (double)n) : 0.0;
  if (err_event) *err_event = first_bad;
  return status;
}

}  // extern "C"

// ======================================== w-parameterised tomography CSDA
namespace {
MatCfg make_matcfg(const int32_t* mi, const double* md) {
  return MatCfg{mi[0], mi[1], mi[2], mi[3], md[0], md[1], md[2], md[3]};
}
// Tape entry of one cell segment: energy partials and log-weight partials.
struct WSeg { double ew, ee, lw, le, vw, ve, hw, he; };
// corr[10] = enabled, kappa_strag, kappa_hard, resid_a..d, material enabled,
// linear, quadratic; NULL = both corrections off.
Corr make_corr(const double* c) {
  Corr k;
  if (!c) return k;
  k.enabled = c[0] != 0.0;
  k.kappa_strag = c[1]; k.kappa_hard = c[2];
  k.resid_a = c[3]; k.resid_b = c[4]; k.resid_c = c[5]; k.resid_d = c[6];
  k.mat_enabled = c[7] != 0.0;
  k.linear = c[8]; k.quadratic = c[9];
  return k;
}
}  // namespace
extern "C" {

// ray_surface_exit_distance (examples/lvd_tomography.jl:219-252) for n
// directions over the map (map_elevation, src/Turtle.jl:523-553).
void orc_ray_surface_exit(int nx, int ny, double x0, double y0, double dx, double dy,
                          const double* data, double xy_scale, double z_ref, int n,
                          const double* dirs, double step, double max_distance, double* dist,
                          uint8_t* found, double* exit_z) {
  Dem m{nx, ny, x0, y0, dx, dy, data};
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    bool f;
    ray_surface_exit(m, xy_scale, z_ref, dirs + 3 * i, step, max_distance, dist[i], f, exit_z[i]);
    found[i] = f;
  }
}
// Gauss-Newton normal equations (src/Tomography_inverse.jl:489-494).
void orc_normal_equations(int n_rows, int n_cols, const int* rp, const int* col, const double* val,
                          const double* weights, const double* residual, int n_act,
                          const int* active, double* normal, double* grad) {
  normal_equations(n_rows, n_cols, rp, col, val, weights, residual, n_act, active, normal, grad);
}
// build_tetra_neighbors (examples/lvd_tomography.jl:431-451), 0-based, -1 = none.
void orc_tet_neighbors(int nt, const int* tets, int* nbr) { build_neighbors(nt, tets, nbr); }
// point_in_tetrahedron for every (tet, point) pair: inside[t * n + i].
void orc_point_in_tet(int np, int nt, const double* pts, const int* tets, int n, const double* p,
                      uint8_t* inside) {
  TetMesh m{np, nt, pts, tets, nullptr};
  for (int t = 0; t < nt; ++t)
    for (int i = 0; i < n; ++i) inside[(size_t)t * n + i] = point_in_tet(m, t, p + 3 * i);
}
// trace_path(::TetraVolume) for n rays.  counts/traveled/start always; with
// cells != NULL the segments are written at row_ptr.
void orc_trace_tetra(int np, int nt, const double* pts, const int* tets, const int* nbr, int n,
                     const double* origin, const double* dir, const double* surface, int* counts,
                     const int* row_ptr, int* cells, double* lens, double* traveled, int* start) {
  TetMesh m{np, nt, pts, tets, nbr};
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < n; ++r) {
    const double* o = origin + 3 * r;
    const double* d = dir + 3 * r;
    const double p0[3] = {o[0] + d[0] * 1e-4, o[1] + d[1] * 1e-4, o[2] + d[2] * 1e-4};
    const int s0 = find_tet(m, p0);
    std::vector<int> c;
    std::vector<double> l;
    const double tr = trace_tet(m, o, d, surface[r], s0, c, l);
    counts[r] = (int)c.size();
    traveled[r] = tr;
    start[r] = s0;
    if (cells)
      for (size_t k = 0; k < c.size(); ++k) {
        cells[row_ptr[r] + k] = c[k];
        lens[row_ptr[r] + k] = l[k];
      }
  }
}
// map_elevation at n points: out = elevation, inside.
void orc_dem_elevation(int nx, int ny, double x0, double y0, double dx, double dy,
                       const double* data, int n, const double* xy, double* z, uint8_t* inside) {
  Dem m{nx, ny, x0, y0, dx, dy, data};
  for (int i = 0; i < n; ++i) {
    bool in;
    z[i] = dem_elevation(m, xy[2 * i], xy[2 * i + 1], in);
    inside[i] = in;
  }
}

// compute_flux_csda_direct / compute_flux_gradient_csda (src/Geometry.jl:
// 1126-1262) for n samples: flux[i] and d flux / d rock_density.
void orc_csda_direct(int M, int K, const double* data, const double* scal, double mass,
                     double ctau, int rock, int air, double primary_altitude, int n,
                     const double* rho, const double* thick, const double* elev,
                     const double* energy, const double* charge, double* flux, double* dflux) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    Dual f = flux_csda_direct(T, rock, air, Dual(rho[i], 1.0), thick[i], elev[i], energy[i],
                              charge[i], primary_altitude);
    flux[i] = f.v;
    if (dflux) dflux[i] = f.d;
  }
}

// One cell segment with the (w, E_in) dual walk of _cell_segment_partials
// (src/Tomography_rules.jl:33-48).  out = E, G, ok, dE/dw, dE/dE_in, dG/dw,
// dG/dE_in, with G the weight factor (unit incoming weight).
void orc_w_cell_segment(int M, int K, const double* data, const double* scal, double mass,
                        double ctau, const int32_t* mi, const double* md, double w, int shallow,
                        double distance, double energy, double* out) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  MatCfg m = make_matcfg(mi, md);
  Dual2 e(energy, 0.0, 1.0), g(1.0);
  bool ok;
  cell_segment(T, Dual2(w, 1.0, 0.0), shallow != 0, m, distance, e, g, ok);
  out[0] = e.v; out[1] = g.v; out[2] = ok; out[3] = e.a; out[4] = e.b; out[5] = g.a; out[6] = g.b;
}

// One atmosphere segment (propagate_csda_air_segment) with its energy tangent.
// out = E, G, ok, dE/dE_in, dG/dE_in.
void orc_w_air_segment(int M, int K, const double* data, const double* scal, double mass,
                       double ctau, int air, const double* site2, double start_z, double cos_t,
                       double distance, double energy, double* out) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Site site{site2[0], site2[1]};
  Dual e(energy, 1.0), g(1.0);
  bool ok;
  air_segment(T, air, site, start_z, cos_t, distance, e, g, ok);
  out[0] = e.v; out[1] = g.v; out[2] = ok; out[3] = e.d; out[4] = g.d;
}

// compute_directional_flux_csda (src/Tomography.jl:667-726) for many paths.
// Invalid paths and cos_t <= 0 give NaN.
void orc_w_flux(int M, int K, const double* data, const double* scal, double mass, double ctau,
                const int32_t* mi, const double* md, const double* site2, int n_rays,
                const int32_t* row_ptr, const int32_t* cells, const double* lens,
                const double* remaining, const double* exit_z, const double* cos_t,
                const uint8_t* valid, const double* w, const uint8_t* shallow, int n_nodes,
                const double* ne, const double* nw, const double* nq, const double* corr10,
                double* flux, int threads) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  MatCfg m = make_matcfg(mi, md);
  Site site{site2[0], site2[1]};
  const Corr corr = make_corr(corr10);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    std::vector<double> wl;
    std::vector<uint8_t> sl;
    std::vector<int> idx;
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < n_rays; ++r) {
      if (!valid[r] || !(cos_t[r] > 0.0)) { flux[r] = NAN; continue; }
      const int b = row_ptr[r], ns = row_ptr[r + 1] - b;
      wl.resize(ns); sl.resize(ns); idx.resize(ns);
      for (int s = 0; s < ns; ++s) { wl[s] = w[cells[b + s]]; sl[s] = shallow[cells[b + s]]; idx[s] = s; }
      flux[r] = directional_flux_w<double>(T, m, site, idx.data(), lens + b, ns, remaining[r],
                                           exit_z[r], cos_t[r], wl.data(), sl.data(), n_nodes, ne,
                                           nw, nq, corr);
    }
  }
}

// Forward-mode derivative of one path's flux with respect to w[cell] (the
// full-vector AD reference of test/test_tomography.jl:89).
double orc_w_dflux(int M, int K, const double* data, const double* scal, double mass, double ctau,
                   const int32_t* mi, const double* md, const double* site2, int ns,
                   const int32_t* cells, const double* lens, double remaining, double exit_z,
                   double cos_t, int n_cells, const double* w, const uint8_t* shallow, int cell,
                   int n_nodes, const double* ne, const double* nw, const double* nq,
                   const double* corr10) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  MatCfg m = make_matcfg(mi, md);
  Site site{site2[0], site2[1]};
  std::vector<Dual> wd(n_cells);
  for (int c = 0; c < n_cells; ++c) wd[c] = Dual(w[c], c == cell ? 1.0 : 0.0);
  Dual f = directional_flux_w<Dual>(T, m, site, cells, lens, ns, remaining, exit_z, cos_t,
                                    wd.data(), shallow, n_nodes, ne, nw, nq, make_corr(corr10));
  return f.d;
}

// directional_flux_and_grad_csda / assemble_forward_and_jacobian
// (src/Tomography.jl:736-820) through the segment pullbacks of
// src/Tomography_rules.jl:63-106.  jac has one slot per path segment: the
// derivative with respect to that segment's cell sits at the cell's first
// occurrence on the path, later occurrences hold 0.
//
// With the corrections on, log flux_n gains log corr(E_s, V, H): per segment
// the tape also holds dv/dw, dv/dE_in, dh/dw, dh/dE_in of the segment's V and H
// terms (total derivatives through E_out), and the reverse sweep adds
// A_V dv + A_H dh with A_V = d log corr / dV, A_H = d log corr / dH.  The
// material multiplier adds M'(q) dq/dw_s times the node sum.
void orc_w_jacobian(int M, int K, const double* data, const double* scal, double mass,
                    double ctau, const int32_t* mi, const double* md, const double* site2,
                    int n_rays, const int32_t* row_ptr, const int32_t* cells, const double* lens,
                    const double* remaining, const double* exit_z, const double* cos_t,
                    const uint8_t* valid, const double* w, const uint8_t* shallow, int n_nodes,
                    const double* ne, const double* nw, const double* nq, const double* corr10,
                    double* flux, double* jac, int threads) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  MatCfg m = make_matcfg(mi, md);
  Site site{site2[0], site2[1]};
  const Corr corr = make_corr(corr10);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    std::vector<WSeg> tape;
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < n_rays; ++r) {
      const int b = row_ptr[r], ns = row_ptr[r + 1] - b;
      for (int s = 0; s < ns; ++s) jac[b + s] = 0.0;
      if (!valid[r] || !(cos_t[r] > 0.0)) { flux[r] = NAN; continue; }
      const double ct = cos_t[r];
      const double rem = remaining[r];
      tape.resize(ns);
      double facc = 0.0;
      for (int n = 0; n < n_nodes; ++n) {
        double E = ne[n], W = 1.0, V = 0.0, H = 0.0;
        bool ok = true;
        for (int s = 0; s < ns && ok; ++s) {
          const int c = cells[b + s];
          const bool sh = shallow[c] != 0;
          const Dual2 wd(w[c], 1.0, 0.0), e_in(E, 0.0, 1.0);
          Dual2 e = e_in, g(W);
          cell_segment(T, wd, sh, m, lens[b + s], e, g, ok);
          WSeg t{e.a, e.b, g.a / g.v, g.b / g.v, 0.0, 0.0, 0.0, 0.0};
          if (ok && corr.enabled) {
            Dual2 e_rep = 0.5 * (e_in + e);
            Dual2 gx = lens[b + s] * cell_density(wd, sh, m);
            Dual2 v = cell_strag(T, wd, sh, m, e_rep) * gx;
            Dual2 h = cell_hard(T, wd, sh, m, e_rep) * gx;
            V += v.v; H += h.v;
            t.vw = v.a; t.ve = v.b; t.hw = h.a; t.he = h.b;
          }
          tape[s] = t;
          E = e.v; W = g.v;
        }
        double rc = 1.0, rl = 0.0, rv = 0.0, rh = 0.0;  // remaining-rock partials
        if (ok && rem > 1e-8) {
          const Dual e_in(E, 1.0);
          Dual e = e_in, g(W);
          uniform_segment(T, m.rock, m.rho_rock, rem, e, g, ok);
          rc = e.d; rl = g.d / g.v;
          if (ok && corr.enabled) {
            Dual e_rep = 0.5 * (e_in + e);
            const double gx = rem * m.rho_rock;
            Dual v = prop_strag(T, m.rock, e_rep) * gx;
            Dual h = hardloss(T, m.rock, e_rep) * gx;
            V += v.v; H += h.v; rv = v.d; rh = h.d;
          }
          E = e.v; W = g.v;
        }
        double ac = 1.0, al = 0.0;  // atmosphere partials
        if (ok) {
          const double air_d = std::max(0.0, site.primary_altitude - exit_z[r]) / std::max(ct, 1e-3);
          Dual e(E, 1.0), g(W);
          air_segment(T, m.air, site, exit_z[r], ct, air_d, e, g, ok);
          ac = e.d; al = g.d / g.v; E = e.v; W = g.v;
        }
        if (!ok) continue;
        Dual fl = flux_gccly(ct, Dual(E, 1.0), nq[n]);
        double cv = 1.0, AE = 0.0, AV = 0.0, AH = 0.0;
        if (corr.enabled) {
          Dual cd = var_correction(Dual(E, 1.0), Dual(V), Dual(H), ct, corr);
          cv = cd.v;
          AE = cd.d / cd.v;
          const double e = std::max(E, 1.0);
          const double bi = spectral_index(e, ct);
          AV = corr.kappa_strag * (0.5 * bi * (bi + 1.0)) / (e * e);
          AH = H >= 0.0 ? -corr.kappa_hard / e : 0.0;
        }
        const double f = 2.0 * nw[n] * (W * cv * fl.v);
        facc += f;
        if (f == 0.0) continue;
        double Tad = fl.d / fl.v + AE;   // d log(flux_n)/dE at the surface
        Tad = al + Tad * ac;             // ... at the rock exit
        if (rem > 1e-8) Tad = rl + Tad * rc + AV * rv + AH * rh;
        for (int s = ns - 1; s >= 0; --s) {
          const WSeg& t = tape[s];
          jac[b + s] += f * (t.lw + Tad * t.ew + AV * t.vw + AH * t.hw);
          Tad = t.le + Tad * t.ee + AV * t.ve + AH * t.he;
        }
      }
      const double G = geom_multiplier(lens + b, ns, rem, ct, corr);
      std::vector<int> idx(ns);
      for (int s = 0; s < ns; ++s) idx[s] = s;
      std::vector<double> wl(ns);
      std::vector<uint8_t> sl(ns);
      for (int s = 0; s < ns; ++s) { wl[s] = w[cells[b + s]]; sl[s] = shallow[cells[b + s]]; }
      const double q = column_water_fraction(m, rem, lens + b, idx.data(), ns, wl.data(), sl.data());
      const double Mq = material_multiplier(q, corr);
      flux[r] = G * Mq * facc / n_nodes;
      // d q / d w_s (the column mass fraction is linear in each cell's mass)
      double cmass = rem * m.rho_rock;
      for (int s = 0; s < ns; ++s) cmass += lens[b + s] * cell_density(wl[s], sl[s] != 0, m);
      const double dM = corr.mat_enabled ? Mq * (corr.linear + 2.0 * corr.quadratic * q) : 0.0;
      for (int s = 0; s < ns; ++s) {
        double v = G * Mq * (jac[b + s] / n_nodes);
        if (dM != 0.0) {
          const Dual gx = lens[b + s] * cell_density(Dual(wl[s], 1.0), sl[s] != 0, m);
          const double dwater = gx.v + wl[s] * gx.d;
          const double dq = cmass >= 1e-30 ? (dwater - q * gx.d) / cmass : dwater / 1e-30;
          v += G * dM * dq * (facc / n_nodes);
        }
        jac[b + s] = v;
      }
      // merge repeated cells into their first occurrence (_path_layout)
      for (int s = 1; s < ns; ++s)
        for (int t = 0; t < s; ++t)
          if (cells[b + t] == cells[b + s]) { jac[b + t] += jac[b + s]; jac[b + s] = 0.0; break; }
    }
  }
}

// One energy node on one path (walk of _directional_flux_core,
// src/Tomography.jl:585-637): out = surface energy, weight, V, H, ok.
void orc_w_node_walk(int M, int K, const double* data, const double* scal, double mass,
                     double ctau, const int32_t* mi, const double* md, const double* site2,
                     int ns, const int32_t* cells, const double* lens, double remaining,
                     double exit_z, double cos_t, const double* w, const uint8_t* shallow,
                     double energy, double* out) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  MatCfg m = make_matcfg(mi, md);
  Site site{site2[0], site2[1]};
  NodeWalk<double> o = walk_node_w<double>(T, m, site, cells, lens, ns, remaining, exit_z, cos_t,
                                           w, shallow, energy, true);
  out[0] = o.energy; out[1] = o.weight; out[2] = o.V; out[3] = o.H; out[4] = o.ok;
}

// local_spectral_index (src/Tomography.jl:367-372) and csda_var_correction (:384-391).
double orc_spectral_index(double energy, double cos_t) { return spectral_index(energy, cos_t); }
double orc_var_correction(double energy, double V, double H, double cos_t, const double* corr10) {
  return var_correction(energy, V, H, cos_t, make_corr(corr10));
}

}  // extern "C"

// ============================== mixed kernel / EHS with scripted draws (pins)
namespace {
MixedKernel kernel_from(int n, const double* e, const double* stop, const double* sigma,
                        const double* range, const double* hazard, const double* proper, int n_el,
                        const double* Z, const double* A, const double* I, const double* f,
                        double onset, double cutoff, double mass, double ctau) {
  MixedKernel c;
  c.mass = mass; c.ctau = ctau; c.cutoff = cutoff; c.onset = onset;
  c.e.assign(e, e + n); c.stop.assign(stop, stop + n); c.sigma.assign(sigma, sigma + n);
  c.range.assign(range, range + n); c.hazard.assign(hazard, hazard + n);
  c.proper.assign(proper, proper + n);
  for (int i = 0; i < n_el; ++i) c.el.push_back({Z[i], A[i], I[i], f[i]});
  return c;
}
Mixture mixture_from(const Tables& T, int n_mix, const int32_t* mats, const double* fracs) {
  return make_mixture(T, std::vector<int>(mats, mats + n_mix),
                      std::vector<double>(fracs, fracs + n_mix));
}
}  // namespace
extern "C" {

// transport_mixed_step (src/MixedTransport.jl:144-200) on an explicit kernel.
// st[0..10] = E, w, t, pos[3], dir[3], grammage, distance (in and out).
// Returns 0, -2 (DomainError), -3 (other error); *used = draws consumed.
int orc_mixed_step(int n, const double* e, const double* stop, const double* sigma,
                   const double* range, const double* hazard, const double* proper, int n_el,
                   const double* Z, const double* A, const double* I, const double* f,
                   double onset, double cutoff, double mass, double ctau, double* st,
                   double density, double distance, double energy_limit, const double* script,
                   int n_script, uint64_t seed, int* used) {
  MixedKernel c = kernel_from(n, e, stop, sigma, range, hazard, proper, n_el, Z, A, I, f, onset,
                              cutoff, mass, ctau);
  Medium md{nullptr, &c};
  McSetup S{};
  McState s;
  s.E = st[0]; s.w = st[1]; s.t = st[2];
  for (int a = 0; a < 3; ++a) { s.pos[a] = st[3 + a]; s.dir[a] = st[6 + a]; }
  s.X = st[9]; s.L = st[10];
  Philox4x32 rng(seed, 0);
  if (script) { rng.script = script; rng.script_n = n_script; }
  int rc = 0;
  try {
    mixed_step(S, md, s, density, distance, energy_limit, rng);
  } catch (const DomainErr&) {
    rc = -2;
  } catch (const std::exception&) {
    rc = -3;
  }
  st[0] = s.E; st[1] = s.w; st[2] = s.t;
  for (int a = 0; a < 3; ++a) st[3 + a] = s.pos[a];
  st[9] = s.X; st[10] = s.L;
  if (used) *used = rng.script_pos;
  return rc;
}

// reverse_del_event (src/MixedTransport.jl:111-142), details = false.
int orc_reverse_del(int n, const double* e, const double* stop, const double* sigma,
                    const double* range, const double* hazard, const double* proper, int n_el,
                    const double* Z, const double* A, const double* I, const double* f,
                    double onset, double cutoff, double mass, double ctau, double k,
                    const double* script, int n_script, double* out) {
  MixedKernel c = kernel_from(n, e, stop, sigma, range, hazard, proper, n_el, Z, A, I, f, onset,
                              cutoff, mass, ctau);
  Philox4x32 rng(0, 0);
  rng.script = script; rng.script_n = n_script;
  try {
    DelOut o = reverse_del_event(c, k, rng, false);
    out[0] = o.k; out[1] = o.w;
  } catch (const std::exception&) {
    return -3;
  }
  return rng.script_pos;
}

// sample_ehs_event, backward (src/Straggling.jl:373-434, :1284-1334).
// out = accepted, X_ehs, k_ehs.  Returns the number of draws consumed.
// fluctuate_energy_loss (backward) on a material or mixture: n seeded streams
// (seed, i), or, with n_script > 0, one call on the scripted draws (the
// return value is then the number of draws consumed).
int orc_fluctuate(int M, int K, const double* data, const double* scal, double mass, double ctau,
                  int n_mix, const int32_t* mats, const double* fracs, double ki, double dX,
                  double Xtot, unsigned long long seed, int n, const double* script, int n_script,
                  double* kout, double* ratio) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Mixture m = mixture_from(T, n_mix, mats, fracs);
  if (n_script > 0) {
    Philox4x32 rng(0, 0);
    rng.script = script; rng.script_n = n_script;
    fluctuate(T, m, ki, dX, Xtot, rng, kout[0], ratio[0]);
    return rng.script_pos;
  }
  for (int i = 0; i < n; ++i) {
    Philox4x32 rng(seed, (uint64_t)i);
    fluctuate(T, m, ki, dX, Xtot, rng, kout[i], ratio[i]);
  }
  return 0;
}

// sample_soft_scattering on n seeded streams (seed, i).
void orc_soft_scattering(int M, int K, const double* data, const double* scal, double mass,
                         double ctau, int n_mix, const int32_t* mats, const double* fracs,
                         double k0, double k1, double step, double density,
                         unsigned long long seed, int n, double* out) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Mixture m = mixture_from(T, n_mix, mats, fracs);
  for (int i = 0; i < n; ++i) {
    Philox4x32 rng(seed, (uint64_t)i);
    out[i] = soft_scattering(T, m, k0, k1, step, density, rng);
  }
}

// sample_scattering_angle at energy k: mode 0 = n lab-frame samples from
// seeded streams; mode 1 = the centre-of-mass draws alone (n samples); mode
// 2 = the lab transform of the n centre-of-mass values in mu1 (no draws).
// kin[5] receives (mu0, A, fspin, gCM, tau).
void orc_scattering_angle(int M, int K, const double* data, const double* scal, double mass,
                          double ctau, int n_mix, const int32_t* mats, const double* fracs,
                          double k, int mode, unsigned long long seed, int n, const double* mu1,
                          double* out, double* kin) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Mixture m = mixture_from(T, n_mix, mats, fracs);
  const WentzelKin w = wentzel_kinematics(T, m, k);
  if (kin) { kin[0] = w.mu0; kin[1] = w.A; kin[2] = w.fspin; kin[3] = w.gCM; kin[4] = w.tau; }
  for (int i = 0; i < n; ++i) {
    Philox4x32 rng(seed, (uint64_t)i);
    if (mode == 0) out[i] = scattering_angle(T, m, k, rng);
    else if (mode == 1) out[i] = wentzel_cm(w, rng);
    else out[i] = wentzel_lab(w, mu1[i]);
  }
}

// DEL polar angles (src/Straggling.jl:746-914): proc 1/2 bremsstrahlung and
// pair production (target charge Z), 3 photonuclear, 4 ionisation; ki is the
// energy before the event, kf after it.  n seeded streams, or one scripted
// call (returns the draws consumed).
int orc_polar(int proc, double m, double ki, double kf, double Z, unsigned long long seed, int n,
              const double* script, int n_script, double* out) {
  auto one = [&](Philox4x32& rng) {
    if (proc == 1 || proc == 2) return polar_brems(m, ki, kf, Z, rng);
    if (proc == 3) return polar_photo(m, ki, kf, rng);
    return polar_ionisation(m, ki, kf);
  };
  if (n_script > 0) {
    Philox4x32 rng(0, 0);
    rng.script = script; rng.script_n = n_script;
    out[0] = one(rng);
    return rng.script_pos;
  }
  for (int i = 0; i < n; ++i) {
    Philox4x32 rng(seed, (uint64_t)i);
    out[i] = one(rng);
  }
  return 0;
}

int orc_ehs_sample(int M, int K, const double* data, const double* scal, double mass, double ctau,
                   int n_mix, const int32_t* mats, const double* fracs, double ki, double kf,
                   double dX, double Xtot, double ratio, double er, const double* script,
                   int n_script, double* out) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Mixture m = mixture_from(T, n_mix, mats, fracs);
  Philox4x32 rng(0, 0);
  rng.script = script; rng.script_n = n_script;
  double X = 0, k = 0;
  bool acc = sample_ehs(T, m, ki, kf, dX, Xtot, ratio, er, rng, X, k);
  out[0] = acc; out[1] = acc ? X : 0.0; out[2] = acc ? k : 0.0;
  return rng.script_pos;
}

// sample_mixture_material, src/Physics.jl:1331-1365: fractions as given
// (MaterialMixture normalises positive sums, src/Types.jl:360-369).
static int sample_mixture_material_o(const Tables& T, int n, const int32_t* mats,
                                     const double* fr, double energy, Philox4x32& rng) {
  if (n == 1) return mats[0];
  double tot = 0.0;
  for (int j = 0; j < n; ++j) tot += fr[j] * property(T, mats[j], T_CROSS_SECTION, energy);
  if (tot <= 0.0) {
    const double u = rng.uniform();
    double c = 0.0;
    for (int j = 0; j < n; ++j) {
      c += fr[j];
      if (u <= c) return mats[j];
    }
    return mats[n - 1];
  }
  const double u = rng.uniform() * tot;
  double c = 0.0;
  for (int j = 0; j < n; ++j) {
    c += fr[j] * property(T, mats[j], T_CROSS_SECTION, energy);
    if (u <= c) return mats[j];
  }
  return mats[n - 1];
}

// compute_del_cross_section, src/Straggling.jl:186-199
static double del_cross_section_o(const Tables& T, int m, double k) {
  if (k < 1e-3) return 0.0;
  return std::max(property(T, m, T_CROSS_SECTION, k), 0.0);
}

// sample_del_event, src/Straggling.jl:217-317 (BMC_ALPHA = 2).
// out: occurred, X_del, k_del, process.
static void sample_del_event_o(const Tables& T, int m, double ki, double kf, double dX,
                               double Xtot, double ratio, double cutoff, bool backward,
                               Philox4x32& rng, double* out) {
  out[0] = out[1] = out[2] = out[3] = 0.0;
  const double sgn = backward ? -1.0 : 1.0;
  const double xs = std::max(del_cross_section_o(T, m, ki), del_cross_section_o(T, m, kf));
  if (xs <= 0.0) return;
  const double zeta = rng.uniform();
  if (zeta <= 1e-12 || zeta >= 1.0 - 1e-12) return;
  const double X = -std::log(zeta) / xs;
  if (X >= dX) return;
  const double kh = prop_kinetic(T, true, m, Xtot - sgn * X);
  const double kat = ki - sgn * ratio * std::fabs(ki - kh);
  if (kat <= 0.0) return;
  const double r = del_cross_section_o(T, m, kat) / xs;
  if (rng.uniform() > r) return;
  const double alpha = 2.0;
  const double u = rng.uniform();
  const double ct = std::pow(cutoff, 1.0 - alpha);
  double nu = std::pow(u * (1.0 - ct) + ct, 1.0 / (1.0 - alpha));
  nu = std::min(std::max(nu, cutoff), 1.0);
  const double kd = backward ? kat / (1.0 - nu) : kat * (1.0 - nu);
  const double b = property(T, m, T_BREMS, kat), p = property(T, m, T_PAIR, kat),
               ph = property(T, m, T_PHOTO, kat);
  const double tot = b + p + ph;
  int proc = 1;
  if (tot > 0.0) {
    const double rp = rng.uniform() * tot;
    proc = rp < b ? 1 : (rp < b + p ? 2 : 3);
  }
  out[0] = 1.0; out[1] = X; out[2] = kd; out[3] = proc;
}

// n seeded draws (sample i from Philox(seed, i)), or one scripted call
// (returns the draws consumed).
int orc_sample_mixture_material(int M, int K, const double* data, const double* scal,
                                int n_mix, const int32_t* mats, const double* fracs, int n,
                                const double* energies, unsigned long long seed,
                                const double* script, int n_script, int32_t* out) {
  Tables T = make_tables(M, K, data, scal, MUON_MASS, MUON_C_TAU);
  double tot = 0.0;
  for (int j = 0; j < n_mix; ++j) tot += fracs[j];
  std::vector<double> fr(n_mix);
  for (int j = 0; j < n_mix; ++j) fr[j] = tot > 0 ? fracs[j] / tot : fracs[j];
  if (n_script > 0) {
    Philox4x32 rng(0, 0);
    rng.script = script; rng.script_n = n_script;
    try {
      out[0] = sample_mixture_material_o(T, n_mix, mats, fr.data(), energies[0], rng);
    } catch (const std::exception&) {
      return -1;   // the script ran out of draws
    }
    return rng.script_pos;
  }
  for (int i = 0; i < n; ++i) {
    Philox4x32 rng(seed, (uint64_t)i);
    out[i] = sample_mixture_material_o(T, n_mix, mats, fr.data(), energies[i], rng);
  }
  return 0;
}

int orc_sample_del_event(int M, int K, const double* data, const double* scal, int m, int n,
                         const double* ki, const double* kf, const double* dX,
                         const double* Xtot, const double* ratio, double cutoff, int backward,
                         unsigned long long seed, const double* script, int n_script,
                         double* out /* 4 per sample */) {
  Tables T = make_tables(M, K, data, scal, MUON_MASS, MUON_C_TAU);
  if (n_script > 0) {
    Philox4x32 rng(0, 0);
    rng.script = script; rng.script_n = n_script;
    try {
      sample_del_event_o(T, m, ki[0], kf[0], dX[0], Xtot[0], ratio[0], cutoff, backward != 0,
                         rng, out);
    } catch (const std::exception&) {
      return -1;   // the script ran out of draws
    }
    return rng.script_pos;
  }
  for (int i = 0; i < n; ++i) {
    Philox4x32 rng(seed, (uint64_t)i);
    sample_del_event_o(T, m, ki[i], kf[i], dX[i], Xtot[i], ratio[i], cutoff, backward != 0, rng,
                       out + 4 * (size_t)i);
  }
  return 0;
}

// property_range (MIXED or CSDA) and compute_ehs_mean_free_path on a
// material or runtime mixture.
double orc_mix_range(int M, int K, const double* data, const double* scal, double mass,
                     double ctau, int n_mix, const int32_t* mats, const double* fracs, int mixed,
                     double k) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Mixture m = mixture_from(T, n_mix, mats, fracs);
  return mix_range(T, mixed != 0, m, k);
}
double orc_ehs_path(int M, int K, const double* data, const double* scal, double mass,
                    double ctau, int n_mix, const int32_t* mats, const double* fracs, double k,
                    double er) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  Mixture m = mixture_from(T, n_mix, mats, fracs);
  return ehs_path(T, m, k, er);
}

}  // extern "C"

// ============================================ fixed-path MC over many paths
extern "C" {

// compute_directional_flux_mc (src/Tomography.jl:979-1041) for every path.
// Cell c uses runtime mixture media[cell_medium[c]] at cell_density[c].
// Event s of path r draws from Philox keyed (seed, r * n_samples + s).
// Returns 0, -2 (DomainError) or -3 (other error) with *err_event the first
// failing event.
int orc_mc_paths(int M, int K, const double* data, const double* scal, double mass, double ctau,
                 const int32_t* n_el, const double* eZ, const double* eA, const double* eI,
                 const double* ef, double cutoff, double elastic_ratio, int n_media,
                 const int32_t* media_n, const int32_t* media_mats, const double* media_fracs,
                 const int32_t* cell_medium, const double* cell_density, int rock, int air,
                 double rho_rock, double det_elev, double palt, int n_rays,
                 const int32_t* row_ptr, const int32_t* cells, const double* lens,
                 const double* remaining, const double* exit_z, const double* cos_t,
                 const uint8_t* valid, int n_samples, const double* se, const double* sw,
                 const double* sq, unsigned long long seed, int straggling, double thr_low,
                 int threads, double* flux, double* sigma, double* w_events,
                 long long* err_event) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  std::vector<std::vector<ElementRec>> rec(M);
  int o = 0;
  for (int m = 0; m < M; ++m)
    for (int i = 0; i < n_el[m]; ++i, ++o) rec[m].push_back({eZ[o], eA[o], eI[o], ef[o]});
  std::vector<Mixture> mixes(n_media + 2);
  std::vector<MixedKernel> kerns(n_media + 2);
  for (int j = 0; j < n_media; ++j) {
    const int n = media_n[j];
    mixes[j] = make_mixture(T, std::vector<int>(media_mats + 4 * j, media_mats + 4 * j + n),
                            std::vector<double>(media_fracs + 4 * j, media_fracs + 4 * j + n));
  }
  mixes[n_media] = make_mixture(T, {rock}, {1.0});
  mixes[n_media + 1] = make_mixture(T, {air}, {1.0});
  for (int j = 0; j < n_media + 2; ++j)
    if (straggling) kerns[j] = prepare_mixed_kernel(T, mixes[j], rec, cutoff);
  std::vector<Medium> media(n_media + 2);
  for (int j = 0; j < n_media + 2; ++j) media[j] = {&mixes[j], &kerns[j]};
  int n_cells = 0;
  for (int s = 0; s < row_ptr[n_rays]; ++s) n_cells = std::max(n_cells, cells[s] + 1);
  std::vector<Medium> cm(n_cells);
  for (int c = 0; c < n_cells; ++c) cm[c] = media[cell_medium[c]];
  McSetup S{};
  S.T = &T; S.cutoff = cutoff; S.elastic_ratio = elastic_ratio;
  double emax = 1e9;
  for (int m = 0; m < M; ++m) emax = std::min(emax, T.arr(m, T_ENERGY)[K - 1]);
  PathMc P{&S, cm.data(), cell_density, &media[n_media], &media[n_media + 1], rho_rock, det_elev,
           palt, thr_low, emax, emax, straggling != 0};
  if (threads > 0) omp_set_num_threads(threads);
  int status = 0;
  long long first_bad = -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < n_rays; ++r) {
    if ((valid && !valid[r]) || !(cos_t[r] > 0.0)) {
      flux[r] = sigma[r] = NAN;
      continue;
    }
    const double st = std::sqrt(std::max(0.0, 1.0 - cos_t[r] * cos_t[r]));
    const double dir[3] = {st, 0.0, cos_t[r]};
    const int b = row_ptr[r], ns = row_ptr[r + 1] - b;
    double s1 = 0, s2 = 0;
    for (int s = 0; s < n_samples; ++s) {
      const long long e = (long long)r * n_samples + s;
      Philox4x32 rng(seed, (uint64_t)e);
      double wi = 0.0;
      try {
        wi = path_event(P, cells + b, lens + b, ns, remaining[r], exit_z[r], cos_t[r], dir, se[s],
                        sq[s], sw[s], rng);
      } catch (const DomainErr&) {
#pragma omp critical
        { if (status == 0 || e < first_bad) { status = -2; first_bad = e; } }
      } catch (const std::exception&) {
#pragma omp critical
        { if (status == 0 || e < first_bad) { status = -3; first_bad = e; } }
      }
      if (w_events) w_events[e] = wi;
      s1 += wi;
      s2 += wi * wi;
    }
    const double n = std::max(n_samples, 1);
    const double f = s1 / n;
    flux[r] = f;
    sigma[r] = std::sqrt(std::max(0.0, (s2 / n - f * f) / std::max(1.0, n - 1.0)));
  }
  if (err_event) *err_event = first_bad;
  return status;
}

}  // extern "C"

extern "C" {
// propagate_mc_uniform_segment on one material (src/Tomography.jl:876-920):
// st = E, w, t, L, X in and out.  Returns 0, -2 or -3.
int orc_mc_uniform_segment(int M, int K, const double* data, const double* scal, double mass,
                           double ctau, const int32_t* n_el, const double* eZ, const double* eA,
                           const double* eI, const double* ef, double cutoff, int material,
                           double density, double distance, int straggling, double thr_low,
                           unsigned long long seed, double* st) {
  Tables T = make_tables(M, K, data, scal, mass, ctau);
  std::vector<std::vector<ElementRec>> rec(M);
  int o = 0;
  for (int m = 0; m < M; ++m)
    for (int i = 0; i < n_el[m]; ++i, ++o) rec[m].push_back({eZ[o], eA[o], eI[o], ef[o]});
  Mixture mx = make_mixture(T, {material}, {1.0});
  MixedKernel k = straggling ? prepare_mixed_kernel(T, mx, rec, cutoff) : MixedKernel{};
  Medium md{&mx, &k};
  McSetup S{};
  S.T = &T; S.cutoff = cutoff;
  McState s;
  s.E = st[0]; s.w = st[1]; s.t = st[2]; s.L = st[3]; s.X = st[4];
  s.pos[0] = s.pos[1] = s.pos[2] = 0.0;
  s.dir[0] = s.dir[1] = 0.0; s.dir[2] = -1.0;
  double emax = 1e9;
  for (int m = 0; m < M; ++m) emax = std::min(emax, T.arr(m, T_ENERGY)[K - 1]);
  Philox4x32 rng(seed, 0);
  int rc = 0;
  try {
    mc_path_segment(S, md, s, density, false, 0.0, 1.0, distance, straggling != 0, thr_low, emax,
                    rng);
  } catch (const DomainErr&) {
    rc = -2;
  } catch (const std::exception&) {
    rc = -3;
  }
  st[0] = s.E; st[1] = s.w; st[2] = s.t; st[3] = s.L; st[4] = s.X;
  return rc;
}
}  // extern "C"

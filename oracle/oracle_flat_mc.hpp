// CPU oracle: the stochastic flat two-layer backward Monte Carlo.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#pragma once
#include "oracle_mc.hpp"
#include "oracle_geometry.hpp"

namespace orc {

// Julia eps(x): spacing of Float64 at x (x >= 0)
inline double ulp(double x) { return std::nextafter(std::fabs(x), INFINITY) - std::fabs(x); }

struct Medium {
  const Mixture* mix;
  const MixedKernel* kernel;
};

struct McState {
  double E, w, t;
  double pos[3], dir[3];
  double X = 0.0, L = 0.0;  // grammage and distance accumulated by mixed steps
};

struct McSetup {
  const Tables* T;
  Medium rock, air, porous;
  bool has_porous;
  double thick, rock_rho, porous_rho, porous_thick, palt;
  double cutoff, elastic_ratio;
};

// transport_backward_step_full, mode = :straggled, src/Geometry.jl:148-357
inline void straggled_step(const McSetup& S, const Medium& md, McState& st, double density,
                           double ds, bool scattering, double energy_limit, Philox4x32& rng) {
  const Tables& T = *S.T;
  const Mixture& mix = *md.mix;
  const MixedKernel& K = *md.kernel;
  const double ki = st.E, mass = T.mass;
  double dX = ds * density;
  const double Xi = mix_range(T, true, mix, ki);
  const double Xtot = Xi + dX;
  double kf = ki, ratio = 1.0;
  if (ki > 0.001) {
    fluctuate(T, mix, ki, dX, Xi, rng, kf, ratio);
    kf = std::max(kf, ki);
  } else {
    kf = mix_kinetic(T, true, mix, Xtot);
  }
  if (energy_limit > 0 && kf >= energy_limit) {
    const double dk = std::fabs(ki - kf);
    if (dk > 0) {
      dX = dX * (std::fabs(energy_limit - ki) / dk);
      kf = energy_limit;
    }
  }
  bool del = false, ehs = false;
  double X_del = INFINITY, k_del = 0, X_ehs = INFINITY, k_ehs = 0;
  {
    const double rate = std::max(klookup(ki, K.e, K.sigma), klookup(kf, K.e, K.sigma));
    if (rate > 0) {
      const double cand = rng.exponential() / rate;
      if (cand < dX) {
        const double kh = mix_kinetic(T, true, mix, Xi + cand);
        const double pre = ki + ratio * std::fabs(kh - ki);
        if (pre > 0 && rng.uniform() <= klookup(pre, K.e, K.sigma) / rate) {
          del = true; X_del = cand; k_del = pre;
        }
      }
    }
  }
  if (scattering) {
    double xe, ke;
    if (sample_ehs(T, mix, ki, kf, dX, Xi, ratio, S.elastic_ratio, rng, xe, ke) && ke > 0) {
      ehs = true; X_ehs = xe; k_ehs = ke;
    }
  }
  if (del && ehs) {
    if (X_del <= X_ehs) ehs = false; else del = false;
  }
  double dw = 1.0, ckf = kf;
  int proc = 0, elem = -1;
  if (del) {
    ckf = k_del;
    DelOut o = reverse_del_event(K, k_del, rng, true);
    kf = o.k; dw = o.w; proc = o.proc; elem = o.elem;
    dX = X_del;
  }
  if (ehs) { kf = k_ehs; ckf = k_ehs; dX = X_ehs; }
  const double step = dX / density;
  double d[3] = {st.dir[0], st.dir[1], st.dir[2]};
  if (scattering) {
    if (del && proc > 0) {
      double mu;
      const ElementRec& e = K.el[elem];
      if (proc == 1 || proc == 2) mu = polar_brems(mass, kf, ckf, e.Z, rng);
      else if (proc == 3) mu = polar_photo(mass, kf, ckf, rng);
      else mu = polar_ionisation(mass, kf, ckf);
      if (mu > 0) rotate(d, mu, rng);
    }
    if (ehs) {
      const double mu = scattering_angle(T, mix, k_ehs, rng);
      if (mu > 0) rotate(d, mu, rng);
    }
    const double ms = soft_scattering(T, mix, ki, ckf, step, density, rng);
    if (ms > 0) rotate(d, ms, rng);
  }
  const double dt = step * mass / 2 * (1.0 / std::sqrt(ki * (ki + 2 * mass)) + 1.0 / std::sqrt(ckf * (ckf + 2 * mass)));
  for (int a = 0; a < 3; ++a) st.pos[a] -= step * st.dir[a];
  st.L += step;
  st.X += dX;
  const double wf = mix_stop(T, true, mix, ckf) / mix_stop(T, true, mix, ki);
  st.w = st.w * wf * std::exp(-dt / T.ctau) * dw;
  st.t += dt;
  st.E = kf;
  for (int a = 0; a < 3; ++a) st.dir[a] = d[a];
}

// transport_mixed_step, src/MixedTransport.jl:144-200
inline void mixed_step(const McSetup& S, const Medium& md, McState& st, double density,
                       double distance, double energy_limit, Philox4x32& rng) {
  const MixedKernel& c = *md.kernel;
  const double m = c.mass;
  if (!(density > 0 && distance >= 0)) throw std::runtime_error("Positive density and nonnegative distance required");
  double k = st.E;
  const double emax = c.e.back();
  if (energy_limit > emax) throw std::runtime_error("Requested mixed energy limit exceeds prepared tables");
  double remaining = distance * density, travelled = 0, proper = 0, weight = st.w;
  const double ceiling = energy_limit > 0 ? std::min(energy_limit, emax) : emax;
  if (st.E > emax) throw DomainErr("Ancestor energy exceeds mixed tables");
  const double tiny = ulp(std::max(1.0, distance * density));
  bool exhausted = true;
  for (int ev = 0; ev < 100000; ++ev) {
    if (remaining <= tiny) { exhausted = false; break; }
    if (k >= ceiling) { exhausted = false; break; }
    const double ri = klookup(k, c.e, c.range), hi = klookup(k, c.e, c.hazard);
    const double rtc = std::max(0.0, klookup(ceiling, c.e, c.range) - ri);
    const double eg = std::min(remaining, rtc);
    const double targetr = ri + eg;
    const double endpoint = eg == rtc ? ceiling : std::max(k, klookup(targetr, c.range, c.e));
    const double od = rng.exponential();
    const double eventh = hi + od;
    const bool occurs = eventh < klookup(endpoint, c.e, c.hazard);
    const double pre = occurs ? std::max(k, klookup(eventh, c.hazard, c.e)) : endpoint;
    const double dx = occurs ? std::max(0.0, klookup(pre, c.e, c.range) - ri) : eg;
    double dt;
    if (dx <= 8 * ulp(ri) || pre - k <= 8 * ulp(k))
      dt = dx / density * m / 2 * (1.0 / std::sqrt(k * (k + 2 * m)) + 1.0 / std::sqrt(pre * (pre + 2 * m)));
    else
      dt = (klookup(pre, c.e, c.proper) - klookup(k, c.e, c.proper)) / density;
    weight *= klookup(pre, c.e, c.stop) / klookup(k, c.e, c.stop) * std::exp(-dt / c.ctau);
    travelled += dx;
    remaining = std::max(0.0, remaining - dx);
    proper += dt;
    k = pre;
    if (occurs) {
      DelOut o = reverse_del_event(c, k, rng, false);
      k = o.k;
      weight *= o.w;
      if (!(std::isfinite(weight) && weight >= 0)) throw std::runtime_error("Invalid adjoint hard-event weight");
      if (weight == 0) { exhausted = false; break; }
    } else {
      exhausted = false;
      break;
    }
  }
  if (exhausted) throw std::runtime_error("Mixed transport exceeded 100000 hard events");
  if (weight > 0 && k >= emax && remaining > tiny && (energy_limit < 0 || k < energy_limit))
    throw DomainErr("Mixed transport exhausted energy tables before segment completion");
  const double dd = travelled / density;
  st.X += travelled;
  st.L += dd;
  st.E = k;
  st.t += proper;
  st.w = weight;
  for (int a = 0; a < 3; ++a) st.pos[a] -= dd * st.dir[a];
}

// CSDA step (transport_backward_step, deterministic) on a mixture
inline void csda_mix_step(const McSetup& S, const Medium& md, McState& st, double density, double ds) {
  const Tables& T = *S.T;
  const Mixture& mix = *md.mix;
  const double ki = st.E, gs = ds * density;
  const double Xi = mix_range(T, false, mix, ki);
  const double kf = mix_kinetic(T, false, mix, Xi + gs);
  const double gi = 1.0 + ki / T.mass, gc = 1.0 + kf / T.mass, ga = (gi + gc) / 2.0;
  const double beta = std::sqrt(std::max(1.0 - 1.0 / (ga * ga), 1e-12));
  const double dt = ds / (beta * ga);
  for (int a = 0; a < 3; ++a) st.pos[a] -= ds * st.dir[a];
  st.L += ds;
  st.X += gs;
  const double wf = mix_stop(T, false, mix, kf) / mix_stop(T, false, mix, ki);
  st.w = st.w * wf * std::exp(-dt / T.ctau);
  st.t += dt;
  st.E = kf;
}

// transport_backward_through_geometry with an RNG, src/Geometry.jl:534-672,
// and the acceptance of compute_flux_single (:739-757).
inline double flat_event(const McSetup& S, double energy, double cos_t, double sin_t, double charge,
                         double threshold, double mode_switch, bool straggling, bool scattering,
                         Philox4x32& rng, McState* out_state = nullptr) {
  const Tables& T = *S.T;
  McState st;
  st.E = energy; st.w = 1.0; st.t = 0.0;
  st.pos[0] = st.pos[1] = st.pos[2] = 0.0;
  st.dir[0] = -sin_t; st.dir[1] = 0.0; st.dir[2] = -cos_t;
  const double palt = S.palt;
  for (int outer = 0; outer < 64; ++outer) {
    if (!(st.E < threshold - EPS64)) break;
    double z = st.pos[2];
    if (z < 0 || z >= palt - EPS64) break;
    if (st.w <= 0 || !std::isfinite(st.w)) break;
    bool use_str, use_mixed, use_scat;
    double limit;
    if (st.E < mode_switch - EPS64) {
      use_str = straggling; use_mixed = false; use_scat = scattering;
      limit = std::min(mode_switch, threshold);
    } else {
      use_str = false; use_mixed = straggling; use_scat = false;
      limit = threshold;
    }
    for (int inner = 0; inner < 50000; ++inner) {
      z = st.pos[2];
      if (z < 0 || z >= palt - EPS64) break;
      if (st.E >= limit - EPS64) break;
      if (st.w <= 0 || !std::isfinite(st.w)) break;
      const double uz = -st.dir[2];
      const double porous_z = S.has_porous ? S.thick - S.porous_thick : S.thick;
      const Medium* md;
      double density, stb;
      const bool in_air = !(z < S.thick);
      if (!in_air) {
        if (S.has_porous && z >= porous_z) { md = &S.porous; density = S.porous_rho; }
        else { md = &S.rock; density = S.rock_rho; }
        if (uz > EPS64) stb = (((S.has_porous && z < porous_z) ? porous_z : S.thick) - z) / uz;
        else if (uz < -EPS64) stb = (((S.has_porous && z >= porous_z) ? porous_z : 0.0) - z) / uz;
        else stb = 1000.0;
      } else {
        md = &S.air;
        density = air_density(z);
        if (uz > EPS64) stb = (palt - z) / uz;
        else if (uz < -EPS64) stb = (S.thick - z) / uz;
        else stb = 1000.0;
      }
      // _backward_geometric_step_limit, src/Geometry.jl:463-476
      const double Xi = mix_range(T, false, *md->mix, st.E);
      double lim = 0.01 * Xi / std::max(density, 1e-6);
      if (use_scat) {
        const double lb1 = mix_transport_path(T, *md->mix, st.E);
        if (lb1 > 0 && density > 0) lim = std::min(lim, 0.01 * lb1 / density);
      }
      if (in_air) lim = std::min(lim, 0.05 * 12e3 / std::max(std::fabs(uz), 0.05));
      double step = std::min(stb + 1e-7, lim);
      step = std::min(std::max(step, STEP_MIN), 1000.0);
      if (use_str) straggled_step(S, *md, st, density, step, use_scat, limit, rng);
      else if (use_mixed) mixed_step(S, *md, st, density, step, limit, rng);
      else csda_mix_step(S, *md, st, density, step);
    }
  }
  if (out_state) *out_state = st;
  if (st.pos[2] >= palt - 1.0 && st.E < threshold) {
    double cf = std::min(std::max(-st.dir[2], 0.0), 1.0);
    if (cf <= 0) return 0.0;
    return st.w * flux_gccly(cf, st.E, charge);
  }
  return 0.0;
}

// ------------------------------------------------ fixed-path (tomography) MC
// _mc_segment_step, src/Tomography.jl:873-874
inline double mc_segment_step(double remaining, double max_step) {
  return std::min(remaining, std::max(1e-6, std::min(remaining + 1e-7, max_step)));
}

// propagate_mc_uniform_segment / propagate_mc_air_segment (src/Tomography.jl:
// 876-969) with scattering off.  air: density from the altitude of the
// current position.  Throws on a step that makes no spatial progress.
inline void mc_path_segment(const McSetup& S, const Medium& md, McState& st, double density,
                            bool air, double det_elev, double cos_t, double distance,
                            bool straggling, double thr_low, double energy_limit,
                            Philox4x32& rng) {
  const Tables& T = *S.T;
  double remaining = distance;
  while (remaining > 1e-8 && st.E < energy_limit - EPS64) {
    if (st.w <= 0.0 || !std::isfinite(st.w)) break;
    const double cur = st.E < thr_low - EPS64 ? std::min(thr_low, energy_limit) : energy_limit;
    double dens = density, max_step;
    const double Xi = mix_range(T, false, *md.mix, st.E);
    if (air) {
      dens = air_density(det_elev + st.pos[2]);
      max_step = 0.01 * Xi / std::max(dens, 1e-8);
      max_step = std::min(max_step, 0.05 * 12e3 / std::max(cos_t, 0.05));
    } else {
      max_step = 0.01 * Xi / std::max(dens, 1e-6);
    }
    const double step = mc_segment_step(remaining, max_step);
    const double oldL = st.L, oldE = st.E;
    if (st.E < thr_low - EPS64) {
      if (straggling) straggled_step(S, md, st, dens, step, false, cur, rng);
      else csda_mix_step(S, md, st, dens, step);
    } else {
      if (straggling) mixed_step(S, md, st, dens, step, cur, rng);
      else csda_mix_step(S, md, st, dens, step);
    }
    const double actual = st.L - oldL;
    if (st.w <= 0.0 || !std::isfinite(st.w)) break;
    if (actual <= 0 && oldE < cur && cur < energy_limit && st.E >= cur) continue;
    if (!(actual > 0)) {
      if (st.E >= energy_limit) break;
      throw std::runtime_error("MC segment made no spatial progress");
    }
    remaining = std::max(0.0, remaining - actual);
  }
}

// One event of compute_directional_flux_mc (src/Tomography.jl:979-1041):
// cells, dry-rock remainder, atmosphere; returns 2 w_sample flux_single.
struct PathMc {
  const McSetup* S;
  const Medium* cell_media;   // per cell
  const double* cell_density; // per cell
  const Medium* rock;
  const Medium* air;
  double rho_rock, det_elev, palt, thr_low, energy_limit, source_max;
  bool straggling;
};
inline double path_event(const PathMc& P, const int32_t* cells, const double* lens, int ns,
                         double remaining, double exit_z, double zenith_cos, const double dir[3],
                         double energy, double charge, double weight, Philox4x32& rng) {
  McState st;
  st.E = energy; st.w = 1.0; st.t = 0.0;
  for (int a = 0; a < 3; ++a) { st.pos[a] = 0.0; st.dir[a] = -dir[a]; }
  for (int s = 0; s < ns; ++s) {
    const int c = cells[s];
    mc_path_segment(*P.S, P.cell_media[c], st, P.cell_density[c], false, P.det_elev, zenith_cos,
                    lens[s], P.straggling, P.thr_low, P.energy_limit, rng);
    if (st.w <= 0.0 || !std::isfinite(st.w)) break;
  }
  if (st.w > 0.0 && std::isfinite(st.w) && remaining > 1e-8)
    mc_path_segment(*P.S, *P.rock, st, P.rho_rock, false, P.det_elev, zenith_cos, remaining,
                    P.straggling, P.thr_low, P.energy_limit, rng);
  if (st.w > 0.0 && std::isfinite(st.w)) {
    const double air_d = std::max(0.0, P.palt - exit_z) / std::max(zenith_cos, 1e-3);
    mc_path_segment(*P.S, *P.air, st, 0.0, true, P.det_elev, zenith_cos, air_d, P.straggling,
                    P.thr_low, P.energy_limit, rng);
  }
  double f = 0.0;
  if (st.w > 0.0 && std::isfinite(st.w) && st.E < P.source_max)
    f = st.w * flux_gccly(zenith_cos, st.E, charge);
  return 2.0 * weight * f;
}

}  // namespace orc

// Located-energy form of the device transport library, for tables whose
// materials share one energy grid (every table built from the PUMAS dE/dx
// files or by create_physics does).
//
// An energy is located on the grid once (clamp flags, interval, interpolation
// weight) and every property of every mixture component at that energy is a
// lerp on the same interval.  Intervals are searched from a guess (the
// previous interval of the event's energy), falling back to bisection, and
// the search returns the same interval as the bisection of the reference.
// The arithmetic of each property value is the one of interpolate_table_fast
// (src/Physics.jl:660-687), so results are bitwise those of the generic
// functions in dp_transport.hpp; only the number of dependent table loads
// changes.
#pragma once
#include "dp_transport.hpp"
#include "dp_del_defer.hpp"

namespace dpk {

// An energy located on the shared grid.  c = -1 below the first knot, +1 at
// or above the last one (interpolate_table_fast clamps), 0 inside.
struct Loc {
  double e, le, t;
  int i, c;
};
__device__ __forceinline__ Loc loc_e(const DevTables& T, double e, double le, int g) {
  const double* es = T.a(0, A_ENERGY);
  const double* les = T.a(0, A_LOG_ENERGY);
  const int K = T.K;
  Loc L;
  L.e = e; L.le = le;
  if (e <= es[0]) { L.c = -1; L.i = 0; L.t = 0.0; }
  else if (e >= es[K - 1]) { L.c = 1; L.i = K - 2; L.t = 1.0; }
  else {
    L.c = 0;
    L.i = locate_idx(les, K, T.pow2, le, g);
    L.t = (le - les[L.i]) / (les[L.i + 1] - les[L.i]);
  }
  return L;
}
__device__ __forceinline__ double at(const double* __restrict__ ys, const Loc& L, int K) {
  if (L.c < 0) return ys[0];
  if (L.c > 0) return ys[K - 1];
  return ys[L.i] + L.t * (ys[L.i + 1] - ys[L.i]);
}

__device__ __forceinline__ double prop_L(const DevTables& T, int m, int arr, const Loc& L) {
  return at(T.a(m, arr), L, T.K);
}
__device__ __forceinline__ double mix_sum_L(const DevTables& T, const DevMix& x, int arr,
                                            const Loc& L) {
  if (x.n == 1) return prop_L(T, x.mat[0], arr, L);
  double v = 0.0;
  for (int j = 0; j < x.n; ++j) v += x.frac[j] * prop_L(T, x.mat[j], arr, L);
  return v;
}
__device__ __forceinline__ double mix_range_L(const DevTables& T, const DevMix& x, bool mixed,
                                              const Loc& L) {
  if (x.n == 1) return prop_L(T, x.mat[0], mixed ? A_MIXED_RANGE : A_CSDA_RANGE, L);
  return at(mixed ? x.mrange : x.crange, L, T.K);
}
// Inverse range from a guessed interval (same arithmetic as inverse_range).
__device__ __forceinline__ double mix_kinetic_g(const DevTables& T, const DevMix& x, bool mixed,
                                                double X, int g) {
  const int K = T.K;
  const double *es, *les, *rt;
  if (x.n == 1) {
    const int m = x.mat[0];
    es = T.a(m, A_ENERGY); les = T.a(m, A_LOG_ENERGY);
    rt = T.a(m, mixed ? A_MIXED_RANGE : A_CSDA_RANGE);
  } else {
    es = x.e; les = x.le; rt = mixed ? x.mrange : x.crange;
  }
  if (X <= rt[0]) return es[0];
  if (X >= rt[K - 1]) return es[K - 1];
  const int i = locate_idx(rt, K, T.pow2, X, g);
  const double dr = rt[i + 1] - rt[i];
  const double t = (X - rt[i]) / dr;
  const double dle = les[i + 1] - les[i];
  return exp(les[i] + t * dle);
}
__device__ __forceinline__ double mix_transport_path_L(const DevTables& T, const DevMix& x,
                                                       const Loc& L) {
  if (x.n == 1) return prop_L(T, x.mat[0], A_TRANSPORT_PATH, L);
  double inv = 0.0;
  for (int j = 0; j < x.n; ++j) {
    const double lp = prop_L(T, x.mat[j], A_TRANSPORT_PATH, L);
    if (lp > 0) inv += x.frac[j] / lp;
  }
  return inv > 0 ? 1.0 / inv : 1e9;
}
__device__ __forceinline__ double ehs_path_one_L(const DevTables& T, int m, const Loc& L) {
  const double k = L.e;
  if (k < 1e-3) return EHS_PATH_MAX;
  const double Lb = prop_L(T, m, A_ELASTIC_PATH, L);
  const double p2 = k * (k + 2.0 * T.mass);
  if (Lb > 0 && Lb < EHS_PATH_MAX * p2) return Lb / p2;
  const double tp = prop_L(T, m, A_TRANSPORT_PATH, L);
  if (tp > 0) return fmin(tp / T.elastic_ratio, EHS_PATH_MAX);
  return EHS_PATH_MAX;
}
__device__ __forceinline__ double ehs_path_L(const DevTables& T, const DevMix& x, const Loc& L) {
  if (x.n == 1) return ehs_path_one_L(T, x.mat[0], L);
  double inv = 0.0;
  for (int j = 0; j < x.n; ++j) {
    const double lp = ehs_path_one_L(T, x.mat[j], L);
    if (lp > 0 && lp < EHS_PATH_MAX) inv += x.frac[j] / lp;
  }
  return inv > 0 ? 1.0 / inv : EHS_PATH_MAX;
}

// Guess of the kernel-grid interval of a located energy: the kernel grid
// refines the shared grid 8 times per interval in log energy.
__device__ __forceinline__ int kguess(const DevKernel& K, const Loc& L) {
  if (L.c < 0) return 0;
  if (L.c > 0) return K.n - 2;
  return 8 * L.i + (int)(8.0 * L.t);
}

// fluctuate_energy_loss (backward), src/Straggling.jl:103-173, :1046-1105.
__device__ __noinline__ static void fluctuate_f(const DevTables& T, const DevMix& x, const Loc& Li,
                                                double dX, double Xtot, Philox& rng, double& kout,
                                                double& ratio) {
  const double ki = Li.e;
  double k1 = mix_kinetic_g(T, x, true, Xtot + dX, Li.i);
  const double dk0 = fabs(ki - k1);
  ratio = 1.0;
  if (k1 > 0 && dk0 > 0) {
    const Loc L1 = loc_e(T, k1, log(k1), Li.i);
    const double O0 = mix_sum_L(T, x, A_STRAGGLING, Li);
    const double O1 = mix_sum_L(T, x, A_STRAGGLING, L1);
    double dk12 = 0.5 * dX * (O0 + O1);
    if (dX / Xtot > 0.05) {
      const double de1 = mix_sum_L(T, x, A_MIXED_STOP, L1);
      const double de0 = mix_sum_L(T, x, A_MIXED_STOP, Li);
      dk12 *= 1.0 - dX / dk0 * (de1 - de0);
    }
    const double dk1 = sqrt(fmax(dk12, 0.0));
    if (dk0 >= 3.0 * dk1 && dk1 > 0) {
      k1 += truncated_randn(rng) * dk1;
    } else if (dk0 >= 1.7320508 * dk1 && dk1 > 0) {
      k1 += 1.7320508 * (1.0 - 2.0 * uniform(rng)) * dk1;
    } else if (dk1 > 0) {
      const double dk32 = 3.0 * dk12, dk02 = dk0 * dk0;
      const double a = 1.0 - (dk32 - dk02) / (dk32 + 3.0 * dk02);
      if (uniform(rng) <= a) {
        const double b = 0.5 * (dk32 + 3.0 * dk02) / dk0;
        k1 = ki + b * uniform(rng);
        if (k1 < 0) k1 = 0;
      } else {
        k1 = ki;
      }
    }
    ratio = fabs(ki - k1) / dk0;
  }
  if (ratio == 0.0) ratio = 1.0;
  kout = fmax(k1, 0.0);
}

// sample_ehs_event (backward), src/Straggling.jl:373-434, :1284-1334.  On
// acceptance Le is the located event energy.
__device__ __noinline__ static bool sample_ehs_f(const DevTables& T, const DevMix& x, const Loc& Li,
                                                 double kf, double dX, double Xtot, double ratio,
                                                 Philox& rng, double& X_ehs, double& k_ehs,
                                                 Loc& Le) {
  const double ki = Li.e;
  double kmin = fmin(ki, kf), kmax = fmax(ki, kf);
  if (kmin < 0.001) { const double t = kmin; kmin = kmax; kmax = t; }
  const Loc Lm = kmin == ki ? Li : loc_e(T, kmin, log(kmin), Li.i);
  const double lb = ehs_path_L(T, x, Lm);
  if (lb <= 0 || lb >= EHS_PATH_MAX) return false;
  const double zeta = uniform(rng);
  if (zeta <= 0 || zeta >= 1) return false;
  X_ehs = -log(zeta) * lb;
  if (X_ehs >= dX) return false;
  const double kh = mix_kinetic_g(T, x, true, Xtot + X_ehs, Li.i);
  k_ehs = ki + ratio * fabs(ki - kh);
  if (k_ehs <= 0) return false;
  Le = loc_e(T, k_ehs, log(k_ehs), Li.i);
  const double lbk = ehs_path_L(T, x, Le);
  if (lbk <= 0 || lbk >= EHS_PATH_MAX) return false;
  const double r = lbk / lb;
  if (r <= 0 || uniform(rng) > 1.0 / r) return false;
  return true;
}

// sample_scattering_angle, src/Straggling.jl:446-517, :1344-1405
__device__ __noinline__ static double scattering_angle_f(const DevTables& T, const DevMix& x,
                                                         const Loc& Lk, Philox& rng) {
  const double mass = T.mass;
  const double k = Lk.e;
  const double p2 = k * (k + 2.0 * mass);
  const double mu0 = fmax(mix_sum_L(T, x, A_SCREENING, Lk), 1e-12);
  double zoa = 0.0;
  for (int j = 0; j < x.n; ++j) zoa += (x.n == 1 ? 1.0 : x.frac[j]) * T.scal[x.mat[j] * 4 + 2];
  const double At = zoa > 0 ? 1.0 / zoa : 22.0;
  const double Mt = At * dpm::NEUTRON_MASS;
  const double s = mass * mass + Mt * Mt + 2.0 * (k + mass) * Mt;
  const double k0 = (s - (mass + Mt) * (mass + Mt)) / (2.0 * sqrt(s));
  const double Ecm = k0 + mass;
  const double pcm = sqrt(fmax(k0 * (k0 + 2.0 * mass), 0.0));
  const double Elab = k + mass;
  const double plab = sqrt(fmax(p2, 0.0));
  const double gCM = (plab > 0 && pcm > 0) ? (Elab * Ecm + plab * pcm) / (mass * sqrt(s)) : 1.0;
  const double tau = (plab > 0 && pcm > 0) ? (Elab * pcm - Ecm * plab) / (plab * pcm) : 0.0;
  const double A = mu0 / 4.0;
  const double fspin = mix_sum_L(T, x, A_SPIN_FACTOR, Lk);
  const double Apm = A + mu0, Ap1 = A + 1.0, omm = 1.0 - mu0;
  double mu1 = mu0;
  for (int t = 0; t < 100; ++t) {
    const double zeta = uniform(rng);
    const double tmp = Ap1 - zeta * omm;
    if (tmp <= 0) continue;
    double trial = Apm * Ap1 / tmp - A;
    trial = fmin(fmax(trial, mu0), 1.0);
    if (uniform(rng) <= 1.0 - fspin * trial) { mu1 = trial; break; }
  }
  double mu;
  if (mu1 > 1e-6) {
    const double a = gCM * (tau + 1.0 - 2.0 * mu1);
    const double cth = a / sqrt(4.0 * mu1 * (1.0 - mu1) + a * a);
    mu = 0.5 * (1.0 - cth);
  } else {
    const double d = gCM * (1.0 + tau);
    mu = mu1 / (d * d);
  }
  return fmin(fmax(mu, 0.0), 1.0);
}

// sample_soft_scattering, src/Straggling.jl:539-626, :1204-1250; lb0 is the
// transport path at the step's initial energy.
__device__ __forceinline__ double soft_scattering_f(const DevTables& T, const DevMix& x,
                                                    double lb0, const Loc& L1, double step,
                                                    double density, Philox& rng) {
  if (step <= 0 || density <= 0) return 0.0;
  const double i0 = lb0 > 0 ? density / lb0 : 0.0;
  double i1;
  if (L1.e > 0) {
    const double l1 = mix_transport_path_L(T, x, L1);
    i1 = l1 > 0 ? density / l1 : 0.0;
  } else {
    i1 = i0;
  }
  if (i0 <= 0 && i1 <= 0) return 0.0;
  double ilb1 = 0.25 * step * (i0 + i1);
  if (ilb1 < 1e-12) return 0.0;
  if (ilb1 > 1.0) ilb1 = 1.0;
  for (int t = 0; t < 1000; ++t) {
    const double u = uniform(rng);
    if (u < 1e-300) continue;
    const double mu = -ilb1 * log(u);
    if (mu <= 1.0) return fmax(mu, 0.0);
  }
  return ilb1;
}

// transport_backward_step_full, mode = :straggled, src/Geometry.jl:148-357.
// Li locates st.E; lb0 is the transport path there (scattering only).  On
// return *gnext guesses the interval of the new energy.
__device__ __noinline__ static void straggled_step_f(const DevTables& T, const Medium& md,
                                                     McState& st, const Loc& Li, double lb0,
                                                     double density, double ds, bool scattering,
                                                     double energy_limit, Philox& rng, int& err,
                                                     int* gnext, DelLog* dlog) {
  const DevMix& x = md.mix;
  const DevKernel& K = md.kern;
  const double ki = Li.e, mass = T.mass;
  double dX = ds * density;
  const double Xi = mix_range_L(T, x, true, Li);
  double kf = ki, ratio = 1.0;
  if (ki > 0.001) {
    fluctuate_f(T, x, Li, dX, Xi, rng, kf, ratio);
    kf = fmax(kf, ki);
  } else {
    kf = mix_kinetic_g(T, x, true, Xi + dX, Li.i);
  }
  if (energy_limit > 0 && kf >= energy_limit) {
    const double dk = fabs(ki - kf);
    if (dk > 0) {
      dX = dX * (fabs(energy_limit - ki) / dk);
      kf = energy_limit;
    }
  }
  bool del = false, ehs = false;
  double X_del = __builtin_huge_val(), k_del = 0, X_ehs = __builtin_huge_val(), k_ehs = 0;
  Loc Le;
  const int kg = kguess(K, Li);  // kernel-grid interval guess near ki
  {
    const double rate = fmax(klookup_g(ki, K.e, K.sigma, K.n, K.pow2, kg, err),
                             klookup_g(kf, K.e, K.sigma, K.n, K.pow2, kg, err));
    if (err) return;
    if (rate > 0) {
      const double cand = exponential(rng) / rate;
      if (cand < dX) {
        const double kh = mix_kinetic_g(T, x, true, Xi + cand, Li.i);
        const double pre = ki + ratio * fabs(kh - ki);
        if (pre > 0 && uniform(rng) <= klookup_g(pre, K.e, K.sigma, K.n, K.pow2, kg, err) / rate) {
          del = true; X_del = cand; k_del = pre;
        }
        if (err) return;
      }
    }
  }
  if (scattering) {
    double xe, ke;
    if (sample_ehs_f(T, x, Li, kf, dX, Xi, ratio, rng, xe, ke, Le) && ke > 0) {
      ehs = true; X_ehs = xe; k_ehs = ke;
    }
  }
  if (del && ehs) {
    if (X_del <= X_ehs) ehs = false; else del = false;
  }
  double dw = 1.0, ckf = kf;
  int proc = 0, elem = -1;
  if (scattering && dlog && dlog->coop) {
    // every lane of this step call takes part in the component evaluation;
    // the call is skipped (uniformly) when no lane has a DEL event, where it
    // would draw nothing and return nothing that is used
    DelOut o{k_del, 0.0, 0, -1};
    if (__ballot(del)) o = coop_del_event(*dlog, dlog->medium, mass, k_del, kg, rng, del, err);
    if (del) {
      if (err) return;
      ckf = k_del;
      kf = o.k; dw = o.w; proc = o.proc; elem = o.elem;
      dX = X_del;
    }
  } else if (del) {
    ckf = k_del;
    if (!scattering && dlog) {
      // the target choice only consumes its draw when nothing is deflected
      const DelSample d = del_sample(K, mass, k_del, kg, rng, err);
      if (err) return;
      kf = d.anc;
      if (d.deferred) {
        if (!del_push(dlog, d, k_del, 1)) dw = del_total(K, mass, d.anc, d.anc - k_del) / d.denom;
        (void)uniform(rng);
      } else {
        dw = d.w;
      }
    } else {
      const DelOut o = reverse_del_event(K, mass, k_del, rng, true, err);
      if (err) return;
      kf = o.k; dw = o.w; proc = o.proc; elem = o.elem;
    }
    dX = X_del;
  }
  if (ehs) { kf = k_ehs; ckf = k_ehs; dX = X_ehs; }
  const Loc Lc = ehs ? Le : loc_e(T, ckf, log(ckf), Li.i);
  const double step = dX / density;
  double d[3] = {st.dir[0], st.dir[1], st.dir[2]};
  if (scattering) {
    if (del && proc > 0) {
      double mu;
      const double Z = K.el[4 * elem];
      if (proc == 1 || proc == 2) mu = polar_brems(mass, kf, ckf, Z, rng);
      else if (proc == 3) mu = polar_photo(mass, kf, ckf, rng);
      else mu = polar_ionisation(mass, kf, ckf);
      if (mu > 0) rotate(d, mu, rng);
    }
    if (ehs) {
      const double mu = scattering_angle_f(T, x, Le, rng);
      if (mu > 0) rotate(d, mu, rng);
    }
    const double ms = soft_scattering_f(T, x, lb0, Lc, step, density, rng);
    if (ms > 0) rotate(d, ms, rng);
  }
  const double dt = step * mass / 2 *
                    (1.0 / sqrt(ki * (ki + 2 * mass)) + 1.0 / sqrt(ckf * (ckf + 2 * mass)));
  for (int a = 0; a < 3; ++a) st.pos[a] -= step * st.dir[a];
  st.L += step;
  const double wf = mix_sum_L(T, x, A_MIXED_STOP, Lc) / mix_sum_L(T, x, A_MIXED_STOP, Li);
  st.w = st.w * wf * exp(-dt / T.ctau) * dw;
  st.t += dt;
  st.E = kf;
  for (int a = 0; a < 3; ++a) st.dir[a] = d[a];
  *gnext = Lc.i;
}

// transport_mixed_step, src/MixedTransport.jl:144-200, with the kernel-grid
// interval of the current energy carried between lookups.
__device__ __noinline__ static void mixed_step_f(const DevTables& T, const Medium& md, McState& st,
                                                 const Loc& Li, double density, double distance,
                                                 double energy_limit, Philox& rng, int& err,
                                                 DelLog* dlog) {
  const DevKernel& c = md.kern;
  const double m = T.mass;
  const int n = c.n, p2 = c.pow2;
  const double emax = c.e[n - 1];
  if (!(density > 0 && distance >= 0) || energy_limit > emax) { err |= ERR_TRANSPORT; return; }
  if (st.E > emax) { err |= ERR_DOMAIN; return; }
  double k = st.E, remaining = distance * density, travelled = 0, proper = 0, weight = st.w;
  const double ceiling = energy_limit > 0 ? fmin(energy_limit, emax) : emax;
  const double tiny = ulp(fmax(1.0, distance * density));
  const double rceil = klookup_g(ceiling, c.e, c.range, n, p2, n - 2, err);
  int ik = kguess(c, Li);  // interval of k on c.e
  bool exhausted = true;
  for (int ev = 0; ev < 100000; ++ev) {
    if (remaining <= tiny) { exhausted = false; break; }
    if (k >= ceiling) { exhausted = false; break; }
    if (!kloc(k, c.e, n, p2, ik, ik, err)) return;
    const double ri = kval(k, c.e, c.range, ik);
    const double hi = kval(k, c.e, c.hazard, ik);
    const double rtc = fmax(0.0, rceil - ri);
    const double eg = fmin(remaining, rtc);
    double endpoint;
    if (eg == rtc) {
      endpoint = ceiling;
    } else {
      int ir;
      if (!kloc(ri + eg, c.range, n, p2, ik, ir, err)) return;
      endpoint = fmax(k, kval(ri + eg, c.range, c.e, ir));
    }
    const double eventh = hi + exponential(rng);
    int iend;
    if (!kloc(endpoint, c.e, n, p2, ik, iend, err)) return;
    const bool occurs = eventh < kval(endpoint, c.e, c.hazard, iend);
    double pre;
    int ip;
    if (occurs) {
      int ih;
      if (!kloc(eventh, c.hazard, n, p2, ik, ih, err)) return;
      pre = fmax(k, kval(eventh, c.hazard, c.e, ih));
      if (!kloc(pre, c.e, n, p2, ik, ip, err)) return;
    } else {
      pre = endpoint;
      ip = iend;
    }
    const double dx = occurs ? fmax(0.0, kval(pre, c.e, c.range, ip) - ri) : eg;
    double dt;
    if (dx <= 8 * ulp(ri) || pre - k <= 8 * ulp(k))
      dt = dx / density * m / 2 * (1.0 / sqrt(k * (k + 2 * m)) + 1.0 / sqrt(pre * (pre + 2 * m)));
    else
      dt = (kval(pre, c.e, c.proper, ip) - kval(k, c.e, c.proper, ik)) / density;
    weight *= kval(pre, c.e, c.stop, ip) / kval(k, c.e, c.stop, ik) * exp(-dt / T.ctau);
    travelled += dx;
    remaining = fmax(0.0, remaining - dx);
    proper += dt;
    k = pre;
    ik = ip;
    if (occurs) {
      if (dlog) {
        const DelSample d = del_sample(c, m, k, ik, rng, err);
        if (err) return;
        if (d.deferred && del_push(dlog, d, k, 0)) {
          k = d.anc;
          continue;
        }
        const double wd = d.deferred ? del_total(c, m, d.anc, d.anc - k) / d.denom : d.w;
        k = d.anc;
        weight *= wd;
      } else {
        const DelOut o = reverse_del_event(c, m, k, rng, false, err);
        if (err) return;
        k = o.k;
        weight *= o.w;
      }
      if (!(isfinite(weight) && weight >= 0)) { err |= ERR_TRANSPORT; return; }
      if (weight == 0) { exhausted = false; break; }
    } else {
      exhausted = false;
      break;
    }
  }
  if (exhausted) { err |= ERR_TRANSPORT; return; }
  if (weight > 0 && k >= emax && remaining > tiny && (energy_limit < 0 || k < energy_limit)) {
    err |= ERR_DOMAIN;
    return;
  }
  const double dd = travelled / density;
  st.L += dd;
  st.E = k;
  st.t += proper;
  st.w = weight;
  for (int a = 0; a < 3; ++a) st.pos[a] -= dd * st.dir[a];
}

// transport_backward_step (deterministic CSDA) on a mixture, src/Geometry.jl:424-457
__device__ __forceinline__ void csda_mix_step_f(const DevTables& T, const Medium& md, McState& st,
                                                const Loc& Li, double density, double ds,
                                                int* gnext) {
  const DevMix& x = md.mix;
  const double ki = Li.e, gs = ds * density;
  const double Xi = mix_range_L(T, x, false, Li);
  const double kf = mix_kinetic_g(T, x, false, Xi + gs, Li.i);
  const double gi = 1.0 + ki / T.mass, gc = 1.0 + kf / T.mass, ga = (gi + gc) / 2.0;
  const double beta = sqrt(fmax(1.0 - 1.0 / (ga * ga), 1e-12));
  const double dt = ds / (beta * ga);
  for (int a = 0; a < 3; ++a) st.pos[a] -= ds * st.dir[a];
  st.L += ds;
  const Loc Lf = loc_e(T, kf, log(kf), Li.i);
  const double wf = mix_sum_L(T, x, A_CSDA_STOP, Lf) / mix_sum_L(T, x, A_CSDA_STOP, Li);
  st.w = st.w * wf * exp(-dt / T.ctau);
  st.t += dt;
  st.E = kf;
  *gnext = Lf.i;
}

}  // namespace dpk

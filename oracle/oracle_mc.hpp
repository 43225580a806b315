// CPU oracle: straggled / mixed backward Monte Carlo of the flat two-layer
// geometry with runtime material mixtures.  TEST INFRASTRUCTURE ONLY (see
// oracle_core.hpp).  Scalar restatement of
//   src/Geometry.jl:135-357, 534-672, 694-757, 818-896 (stepping, driver)
//   src/Straggling.jl (fluctuation, EHS, scattering, rotation, DEL angles)
//   src/MixedTransport.jl (mixed kernel, reverse DEL event, mixed step)
//   src/Physics.jl:996-1365 (mixture properties and integrated tables)
// with the event's draws taken from Philox4x32-10 keyed by (seed, event).
#pragma once
#include "oracle_core.hpp"
#include "oracle_philox.hpp"
#include "oracle_dcs.hpp"
#include <algorithm>
#include <stdexcept>
#include <vector>

namespace orc {

constexpr double EHS_PATH_MAX = 1e9;
constexpr double ELECTRON_MASS_ = 0.510998910e-3;
constexpr double PROTON_MASS_ = 0.938272;
constexpr double NEUTRON_MASS_ = 0.939565;
constexpr double PION_MASS_ = 0.13957018;
constexpr double HBAR_C_ = 1.973269804e-16;
constexpr double PI_ = 3.141592653589793;

struct ElementRec { double Z, A, I, f; };

// A runtime mixture of tabulated materials (MaterialMixture, src/Types.jl:357-391)
// with its integrated range / proper-time tables (MixtureTable,
// src/Physics.jl:1049-1103) built on the grid of material 0.
struct Mixture {
  std::vector<int> mat;
  std::vector<double> frac;
  std::vector<double> e, le, csda_range, mixed_range, time;  // only when mat.size() > 1
  bool single() const { return mat.size() == 1; }
};

inline Mixture make_mixture(const Tables& T, std::vector<int> mats, std::vector<double> fr) {
  Mixture m;
  double tot = 0;
  for (double f : fr) tot += f;
  for (double& f : fr) f = tot > 0 ? f / tot : f;
  m.mat = mats;
  m.frac = fr;
  if (m.single()) return m;
  const int n = T.K;
  const double* E = T.arr(0, T_ENERGY);
  m.e.assign(E, E + n);
  m.le.assign(T.arr(0, T_LOG_ENERGY), T.arr(0, T_LOG_ENERGY) + n);
  m.csda_range.assign(n, 0.0);
  m.mixed_range.assign(n, 0.0);
  m.time.assign(n, 0.0);
  auto stop = [&](bool mixed, double k) {
    double v = 0;
    for (size_t j = 0; j < m.mat.size(); ++j) v += m.frac[j] * prop_stop(T, mixed, m.mat[j], k);
    return v;
  };
  double cp = stop(false, E[0]), mp = stop(true, E[0]);
  m.csda_range[0] = E[0] / cp;
  m.mixed_range[0] = E[0] / mp;
  for (int i = 1; i < n; ++i) {
    const double K = E[i], Kp = E[i - 1], dK = K - Kp;
    const double c = stop(false, K), mm = stop(true, K);
    m.csda_range[i] = m.csda_range[i - 1] + dK / ((c + cp) / 2);
    m.mixed_range[i] = m.mixed_range[i - 1] + dK / ((mm + mp) / 2);
    const double g = 1.0 + K / T.mass, gp = 1.0 + Kp / T.mass;
    const double b = std::sqrt(1.0 - 1.0 / (g * g)), bp = std::sqrt(1.0 - 1.0 / (gp * gp));
    const double avg = (c + cp) / 2;
    m.time[i] = m.time[i - 1] + dK / (avg * ((b + bp) / 2) * ((g + gp) / 2));
    cp = c;
    mp = mm;
  }
  return m;
}

// Mixture property overloads, src/Physics.jl:1152-1323.
inline double mix_stop(const Tables& T, bool mixed, const Mixture& m, double k) {
  if (m.single()) return prop_stop(T, mixed, m.mat[0], k);
  double v = 0;
  for (size_t j = 0; j < m.mat.size(); ++j) v += m.frac[j] * prop_stop(T, mixed, m.mat[j], k);
  return v;
}
inline double mix_prop_sum(const Tables& T, int a, const Mixture& m, double k) {
  if (m.single()) return property(T, m.mat[0], a, k);
  double v = 0;
  for (size_t j = 0; j < m.mat.size(); ++j) v += m.frac[j] * property(T, m.mat[j], a, k);
  return v;
}
inline double mix_range(const Tables& T, bool mixed, const Mixture& m, double k) {
  if (m.single()) return prop_range(T, mixed, m.mat[0], k);
  return interp(m.e.data(), m.le.data(), mixed ? m.mixed_range.data() : m.csda_range.data(),
                (int)m.e.size(), k, std::log(k));
}
inline double mix_kinetic(const Tables& T, bool mixed, const Mixture& m, double x) {
  if (m.single()) return prop_kinetic(T, mixed, m.mat[0], x);
  return inverse_range(m.e.data(), m.le.data(), mixed ? m.mixed_range.data() : m.csda_range.data(),
                       (int)m.e.size(), x);
}
inline double mix_transport_path(const Tables& T, const Mixture& m, double k) {
  if (m.single()) return property(T, m.mat[0], T_TRANSPORT_PATH, k);
  double inv = 0;
  for (size_t j = 0; j < m.mat.size(); ++j) {
    double lp = property(T, m.mat[j], T_TRANSPORT_PATH, k);
    if (lp > 0) inv += m.frac[j] / lp;
  }
  return inv > 0 ? 1.0 / inv : 1e9;
}
inline double mix_screening(const Tables& T, const Mixture& m, double k) {
  if (m.single()) return std::max(property(T, m.mat[0], T_SCREENING, k), 1e-12);
  double v = 0;
  for (size_t j = 0; j < m.mat.size(); ++j) v += m.frac[j] * property(T, m.mat[j], T_SCREENING, k);
  return std::max(v, 1e-12);
}
inline double mix_zoa(const Tables& T, const Mixture& m) {
  if (m.single()) return T.zoa(m.mat[0]);
  double v = 0;
  for (size_t j = 0; j < m.mat.size(); ++j) v += m.frac[j] * T.zoa(m.mat[j]);
  return v;
}

// compute_ehs_mean_free_path, src/Straggling.jl:330-360 and :1025-1038
inline double ehs_path_single(const Tables& T, int mat, double k, double elastic_ratio) {
  if (k < 1e-3) return EHS_PATH_MAX;
  const double Lb = property(T, mat, T_ELASTIC_PATH, k);
  const double p2 = k * (k + 2.0 * T.mass);
  if (Lb > 0 && Lb < EHS_PATH_MAX * p2) return Lb / p2;
  const double tp = property(T, mat, T_TRANSPORT_PATH, k);
  if (tp > 0) return std::min(tp / elastic_ratio, EHS_PATH_MAX);
  return EHS_PATH_MAX;
}
inline double ehs_path(const Tables& T, const Mixture& m, double k, double er) {
  if (m.single()) return ehs_path_single(T, m.mat[0], k, er);
  double inv = 0;
  for (size_t j = 0; j < m.mat.size(); ++j) {
    double lp = ehs_path_single(T, m.mat[j], k, er);
    if (lp > 0 && lp < EHS_PATH_MAX) inv += m.frac[j] / lp;
  }
  return inv > 0 ? 1.0 / inv : EHS_PATH_MAX;
}

// ------------------------------------------------------------- MixedKernel
// _prepare_mixed_kernel, src/MixedTransport.jl:55-97
struct MixedKernel {
  double mass, ctau, cutoff, onset;
  std::vector<double> e, stop, sigma, range, hazard, proper;
  std::vector<ElementRec> el;
};

inline MixedKernel prepare_mixed_kernel(const Tables& T, const Mixture& mx,
                                        const std::vector<std::vector<ElementRec>>& recipes,
                                        double cutoff, int refinement = 8) {
  MixedKernel c;
  c.mass = T.mass; c.ctau = T.ctau; c.cutoff = cutoff;
  double lo = -INFINITY, hi = INFINITY;
  for (int i : mx.mat) {
    lo = std::max(lo, T.arr(i, T_ENERGY)[0]);
    hi = std::min(hi, T.arr(i, T_ENERGY)[T.K - 1]);
  }
  std::vector<double> grid;
  for (int i : mx.mat)
    for (int k = 0; k < T.K; ++k) {
      double e = T.arr(i, T_ENERGY)[k];
      if (lo <= e && e <= hi) grid.push_back(e);
    }
  std::sort(grid.begin(), grid.end());
  grid.erase(std::unique(grid.begin(), grid.end()), grid.end());
  for (size_t j = 0; j + 1 < grid.size(); ++j) {
    const double a = std::log(grid[j]), b = std::log(grid[j + 1]);
    for (int r = 0; r < refinement; ++r) {
      // range(log a, log b; length = refinement + 1) without its last point
      const double t = a + (b - a) * r / refinement;
      c.e.push_back(std::exp(t));
    }
  }
  c.e.push_back(hi);
  c.e[0] = lo;
  const size_t n = c.e.size();
  c.stop.resize(n); c.sigma.resize(n);
  for (size_t j = 0; j < n; ++j) {
    double s = 0, x = 0;
    for (size_t q = 0; q < mx.mat.size(); ++q) {
      s += mx.frac[q] * prop_stop(T, true, mx.mat[q], c.e[j]);
      x += mx.frac[q] * property(T, mx.mat[q], T_CROSS_SECTION, c.e[j]);
    }
    c.stop[j] = s;
    c.sigma[j] = x;
  }
  c.onset = INFINITY;
  for (double g : grid) {
    double x = 0;
    for (size_t q = 0; q < mx.mat.size(); ++q) x += mx.frac[q] * property(T, mx.mat[q], T_CROSS_SECTION, g);
    if (x > 0) {
      c.onset = g;
      for (size_t j = 0; j < n; ++j)
        if (c.e[j] < g) c.sigma[j] = x;
      break;
    }
  }
  c.range.assign(n, 0.0); c.hazard.assign(n, 0.0); c.proper.assign(n, 0.0);
  for (size_t j = 1; j < n; ++j) {
    const double a = c.e[j - 1], b = c.e[j], d = b - a;
    c.range[j] = c.range[j - 1] + d / 2 * (1 / c.stop[j - 1] + 1 / c.stop[j]);
    c.hazard[j] = c.hazard[j - 1] + d / 2 * (c.sigma[j - 1] / c.stop[j - 1] + c.sigma[j] / c.stop[j]);
    c.proper[j] = c.proper[j - 1] + d / 2 * T.mass *
                                        (1 / (std::sqrt(a * (a + 2 * T.mass)) * c.stop[j - 1]) +
                                         1 / (std::sqrt(b * (b + 2 * T.mass)) * c.stop[j]));
  }
  for (size_t q = 0; q < mx.mat.size(); ++q)
    for (const ElementRec& e : recipes[mx.mat[q]]) c.el.push_back({e.Z, e.A, e.I, mx.frac[q] * e.f});
  return c;
}

struct DomainErr : std::runtime_error { using std::runtime_error::runtime_error; };

// lookup, src/MixedTransport.jl:43-49 (natural-variable linear interpolation)
inline double klookup(double x, const std::vector<double>& xs, const std::vector<double>& ys) {
  if (x < xs.front()) throw DomainErr("below prepared mixed kernel domain");
  if (x > xs.back()) throw DomainErr("above prepared mixed kernel domain");
  int i = (int)(std::upper_bound(xs.begin(), xs.end(), x) - xs.begin());  // searchsortedlast (1-based)
  i = std::min(std::max(i, 1), (int)xs.size() - 1) - 1;
  const double d = xs[i + 1] - xs[i];
  return d == 0 ? ys[i] : ys[i] + (x - xs[i]) / d * (ys[i + 1] - ys[i]);
}

// differential_components, src/MixedTransport.jl:99-106: value of each
// (element, process) pair; process 1 brems, 2 pair, 3 photonuclear, 4 ionisation
inline double diff_component(const MixedKernel& c, const ElementRec& e, int proc, double k, double q) {
  double raw;
  switch (proc) {
    case 1: raw = dcs_brems(e.Z, e.A, c.mass, k, q); break;
    case 2: raw = dcs_pair(e.Z, e.A, c.mass, k, q); break;
    case 3: raw = dcs_photo(e.Z, e.A, c.mass, k, q); break;
    default: raw = dcs_ionisation(e.Z, e.I, c.mass, k, q); break;
  }
  return e.f * 6.02214076e23 / (e.A * .001) * k * raw;
}

struct DelOut { double k, w; int proc; int elem; };

// reverse_del_event, src/MixedTransport.jl:111-142 (details selects the
// target element and process with one more draw)
inline DelOut reverse_del_event(const MixedKernel& c, double k, Philox4x32& rng, bool details = true) {
  const double sigma = klookup(k, c.e, c.sigma);
  if (!(sigma > 0)) throw std::runtime_error("Cannot sample a zero-rate hard event");
  const double pcel = k >= c.onset ? 0.0
                                   : -std::log(c.cutoff) /
                                         (-std::log(c.cutoff) + std::log(c.onset / (c.onset - k)));
  if (pcel > 0 && rng.uniform() < pcel) return {k, 1 / pcel, 0, -1};
  const double me = ELECTRON_MASS_;
  const double xmax = k < .5 * (c.mass - me) * (c.mass - me) / me
                          ? std::min(1.0, 2 * me * (k + 2 * c.mass) / ((c.mass + me) * (c.mass + me)))
                          : 1.0;
  const double xmin = k >= c.onset * (1 - c.cutoff) ? c.cutoff : 1 - k / c.onset;
  if (!(xmax > xmin)) return {k, 1.0, 0, -1};
  const double invnorm = 1 / xmin - 1 / xmax;
  const double nu = 1 / (1 / xmin - rng.uniform() * invnorm);
  const double r = 1 - nu;
  const double anc = k / r;
  const double pdf = 1 / (invnorm * (nu * nu));
  double total = 0;
  const int ne = (int)c.el.size();
  std::vector<double> vals(ne * 4);
  for (int i = 0; i < ne; ++i)
    for (int p = 1; p <= 4; ++p) {
      vals[i * 4 + p - 1] = diff_component(c, c.el[i], p, anc, anc - k);
      total += vals[i * 4 + p - 1];
    }
  const double w = total / (pdf * r * sigma * (1 - pcel));
  if (!details || total <= 0) return {anc, w, 0, -1};
  double u = rng.uniform() * total;
  for (int i = 0; i < ne; ++i)
    for (int p = 1; p <= 4; ++p) {
      u -= vals[i * 4 + p - 1];
      if (u <= 0) return {anc, w, p, i};
    }
  return {anc, w, 4, ne - 1};
}

// ------------------------------------------------------------ samplers
// box_muller_randn / truncated_randn, src/Straggling.jl:59-78
inline double box_muller(Philox4x32& rng) {
  const double r = std::sqrt(-2.0 * std::log(rng.uniform()));
  const double phi = 2.0 * PI_ * rng.uniform();
  return r * std::sin(phi);
}
inline double truncated_randn(Philox4x32& rng) {
  double u = box_muller(rng);
  while (std::fabs(u) > 3.0) u = box_muller(rng);
  return u * 0.9848416472;
}

// fluctuate_energy_loss (backward), src/Straggling.jl:103-173, :1046-1105
inline void fluctuate(const Tables& T, const Mixture& m, double ki, double dX, double Xtot,
                      Philox4x32& rng, double& kout, double& ratio) {
  const double sgn = -1.0;
  double k1 = mix_kinetic(T, true, m, Xtot - sgn * dX);
  const double dk0 = std::fabs(ki - k1);
  ratio = 1.0;
  if (k1 > 0 && dk0 > 0) {
    const double O0 = mix_prop_sum(T, T_STRAGGLING, m, ki);
    const double O1 = mix_prop_sum(T, T_STRAGGLING, m, k1);
    double dk12 = 0.5 * dX * (O0 + O1);
    if (dX / Xtot > 0.05) {
      const double de1 = mix_stop(T, true, m, k1), de0 = mix_stop(T, true, m, ki);
      dk12 *= 1.0 + sgn * dX / dk0 * (de1 - de0);
    }
    const double dk1 = std::sqrt(std::max(dk12, 0.0));
    if (dk0 >= 3.0 * dk1 && dk1 > 0) {
      k1 += truncated_randn(rng) * dk1;
    } else if (dk0 >= 1.7320508 * dk1 && dk1 > 0) {
      k1 += 1.7320508 * (1.0 - 2.0 * rng.uniform()) * dk1;
    } else if (dk1 > 0) {
      const double dk32 = 3.0 * dk12, dk02 = dk0 * dk0;
      const double a = 1.0 - (dk32 - dk02) / (dk32 + 3.0 * dk02);
      if (rng.uniform() <= a) {
        const double b = 0.5 * (dk32 + 3.0 * dk02) / dk0;
        k1 = ki - sgn * b * rng.uniform();
        if (k1 < 0) k1 = 0;
      } else {
        k1 = ki;
      }
    }
    ratio = std::fabs(ki - k1) / dk0;
  }
  if (ratio == 0.0) ratio = 1.0;
  kout = std::max(k1, 0.0);
}

// sample_ehs_event (backward), src/Straggling.jl:373-434, :1284-1334
inline bool sample_ehs(const Tables& T, const Mixture& m, double ki, double kf, double dX,
                       double Xtot, double ratio, double er, Philox4x32& rng, double& X_ehs,
                       double& k_ehs) {
  const double sgn = -1.0;
  double kmin = std::min(ki, kf), kmax = std::max(ki, kf);
  if (kmin < 0.001) std::swap(kmin, kmax);
  const double lb = ehs_path(T, m, kmin, er);
  if (lb <= 0 || lb >= EHS_PATH_MAX) return false;
  const double zeta = rng.uniform();
  if (zeta <= 0 || zeta >= 1) return false;
  X_ehs = -std::log(zeta) * lb;
  if (X_ehs >= dX) return false;
  const double kh = mix_kinetic(T, true, m, Xtot - sgn * X_ehs);
  k_ehs = ki - sgn * ratio * std::fabs(ki - kh);
  if (k_ehs <= 0) return false;
  const double lbk = ehs_path(T, m, k_ehs, er);
  if (lbk <= 0 || lbk >= EHS_PATH_MAX) return false;
  const double r = lbk / lb;
  if (r <= 0 || rng.uniform() > 1.0 / r) return false;
  return true;
}

// sample_scattering_angle, src/Straggling.jl:446-517, :1344-1405, in two
// parts: the Wentzel draw with spin rejection in the centre-of-mass frame,
// then the transformation of that angle to the laboratory frame.
struct WentzelKin {
  double mu0, A, fspin, gCM, tau;
};
inline WentzelKin wentzel_kinematics(const Tables& T, const Mixture& m, double k) {
  const double mass = T.mass;
  const double p2 = k * (k + 2.0 * mass);
  WentzelKin w;
  w.mu0 = mix_screening(T, m, k);
  const double zoa = mix_zoa(T, m);
  const double At = zoa > 0 ? 1.0 / zoa : 22.0;
  const double Mt = At * NEUTRON_MASS_;
  const double s = mass * mass + Mt * Mt + 2.0 * (k + mass) * Mt;
  const double k0 = (s - (mass + Mt) * (mass + Mt)) / (2.0 * std::sqrt(s));
  const double Ecm = k0 + mass;
  const double pcm = std::sqrt(std::max(k0 * (k0 + 2.0 * mass), 0.0));
  const double Elab = k + mass;
  const double plab = std::sqrt(std::max(p2, 0.0));
  w.gCM = (plab > 0 && pcm > 0) ? (Elab * Ecm + plab * pcm) / (mass * std::sqrt(s)) : 1.0;
  w.tau = (plab > 0 && pcm > 0) ? (Elab * pcm - Ecm * plab) / (plab * pcm) : 0.0;
  w.A = w.mu0 / 4.0;
  w.fspin = mix_prop_sum(T, T_SPIN_FACTOR, m, k);
  return w;
}
inline double wentzel_cm(const WentzelKin& w, Philox4x32& rng) {
  const double A = w.A, mu0 = w.mu0;
  const double Apm = A + mu0, Ap1 = A + 1.0, omm = 1.0 - mu0;
  double mu1 = mu0;
  for (int t = 0; t < 100; ++t) {
    const double zeta = rng.uniform();
    const double tmp = Ap1 - zeta * omm;
    if (tmp <= 0) continue;
    double trial = Apm * Ap1 / tmp - A;
    trial = std::min(std::max(trial, mu0), 1.0);
    if (rng.uniform() <= 1.0 - w.fspin * trial) { mu1 = trial; break; }
  }
  return mu1;
}
inline double wentzel_lab(const WentzelKin& w, double mu1) {
  double mu;
  if (mu1 > 1e-6) {
    const double a = w.gCM * (w.tau + 1.0 - 2.0 * mu1);
    const double cth = a / std::sqrt(4.0 * mu1 * (1.0 - mu1) + a * a);
    mu = 0.5 * (1.0 - cth);
  } else {
    const double d = w.gCM * (1.0 + w.tau);
    mu = mu1 / (d * d);
  }
  return std::min(std::max(mu, 0.0), 1.0);
}
inline double scattering_angle(const Tables& T, const Mixture& m, double k, Philox4x32& rng) {
  const WentzelKin w = wentzel_kinematics(T, m, k);
  return wentzel_lab(w, wentzel_cm(w, rng));
}

// sample_soft_scattering, src/Straggling.jl:539-626, :1204-1250
inline double soft_scattering(const Tables& T, const Mixture& m, double k0, double k1, double step,
                              double density, Philox4x32& rng) {
  if (step <= 0 || density <= 0) return 0.0;
  const double l0 = mix_transport_path(T, m, k0);
  const double i0 = l0 > 0 ? density / l0 : 0.0;
  double i1;
  if (k1 > 0) {
    const double l1 = mix_transport_path(T, m, k1);
    i1 = l1 > 0 ? density / l1 : 0.0;
  } else {
    i1 = i0;
  }
  if (i0 <= 0 && i1 <= 0) return 0.0;
  double ilb1 = 0.25 * step * (i0 + i1);
  if (ilb1 < 1e-12) return 0.0;
  if (ilb1 > 1.0) ilb1 = 1.0;
  for (int t = 0; t < 1000; ++t) {
    const double u = rng.uniform();
    if (u < 1e-300) continue;
    const double mu = -ilb1 * std::log(u);
    if (mu <= 1.0) return std::max(mu, 0.0);
  }
  return ilb1;
}

// rotate_direction, src/Straggling.jl:645-710
inline void rotate(double d[3], double mu, Philox4x32& rng) {
  const double ct = 1.0 - 2.0 * mu;
  const double st2 = 4.0 * mu * (1.0 - mu);
  if (st2 <= 0) return;
  const double st = std::sqrt(st2);
  const double ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
  double u0x = 0, u0y = 0, u0z = 0;
  if (ax > ay) {
    if (ax > az) {
      const double nrm = 1.0 / std::sqrt(d[0] * d[0] + d[2] * d[2]);
      u0x = -d[2] * nrm; u0z = d[0] * nrm;
    } else {
      const double nrm = 1.0 / std::sqrt(d[1] * d[1] + d[2] * d[2]);
      u0y = d[2] * nrm; u0z = -d[1] * nrm;
    }
  } else {
    if (ay > az) {
      const double nrm = 1.0 / std::sqrt(d[0] * d[0] + d[1] * d[1]);
      u0x = d[1] * nrm; u0y = -d[0] * nrm;
    } else {
      const double nrm = 1.0 / std::sqrt(d[1] * d[1] + d[2] * d[2]);
      u0y = d[2] * nrm; u0z = -d[1] * nrm;
    }
  }
  const double u1x = u0y * d[2] - u0z * d[1];
  const double u1y = u0z * d[0] - u0x * d[2];
  const double u1z = u0x * d[1] - u0y * d[0];
  const double phi = PI_ * (1.0 - 2.0 * rng.uniform());
  const double cp = std::cos(phi), sp = std::sin(phi);
  double nx = ct * d[0] + st * (cp * u0x + sp * u1x);
  double ny = ct * d[1] + st * (cp * u0y + sp * u1y);
  double nz = ct * d[2] + st * (cp * u0z + sp * u1z);
  const double nrm = std::sqrt(nx * nx + ny * ny + nz * nz);
  if (nrm > 0) { nx /= nrm; ny /= nrm; nz /= nrm; }
  d[0] = nx; d[1] = ny; d[2] = nz;
}

// nuclear_radius of src/Straggling.jl:726-729 (table entry Z, 1-based)
inline double straggling_nuclear_radius(double Z) {
  static const double R[120] = {
      2.098, 0.858, 1.680, 2.400, 2.518, 2.405, 2.470, 2.548, 2.734, 2.900, 2.993, 2.940, 3.043,
      3.035, 3.098, 3.187, 3.245, 3.360, 3.413, 3.408, 3.477, 3.443, 3.595, 3.600, 3.644, 3.681,
      3.748, 3.843, 3.776, 3.943, 3.942, 4.032, 4.065, 4.078, 4.123, 4.135, 4.188, 4.209, 4.237,
      4.249, 4.306, 4.318, 4.363, 4.388, 4.432, 4.435, 4.479, 4.520, 4.612, 4.646, 4.640, 4.630,
      4.714, 4.706, 4.756, 4.774, 4.823, 4.848, 4.878, 4.897, 4.927, 4.948, 5.054, 5.093, 5.133,
      5.177, 5.197, 5.210, 5.264, 5.301, 5.390, 5.371, 5.429, 5.479, 5.423, 5.400, 5.409, 5.411,
      5.387, 5.318, 5.404, 5.471, 5.498, 5.520, 5.520, 5.529, 5.629, 5.637, 5.661, 5.669, 5.710,
      5.701, 5.784, 5.825, 5.824, 5.817, 5.843, 5.843, 5.868, 5.875, 5.906, 5.912, 5.918, 5.937,
      5.967, 5.973, 5.979, 5.985, 5.980, 6.033, 6.051, 6.056, 6.074, 6.079, 6.097, 6.097, 6.119,
      6.125, 6.125, 2.958};
  int i = (int)std::lround(Z);
  i = std::min(std::max(i, 1), 120);
  return R[i - 1] * 1e-15;
}

// polar_bremsstrahlung (also pair production), src/Straggling.jl:746-835
inline double polar_brems(double m, double ki, double kf, double Z, Philox4x32& rng) {
  const double E = ki + m, nu = ki - kf;
  if (nu <= 0 || E <= m) return 0.0;
  const double y = nu / E;
  const double tmp = 0.5 * m * nu / (E * (E - nu));
  const double mu0 = tmp * tmp;
  const double c1 = (2.0 * (1.0 - y) + y * y) * mu0;
  const double c2 = 4.0 * (1.0 - y) * mu0 * mu0;
  const double RN = straggling_nuclear_radius(Z);
  const double t2 = m * RN / HBAR_C_;
  const double muni = t2 * t2 / 6.0;
  const double muc = std::sqrt(mu0) - mu0;
  if (muc < 0) return 0.0;
  const double q0 = muni, r0 = mu0, qr0 = q0 * r0, den0 = r0 + qr0;
  if (den0 <= 0 || (1.0 + qr0) <= 0) return 0.0;
  const double x0 = (1.0 + 2.0 * qr0) * std::log((1.0 + qr0) / den0) - (1.0 - r0) * (1.0 + 2.0 * q0) / (1.0 + q0);
  if (x0 <= 0) return 0.0;
  for (int t = 0; t < 100; ++t) {
    const double rr = rng.uniform();
    const double r1 = rng.uniform();
    const double mu = rr * mu0 / (mu0 + 1.0 - rr);
    const double mus = mu0 + mu, mu2 = mus * mus;
    const double d1 = 1.0 / mu2, d2 = d1 * d1;
    const double score = r1 * c1 * d1;
    double pdf = c1 * d1 - mu * c2 * d2;
    if (score > pdf) continue;
    const double l2 = mu2 / (mu0 * mu0);
    const double q = muni * l2, r = mu0 * l2, qr = q * r;
    if (qr > 1e5) continue;
    const double den = r + qr;
    if (den <= 0 || (1.0 + qr) <= 0) continue;
    const double x = (1.0 + 2.0 * qr) * std::log((1.0 + qr) / den) - (1.0 - r) * (1.0 + 2.0 * q) / (1.0 + q);
    pdf *= x / x0;
    if (score <= pdf) return mu;
  }
  return 0.0;
}

// polar_photonuclear, src/Straggling.jl:846-892
inline double polar_photo(double ml, double ki, double kf, Philox4x32& rng) {
  const double M = 0.5 * (PROTON_MASS_ + NEUTRON_MASS_);
  const double q = ki - kf, E = ki + ml, ml2 = ml * ml;
  const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
  const double Q2max = 2.0 * M * (q - PION_MASS_) - PION_MASS_ * PION_MASS_;
  if (Q2max < Q2min || Q2min < 0) return 0.0;
  const double rQ2 = std::log(Q2max) - std::log(Q2min);
  const double Q2 = Q2min * std::exp(rQ2 * rng.uniform());
  const double p = std::sqrt(ki * (ki + 2.0 * ml));
  const double E1 = E - q;
  const double er = ml / E1;
  const double p1 = E1 * std::sqrt(std::max((1.0 + er) * (1.0 - er), 0.0));
  const double pp = p * p1;
  if (pp <= 0) return 0.0;
  double tv = pp + ml2 - E * E1;
  if (std::fabs(tv) <= 3.0 * EPS64 * pp) tv = 0.0;
  const double mu = 0.5 * tv / pp + 0.25 / pp * Q2;
  return std::min(std::max(mu, 0.0), 1.0);
}

// polar_ionisation, src/Straggling.jl:902-914
inline double polar_ionisation(double m, double ki, double kf) {
  const double nu = ki - kf, p2 = ki * (ki + 2.0 * m), E = ki + m;
  const double den = std::sqrt(p2 * (p2 + nu * nu - 2.0 * nu * E));
  if (den <= 0) return 0.0;
  const double c = (p2 - nu * (E + ELECTRON_MASS_)) / den;
  return 0.5 * (1.0 - c);
}

}  // namespace orc

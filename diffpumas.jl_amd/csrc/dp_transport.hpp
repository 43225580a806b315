// Device transport library for the stochastic backward Monte Carlo: one
// event per lane, all state in registers, tables read through L2.
//
// Restates, for gfx950:
//   transport_backward_step_full (straggled)    src/Geometry.jl:148-357
//   transport_mixed_step / reverse_del_event     src/MixedTransport.jl:111-200
//   fluctuate_energy_loss, sample_ehs_event,
//   sample_scattering_angle, sample_soft_scattering,
//   rotate_direction, polar_* DEL angles         src/Straggling.jl
//   mixture property overloads                   src/Physics.jl:1152-1365
// Every function reports DomainError / ErrorException conditions through the
// lane's error code instead of throwing (the host maps them to DP_E_*).
#pragma once
#include "dp_device.hpp"
#include "dp_models.hpp"

namespace dpk {

constexpr double EHS_PATH_MAX = 1e9;
constexpr double PI_D = 3.141592653589793;
constexpr int ERR_DOMAIN = 1, ERR_TRANSPORT = 2, ERR_CAP = 4;

// Table arrays (MaterialTable order, src/Physics.jl:64-102)
enum {
  A_ENERGY = 0, A_LOG_ENERGY = 1, A_CSDA_STOP = 2, A_CSDA_RANGE = 3, A_MIXED_STOP = 5,
  A_MIXED_RANGE = 6, A_STRAGGLING = 7, A_CROSS_SECTION = 12, A_TRANSPORT_PATH = 13,
  A_ELASTIC_PATH = 14, A_SCREENING = 17, A_SPIN_FACTOR = 18
};

struct DevTables {
  const double* tab;
  const double* scal;  // density, I, ZoA, 0 per material
  int K, pow2;
  double mass, ctau, elastic_ratio;
  __device__ __forceinline__ const double* a(int m, int arr) const {
    return tab + ((size_t)m * 19 + arr) * K;
  }
};

// A medium: runtime mixture (1-4 components) + its integrated tables (for
// n > 1, on the grid of material 0) + its mixed kernel.
struct DevMix {
  int n;
  int mat[4];
  double frac[4];
  const double *e, *le, *crange, *mrange;  // n > 1 only
};
struct DevKernel {
  int n, pow2, n_el;
  const double *e, *stop, *sigma, *range, *hazard, *proper;
  const double* el;  // n_el * 4: Z, A, I, fraction
  double onset, cutoff;
};
struct Medium {
  DevMix mix;
  DevKernel kern;
};

// ---------------------------------------------------------- lookups
__device__ __forceinline__ double prop(const DevTables& T, int m, int arr, double e, double le) {
  return interp_e(T.a(m, A_ENERGY), T.a(m, A_LOG_ENERGY), T.a(m, arr), T.K, T.pow2, e, le).y;
}
__device__ __forceinline__ double mix_sum(const DevTables& T, const DevMix& x, int arr, double e,
                                          double le) {
  if (x.n == 1) return prop(T, x.mat[0], arr, e, le);
  double v = 0.0;
  for (int j = 0; j < x.n; ++j) v += x.frac[j] * prop(T, x.mat[j], arr, e, le);
  return v;
}
__device__ __forceinline__ double mix_range(const DevTables& T, const DevMix& x, bool mixed,
                                            double e, double le) {
  if (x.n == 1) return prop(T, x.mat[0], mixed ? A_MIXED_RANGE : A_CSDA_RANGE, e, le);
  return interp_e(x.e, x.le, mixed ? x.mrange : x.crange, T.K, T.pow2, e, le).y;
}
__device__ __forceinline__ double mix_kinetic(const DevTables& T, const DevMix& x, bool mixed,
                                              double X) {
  if (x.n == 1) {
    const int m = x.mat[0];
    return inverse_range(T.a(m, A_ENERGY), T.a(m, A_LOG_ENERGY),
                         T.a(m, mixed ? A_MIXED_RANGE : A_CSDA_RANGE), T.K, T.pow2, X).y;
  }
  return inverse_range(x.e, x.le, mixed ? x.mrange : x.crange, T.K, T.pow2, X).y;
}
__device__ __forceinline__ double mix_transport_path(const DevTables& T, const DevMix& x, double e,
                                                     double le) {
  if (x.n == 1) return prop(T, x.mat[0], A_TRANSPORT_PATH, e, le);
  double inv = 0.0;
  for (int j = 0; j < x.n; ++j) {
    const double lp = prop(T, x.mat[j], A_TRANSPORT_PATH, e, le);
    if (lp > 0) inv += x.frac[j] / lp;
  }
  return inv > 0 ? 1.0 / inv : 1e9;
}
__device__ __forceinline__ double ehs_path_one(const DevTables& T, int m, double k, double lk) {
  if (k < 1e-3) return EHS_PATH_MAX;
  const double Lb = prop(T, m, A_ELASTIC_PATH, k, lk);
  const double p2 = k * (k + 2.0 * T.mass);
  if (Lb > 0 && Lb < EHS_PATH_MAX * p2) return Lb / p2;
  const double tp = prop(T, m, A_TRANSPORT_PATH, k, lk);
  if (tp > 0) return fmin(tp / T.elastic_ratio, EHS_PATH_MAX);
  return EHS_PATH_MAX;
}
__device__ __forceinline__ double ehs_path(const DevTables& T, const DevMix& x, double k) {
  const double lk = log(k);
  if (x.n == 1) return ehs_path_one(T, x.mat[0], k, lk);
  double inv = 0.0;
  for (int j = 0; j < x.n; ++j) {
    const double lp = ehs_path_one(T, x.mat[j], k, lk);
    if (lp > 0 && lp < EHS_PATH_MAX) inv += x.frac[j] / lp;
  }
  return inv > 0 ? 1.0 / inv : EHS_PATH_MAX;
}

// lookup, src/MixedTransport.jl:43-49: searchsortedlast clamped to an
// interval, linear in the natural variable, DomainError outside the grid.
__device__ __forceinline__ double klookup(double x, const double* xs, const double* ys, int n,
                                          int pow2, int& err) {
  if (!(x >= xs[0] && x <= xs[n - 1])) {
    err |= ERR_DOMAIN;
    return 0.0;
  }
  const int i = bracket(xs, n, pow2, x);
  const double d = xs[i + 1] - xs[i];
  return d == 0 ? ys[i] : ys[i] + (x - xs[i]) / d * (ys[i + 1] - ys[i]);
}

// The same lookup from an interval guess g (the search returns the interval
// the bisection returns, so values are those of klookup).
__device__ __forceinline__ bool kloc(double x, const double* __restrict__ xs, int n, int pow2,
                                     int g, int& i, int& err) {
  if (!(x >= xs[0] && x <= xs[n - 1])) { err |= ERR_DOMAIN; i = 0; return false; }
  i = locate_idx(xs, n, pow2, x, g);
  return true;
}
__device__ __forceinline__ double kval(double x, const double* __restrict__ xs,
                                       const double* __restrict__ ys, int i) {
  const double d = xs[i + 1] - xs[i];
  return d == 0 ? ys[i] : ys[i] + (x - xs[i]) / d * (ys[i + 1] - ys[i]);
}
__device__ __forceinline__ double klookup_g(double x, const double* xs, const double* ys, int n,
                                            int pow2, int g, int& err) {
  int i;
  if (!kloc(x, xs, n, pow2, g, i, err)) return 0.0;
  return kval(x, xs, ys, i);
}

// ------------------------------------------------------ random variates
__device__ __forceinline__ double uniform(Philox& r) { return r.uniform(); }
__device__ __forceinline__ double exponential(Philox& r) { return -log(1.0 - r.uniform()); }
__device__ __forceinline__ double truncated_randn(Philox& r) {
  for (;;) {
    const double rr = sqrt(-2.0 * log(r.uniform()));
    const double phi = 2.0 * PI_D * r.uniform();
    const double u = rr * sin(phi);
    if (fabs(u) <= 3.0) return u * 0.9848416472;
  }
}

// fluctuate_energy_loss (backward), src/Straggling.jl:103-173, :1046-1105
__device__ __noinline__ static void fluctuate(const DevTables& T, const DevMix& x, double ki, double lki, double dX,
                          double Xtot, Philox& rng, double& kout, double& ratio) {
  double k1 = mix_kinetic(T, x, true, Xtot + dX);
  const double dk0 = fabs(ki - k1);
  ratio = 1.0;
  if (k1 > 0 && dk0 > 0) {
    const double lk1 = log(k1);
    const double O0 = mix_sum(T, x, A_STRAGGLING, ki, lki);
    const double O1 = mix_sum(T, x, A_STRAGGLING, k1, lk1);
    double dk12 = 0.5 * dX * (O0 + O1);
    if (dX / Xtot > 0.05) {
      const double de1 = mix_sum(T, x, A_MIXED_STOP, k1, lk1);
      const double de0 = mix_sum(T, x, A_MIXED_STOP, ki, lki);
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

// sample_ehs_event (backward), src/Straggling.jl:373-434, :1284-1334
__device__ __noinline__ static bool sample_ehs(const DevTables& T, const DevMix& x, double ki, double kf, double dX,
                           double Xtot, double ratio, Philox& rng, double& X_ehs, double& k_ehs) {
  double kmin = fmin(ki, kf), kmax = fmax(ki, kf);
  if (kmin < 0.001) { const double t = kmin; kmin = kmax; kmax = t; }
  const double lb = ehs_path(T, x, kmin);
  if (lb <= 0 || lb >= EHS_PATH_MAX) return false;
  const double zeta = uniform(rng);
  if (zeta <= 0 || zeta >= 1) return false;
  X_ehs = -log(zeta) * lb;
  if (X_ehs >= dX) return false;
  const double kh = mix_kinetic(T, x, true, Xtot + X_ehs);
  k_ehs = ki + ratio * fabs(ki - kh);
  if (k_ehs <= 0) return false;
  const double lbk = ehs_path(T, x, k_ehs);
  if (lbk <= 0 || lbk >= EHS_PATH_MAX) return false;
  const double r = lbk / lb;
  if (r <= 0 || uniform(rng) > 1.0 / r) return false;
  return true;
}

// sample_scattering_angle, src/Straggling.jl:446-517, :1344-1405
__device__ __noinline__ static double scattering_angle(const DevTables& T, const DevMix& x, double k, Philox& rng) {
  const double mass = T.mass;
  const double lk = log(k);
  const double p2 = k * (k + 2.0 * mass);
  const double mu0 = fmax(mix_sum(T, x, A_SCREENING, k, lk), 1e-12);
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
  const double fspin = mix_sum(T, x, A_SPIN_FACTOR, k, lk);
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

// sample_soft_scattering, src/Straggling.jl:539-626, :1204-1250
__device__ __noinline__ static double soft_scattering(const DevTables& T, const DevMix& x, double k0, double k1,
                                  double step, double density, Philox& rng) {
  if (step <= 0 || density <= 0) return 0.0;
  const double l0 = mix_transport_path(T, x, k0, log(k0));
  const double i0 = l0 > 0 ? density / l0 : 0.0;
  double i1;
  if (k1 > 0) {
    const double l1 = mix_transport_path(T, x, k1, log(k1));
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

// rotate_direction, src/Straggling.jl:645-710.  Inlined at its call sites:
// as an out-of-line call its caller's live registers went through the stack
// on every scattered step (measured 2.8 % of the straggled MC).
__device__ __forceinline__ static void rotate(double d[3], double mu, Philox& rng) {
  const double ct = 1.0 - 2.0 * mu;
  const double st2 = 4.0 * mu * (1.0 - mu);
  if (st2 <= 0) return;
  const double st = sqrt(st2);
  const double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
  double u0x = 0, u0y = 0, u0z = 0;
  if (ax > ay) {
    if (ax > az) {
      const double nrm = 1.0 / sqrt(d[0] * d[0] + d[2] * d[2]);
      u0x = -d[2] * nrm; u0z = d[0] * nrm;
    } else {
      const double nrm = 1.0 / sqrt(d[1] * d[1] + d[2] * d[2]);
      u0y = d[2] * nrm; u0z = -d[1] * nrm;
    }
  } else {
    if (ay > az) {
      const double nrm = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1]);
      u0x = d[1] * nrm; u0y = -d[0] * nrm;
    } else {
      const double nrm = 1.0 / sqrt(d[1] * d[1] + d[2] * d[2]);
      u0y = d[2] * nrm; u0z = -d[1] * nrm;
    }
  }
  const double u1x = u0y * d[2] - u0z * d[1];
  const double u1y = u0z * d[0] - u0x * d[2];
  const double u1z = u0x * d[1] - u0y * d[0];
  const double phi = PI_D * (1.0 - 2.0 * uniform(rng));
  double sp, cp;
  sincos(phi, &sp, &cp);
  double nx = ct * d[0] + st * (cp * u0x + sp * u1x);
  double ny = ct * d[1] + st * (cp * u0y + sp * u1y);
  double nz = ct * d[2] + st * (cp * u0z + sp * u1z);
  const double nrm = sqrt(nx * nx + ny * ny + nz * nz);
  if (nrm > 0) { nx /= nrm; ny /= nrm; nz /= nrm; }
  d[0] = nx; d[1] = ny; d[2] = nz;
}

// nuclear_radius of src/Straggling.jl:726-729 (entry Z of the 1-based
// table of nuclear RMS radii, fm)
__constant__ const double DEL_NUCLEAR_RADII[120] = {
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
__device__ __forceinline__ double del_nuclear_radius(double Z) {
  int i = (int)rint(Z);
  i = min(max(i, 1), 120);
  return DEL_NUCLEAR_RADII[i - 1] * 1e-15;
}

// polar_bremsstrahlung (also pair production), src/Straggling.jl:746-835
__device__ __noinline__ static double polar_brems(double m, double ki, double kf, double Z, Philox& rng) {
  const double E = ki + m, nu = ki - kf;
  if (nu <= 0 || E <= m) return 0.0;
  const double y = nu / E;
  const double tmp = 0.5 * m * nu / (E * (E - nu));
  const double mu0 = tmp * tmp;
  const double c1 = (2.0 * (1.0 - y) + y * y) * mu0;
  const double c2 = 4.0 * (1.0 - y) * mu0 * mu0;
  const double RN = del_nuclear_radius(Z);
  const double t2 = m * RN / dpm::HBAR_C;
  const double muni = t2 * t2 / 6.0;
  const double muc = sqrt(mu0) - mu0;
  if (muc < 0) return 0.0;
  const double q0 = muni, r0 = mu0, qr0 = q0 * r0, den0 = r0 + qr0;
  if (den0 <= 0 || (1.0 + qr0) <= 0) return 0.0;
  const double x0 = (1.0 + 2.0 * qr0) * log((1.0 + qr0) / den0) - (1.0 - r0) * (1.0 + 2.0 * q0) / (1.0 + q0);
  if (x0 <= 0) return 0.0;
  for (int t = 0; t < 100; ++t) {
    const double rr = uniform(rng);
    const double r1 = uniform(rng);
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
    const double x = (1.0 + 2.0 * qr) * log((1.0 + qr) / den) - (1.0 - r) * (1.0 + 2.0 * q) / (1.0 + q);
    pdf *= x / x0;
    if (score <= pdf) return mu;
  }
  return 0.0;
}

// polar_photonuclear, src/Straggling.jl:846-892
__device__ __noinline__ static double polar_photo(double ml, double ki, double kf, Philox& rng) {
  const double M = 0.5 * (dpm::PROTON_MASS + dpm::NEUTRON_MASS);
  const double q = ki - kf, E = ki + ml, ml2 = ml * ml;
  const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
  const double Q2max = 2.0 * M * (q - dpm::PION_MASS) - dpm::PION_MASS * dpm::PION_MASS;
  if (Q2max < Q2min || Q2min < 0) return 0.0;
  const double rQ2 = log(Q2max) - log(Q2min);
  const double Q2 = Q2min * exp(rQ2 * uniform(rng));
  const double p = sqrt(ki * (ki + 2.0 * ml));
  const double E1 = E - q;
  const double er = ml / E1;
  const double p1 = E1 * sqrt(fmax((1.0 + er) * (1.0 - er), 0.0));
  const double pp = p * p1;
  if (pp <= 0) return 0.0;
  double tv = pp + ml2 - E * E1;
  if (fabs(tv) <= 3.0 * EPS64 * pp) tv = 0.0;
  const double mu = 0.5 * tv / pp + 0.25 / pp * Q2;
  return fmin(fmax(mu, 0.0), 1.0);
}

// polar_ionisation, src/Straggling.jl:902-914
__device__ __forceinline__ double polar_ionisation(double m, double ki, double kf) {
  const double nu = ki - kf, p2 = ki * (ki + 2.0 * m), E = ki + m;
  const double den = sqrt(p2 * (p2 + nu * nu - 2.0 * nu * E));
  if (den <= 0) return 0.0;
  const double c = (p2 - nu * (E + dpm::ELECTRON_MASS)) / den;
  return 0.5 * (1.0 - c);
}

// differential_components, src/MixedTransport.jl:99-106
__device__ __forceinline__ double diff_component(const DevKernel& c, double mass, int i, int proc,
                                                 double k, double q) {
  const double Z = c.el[4 * i], A = c.el[4 * i + 1], I = c.el[4 * i + 2], f = c.el[4 * i + 3];
  double raw;
  if (proc == 1) raw = dpm::dcs_brems(Z, A, mass, k, q);
  else if (proc == 2) raw = dpm::dcs_pair(Z, A, mass, k, q);
  else if (proc == 3) raw = dpm::dcs_photo(Z, A, mass, k, q);
  else raw = dpm::dcs_ionisation(Z, I, mass, k, q);
  return f * dpm::AVOGADRO_NUMBER / (A * .001) * k * raw;
}

struct DelOut {
  double k, w;
  int proc, elem;
};

// reverse_del_event, src/MixedTransport.jl:111-142.  The components are
// evaluated twice when a target is needed (sum, then selection) instead of
// being stored: it keeps the lane's register footprint small.
__device__ __noinline__ static DelOut reverse_del_event(const DevKernel& c, double mass, double k, Philox& rng,
                                    bool details, int& err) {
  const double sigma = klookup(k, c.e, c.sigma, c.n, c.pow2, err);
  if (err) return {k, 0.0, 0, -1};
  if (!(sigma > 0)) { err |= ERR_TRANSPORT; return {k, 0.0, 0, -1}; }
  const double pcel = k >= c.onset ? 0.0
                                   : -log(c.cutoff) / (-log(c.cutoff) + log(c.onset / (c.onset - k)));
  if (pcel > 0 && uniform(rng) < pcel) return {k, 1 / pcel, 0, -1};
  const double me = dpm::ELECTRON_MASS;
  const double xmax = k < .5 * (mass - me) * (mass - me) / me
                          ? fmin(1.0, 2 * me * (k + 2 * mass) / ((mass + me) * (mass + me)))
                          : 1.0;
  const double xmin = k >= c.onset * (1 - c.cutoff) ? c.cutoff : 1 - k / c.onset;
  if (!(xmax > xmin)) return {k, 1.0, 0, -1};
  const double invnorm = 1 / xmin - 1 / xmax;
  const double nu = 1 / (1 / xmin - uniform(rng) * invnorm);
  const double r = 1 - nu;
  const double anc = k / r;
  const double pdf = 1 / (invnorm * (nu * nu));
  double total = 0.0;
  for (int i = 0; i < c.n_el; ++i)
    for (int p = 1; p <= 4; ++p) total += diff_component(c, mass, i, p, anc, anc - k);
  const double w = total / (pdf * r * sigma * (1 - pcel));
  if (!details || total <= 0) return {anc, w, 0, -1};
  double u = uniform(rng) * total;
  for (int i = 0; i < c.n_el; ++i)
    for (int p = 1; p <= 4; ++p) {
      u -= diff_component(c, mass, i, p, anc, anc - k);
      if (u <= 0) return {anc, w, p, i};
    }
  return {anc, w, 4, c.n_el - 1};
}

// ------------------------------------------------------------ the state
struct McState {
  double E, w, t;
  double pos[3], dir[3];
  double L;  // distance travelled (State.distance)
};

// transport_backward_step_full, mode = :straggled, src/Geometry.jl:148-357
__device__ __noinline__ static void straggled_step(const DevTables& T, const Medium& md, McState& st, double density,
                               double ds, bool scattering, double energy_limit, Philox& rng,
                               int& err) {
  const DevMix& x = md.mix;
  const DevKernel& K = md.kern;
  const double ki = st.E, mass = T.mass, lki = log(ki);
  double dX = ds * density;
  const double Xi = mix_range(T, x, true, ki, lki);
  double kf = ki, ratio = 1.0;
  if (ki > 0.001) {
    fluctuate(T, x, ki, lki, dX, Xi, rng, kf, ratio);
    kf = fmax(kf, ki);
  } else {
    kf = mix_kinetic(T, x, true, Xi + dX);
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
  {
    const double rate = fmax(klookup(ki, K.e, K.sigma, K.n, K.pow2, err),
                             klookup(kf, K.e, K.sigma, K.n, K.pow2, err));
    if (err) return;
    if (rate > 0) {
      const double cand = exponential(rng) / rate;
      if (cand < dX) {
        const double kh = mix_kinetic(T, x, true, Xi + cand);
        const double pre = ki + ratio * fabs(kh - ki);
        if (pre > 0 && uniform(rng) <= klookup(pre, K.e, K.sigma, K.n, K.pow2, err) / rate) {
          del = true; X_del = cand; k_del = pre;
        }
        if (err) return;
      }
    }
  }
  if (scattering) {
    double xe, ke;
    if (sample_ehs(T, x, ki, kf, dX, Xi, ratio, rng, xe, ke) && ke > 0) {
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
    const DelOut o = reverse_del_event(K, mass, k_del, rng, true, err);
    if (err) return;
    kf = o.k; dw = o.w; proc = o.proc; elem = o.elem;
    dX = X_del;
  }
  if (ehs) { kf = k_ehs; ckf = k_ehs; dX = X_ehs; }
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
      const double mu = scattering_angle(T, x, k_ehs, rng);
      if (mu > 0) rotate(d, mu, rng);
    }
    const double ms = soft_scattering(T, x, ki, ckf, step, density, rng);
    if (ms > 0) rotate(d, ms, rng);
  }
  const double dt = step * mass / 2 *
                    (1.0 / sqrt(ki * (ki + 2 * mass)) + 1.0 / sqrt(ckf * (ckf + 2 * mass)));
  for (int a = 0; a < 3; ++a) st.pos[a] -= step * st.dir[a];
  st.L += step;
  const double wf = mix_sum(T, x, A_MIXED_STOP, ckf, log(ckf)) / mix_sum(T, x, A_MIXED_STOP, ki, lki);
  st.w = st.w * wf * exp(-dt / T.ctau) * dw;
  st.t += dt;
  st.E = kf;
  for (int a = 0; a < 3; ++a) st.dir[a] = d[a];
}

// Julia eps(x) for x >= 0
__device__ __forceinline__ double ulp(double x) {
  x = fabs(x);
  return nextafter(x, __builtin_huge_val()) - x;
}

// transport_mixed_step, src/MixedTransport.jl:144-200
__device__ __noinline__ static void mixed_step(const DevTables& T, const Medium& md, McState& st, double density,
                           double distance, double energy_limit, Philox& rng, int& err) {
  const DevKernel& c = md.kern;
  const double m = T.mass;
  const double emax = c.e[c.n - 1];
  if (!(density > 0 && distance >= 0) || energy_limit > emax) { err |= ERR_TRANSPORT; return; }
  if (st.E > emax) { err |= ERR_DOMAIN; return; }
  double k = st.E, remaining = distance * density, travelled = 0, proper = 0, weight = st.w;
  const double ceiling = energy_limit > 0 ? fmin(energy_limit, emax) : emax;
  const double tiny = ulp(fmax(1.0, distance * density));
  const double rceil = klookup(ceiling, c.e, c.range, c.n, c.pow2, err);
  bool exhausted = true;
  for (int ev = 0; ev < 100000; ++ev) {
    if (remaining <= tiny) { exhausted = false; break; }
    if (k >= ceiling) { exhausted = false; break; }
    const double ri = klookup(k, c.e, c.range, c.n, c.pow2, err);
    const double hi = klookup(k, c.e, c.hazard, c.n, c.pow2, err);
    const double rtc = fmax(0.0, rceil - ri);
    const double eg = fmin(remaining, rtc);
    const double endpoint = eg == rtc ? ceiling : fmax(k, klookup(ri + eg, c.range, c.e, c.n, c.pow2, err));
    const double eventh = hi + exponential(rng);
    const bool occurs = eventh < klookup(endpoint, c.e, c.hazard, c.n, c.pow2, err);
    const double pre = occurs ? fmax(k, klookup(eventh, c.hazard, c.e, c.n, c.pow2, err)) : endpoint;
    const double dx = occurs ? fmax(0.0, klookup(pre, c.e, c.range, c.n, c.pow2, err) - ri) : eg;
    double dt;
    if (dx <= 8 * ulp(ri) || pre - k <= 8 * ulp(k))
      dt = dx / density * m / 2 * (1.0 / sqrt(k * (k + 2 * m)) + 1.0 / sqrt(pre * (pre + 2 * m)));
    else
      dt = (klookup(pre, c.e, c.proper, c.n, c.pow2, err) - klookup(k, c.e, c.proper, c.n, c.pow2, err)) / density;
    weight *= klookup(pre, c.e, c.stop, c.n, c.pow2, err) / klookup(k, c.e, c.stop, c.n, c.pow2, err) *
              exp(-dt / T.ctau);
    if (err) return;
    travelled += dx;
    remaining = fmax(0.0, remaining - dx);
    proper += dt;
    k = pre;
    if (occurs) {
      const DelOut o = reverse_del_event(c, m, k, rng, false, err);
      if (err) return;
      k = o.k;
      weight *= o.w;
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
__device__ static void csda_mix_step(const DevTables& T, const Medium& md, McState& st, double density,
                              double ds) {
  const DevMix& x = md.mix;
  const double ki = st.E, lki = log(ki), gs = ds * density;
  const double Xi = mix_range(T, x, false, ki, lki);
  const double kf = mix_kinetic(T, x, false, Xi + gs);
  const double gi = 1.0 + ki / T.mass, gc = 1.0 + kf / T.mass, ga = (gi + gc) / 2.0;
  const double beta = sqrt(fmax(1.0 - 1.0 / (ga * ga), 1e-12));
  const double dt = ds / (beta * ga);
  for (int a = 0; a < 3; ++a) st.pos[a] -= ds * st.dir[a];
  st.L += ds;
  const double wf = mix_sum(T, x, A_CSDA_STOP, kf, log(kf)) / mix_sum(T, x, A_CSDA_STOP, ki, lki);
  st.w = st.w * wf * exp(-dt / T.ctau);
  st.t += dt;
  st.E = kf;
}

}  // namespace dpk

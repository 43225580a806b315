// CPU oracle: differential cross-sections of the hard processes and their
// energy-moment integrals.  TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// An independent restatement of the reference models, kept in the
// reference's own operation order (the shipped models in
// diffpumas.jl_amd/csrc/dp_models.hpp hoist per-call invariants and share
// reciprocals; this file does neither):
//   bremsstrahlung  SSR              src/Materials.jl:631-695
//   pair production SSR (+ 12-pt GQ) src/Materials.jl:701-832
//   photonuclear    DRSS / ALLM97    src/Materials.jl:838-926
//   ionisation      Salvat 2013      src/Materials.jl:534-586
//   dilogarithm                      src/Materials.jl:286-339
//   moments         compute_dcs_integral, src/Materials.jl:939-1007
// The Gauss-Legendre nodes and weights are computed here from the Legendre
// polynomials (Newton iteration) instead of the reference's 16-digit tables
// (src/Materials.jl:141-235); they agree to rounding.
#pragma once
#include <algorithm>
#include <cmath>

namespace orc {
namespace dcs {

// src/Constants.jl:15-62
constexpr double ALPHA_EM = 7.2973525693e-3;
constexpr double HBAR_C = 1.973269804e-16;
constexpr double ELECTRON_MASS = 0.510998910e-3;
constexpr double ELECTRON_RADIUS = 2.817940285e-15;
constexpr double PROTON_MASS = 0.938272;
constexpr double NEUTRON_MASS = 0.939565;
constexpr double PION_MASS = 0.13957018;
constexpr double AVOGADRO_NUMBER = 6.02214076e23;
constexpr double PI = 3.141592653589793;

// n-point Gauss-Legendre rule on [0, 1]: nodes x (ascending) and weights w,
// from Newton iteration on P_n with the three-term recurrence.
struct GaussRule {
  double x[16], w[16];
  int n;
  explicit GaussRule(int n_) : n(n_) {
    for (int i = 0; i < n; ++i) {
      // i-th root of P_n on [-1, 1], descending from the largest
      double z = std::cos(PI * (i + 0.75) / (n + 0.5));
      double dp = 0;
      for (int it = 0; it < 100; ++it) {
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
          const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        if (n == 1) { p1 = z; p0 = 1.0; }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        const double dz = p1 / dp;
        z -= dz;
        if (std::fabs(dz) < 1e-16) break;
      }
      {
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
          const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
      }
      // map [-1, 1] -> [0, 1]: ascending order
      x[i] = 0.5 * (1.0 - z);
      w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
  }
};
inline const GaussRule& gauss(int n) {
  static const GaussRule g9(9), g12(12);
  return n == 9 ? g9 : g12;
}

// math_dilog, src/Materials.jl:286-339 (Chebyshev expansion)
inline double dilog(double x) {
  static const double C[20] = {
      0.42996693560813697, 0.40975987533077105, -0.01858843665014592, 0.00145751084062268,
      -0.00014304184442340, 0.00001588415541880, -0.00000190784959387, 0.00000024195180854,
      -0.00000003193341274, 0.00000000434545063, -0.00000000060578480, 0.00000000008612098,
      -0.00000000001244332, 0.00000000000182256, -0.00000000000027007, 0.00000000000004042,
      -0.00000000000000610, 0.00000000000000093, -0.00000000000000014, 0.00000000000000002};
  const double PI6 = PI * PI / 6.0, PI12 = PI * PI / 12.0, PI3 = PI * PI / 3.0;
  if (x == 1.0) return PI6;
  if (x == -1.0 || (1.0 - x) == 0.0) return -PI12;
  const double T = -x;
  double Y, S, Av;
  if (T <= -2.0) {
    Y = -1.0 / (1.0 + T);
    S = 1.0;
    const double a = std::log(-T), b = std::log(1.0 + 1.0 / T);
    Av = -PI3 + 0.5 * (a * a - b * b);
  } else if (T < -1.0) {
    Y = -1.0 - T;
    S = -1.0;
    const double la = std::log(-T);
    Av = -PI6 + la * (la + std::log(1.0 + 1.0 / T));
  } else if (T <= -0.5) {
    Y = -(1.0 + T) / T;
    S = 1.0;
    const double la = std::log(-T);
    Av = -PI6 + la * (-0.5 * la + std::log(1.0 + T));
  } else if (T < 0.0) {
    Y = -T / (1.0 + T);
    S = -1.0;
    const double l = std::log(1.0 + T);
    Av = 0.5 * l * l;
  } else if (T <= 1.0) {
    Y = T;
    S = 1.0;
    Av = 0.0;
  } else {
    Y = 1.0 / T;
    S = -1.0;
    const double l = std::log(T);
    Av = PI6 + 0.5 * l * l;
  }
  const double H = Y + Y - 1.0, ALFA = H + H;
  double B1 = 0.0, B2 = 0.0;
  for (int i = 19; i >= 0; --i) {
    const double B0 = C[i] + ALFA * B1 - B2;
    B2 = B1;
    B1 = B0;
  }
  return -(S * (B1 - H * B2) + Av);
}

// radiation_logarithm, src/Materials.jl:592-620
inline double radiation_logarithm(double Z) {
  static const double L[92] = {
      202.4, 151.9, 159.9, 172.3, 177.9, 178.3, 176.6, 173.4, 170.0, 165.8, 165.8, 167.1,
      169.1, 170.8, 172.2, 173.4, 174.3, 174.8, 175.1, 175.6, 176.2, 176.8, 0.0,   0.0,
      0.0,   175.8, 0.0,   0.0,   173.1, 0.0,   0.0,   173.0, 0.0,   0.0,   173.5, 0.0,
      0.0,   0.0,   0.0,   0.0,   0.0,   175.9, 0.0,   0.0,   0.0,   0.0,   0.0,   0.0,
      0.0,   177.4, 0.0,   0.0,   178.6, 0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,
      0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,
      0.0,   177.6, 0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   178.0, 0.0,   0.0,
      0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   179.8};
  const long i = std::lround(Z);
  if (i >= 1 && i <= 92 && L[i - 1] > 0.0) return L[i - 1];
  return 182.7;
}

// dcs_bremsstrahlung_ssr, src/Materials.jl:631-695
inline double brems(double Z, double A, double m, double K, double q) {
  const double SQRTE = 1.648721270700128, ME_MEV = 0.5109989461, RE_CM = 2.8179403227e-13,
               MMU_MEV = 105.6583745;
  if (Z <= 0 || A <= 0 || m <= 0 || K <= 0 || q <= 0) return 0.0;
  const double z13 = std::pow(Z, 1.0 / 3.0);
  if (q >= K + m * (1.0 - 0.75 * SQRTE * z13)) return 0.0;
  const double a_c[3] = {-0.00349, 148.84, -987.531};
  const double b_c[4] = {0.1642, 132.573, -585.361, 1407.77};
  const double c_c[6] = {-2.8922, -19.0156, 57.698, -63.418, 14.1166, 1.84206};
  const double d_c[6] = {2134.19, 581.823, -2708.85, 4767.05, 1.52918, 0.361933};
  const double energy = (K + m) * 1e3;
  const double v = q * 1e3 / energy;
  const double m_mev = m * 1e3;
  const double Z13 = 1.0 / z13;
  const double rad_log = radiation_logarithm(Z);
  const double rad_log_inel = Z == 1.0 ? 446.0 : 1429.0;
  const double Dn = 1.54 * std::pow(A, 0.27);
  const double mu_qc = m_mev / (MMU_MEV * std::exp(1.0) / Dn);
  const double rho = std::sqrt(1.0 + 4.0 * mu_qc * mu_qc);
  const double log_rho = std::log((rho + 1.0) / (rho - 1.0));
  const double delta1 = std::log(mu_qc) + 0.5 * rho * log_rho;
  const double delta2 = std::log(mu_qc) + 0.25 * (3.0 * rho - rho * rho * rho) * log_rho +
                        2.0 * mu_qc * mu_qc;
  const double delta = m_mev * m_mev * v / (2.0 * energy * (1.0 - v));
  double phi1 = std::log(rad_log * Z13 * (m_mev / ME_MEV) /
                         (1.0 + rad_log * Z13 * std::exp(0.5) * delta / ME_MEV));
  double phi2 = std::log(rad_log * Z13 * std::exp(-1.0 / 6.0) * (m_mev / ME_MEV) /
                         (1.0 + rad_log * Z13 * std::exp(1.0 / 3.0) * delta / ME_MEV));
  phi1 -= delta1 * (1.0 - 1.0 / Z);
  phi2 -= delta2 * (1.0 - 1.0 / Z);
  const double s1 = std::log(m_mev / delta / (m_mev * delta / (ME_MEV * ME_MEV) + SQRTE));
  const double s2 = std::log(1.0 + ME_MEV / (delta * rad_log_inel * Z13 * Z13 * SQRTE));
  const double s_atomic = (4.0 / 3.0 * (1.0 - v) + v * v) * (s1 - s2);
  double s_rad;
  if (v < 0.0 || v > 1.0) {
    s_rad = 0.0;
  } else if (v < 0.02) {
    s_rad = a_c[0] + a_c[1] * v + a_c[2] * v * v;
  } else if (v < 0.1) {
    s_rad = b_c[0] + b_c[1] * v + b_c[2] * v * v + b_c[3] * v * v * v;
  } else if (v < 0.9) {
    s_rad = c_c[0] + c_c[1] * v + c_c[2] * v * v;
    const double t = std::log(1.0 - v);
    s_rad += c_c[3] * v * std::log(v) + c_c[4] * t + c_c[5] * t * t;
  } else {
    s_rad = d_c[0] + d_c[1] * v + d_c[2] * v * v;
    const double t = std::log(1.0 - v);
    s_rad += d_c[3] * v * std::log(v) + d_c[4] * t + d_c[5] * t * t;
  }
  const double result = ((2.0 - 2.0 * v + v * v) * phi1 - 2.0 / 3.0 * (1.0 - v) * phi2) +
                        1.0 / Z * s_atomic + 0.25 * ALPHA_EM * phi1 * s_rad;
  if (result <= 0.0) return 0.0;
  const double aux = 2.0 * (ME_MEV / m_mev) * RE_CM * Z;
  return aux * aux * (ALPHA_EM / q) * result * 1e-4;
}

// _dcs_pair_production_d2_ssr, src/Materials.jl:701-800
inline double pair_d2(double Z, double A, double mass, double K, double q, double rho_in) {
  const double ME_MEV = 0.5109989461, RE_CM = 2.8179403227e-13;
  const double energy = (K + mass) * 1e3;
  const double v = q / (K + mass);
  const double m = mass * 1e3;
  const double rad_log = radiation_logarithm(Z);
  const double const_pf = 4.0 / (3.0 * PI) * Z * (ALPHA_EM * RE_CM) * (ALPHA_EM * RE_CM);
  const double Z13 = std::pow(Z, -1.0 / 3.0);
  const double d_n = 1.54 * std::pow(A, 0.27);
  const double rho = 1.0 - rho_in;
  const double rho2 = rho * rho;
  const double g1 = Z == 1.0 ? 4.4e-5 : 1.95e-5;
  const double g2 = Z == 1.0 ? 4.8e-5 : 5.3e-5;
  const double zeta1 =
      0.073 * std::log(energy / m / (1.0 + g1 * std::pow(Z, 2.0 / 3.0) * energy / m)) - 0.26;
  const double zeta2 = 0.058 * std::log(energy / m / (1.0 + g2 / Z13 * energy / m)) - 0.14;
  const double zeta = (zeta1 > 0.0 && zeta2 > 0.0) ? zeta1 / zeta2 : 0.0;
  const double beta = v * v / (2.0 * (1.0 - v));
  const double t0 = m * v / (2.0 * ME_MEV);
  const double xi = t0 * t0 * (1.0 - rho2) / (1.0 - v);
  const double Be = ((2.0 + rho2) * (1.0 + beta) + xi * (3.0 + rho2)) * std::log(1.0 + 1.0 / xi) +
                    (1.0 - rho2 - beta) / (1.0 + xi) - (3.0 + rho2);
  const double Ce2 =
      ((1.0 - rho2) * (1.0 + beta) + xi * (3.0 - rho2)) * std::log(1.0 + 1.0 / xi) +
      2.0 * (1.0 - beta - rho2) / (1.0 + xi) - (3.0 - rho2);
  const double Ce1 = Be - Ce2;
  const double De = ((2.0 + rho2) * (1.0 + beta) + xi * (3.0 + rho2)) * dilog(1.0 / (1.0 + xi)) -
                    (2.0 + rho2) * xi * std::log(1.0 + 1.0 / xi) -
                    (xi + rho2 + beta) / (1.0 + xi);
  const double mn = ME_MEV / m * d_n;
  double Le1, Le2;
  if (De / Be > 0.0) {
    const double tmp = De / Be;
    const double Xe = tmp < 100.0 ? std::exp(-tmp) : 0.0;
    Le1 = std::log(rad_log * Z13 * std::sqrt(1.0 + xi) /
                   (Xe + 2.0 * ME_MEV * std::exp(0.5) * rad_log * Z13 * (1.0 + xi) /
                             (energy * v * (1.0 - rho2)))) -
          tmp - 0.5 * std::log(Xe + mn * mn * (1.0 + xi));
    Le2 = std::log(rad_log * Z13 * std::exp(-1.0 / 6.0) * std::sqrt(1.0 + xi) /
                   (Xe + 2.0 * ME_MEV * std::exp(1.0 / 3.0) * rad_log * Z13 * (1.0 + xi) /
                             (energy * v * (1.0 - rho2)))) -
          tmp - 0.5 * std::log(Xe + mn * mn * std::exp(-1.0 / 3.0) * (1.0 + xi));
  } else {
    const double tmp = De / Be;
    const double Xi = tmp > -100.0 ? std::exp(tmp) : 0.0;
    Le1 = std::log(rad_log * Z13 * std::sqrt(1.0 + xi) /
                   (1.0 + Xi * 2.0 * ME_MEV * std::exp(0.5) * rad_log * Z13 * (1.0 + xi) /
                              (energy * v * (1.0 - rho2)))) -
          0.5 * tmp - 0.5 * std::log(1.0 + Xi * mn * mn * (1.0 + xi));
    Le2 = std::log(rad_log * Z13 * std::exp(-1.0 / 6.0) * std::sqrt(1.0 + xi) /
                   (1.0 + Xi * 2.0 * ME_MEV * std::exp(1.0 / 3.0) * rad_log * Z13 * (1.0 + xi) /
                              (energy * v * (1.0 - rho2)))) -
          0.5 * tmp - 0.5 * std::log(1.0 + Xi * mn * mn * std::exp(-1.0 / 3.0) * (1.0 + xi));
  }
  double diagram_e = const_pf * (Z + zeta) * (1.0 - v) / v * (Ce1 * Le1 + Ce2 * Le2);
  if (diagram_e < 0.0) diagram_e = 0.0;
  const double Bm = ((1.0 + rho2) * (1.0 + 1.5 * beta) -
                     1.0 / xi * (1.0 + 2.0 * beta) * (1.0 - rho2)) * std::log(1.0 + xi) +
                    xi * (1.0 - rho2 - beta) / (1.0 + xi) + (1.0 + 2.0 * beta) * (1.0 - rho2);
  const double Cm2 = ((1.0 - beta) * (1.0 - rho2) - xi * (1.0 + rho2)) * std::log(1.0 + xi) / xi -
                     2.0 * (1.0 - beta - rho2) / (1.0 + xi) + 1.0 - beta - (1.0 + beta) * rho2;
  const double Cm1 = Bm - Cm2;
  const double Dm = ((1.0 + rho2) * (1.0 + 1.5 * beta) -
                     1.0 / xi * (1.0 + 2.0 * beta) * (1.0 - rho2)) * dilog(xi / (1.0 + xi)) +
                    (1.0 + 1.5 * beta) * (1.0 - rho2) / xi * std::log(1.0 + xi) +
                    (1.0 - rho2 - 0.5 * beta * (1.0 + rho2) + (1.0 - rho2) / (2.0 * xi) * beta) *
                        xi / (1.0 + xi);
  double Lm1, Lm2;
  if (Dm / Bm > 0.0) {
    const double tmp = Dm / Bm;
    const double Xm = tmp < 100.0 ? std::exp(-tmp) : 0.0;
    Lm1 = std::log(m / ME_MEV * rad_log * Z13 / d_n /
                   (Xm + 2.0 * ME_MEV * std::exp(0.5) * rad_log * Z13 * (1.0 + xi) /
                             (energy * v * (1.0 - rho2)))) -
          tmp;
    Lm2 = std::log(m / ME_MEV * rad_log * Z13 / d_n /
                   (Xm + 2.0 * ME_MEV * std::exp(1.0 / 3.0) * rad_log * Z13 * (1.0 + xi) /
                             (energy * v * (1.0 - rho2)))) -
          tmp;
  } else {
    const double tmp = Dm / Bm;
    const double Xmv = tmp > -100.0 ? std::exp(tmp) : 0.0;
    Lm1 = std::log(m / ME_MEV * rad_log * Z13 / d_n /
                   (1.0 + 2.0 * ME_MEV * std::exp(0.5) * rad_log * Z13 * (1.0 + xi) /
                              (energy * v * (1.0 - rho2)) * Xmv));
    Lm2 = std::log(m / ME_MEV * rad_log * Z13 / d_n /
                   (1.0 + 2.0 * ME_MEV * std::exp(1.0 / 3.0) * rad_log * Z13 * (1.0 + xi) /
                              (energy * v * (1.0 - rho2)) * Xmv));
  }
  double diagram_mu = const_pf * (Z + zeta) * (1.0 - v) / v * (ME_MEV / m) * (ME_MEV / m) *
                      (Cm1 * Lm1 + Cm2 * Lm2);
  if (diagram_mu < 0.0) diagram_mu = 0.0;
  return (diagram_e + diagram_mu) * 0.1 / energy;
}

// dcs_pair_production_ssr, src/Materials.jl:802-832
inline double pair(double Z, double A, double mass, double K, double q) {
  if (Z <= 0 || A <= 0 || mass <= 0 || K <= 0 || q <= 0) return 0.0;
  if (q <= 4.0 * ELECTRON_MASS) return 0.0;
  const double sqrte = 1.6487212707;
  const double Z13 = std::pow(Z, 1.0 / 3.0);
  if (q >= K + mass * (1.0 - 0.75 * sqrte * Z13)) return 0.0;
  const double gamma = 1.0 + K / mass;
  const double x0 = 1.0 - 4.0 * ELECTRON_MASS / q;
  const double x1 = 1.0 - 6.0 / (gamma * (gamma - q / mass));
  if (!(x0 > 0.0 && x1 > 0.0)) return 0.0;
  const double rmax = std::sqrt(x0) * x1;
  const double ri = 1.0 - rmax;
  if (ri <= 0.0 || ri >= 1.0) return 0.0;
  const double tmin = std::log(ri);
  const GaussRule& g = gauss(12);
  double I = 0.0;
  for (int i = 0; i < 12; ++i) {
    const double r = std::exp(g.x[i] * tmin);
    I -= pair_d2(Z, A, mass, K, q, r) * r * g.w[i] * tmin;
  }
  return I < 0.0 ? 0.0 : I;
}

// _dcs_photonuclear_f2p_ALLM97, src/Materials.jl:838-863
inline double f2p_allm97(double x, double Q2) {
  const double m02 = 0.31985, mP2 = 49.457, mR2 = 0.15052, Q02 = 0.52544, Lambda2 = 0.06527;
  const double cP1 = 0.28067, cP2 = 0.22291, cP3 = 2.1979;
  const double aP1 = -0.0808, aP2 = -0.44812, aP3 = 1.1709;
  const double bP1 = 0.36292, bP2 = 1.8917, bP3 = 1.8439;
  const double cR1 = 0.80107, cR2 = 0.97307, cR3 = 3.4942;
  const double aR1 = 0.58400, aR2 = 0.37888, aR3 = 2.6063;
  const double bR1 = 0.01147, bR2 = 3.7582, bR3 = 0.49338;
  const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
  const double M2 = M * M;
  const double W2 = M2 + Q2 * (1.0 / x - 1.0);
  const double t = std::log(std::log((Q2 + Q02) / Lambda2) / std::log(Q02 / Lambda2));
  const double xP = (Q2 + mP2) / (Q2 + mP2 + W2 - M2);
  const double xR = (Q2 + mR2) / (Q2 + mR2 + W2 - M2);
  const double cP = cP1 + (cP1 - cP2) * (1.0 / (1.0 + std::pow(t, cP3)) - 1.0);
  const double aP = aP1 + (aP1 - aP2) * (1.0 / (1.0 + std::pow(t, aP3)) - 1.0);
  const double bP = bP1 + bP2 * std::pow(t, bP3);
  const double cR = cR1 + cR2 * std::pow(t, cR3);
  const double aR = aR1 + aR2 * std::pow(t, aR3);
  const double bR = bR1 + bR2 * std::pow(t, bR3);
  const double F2P = cP * std::exp(aP * std::log(xP) + bP * std::log(1.0 - x));
  const double F2R = cR * std::exp(aR * std::log(xR) + bR * std::log(1.0 - x));
  return Q2 / (Q2 + m02) * (F2P + F2R);
}

// _dcs_photonuclear_shadowing_DRSS, _f2a_DRSS, _d2_DRSS, src/Materials.jl:865-895
inline double shadowing(double Z, double A, double x) {
  if (Z == 1.0) return 1.0;
  if (x < 0.0014) return std::exp(-0.1 * std::log(A));
  if (x < 0.04) return std::exp((0.069 * std::log10(x) + 0.097) * std::log(A));
  return 1.0;
}
inline double photo_d2(double Z, double A, double ml, double K, double q, double Q2) {
  const double cf = 4.0 * PI * ALPHA_EM * ALPHA_EM * HBAR_C * HBAR_C;
  const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
  const double E = K + ml;
  const double y = q / E;
  const double x = 0.5 * Q2 / (M * q);
  const double F2p = f2p_allm97(x, Q2);
  const double sh = shadowing(Z, A, x);
  const double F2A = F2p * sh * (Z + (A - Z) * (1.0 + x * (-1.85 + x * (2.45 + x * (-2.35 + x)))));
  const double dds = (1.0 - y + 0.5 * (1.0 - 2.0 * ml * ml / Q2) * (y * y + Q2 / (E * E))) /
                         (Q2 * Q2) -
                     0.25 / (E * E * Q2);
  return cf * F2A * dds / q;
}

// dcs_photonuclear_drss, src/Materials.jl:897-926
inline double photo(double Z, double A, double m, double K, double q) {
  if (Z <= 0 || A <= 0 || m <= 0 || K <= 0 || q <= 0) return 0.0;
  const double M = 0.5 * (NEUTRON_MASS + PROTON_MASS);
  const double mpi = PION_MASS;
  const double qmin = mpi + 0.5 * mpi * mpi / M;
  const double qmax = K + m - 0.5 * (M + m * m / M);
  if (q <= qmin || q >= qmax) return 0.0;
  const double E = K + m;
  const double ml2 = m * m;
  const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
  const double Q2max = 2.0 * M * (q - mpi) - mpi * mpi;
  if (Q2max < Q2min || Q2min < 0.0) return 0.0;
  const GaussRule& g = gauss(9);
  const double lmin = std::log(Q2min);
  const double dx = std::log(Q2max) - lmin;
  double ds = 0.0;
  for (int i = 0; i < 9; ++i) {
    const double Q2 = std::exp(lmin + dx * g.x[i]);
    const double val = photo_d2(Z, A, m, K, q, Q2);
    if (val > 0.0) ds += val * (g.w[i] * dx * Q2);
  }
  return ds < 0.0 ? 0.0 : ds;
}

// electronic_dcs, src/Materials.jl:534-548
inline double ionisation(double Z, double I, double m, double K, double q) {
  const double P2 = K * (K + 2.0 * m);
  const double E = K + m;
  const double Wmax = 2.0 * ELECTRON_MASS * P2 / (m * m + ELECTRON_MASS * (ELECTRON_MASS + 2.0 * E));
  if (q > Wmax) return 0.0;
  const double Wmin = (13.6 / 19.2) * I;
  if (q <= Wmin) return 0.0;
  const double E2 = E * E;
  const double beta2 = P2 / E2;
  const double a0 = 0.5 / E2;
  const double a1 = beta2 / Wmax;
  return 2.0 * PI * ELECTRON_MASS * ELECTRON_RADIUS * ELECTRON_RADIUS * Z *
         (a0 + 1.0 / q * (1.0 / q - a1)) / beta2;
}

// dcs_ionisation_integrate, src/Materials.jl:556-586
inline double ionisation_integral(double Z, double A, double I, double mass, double K,
                                  double xlow, double xhigh, int mode) {
  const double P2 = K * (K + 2.0 * mass);
  const double E = K + mass;
  const double Wr = 2.0 * ELECTRON_MASS * P2 / (mass * mass + ELECTRON_MASS * (ELECTRON_MASS + 2.0 * E));
  const double ql = K * xlow, qh = K * xhigh;
  double Wmax = Wr;
  if (Wmax < ql) return 0.0;
  if (Wmax > qh) Wmax = qh;
  double Wmin = (13.6 / 19.2) * I;
  if (ql >= Wmin) Wmin = ql;
  if (Wmax <= Wmin) return 0.0;
  const double E2 = E * E, beta2 = P2 / E2, a0 = 0.5 / E2, a1 = beta2 / Wr;
  double Iv;
  if (mode == 0)
    Iv = a0 * (Wmax - Wmin) - a1 * std::log(Wmax / Wmin) + 1.0 / Wmin - 1.0 / Wmax;
  else if (mode == 1)
    Iv = 0.5 * a0 * (Wmax * Wmax - Wmin * Wmin) - a1 * (Wmax - Wmin) + std::log(Wmax / Wmin);
  else
    Iv = a0 * (Wmax * Wmax * Wmax - Wmin * Wmin * Wmin) / 3.0 -
         0.5 * a1 * (Wmax * Wmax - Wmin * Wmin) + Wmax - Wmin;
  return 2.0 * PI * ELECTRON_MASS * ELECTRON_RADIUS * ELECTRON_RADIUS * Z * AVOGADRO_NUMBER /
         (A * 1e-3 * beta2) * Iv;
}

// dσ/dq per atom of process p (0 brems, 1 pair, 2 photonuclear, 3 ionisation)
inline double dcs(int p, double Z, double A, double I, double m, double K, double q) {
  switch (p) {
    case 0: return brems(Z, A, m, K, q);
    case 1: return pair(Z, A, m, K, q);
    case 2: return photo(Z, A, m, K, q);
    case 3: return ionisation(Z, I, m, K, q);
    default: return 0.0;
  }
}

// compute_dcs_integral, src/Materials.jl:939-1007
inline double integral(int p, double Z, double A, double I, double mass, double K, double xlow,
                       double xhigh, int mode, int npoints) {
  if (xlow <= 0.0) xlow = 1e-6;
  if (p == 3) return ionisation_integral(Z, A, I, mass, K, xlow, xhigh, mode);
  const double M = 0.5 * (NEUTRON_MASS + PROTON_MASS);
  const double mpi = PION_MASS;
  double qmin_abs, qmax_abs;
  if (p == 2) {
    qmin_abs = mpi + 0.5 * mpi * mpi / M;
    qmax_abs = K + mass - 0.5 * (M + mass * mass / M);
  } else {
    const double Z13 = std::pow(Z, 1.0 / 3.0);
    qmax_abs = K + mass * (1.0 - 0.75 * 1.648721271 * Z13);
    qmin_abs = p == 1 ? 4.0 * ELECTRON_MASS : 0.0;
  }
  const double qlow = std::max(xlow * K, qmin_abs);
  const double qhigh = std::min(xhigh * K, qmax_abs);
  if (qlow >= qhigh) return 0.0;
  const int n_int = (npoints + 8) / 9;
  const GaussRule& g = gauss(9);
  double s = 0.0;
  const double lql = std::log(qlow), lqh = std::log(qhigh);
  for (int itv = 1; itv <= n_int; ++itv) {
    const double a = lql + (itv - 1) * (lqh - lql) / n_int;
    const double b = lql + itv * (lqh - lql) / n_int;
    const double dx = b - a;
    for (int j = 0; j < 9; ++j) {
      const double xi = std::exp(a + dx * g.x[j]);
      const double wi = g.w[j] * dx * xi;
      double y = dcs(p, Z, A, 0.0, mass, K, xi);
      if (mode >= 1) y *= xi;
      if (mode >= 2) y *= xi;
      s += y * wi;
    }
  }
  return s * AVOGADRO_NUMBER / (A * 1e-3);
}

}  // namespace dcs

using dcs::dilog;
inline double dcs_brems(double Z, double A, double m, double K, double q) { return dcs::brems(Z, A, m, K, q); }
inline double dcs_pair(double Z, double A, double m, double K, double q) { return dcs::pair(Z, A, m, K, q); }
inline double dcs_photo(double Z, double A, double m, double K, double q) { return dcs::photo(Z, A, m, K, q); }
inline double dcs_ionisation(double Z, double I, double m, double K, double q) {
  return dcs::ionisation(Z, I, m, K, q);
}

}  // namespace orc

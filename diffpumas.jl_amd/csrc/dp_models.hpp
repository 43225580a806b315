// Physics models shared by the host table builder and the device kernels:
// differential cross-sections of the radiative and ionisation processes and
// their energy-moment integrals.
//
// Models (PUMAS 1.2 physics as restated by the reference):
//   bremsstrahlung  SSR              src/Materials.jl:631-695
//   pair production SSR (+ 12-pt GQ) src/Materials.jl:701-832
//   photonuclear    DRSS / ALLM97    src/Materials.jl:838-926
//   ionisation      Salvat 2013      src/Materials.jl:534-586
//   moments         compute_dcs_integral, src/Materials.jl:939-1007
// Units: GeV, m, kg.  dσ/dq per atom in m^2/GeV.
#pragma once
#include <stdint.h>
#include <math.h>
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define DPM_HD __host__ __device__ inline
#define DPM_HD_COLD __host__ __device__ inline __attribute__((noinline))
#define DPM_HD_HOT __host__ __device__ inline __attribute__((always_inline))
#else
#define DPM_HD_HOT inline
#define DPM_HD inline
#define DPM_HD_COLD inline
#endif

namespace dpm {

// src/Constants.jl:15-62
constexpr double ALPHA_EM = 7.2973525693e-3;
constexpr double HBAR_C = 1.973269804e-16;
constexpr double BOHR_RADIUS = 0.529177210903e-10;
constexpr double ELECTRON_MASS = 0.510998910e-3;
constexpr double ELECTRON_RADIUS = 2.817940285e-15;
constexpr double PROTON_MASS = 0.938272;
constexpr double NEUTRON_MASS = 0.939565;
constexpr double PION_MASS = 0.13957018;
constexpr double AVOGADRO_NUMBER = 6.02214076e23;
constexpr double PI = 3.141592653589793;

enum Process { BREMS = 0, PAIR = 1, PHOTO = 2, IONISATION = 3 };

// ---------------------------------------------- Gauss-Legendre on [0, 1]
// PUMAS coefficients for N = 9 and N = 12 (src/Materials.jl:141-235); only
// these orders are used by the models and integrals.
DPM_HD double gq9_x(int i) {
  const double x[9] = {0.0159198802461870, 0.0819844463366821, 0.1933142836497048,
                       0.3378732882980955, 0.5000000000000000, 0.6621267117019045,
                       0.8066857163502952, 0.9180155536633179, 0.9840801197538130};
  return x[i];
}
DPM_HD double gq9_w(int i) {
  const double w[9] = {0.0406371941807872, 0.0903240803474287, 0.1303053482014677,
                       0.1561735385200015, 0.1651196775006299, 0.1561735385200015,
                       0.1303053482014677, 0.0903240803474287, 0.0406371941807872};
  return w[i];
}
DPM_HD double gq12_x(int i) {
  const double x[12] = {0.0092196828766404, 0.0479413718147625, 0.1150486629028477,
                        0.2063410228566913, 0.3160842505009099, 0.4373832957442655,
                        0.5626167042557344, 0.6839157494990901, 0.7936589771433087,
                        0.8849513370971523, 0.9520586281852375, 0.9907803171233596};
  return x[i];
}
DPM_HD double gq12_w(int i) {
  const double w[12] = {0.0235876681932559, 0.0534696629976592, 0.0800391642716731,
                        0.1015837133615330, 0.1167462682691774, 0.1245735229067014,
                        0.1245735229067014, 0.1167462682691774, 0.1015837133615330,
                        0.0800391642716731, 0.0534696629976592, 0.0235876681932559};
  return w[i];
}

// ------------------------------------------------------ dilogarithm
// Chebyshev expansion of PUMAS math_dilog (src/Materials.jl:272-339).
DPM_HD_HOT double dilog(double x) {
  const double c[20] = {0.42996693560813697, 0.40975987533077105, -0.01858843665014592,
                        0.00145751084062268, -0.00014304184442340, 0.00001588415541880,
                        -0.00000190784959387, 0.00000024195180854, -0.00000003193341274,
                        0.00000000434545063, -0.00000000060578480, 0.00000000008612098,
                        -0.00000000001244332, 0.00000000000182256, -0.00000000000027007,
                        0.00000000000004042, -0.00000000000000610, 0.00000000000000093,
                        -0.00000000000000014, 0.00000000000000002};
  const double PI6 = PI * PI / 6.0, PI12 = PI * PI / 12.0, PI3 = PI * PI / 3.0;
  if (x == 1.0) return PI6;
  if (x == -1.0 || (1.0 - x) == 0.0) return -PI12;
  double T = -x, Y, S, A;
  if (T <= -2.0) {
    Y = -1.0 / (1.0 + T); S = 1.0;
    double l1 = log(-T), l2 = log(1.0 + 1.0 / T);
    A = -PI3 + 0.5 * (l1 * l1 - l2 * l2);
  } else if (T < -1.0) {
    Y = -1.0 - T; S = -1.0;
    double la = log(-T);
    A = -PI6 + la * (la + log(1.0 + 1.0 / T));
  } else if (T <= -0.5) {
    Y = -(1.0 + T) / T; S = 1.0;
    double la = log(-T);
    A = -PI6 + la * (-0.5 * la + log(1.0 + T));
  } else if (T < 0.0) {
    Y = -T / (1.0 + T); S = -1.0;
    double l = log(1.0 + T);
    A = 0.5 * l * l;
  } else if (T <= 1.0) {
    Y = T; S = 1.0; A = 0.0;
  } else {
    Y = 1.0 / T; S = -1.0;
    double l = log(T);
    A = PI6 + 0.5 * l * l;
  }
  double H = Y + Y - 1.0, ALFA = H + H, B1 = 0.0, B2 = 0.0;
#pragma unroll
  for (int i = 19; i >= 0; --i) {
    double B0 = c[i] + ALFA * B1 - B2;
    B2 = B1;
    B1 = B0;
  }
  return -(S * (B1 - H * B2) + A);
}

// radiation_logarithm, src/Materials.jl:592-620 (elements without a tabulated
// value use 182.7)
DPM_HD double radiation_logarithm(double Z) {
  switch ((int)rint(Z)) {
    case 1: return 202.4; case 2: return 151.9; case 3: return 159.9; case 4: return 172.3;
    case 5: return 177.9; case 6: return 178.3; case 7: return 176.6; case 8: return 173.4;
    case 9: return 170.0; case 10: return 165.8; case 11: return 165.8; case 12: return 167.1;
    case 13: return 169.1; case 14: return 170.8; case 15: return 172.2; case 16: return 173.4;
    case 17: return 174.3; case 18: return 174.8; case 19: return 175.1; case 20: return 175.6;
    case 21: return 176.2; case 22: return 176.8; case 26: return 175.8; case 29: return 173.1;
    case 32: return 173.0; case 35: return 173.5; case 42: return 175.9; case 50: return 177.4;
    case 53: return 178.6; case 74: return 177.6; case 82: return 178.0; case 92: return 179.8;
    default: return 182.7;
  }
}

// ------------------------------------------------------ bremsstrahlung
// dcs_bremsstrahlung_ssr, src/Materials.jl:631-695
DPM_HD_COLD double dcs_brems(double Z, double A, double m, double K, double q) {
  const double SQRTE = 1.648721270700128, ME = 0.5109989461, RE = 2.8179403227e-13,
               MMU = 105.6583745;
  if (Z <= 0 || A <= 0 || m <= 0 || K <= 0 || q <= 0) return 0.0;
  const double z13 = pow(Z, 1.0 / 3.0);
  if (q >= K + m * (1.0 - 0.75 * SQRTE * z13)) return 0.0;
  const double energy = (K + m) * 1e3;
  const double v = q * 1e3 / energy;
  const double mm = m * 1e3;
  const double Z13 = 1.0 / z13;
  const double rl = radiation_logarithm(Z);
  const double rli = Z == 1.0 ? 446.0 : 1429.0;
  const double Dn = 1.54 * pow(A, 0.27);
  const double mu_qc = mm / (MMU * exp(1.0) / Dn);
  const double rho = sqrt(1.0 + 4.0 * mu_qc * mu_qc);
  const double log_rho = log((rho + 1.0) / (rho - 1.0));
  const double delta1 = log(mu_qc) + 0.5 * rho * log_rho;
  const double delta2 = log(mu_qc) + 0.25 * (3.0 * rho - rho * rho * rho) * log_rho + 2.0 * mu_qc * mu_qc;
  const double delta = mm * mm * v / (2.0 * energy * (1.0 - v));
  double phi1 = log(rl * Z13 * (mm / ME) / (1.0 + rl * Z13 * exp(0.5) * delta / ME));
  double phi2 = log(rl * Z13 * exp(-1.0 / 6.0) * (mm / ME) /
                    (1.0 + rl * Z13 * exp(1.0 / 3.0) * delta / ME));
  phi1 -= delta1 * (1.0 - 1.0 / Z);
  phi2 -= delta2 * (1.0 - 1.0 / Z);
  const double sa1 = log(mm / delta / (mm * delta / (ME * ME) + SQRTE));
  const double sa2 = log(1.0 + ME / (delta * rli * Z13 * Z13 * SQRTE));
  const double s_atomic = (4.0 / 3.0 * (1.0 - v) + v * v) * (sa1 - sa2);
  double s_rad;
  if (v < 0.0 || v > 1.0) {
    s_rad = 0.0;
  } else if (v < 0.02) {
    s_rad = -0.00349 + 148.84 * v + -987.531 * v * v;
  } else if (v < 0.1) {
    s_rad = 0.1642 + 132.573 * v + -585.361 * v * v + 1407.77 * v * v * v;
  } else if (v < 0.9) {
    s_rad = -2.8922 + -19.0156 * v + 57.698 * v * v;
    double tmp = log(1.0 - v);
    s_rad += -63.418 * v * log(v) + 14.1166 * tmp + 1.84206 * tmp * tmp;
  } else {
    s_rad = 2134.19 + 581.823 * v + -2708.85 * v * v;
    double tmp = log(1.0 - v);
    s_rad += 4767.05 * v * log(v) + 1.52918 * tmp + 0.361933 * tmp * tmp;
  }
  const double result = ((2.0 - 2.0 * v + v * v) * phi1 - 2.0 / 3.0 * (1.0 - v) * phi2) +
                        1.0 / Z * s_atomic + 0.25 * ALPHA_EM * phi1 * s_rad;
  if (result <= 0.0) return 0.0;
  const double aux = 2.0 * (ME / mm) * RE * Z;
  return aux * aux * (ALPHA_EM / q) * result * 1e-4;
}

// ----------------------------------------------------- pair production
// _dcs_pair_production_d2_ssr, src/Materials.jl:701-800, split into the terms
// that do not depend on the quadrature variable (PairInv, evaluated once per
// dcs_pair call) and the per-node part.  Each term is the same expression as
// in the single-function form, so the values are unchanged.
struct PairInv {
  double Z, v, rl, cpf, Z13, d_n, zeta, beta;
  // products of the above that every node uses (the reference evaluates this
  // model under @fastmath, which permits regrouping them)
  double kxi;       // (m v / 2 m_e)^2 / (1 - v): xi = kxi (1 - rho^2)
  double kq1, kq2;  // 2 m_e e^{1/2} rl Z^{-1/3} / (E v) and the same with e^{1/3}
  double rz;        // rl Z^{-1/3}
  double lm0;       // (m / m_e) rl Z^{-1/3} / d_n
  double dmm;       // (m_e / m d_n)^2
  double pe, pm;    // cpf (Z + zeta) (1 - v) / v, and times (m_e / m)^2
  double scale;     // 0.1 / E
};
DPM_HD PairInv pair_invariants(double Z, double A, double mass, double K, double q) {
  const double ME = 0.5109989461, RE = 2.8179403227e-13;
  PairInv c;
  c.Z = Z;
  const double energy = (K + mass) * 1e3;
  c.v = q / (K + mass);
  const double m = mass * 1e3;
  c.rl = radiation_logarithm(Z);
  c.cpf = 4.0 / (3.0 * PI) * Z * (ALPHA_EM * RE) * (ALPHA_EM * RE);
  c.Z13 = pow(Z, -1.0 / 3.0);
  c.d_n = 1.54 * pow(A, 0.27);
  const double Z13 = c.Z13, v = c.v;
  const double g1 = Z == 1.0 ? 4.4e-5 : 1.95e-5;
  const double g2 = Z == 1.0 ? 4.8e-5 : 5.3e-5;
  const double zeta1 = 0.073 * log(energy / m / (1.0 + g1 * pow(Z, 2.0 / 3.0) * energy / m)) - 0.26;
  const double zeta2 = 0.058 * log(energy / m / (1.0 + g2 / Z13 * energy / m)) - 0.14;
  c.zeta = (zeta1 > 0.0 && zeta2 > 0.0) ? zeta1 / zeta2 : 0.0;
  c.beta = v * v / (2.0 * (1.0 - v));
  const double mv = m * v / (2.0 * ME);
  c.kxi = mv * mv / (1.0 - v);
  c.rz = c.rl * Z13;
  const double ev = energy * v;
  c.kq1 = 2.0 * ME * exp(0.5) * c.rz / ev;
  c.kq2 = 2.0 * ME * exp(1.0 / 3.0) * c.rz / ev;
  c.lm0 = m / ME * c.rz / c.d_n;
  const double md = ME / m * c.d_n;
  c.dmm = md * md;
  c.pe = c.cpf * (Z + c.zeta) * (1.0 - v) / v;
  c.pm = c.pe * (ME / m) * (ME / m);
  c.scale = 0.1 / energy;
  return c;
}
// _dcs_pair_production_d2_ssr at rho = 1 - rho_in, with the terms shared by
// several expressions (1/(1+xi), 1/xi, 1/(1-rho^2)) formed once.
DPM_HD double dcs_pair_d2(const PairInv& c, double rho_in) {
  const double beta = c.beta;
  const double rho = 1.0 - rho_in;
  const double rho2 = rho * rho;
  const double omr = 1.0 - rho2;
  const double xi = c.kxi * omr;
  const double ixi = 1.0 / xi;
  const double opx = 1.0 + xi;
  const double iopx = 1.0 / opx;
  const double iomr = 1.0 / omr;
  const double lxi = log(1.0 + ixi);
  const double Be = ((2.0 + rho2) * (1.0 + beta) + xi * (3.0 + rho2)) * lxi +
                    (omr - beta) * iopx - (3.0 + rho2);
  const double Ce2 = (omr * (1.0 + beta) + xi * (3.0 - rho2)) * lxi +
                     2.0 * (1.0 - beta - rho2) * iopx - (3.0 - rho2);
  const double Ce1 = Be - Ce2;
  const double De = ((2.0 + rho2) * (1.0 + beta) + xi * (3.0 + rho2)) * dilog(iopx) -
                    (2.0 + rho2) * xi * lxi - (xi + rho2 + beta) * iopx;
  const double q1 = c.kq1 * opx * iomr, q2 = c.kq2 * opx * iomr;   // 2 m_e e^k rl Z13 (1+xi)/den
  const double sq = c.rz * sqrt(opx);
  const double tmpe = De / Be;
  double Le1, Le2;
  if (tmpe > 0.0) {
    const double Xe = tmpe < 100.0 ? exp(-tmpe) : 0.0;
    Le1 = log(sq / (Xe + q1)) - tmpe - 0.5 * log(Xe + c.dmm * opx);
    Le2 = log(exp(-1.0 / 6.0) * sq / (Xe + q2)) - tmpe -
          0.5 * log(Xe + c.dmm * exp(-1.0 / 3.0) * opx);
  } else {
    const double Xi = tmpe > -100.0 ? exp(tmpe) : 0.0;
    Le1 = log(sq / (1.0 + Xi * q1)) - 0.5 * tmpe - 0.5 * log(1.0 + Xi * c.dmm * opx);
    Le2 = log(exp(-1.0 / 6.0) * sq / (1.0 + Xi * q2)) - 0.5 * tmpe -
          0.5 * log(1.0 + Xi * c.dmm * exp(-1.0 / 3.0) * opx);
  }
  double diag_e = c.pe * (Ce1 * Le1 + Ce2 * Le2);
  if (diag_e < 0.0) diag_e = 0.0;
  const double l1xi = log(opx);
  const double bm0 = (1.0 + rho2) * (1.0 + 1.5 * beta) - ixi * (1.0 + 2.0 * beta) * omr;
  const double Bm = bm0 * l1xi + xi * (omr - beta) * iopx + (1.0 + 2.0 * beta) * omr;
  const double Cm2 = ((1.0 - beta) * omr - xi * (1.0 + rho2)) * l1xi * ixi -
                     2.0 * (1.0 - beta - rho2) * iopx + 1.0 - beta - (1.0 + beta) * rho2;
  const double Cm1 = Bm - Cm2;
  const double Dm = bm0 * dilog(xi * iopx) + (1.0 + 1.5 * beta) * omr * ixi * l1xi +
                    (omr - 0.5 * beta * (1.0 + rho2) + 0.5 * omr * ixi * beta) * xi * iopx;
  const double tmpm = Dm / Bm;
  double Lm1, Lm2;
  if (tmpm > 0.0) {
    const double Xm = tmpm < 100.0 ? exp(-tmpm) : 0.0;
    Lm1 = log(c.lm0 / (Xm + q1)) - tmpm;
    Lm2 = log(c.lm0 / (Xm + q2)) - tmpm;
  } else {
    const double Xmv = tmpm > -100.0 ? exp(tmpm) : 0.0;
    Lm1 = log(c.lm0 / (1.0 + q1 * Xmv));
    Lm2 = log(c.lm0 / (1.0 + q2 * Xmv));
  }
  double diag_mu = c.pm * (Cm1 * Lm1 + Cm2 * Lm2);
  if (diag_mu < 0.0) diag_mu = 0.0;
  return (diag_e + diag_mu) * c.scale;
}

// dcs_pair_production_ssr, src/Materials.jl:802-832, split into the
// quadrature's range (ok = false: the cross-section is 0) and its node terms,
// so that the nodes can also be evaluated apart and summed in node order.
struct PairQuad {
  double tmin;
  bool ok;
};
DPM_HD PairQuad pair_quad(double Z, double A, double mass, double K, double q) {
  PairQuad p{0.0, false};
  if (Z <= 0 || A <= 0 || mass <= 0 || K <= 0 || q <= 0) return p;
  if (q <= 4.0 * ELECTRON_MASS) return p;
  const double sqrte = 1.6487212707;
  const double Z13 = pow(Z, 1.0 / 3.0);
  if (q >= K + mass * (1.0 - 0.75 * sqrte * Z13)) return p;
  const double gamma = 1.0 + K / mass;
  const double x0 = 1.0 - 4.0 * ELECTRON_MASS / q;
  const double x1 = 1.0 - 6.0 / (gamma * (gamma - q / mass));
  if (!(x0 > 0.0 && x1 > 0.0)) return p;
  const double rmax = sqrt(x0) * x1;
  const double ri = 1.0 - rmax;
  if (ri <= 0.0 || ri >= 1.0) return p;
  p.tmin = log(ri);
  p.ok = true;
  return p;
}
// Term of node i: the quadrature is I = 0 - term(0) - term(1) - ... - term(11).
DPM_HD double pair_node(const PairInv& c, double tmin, int i) {
  const double r = exp(gq12_x(i) * tmin);
  return dcs_pair_d2(c, r) * r * gq12_w(i) * tmin;
}
DPM_HD_COLD double dcs_pair(double Z, double A, double mass, double K, double q) {
  const PairQuad p = pair_quad(Z, A, mass, K, q);
  if (!p.ok) return 0.0;
  const PairInv c = pair_invariants(Z, A, mass, K, q);
  double I = 0.0;
  // Rolled on purpose: twelve inlined copies of the DDCS overflow the device
  // register file and spill around every node; one rolled body with the
  // dilogarithm inlined keeps its state in registers.
#pragma unroll 1
  for (int i = 0; i < 12; ++i) I -= pair_node(c, p.tmin, i);
  return I < 0.0 ? 0.0 : I;
}
// One node term of dcs_pair on its own (0 where the cross-section is 0).
DPM_HD_COLD double dcs_pair_node(double Z, double A, double mass, double K, double q, int i) {
  const PairQuad p = pair_quad(Z, A, mass, K, q);
  if (!p.ok) return 0.0;
  return pair_node(pair_invariants(Z, A, mass, K, q), p.tmin, i);
}

// -------------------------------------------------------- photonuclear
// ALLM97 proton structure function, src/Materials.jl:838-863
DPM_HD double f2p_allm97(double x, double Q2) {
  const double m02 = 0.31985, mP2 = 49.457, mR2 = 0.15052, Q02 = 0.52544, L2 = 0.06527;
  const double cP1 = 0.28067, cP2 = 0.22291, cP3 = 2.1979;
  const double aP1 = -0.0808, aP2 = -0.44812, aP3 = 1.1709;
  const double bP1 = 0.36292, bP2 = 1.8917, bP3 = 1.8439;
  const double cR1 = 0.80107, cR2 = 0.97307, cR3 = 3.4942;
  const double aR1 = 0.58400, aR2 = 0.37888, aR3 = 2.6063;
  const double bR1 = 0.01147, bR2 = 3.7582, bR3 = 0.49338;
  const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
  const double M2 = M * M;
  const double W2 = M2 + Q2 * (1.0 / x - 1.0);
  const double t = log(log((Q2 + Q02) / L2) / log(Q02 / L2));
  const double xP = (Q2 + mP2) / (Q2 + mP2 + W2 - M2);
  const double xR = (Q2 + mR2) / (Q2 + mR2 + W2 - M2);
  // t^c as exp(c log t): the reference evaluates these powers under @fastmath,
  // which permits this form; t >= 0, and t = 0 gives 0 either way.
  const double lt = log(t);
  const double cP = cP1 + (cP1 - cP2) * (1.0 / (1.0 + exp(cP3 * lt)) - 1.0);
  const double aP = aP1 + (aP1 - aP2) * (1.0 / (1.0 + exp(aP3 * lt)) - 1.0);
  const double bP = bP1 + bP2 * exp(bP3 * lt);
  const double cR = cR1 + cR2 * exp(cR3 * lt);
  const double aR = aR1 + aR2 * exp(aR3 * lt);
  const double bR = bR1 + bR2 * exp(bR3 * lt);
  const double F2P = cP * exp(aP * log(xP) + bP * log(1.0 - x));
  const double F2R = cR * exp(aR * log(xR) + bR * log(1.0 - x));
  return Q2 / (Q2 + m02) * (F2P + F2R);
}
// DRSS shadowing and nuclear structure function, src/Materials.jl:865-879
DPM_HD double shadowing_drss(double Z, double A, double x) {
  if (Z == 1.0) return 1.0;
  if (x < 0.0014) return exp(-0.1 * log(A));
  if (x < 0.04) return exp((0.069 * log10(x) + 0.097) * log(A));
  return 1.0;
}
// src/Materials.jl:881-895
DPM_HD double photo_d2(double Z, double A, double ml, double K, double q, double Q2) {
  const double cf = 4.0 * PI * ALPHA_EM * ALPHA_EM * HBAR_C * HBAR_C;
  const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
  const double E = K + ml;
  const double y = q / E;
  const double x = 0.5 * Q2 / (M * q);
  const double F2p = f2p_allm97(x, Q2);
  const double sh = shadowing_drss(Z, A, x);
  const double F2A = F2p * sh * (Z + (A - Z) * (1.0 + x * (-1.85 + x * (2.45 + x * (-2.35 + x)))));
  const double dds = (1.0 - y + 0.5 * (1.0 - 2.0 * ml * ml / Q2) * (y * y + Q2 / (E * E))) / (Q2 * Q2) -
                     0.25 / (E * E * Q2);
  return cf * F2A * dds / q;
}
// dcs_photonuclear_drss, src/Materials.jl:897-926, split like dcs_pair.
struct PhotoQuad {
  double lmin, dx;
  bool ok;
};
DPM_HD PhotoQuad photo_quad(double Z, double A, double m, double K, double q) {
  PhotoQuad p{0.0, 0.0, false};
  if (Z <= 0 || A <= 0 || m <= 0 || K <= 0 || q <= 0) return p;
  const double M = 0.5 * (NEUTRON_MASS + PROTON_MASS);
  const double mpi = PION_MASS;
  const double qmin = mpi + 0.5 * mpi * mpi / M;
  const double qmax = K + m - 0.5 * (M + m * m / M);
  if (q <= qmin || q >= qmax) return p;
  const double E = K + m;
  const double ml2 = m * m;
  const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
  const double Q2max = 2.0 * M * (q - mpi) - mpi * mpi;
  if (Q2max < Q2min || Q2min < 0.0) return p;
  p.lmin = log(Q2min);
  p.dx = log(Q2max) - p.lmin;
  p.ok = true;
  return p;
}
// Term of node i (the reference skips a node whose DDCS is not positive; its
// term here is +0, and adding +0 to the non-negative running sum changes
// nothing): ds = 0 + term(0) + ... + term(8).
DPM_HD double photo_node(double Z, double A, double m, double K, double q, const PhotoQuad& p,
                         int i) {
  const double Q2 = exp(p.lmin + p.dx * gq9_x(i));
  const double v = photo_d2(Z, A, m, K, q, Q2);
  return v > 0.0 ? v * (gq9_w(i) * p.dx * Q2) : 0.0;
}
DPM_HD_COLD double dcs_photo(double Z, double A, double m, double K, double q) {
  const PhotoQuad p = photo_quad(Z, A, m, K, q);
  if (!p.ok) return 0.0;
  double ds = 0.0;
#pragma unroll 1
  for (int i = 0; i < 9; ++i) ds += photo_node(Z, A, m, K, q, p, i);
  return ds < 0.0 ? 0.0 : ds;
}
// One node term of dcs_photo on its own (0 where the cross-section is 0).
DPM_HD_COLD double dcs_photo_node(double Z, double A, double m, double K, double q, int i) {
  const PhotoQuad p = photo_quad(Z, A, m, K, q);
  if (!p.ok) return 0.0;
  return photo_node(Z, A, m, K, q, p, i);
}

// -------------------------------------------------------- ionisation
// electronic_dcs, src/Materials.jl:534-548
DPM_HD_COLD double dcs_ionisation(double Z, double I, double m, double K, double q) {
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

// dcs_ionisation_integrate, src/Materials.jl:556-586 (analytic moments)
DPM_HD double ionisation_integral(double Z, double A, double I, double mass, double K,
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
  const double E2 = E * E;
  const double beta2 = P2 / E2;
  const double a0 = 0.5 / E2;
  const double a1 = beta2 / Wr;
  double Iv;
  if (mode == 0)
    Iv = a0 * (Wmax - Wmin) - a1 * log(Wmax / Wmin) + 1.0 / Wmin - 1.0 / Wmax;
  else if (mode == 1)
    Iv = 0.5 * a0 * (Wmax * Wmax - Wmin * Wmin) - a1 * (Wmax - Wmin) + log(Wmax / Wmin);
  else
    Iv = a0 * (Wmax * Wmax * Wmax - Wmin * Wmin * Wmin) / 3.0 - 0.5 * a1 * (Wmax * Wmax - Wmin * Wmin) + Wmax - Wmin;
  return 2.0 * PI * ELECTRON_MASS * ELECTRON_RADIUS * ELECTRON_RADIUS * Z * AVOGADRO_NUMBER /
         (A * 1e-3 * beta2) * Iv;
}

// Per-atom dσ/dq of a radiative process.
DPM_HD double dcs_radiative(int proc, double Z, double A, double m, double K, double q) {
  if (proc == BREMS) return dcs_brems(Z, A, m, K, q);
  if (proc == PAIR) return dcs_pair(Z, A, m, K, q);
  if (proc == PHOTO) return dcs_photo(Z, A, m, K, q);
  return 0.0;
}

// compute_dcs_integral, src/Materials.jl:939-1007: moment `mode` (0, 1, 2)
// of dσ/dq on q in [xlow K, xhigh K] (clipped to kinematic limits), per kg
// of the element.  Log-spaced 9-point Gauss-Legendre on ceil(npoints/9)
// intervals.
DPM_HD double dcs_integral(int proc, double Z, double A, double I, double mass, double K,
                           double xlow, double xhigh, int mode, int npoints) {
  if (xlow <= 0.0) xlow = 1e-6;
  if (proc == IONISATION) return ionisation_integral(Z, A, I, mass, K, xlow, xhigh, mode);
  const double M = 0.5 * (NEUTRON_MASS + PROTON_MASS);
  const double mpi = PION_MASS;
  double qmin, qmax;
  if (proc == PHOTO) {
    qmin = mpi + 0.5 * mpi * mpi / M;
    qmax = K + mass - 0.5 * (M + mass * mass / M);
  } else {
    const double Z13 = pow(Z, 1.0 / 3.0);
    qmax = K + mass * (1.0 - 0.75 * 1.648721271 * Z13);
    qmin = proc == PAIR ? 4.0 * ELECTRON_MASS : 0.0;
  }
  const double qlow = fmax(xlow * K, qmin);
  const double qhigh = fmin(xhigh * K, qmax);
  if (qlow >= qhigh) return 0.0;
  const int nint = (npoints + 8) / 9;
  double s = 0.0;
  const double lql = log(qlow), lqh = log(qhigh);
  for (int it = 1; it <= nint; ++it) {
    const double a = lql + (it - 1) * (lqh - lql) / nint;
    const double b = lql + it * (lqh - lql) / nint;
    const double dx = b - a;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const double xi = exp(a + dx * gq9_x(j));
      const double wi = gq9_w(j) * dx * xi;
      double y = dcs_radiative(proc, Z, A, mass, K, xi);
      if (mode >= 1) y *= xi;
      if (mode >= 2) y *= xi;
      s += y * wi;
    }
  }
  return s * AVOGADRO_NUMBER / (A * 1e-3);
}

}  // namespace dpm

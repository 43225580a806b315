// Host-side table builder: MaterialTable arrays from a PUMAS dE/dx table and
// an element recipe (create_material_table_from_dedx, src/Physics.jl:146-273),
// composite tables (create_composite_table, :421-525) and the Coulomb
// scattering tables (compute_coulomb_tables!, src/Coulomb.jl:688-826).
// Runs once per physics object; the transport kernels only read its output.
#include "dp_internal.hpp"
#include "dp_models.hpp"
#include "dp_atomic_shells.hpp"
#include <algorithm>
#include <cmath>
#include <vector>

using namespace dpm;

namespace {

constexpr double LARMOR_FACTOR = 0.299792458;   // src/Constants.jl:29
constexpr double EHS_PATH_MAX = 1e9;            // src/Constants.jl:80
constexpr double MAX_SOFT_ANGLE = 1.0;          // src/Constants.jl:89

struct Elem {
  double Z, A, I, f;
};

// ------------------------------------------------------------ Coulomb data
// Salvat et al. 1987 screening functions and nuclear radii, as tabulated by
// PUMAS (src/Coulomb.jl:25-160).
const double SALVAT_A1[103] = {
    1.00000e+00, -2.25920e-01, 6.04537e-01, 3.27766e-01, 2.32684e-01, 1.53676e-01, 9.95750e-02,
    6.25130e-02, 3.68040e-02, 1.88410e-02, 7.44440e-01, 6.42349e-01, 6.00152e-01, 5.15971e-01,
    4.38675e-01, 5.45871e-01, 7.24889e-01, 2.19124e+00, 4.85607e-02, 5.80017e-01, 5.54340e-01,
    1.11950e-02, 3.18350e-02, 1.07503e-01, 4.97556e-02, 5.11841e-02, 5.00039e-02, 4.73509e-02,
    7.70967e-02, 4.00041e-02, 1.08344e-01, 6.09767e-02, 2.11561e-02, 4.83575e-01, 4.50364e-01,
    4.19036e-01, 1.73438e-01, 3.35694e-02, 6.88939e-02, 1.17552e-01, 2.55689e-01, 2.69313e-01,
    2.20138e-01, 2.75057e-01, 2.71053e-01, 2.78363e-01, 2.56210e-01, 2.27100e-01, 2.49215e-01,
    2.15313e-01, 1.80560e-01, 1.30772e-01, 5.88293e-02, 4.45145e-01, 2.70796e-01, 1.72814e-01,
    1.94726e-01, 1.91338e-01, 1.86776e-01, 1.66461e-01, 1.62350e-01, 1.58016e-01, 1.53759e-01,
    1.58729e-01, 1.45327e-01, 1.41260e-01, 1.37360e-01, 1.33614e-01, 1.29853e-01, 1.26659e-01,
    1.28806e-01, 1.30256e-01, 1.38420e-01, 1.50030e-01, 1.60803e-01, 1.72164e-01, 1.83411e-01,
    2.23043e-01, 2.28909e-01, 2.09753e-01, 2.70821e-01, 2.37958e-01, 2.28771e-01, 1.94059e-01,
    1.49995e-01, 9.55262e-02, 3.19155e-01, 2.40406e-01, 2.26579e-01, 2.17619e-01, 2.41294e-01,
    2.44758e-01, 2.46231e-01, 2.55572e-01, 2.53567e-01, 2.43832e-01, 2.41898e-01, 2.44050e-01,
    2.40237e-01, 2.34997e-01, 2.32114e-01, 2.27937e-01, 2.29571e-01};
const double SALVAT_A2[103] = {
    0.00000e+00, 1.22592e+00, 3.95463e-01, 6.72234e-01, 7.67316e-01, 8.46324e-01, 9.00425e-01,
    9.37487e-01, 9.63196e-01, 9.81159e-01, 2.55560e-01, 3.57651e-01, 3.99848e-01, 4.84029e-01,
    5.61325e-01, -5.33329e-01, -7.54809e-01, -2.28520e+00, 7.75935e-01, 4.19983e-01, 4.45660e-01,
    6.83176e-01, 6.75303e-01, 7.16172e-01, 6.86632e-01, 6.99533e-01, 7.14201e-01, 7.29404e-01,
    7.95083e-01, 7.59034e-01, 7.48941e-01, 7.15671e-01, 6.70932e-01, 5.16425e-01, 5.49636e-01,
    5.80964e-01, 7.25336e-01, 7.81581e-01, 7.20203e-01, 6.58088e-01, 5.82051e-01, 5.75262e-01,
    5.61797e-01, 5.94338e-01, 6.11921e-01, 6.06653e-01, 6.50520e-01, 6.15496e-01, 6.43990e-01,
    6.11497e-01, 5.76688e-01, 5.50366e-01, 5.48174e-01, 5.54855e-01, 6.52415e-01, 6.84485e-01,
    6.38429e-01, 6.46684e-01, 6.55810e-01, 7.05677e-01, 7.13311e-01, 7.20978e-01, 7.28385e-01,
    7.02414e-01, 7.42619e-01, 7.49352e-01, 7.55797e-01, 7.61947e-01, 7.68005e-01, 7.73365e-01,
    7.52781e-01, 7.32428e-01, 7.09596e-01, 6.87141e-01, 6.65932e-01, 6.46849e-01, 6.30598e-01,
    6.17575e-01, 6.11402e-01, 6.00426e-01, 6.42829e-01, 6.30789e-01, 6.21959e-01, 6.10455e-01,
    6.03147e-01, 6.05994e-01, 6.23324e-01, 6.56665e-01, 6.42246e-01, 6.24013e-01, 6.30394e-01,
    6.29816e-01, 6.31596e-01, 6.49005e-01, 6.53604e-01, 6.43738e-01, 6.48850e-01, 6.70318e-01,
    6.76319e-01, 6.65571e-01, 6.88406e-01, 6.94394e-01, 6.82014e-01};
const double SALVAT_B1[103] = {
    1.11728e+00, 5.52725e+00, 2.81741e+00, 4.54302e+00, 5.99006e+00, 8.04043e+00, 1.08122e+01,
    1.48233e+01, 2.14001e+01, 3.49994e+01, 4.12050e+00, 4.72663e+00, 5.14051e+00, 5.84918e+00,
    6.67070e+00, 6.37029e+00, 6.21183e+00, 5.54701e+00, 3.02597e+01, 6.32184e+00, 6.63280e+00,
    9.97569e+01, 4.25330e+01, 1.89587e+01, 3.18642e+01, 3.18251e+01, 3.29153e+01, 3.47580e+01,
    2.53264e+01, 4.03429e+01, 2.01922e+01, 2.91996e+01, 6.24873e+01, 8.78242e+00, 9.33480e+00,
    9.91420e+00, 1.71659e+01, 5.52077e+01, 3.13659e+01, 2.20537e+01, 1.42403e+01, 1.40442e+01,
    1.59176e+01, 1.43137e+01, 1.46537e+01, 1.46455e+01, 1.55878e+01, 1.69141e+01, 1.61552e+01,
    1.77931e+01, 1.98751e+01, 2.41540e+01, 3.99955e+01, 1.18053e+01, 1.65915e+01, 2.23966e+01,
    2.07637e+01, 2.12350e+01, 2.18033e+01, 2.39492e+01, 2.45984e+01, 2.52966e+01, 2.60169e+01,
    2.54973e+01, 2.75466e+01, 2.83460e+01, 2.91604e+01, 2.99904e+01, 3.08345e+01, 3.16806e+01,
    3.13526e+01, 3.12166e+01, 3.00767e+01, 2.86302e+01, 2.75684e+01, 2.65861e+01, 2.57339e+01,
    2.29939e+01, 2.28644e+01, 2.44080e+01, 2.09409e+01, 2.29872e+01, 2.37917e+01, 2.66951e+01,
    3.18397e+01, 4.34890e+01, 2.00150e+01, 2.45012e+01, 2.56843e+01, 2.65542e+01, 2.51930e+01,
    2.52522e+01, 2.54271e+01, 2.51526e+01, 2.55959e+01, 2.65567e+01, 2.70360e+01, 2.72673e+01,
    2.79152e+01, 2.86446e+01, 2.93353e+01, 3.01040e+01, 3.02650e+01};
const double SALVAT_B2[103] = {
    1.00000e+00, 2.39924e+00, 6.62463e-01, 9.85154e-01, 1.21347e+00, 1.49129e+00, 1.76868e+00,
    2.04035e+00, 2.30601e+00, 2.56621e+00, 8.71798e-01, 1.00247e+00, 1.01529e+00, 1.17314e+00,
    1.34102e+00, 2.55169e+00, 3.38827e+00, 4.56873e+00, 3.12426e+00, 1.00935e+00, 1.10227e+00,
    4.12865e+00, 3.94043e+00, 3.06375e+00, 3.78110e+00, 3.77161e+00, 3.79085e+00, 3.82989e+00,
    3.39276e+00, 3.94645e+00, 3.47325e+00, 4.12525e+00, 4.95015e+00, 1.69671e+00, 1.79002e+00,
    1.88354e+00, 3.11025e+00, 4.28418e+00, 4.24121e+00, 4.03254e+00, 2.97020e+00, 2.86107e+00,
    3.36719e+00, 2.73701e+00, 2.71828e+00, 2.61549e+00, 2.74124e+00, 3.08408e+00, 2.88189e+00,
    3.29372e+00, 3.80921e+00, 4.61191e+00, 5.91318e+00, 1.79673e+00, 2.69645e+00, 3.45951e+00,
    3.46574e+00, 3.48193e+00, 3.50982e+00, 3.51987e+00, 3.55603e+00, 3.59628e+00, 3.63834e+00,
    3.73639e+00, 3.72882e+00, 3.77625e+00, 3.82444e+00, 3.87344e+00, 3.92327e+00, 3.97271e+00,
    4.09040e+00, 4.20492e+00, 4.24918e+00, 4.24261e+00, 4.23412e+00, 4.19992e+00, 4.14615e+00,
    3.73461e+00, 3.69138e+00, 3.96429e+00, 3.24563e+00, 3.62172e+00, 3.77959e+00, 4.25824e+00,
    4.92848e+00, 5.85205e+00, 2.90906e+00, 3.55241e+00, 3.79223e+00, 4.00437e+00, 3.67795e+00,
    3.63966e+00, 3.61328e+00, 3.43021e+00, 3.43474e+00, 3.59089e+00, 3.59411e+00, 3.48061e+00,
    3.50331e+00, 3.61870e+00, 3.55697e+00, 3.58685e+00, 3.64085e+00};
const double SALVAT_B3[103] = {
    1.00000e+00, 1.00000e+00, 1.00000e+00, 1.00000e+00, 1.00000e+00, 1.00000e+00, 1.00000e+00,
    1.00000e+00, 1.00000e+00, 1.00000e+00, 1.00000e+00, 1.00000e+00, 1.00000e+00, 1.00000e+00,
    1.00000e+00, 1.67534e+00, 1.85964e+00, 2.04455e+00, 7.32637e-01, 1.00000e+00, 1.00000e+00,
    1.00896e+00, 1.05333e+00, 1.00137e+00, 1.12787e+00, 1.16064e+00, 1.19152e+00, 1.22089e+00,
    1.14261e+00, 1.27594e+00, 1.00643e+00, 1.18447e+00, 1.35819e+00, 1.00000e+00, 1.00000e+00,
    1.00000e+00, 7.17673e-01, 8.57842e-01, 9.47152e-01, 1.01806e+00, 1.01699e+00, 1.05906e+00,
    1.15477e+00, 1.10923e+00, 1.12336e+00, 1.43183e+00, 1.14079e+00, 1.26189e+00, 9.94156e-01,
    1.14781e+00, 1.28288e+00, 1.41954e+00, 1.54707e+00, 1.00000e+00, 6.81361e-01, 8.07311e-01,
    8.91057e-01, 9.01112e-01, 9.10636e-01, 8.48620e-01, 8.56929e-01, 8.65025e-01, 8.73083e-01,
    9.54998e-01, 8.88981e-01, 8.96917e-01, 9.04803e-01, 9.12768e-01, 9.20306e-01, 9.28838e-01,
    1.00717e+00, 1.09456e+00, 1.16966e+00, 1.23403e+00, 1.29699e+00, 1.35350e+00, 1.40374e+00,
    1.44284e+00, 1.48856e+00, 1.53432e+00, 1.11214e+00, 1.23735e+00, 1.25338e+00, 1.35772e+00,
    1.46828e+00, 1.57359e+00, 7.20714e-01, 8.37599e-01, 9.33468e-01, 1.02385e+00, 9.69895e-01,
    9.82474e-01, 9.92527e-01, 9.32751e-01, 9.41671e-01, 1.01827e+00, 1.02554e+00, 9.66447e-01,
    9.74347e-01, 1.04137e+00, 9.90568e-01, 9.98878e-01, 1.04473e+00};
const double NUCLEAR_RADII[120] = {
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

// nuclear_radius, src/Coulomb.jl:163-177 (index 119: fictitious "Rockium")
double nuclear_radius(double Z, double A) {
  int iN = (int)Z;
  const int nN = 120;
  if (iN >= nN - 1) iN = nN - 2;
  else if (iN == 1 && A >= 1.5) iN = 0;
  else if (iN == 11 && A == 22.0) iN = nN - 1;
  return NUCLEAR_RADII[iN] * 1e-15;
}

// frame_parameters, src/Coulomb.jl:194-210
void frame_parameters(double Z, double A, double mass, double K, double& K0, double& gCM,
                      double& tau) {
  const double Ma = Z * PROTON_MASS + (A - Z) * NEUTRON_MASS;
  const double M2 = (mass + Ma) * (mass + Ma);
  const double s = 1.0 / std::sqrt(M2 + 2.0 * Ma * K);
  gCM = (K + mass + Ma) * s;
  K0 = (K * Ma + mass * (mass + Ma)) * s - mass;
  if (K0 < 1e-9) K0 = 1e-9;
  const double et = K + mass + Ma;
  const double b2 = K * (K + 2.0 * mass) / (et * et);
  const double rM2 = (mass / Ma) * (mass / Ma);
  tau = std::sqrt(rM2 * (1.0 - b2) + b2);
}

// normalisation, src/Coulomb.jl:215-221
double normalisation(double Z, double A, double mass, double K, double K0) {
  const double p2 = K * (K + 2.0 * mass);
  const double p02 = K0 * (K0 + 2.0 * mass);
  const double d = ALPHA_EM * Z * HBAR_C * (K + mass);
  return A * 1e-3 * p2 * p02 / (d * d * PI * AVOGADRO_NUMBER);
}

struct Screen {
  int n;            // atomic terms + nuclear term
  double amp[4];
  double scr[5];
  double a[4], b[4], c[4];
  double fspin, norm, d2;
};

// screening_parameters, src/Coulomb.jl:228-291
void screening_parameters(double Z, double A, double mass, double K, double K0, Screen& s) {
  const double aZE = ALPHA_EM * Z * (K + mass);
  const double p2 = K * (K + 2.0 * mass);
  const double zeta2 = aZE * aZE / p2;
  double f;
  if (zeta2 < 1.0) {
    f = zeta2 * (1.0 / (1.0 + zeta2) + 0.20205690315959424 - 0.03692775514336999 * zeta2 +
                 0.008349277381922926 * zeta2 * zeta2);
  } else {
    const double r2i = 1.0 / (1.0 + zeta2);
    const double phi = -std::atan(std::sqrt(zeta2));
    f = -0.5 * std::log(r2i) + 0.5 * (1.0 - r2i) + (1.0 / 12.0) * (1.0 - std::cos(2 * phi) * r2i) -
        (1.0 / 120.0) * (1.0 - std::cos(4 * phi) * r2i * r2i) +
        (1.0 / 252.0) * (1.0 - std::cos(6 * phi) * r2i * r2i * r2i);
  }
  const double ktv = std::exp(2.0 * f);
  const double p02 = K0 * (K0 + 2.0 * mass);
  const double d = 0.25 * HBAR_C * HBAR_C / p02 * ktv;
  const int iZ = std::min(std::max((int)Z - 1, 0), 102);
  for (int i = 0; i < 4; ++i) s.amp[i] = 0.0;
  for (int i = 0; i < 5; ++i) s.scr[i] = 0.0;
  int n;
  if (SALVAT_B3[iZ] == 1.0) {
    if (iZ == 0) { n = 1; s.amp[0] = 1.0; }
    else { n = 2; s.amp[0] = SALVAT_A1[iZ]; s.amp[1] = 1.0 - s.amp[0]; }
  } else {
    n = 3;
    s.amp[0] = SALVAT_A1[iZ];
    s.amp[1] = SALVAT_A2[iZ];
    s.amp[2] = 1.0 - (s.amp[0] + s.amp[1]);
  }
  const double* B[3] = {SALVAT_B1, SALVAT_B2, SALVAT_B3};
  for (int i = 0; i < n; ++i) {
    const double a = B[i][iZ] / BOHR_RADIUS;
    s.scr[i] = d * a * a;
  }
  const double RN = nuclear_radius(Z, A);
  s.scr[n] = 12.0 * d / (RN * RN);
  s.n = n + 1;
}

// pole_reduction, src/Coulomb.jl:298-347
void pole_reduction(Screen& s) {
  const int n = s.n - 1;
  for (int i = 0; i < 4; ++i) { s.a[i] = 0.0; s.b[i] = 0.0; s.c[i] = 0.0; }
  for (int i = 0; i < n; ++i) {
    const double Ai = s.amp[i], Bi = s.scr[i];
    s.b[i] = Ai * Ai;
    for (int j = i + 1; j < n; ++j) {
      const double Aj = s.amp[j], Bj = s.scr[j];
      const double d = 2.0 * Ai * Aj / (Bj - Bi);
      s.a[i] += d;
      s.a[j] -= d;
    }
  }
  const double N = s.scr[n];
  double Sa[4] = {0, 0, 0, 0}, Sb[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const double x = 1.0 - s.scr[i] / N;
    double r = x;
    for (int j = 0; j < 4; ++j) {
      Sa[j] += s.a[i] / r;
      r *= x;
      Sb[j] += s.b[i] / r;
    }
  }
  double tmp = 1.0;
  for (int i = 0; i < 4; ++i) {
    s.c[i] = ((4 - i) * Sb[3 - i] / N - Sa[3 - i]) * tmp;
    tmp *= N;
  }
  for (int i = 0; i < n; ++i) {
    const double Bi = s.scr[i];
    const double d = 1.0 - Bi / N;
    const double d4 = d * d * d * d;
    s.a[i] = (s.a[i] - 4.0 * s.b[i] / (N - Bi)) / d4;
    s.b[i] = s.b[i] / d4;
  }
}

// transport_coefficients, src/Coulomb.jl:353-404: (cs0, 2 cs1)
void transport_coefficients(double mu, const Screen& s, double& cs0, double& cs1x2) {
  const double S = s.fspin;
  const int n = s.n - 1;
  double c0 = 0.0, c1 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double alp = s.scr[i];
    const double r = mu / (mu + alp);
    const double L = std::log(1.0 + mu / alp);
    const double I0 = r / alp, J0 = L, I1 = L - r, J1 = -alp * L;
    const double I2 = mu + alp * (r - 2.0 * L), J2 = alp * (alp * L - mu);
    c0 += s.a[i] * (J0 - S * J1) + s.b[i] * (I0 - S * I1);
    c1 += s.a[i] * (J1 - S * J2) + s.b[i] * (I1 - S * I2);
  }
  const double N = s.scr[n];
  const double r = mu / (mu + N);
  const double L = std::log(1.0 + mu / N);
  const double I0 = r / N, J0 = L, I1 = L - r, J1 = -N * L;
  const double I2 = mu + N * (r - 2.0 * L), J2 = N * (N * L - mu);
  const double rn = 1.0 / (1.0 + mu / N);
  const double K0 = (1.0 - rn) * (1.0 + rn) / (2.0 * N * N);
  const double L0 = (1.0 - rn) * ((1.0 + rn) * (1.0 + rn) - rn) / (3.0 * N * N * N);
  const double K1 = I0 - N * K0, L1 = K0 - N * L0;
  const double K2 = I1 - N * K1, L2 = K1 - N * L1;
  c0 += s.c[0] * (J0 - S * J1) + s.c[1] * (I0 - S * I1) + s.c[2] * (K0 - S * K1) + s.c[3] * (L0 - S * L1);
  c1 += s.c[0] * (J1 - S * J2) + s.c[1] * (I1 - S * I2) + s.c[2] * (K1 - S * K2) + s.c[3] * (L1 - S * L2);
  cs0 = c0;
  cs1x2 = 2.0 * c1;
}

// restricted_cs, src/Coulomb.jl:409-418
double restricted_cs(double mu, const Screen& s) {
  if (mu >= 1.0 || mu >= 1e6 * s.scr[s.n - 1]) return 0.0;
  double f0, f1, m0, m1;
  transport_coefficients(1.0, s, f0, f1);
  transport_coefficients(mu, s, m0, m1);
  return m0 > f0 ? 0.0 : f0 - m0;
}

// ridder_find_root, src/Coulomb.jl:430-468
template <class F>
double ridder(F f, double xa, double xb, double fa, double fb, double xtol, double rtol, int maxiter) {
  if (fa * fb > 0) return 0.0;
  if (fa == 0) return xa;
  if (fb == 0) return xb;
  const double tol = xtol + rtol * std::min(std::fabs(xa), std::fabs(xb));
  double xn = 0.0;
  for (int it = 0; it < maxiter; ++it) {
    const double dm = 0.5 * (xb - xa);
    const double xm = xa + dm;
    const double fm = f(xm);
    const double sgn = fb > fa ? 1.0 : -1.0;
    double dn = sgn * dm * fm / std::sqrt(fm * fm - fa * fb);
    const double sgn2 = dn > 0.0 ? 1.0 : -1.0;
    dn = std::fabs(dn);
    double dma = std::fabs(dm) - 0.5 * tol;
    if (dn < dma) dma = dn;
    xn = xm - sgn2 * dma;
    const double fn = f(xn);
    if (fn * fm < 0.0) { xa = xn; fa = fn; xb = xm; fb = fm; }
    else if (fn * fa < 0.0) { xb = xn; fb = fn; }
    else { xa = xn; fa = fn; }
    if (fn == 0.0 || std::fabs(xb - xa) < tol) return xn;
  }
  return xn;
}

// transverse_transport_electronic, src/Coulomb.jl:489-508
double transverse_transport_electronic(double ZoA, double I, double aS, double mass, double K,
                                       double nu) {
  const double p2 = K * (K + 2.0 * mass);
  const double E = K + mass;
  const double Wr = 2.0 * ELECTRON_MASS * p2 / (mass * mass + ELECTRON_MASS * (ELECTRON_MASS + 2.0 * E));
  const double Wmax = std::min(nu, Wr);
  const double beta2 = p2 / (E * E);
  const double J = std::log(aS * Wmax / I) - beta2 * Wmax / Wr + 0.25 * Wmax * Wmax / (E * E) + 1.0 / aS;
  const double gamma = E / mass;
  const double lQ = std::log(1.0 + 2.0 * Wr / ELECTRON_MASS);
  const double Delta = ALPHA_EM / (2 * PI) * (std::log(2.0 * gamma) - lQ / 3.0) * lQ * lQ;
  return 2 * PI * ELECTRON_RADIUS * ELECTRON_RADIUS * ELECTRON_MASS * ELECTRON_MASS *
         AVOGADRO_NUMBER * ZoA / (beta2 * 1e-3 * p2) * (J + Delta);
}

// 20-point Gauss-Legendre on [-1, 1] (src/Coulomb.jl:525-545)
const double GL20_X[20] = {
    -0.9931285991850949, -0.9639719272779138, -0.9122344282513259, -0.8391169718222188,
    -0.7463319064601508, -0.6360536807265150, -0.5108670019508271, -0.3737060887154196,
    -0.2277858511416451, -0.0765265211334973, 0.0765265211334973, 0.2277858511416451,
    0.3737060887154196, 0.5108670019508271, 0.6360536807265150, 0.7463319064601508,
    0.8391169718222188, 0.9122344282513259, 0.9639719272779138, 0.9931285991850949};
const double GL20_W[20] = {
    0.0176140071391521, 0.0406014298003869, 0.0626720483341091, 0.0832767415767048,
    0.1019301198172404, 0.1181945319615184, 0.1316886384491766, 0.1420961093183821,
    0.1491729864726037, 0.1527533871307259, 0.1527533871307259, 0.1491729864726037,
    0.1420961093183821, 0.1316886384491766, 0.1181945319615184, 0.1019301198172404,
    0.0832767415767048, 0.0626720483341091, 0.0406014298003869, 0.0176140071391521};

// dcs_transport_integral, src/Coulomb.jl:554-580
double dcs_transport_integral(int proc, double Z, double A, double mass, double K, double qmin,
                              double qmax) {
  if (qmin >= qmax || qmin <= 0.0) return 0.0;
  const double p2 = K * (K + 2.0 * mass);
  const double lo = std::log(qmin), hi = std::log(qmax);
  const double mid = 0.5 * (lo + hi), half = 0.5 * (hi - lo);
  double r = 0.0;
  for (int i = 0; i < 20; ++i) {
    const double q = std::exp(mid + half * GL20_X[i]);
    const double v = dcs_radiative(proc, Z, A, mass, K, q);
    if (v > 0.0) r += v * (q * q / p2) * q * (half * GL20_W[i]);
  }
  return r;
}

// dcs_*_transport and compute_element_soft_scattering, src/Coulomb.jl:586-652:
// the restricted (mixed-mode) radiative contribution to 1/lambda_1.
double element_soft_scattering_mixed(double Z, double A, double mass, double K, double cutoff) {
  const double sqrte = 1.6487212707;
  const double Z13 = std::pow(Z, 1.0 / 3.0);
  const double qmax_phys = K + mass * (1.0 - 0.75 * sqrte * Z13);
  double s = 0.0;
  {  // bremsstrahlung
    const double qmax = std::min(cutoff * K, qmax_phys);
    const double qmin = 1e-3 * qmax;
    if (!(qmin >= qmax || qmin <= 0.0)) s += dcs_transport_integral(BREMS, Z, A, mass, K, qmin, qmax);
  }
  {  // pair production
    const double qmax = std::min(cutoff * K, qmax_phys);
    const double qmin = std::max(4.0 * ELECTRON_MASS, 1e-3 * qmax);
    if (!(qmin >= qmax || qmin <= 0.0)) s += dcs_transport_integral(PAIR, Z, A, mass, K, qmin, qmax);
  }
  {  // photonuclear
    const double qmax = cutoff * K;
    const double qmin = std::max(0.0, 1e-3 * qmax);
    if (!(qmin >= qmax || qmin <= 0.0)) s += dcs_transport_integral(PHOTO, Z, A, mass, K, qmin, qmax);
  }
  return s;
}

// compute_coulomb_tables!, src/Coulomb.jl:688-826
void coulomb_tables(const std::vector<Elem>& el, double mass, double elastic_ratio, double cutoff,
                    const double* E, const double* range, int n, double ZoA, double I,
                    double density, double* tpath, double* epath, double* screen) {
  const double aS = (I * 1e9) / (28.816 * std::sqrt(ZoA * (density * 1e-3)));
  const int ne = (int)el.size();
  std::vector<Screen> sd(ne);
  for (int ie = 0; ie < n; ++ie) {
    const double K = E[ie];
    if (K <= 0.0) { tpath[ie] = epath[ie] = screen[ie] = 0.0; continue; }
    double cs_m = 0.0, cs1_m = 0.0, A_min = INFINITY, cs_A = 0.0;
    for (int k = 0; k < ne; ++k) {
      double K0, gCM, tau;
      frame_parameters(el[k].Z, el[k].A, mass, K, K0, gCM, tau);
      Screen& s = sd[k];
      s.fspin = K * ((K + mass) + mass) / ((K + mass) * (K + mass));
      screening_parameters(el[k].Z, el[k].A, mass, K, K0, s);
      pole_reduction(s);
      double G0, G1;
      transport_coefficients(1.0, s, G0, G1);
      s.norm = el[k].f / normalisation(el[k].Z, el[k].A, mass, K, K0);
      cs_m += s.norm * G0;
      const double d = 1.0 / (gCM * (1.0 + tau));
      s.d2 = d * d;
      cs1_m += s.norm * G1 * s.d2;
      for (int j = 0; j < s.n - 1; ++j) A_min = std::min(A_min, s.scr[j]);
      cs_A += s.norm;
    }
    cs_A /= A_min * (A_min + 1.0);
    const double range_i = range[ie] > 0.0 ? 1.0 / range[ie] : 0.0;
    double cs_h = std::max(cs1_m, range_i) / elastic_ratio;
    if (cs_h < 1.0 / EHS_PATH_MAX) cs_h = 1.0 / EHS_PATH_MAX;
    const double max_mu0 = 0.5 * (1.0 - std::cos(MAX_SOFT_ANGLE * PI / 180.0));
    double mu0, lb_h;
    if (cs_h < cs_m) {
      const double zeta = cs_h / cs_A;
      double mu_max = A_min * (1.0 - zeta) / (A_min + zeta);
      if (mu_max > max_mu0) mu_max = max_mu0;
      auto obj = [&](double mu) {
        double t = 0.0;
        for (int k = 0; k < ne; ++k) t += sd[k].norm * restricted_cs(mu, sd[k]);
        return t - cs_h;
      };
      const double fmax = obj(mu_max);
      const double fmin = cs_m - cs_h;
      mu0 = ridder(obj, 0.0, mu_max, fmin, fmax, 1e-6 * mu_max, 1e-6, 100);
      if (mu0 > max_mu0) mu0 = max_mu0;
      cs_h += obj(mu0);
      lb_h = cs_h <= 1.0 / EHS_PATH_MAX ? EHS_PATH_MAX : 1.0 / cs_h;
    } else {
      mu0 = 0.0;
      lb_h = 1.0 / cs_m;
    }
    screen[ie] = mu0;
    const double p2 = K * (K + 2.0 * mass);
    epath[ie] = lb_h * p2;
    double inv = 0.0;
    for (int k = 0; k < ne; ++k) {
      double G0, G1;
      transport_coefficients(mu0, sd[k], G0, G1);
      inv += sd[k].norm * sd[k].d2 * G1;
    }
    double inv_rad = 0.0;
    for (int k = 0; k < ne; ++k)
      inv_rad += element_soft_scattering_mixed(el[k].Z, el[k].A, mass, K, cutoff) * el[k].f;
    const double inv_el = transverse_transport_electronic(ZoA, I, aS, mass, K, cutoff * K);
    const double tot = inv + inv_rad + inv_el;
    tpath[ie] = tot > 0.0 ? 1.0 / tot : 0.0;
  }
}

// Proper time, spin factor, elastic cutoff, magnetic rotation of a table
// whose energies, stopping powers, ranges and screening are filled
// (src/Physics.jl:158-181, :256-263).
void kinematic_columns(double* t, int K, double mass, bool proper_time) {
  double* E = t + (size_t)DP_T_ENERGY * K;
  double* S = t + (size_t)DP_T_CSDA_STOP * K;
  double* tm = t + (size_t)DP_T_CSDA_TIME * K;
  if (proper_time) {
    tm[0] = 0.0;
    for (int i = 1; i < K; ++i) {
      const double Ki = E[i], Kp = E[i - 1];
      const double g = 1.0 + Ki / mass, gp = 1.0 + Kp / mass;
      const double b = std::sqrt(1.0 - 1.0 / (g * g)), bp = std::sqrt(1.0 - 1.0 / (gp * gp));
      const double gm = (g + gp) / 2, bm = (b + bp) / 2;
      const double avg = (S[i] + S[i - 1]) / 2;
      tm[i] = tm[i - 1] + (Ki - Kp) / (avg * bm * gm);
    }
  }
  for (int i = 0; i < K; ++i) {
    const double Ki = E[i];
    const double g = 1.0 + Ki / mass;
    const double b = std::sqrt(1.0 - 1.0 / (g * g));
    const double p = std::sqrt(Ki * (Ki + 2 * mass));
    const double e = Ki + mass;
    t[(size_t)DP_T_SPIN_FACTOR * K + i] = Ki * (e + mass) / (e * e);
    const double mu0 = t[(size_t)DP_T_SCREENING * K + i];
    t[(size_t)DP_T_ELASTIC_CUTOFF * K + i] =
        mu0 < 1e-8 ? 2.0 * std::sqrt(std::max(mu0, 0.0)) : std::acos(std::max(1.0 - 2.0 * mu0, -1.0));
    t[(size_t)DP_T_MAGNETIC_ROTATION * K + i] = LARMOR_FACTOR / (b * p);
    t[(size_t)DP_T_LOG_ENERGY * K + i] = std::log(Ki);
  }
}

void integrate_range(const double* E, const double* S, double* R, int K) {
  R[0] = E[0] / S[0];
  for (int i = 1; i < K; ++i) R[i] = R[i - 1] + (E[i] - E[i - 1]) / ((S[i] + S[i - 1]) / 2);
}

std::vector<Elem> elements_of(const dp_material_recipe* r) {
  std::vector<Elem> el;
  double tot = 0.0;
  for (int i = 0; i < r->n_elements; ++i) tot += r->fraction[i];
  for (int i = 0; i < r->n_elements; ++i)
    el.push_back({r->Z[i], r->A[i], r->I[i], r->fraction[i] / tot});
  return el;
}

int check_recipe(const dp_material_recipe* r) {
  if (!r || r->n_elements <= 0 || !r->Z || !r->A || !r->I || !r->fraction)
    return dph::fail(DP_E_ARGUMENT, "material recipe needs >= 1 element with Z, A, I, fraction");
  if (!(r->density > 0) || !(r->I_material > 0) || !(r->ZoA > 0))
    return dph::fail(DP_E_ARGUMENT, "material recipe needs positive density, I and Z/A");
  return DP_OK;
}

// ------------------------------------------------- electronic stopping
// Atomic shells of a material (atomic_shell_create, src/Materials.jl:378-436):
// every element's Geant4 shells with occupancy weighted by its mass fraction
// over A (the fictitious Z = 11, A = 22 standard-rock element takes the
// shells of C, 3 O and Ca scaled by 11/50), then oscillator strengths
// normalised to 1 and binding energies scaled so that the strength-weighted
// mean log energy equals log(I) in units of the plasma energy.
struct Shell {
  double f, E;
};
int shell_count(int Z) {
  const int z = std::min(std::max(Z, 1), dpshell::SHELL_N_Z);
  return dpshell::SHELL_INDEX[z] - dpshell::SHELL_INDEX[z - 1];
}
void add_shells(std::vector<Shell>& out, int Z, double w) {
  const int z = std::min(std::max(Z, 1), dpshell::SHELL_N_Z);
  for (int i = dpshell::SHELL_INDEX[z - 1]; i < dpshell::SHELL_INDEX[z]; ++i)
    out.push_back({w * (double)dpshell::SHELL_OCCUPANCY[i], (double)dpshell::SHELL_ENERGY_EV[i]});
}
std::vector<Shell> atomic_shells(const std::vector<Elem>& el, double ZoA, double I,
                                 double density) {
  std::vector<Shell> sh;
  for (const Elem& e : el) {
    const double w = e.f / e.A;
    if (e.Z == 11.0 && e.A == 22.0) {
      const double ws = w * 11.0 / (6.0 + 3 * 8.0 + 20.0);
      add_shells(sh, 6, ws);
      add_shells(sh, 8, 3 * ws);
      add_shells(sh, 20, ws);
    } else {
      add_shells(sh, (int)std::lround(e.Z), w);
    }
  }
  double ftot = 0.0, lnI = 0.0;
  for (const Shell& s : sh) {
    ftot += s.f;
    lnI += s.f * std::log(s.E);
  }
  const double inv_ftot = 1.0 / ftot;
  lnI *= inv_ftot;
  const double wp = 28.816 * std::sqrt(ZoA * density * 1e-3);
  const double aS = I * 1e9 / std::exp(lnI);
  const double r = aS / wp;
  for (Shell& s : sh) {
    s.f *= inv_ftot;
    s.E *= r;
  }
  return sh;
}

// Fano density effect delta(gamma) by the bisection of PUMAS
// (electronic_density_effect, src/Materials.jl:442-475): the root l2 of
// sum f/(E^2 + l2) = 1/gamma^2, then delta = sum f log(1 + l2/E^2) - l2/gamma^2.
double density_effect(const std::vector<Shell>& sh, double gamma) {
  const double y = 1.0 / (gamma * gamma);
  double ymax = 0.0;
  for (const Shell& s : sh) ymax += s.f / (s.E * s.E);
  if (ymax <= y) return 0.0;
  const double eps = 2.220446049250313e-16;
  double l2min = 0.0, l2max = 1.0 / y, l2 = 0.0;
  for (int it = 0; it < 200; ++it) {
    l2 = 0.5 * (l2min + l2max);
    double yi = 0.0;
    for (const Shell& s : sh) yi += s.f / (s.E * s.E + l2);
    if (std::fabs(yi - y) <= eps || (l2max - l2min) <= eps) break;
    if (yi > y) l2min = l2;
    else l2max = l2;
  }
  double delta = -l2 * y;
  for (const Shell& s : sh) delta += s.f * std::log(1.0 + l2 / (s.E * s.E));
  return delta;
}

// Electronic stopping power, GeV m2/kg: Bethe with the maximum energy
// transfer Qmax, the radiative correction Delta and the Fano density effect
// (_electronic_energy_loss, src/Materials.jl:500-520).
double electronic_loss(double ZoA, double I, const std::vector<Shell>& sh, double mass,
                       double K) {
  const double E = K + mass;
  const double P2 = K * (K + 2.0 * mass);
  const double beta2 = P2 / (E * E);
  const double gamma = E / mass;
  const double r = ELECTRON_MASS / mass;
  const double Qmax = 2.0 * r * P2 / (mass * (1.0 + r * r) + 2.0 * r * E);
  const double lQ = std::log(1.0 + 2.0 * Qmax / ELECTRON_MASS);
  const double Delta = ALPHA_EM / (2 * PI) * (std::log(2.0 * gamma) - lQ / 3.0) * lQ * lQ;
  const double delta = density_effect(sh, gamma);
  return 2 * PI * ELECTRON_RADIUS * ELECTRON_RADIUS * ELECTRON_MASS * AVOGADRO_NUMBER * ZoA /
         (beta2 * 1e-3) *
         (std::log(2.0 * ELECTRON_MASS * beta2 * gamma * gamma * Qmax / (I * I)) - 2.0 * beta2 -
          delta + 0.25 * Qmax * Qmax / (E * E) + Delta);
}

}  // namespace

extern "C" {

double dp_electronic_stopping_power(const dp_material_recipe* r, double mass, double K) {
  if (check_recipe(r)) return NAN;
  const std::vector<Elem> el = elements_of(r);
  return electronic_loss(r->ZoA, r->I_material,
                         atomic_shells(el, r->ZoA, r->I_material, r->density), mass, K);
}

double dp_density_effect(const dp_material_recipe* r, double gamma) {
  if (check_recipe(r)) return NAN;
  const std::vector<Elem> el = elements_of(r);
  return density_effect(atomic_shells(el, r->ZoA, r->I_material, r->density), gamma);
}

int dp_build_material_table_empirical(const dp_material_recipe* r, const double* energies,
                                      int32_t n, const dp_physics_settings* st, double* out) {
  int rc = check_recipe(r);
  if (rc) return rc;
  if (!energies || n <= 2 || !st || !out)
    return dph::fail(DP_E_ARGUMENT, "dp_build_material_table_empirical: bad arguments");
  for (int i = 0; i < n; ++i)
    if (!(energies[i] > 0) || (i > 0 && !(energies[i] > energies[i - 1])))
      return dph::fail(DP_E_ARGUMENT,
                       "dp_build_material_table_empirical: energies must be positive and increasing");
  const int K = n;
  const double mass = st->mass, cutoff = st->cutoff;
  for (size_t i = 0; i < (size_t)DP_T_NARRAYS * K; ++i) out[i] = 0.0;
  auto col = [&](int a) { return out + (size_t)a * K; };
  const std::vector<Elem> el = elements_of(r);
  const std::vector<Shell> sh = atomic_shells(el, r->ZoA, r->I_material, r->density);
  const int npts = 180;
  const double x_low = 1e-6;
  for (int i = 0; i < K; ++i) {
    const double Ki = energies[i];
    col(DP_T_ENERGY)[i] = Ki;
    double brem = 0.0, pair = 0.0, photo = 0.0, hard = 0.0, cs = 0.0, strag = 0.0;
    if (Ki >= 1e-6) {
      for (const Elem& e : el) {
        brem += e.f * dcs_integral(BREMS, e.Z, e.A, e.I, mass, Ki, x_low, 1.0, 1, npts);
        pair += e.f * dcs_integral(PAIR, e.Z, e.A, e.I, mass, Ki, x_low, 1.0, 1, npts);
        photo += e.f * dcs_integral(PHOTO, e.Z, e.A, e.I, mass, Ki, x_low, 1.0, 1, npts);
        for (int p : {BREMS, PAIR, PHOTO, IONISATION}) {
          hard += e.f * dcs_integral(p, e.Z, e.A, e.I, mass, Ki, cutoff, 1.0, 1, npts);
          cs += e.f * dcs_integral(p, e.Z, e.A, e.I, mass, Ki, cutoff, 1.0, 0, npts);
        }
        strag += e.f * dcs_integral(IONISATION, e.Z, e.A, e.I, mass, Ki, 0.0, cutoff, 2, npts);
      }
    }
    const double elec = electronic_loss(r->ZoA, r->I_material, sh, mass, Ki);
    col(DP_T_IONIZATION)[i] = elec;
    col(DP_T_BREMS)[i] = brem;
    col(DP_T_PAIR)[i] = pair;
    col(DP_T_PHOTO)[i] = photo;
    col(DP_T_CSDA_STOP)[i] = elec + brem + pair + photo;
    col(DP_T_MIXED_STOP)[i] = std::max(col(DP_T_CSDA_STOP)[i] - hard, 1e-20);
    col(DP_T_STRAGGLING)[i] = strag;
    col(DP_T_CROSS_SECTION)[i] = cs;
  }
  integrate_range(col(DP_T_ENERGY), col(DP_T_CSDA_STOP), col(DP_T_CSDA_RANGE), K);
  integrate_range(col(DP_T_ENERGY), col(DP_T_MIXED_STOP), col(DP_T_MIXED_RANGE), K);
  coulomb_tables(el, mass, st->elastic_ratio, cutoff, col(DP_T_ENERGY), col(DP_T_CSDA_RANGE), K,
                 r->ZoA, r->I_material, r->density, col(DP_T_TRANSPORT_PATH),
                 col(DP_T_ELASTIC_PATH), col(DP_T_SCREENING));
  kinematic_columns(out, K, mass, true);
  return DP_OK;
}


double dp_dcs(int32_t process, double Z, double A, double I, double mass, double K, double q) {
  if (process == IONISATION) return dcs_ionisation(Z, I, mass, K, q);
  return dcs_radiative(process, Z, A, mass, K, q);
}

double dp_dcs_integral(int32_t process, double Z, double A, double I, double mass, double K,
                       double xlow, double xhigh, int32_t mode, int32_t npoints) {
  return dcs_integral(process, Z, A, I, mass, K, xlow, xhigh, mode, npoints);
}

int dp_build_material_table(const dp_material_recipe* r, const dp_dedx_columns* d,
                            const dp_physics_settings* st, double* out) {
  int rc = check_recipe(r);
  if (rc) return rc;
  if (!d || !st || !out || d->n <= 2 || !d->kinetic_energy || !d->stopping_power ||
      !d->csda_range)
    return dph::fail(DP_E_ARGUMENT, "dp_build_material_table: null or short dE/dx columns");
  const int K = d->n;
  const double mass = st->mass, cutoff = st->cutoff;
  for (size_t i = 0; i < (size_t)DP_T_NARRAYS * K; ++i) out[i] = 0.0;
  auto col = [&](int a) { return out + (size_t)a * K; };
  for (int i = 0; i < K; ++i) {
    col(DP_T_ENERGY)[i] = d->kinetic_energy[i];
    col(DP_T_CSDA_STOP)[i] = d->stopping_power[i];
    col(DP_T_CSDA_RANGE)[i] = d->csda_range[i];
    if (d->ionization) col(DP_T_IONIZATION)[i] = d->ionization[i];
    if (d->bremsstrahlung) col(DP_T_BREMS)[i] = d->bremsstrahlung[i];
    if (d->pair_production) col(DP_T_PAIR)[i] = d->pair_production[i];
    if (d->photonuclear) col(DP_T_PHOTO)[i] = d->photonuclear[i];
  }
  const std::vector<Elem> el = elements_of(r);
  const int npts = 180;
  // mixed stopping power, straggling and hard cross-section per knot
  // (src/Physics.jl:183-237); knots below 1 MeV keep zero radiative terms.
  for (int i = 0; i < K; ++i) {
    const double Ki = col(DP_T_ENERGY)[i];
    double hard = 0.0, strag = 0.0, cs = 0.0;
    if (Ki >= 1e-3) {
      for (const Elem& e : el) {
        for (int p : {BREMS, PAIR, PHOTO}) hard += e.f * dcs_integral(p, e.Z, e.A, e.I, mass, Ki, cutoff, 1.0, 1, npts);
        hard += e.f * dcs_integral(IONISATION, e.Z, e.A, e.I, mass, Ki, cutoff, 1.0, 1, npts);
        strag += e.f * dcs_integral(IONISATION, e.Z, e.A, e.I, mass, Ki, 0.0, cutoff, 2, npts);
        for (int p : {BREMS, PAIR, PHOTO}) cs += e.f * dcs_integral(p, e.Z, e.A, e.I, mass, Ki, cutoff, 1.0, 0, npts);
        cs += e.f * dcs_integral(IONISATION, e.Z, e.A, e.I, mass, Ki, cutoff, 1.0, 0, npts);
      }
    }
    col(DP_T_MIXED_STOP)[i] = std::max(col(DP_T_CSDA_STOP)[i] - hard, 1e-20);
    col(DP_T_STRAGGLING)[i] = strag;
    col(DP_T_CROSS_SECTION)[i] = cs;
  }
  integrate_range(col(DP_T_ENERGY), col(DP_T_MIXED_STOP), col(DP_T_MIXED_RANGE), K);
  // the Coulomb tables use the Z/A of the dE/dx header (src/Physics.jl:244-248)
  coulomb_tables(el, mass, st->elastic_ratio, cutoff, col(DP_T_ENERGY), col(DP_T_CSDA_RANGE), K,
                 d->ZoA > 0 ? d->ZoA : r->ZoA, r->I_material, r->density, col(DP_T_TRANSPORT_PATH),
                 col(DP_T_ELASTIC_PATH), col(DP_T_SCREENING));
  kinematic_columns(out, K, mass, true);
  return DP_OK;
}

int dp_build_composite_table(int32_t n_components, const double* const* tables,
                             const double* fractions, int32_t n_knots,
                             const dp_material_recipe* flat, const dp_physics_settings* st,
                             double* out) {
  int rc = check_recipe(flat);
  if (rc) return rc;
  if (n_components <= 0 || !tables || !fractions || n_knots <= 2 || !st || !out)
    return dph::fail(DP_E_ARGUMENT, "dp_build_composite_table: bad arguments");
  const int K = n_knots;
  const double mass = st->mass;
  for (size_t i = 0; i < (size_t)DP_T_NARRAYS * K; ++i) out[i] = 0.0;
  auto col = [&](int a) { return out + (size_t)a * K; };
  for (int i = 0; i < K; ++i) {
    col(DP_T_ENERGY)[i] = tables[0][(size_t)DP_T_ENERGY * K + i];
  }
  const int summed[] = {DP_T_CSDA_STOP, DP_T_MIXED_STOP, DP_T_STRAGGLING, DP_T_CROSS_SECTION,
                        DP_T_IONIZATION, DP_T_BREMS, DP_T_PAIR, DP_T_PHOTO};
  for (int c = 0; c < n_components; ++c)
    for (int a : summed)
      for (int i = 0; i < K; ++i) col(a)[i] += fractions[c] * tables[c][(size_t)a * K + i];
  integrate_range(col(DP_T_ENERGY), col(DP_T_CSDA_STOP), col(DP_T_CSDA_RANGE), K);
  integrate_range(col(DP_T_ENERGY), col(DP_T_MIXED_STOP), col(DP_T_MIXED_RANGE), K);
  const std::vector<Elem> el = elements_of(flat);
  coulomb_tables(el, mass, st->elastic_ratio, st->cutoff, col(DP_T_ENERGY), col(DP_T_CSDA_RANGE),
                 K, flat->ZoA, flat->I_material, flat->density, col(DP_T_TRANSPORT_PATH),
                 col(DP_T_ELASTIC_PATH), col(DP_T_SCREENING));
  kinematic_columns(out, K, mass, true);
  return DP_OK;
}

}  // extern "C"

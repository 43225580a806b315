// CPU oracle for the DiffPumas backward-transport hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a scalar restatement of the reference
// algorithms, used by tests/ and by bench.py's cpu_baseline leg as the checker.
// The shipped library never links, loads or calls anything in this folder.
//
// Every routine is templated on the arithmetic type R (double, or the Dual
// forward-mode number below) so one source yields both primal values and
// exact derivatives of the same discretised program, the way the reference
// differentiates its Julia code with ForwardDiff/Zygote.
//
// Citations are into the reference tree (grinisrit/diffpumas.jl).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <limits>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- constants
// src/Constants.jl:23-75
constexpr double MUON_MASS = 0.10565839;
constexpr double MUON_C_TAU = 658.654;
constexpr double STEP_MIN = 1e-7;
constexpr double EPS64 = 2.220446049250313e-16;  // Julia eps(Float64)
// src/Geometry.jl:36
constexpr double PRIMARY_ALTITUDE = 3e4;

// Table array order of MaterialTable (src/Physics.jl:64-102).
enum TabArr {
  T_ENERGY = 0, T_LOG_ENERGY, T_CSDA_STOP, T_CSDA_RANGE, T_CSDA_TIME,
  T_MIXED_STOP, T_MIXED_RANGE, T_STRAGGLING, T_IONIZATION, T_BREMS, T_PAIR,
  T_PHOTO, T_CROSS_SECTION, T_TRANSPORT_PATH, T_ELASTIC_PATH, T_ELASTIC_CUTOFF,
  T_MAGNETIC_ROTATION, T_SCREENING, T_SPIN_FACTOR, T_NARR
};

// ------------------------------------------------------------ dual numbers
struct Dual {
  double v, d;
  Dual() : v(0), d(0) {}
  Dual(double x) : v(x), d(0) {}
  Dual(double x, double dx) : v(x), d(dx) {}
};
inline double val(double x) { return x; }
inline double val(const Dual& x) { return x.v; }
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
inline Dual operator/(double a, Dual b) {
  double q = a / b.v;
  return {q, -q * b.d / b.v};
}
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
inline Dual log(Dual a) { return {std::log(a.v), a.d / a.v}; }
inline Dual exp(Dual a) { double e = std::exp(a.v); return {e, e * a.d}; }
inline Dual sqrt(Dual a) { double s = std::sqrt(a.v); return {s, a.d / (2.0 * s)}; }
inline Dual pow(Dual a, double p) {
  return {std::pow(a.v, p), p * std::pow(a.v, p - 1.0) * a.d};
}
inline Dual abs(Dual a) { return a.v < 0 ? -a : a; }
// Forward mode with two directions (used for segment Jacobians in (w, E)).
struct Dual2 {
  double v, a, b;
  Dual2() : v(0), a(0), b(0) {}
  Dual2(double x) : v(x), a(0), b(0) {}
  Dual2(double x, double da, double db) : v(x), a(da), b(db) {}
};
inline double val(const Dual2& x) { return x.v; }
inline Dual2 operator+(Dual2 x, Dual2 y) { return {x.v + y.v, x.a + y.a, x.b + y.b}; }
inline Dual2 operator-(Dual2 x, Dual2 y) { return {x.v - y.v, x.a - y.a, x.b - y.b}; }
inline Dual2 operator-(Dual2 x) { return {-x.v, -x.a, -x.b}; }
inline Dual2 operator*(Dual2 x, Dual2 y) { return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b}; }
inline Dual2 operator/(Dual2 x, Dual2 y) {
  double q = x.v / y.v;
  return {q, (x.a - q * y.a) / y.v, (x.b - q * y.b) / y.v};
}
inline Dual2 operator+(Dual2 x, double y) { return {x.v + y, x.a, x.b}; }
inline Dual2 operator+(double x, Dual2 y) { return {x + y.v, y.a, y.b}; }
inline Dual2 operator-(Dual2 x, double y) { return {x.v - y, x.a, x.b}; }
inline Dual2 operator-(double x, Dual2 y) { return {x - y.v, -y.a, -y.b}; }
inline Dual2 operator*(Dual2 x, double y) { return {x.v * y, x.a * y, x.b * y}; }
inline Dual2 operator*(double x, Dual2 y) { return {x * y.v, x * y.a, x * y.b}; }
inline Dual2 operator/(Dual2 x, double y) { return {x.v / y, x.a / y, x.b / y}; }
inline Dual2 operator/(double x, Dual2 y) {
  double q = x / y.v;
  return {q, -q * y.a / y.v, -q * y.b / y.v};
}
inline Dual2 log(Dual2 x) { return {std::log(x.v), x.a / x.v, x.b / x.v}; }
inline Dual2 exp(Dual2 x) { double e = std::exp(x.v); return {e, e * x.a, e * x.b}; }
inline Dual2 sqrt(Dual2 x) { double s = std::sqrt(x.v); return {s, x.a / (2.0 * s), x.b / (2.0 * s)}; }
inline Dual2 pow(Dual2 x, double p) {
  double d = p * std::pow(x.v, p - 1.0);
  return {std::pow(x.v, p), d * x.a, d * x.b};
}
inline bool isfinite_(const Dual2& x) { return std::isfinite(x.v); }
inline bool isfinite_(double x) { return std::isfinite(x); }
inline bool isfinite_(const Dual& x) { return std::isfinite(x.v) && std::isfinite(x.d); }
using std::log; using std::exp; using std::sqrt; using std::pow; using std::abs;

// Julia max/min on reals compare primal values only.
template <class A, class B> inline A vmax(A a, B b) { return val(a) < val(b) ? A(b) : a; }
template <class A, class B> inline A vmin(A a, B b) { return val(b) < val(a) ? A(b) : a; }

// ------------------------------------------------------------------ tables
struct Tables {
  int K = 0;            // knots per material
  int M = 0;            // materials
  const double* data = nullptr;     // M * T_NARR * K
  const double* scalars = nullptr;  // M * 4 : density, I, ZoA, reserved
  double mass = MUON_MASS;
  double ctau = MUON_C_TAU;
  const double* arr(int m, int a) const { return data + ((size_t)m * T_NARR + a) * (size_t)K; }
  double zoa(int m) const { return scalars[m * 4 + 2]; }
};

// interpolate_table_fast, src/Physics.jl:660-708 (the Dual overload :691-708
// compares primal values; outside the table the result is constant).
template <class R>
inline R interp(const double* xs, const double* lxs, const double* ys, int n, R x, R lx) {
  if (val(x) <= xs[0]) return R(ys[0]);
  if (val(x) >= xs[n - 1]) return R(ys[n - 1]);
  int lo = 0, hi = n - 1;
  while (lo < hi - 1) {
    int mid = (lo + hi) >> 1;
    if (val(lx) < lxs[mid]) hi = mid; else lo = mid;
  }
  R t = (lx - lxs[lo]) / (lxs[lo + 1] - lxs[lo]);
  return ys[lo] + t * (ys[lo + 1] - ys[lo]);
}

template <class R>
inline R property(const Tables& T, int m, int a, R energy) {
  return interp(T.arr(m, T_ENERGY), T.arr(m, T_LOG_ENERGY), T.arr(m, a), T.K, energy, log(energy));
}
// property_range / _stopping_power, src/Physics.jl:783-810 (mixed=true
// selects the restricted MIXED tables).
template <class R> inline R prop_range(const Tables& T, bool mixed, int m, R e) {
  return property(T, m, mixed ? T_MIXED_RANGE : T_CSDA_RANGE, e);
}
template <class R> inline R prop_stop(const Tables& T, bool mixed, int m, R e) {
  return property(T, m, mixed ? T_MIXED_STOP : T_CSDA_STOP, e);
}
// Inverse range on a (energies, log_energies, range) triple,
// src/Physics.jl:833-867 and the mixture form :1291-1323.
template <class R>
inline R inverse_range(const double* es, const double* les, const double* rt, int n, R range) {
  if (val(range) <= rt[0]) return R(es[0]);
  if (val(range) >= rt[n - 1]) return R(es[n - 1]);
  int lo = 0, hi = n - 1;
  while (lo < hi - 1) {
    int mid = (lo + hi) >> 1;
    if (val(range) < rt[mid]) hi = mid; else lo = mid;
  }
  R t = (range - rt[lo]) / (rt[lo + 1] - rt[lo]);
  R lE = les[lo] + t * (les[lo + 1] - les[lo]);
  return exp(lE);
}
template <class R> inline R prop_kinetic(const Tables& T, bool mixed, int m, R range) {
  return inverse_range(T.arr(m, T_ENERGY), T.arr(m, T_LOG_ENERGY),
                       T.arr(m, mixed ? T_MIXED_RANGE : T_CSDA_RANGE), T.K, range);
}

// ------------------------------------------------------------- flux model
// src/GaisserFlux.jl:36-46
inline double charge_fraction(double q) {
  const double r = 1.2766;
  if (q < 0) return 1.0 / (1.0 + r);
  if (q > 0) return r / (1.0 + r);
  return 1.0;
}
// src/GaisserFlux.jl:62-76 (cos_theta is geometry, never differentiated)
inline double cos_theta_star(double c) {
  const double p0 = 0.102573, p1 = -0.068287, p2 = 0.958633, p3 = 0.0407253, p4 = 0.817285;
  double cs2 = (c * c + p0 * p0 + p1 * std::pow(c, p2) + p3 * std::pow(c, p4)) /
               (1.0 + p0 * p0 + p1 + p3);
  return cs2 > 0 ? std::sqrt(cs2) : 0.0;
}
// src/GaisserFlux.jl:93-105
template <class R> inline R flux_gaisser(double c, R K, double q) {
  R Emu = K + MUON_MASS;
  R ec = 1.1 * Emu * c;
  R rpi = 1.0 + ec / 115.0;
  R rK = 1.0 + ec / 850.0;
  return 1.4e3 * pow(Emu, -2.7) * (1.0 / rpi + 0.054 / rK) * charge_fraction(q);
}
// src/GaisserFlux.jl:125-133
template <class R> inline R flux_gccly(double c, R K, double q) {
  R Emu = K + MUON_MASS;
  double cs = cos_theta_star(c);
  R f = pow(1.0 + 3.64 / (Emu * std::pow(cs, 1.29)), -2.7);
  return f * flux_gaisser(cs, K, q);
}

// ------------------------------------------------------ air / decay helpers
// src/Geometry.jl:1043-1055
inline double air_grammage(double z0, double z1, double c) {
  const double rho0 = 1.205, h = 12e3;
  double cs = std::max(std::fabs(c), 0.01);
  double g = (rho0 * h / cs) * (std::exp(-z0 / h) - std::exp(-z1 / h));
  return std::max(g, 0.0);
}
// src/Geometry.jl:95-106
inline double air_density(double z) { return 1.205 * std::exp(-z / 12e3); }
// src/Geometry.jl:1083-1107 (Dual overload keeps the primal floor branch)
template <class R, class P>
inline auto decay_weight(const Tables& T, P path, R Eavg) {
  auto g = 1.0 + Eavg / T.mass;
  auto b2 = 1.0 - 1.0 / (g * g);
  auto beta = sqrt(val(b2) >= 1e-12 ? b2 : decltype(b2)(1e-12));
  auto L = beta * g * T.ctau;
  return exp(-path / L);
}

// -------------------------------------------------------- energy sampling
// src/Pumas.jl:195-207 and the inline copy src/Geometry.jl:856-875
inline void sample_loguniform(double emin, double emax, double u, double& k, double& w) {
  double rk = std::log(emax / emin);
  if (rk > 0) { k = emin * std::exp(rk * u); w = k * rk; }
  else { k = emin; w = 1.0; }
}

}  // namespace orc

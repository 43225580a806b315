// Device-side building blocks shared by the transport kernels (gfx950).
//
// Table lookups follow the reference interpolation conventions exactly
// (src/Physics.jl:660-708, :833-867): clamp on the energy array, search the
// log-energy (or range) grid, linear interpolation in log energy.  The search
// returns the same bracket as the reference bisection for every input, and
// each lookup also returns its slope so kernels can form exact derivatives of
// the discretised program (zero outside the table, right interval at knots).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpk {

constexpr double EPS64 = 2.220446049250313e-16;  // Julia eps(Float64)
constexpr double MUON_MASS = 0.10565839;         // src/Constants.jl:47

// ------------------------------------------------------------ bracketing
// Largest lo in [0, n-2] with a[lo] <= x, or 0.  For a strictly increasing
// grid this is the interval the reference bisection (lo, hi = 1, n) selects,
// including the right-interval choice at knots.  pow2 = largest power of two
// <= n-2 (precomputed per table).
__device__ __forceinline__ int bracket(const double* __restrict__ a, int n, int pow2, double x) {
  int lo = 0;
#pragma unroll 1
  for (int step = pow2; step > 0; step >>= 1) {
    int c = lo + step;
    if (c <= n - 2 && a[c] <= x) lo = c;
  }
  return lo;
}

// Largest i in [0, n-2] with xs[i] <= x (0 if none): checks the guessed
// interval and its neighbours, then bisects.
__device__ __forceinline__ int locate_idx(const double* __restrict__ xs, int n, int pow2, double x,
                                          int g) {
  g = min(max(g, 0), n - 2);
  if (xs[g] <= x) {
    if (g == n - 2 || x < xs[g + 1]) return g;
    if (g + 1 == n - 2 || x < xs[g + 2]) return g + 1;
  } else if (g > 0 && xs[g - 1] <= x) {
    return g - 1;
  }
  return bracket(xs, n, pow2, x);
}

// The same two searches over FP32 grids (the FP32 stepping kernels).
__device__ __forceinline__ int bracket(const float* __restrict__ a, int n, int pow2, float x) {
  int lo = 0;
#pragma unroll 1
  for (int step = pow2; step > 0; step >>= 1) {
    int c = lo + step;
    if (c <= n - 2 && a[c] <= x) lo = c;
  }
  return lo;
}
__device__ __forceinline__ int locate_idx(const float* __restrict__ xs, int n, int pow2, float x,
                                          int g) {
  g = min(max(g, 0), n - 2);
  if (xs[g] <= x) {
    if (g == n - 2 || x < xs[g + 1]) return g;
    if (g + 1 == n - 2 || x < xs[g + 2]) return g + 1;
  } else if (g > 0 && xs[g - 1] <= x) {
    return g - 1;
  }
  return bracket(xs, n, pow2, x);
}

// Interpolated property with slope dy/dE (0 when clamped).
struct Lookup {
  double y, dy;
};
__device__ __forceinline__ Lookup interp_e(const double* __restrict__ es,
                                           const double* __restrict__ les,
                                           const double* __restrict__ ys, int n, int pow2,
                                           double e, double le) {
  Lookup r;
  if (e <= es[0]) { r.y = ys[0]; r.dy = 0.0; return r; }
  if (e >= es[n - 1]) { r.y = ys[n - 1]; r.dy = 0.0; return r; }
  int i = bracket(les, n, pow2, le);
  double dl = les[i + 1] - les[i];
  double t = (le - les[i]) / dl;
  double dyy = ys[i + 1] - ys[i];
  r.y = ys[i] + t * dyy;
  r.dy = (dyy / dl) / e;
  return r;
}
// Two properties on the same bracket (same energy, same search).
struct Lookup2 {
  double y0, dy0, y1, dy1;
};
__device__ __forceinline__ Lookup2 interp_e2(const double* __restrict__ es,
                                             const double* __restrict__ les,
                                             const double* __restrict__ y0,
                                             const double* __restrict__ y1, int n, int pow2,
                                             double e, double le) {
  Lookup2 r;
  if (e <= es[0]) { r.y0 = y0[0]; r.y1 = y1[0]; r.dy0 = r.dy1 = 0.0; return r; }
  if (e >= es[n - 1]) { r.y0 = y0[n - 1]; r.y1 = y1[n - 1]; r.dy0 = r.dy1 = 0.0; return r; }
  int i = bracket(les, n, pow2, le);
  double dl = les[i + 1] - les[i];
  double t = (le - les[i]) / dl;
  double d0 = y0[i + 1] - y0[i], d1 = y1[i + 1] - y1[i];
  r.y0 = y0[i] + t * d0;
  r.y1 = y1[i] + t * d1;
  r.dy0 = (d0 / dl) / e;
  r.dy1 = (d1 / dl) / e;
  return r;
}
// Inverse range: energy for a grammage, with dE/dX (0 when clamped).
__device__ __forceinline__ Lookup inverse_range(const double* __restrict__ es,
                                                const double* __restrict__ les,
                                                const double* __restrict__ rt, int n, int pow2,
                                                double x) {
  Lookup r;
  if (x <= rt[0]) { r.y = es[0]; r.dy = 0.0; return r; }
  if (x >= rt[n - 1]) { r.y = es[n - 1]; r.dy = 0.0; return r; }
  int i = bracket(rt, n, pow2, x);
  double dr = rt[i + 1] - rt[i];
  double t = (x - rt[i]) / dr;
  double dle = les[i + 1] - les[i];
  double le = les[i] + t * dle;
  r.y = exp(le);
  r.dy = r.y * (dle / dr);
  return r;
}

// ------------------------------------------------------- forward-mode dual
struct Dd {
  double v, d;
};
__device__ __forceinline__ Dd dmul(Dd a, Dd b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dd ddiv(Dd a, Dd b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dd dadd(Dd a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dd dpow(Dd a, double p) {
  return {pow(a.v, p), p * pow(a.v, p - 1.0) * a.d};
}

// ------------------------------------------------------------- flux model
// charge_fraction, cos_theta_star, flux_gaisser, flux_gccly:
// src/GaisserFlux.jl:36-133.
__device__ __forceinline__ double charge_fraction(double q) {
  const double r = 1.2766;
  if (q < 0) return 1.0 / (1.0 + r);
  if (q > 0) return r / (1.0 + r);
  return 1.0;
}
__device__ __forceinline__ double cos_theta_star(double c) {
  const double p0 = 0.102573, p1 = -0.068287, p2 = 0.958633, p3 = 0.0407253, p4 = 0.817285;
  double cs2 = (c * c + p0 * p0 + p1 * pow(c, p2) + p3 * pow(c, p4)) / (1.0 + p0 * p0 + p1 + p3);
  return cs2 > 0 ? sqrt(cs2) : 0.0;
}
// Value and d/dK of flux_gccly at fixed cos(theta) and charge.
__device__ __forceinline__ Dd flux_gccly_d(double c, double K, double q) {
  Dd Emu = {K + MUON_MASS, 1.0};
  double cs = cos_theta_star(c);
  double cs129 = pow(cs, 1.29);
  Dd f = dpow(dadd(ddiv({3.64, 0.0}, dmul(Emu, {cs129, 0.0})), 1.0), -2.7);
  Dd ec = {1.1 * Emu.v * cs, 1.1 * cs};
  Dd rpi = {1.0 + ec.v / 115.0, ec.d / 115.0};
  Dd rK = {1.0 + ec.v / 850.0, ec.d / 850.0};
  Dd t = ddiv({1.0, 0.0}, rpi);
  Dd u = ddiv({0.054, 0.0}, rK);
  Dd s = {t.v + u.v, t.d + u.d};
  Dd p = dpow(Emu, -2.7);
  Dd g = dmul({1.4e3 * p.v, 1.4e3 * p.d}, s);
  double cf = charge_fraction(q);
  g.v *= cf;
  g.d *= cf;
  return dmul(f, g);
}
__device__ __forceinline__ double flux_gccly(double c, double K, double q) {
  double Emu = K + MUON_MASS;
  double cs = cos_theta_star(c);
  double f = pow(1.0 + 3.64 / (Emu * pow(cs, 1.29)), -2.7);
  double ec = 1.1 * Emu * cs;
  double rpi = 1.0 + ec / 115.0;
  double rK = 1.0 + ec / 850.0;
  return f * (1.4e3 * pow(Emu, -2.7) * (1.0 / rpi + 0.054 / rK) * charge_fraction(q));
}

// compute_air_grammage, src/Geometry.jl:1043-1055
__device__ __forceinline__ double air_grammage(double z0, double z1, double c) {
  const double rho0 = 1.205, h = 12e3;
  double cs = fmax(fabs(c), 0.01);
  double g = (rho0 * h / cs) * (exp(-z0 / h) - exp(-z1 / h));
  return fmax(g, 0.0);
}

// compute_decay_weight_from_path, src/Geometry.jl:1083-1107: value and
// d log(decay) / d E_avg.
__device__ __forceinline__ double decay_weight_dlog(double path, double eavg, double mass,
                                                    double ctau, double* dlog) {
  double g = 1.0 + eavg / mass;
  double b2 = 1.0 - 1.0 / (g * g);
  double beta, dbg_dg;
  if (b2 >= 1e-12) {
    beta = sqrt(b2);
    dbg_dg = 1.0 / beta;  // d(beta*gamma)/d gamma = gamma / sqrt(gamma^2 - 1)
  } else {
    beta = 1e-6;
    dbg_dg = beta;
  }
  double L = beta * g * ctau;
  double w = exp(-path / L);
  *dlog = path / (L * L) * (ctau * dbg_dg / mass);
  return w;
}

// ------------------------------------------------------------ Philox4x32-10
// Salmon et al. (SC'11).  key = (seed lo, seed hi), counter = (event lo,
// event hi, block, 0).  A uniform double on [0,1) takes two words.
struct Philox {
  uint32_t k0, k1, e0, e1, block;
  uint32_t buf[4];
  int pos;
  __device__ __forceinline__ void init(uint64_t seed, uint64_t event) {
    k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
    e0 = (uint32_t)event; e1 = (uint32_t)(event >> 32);
    block = 0; pos = 4;
  }
  __device__ __forceinline__ void refill() {
    uint32_t c0 = e0, c1 = e1, c2 = block++, c3 = 0, a = k0, b = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
      uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
      uint32_t o0 = hi1 ^ c1 ^ a, o2 = hi0 ^ c3 ^ b;
      c0 = o0; c1 = lo1; c2 = o2; c3 = lo0;
      a += 0x9E3779B9u; b += 0xBB67AE85u;
    }
    buf[0] = c0; buf[1] = c1; buf[2] = c2; buf[3] = c3;
    pos = 0;
  }
  __device__ __forceinline__ uint32_t u32() {
    if (pos == 4) refill();
    uint32_t v = buf[0];
    // rotate the buffer so the index stays a compile-time constant
    buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3];
    ++pos;
    return v;
  }
  __device__ __forceinline__ double uniform() {
    uint64_t a = u32();
    uint64_t b = u32();
    return (double)(((a << 32) | b) >> 11) * 0x1.0p-53;
  }
};

// ------------------------------------------------------ wave reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Value of lane Q of the caller's quad of lanes (DPP quad_perm broadcast);
// every lane of the quad must be active.
template <int Q>
__device__ __forceinline__ int quad_bcast(int v) {
  return __builtin_amdgcn_update_dpp(0, v, Q * 0x55, 0xf, 0xf, false);
}
template <int Q>
__device__ __forceinline__ double quad_bcast(double v) {
  const int hi = quad_bcast<Q>(__double2hiint(v)), lo = quad_bcast<Q>(__double2loint(v));
  return __hiloint2double(hi, lo);
}

}  // namespace dpk

namespace dpk {
// ------------------------------------------- generic (value or dual) lookups
// Used where a tangent must flow through long adaptive loops (the
// deterministic CSDA stepping): the same code runs on double or on DualD.
struct DualD {
  double v, d;
  __device__ DualD() : v(0), d(0) {}
  __device__ DualD(double x) : v(x), d(0) {}
  __device__ DualD(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(const DualD& x) { return x.v; }
__device__ __forceinline__ DualD operator+(DualD a, DualD b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ DualD operator-(DualD a, DualD b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ DualD operator-(DualD a) { return {-a.v, -a.d}; }
__device__ __forceinline__ DualD operator*(DualD a, DualD b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ DualD operator/(DualD a, DualD b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ DualD operator+(DualD a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ DualD operator+(double a, DualD b) { return {a + b.v, b.d}; }
__device__ __forceinline__ DualD operator-(DualD a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ DualD operator-(double a, DualD b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ DualD operator*(DualD a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ DualD operator*(double a, DualD b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ DualD operator/(DualD a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ DualD operator/(double a, DualD b) {
  double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__device__ __forceinline__ DualD log(DualD a) { return {::log(a.v), a.d / a.v}; }
__device__ __forceinline__ DualD exp(DualD a) { double e = ::exp(a.v); return {e, e * a.d}; }
__device__ __forceinline__ DualD sqrt(DualD a) { double s = ::sqrt(a.v); return {s, a.d / (2.0 * s)}; }
__device__ __forceinline__ DualD pow(DualD a, double p) {
  return {::pow(a.v, p), p * ::pow(a.v, p - 1.0) * a.d};
}
// Two tangent directions (segment Jacobians in (w, E_in)).
struct D2 {
  double v, a, b;
  __device__ D2() : v(0), a(0), b(0) {}
  __device__ D2(double x) : v(x), a(0), b(0) {}
  __device__ D2(double x, double da, double db) : v(x), a(da), b(db) {}
};
__device__ __forceinline__ double val(const D2& x) { return x.v; }
__device__ __forceinline__ D2 operator+(D2 x, D2 y) { return {x.v + y.v, x.a + y.a, x.b + y.b}; }
__device__ __forceinline__ D2 operator-(D2 x, D2 y) { return {x.v - y.v, x.a - y.a, x.b - y.b}; }
__device__ __forceinline__ D2 operator-(D2 x) { return {-x.v, -x.a, -x.b}; }
__device__ __forceinline__ D2 operator*(D2 x, D2 y) {
  return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b};
}
__device__ __forceinline__ D2 operator/(D2 x, D2 y) {
  double q = x.v / y.v;
  return {q, (x.a - q * y.a) / y.v, (x.b - q * y.b) / y.v};
}
__device__ __forceinline__ D2 operator+(D2 x, double y) { return {x.v + y, x.a, x.b}; }
__device__ __forceinline__ D2 operator+(double x, D2 y) { return {x + y.v, y.a, y.b}; }
__device__ __forceinline__ D2 operator-(D2 x, double y) { return {x.v - y, x.a, x.b}; }
__device__ __forceinline__ D2 operator-(double x, D2 y) { return {x - y.v, -y.a, -y.b}; }
__device__ __forceinline__ D2 operator*(D2 x, double y) { return {x.v * y, x.a * y, x.b * y}; }
__device__ __forceinline__ D2 operator*(double x, D2 y) { return {x * y.v, x * y.a, x * y.b}; }
__device__ __forceinline__ D2 operator/(D2 x, double y) { return {x.v / y, x.a / y, x.b / y}; }
__device__ __forceinline__ D2 operator/(double x, D2 y) {
  double q = x / y.v;
  return {q, -q * y.a / y.v, -q * y.b / y.v};
}
__device__ __forceinline__ D2 log(D2 x) { return {::log(x.v), x.a / x.v, x.b / x.v}; }
__device__ __forceinline__ D2 exp(D2 x) { double e = ::exp(x.v); return {e, e * x.a, e * x.b}; }
__device__ __forceinline__ D2 sqrt(D2 x) {
  double s = ::sqrt(x.v);
  return {s, x.a / (2.0 * s), x.b / (2.0 * s)};
}
__device__ __forceinline__ D2 pow(D2 x, double p) {
  double d = p * ::pow(x.v, p - 1.0);
  return {::pow(x.v, p), d * x.a, d * x.b};
}
// Reciprocal and reciprocal square root of a positive normal double: the
// hardware estimate (about 5e-8 relative on gfx950) refined by two Newton
// steps, in 5 operations instead of the 10 of a correctly rounded quotient.
// Measured over 2^20 values spread over 26 decades: the reciprocal equal to
// 1.0 / x, the reciprocal square root within 2.6e-16 of 1.0 / sqrt(x) (one
// Newton step leaves 2.2e-15 and 4.1e-15).  A quotient a * rcp_nr(b) is two
// roundings instead of one.  For the deterministic CSDA kernels, compared with
// the oracle at 1e-10; the MC kernels keep correctly rounded quotients.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rsqrt_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  r = fma(r, fma(-hx * r, r, 0.5), r);
  return fma(r, fma(-hx * r, r, 0.5), r);
}
// The same for the FP32 stepping map (compared at 1e-5).
__device__ __forceinline__ float rcpf_nr(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.f), r);
}

// Quotients, logarithm and square root built on them, for values and duals.
__device__ __forceinline__ double qdiv(double a, double b) { return a * rcp_nr(b); }
__device__ __forceinline__ DualD qdiv(DualD a, DualD b) {
  const double r = rcp_nr(b.v), q = a.v * r;
  return {q, (a.d - q * b.d) * r};
}
__device__ __forceinline__ DualD qdiv(DualD a, double b) {
  const double r = rcp_nr(b);
  return {a.v * r, a.d * r};
}
__device__ __forceinline__ DualD qdiv(double a, DualD b) {
  const double r = rcp_nr(b.v), q = a * r;
  return {q, -q * b.d * r};
}
__device__ __forceinline__ double qlog(double x) { return ::log(x); }
__device__ __forceinline__ DualD qlog(DualD a) { return {::log(a.v), a.d * rcp_nr(a.v)}; }
__device__ __forceinline__ D2 qdiv(D2 x, D2 y) {
  const double r = rcp_nr(y.v), q = x.v * r;
  return {q, (x.a - q * y.a) * r, (x.b - q * y.b) * r};
}
__device__ __forceinline__ D2 qdiv(D2 x, double y) {
  const double r = rcp_nr(y);
  return {x.v * r, x.a * r, x.b * r};
}
__device__ __forceinline__ D2 qdiv(double x, D2 y) {
  const double r = rcp_nr(y.v), q = x * r;
  return {q, -q * y.a * r, -q * y.b * r};
}
__device__ __forceinline__ D2 qlog(D2 x) {
  const double r = rcp_nr(x.v);
  return {::log(x.v), x.a * r, x.b * r};
}
__device__ __forceinline__ D2 qsqrt(D2 x) {
  const double s = ::sqrt(x.v), h = 0.5 * rcp_nr(s);
  return {s, x.a * h, x.b * h};
}
__device__ __forceinline__ double qsqrt(double x) { return ::sqrt(x); }
__device__ __forceinline__ DualD qsqrt(DualD a) {
  const double s = ::sqrt(a.v);
  return {s, a.d * (0.5 * rcp_nr(s))};
}

using ::log;
using ::exp;
using ::sqrt;
using ::pow;

template <class R>
__device__ __forceinline__ R interp_r(const double* __restrict__ es, const double* __restrict__ les,
                                      const double* __restrict__ ys, int n, int pow2, R x, R lx) {
  if (val(x) <= es[0]) return R(ys[0]);
  if (val(x) >= es[n - 1]) return R(ys[n - 1]);
  int i = bracket(les, n, pow2, val(lx));
  R t = (lx - les[i]) / (les[i + 1] - les[i]);
  return ys[i] + t * (ys[i + 1] - ys[i]);
}
template <class R>
__device__ __forceinline__ R inverse_range_r(const double* __restrict__ es,
                                             const double* __restrict__ les,
                                             const double* __restrict__ rt, int n, int pow2, R x) {
  if (val(x) <= rt[0]) return R(es[0]);
  if (val(x) >= rt[n - 1]) return R(es[n - 1]);
  int i = bracket(rt, n, pow2, val(x));
  R t = (x - rt[i]) / (rt[i + 1] - rt[i]);
  R le = les[i] + t * (les[i + 1] - les[i]);
  return exp(le);
}
// flux_gccly on a generic energy type (cos and charge fixed).
template <class R>
__device__ __forceinline__ R flux_gccly_r(double c, R K, double q) {
  R Emu = K + MUON_MASS;
  double cs = cos_theta_star(c);
  R f = pow(1.0 + 3.64 / (Emu * ::pow(cs, 1.29)), -2.7);
  R ec = 1.1 * Emu * cs;
  R rpi = 1.0 + ec / 115.0;
  R rK = 1.0 + ec / 850.0;
  return f * (1.4e3 * pow(Emu, -2.7) * (1.0 / rpi + 0.054 / rK) * charge_fraction(q));
}
}  // namespace dpk

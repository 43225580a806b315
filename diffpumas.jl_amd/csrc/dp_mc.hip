// Flat two-layer backward Monte Carlo (run_backward_mc), one event per lane.
//
// Event e draws from Philox4x32-10 keyed by (seed, e): its first uniform
// samples the energy (log-uniform), the second the charge, later ones feed
// the transport.  Each lane writes its event's contribution; a fixed-shape
// tree then reduces the per-event values, so flux and sigma are
// bitwise independent of the launch shape.
//
// CSDA mode (straggling = scattering = false) walks the deterministic CSDA
// stepping of transport_backward_through_geometry (src/Geometry.jl:534-672)
// and, with compute_gradient, carries the forward-mode tangent d/d rho
// through the adaptive step sizes (src/Geometry.jl:933-1028).
#include "dp_internal.hpp"
#include "dp_device.hpp"
#include <cmath>

using namespace dpk;

namespace {

constexpr int MC_ARRS = 5;  // energy, log energy, csda range, csda stopping, transport path

struct FlatGeo {
  double thick, rho;
  int rock, air, porous;
  double porous_rho, porous_thick, palt;
};

struct MatView {
  const double *E, *LE, *R, *S, *TP;
};

template <class R>
struct Walk {
  R energy, weight, z;
};

// transport_backward_step, src/Geometry.jl:424-457 (deterministic branch)
template <class R, class D>
__device__ __forceinline__ void csda_step(const MatView& m, int K, int pow2, double mass,
                                          double ctau, Walk<R>& s, D density, R ds, double uz) {
  R ki = s.energy;
  R gs = ds * density;
  R lki = log(ki);
  R Xi = interp_r(m.E, m.LE, m.R, K, pow2, ki, lki);
  R kf = inverse_range_r(m.E, m.LE, m.R, K, pow2, Xi + gs);
  R gi = 1.0 + ki / mass;
  R gc = 1.0 + kf / mass;
  R ga = (gi + gc) / 2.0;
  R b2 = 1.0 - 1.0 / (ga * ga);
  R beta = sqrt(val(b2) >= 1e-12 ? b2 : R(1e-12));
  R dt = ds / (beta * ga);
  s.z = s.z - ds * (-uz);
  R si = interp_r(m.E, m.LE, m.S, K, pow2, ki, lki);
  R sc = interp_r(m.E, m.LE, m.S, K, pow2, kf, log(kf));
  R wf = sc / si;
  R decay = exp(-dt / ctau);
  s.weight = s.weight * wf * decay;
  s.energy = kf;
}

// transport_backward_through_geometry with deterministic CSDA steps,
// src/Geometry.jl:534-672, and the final acceptance of compute_flux_single
// (:739-757).  Returns the weighted flux (0 when rejected).
template <class R>
__device__ R flat_csda_flux(const MatView* mv, int K, int pow2, double mass, double ctau,
                            const FlatGeo& g, R rock_density, double energy0, double uz,
                            double charge, double threshold, double mode_switch, bool scat_limit,
                            int* err) {
  Walk<R> s{R(energy0), R(1.0), R(0.0)};
  const double palt = g.palt;
  const bool has_porous = g.porous >= 0 && g.porous_thick > 0;
  const double porous_z = has_porous ? g.thick - g.porous_thick : g.thick;
  int outer = 0;
  for (; outer < 64; ++outer) {
    if (!(val(s.energy) < threshold - EPS64)) break;
    double z = val(s.z);
    if (z < 0 || z >= palt - EPS64) break;
    if (val(s.weight) <= 0 || !isfinite(val(s.weight))) break;
    const bool below = val(s.energy) < mode_switch - EPS64;
    const bool use_scat = below ? scat_limit : false;
    const double limit = below ? fmin(mode_switch, threshold) : threshold;
    for (int inner = 0; inner < 50000; ++inner) {
      z = val(s.z);
      if (z < 0 || z >= palt - EPS64) break;
      if (val(s.energy) >= limit - EPS64) break;
      if (val(s.weight) <= 0 || !isfinite(val(s.weight))) break;
      const bool in_air = !(z < g.thick);
      int mi;
      R density, stb;
      if (!in_air) {
        if (has_porous && z >= porous_z) { mi = 2; density = R(g.porous_rho); }
        else { mi = 0; density = rock_density; }
        if (uz > EPS64) stb = (((has_porous && z < porous_z) ? porous_z : g.thick) - s.z) / uz;
        else if (uz < -EPS64) stb = (((has_porous && z >= porous_z) ? porous_z : 0.0) - s.z) / uz;
        else stb = R(1000.0);
      } else {
        mi = 1;
        density = 1.205 * exp(-s.z / 12e3);
        if (uz > EPS64) stb = (palt - s.z) / uz;
        else if (uz < -EPS64) stb = (g.thick - s.z) / uz;
        else stb = R(1000.0);
      }
      const MatView& m = mv[mi];
      // _backward_geometric_step_limit, src/Geometry.jl:463-476
      R Xi = interp_r(m.E, m.LE, m.R, K, pow2, s.energy, log(s.energy));
      R lim = 0.01 * Xi / (val(density) < 1e-6 ? R(1e-6) : density);
      if (use_scat) {
        R lb1 = interp_r(m.E, m.LE, m.TP, K, pow2, s.energy, log(s.energy));
        if (val(lb1) > 0 && val(density) > 0) {
          R sl = 0.01 * lb1 / density;
          if (val(sl) < val(lim)) lim = sl;
        }
      }
      if (in_air) {
        double al = 0.05 * 12e3 / fmax(fabs(uz), 0.05);
        if (al < val(lim)) lim = R(al);
      }
      R step = stb + 1e-7;
      if (val(lim) < val(step)) step = lim;
      if (val(step) > 1000.0) step = R(1000.0);
      else if (val(step) < 1e-7) step = R(1e-7);
      csda_step(m, K, pow2, mass, ctau, s, density, step, uz);
    }
  }
  if (outer == 64) atomicOr(err, 2);
  if (val(s.z) >= palt - 1.0 && val(s.energy) < threshold) {
    double cf = fmin(fmax(uz, 0.0), 1.0);
    if (cf <= 0) return R(0.0);
    return s.weight * flux_gccly_r(cf, s.energy, charge);
  }
  return R(0.0);
}

struct McArgs {
  const double* tab;
  int K, pow2;
  double mass, ctau;
  FlatGeo g, gg;  // MC geometry, gradient-surrogate geometry
  bool same_geo;
  double cos_t, emin, emax, thr, sw;
  int64_t n;
  uint64_t seed;
  bool grad, scat_limit;
  double* out_w;   // n
  double* out_gf;  // n (gradient flux contribution)
  double* out_gd;  // n (gradient contribution)
  int* err;
};

__global__ __launch_bounds__(256) void mc_flat_csda_kernel(McArgs a) {
  extern __shared__ double lds[];
  const int K = a.K;
  MatView mv[3];
  const int mats[3] = {a.g.rock, a.g.air, a.g.porous >= 0 ? a.g.porous : a.g.rock};
  for (int j = 0; j < 3; ++j) {
    double* base = lds + (size_t)j * MC_ARRS * K;
    const double* t = a.tab + (size_t)mats[j] * 19 * K;
    for (int i = threadIdx.x; i < K; i += blockDim.x) {
      base[i] = t[i];                 // energy
      base[K + i] = t[K + i];         // log energy
      base[2 * K + i] = t[3 * K + i]; // csda range
      base[3 * K + i] = t[2 * K + i]; // csda stopping power
      base[4 * K + i] = t[13 * K + i];// transport path
    }
    mv[j] = {base, base + K, base + 2 * K, base + 3 * K, base + 4 * K};
  }
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.n; e += stride) {
    Philox rng;
    rng.init(a.seed, (uint64_t)e);
    double u = rng.uniform();
    double k, w;
    double rk = log(a.emax / a.emin);
    if (rk > 0) { k = a.emin * exp(rk * u); w = k * rk; }
    else { k = a.emin; w = 1.0; }
    double q = rng.uniform() > 0.5 ? 1.0 : -1.0;
    w *= 2.0;
    if (a.grad) {
      DualD fd = flat_csda_flux<DualD>(mv, K, a.pow2, a.mass, a.ctau, a.gg, DualD(a.g.rho, 1.0), k,
                                       a.cos_t, q, a.thr, a.sw, false, a.err);
      a.out_gf[e] = w * fd.v;
      a.out_gd[e] = w * fd.d;
      if (a.same_geo) {
        a.out_w[e] = w * fd.v;
        continue;
      }
    }
    double f = flat_csda_flux<double>(mv, K, a.pow2, a.mass, a.ctau, a.g, a.g.rho, k, a.cos_t, q,
                                      a.thr, a.sw, a.scat_limit, a.err);
    a.out_w[e] = w * f;
  }
}

// Fixed-shape deterministic reduction of per-event values:
// each block sums CHUNK consecutive events (every thread a contiguous run,
// then a fixed tree), producing one partial per chunk.
constexpr int RED_THREADS = 256;
constexpr int RED_PER_THREAD = 16;
constexpr int64_t RED_CHUNK = (int64_t)RED_THREADS * RED_PER_THREAD;

// mode 0: (sum x, sum x^2) of x = in0; mode 1: sum of in0, in1 (two sums)
__global__ __launch_bounds__(RED_THREADS) void reduce_kernel(const double* __restrict__ in0,
                                                             const double* __restrict__ in1,
                                                             int64_t n, int square,
                                                             double* __restrict__ out0,
                                                             double* __restrict__ out1) {
  __shared__ double s0[RED_THREADS], s1[RED_THREADS];
  const int64_t base = (int64_t)blockIdx.x * RED_CHUNK + (int64_t)threadIdx.x * RED_PER_THREAD;
  double a0 = 0.0, a1 = 0.0;
  for (int i = 0; i < RED_PER_THREAD; ++i) {
    int64_t j = base + i;
    if (j < n) {
      double x = in0[j];
      a0 += x;
      a1 += square ? x * x : in1[j];
    }
  }
  s0[threadIdx.x] = a0;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      s0[threadIdx.x] += s0[threadIdx.x + off];
      s1[threadIdx.x] += s1[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out0[blockIdx.x] = s0[0];
    out1[blockIdx.x] = s1[0];
  }
}

// Reduce pairs of arrays (x, y) of length n to two scalars; `square` makes
// y = x^2 on the first pass.  buf holds 4 * ceil(n / CHUNK) doubles.
int reduce_pair(hipStream_t st, const double* x, const double* y, int64_t n, bool square,
                double* buf, double* res /*device, 2*/) {
  const double* a = x;
  const double* b = y;
  int sq = square ? 1 : 0;
  double* o = buf;
  while (true) {
    int64_t blocks = (n + RED_CHUNK - 1) / RED_CHUNK;
    if (blocks < 1) blocks = 1;
    double* o0 = blocks == 1 ? res : o;
    double* o1 = blocks == 1 ? res + 1 : o + blocks;
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)blocks), dim3(RED_THREADS), 0, st, a, b, n,
                       sq, o0, o1);
    if (hipGetLastError() != hipSuccess) return dph::fail(DP_E_HIP, "reduce_kernel launch failed");
    if (blocks == 1) break;
    a = o0;
    b = o1;
    n = blocks;
    sq = 0;
    o = (o == buf) ? buf + 2 * blocks : buf;  // ping-pong: second half never overlaps
  }
  return DP_OK;
}

__global__ void reset_flags(int* f) { if (threadIdx.x < 4) f[threadIdx.x] = 0; }

}  // namespace

extern "C" {

int dp_mc_flat_backward(dp_ctx c, dp_tables t, const dp_flat_geometry* geo,
                        const dp_mc_config* cfg, dp_mc_result* out) {
  if (!c || !t || !geo || !cfg || !out) return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: null argument");
  if (cfg->n_events <= 0) return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: n_events must be positive");
  const int M = t->n_materials;
  if (geo->rock_material < 0 || geo->rock_material >= M || geo->air_material < 0 ||
      geo->air_material >= M || geo->porous_material >= M)
    return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: material index out of range");
  if (!(cfg->energy_min > 0) || !(cfg->energy_max >= cfg->energy_min))
    return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: need 0 < energy_min <= energy_max");
  double supported = t->last_energy[0];
  for (double e : t->last_energy) supported = e < supported ? e : supported;
  double thr = cfg->energy_threshold > 0 ? cfg->energy_threshold : fmin(1e9, supported);
  // compute_flux_single, src/Geometry.jl:725-727
  if (!(thr > 0 && thr <= supported))
    return dph::fail(DP_E_ARGUMENT, "source_energy_max must be positive and within the physics energy grid");
  if (cfg->straggling || cfg->scattering)
    return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: straggled/scattering transport needs prepared mixed kernels (dp_mc_flat_straggled)");
  DP_HIP(hipSetDevice(c->device));
  const int64_t n = cfg->n_events;
  const int64_t nparts = (n + RED_CHUNK - 1) / RED_CHUNK + 1;
  size_t o_w = 0, o_gf = dph::align_up(8 * (size_t)n, 256);
  size_t o_gd = dph::align_up(o_gf + 8 * (size_t)n, 256);
  size_t o_buf = dph::align_up(o_gd + 8 * (size_t)n, 256);
  size_t o_res = dph::align_up(o_buf + 8 * 4 * (size_t)nparts, 256);
  size_t o_flag = o_res + 256;
  int rc = dph::ensure(c->scratch, o_flag + 256);
  if (rc) return rc;
  char* b = (char*)c->scratch.ptr;
  McArgs a;
  a.tab = t->data; a.K = t->n_knots; a.pow2 = t->pow2; a.mass = t->mass; a.ctau = t->ctau;
  a.g = {geo->rock_thickness, geo->rock_density, geo->rock_material, geo->air_material,
         geo->porous_material, geo->porous_density, geo->porous_thickness, geo->primary_altitude};
  a.gg = {geo->rock_thickness, geo->rock_density, geo->rock_material, geo->air_material, -1, 0.0,
          0.0, 3e4};
  a.same_geo = !(geo->porous_material >= 0 && geo->porous_thickness > 0) &&
               geo->primary_altitude == 3e4 && !cfg->scattering;
  a.cos_t = cfg->cos_theta; a.emin = cfg->energy_min; a.emax = cfg->energy_max;
  a.thr = thr; a.sw = cfg->mode_switch; a.n = n; a.seed = cfg->seed;
  a.grad = cfg->compute_gradient != 0; a.scat_limit = cfg->scattering != 0;
  a.out_w = (double*)(b + o_w); a.out_gf = (double*)(b + o_gf); a.out_gd = (double*)(b + o_gd);
  int* flags = (int*)(b + o_flag);
  a.err = flags;
  size_t lds = (size_t)3 * MC_ARRS * t->n_knots * sizeof(double);
  int occ = 0;
  DP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, mc_flat_csda_kernel, 256, lds));
  if (occ < 1) occ = 1;
  int64_t want = (n + 255) / 256;
  int64_t grid = (int64_t)c->n_cu * occ * 4;
  if (grid > want) grid = want;
  hipLaunchKernelGGL(reset_flags, dim3(1), dim3(64), 0, c->stream, flags);
  (void)hipEventRecord(c->ev0, c->stream);
  hipLaunchKernelGGL(mc_flat_csda_kernel, dim3((unsigned)grid), dim3(256), lds, c->stream, a);
  (void)hipEventRecord(c->ev1, c->stream);
  DP_HIP(hipGetLastError());
  double* res = (double*)(b + o_res);
  double* buf = (double*)(b + o_buf);
  rc = reduce_pair(c->stream, a.out_w, nullptr, n, true, buf, res);
  if (rc) return rc;
  if (a.grad) {
    rc = reduce_pair(c->stream, a.out_gf, a.out_gd, n, false, buf, res + 2);
    if (rc) return rc;
  }
  double h[4] = {0, 0, 0, 0};
  int hf[4] = {0, 0, 0, 0};
  DP_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipStreamSynchronize(c->stream));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  if (hf[0] & 2) return dph::fail(DP_E_EXHAUSTED, "dp_mc_flat_backward: outer stepping cap reached");
  const double dn = (double)n;
  out->flux = h[0] / dn;
  double var = h[1] / dn - out->flux * out->flux;
  out->sigma = geo->rock_thickness > 0 ? sqrt(fmax(var, 0.0) / dn) : 0.0;
  out->gradient_flux = a.grad ? h[2] / dn : NAN;
  out->gradient = a.grad ? h[3] / dn : NAN;
  out->source_energy_max = thr;
  out->n_events = n;
  return DP_OK;
}

}  // extern "C"

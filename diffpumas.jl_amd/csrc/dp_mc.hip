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
#include "dp_transport.hpp"
#include "dp_transport_fast.hpp"
#include "dp_del_defer.hpp"
#include "dp_media.hpp"
#include <algorithm>
#include <cmath>

using namespace dpk;

namespace {

constexpr int MC_ARRS = 5;  // energy, log energy, csda range, csda stopping, transport path

// Stochastic MC blocks: one wave, with a pool of MC_POOL deferred hard-event
// records evaluated by the whole wave after each round of 64 events.
constexpr int MC_THREADS = 64;
#ifndef DP_MC_POOL
#define DP_MC_POOL 192  // a smaller pool (host emulation builds) exercises the inline fallback
#endif
constexpr int MC_POOL = DP_MC_POOL;
#ifndef DP_MC_WAVES
#define DP_MC_WAVES 4  // waves per SIMD the register budget is sized for (128 VGPRs)
#endif
constexpr int MC_WAVES = DP_MC_WAVES;
constexpr int MC_BUCKETS = 256;  // source-energy buckets of the event order

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
  if (outer == 64) atomicOr(err, ERR_CAP);
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
  HIP_DYNAMIC_SHARED(double, lds)
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
      if (a.same_geo || !a.out_w) {
        if (a.out_w) a.out_w[e] = w * fd.v;
        continue;
      }
    }
    if (!a.out_w) continue;
    double f = flat_csda_flux<double>(mv, K, a.pow2, a.mass, a.ctau, a.g, a.g.rho, k, a.cos_t, q,
                                      a.thr, a.sw, a.scat_limit, a.err);
    a.out_w[e] = w * f;
  }
}

// ------------------------------------------------- stochastic transport
// The table and medium descriptors that the out-of-line step functions take
// by reference live in device memory (Tg, media), not in the kernel
// arguments: a reference into the arguments makes the compiler copy them to
// every lane's stack, and each descriptor read in a step becomes a per-lane
// scratch load instead of a load that all lanes of a wave share.
struct StochArgs {
  DevTables T;
  const DevTables* Tg;  // T in device memory
  bool has_porous;
  double thick, rock_rho, porous_rho, porous_thick, palt;
  double cos_t, sin_t, emin, emax, thr, sw;
  bool straggling, scattering;
  int64_t n;
  uint64_t seed;
  double* out_w;
  int* err;            // [0] error bits, [1] spare
  unsigned long long* bad_event;
  const Medium* media;  // rock, air, porous in device memory (DelLog::medium indexes it)
  const uint32_t* order;  // events in order of source-energy bucket (null: identity)
};


// transport_backward_through_geometry with an RNG, src/Geometry.jl:534-672,
// and the acceptance of compute_flux_single (:739-757).
__device__ double flat_event(const StochArgs& a, double energy, double charge, Philox& rng,
                             int& err) {
  const DevTables& T = *a.Tg;
  McState st;
  st.E = energy; st.w = 1.0; st.t = 0.0; st.L = 0.0;
  st.pos[0] = st.pos[1] = st.pos[2] = 0.0;
  st.dir[0] = -a.sin_t; st.dir[1] = 0.0; st.dir[2] = -a.cos_t;
  const double palt = a.palt;
  int outer = 0;
  for (; outer < 64; ++outer) {
    if (!(st.E < a.thr - EPS64)) break;
    double z = st.pos[2];
    if (z < 0 || z >= palt - EPS64) break;
    if (st.w <= 0 || !isfinite(st.w)) break;
    const bool below = st.E < a.sw - EPS64;
    const bool use_str = below && a.straggling;
    const bool use_mixed = !below && a.straggling;
    const bool use_scat = below && a.scattering;
    const double limit = below ? fmin(a.sw, a.thr) : a.thr;
    for (int inner = 0; inner < 50000; ++inner) {
      z = st.pos[2];
      if (z < 0 || z >= palt - EPS64) break;
      if (st.E >= limit - EPS64) break;
      if (st.w <= 0 || !isfinite(st.w)) break;
      const double uz = -st.dir[2];
      const double porous_z = a.has_porous ? a.thick - a.porous_thick : a.thick;
      const Medium* md;
      double density, stb;
      const bool in_air = !(z < a.thick);
      if (!in_air) {
        if (a.has_porous && z >= porous_z) { md = &a.media[2]; density = a.porous_rho; }
        else { md = &a.media[0]; density = a.rock_rho; }
        if (uz > EPS64) stb = (((a.has_porous && z < porous_z) ? porous_z : a.thick) - z) / uz;
        else if (uz < -EPS64) stb = (((a.has_porous && z >= porous_z) ? porous_z : 0.0) - z) / uz;
        else stb = 1000.0;
      } else {
        md = &a.media[1];
        density = 1.205 * exp(-z / 12e3);
        if (uz > EPS64) stb = (palt - z) / uz;
        else if (uz < -EPS64) stb = (a.thick - z) / uz;
        else stb = 1000.0;
      }
      // _backward_geometric_step_limit, src/Geometry.jl:463-476
      const double le = log(st.E);
      const double Xi = mix_range(T, md->mix, false, st.E, le);
      double lim = 0.01 * Xi / fmax(density, 1e-6);
      if (use_scat) {
        const double lb1 = mix_transport_path(T, md->mix, st.E, le);
        if (lb1 > 0 && density > 0) lim = fmin(lim, 0.01 * lb1 / density);
      }
      if (in_air) lim = fmin(lim, 0.05 * 12e3 / fmax(fabs(uz), 0.05));
      double step = fmin(stb + 1e-7, lim);
      step = fmin(fmax(step, 1e-7), 1000.0);
      if (use_str) straggled_step(T, *md, st, density, step, use_scat, limit, rng, err);
      else if (use_mixed) mixed_step(T, *md, st, density, step, limit, rng, err);
      else csda_mix_step(T, *md, st, density, step);
      if (err) return 0.0;
    }
  }
  if (outer == 64) { err |= ERR_CAP; return 0.0; }
  if (st.pos[2] >= palt - 1.0 && st.E < a.thr) {
    const double cf = fmin(fmax(-st.dir[2], 0.0), 1.0);
    if (cf <= 0) return 0.0;
    return st.w * flux_gccly(cf, st.E, charge);
  }
  return 0.0;
}

// flat_event on tables with one shared energy grid: each step locates the
// event's energy once and hands it to the step functions.
__device__ double flat_event_f(const StochArgs& a, double energy, double charge, Philox& rng,
                               int& err, DelLog* dlog) {
  const DevTables& T = *a.Tg;
  McState st;
  st.E = energy; st.w = 1.0; st.t = 0.0; st.L = 0.0;
  st.pos[0] = st.pos[1] = st.pos[2] = 0.0;
  st.dir[0] = -a.sin_t; st.dir[1] = 0.0; st.dir[2] = -a.cos_t;
  const double palt = a.palt;
  int g = T.K / 2;  // interval guess of st.E
  int outer = 0;
  for (; outer < 64; ++outer) {
    if (!(st.E < a.thr - EPS64)) break;
    double z = st.pos[2];
    if (z < 0 || z >= palt - EPS64) break;
    if (st.w <= 0 || !isfinite(st.w)) break;
    const bool below = st.E < a.sw - EPS64;
    const bool use_str = below && a.straggling;
    const bool use_mixed = !below && a.straggling;
    const bool use_scat = below && a.scattering;
    const double limit = below ? fmin(a.sw, a.thr) : a.thr;
    for (int inner = 0; inner < 50000; ++inner) {
      z = st.pos[2];
      if (z < 0 || z >= palt - EPS64) break;
      if (st.E >= limit - EPS64) break;
      if (st.w <= 0 || !isfinite(st.w)) break;
      const double uz = -st.dir[2];
      const double porous_z = a.has_porous ? a.thick - a.porous_thick : a.thick;
      const Medium* md;
      double density, stb;
      const bool in_air = !(z < a.thick);
      if (!in_air) {
        if (a.has_porous && z >= porous_z) { md = &a.media[2]; density = a.porous_rho; }
        else { md = &a.media[0]; density = a.rock_rho; }
        if (uz > EPS64) stb = (((a.has_porous && z < porous_z) ? porous_z : a.thick) - z) / uz;
        else if (uz < -EPS64) stb = (((a.has_porous && z >= porous_z) ? porous_z : 0.0) - z) / uz;
        else stb = 1000.0;
      } else {
        md = &a.media[1];
        density = 1.205 * exp(-z / 12e3);
        if (uz > EPS64) stb = (palt - z) / uz;
        else if (uz < -EPS64) stb = (a.thick - z) / uz;
        else stb = 1000.0;
      }
      // _backward_geometric_step_limit, src/Geometry.jl:463-476
      const Loc L = loc_e(a.T, st.E, log(st.E), g);
      const double Xi = mix_range_L(a.T, md->mix, false, L);
      double lim = 0.01 * Xi / fmax(density, 1e-6);
      double lb0 = 0.0;
      if (use_scat) {
        lb0 = mix_transport_path_L(a.T, md->mix, L);
        if (lb0 > 0 && density > 0) lim = fmin(lim, 0.01 * lb0 / density);
      }
      if (in_air) lim = fmin(lim, 0.05 * 12e3 / fmax(fabs(uz), 0.05));
      double step = fmin(stb + 1e-7, lim);
      step = fmin(fmax(step, 1e-7), 1000.0);
      g = L.i;
      if (dlog) dlog->medium = (int)(md - a.media);
      if (use_str) straggled_step_f(T, *md, st, L, lb0, density, step, use_scat, limit, rng, err, &g, dlog);
      else if (use_mixed) mixed_step_f(T, *md, st, L, density, step, limit, rng, err, dlog);
      else csda_mix_step_f(T, *md, st, L, density, step, &g);
      if (err) return 0.0;
    }
  }
  if (outer == 64) { err |= ERR_CAP; return 0.0; }
  if (st.pos[2] >= palt - 1.0 && st.E < a.thr) {
    const double cf = fmin(fmax(-st.dir[2], 0.0), 1.0);
    if (cf <= 0) return 0.0;
    return st.w * flux_gccly(cf, st.E, charge);
  }
  return 0.0;
}

// Events in order of source energy: a counting sort of the event indices into
// MC_BUCKETS buckets of the first Philox draw, which fixes the log-uniform
// source energy.  The lanes of a block then start in the same transport mode
// and take similar numbers of steps.  Each event keeps its own stream and
// output slot, so results do not depend on the order.
__device__ __forceinline__ int energy_bucket(uint64_t seed, int64_t e) {
  Philox rng;
  rng.init(seed, (uint64_t)e);
  const double u = rng.uniform();
  return min(MC_BUCKETS - 1, max(0, (int)(u * MC_BUCKETS)));
}
__global__ void bucket_clear(int* cnt) {
  for (int i = threadIdx.x; i < MC_BUCKETS; i += blockDim.x) cnt[i] = 0;
}
__global__ __launch_bounds__(256) void bucket_count(uint64_t seed, int64_t n, int* cnt) {
  __shared__ int h[MC_BUCKETS];
  for (int i = threadIdx.x; i < MC_BUCKETS; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
    atomicAdd(&h[energy_bucket(seed, e)], 1);
  __syncthreads();
  for (int i = threadIdx.x; i < MC_BUCKETS; i += blockDim.x)
    if (h[i]) atomicAdd(&cnt[i], h[i]);
}
// counts -> first position of each bucket (one thread)
__global__ void bucket_offsets(int* cnt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int acc = 0;
  for (int i = 0; i < MC_BUCKETS; ++i) {
    const int c = cnt[i];
    cnt[i] = acc;
    acc += c;
  }
}
// Each block takes rounds of 256 x SCATTER_ITEMS events: ranks inside the
// round come from an LDS histogram, and one global atomic per bucket and
// round reserves the round's range in that bucket (the order inside a bucket
// is arbitrary; results do not depend on it).
constexpr int SCATTER_ITEMS = 8;
__global__ __launch_bounds__(256) void bucket_scatter(uint64_t seed, int64_t n, int* cur,
                                                      uint32_t* order) {
  __shared__ int h[MC_BUCKETS];
  __shared__ int base[MC_BUCKETS];
  const int tid = threadIdx.x;
  const int64_t per = (int64_t)256 * SCATTER_ITEMS;
  for (int64_t b0 = (int64_t)blockIdx.x * per; b0 < n; b0 += (int64_t)gridDim.x * per) {
    for (int i = tid; i < MC_BUCKETS; i += 256) h[i] = 0;
    __syncthreads();
    int bk[SCATTER_ITEMS], rk[SCATTER_ITEMS];
#pragma unroll
    for (int k = 0; k < SCATTER_ITEMS; ++k) {
      const int64_t e = b0 + (int64_t)k * 256 + tid;
      bk[k] = 0;
      rk[k] = -1;
      if (e < n) {
        bk[k] = energy_bucket(seed, e);
        rk[k] = atomicAdd(&h[bk[k]], 1);
      }
    }
    __syncthreads();
    for (int i = tid; i < MC_BUCKETS; i += 256) base[i] = h[i] ? atomicAdd(&cur[i], h[i]) : 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCATTER_ITEMS; ++k) {
      const int64_t e = b0 + (int64_t)k * 256 + tid;
      if (e < n && rk[k] >= 0) {
        const int pos = base[bk[k]] + rk[k];
        if (pos >= 0 && pos < n) order[pos] = (uint32_t)e;
      }
    }
    __syncthreads();   // h and base are reused by the next round
  }
}

// One event per lane, rounds of MC_THREADS events per block (the loop bound is
// uniform over the block).  On tables with a shared energy grid (FAST) the
// deferrable hard-event weights of a round go to the block's pool and are
// evaluated by all its lanes once every lane has finished its event.
template <bool FAST>
__global__ __launch_bounds__(MC_THREADS, MC_WAVES) void mc_flat_stochastic_kernel(StochArgs a) {
  __shared__ PoolRec pool[FAST ? MC_POOL : 1];
  __shared__ int pool_n;
  __shared__ CoopBuf coop;
  const int tid = threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * MC_THREADS;
  for (int64_t base = (int64_t)blockIdx.x * MC_THREADS; base < a.n; base += stride) {
    if (FAST) {
      if (tid == 0) pool_n = 0;
      __syncthreads();  // the previous round's records are applied; the pool is empty
    }
    const int64_t p = base + tid;
    const int64_t e = p < a.n ? (a.order ? (int64_t)a.order[p] : p) : a.n;
    const bool active = e < a.n;
    double v = 0.0;
    int err = 0;
    DelLog dlg{pool, &pool_n, MC_POOL, tid, 0, 0, &coop, a.media, 3};
    if (active) {
      Philox rng;
      rng.init(a.seed, (uint64_t)e);
      const double u = rng.uniform();
      double k, w;
      const double rk = log(a.emax / a.emin);
      if (rk > 0) { k = a.emin * exp(rk * u); w = k * rk; }
      else { k = a.emin; w = 1.0; }
      const double q = rng.uniform() > 0.5 ? 1.0 : -1.0;
      w *= 2.0;
      const double f = FAST ? flat_event_f(a, k, q, rng, err, &dlg) : flat_event(a, k, q, rng, err);
      v = err ? 0.0 : w * f;
    }
    int m = 0;
    if (FAST) {
      __syncthreads();  // every record of the round is in the pool
      m = min(pool_n, MC_POOL);
      pool_eval(pool, m, a.T.mass, [&](int md) -> const DevKernel& {
        return a.media[md < 0 ? 0 : (md > 2 ? 2 : md)].kern;
      });
      __syncthreads();  // every factor is in the pool
    }
    if (active) {
      int fail = err;
      if (FAST && dlg.n > 0) v = pool_apply(pool, m, tid, dlg.n, v, err, fail);
      if (fail) {
        atomicOr(a.err, fail);
        atomicMin(a.bad_event, (unsigned long long)e);
        a.out_w[e] = 0.0;
      } else {
        a.out_w[e] = v;
      }
    }
  }
}

}  // namespace

namespace dph {
// ---------------------------------------------------- host preparation
// Mixture tables (src/Physics.jl:1049-1103) and MixedKernel snapshots
// (src/MixedTransport.jl:55-97) for a medium, built once per physics object
// and (components, fractions), kept in device memory.
struct HostT {
  const dp_tables_s* t;
  const double* a(int m, int arr) const {
    return t->host_data.data() + ((size_t)m * DP_T_NARRAYS + arr) * t->n_knots;
  }
  double interp(int m, int arr, double e) const {
    const int K = t->n_knots;
    const double* es = a(m, DP_T_ENERGY);
    const double* les = a(m, DP_T_LOG_ENERGY);
    const double* ys = a(m, arr);
    if (e <= es[0]) return ys[0];
    if (e >= es[K - 1]) return ys[K - 1];
    const double le = std::log(e);
    int lo = 0, hi = K - 1;
    while (lo < hi - 1) {
      const int mid = (lo + hi) >> 1;
      if (le < les[mid]) hi = mid; else lo = mid;
    }
    const double tt = (le - les[lo]) / (les[lo + 1] - les[lo]);
    return ys[lo] + tt * (ys[lo + 1] - ys[lo]);
  }
};

std::vector<double> medium_key(const std::vector<int>& mats, const std::vector<double>& fr) {
  std::vector<double> k;
  for (size_t i = 0; i < mats.size(); ++i) { k.push_back(mats[i]); k.push_back(fr[i]); }
  return k;
}

int upload_prepared(dph::Prepared& p) {
  const size_t nb = p.host.size() * sizeof(double);
  if (hipMalloc(&p.dev, nb) != hipSuccess) { p.dev = nullptr; return dph::fail(DP_E_HIP, "hipMalloc (prepared tables)"); }
  if (hipMemcpy(p.dev, p.host.data(), nb, hipMemcpyHostToDevice) != hipSuccess)
    return dph::fail(DP_E_HIP, "hipMemcpy (prepared tables)");
  return DP_OK;
}

// Integrated mixture tables on the grid of material 0: [e, le, crange, mrange].
int prepare_mixture(dp_tables_s* t, const std::vector<int>& mats, const std::vector<double>& fr,
                    dph::Prepared** out) {
  auto key = medium_key(mats, fr);
  auto it = t->mixtures.find(key);
  if (it != t->mixtures.end()) { *out = it->second.get(); return DP_OK; }
  HostT H{t};
  const int K = t->n_knots;
  auto p = std::make_unique<dph::Prepared>();
  p->n = K;
  p->host.assign((size_t)4 * K, 0.0);
  double* e = p->host.data();
  double* le = e + K;
  double* cr = e + 2 * K;
  double* mr = e + 3 * K;
  for (int i = 0; i < K; ++i) { e[i] = H.a(0, DP_T_ENERGY)[i]; le[i] = H.a(0, DP_T_LOG_ENERGY)[i]; }
  auto stop = [&](bool mixed, double k) {
    double v = 0;
    for (size_t j = 0; j < mats.size(); ++j) v += fr[j] * H.interp(mats[j], mixed ? DP_T_MIXED_STOP : DP_T_CSDA_STOP, k);
    return v;
  };
  double cp = stop(false, e[0]), mp = stop(true, e[0]);
  cr[0] = e[0] / cp;
  mr[0] = e[0] / mp;
  for (int i = 1; i < K; ++i) {
    const double dK = e[i] - e[i - 1];
    const double c = stop(false, e[i]), m = stop(true, e[i]);
    cr[i] = cr[i - 1] + dK / ((c + cp) / 2);
    mr[i] = mr[i - 1] + dK / ((m + mp) / 2);
    cp = c;
    mp = m;
  }
  int rc = upload_prepared(*p);
  if (rc) return rc;
  *out = p.get();
  t->mixtures[key] = std::move(p);
  return DP_OK;
}

// MixedKernel: [e, stop, sigma, range, hazard, proper, el(4 per element)].
int prepare_kernel(dp_tables_s* t, const std::vector<int>& mats, const std::vector<double>& fr,
                   dph::Prepared** out) {
  auto key = medium_key(mats, fr);
  auto it = t->kernels.find(key);
  if (it != t->kernels.end()) { *out = it->second.get(); return DP_OK; }
  HostT H{t};
  const int K = t->n_knots;
  const double mass = t->mass;
  for (int m : mats)
    if (t->recipes[m].empty())
      return dph::fail(DP_E_ARGUMENT, "stochastic transport needs the element recipe of material " + std::to_string(m));
  double lo = -INFINITY, hi = INFINITY;
  for (int m : mats) { lo = std::max(lo, H.a(m, DP_T_ENERGY)[0]); hi = std::min(hi, H.a(m, DP_T_ENERGY)[K - 1]); }
  std::vector<double> grid;
  for (int m : mats)
    for (int i = 0; i < K; ++i) {
      const double x = H.a(m, DP_T_ENERGY)[i];
      if (lo <= x && x <= hi) grid.push_back(x);
    }
  std::sort(grid.begin(), grid.end());
  grid.erase(std::unique(grid.begin(), grid.end()), grid.end());
  const int refinement = 8;
  std::vector<double> e;
  for (size_t j = 0; j + 1 < grid.size(); ++j) {
    const double a = std::log(grid[j]), b = std::log(grid[j + 1]);
    for (int r = 0; r < refinement; ++r) e.push_back(std::exp(a + (b - a) * r / refinement));
  }
  e.push_back(hi);
  e[0] = lo;
  const int n = (int)e.size();
  std::vector<double> st(n), sg(n), rg(n, 0.0), hz(n, 0.0), pr(n, 0.0);
  for (int j = 0; j < n; ++j) {
    double s = 0, x = 0;
    for (size_t q = 0; q < mats.size(); ++q) {
      s += fr[q] * H.interp(mats[q], DP_T_MIXED_STOP, e[j]);
      x += fr[q] * H.interp(mats[q], DP_T_CROSS_SECTION, e[j]);
    }
    if (!(std::isfinite(s) && s > 0)) return dph::fail(DP_E_ARGUMENT, "invalid restricted stopping power");
    if (!(std::isfinite(x) && x >= 0)) return dph::fail(DP_E_ARGUMENT, "invalid hard cross section");
    st[j] = s;
    sg[j] = x;
  }
  double onset = INFINITY;
  for (double g : grid) {
    double x = 0;
    for (size_t q = 0; q < mats.size(); ++q) x += fr[q] * H.interp(mats[q], DP_T_CROSS_SECTION, g);
    if (x > 0) {
      onset = g;
      for (int j = 0; j < n; ++j)
        if (e[j] < g) sg[j] = x;
      break;
    }
  }
  for (int j = 1; j < n; ++j) {
    const double a = e[j - 1], b = e[j], d = b - a;
    rg[j] = rg[j - 1] + d / 2 * (1 / st[j - 1] + 1 / st[j]);
    hz[j] = hz[j - 1] + d / 2 * (sg[j - 1] / st[j - 1] + sg[j] / st[j]);
    pr[j] = pr[j - 1] + d / 2 * mass * (1 / (std::sqrt(a * (a + 2 * mass)) * st[j - 1]) +
                                        1 / (std::sqrt(b * (b + 2 * mass)) * st[j]));
  }
  auto p = std::make_unique<dph::Prepared>();
  p->n = n;
  p->pow2 = dph::pow2_for(n);
  p->onset = onset;
  p->cutoff = t->cutoff;
  for (auto* v : {&e, &st, &sg, &rg, &hz, &pr}) p->host.insert(p->host.end(), v->begin(), v->end());
  int n_el = 0;
  for (size_t q = 0; q < mats.size(); ++q)
    for (const dph::Element& el : t->recipes[mats[q]]) {
      p->host.push_back(el.Z); p->host.push_back(el.A); p->host.push_back(el.I);
      p->host.push_back(fr[q] * el.f);
      ++n_el;
    }
  p->n_el = n_el;
  int rc = upload_prepared(*p);
  if (rc) return rc;
  *out = p.get();
  t->kernels[key] = std::move(p);
  return DP_OK;
}

int make_medium(dp_tables_s* t, std::vector<int> mats, std::vector<double> fr, Medium& md,
                bool with_kernel) {
  double tot = 0;
  for (double f : fr) tot += f;
  if (!(tot > 0)) return dph::fail(DP_E_ARGUMENT, "mixture fractions must sum to a positive value");
  for (double& f : fr) f /= tot;
  md.mix.n = (int)mats.size();
  for (int j = 0; j < 4; ++j) { md.mix.mat[j] = j < md.mix.n ? mats[j] : 0; md.mix.frac[j] = j < md.mix.n ? fr[j] : 0.0; }
  md.mix.e = md.mix.le = md.mix.crange = md.mix.mrange = nullptr;
  if (md.mix.n > 1) {
    dph::Prepared* p;
    int rc = prepare_mixture(t, mats, fr, &p);
    if (rc) return rc;
    const double* b = (const double*)p->dev;
    const int K = t->n_knots;
    md.mix.e = b; md.mix.le = b + K; md.mix.crange = b + 2 * K; md.mix.mrange = b + 3 * K;
  }
  md.kern = DevKernel{};
  if (!with_kernel) return DP_OK;
  dph::Prepared* k;
  int rc = prepare_kernel(t, mats, fr, &k);
  if (rc) return rc;
  const double* b = (const double*)k->dev;
  const int n = k->n;
  md.kern.n = n; md.kern.pow2 = k->pow2; md.kern.n_el = k->n_el;
  md.kern.e = b; md.kern.stop = b + n; md.kern.sigma = b + 2 * n; md.kern.range = b + 3 * n;
  md.kern.hazard = b + 4 * n; md.kern.proper = b + 5 * n; md.kern.el = b + 6 * n;
  md.kern.onset = k->onset; md.kern.cutoff = k->cutoff;
  return DP_OK;
}

}  // namespace dph

namespace {
using dph::make_medium;

__global__ void reset_bad(unsigned long long* b) { if (threadIdx.x == 0) *b = ~0ull; }

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

namespace {

int mc_run(dp_ctx c, dp_tables t, const dp_flat_geometry* geo, const dp_mc_config* cfg,
           dp_mc_result* out, double* host_events) {
  if (!c || !t || !geo || !cfg || !out) return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: null argument");
  if (cfg->n_events <= 0) return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: n_events must be positive");
  const int M = t->n_materials;
  const bool mixture = geo->rock_n_components > 0;
  if (mixture) {
    if (geo->rock_n_components > 4) return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: at most 4 rock components");
    for (int j = 0; j < geo->rock_n_components; ++j)
      if (geo->rock_components[j] < 0 || geo->rock_components[j] >= M || !(geo->rock_fractions[j] >= 0))
        return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: invalid rock mixture component");
  } else if (geo->rock_material < 0 || geo->rock_material >= M) {
    return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: material index out of range");
  }
  if (geo->air_material < 0 || geo->air_material >= M || geo->porous_material >= M)
    return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: material index out of range");
  if (!(cfg->energy_min > 0) || !(cfg->energy_max >= cfg->energy_min))
    return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: need 0 < energy_min <= energy_max");
  if (!(geo->rock_density > 0) || !(geo->primary_altitude > geo->rock_thickness))
    return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: need a positive density and primary altitude above the rock");
  double supported = t->last_energy[0];
  for (double e : t->last_energy) supported = e < supported ? e : supported;
  const double thr = cfg->energy_threshold > 0 ? cfg->energy_threshold : fmin(1e9, supported);
  // compute_flux_single, src/Geometry.jl:725-727
  if (!(thr > 0 && thr <= supported))
    return dph::fail(DP_E_ARGUMENT, "source_energy_max must be positive and within the physics energy grid");
  const bool stochastic = cfg->straggling || cfg->scattering;
  if (cfg->compute_gradient && mixture)
    return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward: the CSDA gradient surrogate needs a single rock material");
  DP_HIP(hipSetDevice(c->device));
  const int64_t n = cfg->n_events;
  const int64_t nparts = (n + RED_CHUNK - 1) / RED_CHUNK + 1;
  size_t o_w = 0, o_gf = dph::align_up(8 * (size_t)n, 256);
  size_t o_gd = dph::align_up(o_gf + 8 * (size_t)n, 256);
  size_t o_buf = dph::align_up(o_gd + 8 * (size_t)n, 256);
  size_t o_res = dph::align_up(o_buf + 8 * 4 * (size_t)nparts, 256);
  size_t o_flag = o_res + 256;
  size_t o_bad = o_flag + 256;
  size_t o_med = o_bad + 256;
  size_t o_tab = dph::align_up(o_med + 3 * sizeof(Medium), 256);
  size_t o_cnt = dph::align_up(o_tab + sizeof(DevTables), 256);
  size_t o_ord = dph::align_up(o_cnt + 4 * MC_BUCKETS, 256);
  int rc = dph::ensure(c->scratch, dph::align_up(o_ord + 4 * (size_t)n, 256));
  if (rc) return rc;
  char* b = (char*)c->scratch.ptr;
  int* flags = (int*)(b + o_flag);
  unsigned long long* bad = (unsigned long long*)(b + o_bad);
  double* out_w = (double*)(b + o_w);
  const double th = (90.0 - cfg->elevation) * M_PI / 180.0;
  const double cos_t = cfg->cos_theta != 0 || cfg->sin_theta != 0 ? cfg->cos_theta : std::cos(th);
  const double sin_t = cfg->cos_theta != 0 || cfg->sin_theta != 0 ? cfg->sin_theta : std::sin(th);
  hipLaunchKernelGGL(reset_flags, dim3(1), dim3(64), 0, c->stream, flags);
  hipLaunchKernelGGL(reset_bad, dim3(1), dim3(64), 0, c->stream, bad);
  const int rock = mixture ? geo->rock_components[0] : geo->rock_material;
  McArgs a;
  a.tab = t->data; a.K = t->n_knots; a.pow2 = t->pow2; a.mass = t->mass; a.ctau = t->ctau;
  a.g = {geo->rock_thickness, geo->rock_density, rock, geo->air_material,
         geo->porous_material, geo->porous_density, geo->porous_thickness, geo->primary_altitude};
  a.gg = {geo->rock_thickness, geo->rock_density, rock, geo->air_material, -1, 0.0, 0.0, 3e4};
  a.same_geo = !stochastic && !(geo->porous_material >= 0 && geo->porous_thickness > 0) &&
               geo->primary_altitude == 3e4;
  a.cos_t = cos_t; a.emin = cfg->energy_min; a.emax = cfg->energy_max;
  a.thr = thr; a.sw = cfg->mode_switch; a.n = n; a.seed = cfg->seed;
  a.grad = cfg->compute_gradient != 0; a.scat_limit = false;
  a.out_w = stochastic ? nullptr : out_w;
  a.out_gf = (double*)(b + o_gf); a.out_gd = (double*)(b + o_gd);
  a.err = flags;
  Medium hm[3];  // host copies of the media and tables, read by async copies before the
  DevTables ht;   // final sync
  if (stochastic) {
    StochArgs s;
    s.T = {t->data, t->scalars, t->n_knots, t->pow2, t->mass, t->ctau, t->elastic_ratio};
    ht = s.T;
    std::vector<int> rm;
    std::vector<double> rf;
    if (mixture) {
      for (int j = 0; j < geo->rock_n_components; ++j) { rm.push_back(geo->rock_components[j]); rf.push_back(geo->rock_fractions[j]); }
    } else {
      rm = {geo->rock_material}; rf = {1.0};
    }
    if ((rc = make_medium(t, rm, rf, hm[0]))) return rc;
    if ((rc = make_medium(t, {geo->air_material}, {1.0}, hm[1]))) return rc;
    s.has_porous = geo->porous_material >= 0 && geo->porous_thickness > 0;
    if ((rc = make_medium(t, {s.has_porous ? geo->porous_material : rm[0]}, {1.0}, hm[2]))) return rc;
    s.thick = geo->rock_thickness; s.rock_rho = geo->rock_density;
    s.porous_rho = geo->porous_density; s.porous_thick = geo->porous_thickness;
    s.palt = geo->primary_altitude;
    s.cos_t = cos_t; s.sin_t = sin_t; s.emin = cfg->energy_min; s.emax = cfg->energy_max;
    s.thr = thr; s.sw = cfg->mode_switch;
    s.straggling = cfg->straggling != 0; s.scattering = cfg->scattering != 0;
    s.n = n; s.seed = cfg->seed; s.out_w = out_w; s.err = flags; s.bad_event = bad;
    DP_HIP(hipMemcpyAsync(b + o_med, hm, sizeof(hm), hipMemcpyHostToDevice, c->stream));
    DP_HIP(hipMemcpyAsync(b + o_tab, &ht, sizeof(ht), hipMemcpyHostToDevice, c->stream));
    s.media = (const Medium*)(b + o_med);
    s.Tg = (const DevTables*)(b + o_tab);
    dph::record(c, c->ev0);
    s.order = nullptr;
    if (n < 0x7fffffff) {
      int* cnt = (int*)(b + o_cnt);
      uint32_t* ord = (uint32_t*)(b + o_ord);
      const unsigned gb = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)c->n_cu * 8);
      hipLaunchKernelGGL(bucket_clear, dim3(1), dim3(256), 0, c->stream, cnt);
      hipLaunchKernelGGL(bucket_count, dim3(gb), dim3(256), 0, c->stream, cfg->seed, n, cnt);
      hipLaunchKernelGGL(bucket_offsets, dim3(1), dim3(64), 0, c->stream, cnt);
      hipLaunchKernelGGL(bucket_scatter, dim3(gb), dim3(256), 0, c->stream, cfg->seed, n, cnt, ord);
      DP_HIP(hipGetLastError());
      s.order = ord;
    }
    const bool fast = t->shared_grid;
    int occ = 0;
    if (fast) DP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, mc_flat_stochastic_kernel<true>, MC_THREADS, 0));
    else DP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, mc_flat_stochastic_kernel<false>, MC_THREADS, 0));
    if (occ < 1) occ = 1;
    // Blocks take 64-event rounds grid-stride; the tuned knob k0 is the
    // number of blocks per resident block slot, capped at one block per round
    // (default 64, which at 1e7 events is that cap: every block takes one
    // round and the hardware hands out blocks as slots free up; measured at
    // 1e7 events 858 ms, against 860 ms at 32, 866 ms at 16 and 878 ms at 4).
    const int64_t want = (n + MC_THREADS - 1) / MC_THREADS;
    auto launch = [&](const dph::TuneCfg& cfg) -> int {
      int64_t grid = (int64_t)c->n_cu * occ * std::max(1, cfg.k0);
      if (grid > want) grid = want;
      if (fast) hipLaunchKernelGGL(mc_flat_stochastic_kernel<true>, dim3((unsigned)grid), dim3(MC_THREADS), 0, c->stream, s);
      else hipLaunchKernelGGL(mc_flat_stochastic_kernel<false>, dim3((unsigned)grid), dim3(MC_THREADS), 0, c->stream, s);
      DP_HIP(hipGetLastError());
      return DP_OK;
    };
    static const std::vector<dph::TuneCfg> cands = {{64, 0, 0}, {32, 0, 0}, {16, 0, 0}, {8, 0, 0}, {4, 0, 0}};
    const std::string key = std::string("mc_flat_stochastic fast=") + (fast ? "1" : "0") +
                            " events=" + std::to_string(n);
    dph::TuneCfg cfg;
    if ((rc = dph::tuned(c, key, cands, launch, &cfg))) return rc;
    if ((rc = launch(cfg))) return rc;
    dph::record(c, c->ev1);
  }
  if (!stochastic || a.grad) {
    size_t lds = (size_t)3 * MC_ARRS * t->n_knots * sizeof(double);
    int occ = 0;
    DP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, mc_flat_csda_kernel, 256, lds));
    if (occ < 1) occ = 1;
    int64_t want = (n + 255) / 256;
    int64_t grid = (int64_t)c->n_cu * occ * 4;
    if (grid > want) grid = want;
    if (!stochastic) dph::record(c, c->ev0);
    hipLaunchKernelGGL(mc_flat_csda_kernel, dim3((unsigned)grid), dim3(256), lds, c->stream, a);
    if (!stochastic) dph::record(c, c->ev1);
    DP_HIP(hipGetLastError());
  }
  double* res = (double*)(b + o_res);
  double* buf = (double*)(b + o_buf);
  rc = reduce_pair(c->stream, out_w, nullptr, n, true, buf, res);
  if (rc) return rc;
  if (a.grad) {
    rc = reduce_pair(c->stream, a.out_gf, a.out_gd, n, false, buf, res + 2);
    if (rc) return rc;
  }
  double h[4] = {0, 0, 0, 0};
  int hf[4] = {0, 0, 0, 0};
  unsigned long long hb = 0;
  DP_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipMemcpyAsync(&hb, bad, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
  if (host_events) DP_HIP(hipMemcpyAsync(host_events, out_w, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipStreamSynchronize(c->stream));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  if (hf[0] & ERR_DOMAIN)
    return dph::fail(DP_E_DOMAIN, "energy outside a prepared mixed kernel or table (first failing event " + std::to_string(hb) + ")");
  if (hf[0] & (ERR_TRANSPORT | ERR_CAP))
    return dph::fail(DP_E_EXHAUSTED, "transport cap or invalid hard-event weight (first failing event " + std::to_string(hb) + ")");
  const double dn = (double)n;
  out->flux = h[0] / dn;
  const double var = h[1] / dn - out->flux * out->flux;
  out->sigma = geo->rock_thickness > 0 ? sqrt(fmax(var, 0.0) / dn) : 0.0;
  out->gradient_flux = a.grad ? h[2] / dn : NAN;
  out->gradient = a.grad ? h[3] / dn : NAN;
  out->source_energy_max = thr;
  out->n_events = n;
  return DP_OK;
}

}  // namespace

extern "C" {

int dp_mc_flat_backward(dp_ctx c, dp_tables t, const dp_flat_geometry* geo,
                        const dp_mc_config* cfg, dp_mc_result* out) {
  return mc_run(c, t, geo, cfg, out, nullptr);
}

int dp_mc_flat_backward_events(dp_ctx c, dp_tables t, const dp_flat_geometry* geo,
                               const dp_mc_config* cfg, dp_mc_result* out, double* ev) {
  if (!ev) return dph::fail(DP_E_ARGUMENT, "dp_mc_flat_backward_events: null event buffer");
  return mc_run(c, t, geo, cfg, out, ev);
}

}  // extern "C"

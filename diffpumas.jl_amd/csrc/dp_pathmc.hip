// Fixed-path (tomography) backward Monte Carlo: compute_directional_flux_mc
// for many paths in one persistent launch.
//
// Work items are (path, chunk of 64 samples) in path-major order, taken from
// an atomic queue, so the lanes of a block walk the same segment list (cell,
// length, medium and density loads are wave-uniform) and paths of any length
// balance across the chip.  One event per lane; Philox keyed by
// (seed, path * n_samples + sample).  Hard-event weights of an item go to the
// block's pool and are evaluated by the block's 64 lanes once the item's events are
// done.  Each item writes (sum w, sum w^2) of its chunk through a fixed
// wave/LDS tree; a second kernel adds a path's chunks in order, so flux and
// sigma do not depend on the launch shape.
#include "dp_internal.hpp"
#include "dp_device.hpp"
#include "dp_transport.hpp"
#include "dp_transport_fast.hpp"
#include "dp_media.hpp"
#include <algorithm>
#include <cmath>

using namespace dpk;

namespace {

// One-wave blocks: an item waits only for the slowest of its own 64 events
// (measured at 1024 paths x 1e4 events: 674 ms, against 700 ms with 128-lane
// blocks, 128-sample items and a 384-record pool).
constexpr int PM_THREADS = 64;
constexpr int PM_POOL = 192;  // pooled hard-event records per block

// The out-of-line step functions take the tables by reference from Tg, a copy
// in device memory: a reference into the kernel arguments would make the
// compiler copy them to every lane's stack.
struct PathArgs {
  DevTables T;
  const DevTables* Tg;  // T in device memory
  const Medium* media;  // n_media cell media, then rock, then air
  const int32_t* cell_medium;
  const double* cell_density;
  int rock_medium, air_medium;
  double rho_rock, det_elev, palt, thr_low, elimit;
  bool straggling;
  int n_rays;
  const int32_t* row_ptr;
  const int32_t* cell;
  const double* len;
  const double* cos_t;
  const double* remaining;
  const double* exit_z;
  const uint8_t* valid;
  int n_samples;
  const double* se;
  const double* sw;
  const double* sq;
  const uint32_t* order;  // samples in order of energy
  uint64_t seed;
  int chunks;          // chunks per path
  int r0;              // first path of this launch
  int64_t n_items;     // paths of this launch * chunks
  double* part;        // 2 per item (indexed from path r0)
  double* events;      // optional, n_rays * n_samples
  double* w_local;     // per event of this launch (with deferred DEL weights)
  int* queue;
  int* err;
  unsigned long long* bad;
  // Second sample set of the same launch: deterministic CSDA nodes (charge
  // -summed allowed), items after the MC items.  n_nodes = 0: none.
  int n_nodes;
  const double* ne;
  const double* nw;
  const double* nq;
  const uint32_t* order_n;
  int chunks_n;
  int64_t n_items_mc;  // items [0, n_items_mc) of this launch are MC chunks
  double* part_n;      // 2 per CSDA item (indexed from path r0)
  double* w_local_n;   // per CSDA node of this launch
};

// The sample set an item of this launch draws from.
struct ItemSet {
  bool csda;           // deterministic CSDA steps (no straggling, no draws)
  int n, chunks;
  const double *e, *w, *q;
  const uint32_t* order;
  double* w_local;
  int64_t it;          // item index within its set
};
__device__ __forceinline__ ItemSet item_set(const PathArgs& a, int64_t item) {
  ItemSet S;
  S.csda = item >= a.n_items_mc;
  if (S.csda) {
    S.n = a.n_nodes; S.chunks = a.chunks_n; S.e = a.ne; S.w = a.nw; S.q = a.nq;
    S.order = a.order_n; S.w_local = a.w_local_n; S.it = item - a.n_items_mc;
  } else {
    S.n = a.n_samples; S.chunks = a.chunks; S.e = a.se; S.w = a.sw; S.q = a.sq;
    S.order = a.order; S.w_local = a.w_local; S.it = item;
  }
  return S;
}

// propagate_mc_uniform_segment / propagate_mc_air_segment, src/Tomography.jl:
// 876-969 (scattering off).  _mc_segment_step (:873-874) bounds the step by
// the remaining distance.  A step without spatial progress below the energy
// limit is the reference's ErrorException (ERR_TRANSPORT).
template <bool FAST>
__device__ void path_segment(const PathArgs& a, const Medium& md, McState& st, double density,
                             bool air, double cos_t, double distance, Philox& rng, int& err,
                             int& g, DelLog* dlog, bool straggling) {
  double remaining = distance;
  const double elimit = a.elimit;
  const DevTables& Tc = *a.Tg;  // for the out-of-line steps
  int it = 0;
  while (remaining > 1e-8 && st.E < elimit - EPS64) {
    if (++it > 1000000) { err |= ERR_CAP; return; }
    if (st.w <= 0.0 || !isfinite(st.w)) break;
    const double cur = st.E < a.thr_low - EPS64 ? fmin(a.thr_low, elimit) : elimit;
    Loc L;
    double Xi;
    if (FAST) {
      L = loc_e(a.T, st.E, log(st.E), g);
      Xi = mix_range_L(a.T, md.mix, false, L);
      g = L.i;
    } else {
      Xi = mix_range(a.T, md.mix, false, st.E, log(st.E));
    }
    double dens = density, max_step;
    if (air) {
      dens = 1.205 * exp(-(a.det_elev + st.pos[2]) / 12e3);
      max_step = 0.01 * Xi / fmax(dens, 1e-8);
      max_step = fmin(max_step, 0.05 * 12e3 / fmax(cos_t, 0.05));
    } else {
      max_step = 0.01 * Xi / fmax(dens, 1e-6);
    }
    const double step = fmin(remaining, fmax(1e-6, fmin(remaining + 1e-7, max_step)));
    const double oldL = st.L, oldE = st.E;
    if (FAST) {
      if (st.E < a.thr_low - EPS64 && straggling)
        straggled_step_f(Tc, md, st, L, 0.0, dens, step, false, cur, rng, err, &g, dlog);
      else if (straggling)
        mixed_step_f(Tc, md, st, L, dens, step, cur, rng, err, dlog);
      else
        csda_mix_step_f(a.T, md, st, L, dens, step, &g);
    } else if (st.E < a.thr_low - EPS64) {
      if (straggling) straggled_step(Tc, md, st, dens, step, false, cur, rng, err);
      else csda_mix_step(a.T, md, st, dens, step);
    } else {
      if (straggling) mixed_step(Tc, md, st, dens, step, cur, rng, err);
      else csda_mix_step(a.T, md, st, dens, step);
    }
    if (err) return;
    const double actual = st.L - oldL;
    if (st.w <= 0.0 || !isfinite(st.w)) break;
    if (actual <= 0 && oldE < cur && cur < elimit && st.E >= cur) continue;
    if (!(actual > 0)) {
      if (st.E >= elimit) break;
      err |= ERR_TRANSPORT;
      return;
    }
    remaining = fmax(0.0, remaining - actual);
  }
}

template <bool FAST>
__global__ __launch_bounds__(PM_THREADS, 4) void path_mc_kernel(PathArgs a) {
  __shared__ int64_t s_item;
  __shared__ PoolRec pool[FAST ? PM_POOL : 1];
  __shared__ int pool_n;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) {
      s_item = atomicAdd((unsigned long long*)a.queue, 1ull);
      pool_n = 0;
    }
    __syncthreads();
    const int64_t item = s_item;
    if (item >= a.n_items) break;
    const ItemSet S = item_set(a, item);
    const bool strag = a.straggling && !S.csda;
    const int r = a.r0 + (int)(S.it / S.chunks);
    const int sl = (int)(S.it % S.chunks) * PM_THREADS + tid;
    const double ct = a.cos_t[r];
    const bool live = (a.valid ? a.valid[r] != 0 : true) && ct > 0.0;
    const bool ran = live && sl < S.n;
    if (!live && sl < S.n) {
      const int64_t el = (int64_t)(r - a.r0) * S.n + (int)S.order[sl];
      S.w_local[el] = 0.0;
    }
    double wi = 0.0;
    int err = 0;
    int64_t e = 0, el = 0;
    DelLog dlg{pool, &pool_n, PM_POOL, tid, 0, 0, nullptr, a.media, a.air_medium + 1};
    if (ran) {
      const int s = (int)S.order[sl];
      e = (int64_t)r * S.n + s;
      el = (int64_t)(r - a.r0) * S.n + s;  // sample index in this launch's set
      Philox rng;
      rng.init(a.seed, (uint64_t)e);
      McState st;
      st.E = S.e[s]; st.w = 1.0; st.t = 0.0; st.L = 0.0;
      st.pos[0] = st.pos[1] = st.pos[2] = 0.0;
      st.dir[0] = -sqrt(fmax(0.0, 1.0 - ct * ct)); st.dir[1] = 0.0; st.dir[2] = -ct;
      int g = a.T.K / 2;
      DelLog* dl = (FAST && strag) ? &dlg : nullptr;
      const int b0 = a.row_ptr[r], ns = a.row_ptr[r + 1] - b0;
      for (int k = 0; k < ns; ++k) {
        const int c = a.cell[b0 + k];
        dlg.medium = a.cell_medium[c];
        path_segment<FAST>(a, a.media[dlg.medium], st, a.cell_density[c], false, ct,
                           a.len[b0 + k], rng, err, g, dl, strag);
        if (err || st.w <= 0.0 || !isfinite(st.w)) break;
      }
      if (!err && st.w > 0.0 && isfinite(st.w) && a.remaining[r] > 1e-8) {
        dlg.medium = a.rock_medium;
        path_segment<FAST>(a, a.media[a.rock_medium], st, a.rho_rock, false, ct, a.remaining[r],
                           rng, err, g, dl, strag);
      }
      if (!err && st.w > 0.0 && isfinite(st.w)) {
        const double air_d = fmax(0.0, a.palt - a.exit_z[r]) / fmax(ct, 1e-3);
        dlg.medium = a.air_medium;
        path_segment<FAST>(a, a.media[a.air_medium], st, 0.0, true, ct, air_d, rng, err, g, dl,
                           strag);
      }
      double f = 0.0;
      if (!err && st.w > 0.0 && isfinite(st.w) && st.E < a.elimit)
        f = st.w * flux_gccly(ct, st.E, S.q[s]);
      wi = err ? 0.0 : 2.0 * S.w[s] * f;
    }
    int m = 0;
    if (FAST) {
      __syncthreads();  // every record of the item is in the pool
      m = min(pool_n, PM_POOL);
      pool_eval(pool, m, a.T.mass, [&](int md) -> const DevKernel& { return a.media[md].kern; });
      __syncthreads();  // every factor is in the pool
    }
    if (ran) {
      int fail = err;
      if (FAST && dlg.n > 0) wi = pool_apply(pool, m, tid, dlg.n, wi, err, fail);
      if (fail) {
        atomicOr(a.err, fail);
        atomicMin(a.bad, (unsigned long long)e);
        wi = 0.0;
      }
      S.w_local[el] = wi;
    }
    __syncthreads();  // every thread has read s_item and the pool before thread 0 resets them
  }
}

// (sum w, sum w^2) of each item's samples,
// in the sorted sample order of the item (fixed tree).
__global__ __launch_bounds__(PM_THREADS) void path_items_kernel(PathArgs a) {
  __shared__ double s_sum[2][PM_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int64_t item = blockIdx.x; item < a.n_items; item += gridDim.x) {
    const ItemSet S = item_set(a, item);
    const int r = a.r0 + (int)(S.it / S.chunks);
    const int sl = (int)(S.it % S.chunks) * PM_THREADS + tid;
    double wi = 0.0;
    if (sl < S.n) {
      const int s = (int)S.order[sl];
      const int64_t el = (int64_t)(r - a.r0) * S.n + s;
      wi = S.w_local[el];
      if (a.events && !S.csda) a.events[(int64_t)r * S.n + s] = wi;
    }
    const double s1 = wave_sum(wi), s2 = wave_sum(wi * wi);
    if (lane == 0) { s_sum[0][wave] = s1; s_sum[1][wave] = s2; }
    __syncthreads();
    if (tid == 0) {
      double t1 = 0.0, t2 = 0.0;
      for (int j = 0; j < PM_THREADS / 64; ++j) { t1 += s_sum[0][j]; t2 += s_sum[1][j]; }
      const int64_t gi = (int64_t)a.r0 * S.chunks + S.it;
      double* part = S.csda ? a.part_n : a.part;
      part[2 * gi] = t1;
      part[2 * gi + 1] = t2;
    }
    __syncthreads();
  }
}

__global__ void path_reduce(const double* __restrict__ part, int n_rays, int chunks,
                            int n_samples, const double* __restrict__ cos_t,
                            const uint8_t* __restrict__ valid, double* __restrict__ flux,
                            double* __restrict__ sigma) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rays) return;
  if ((valid && !valid[r]) || !(cos_t[r] > 0.0)) {
    flux[r] = sigma[r] = __builtin_nan("");
    return;
  }
  double s1 = 0.0, s2 = 0.0;
  for (int c = 0; c < chunks; ++c) {
    s1 += part[2 * ((int64_t)r * chunks + c)];
    s2 += part[2 * ((int64_t)r * chunks + c) + 1];
  }
  const double n = n_samples > 1 ? (double)n_samples : 1.0;
  const double f = s1 / n;
  flux[r] = f;
  sigma[r] = sqrt(fmax(0.0, (s2 / n - f * f) / fmax(1.0, n - 1.0)));
}

__global__ void reset_path_flags(int* q, unsigned long long* bad) {
  if (threadIdx.x < 4) q[threadIdx.x] = 0;
  if (threadIdx.x == 0) *bad = ~0ull;
}
__global__ void reset_queue(int* q) {
  if (threadIdx.x < 2) q[threadIdx.x] = 0;
}

}  // namespace

namespace {

// The fixed-path MC of every path for the samples `es` and, when `nodes` is
// given, the deterministic CSDA flux of every path over those nodes, in one
// persistent launch per batch of paths (MC items first, then CSDA items).
int run_paths(const std::string& fn, dp_ctx c, dp_tables t, const dp_material_config* mat,
              const dp_site_config* site, const dp_directional_paths* p, int32_t n_cells,
              int32_t n_media, const dp_mixture* media, const int32_t* cell_medium,
              const double* cell_density, const dp_energy_samples* es,
              const dp_energy_samples* nodes, uint64_t seed, int32_t straggling, double thr_low,
              double* flux, double* sigma, double* event_weights, double* csda_flux) {
  if (!c || !t || !mat || !site || !p || !media || !cell_medium || !cell_density || !es || !flux ||
      !sigma || (nodes && !csda_flux))
    return dph::fail(DP_E_ARGUMENT, fn + ": null argument");
  const int M = t->n_materials, nr = p->n_rays, ns = p->n_segments, nn = es->n;
  const int nq = nodes ? nodes->n : 0;
  if (nr < 0 || ns < 0 || n_cells <= 0 || n_media <= 0 || nn <= 0 || (nodes && nq <= 0))
    return dph::fail(DP_E_ARGUMENT, fn + ": sizes must be positive");
  for (int i = 0; i < nq; ++i)
    if (!(nodes->energy[i] > 0)) return dph::fail(DP_E_ARGUMENT, fn + ": node energies must be positive");
  if (mat->rock_material < 0 || mat->rock_material >= M || mat->air_material < 0 ||
      mat->air_material >= M || !(mat->rock_density > 0))
    return dph::fail(DP_E_ARGUMENT, fn + ": invalid rock/air material or rock density");
  for (int i = 0; i < nn; ++i) {
    if (es->charge[i] != 1.0 && es->charge[i] != -1.0)
      return dph::fail(DP_E_ARGUMENT,
                       "MC requires physical charges +-1; CSDA quadrature nodes are not MC samples");
    if (!(es->energy[i] > 0)) return dph::fail(DP_E_ARGUMENT, fn + ": sample energies must be positive");
  }
  for (int j = 0; j < n_media; ++j) {
    const dp_mixture& m = media[j];
    if (m.n_components < 1 || m.n_components > 4)
      return dph::fail(DP_E_ARGUMENT, fn + ": a medium needs 1 to 4 components");
    for (int k = 0; k < m.n_components; ++k)
      if (m.components[k] < 0 || m.components[k] >= M || !(m.fractions[k] >= 0))
        return dph::fail(DP_E_ARGUMENT, fn + ": invalid medium component");
  }
  for (int i = 0; i < n_cells; ++i)
    if (cell_medium[i] < 0 || cell_medium[i] >= n_media || !(cell_density[i] > 0))
      return dph::fail(DP_E_ARGUMENT, fn + ": invalid cell medium or density at cell " + std::to_string(i));
  if (nr == 0) return DP_OK;
  if (p->row_ptr[0] != 0 || p->row_ptr[nr] != ns)
    return dph::fail(DP_E_ARGUMENT, fn + ": row_ptr must start at 0 and end at n_segments");
  for (int r = 0; r < nr; ++r)
    if (p->row_ptr[r + 1] < p->row_ptr[r]) return dph::fail(DP_E_ARGUMENT, fn + ": row_ptr must be non-decreasing");
  for (int s = 0; s < ns; ++s)
    if (p->cell[s] < 0 || p->cell[s] >= n_cells)
      return dph::fail(DP_E_ARGUMENT, fn + ": cell index out of range at segment " + std::to_string(s));
  DP_HIP(hipSetDevice(c->device));
  // media on the device: cell media, then the dry-rock and air materials
  std::vector<Medium> hm(n_media + 2);
  const bool kern = straggling != 0;
  int rc;
  for (int j = 0; j < n_media; ++j) {
    std::vector<int> mm(media[j].components, media[j].components + media[j].n_components);
    std::vector<double> ff(media[j].fractions, media[j].fractions + media[j].n_components);
    if ((rc = dph::make_medium(t, mm, ff, hm[j], kern))) return rc;
  }
  if ((rc = dph::make_medium(t, {mat->rock_material}, {1.0}, hm[n_media], kern))) return rc;
  if ((rc = dph::make_medium(t, {mat->air_material}, {1.0}, hm[n_media + 1], kern))) return rc;
  const int chunks = (nn + PM_THREADS - 1) / PM_THREADS;
  const int chunks_n = (nq + PM_THREADS - 1) / PM_THREADS;
  const int64_t n_items = (int64_t)nr * chunks;
  const int64_t n_ev = (int64_t)nr * nn;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = dph::align_up(o + bytes, 256); return at; };
  const size_t o_med = take(sizeof(Medium) * hm.size()), o_tab = take(sizeof(DevTables)),
               o_cm = take(4 * (size_t)n_cells),
               o_cd = take(8 * (size_t)n_cells), o_rp = take(4 * (size_t)(nr + 1)),
               o_cell = take(4 * (size_t)ns), o_len = take(8 * (size_t)ns),
               o_cos = take(8 * (size_t)nr), o_rem = take(8 * (size_t)nr),
               o_zex = take(8 * (size_t)nr), o_val = take((size_t)nr),
               o_se = take(8 * (size_t)nn), o_sw = take(8 * (size_t)nn), o_sq = take(8 * (size_t)nn),
               o_ord = take(4 * (size_t)nn),
               o_ne = take(8 * (size_t)nq), o_nw = take(8 * (size_t)nq), o_nq = take(8 * (size_t)nq),
               o_ordn = take(4 * (size_t)nq), o_partn = take(16 * (size_t)nr * chunks_n),
               o_fluxn = take(8 * (size_t)nr), o_sign = take(8 * (size_t)nr),
               o_part = take(16 * (size_t)n_items), o_flux = take(8 * (size_t)nr),
               o_sig = take(8 * (size_t)nr), o_ev = take(event_weights ? 8 * (size_t)n_ev : 8),
               o_q = take(256);
  rc = dph::ensure(c->staging, o);
  if (rc) return rc;
  char* b = (char*)c->staging.ptr;
  hipStream_t st = c->stream;
  DP_HIP(hipMemcpyAsync(b + o_med, hm.data(), sizeof(Medium) * hm.size(), hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_cm, cell_medium, 4 * (size_t)n_cells, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_cd, cell_density, 8 * (size_t)n_cells, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_rp, p->row_ptr, 4 * (size_t)(nr + 1), hipMemcpyHostToDevice, st));
  if (ns) {
    DP_HIP(hipMemcpyAsync(b + o_cell, p->cell, 4 * (size_t)ns, hipMemcpyHostToDevice, st));
    DP_HIP(hipMemcpyAsync(b + o_len, p->length, 8 * (size_t)ns, hipMemcpyHostToDevice, st));
  }
  DP_HIP(hipMemcpyAsync(b + o_cos, p->cos_theta, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_rem, p->remaining_rock, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_zex, p->exit_z, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  if (p->valid) DP_HIP(hipMemcpyAsync(b + o_val, p->valid, (size_t)nr, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_se, es->energy, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_sw, es->weight, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_sq, es->charge, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  // samples in order of energy: the lanes of a chunk share the transport mode
  std::vector<uint32_t> ord(nn);
  for (int i = 0; i < nn; ++i) ord[i] = (uint32_t)i;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](uint32_t x, uint32_t y) { return es->energy[x] < es->energy[y]; });
  DP_HIP(hipMemcpyAsync(b + o_ord, ord.data(), 4 * (size_t)nn, hipMemcpyHostToDevice, st));
  std::vector<uint32_t> ordn(nq);
  if (nq) {
    DP_HIP(hipMemcpyAsync(b + o_ne, nodes->energy, 8 * (size_t)nq, hipMemcpyHostToDevice, st));
    DP_HIP(hipMemcpyAsync(b + o_nw, nodes->weight, 8 * (size_t)nq, hipMemcpyHostToDevice, st));
    DP_HIP(hipMemcpyAsync(b + o_nq, nodes->charge, 8 * (size_t)nq, hipMemcpyHostToDevice, st));
    for (int i = 0; i < nq; ++i) ordn[i] = (uint32_t)i;
    std::stable_sort(ordn.begin(), ordn.end(),
                     [&](uint32_t x, uint32_t y) { return nodes->energy[x] < nodes->energy[y]; });
    DP_HIP(hipMemcpyAsync(b + o_ordn, ordn.data(), 4 * (size_t)nq, hipMemcpyHostToDevice, st));
  }
  double supported = t->last_energy[0];
  for (double e : t->last_energy) supported = e < supported ? e : supported;
  PathArgs a;
  a.T = {t->data, t->scalars, t->n_knots, t->pow2, t->mass, t->ctau, t->elastic_ratio};
  DP_HIP(hipMemcpyAsync(b + o_tab, &a.T, sizeof(DevTables), hipMemcpyHostToDevice, st));
  a.Tg = (const DevTables*)(b + o_tab);
  a.media = (const Medium*)(b + o_med);
  a.cell_medium = (const int32_t*)(b + o_cm);
  a.cell_density = (const double*)(b + o_cd);
  a.rock_medium = n_media; a.air_medium = n_media + 1;
  a.rho_rock = mat->rock_density; a.det_elev = site->detector_elevation;
  a.palt = site->primary_altitude; a.thr_low = thr_low; a.elimit = fmin(1e9, supported);
  a.straggling = straggling != 0;
  a.n_rays = nr;
  a.row_ptr = (const int32_t*)(b + o_rp); a.cell = (const int32_t*)(b + o_cell);
  a.len = (const double*)(b + o_len); a.cos_t = (const double*)(b + o_cos);
  a.remaining = (const double*)(b + o_rem); a.exit_z = (const double*)(b + o_zex);
  a.valid = p->valid ? (const uint8_t*)(b + o_val) : nullptr;
  a.n_samples = nn; a.se = (const double*)(b + o_se); a.sw = (const double*)(b + o_sw);
  a.sq = (const double*)(b + o_sq);
  a.order = (const uint32_t*)(b + o_ord);
  a.seed = seed; a.chunks = chunks; a.n_items = n_items; a.r0 = 0;
  a.n_nodes = nq; a.chunks_n = chunks_n;
  a.ne = (const double*)(b + o_ne); a.nw = (const double*)(b + o_nw); a.nq = (const double*)(b + o_nq);
  a.order_n = (const uint32_t*)(b + o_ordn);
  a.part_n = (double*)(b + o_partn);
  a.part = (double*)(b + o_part);
  a.events = event_weights ? (double*)(b + o_ev) : nullptr;
  int* q = (int*)(b + o_q);
  a.queue = q; a.err = q + 2;
  a.bad = (unsigned long long*)(q + 4);
  hipLaunchKernelGGL(reset_path_flags, dim3(1), dim3(64), 0, st, q, a.bad);
  const bool fast = t->shared_grid;
  int occ = 0;
  if (fast) DP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, path_mc_kernel<true>, PM_THREADS, 0));
  else DP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, path_mc_kernel<false>, PM_THREADS, 0));
  if (occ < 1) occ = 1;
  // Paths are processed in batches of at most 2^22 events, which bounds the
  // per-event weight buffer.
  const int64_t batch_ev = (int64_t)1 << 22;
  int rays_per_batch = (int)std::max<int64_t>(1, batch_ev / (nn + nq));
  if (rays_per_batch > nr) rays_per_batch = nr;
  const int64_t ev_batch = (int64_t)rays_per_batch * nn;
  const size_t o_wn = dph::align_up(8 * (size_t)ev_batch, 256);
  rc = dph::ensure(c->scratch, o_wn + 8 * (size_t)rays_per_batch * nq + 256);
  if (rc) return rc;
  a.w_local = (double*)c->scratch.ptr;
  a.w_local_n = (double*)((char*)c->scratch.ptr + o_wn);
  // One batch of paths: the persistent queue kernel with k0 blocks per
  // resident block slot (the tuned knob, default 2: 277 ms against 282 ms at
  // 1 on 4096 paths x 1e4 events), then the per-item sums.
  auto batch = [&](int r0, const dph::TuneCfg& cfg) -> int {
    const int r1 = std::min(nr, r0 + rays_per_batch);
    a.r0 = r0;
    a.n_items_mc = (int64_t)(r1 - r0) * chunks;
    a.n_items = a.n_items_mc + (int64_t)(r1 - r0) * chunks_n;
    int64_t grid = (int64_t)c->n_cu * occ * std::max(1, cfg.k0);
    if (grid > a.n_items) grid = a.n_items;
    hipLaunchKernelGGL(reset_queue, dim3(1), dim3(64), 0, st, q);
    if (fast) hipLaunchKernelGGL(path_mc_kernel<true>, dim3((unsigned)grid), dim3(PM_THREADS), 0, st, a);
    else hipLaunchKernelGGL(path_mc_kernel<false>, dim3((unsigned)grid), dim3(PM_THREADS), 0, st, a);
    DP_HIP(hipGetLastError());
    int64_t gi = a.n_items < (int64_t)c->n_cu * 8 ? a.n_items : (int64_t)c->n_cu * 8;
    hipLaunchKernelGGL(path_items_kernel, dim3((unsigned)gi), dim3(PM_THREADS), 0, st, a);
    DP_HIP(hipGetLastError());
    return DP_OK;
  };
  // Tuned on the first batch (every batch has the same shape but the last).
  static const std::vector<dph::TuneCfg> cands = {{2, 0, 0}, {1, 0, 0}, {4, 0, 0}};
  const std::string key = std::string("path_mc fast=") + (fast ? "1" : "0") + " rays=" +
                          std::to_string(nr) + " events=" + std::to_string(nn) + " nodes=" +
                          std::to_string(nq) + " segments=" + std::to_string(p->n_segments);
  dph::TuneCfg cfg;
  rc = dph::tuned(c, key, cands, [&](const dph::TuneCfg& k) { return batch(0, k); }, &cfg);
  if (rc) return rc;
  dph::record(c, c->ev0);
  for (int r0 = 0; r0 < nr; r0 += rays_per_batch)
    if ((rc = batch(r0, cfg))) return rc;
  dph::record(c, c->ev1);
  hipLaunchKernelGGL(path_reduce, dim3((nr + 255) / 256), dim3(256), 0, st, a.part, nr, chunks,
                     nn, a.cos_t, a.valid, (double*)(b + o_flux), (double*)(b + o_sig));
  if (nq)
    hipLaunchKernelGGL(path_reduce, dim3((nr + 255) / 256), dim3(256), 0, st, a.part_n, nr,
                       chunks_n, nq, a.cos_t, a.valid, (double*)(b + o_fluxn),
                       (double*)(b + o_sign));
  DP_HIP(hipGetLastError());
  int herr[2] = {0, 0};
  unsigned long long hb = 0;
  DP_HIP(hipMemcpyAsync(flux, b + o_flux, 8 * (size_t)nr, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(sigma, b + o_sig, 8 * (size_t)nr, hipMemcpyDeviceToHost, st));
  if (nq) DP_HIP(hipMemcpyAsync(csda_flux, b + o_fluxn, 8 * (size_t)nr, hipMemcpyDeviceToHost, st));
  if (event_weights)
    DP_HIP(hipMemcpyAsync(event_weights, b + o_ev, 8 * (size_t)n_ev, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(herr, a.err, 8, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(&hb, a.bad, 8, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  if (herr[0] & ERR_DOMAIN)
    return dph::fail(DP_E_DOMAIN, fn + ": energy outside a prepared mixed kernel or table (first failing event " +
                                      std::to_string(hb) + ")");
  if (herr[0] & (ERR_TRANSPORT | ERR_CAP))
    return dph::fail(DP_E_EXHAUSTED, fn + ": MC segment made no spatial progress or hit a cap (first failing event " +
                                         std::to_string(hb) + ")");
  return DP_OK;
}

}  // namespace

extern "C" {

int dp_mc_fixed_paths(dp_ctx c, dp_tables t, const dp_material_config* mat,
                      const dp_site_config* site, const dp_directional_paths* p, int32_t n_cells,
                      int32_t n_media, const dp_mixture* media, const int32_t* cell_medium,
                      const double* cell_density, const dp_energy_samples* es, uint64_t seed,
                      int32_t straggling, double thr_low, double* flux, double* sigma,
                      double* event_weights) {
  return run_paths("dp_mc_fixed_paths", c, t, mat, site, p, n_cells, n_media, media, cell_medium,
                   cell_density, es, nullptr, seed, straggling, thr_low, flux, sigma,
                   event_weights, nullptr);
}

int dp_mc_fixed_paths_csda(dp_ctx c, dp_tables t, const dp_material_config* mat,
                           const dp_site_config* site, const dp_directional_paths* p,
                           int32_t n_cells, int32_t n_media, const dp_mixture* media,
                           const int32_t* cell_medium, const double* cell_density,
                           const dp_energy_samples* samples, const dp_energy_samples* nodes,
                           uint64_t seed, double thr_low, double* flux, double* sigma,
                           double* csda_flux) {
  if (!nodes) return dph::fail(DP_E_ARGUMENT, "dp_mc_fixed_paths_csda: null nodes");
  return run_paths("dp_mc_fixed_paths_csda", c, t, mat, site, p, n_cells, n_media, media,
                   cell_medium, cell_density, samples, nodes, seed, 1, thr_low, flux, sigma,
                   nullptr, csda_flux);
}

}  // extern "C"

// Hard-event (DEL) weights pooled per block and evaluated by all its lanes.
//
// reverse_del_event (src/MixedTransport.jl:111-142) draws the ancestor energy
// from a nu^-2 proposal and weights it by the differential kernel summed over
// every element and process of the medium (4 DCS models per element, about
// 2e4 FP64 operations per element).  Evaluated where it occurs, that sum runs
// on one lane while the rest of the wave waits.  Where the weight only
// multiplies the event weight - the mixed step, and the straggled step
// without scattering, where the element/process choice only consumes its
// draw - the transport keeps going with the ancestor energy and records
// (ancestor, q, denominator, medium) in the block's LDS pool.  When every lane
// of the block has finished its event, each lane evaluates pooled records
// (one record per lane and round), and each lane then multiplies its own
// factors in the order it recorded them.  The draws consumed, the energies
// and the factors are those of the inline evaluation; only the product order
// of one event's weight changes (rounding).  A record that finds the pool
// full is evaluated inline.
//
// A factor of 0 kills the event (weight 0), as inline.  An error raised by
// the transport after some records is void when one of those records killed
// the event; otherwise it stands.
#pragma once
#include "dp_transport.hpp"

namespace dpk {

struct PoolRec {
  double anc, q;
  double denom;  // replaced by the factor once evaluated
  int owner;     // lane (thread index in the block) that recorded it
  int mk;        // medium << 1 | kind; kind 0: mixed step (non-finite weight is an
                 // error), 1: straggled step (non-finite weight kills the event)
};

struct CoopBuf;

struct DelLog {
  PoolRec* pool;  // the block's pool (LDS)
  int* count;     // records reserved so far in the block (LDS)
  int cap;        // pool capacity
  int owner;      // this lane
  int n;          // records this lane's event has pooled
  int medium;     // medium index of the current step
  CoopBuf* coop;  // this wave's buffer for target-selecting hard events (null: inline)
  const Medium* media;  // media by index (device memory)
  int n_media;
};

struct DelSample {
  double anc, w, denom;
  bool deferred;  // weight = (sum of components at (anc, anc - k)) / denom
};

// The draws of reverse_del_event up to the ancestor energy; g guesses the
// kernel-grid interval of k.
__device__ __forceinline__ DelSample del_sample(const DevKernel& c, double mass, double k, int g,
                                                Philox& rng, int& err) {
  DelSample d;
  d.anc = k; d.w = 0.0; d.denom = 1.0; d.deferred = false;
  const double sigma = klookup_g(k, c.e, c.sigma, c.n, c.pow2, g, err);
  if (err) return d;
  if (!(sigma > 0)) { err |= ERR_TRANSPORT; return d; }
  const double pcel = k >= c.onset ? 0.0
                                   : -log(c.cutoff) / (-log(c.cutoff) + log(c.onset / (c.onset - k)));
  if (pcel > 0 && uniform(rng) < pcel) { d.w = 1 / pcel; return d; }
  const double me = dpm::ELECTRON_MASS;
  const double xmax = k < .5 * (mass - me) * (mass - me) / me
                          ? fmin(1.0, 2 * me * (k + 2 * mass) / ((mass + me) * (mass + me)))
                          : 1.0;
  const double xmin = k >= c.onset * (1 - c.cutoff) ? c.cutoff : 1 - k / c.onset;
  if (!(xmax > xmin)) { d.w = 1.0; return d; }
  const double invnorm = 1 / xmin - 1 / xmax;
  const double nu = 1 / (1 / xmin - uniform(rng) * invnorm);
  const double r = 1 - nu;
  d.anc = k / r;
  const double pdf = 1 / (invnorm * (nu * nu));
  d.denom = pdf * r * sigma * (1 - pcel);
  d.deferred = true;
  return d;
}

// Reserves a pool slot; false when the pool is full (evaluate inline).
__device__ __forceinline__ bool del_push(DelLog* log, const DelSample& d, double k, int kind) {
  const int i = atomicAdd(log->count, 1);
  if (i >= log->cap) return false;
  PoolRec& r = log->pool[i];
  r.anc = d.anc;
  r.q = d.anc - k;
  r.denom = d.denom;
  r.owner = log->owner;
  r.mk = log->medium << 1 | kind;
  ++log->n;
  return true;
}

// differential_kernel (src/MixedTransport.jl:107-109) in component order.
__device__ __noinline__ static double del_total(const DevKernel& c, double mass, double anc,
                                                double q) {
  double total = 0.0;
  for (int i = 0; i < c.n_el; ++i)
    for (int p = 1; p <= 4; ++p) total += diff_component(c, mass, i, p, anc, q);
  return total;
}

// All threads of the block, after a barrier that follows the last push:
// factor of every pooled record.  kern(m) gives the DevKernel of medium m.
template <class KernOf>
__device__ __forceinline__ void pool_eval(PoolRec* pool, int m, double mass, KernOf kern) {
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    PoolRec& r = pool[j];
    r.denom = del_total(kern(r.mk >> 1), mass, r.anc, r.q) / r.denom;
  }
}

// After a barrier that follows pool_eval: this lane's factors in recording
// order applied to its event weight w.  pend is the error the transport
// raised after the lane's nown records (0: none).  Returns the final weight
// and sets fail to the error that stands.
__device__ __forceinline__ double pool_apply(const PoolRec* pool, int m, int owner, int nown,
                                             double w, int pend, int& fail) {
  bool dead = false;
  fail = 0;
  int seen = 0;
  for (int j = 0; j < m && seen < nown; ++j) {
    if (pool[j].owner != owner) continue;
    ++seen;
    const double f = pool[j].denom;
    if (f == 0.0) { dead = true; break; }
    w *= f;
    if (!(isfinite(w) && w >= 0)) {
      if (pool[j].mk & 1) dead = true;
      else fail = ERR_TRANSPORT;
      break;
    }
  }
  if (!dead && !fail && pend) fail = pend;
  return (fail || dead) ? 0.0 : w;
}

// ---------------------------------------------------------------------------
// A hard event that selects a target (straggled step with scattering,
// src/Geometry.jl:261-278, reverse_del_event with details = true) cannot wait
// for the pool: the chosen process deflects the muon in the same step.  Its
// components are evaluated at once by every lane of the wave that is in the
// same straggled step; the owner sums them in order and picks the target.
//
// The lanes split the work below the component level: one item per
// quadrature node of the pair-production (12) and photonuclear (9) cross
// sections of every element, then one per bremsstrahlung and ionisation
// value.  Items of one kind run together, so the wave evaluates one DDCS
// node per pass instead of a whole quadrature per component.  Each component
// is then assembled by one lane from its node terms in node order (the
// arithmetic of dcs_pair / dcs_photo), and the owner sums the components in
// order, so every value equals the inline evaluation.
constexpr int COOP_SLOTS = 2;  // events per round (rarely more than two hard events per step)
constexpr int COOP_EL = 8;     // elements per medium (more: the owner evaluates alone)
constexpr int COOP_LEAF = 23;  // per element: bremsstrahlung, 12 pair nodes, 9 photonuclear nodes, ionisation

struct CoopBuf {
  double leaf[COOP_SLOTS][COOP_EL][COOP_LEAF];
  double comp[COOP_SLOTS][4 * COOP_EL];
  double anc[COOP_SLOTS], q[COOP_SLOTS];
  int mi[COOP_SLOTS];
};

// Slot and element of the e-th element of a round (slots in order).
__device__ __forceinline__ void coop_elem(int e, const int* ne, int ng, int& s, int& i) {
  s = 0;
  i = e;
  while (s + 1 < ng && i >= ne[s]) { i -= ne[s]; ++s; }
}

// Orders the LDS accesses of the lanes of one wave (no hardware barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Owner-only evaluation after del_sample: total, weight and target.
__device__ __forceinline__ DelOut del_finish(const DevKernel& c, double mass, double k,
                                             const DelSample& d, Philox& rng) {
  const double q = d.anc - k;
  const double total = del_total(c, mass, d.anc, q);
  DelOut o{d.anc, total / d.denom, 0, -1};
  if (total <= 0) return o;
  double u = uniform(rng) * total;
  for (int i = 0; i < c.n_el; ++i)
    for (int p = 1; p <= 4; ++p) {
      u -= diff_component(c, mass, i, p, d.anc, q);
      if (u <= 0) { o.proc = p; o.elem = i; return o; }
    }
  o.proc = 4; o.elem = c.n_el - 1;
  return o;
}

// Called by every lane active in the caller (the branch to it is uniform over
// those lanes).  Lanes with del get reverse_del_event(details = true) at
// energy k in medium mi (g guesses k's kernel-grid interval): the same draws,
// ancestor energy, weight and target as the inline evaluation.
//
// Convergence: the routine does not assume that the whole wave, or any
// particular set of lanes, is present.  It is reached from per-lane step
// loops, so the lanes executing one call are whatever subset the EXEC mask
// holds there; the routine takes that subset from its own __ballot(1) and
// exchanges data only among those lanes: every loop that contains a
// wave_sync() is bounded by ballot results (uniform over the subset), and the
// per-lane work loop between two wave_sync() calls reconverges before the
// next one.  A lane's result depends only on its own event and draws; the
// subset only decides which lanes evaluate which DCS components, so splitting
// the lanes over several calls (by the hardware or by the compiler) gives the
// same results.  The LDS buffer belongs to the block's single wave, and
// successive calls are ordered by the final wave_sync().  A host emulation
// that requires all 64 lanes at every ballot therefore cannot run this
// routine; the per-event GPU tests against the oracle cover it.
__device__ __noinline__ static DelOut coop_del_event(const DelLog& lg, int mi, double mass,
                                                     double k, int g, Philox& rng, bool del,
                                                     int& err) {
  CoopBuf* cb = lg.coop;
  DelOut out{k, 0.0, 0, -1};
  DelSample d{k, 0.0, 1.0, false};
  bool need = false;
  if (del) {
    const DevKernel& c = lg.media[mi].kern;
    d = del_sample(c, mass, k, g, rng, err);
    if (!err) {
      if (!d.deferred) out = DelOut{d.anc, d.w, 0, -1};
      else if (c.n_el > COOP_EL) out = del_finish(c, mass, k, d, rng);
      else need = true;
    }
  }
  const uint64_t act = __ballot(1);
  const int lane = __lane_id();
  const uint64_t below = lane ? (act & (~0ull >> (64 - lane))) : 0ull;
  const int rank = __popcll(below);
  const int A = __popcll(act);
  uint64_t left = __ballot(need);
  while (left) {
    uint64_t grp = 0;
    int ng = 0;
    while (left && ng < COOP_SLOTS) {
      const uint64_t b = left & (~left + 1);
      grp |= b;
      left ^= b;
      ++ng;
    }
    const bool mine = (grp >> lane) & 1;
    const int slot = __popcll(lane ? (grp & (~0ull >> (64 - lane))) : 0ull);
    if (mine) {
      cb->anc[slot] = d.anc;
      cb->q[slot] = d.anc - k;
      cb->mi[slot] = mi;
    }
    wave_sync();
    int ms[COOP_SLOTS], ne[COOP_SLOTS];
    int nt = 0;  // elements of the round
    for (int s = 0; s < ng; ++s) {
      ms[s] = min(max(cb->mi[s], 0), lg.n_media - 1);
      ne[s] = min(lg.media[ms[s]].kern.n_el, COOP_EL);
      nt += ne[s];
    }
    // node terms and single values (differential_components, src/MixedTransport.jl:99-106)
    for (int w = rank; w < 12 * nt; w += A) {
      const int e = w / 12, node = w - 12 * e;
      int s, i;
      coop_elem(e, ne, ng, s, i);
      const double* el = lg.media[ms[s]].kern.el + 4 * i;
      cb->leaf[s][i][1 + node] = dpm::dcs_pair_node(el[0], el[1], mass, cb->anc[s], cb->q[s], node);
    }
    for (int w = rank; w < 9 * nt; w += A) {
      const int e = w / 9, node = w - 9 * e;
      int s, i;
      coop_elem(e, ne, ng, s, i);
      const double* el = lg.media[ms[s]].kern.el + 4 * i;
      cb->leaf[s][i][13 + node] = dpm::dcs_photo_node(el[0], el[1], mass, cb->anc[s], cb->q[s], node);
    }
    for (int w = rank; w < 2 * nt; w += A) {
      int s, i;
      coop_elem(w >> 1, ne, ng, s, i);
      const double* el = lg.media[ms[s]].kern.el + 4 * i;
      if (w & 1) cb->leaf[s][i][22] = dpm::dcs_ionisation(el[0], el[2], mass, cb->anc[s], cb->q[s]);
      else cb->leaf[s][i][0] = dpm::dcs_brems(el[0], el[1], mass, cb->anc[s], cb->q[s]);
    }
    wave_sync();
    // components in (element, process) order: dcs_pair and dcs_photo sums in
    // node order, then the factor of diff_component
    for (int w = rank; w < 4 * nt; w += A) {
      int s, i;
      coop_elem(w >> 2, ne, ng, s, i);
      const int p = w & 3;
      const double* lf = cb->leaf[s][i];
      double raw;
      if (p == 0) {
        raw = lf[0];
      } else if (p == 1) {
        double I = 0.0;
        for (int n = 0; n < 12; ++n) I -= lf[1 + n];
        raw = I < 0.0 ? 0.0 : I;
      } else if (p == 2) {
        double ds = 0.0;
        for (int n = 0; n < 9; ++n) ds += lf[13 + n];
        raw = ds < 0.0 ? 0.0 : ds;
      } else {
        raw = lf[22];
      }
      const double* el = lg.media[ms[s]].kern.el + 4 * i;
      cb->comp[s][4 * i + p] = el[3] * dpm::AVOGADRO_NUMBER / (el[1] * .001) * cb->anc[s] * raw;
    }
    wave_sync();
    if (mine) {
      const int nc = 4 * lg.media[mi].kern.n_el;
      double total = 0.0;
      for (int j = 0; j < nc; ++j) total += cb->comp[slot][j];
      out = DelOut{d.anc, total / d.denom, 0, -1};
      if (total > 0) {
        double u = uniform(rng) * total;
        out.proc = 4;
        out.elem = nc / 4 - 1;
        for (int j = 0; j < nc; ++j) {
          u -= cb->comp[slot][j];
          if (u <= 0) { out.proc = (j & 3) + 1; out.elem = j >> 2; break; }
        }
      }
    }
    wave_sync();  // the buffer is read before the next round writes it
  }
  return out;
}

}  // namespace dpk

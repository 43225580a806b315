// Deferred evaluation of hard-event (DEL) weights.
//
// reverse_del_event (src/MixedTransport.jl:111-142) draws the ancestor energy
// from a nu^-2 proposal and weights it by the differential kernel summed over
// every element and process of the medium (4 DCS models per element).  Where
// the weight only multiplies the event weight - the mixed step, and the
// straggled step without scattering, where the element/process choice only
// consumes its draw - the transport keeps going with the ancestor energy and
// records (ancestor, q, denominator, medium).  A separate pass then evaluates
// every record with all lanes doing the same DCS sequence, and multiplies each
// event's factors in order.  The draws consumed, the energies and the factors
// are those of the inline evaluation; only the product order of one event's
// weight changes (rounding).
//
// A record whose factor is 0 kills the event (weight 0), as in the inline
// path; an error raised by the transport after such a record is void, so
// errors of events with records are resolved in the same pass.
#pragma once
#include "dp_transport.hpp"

namespace dpk {

constexpr int DEL_MAXR = 8;  // records kept per event; more are evaluated inline

struct DelRec {
  double anc, q, denom;
  int medium;
  int kind;  // 0: mixed step (a non-finite weight is an error), 1: straggled (kills)
};
struct DelLog {
  DelRec* slot;  // this event's DEL_MAXR slots
  int n;         // records used
  int medium;    // medium index of the current step
};

struct DelSample {
  double anc, w, denom;
  bool deferred;  // weight = (sum of components at (anc, anc - k)) / denom
};

// The draws of reverse_del_event up to the ancestor energy.
__device__ __forceinline__ DelSample del_sample(const DevKernel& c, double mass, double k,
                                                Philox& rng, int& err) {
  DelSample d;
  d.anc = k; d.w = 0.0; d.denom = 1.0; d.deferred = false;
  const double sigma = klookup(k, c.e, c.sigma, c.n, c.pow2, err);
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

__device__ __forceinline__ void del_push(DelLog* log, const DelSample& d, double k, int kind) {
  DelRec& r = log->slot[log->n++];
  r.anc = d.anc; r.q = d.anc - k; r.denom = d.denom; r.medium = log->medium; r.kind = kind;
}

// differential_kernel (src/MixedTransport.jl:107-109) in component order.
__device__ __noinline__ static double del_total(const DevKernel& c, double mass, double anc,
                                                double q) {
  double total = 0.0;
  for (int i = 0; i < c.n_el; ++i)
    for (int p = 1; p <= 4; ++p) total += diff_component(c, mass, i, p, anc, q);
  return total;
}

// Records of event e: list[off[e] + j] = e * DEL_MAXR + j.
static __global__ void del_list_kernel_i(int64_t n, const int* __restrict__ nrec,
                                  const int* __restrict__ off, int64_t* __restrict__ list) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  for (int j = 0; j < nrec[e]; ++j) list[off[e] + j] = e * DEL_MAXR + j;
}

// factor of each record (persistent grid over the record list).
static __global__ __launch_bounds__(128, 4) void del_eval_kernel(const int64_t* __restrict__ count,
                                                       const int64_t* __restrict__ list,
                                                       const DelRec* __restrict__ rec,
                                                       const Medium* __restrict__ media,
                                                       double mass, double* __restrict__ fac,
                                                       int* __restrict__ kind) {
  const int64_t total = *count;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int64_t slot = list[t];
    const DelRec r = rec[slot];
    const DevKernel& c = media[r.medium].kern;
    fac[slot] = del_total(c, mass, r.anc, r.q) / r.denom;
    kind[slot] = r.kind;
  }
}

// Apply each event's factors in order.  err_pend[e] holds an error raised
// after err_nrec[e] records; it stands unless one of those records killed
// the event.  e0 is the global index of event 0 (error reports).
static __global__ void del_apply_kernel_i(int64_t n, const int* __restrict__ nrec,
                                 const double* __restrict__ fac, const int* __restrict__ kind,
                                 const uint8_t* __restrict__ err_pend,
                                 const uint8_t* __restrict__ err_nrec, double* __restrict__ out_w,
                                 int* __restrict__ err, unsigned long long* __restrict__ bad,
                                 int64_t e0) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int m = nrec[e];
  if (m == 0) return;
  double w = out_w[e];
  const int pend = err_pend[e];
  const int upto = pend ? err_nrec[e] : m;
  bool dead = false;
  int fail = 0;
  for (int j = 0; j < upto; ++j) {
    const double f = fac[e * DEL_MAXR + j];
    if (f == 0.0) { dead = true; break; }
    w *= f;
    if (!(isfinite(w) && w >= 0)) {
      if (kind[e * DEL_MAXR + j]) dead = true;
      else fail = ERR_TRANSPORT;
      break;
    }
  }
  if (!dead && !fail && pend) fail = pend;
  if (fail) {
    atomicOr(err, fail);
    atomicMin(bad, (unsigned long long)(e0 + e));
    out_w[e] = 0.0;
  } else {
    out_w[e] = dead ? 0.0 : w;
  }
}

}  // namespace dpk

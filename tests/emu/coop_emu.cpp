// Lockstep check of the wave-cooperative hard-event evaluation
// (coop_del_event, diffpumas.jl_amd/csrc/dp_del_defer.hpp) on the host.
//
// A block of B threads stands for the B lanes that execute one call
// together (B = 64: the whole wave; smaller B: a subset of it, as when some
// lanes of the wave are elsewhere).  Every lane draws an energy, a medium and
// whether it has a hard event, calls coop_del_event, and compares the result
// with the inline evaluation reverse_del_event(details = true) run on a copy
// of its generator: ancestor energy, weight, process, element, error bits and
// the generator state afterwards must be identical (bitwise).  Lanes without
// a hard event must get their energy back with weight 0 and an untouched
// generator.  Media: a 7-element mixture (rock + water + air), a 4-element
// air, and a 9-element medium that exceeds the cooperative buffer (its owner
// evaluates alone).  Prints "<n> failures".
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "dp_del_defer.hpp"

using namespace dpk;

namespace {

struct ElementDef {
  double Z, A, I, f;
};

struct HostKernel {
  std::vector<double> e, sigma, el;
  DevKernel k;
};

// A mixed-kernel snapshot with a smooth positive hard cross-section on a
// log grid from 0.05 GeV to 1e7 GeV (only the grid, sigma, onset, cutoff and
// elements enter the hard-event sampler).
void make_kernel(HostKernel& h, const std::vector<ElementDef>& els) {
  const int n = 400;
  h.e.resize(n);
  h.sigma.resize(n);
  for (int j = 0; j < n; ++j) {
    h.e[j] = 0.05 * std::pow(10.0, 8.3 * j / (n - 1));
    h.sigma[j] = 2e-4 * (1.0 + 0.3 * std::log(h.e[j] / 0.05));
  }
  for (const ElementDef& d : els) {
    h.el.push_back(d.Z); h.el.push_back(d.A); h.el.push_back(d.I); h.el.push_back(d.f);
  }
  int p2 = 1;
  while (p2 * 2 <= n - 2) p2 *= 2;
  h.k = DevKernel{};
  h.k.n = n;
  h.k.pow2 = p2;
  h.k.n_el = (int)els.size();
  h.k.e = h.e.data();
  h.k.sigma = h.sigma.data();
  h.k.stop = h.k.range = h.k.hazard = h.k.proper = h.sigma.data();  // not read by the sampler
  h.k.el = h.el.data();
  h.k.onset = 1.0;
  h.k.cutoff = 0.05;
}

constexpr double MASS = 0.10565839;

struct Args {
  const Medium* media;
  int n_media;
  uint64_t seed;
  double p_del;
  int* failures;
  int* hard;  // lanes that evaluated a hard event
};

bool same_rng(const Philox& a, const Philox& b) {
  if (a.k0 != b.k0 || a.k1 != b.k1 || a.e0 != b.e0 || a.e1 != b.e1 || a.block != b.block ||
      a.pos != b.pos)
    return false;
  for (int i = a.pos; i < 4; ++i)
    if (a.buf[i - a.pos] != b.buf[i - b.pos]) return false;
  return true;
}

void lane_kernel(Args a) {
  __shared__ CoopBuf coop;
  const int t = threadIdx.x;
  Philox rng;
  rng.init(a.seed, (uint64_t)(blockIdx.x * 1000 + t));
  // energy 0.2 GeV .. 1e4 GeV, medium, hard event or not
  const double k = 0.2 * std::exp(std::log(5e4) * rng.uniform());
  const int mi = (int)(rng.uniform() * a.n_media) % a.n_media;
  const bool del = rng.uniform() < a.p_del;
  DelLog lg{nullptr, nullptr, 0, t, 0, mi, &coop, a.media, a.n_media};
  Philox ref = rng;
  int err = 0;
  // any interval guess gives the bisection's interval
  const int g = (int)(a.media[mi].kern.n * ((t * 37 + 11) % 64) / 64.0);
  const DelOut o = coop_del_event(lg, mi, MASS, k, g, rng, del, err);
  bool ok;
  if (del) {
    int err_ref = 0;
    const DelOut r = reverse_del_event(a.media[mi].kern, MASS, k, ref, true, err_ref);
    ok = o.k == r.k && o.w == r.w && o.proc == r.proc && o.elem == r.elem && err == err_ref &&
         same_rng(rng, ref);
    __atomic_fetch_add(a.hard, 1, __ATOMIC_SEQ_CST);
    if (!ok)
      std::printf("mismatch lane %d medium %d k %.17g: coop (%.17g %.17g %d %d err %d) inline "
                  "(%.17g %.17g %d %d err %d)\n",
                  t, mi, k, o.k, o.w, o.proc, o.elem, err, r.k, r.w, r.proc, r.elem, err_ref);
  } else {
    ok = o.k == k && o.w == 0.0 && o.proc == 0 && o.elem == -1 && err == 0 && same_rng(rng, ref);
    if (!ok) std::printf("lane %d without a hard event changed\n", t);
  }
  if (!ok) __atomic_fetch_add(a.failures, 1, __ATOMIC_SEQ_CST);
}

}  // namespace

int main() {
  const std::vector<ElementDef> mixture = {
      {11, 22, 136.4e-9, 0.85},      {1, 1.008, 19.2e-9, 0.0157}, {8, 15.999, 95e-9, 0.1243},
      {6, 12.011, 78e-9, 0.0001},    {7, 14.007, 82e-9, 0.0075},  {8, 15.999, 95e-9, 0.0023},
      {18, 39.948, 188e-9, 0.0001}};
  const std::vector<ElementDef> air = {{6, 12.011, 78e-9, 0.000124},
                                       {7, 14.007, 82e-9, 0.755267},
                                       {8, 15.999, 95e-9, 0.231781},
                                       {18, 39.948, 188e-9, 0.012827}};
  std::vector<ElementDef> big;
  for (int i = 0; i < 9; ++i) big.push_back({6.0 + i, 12.0 + 2.1 * i, (70.0 + 10 * i) * 1e-9, 1.0 / 9});
  HostKernel hk[3];
  make_kernel(hk[0], mixture);
  make_kernel(hk[1], air);
  make_kernel(hk[2], big);
  Medium media[3];
  for (int m = 0; m < 3; ++m) {
    media[m] = Medium{};
    media[m].kern = hk[m].k;
  }
  int failures = 0, hard = 0;
  const int widths[] = {64, 37, 8, 1};
  const double probs[] = {0.03, 0.3, 1.0};
  int cases = 0;
  for (int nm = 2; nm <= 3; ++nm)
    for (int B : widths)
      for (double p : probs) {
        Args a{media, nm, 1234567ull + (uint64_t)cases, p, &failures, &hard};
        emu_launch(lane_kernel, dim3(6), dim3((unsigned)B), a);
        ++cases;
      }
  std::printf("%d cases, %d hard events, %d failures\n", cases, hard, failures);
  return failures == 0 && hard > 0 ? 0 : 1;
}

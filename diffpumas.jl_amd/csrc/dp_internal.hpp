// Host-side internals shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/diffpumas_hip.h"

namespace dph {

// Per-thread error message of the last failing call.
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define DP_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return dph::fail(DP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Growable device scratch buffer owned by a context.  Every reallocation
// bumps the owning context's generation (captured graphs hold raw pointers
// into these buffers and are refused once any of them moved), and a buffer
// cannot grow while its context captures a graph.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  uint64_t* generation = nullptr;  // the owning context's counter
  const bool* frozen = nullptr;    // the owning context's capture flag
};
int ensure(DevBuf& b, size_t bytes);
void release(DevBuf& b);

}  // namespace dph

#include <functional>
#include <map>
#include <memory>

namespace dph {
// Launch knobs of one kernel configuration (their meaning is per kernel).
struct TuneCfg {
  int k0 = 0, k1 = 0, k2 = 0;
};
// What the autotuner measured for one (kernel, workload shape) key.
struct TuneEntry {
  TuneCfg best;
  float best_ms = 0.f;
  std::vector<std::pair<TuneCfg, float>> tried;
};
// Element of a material recipe (Z, A, I in GeV, mass fraction).
struct Element { double Z, A, I, f; };
// A prepared object living in device memory (mixture tables or mixed kernel).
struct Prepared {
  void* dev = nullptr;        // one allocation holding all arrays
  std::vector<double> host;   // host copy of the same arrays
  int n = 0, pow2 = 0, n_el = 0;
  double onset = 0, cutoff = 0;
  ~Prepared() { if (dev) (void)hipFree(dev); }
};
}  // namespace dph

struct dp_tables_s {
  int32_t n_materials = 0, n_knots = 0, pow2 = 0;
  bool shared_grid = false;    // all materials on one energy grid
  double mass = 0, ctau = 0, cutoff = 0.05, elastic_ratio = 0.05;
  double* data = nullptr;      // device: M * NARR * K
  double* scalars = nullptr;   // device: M * 4
  std::vector<double> host_data;     // M * NARR * K
  std::vector<double> host_scalars;  // M * 4
  std::vector<double> last_energy;   // per material, top of the table
  std::vector<std::vector<dph::Element>> recipes;  // per material (may be empty)
  // Lazily prepared per (components, fractions) key: mixture range tables
  // and mixed kernels (MixedKernel snapshots).
  std::map<std::vector<double>, std::unique_ptr<dph::Prepared>> mixtures, kernels;
};

struct dp_ctx_s {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int n_cu = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
  dph::DevBuf scratch;       // kernel scratch (partials, queues)
  dph::DevBuf scratch2;      // second scratch (sort buffers)
  dph::DevBuf counter;       // small: work-queue counters and the sticky error word
  dph::DevBuf staging;       // host-API staging of inputs/outputs
  dp_csda_correction corr{};            // w-model corrections (off)
  dp_csda_material_correction mcorr{};
  bool capturing = false;               // between dp_capture_begin and dp_capture_end
  uint64_t generation = 0;              // bumped by every scratch reallocation
  void* noalloc_scratch = nullptr;      // this device's static scratch (no-allocation VJP)
  bool autotune = false;                // time launch candidates on first use
  bool fp32_stepping = false;           // density adjoint steps in FP32
  std::map<std::string, dph::TuneEntry> tune;  // chosen launch knobs per key
};

namespace dph {
// Largest power of two <= n - 2 (bracket search start).
inline int pow2_for(int n) {
  int p = 1;
  while (p * 2 <= n - 2) p *= 2;
  return n - 2 >= 1 ? p : 0;
}
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
// Int slots of the context's counter buffer: [COUNTER_QUEUE] the work queue
// of the persistent kernels, [COUNTER_STICKY] the error bits the _dev entry
// points accumulate until dp_check_errors reads and clears them.
constexpr int COUNTER_QUEUE = 0;
constexpr int COUNTER_STICKY = 16;
// Error bits of the deterministic kernels (both the per-call and the sticky
// word): a path longer than the scratch sized for it, a cell index outside
// the field.
constexpr int KERR_CAPACITY = 1;
constexpr int KERR_CELL = 2;
// Device address of the no-allocation VJP's static scratch on `device`.
int noalloc_scratch(int device, void** out);
// Launch knobs for `key`: the cached choice if there is one; otherwise, with
// autotuning on and no capture open, the fastest of `cands` (each timed once
// by run(cfg), which enqueues the work on the context's stream), cached; else
// cands[0], the default.  The kernels' outputs do not depend on the knobs, so
// the timed runs only overwrite outputs with the same values.
int tuned(dp_ctx c, const std::string& key, const std::vector<TuneCfg>& cands,
          const std::function<int(const TuneCfg&)>& run, TuneCfg* out);
// Timing event on the context's stream; not recorded while the stream is
// being captured into a graph.
inline void record(dp_ctx c, hipEvent_t e) {
  if (!c->capturing) (void)hipEventRecord(e, c->stream);
}
}  // namespace dph

// Context lifecycle, error reporting, table upload.
#include "dp_internal.hpp"
#include <algorithm>
#include <cstring>
#include <mutex>

namespace dph {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
int ensure(DevBuf& b, size_t bytes) {
  if (bytes <= b.bytes && b.ptr) return DP_OK;
  if (b.frozen && *b.frozen)
    return fail(DP_E_ARGUMENT, "a device buffer would have to grow during a graph capture: call "
                               "the entry point once with the same sizes before capturing");
  if (b.generation) ++*b.generation;
  if (b.ptr) {
    hipError_t e = hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return fail(DP_E_HIP, std::string("hipFree: ") + hipGetErrorString(e));
  }
  size_t want = align_up(bytes < 256 ? 256 : bytes, 256);
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    return fail(DP_E_HIP, std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
  }
  b.bytes = want;
  return DP_OK;
}
void release(DevBuf& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

int tuned(dp_ctx c, const std::string& key, const std::vector<TuneCfg>& cands,
          const std::function<int(const TuneCfg&)>& run, TuneCfg* out) {
  auto it = c->tune.find(key);
  if (it != c->tune.end()) { *out = it->second.best; return DP_OK; }
  *out = cands.front();
  if (!c->autotune || c->capturing || cands.size() < 2) return DP_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail(DP_E_HIP, "autotune: hipEventCreate");
  }
  TuneEntry E;
  E.best_ms = 3.0e38f;
  // one untimed run first, so no candidate is timed on cold caches and pages
  int rc = run(cands.front());
  if (rc) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  for (const TuneCfg& cfg : cands) {
    if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(DP_E_HIP, "autotune: hipEventRecord"); break; }
    if ((rc = run(cfg)) != DP_OK) break;
    if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      rc = fail(DP_E_HIP, "autotune: timing the candidate failed");
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    E.tried.push_back({cfg, ms});
    if (ms < E.best_ms) { E.best_ms = ms; E.best = cfg; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  c->tune[key] = E;
  *out = E.best;
  return DP_OK;
}
}  // namespace dph

extern "C" {

const char* dp_last_error(void) { return dph::g_last_error.c_str(); }

int dp_create(dp_ctx* out, int device) {
  if (!out) return dph::fail(DP_E_ARGUMENT, "dp_create: null output");
  *out = nullptr;
  int n = 0;
  DP_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return dph::fail(DP_E_ARGUMENT, "dp_create: device " + std::to_string(device) + " not present");
  DP_HIP(hipSetDevice(device));
  dp_ctx c = new dp_ctx_s();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return dph::fail(DP_E_HIP, "hipGetDeviceProperties"); }
  c->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return dph::fail(DP_E_HIP, "dp_create: stream/event creation failed");
  }
  c->stream = c->own_stream;
  // scratch, scratch2 and counter are what _dev entry points (and so captured
  // graphs) use; staging only serves the host entry points
  for (dph::DevBuf* b : {&c->scratch, &c->scratch2, &c->counter}) b->generation = &c->generation;
  for (dph::DevBuf* b : {&c->scratch, &c->scratch2, &c->counter, &c->staging}) b->frozen = &c->capturing;
  // the counter buffer exists from the start with a zero sticky error word
  if (dph::ensure(c->counter, 256) != DP_OK || hipMemset(c->counter.ptr, 0, 256) != hipSuccess) {
    const std::string msg = dp_last_error();
    dph::release(c->counter);
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return dph::fail(DP_E_HIP, "dp_create: counter buffer: " + msg);
  }
  if (dph::noalloc_scratch(device, &c->noalloc_scratch) != DP_OK) {
    const std::string msg = dp_last_error();
    dph::release(c->counter);
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return dph::fail(DP_E_HIP, "dp_create: " + msg);
  }
  *out = c;
  return DP_OK;
}

int dp_destroy(dp_ctx c) {
  if (!c) return DP_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  dph::release(c->scratch);
  dph::release(c->scratch2);
  dph::release(c->counter);
  dph::release(c->staging);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return DP_OK;
}

int dp_set_stream(dp_ctx c, void* stream) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_set_stream: null context");
  c->stream = stream ? (hipStream_t)stream : c->own_stream;
  return DP_OK;
}

int dp_synchronize(dp_ctx c) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_synchronize: null context");
  DP_HIP(hipStreamSynchronize(c->stream));
  return DP_OK;
}

int dp_set_autotune(dp_ctx c, int32_t on) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_set_autotune: null context");
  c->autotune = on != 0;
  return DP_OK;
}

int dp_set_fp32_stepping(dp_ctx c, int32_t on) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_set_fp32_stepping: null context");
  c->fp32_stepping = on != 0;
  return DP_OK;
}

int dp_set_tuning(dp_ctx c, const char* key, int32_t k0, int32_t k1, int32_t k2) {
  if (!c || !key) return dph::fail(DP_E_ARGUMENT, "dp_set_tuning: null argument");
  dph::TuneEntry E;
  E.best = {k0, k1, k2};
  c->tune[key] = E;
  return DP_OK;
}

int dp_clear_tuning(dp_ctx c) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_clear_tuning: null context");
  c->tune.clear();
  return DP_OK;
}

int64_t dp_tuning_report(dp_ctx c, char* buf, int64_t size) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_tuning_report: null context");
  std::string r;
  for (const auto& kv : c->tune) {
    const dph::TuneEntry& E = kv.second;
    r += kv.first + " | best " + std::to_string(E.best.k0) + "," + std::to_string(E.best.k1) + "," +
         std::to_string(E.best.k2) + " " + std::to_string(E.best_ms) + " ms | tried";
    for (const auto& t : E.tried)
      r += " " + std::to_string(t.first.k0) + "," + std::to_string(t.first.k1) + "," +
           std::to_string(t.first.k2) + ":" + std::to_string(t.second);
    r += "\n";
  }
  if (buf && size > 0) {
    const size_t n = std::min<size_t>(r.size(), (size_t)size - 1);
    std::memcpy(buf, r.data(), n);
    buf[n] = 0;
  }
  return (int64_t)r.size();
}

int dp_check_errors(dp_ctx c) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_check_errors: null context");
  if (c->capturing) return dph::fail(DP_E_ARGUMENT, "dp_check_errors: a capture is open");
  DP_HIP(hipSetDevice(c->device));
  DP_HIP(hipStreamSynchronize(c->stream));
  int* w = (int*)c->counter.ptr + dph::COUNTER_STICKY;
  int bits = 0;
  DP_HIP(hipMemcpy(&bits, w, sizeof(int), hipMemcpyDeviceToHost));
  if (!bits) return DP_OK;
  DP_HIP(hipMemset(w, 0, sizeof(int)));
  if (bits & dph::KERR_CELL)
    return dph::fail(DP_E_ARGUMENT, "a _dev call met a cell index outside the field");
  return dph::fail(DP_E_CAPACITY, "a _dev call met a path longer than its max_segments");
}

// ------------------------------------------------------------ graph capture
struct dp_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  dp_ctx ctx = nullptr;     // context whose buffers the graph's nodes point into
  uint64_t generation = 0;  // that context's buffer generation at capture
};

int dp_capture_begin(dp_ctx c) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_capture_begin: null context");
  if (c->capturing) return dph::fail(DP_E_ARGUMENT, "dp_capture_begin: a capture is already open");
  if (!c->stream) return dph::fail(DP_E_ARGUMENT, "dp_capture_begin: the null stream cannot be captured");
  DP_HIP(hipSetDevice(c->device));
  DP_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  return DP_OK;
}

int dp_capture_end(dp_ctx c, dp_graph* out) {
  if (!c || !out) return dph::fail(DP_E_ARGUMENT, "dp_capture_end: null argument");
  if (!c->capturing) return dph::fail(DP_E_ARGUMENT, "dp_capture_end: no capture is open");
  *out = nullptr;
  c->capturing = false;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return dph::fail(DP_E_HIP, std::string("dp_capture_end: the capture failed (only _dev entry "
                                           "points, with buffers already sized, can be "
                                           "captured): ") + hipGetErrorString(e));
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return dph::fail(DP_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  dp_graph h = new dp_graph_s;
  h->graph = g;
  h->exec = x;
  h->ctx = c;
  h->generation = c->generation;
  *out = h;
  return DP_OK;
}

int dp_graph_launch(dp_ctx c, dp_graph g) {
  if (!c || !g) return dph::fail(DP_E_ARGUMENT, "dp_graph_launch: null argument");
  if (c->capturing) return dph::fail(DP_E_ARGUMENT, "dp_graph_launch: a capture is open");
  if (g->ctx != c) return dph::fail(DP_E_ARGUMENT, "dp_graph_launch: the graph was captured on another context");
  if (g->generation != c->generation)
    return dph::fail(DP_E_ARGUMENT, "dp_graph_launch: a later call reallocated the context's device "
                                    "buffers that the graph uses; capture it again");
  DP_HIP(hipSetDevice(c->device));
  DP_HIP(hipGraphLaunch(g->exec, c->stream));
  return DP_OK;
}

int dp_graph_destroy(dp_graph g) {
  if (!g) return DP_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return DP_OK;
}

int dp_device_info(dp_ctx c, int32_t* n_cu, int32_t* wave) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_device_info: null context");
  if (n_cu) *n_cu = c->n_cu;
  if (wave) *wave = 64;
  return DP_OK;
}

double dp_last_kernel_ms(dp_ctx c) { return c ? (double)c->last_ms : 0.0; }

int dp_upload_tables(dp_ctx c, const dp_table_desc* d, dp_tables* out) {
  if (!c || !d || !out) return dph::fail(DP_E_ARGUMENT, "dp_upload_tables: null argument");
  *out = nullptr;
  if (d->n_materials <= 0 || d->n_knots < 3 || !d->data || !d->scalars)
    return dph::fail(DP_E_ARGUMENT, "dp_upload_tables: need >=1 material, >=3 knots, data and scalars");
  if (!(d->mass > 0) || !(d->ctau > 0))
    return dph::fail(DP_E_ARGUMENT, "dp_upload_tables: mass and ctau must be positive");
  const int M = d->n_materials, K = d->n_knots;
  // Energy grids must be strictly increasing and log_energies consistent.
  for (int m = 0; m < M; ++m) {
    const double* e = d->data + ((size_t)m * DP_T_NARRAYS + DP_T_ENERGY) * K;
    for (int i = 1; i < K; ++i)
      if (!(e[i] > e[i - 1]))
        return dph::fail(DP_E_ARGUMENT, "dp_upload_tables: energy grid of material " +
                                            std::to_string(m) + " is not strictly increasing");
  }
  DP_HIP(hipSetDevice(c->device));
  dp_tables t = new dp_tables_s();
  t->n_materials = M;
  t->n_knots = K;
  t->pow2 = dph::pow2_for(K);
  t->mass = d->mass;
  t->ctau = d->ctau;
  t->host_scalars.assign(d->scalars, d->scalars + (size_t)M * 4);
  t->host_data.assign(d->data, d->data + (size_t)M * DP_T_NARRAYS * K);
  t->cutoff = d->cutoff > 0 ? d->cutoff : 0.05;
  t->elastic_ratio = d->elastic_ratio > 0 ? d->elastic_ratio : 0.05;
  t->recipes.assign(M, {});
  if (d->n_elements) {
    size_t o = 0;
    for (int m = 0; m < M; ++m) {
      if (d->n_elements[m] < 0) { delete t; return dph::fail(DP_E_ARGUMENT, "dp_upload_tables: negative element count"); }
      for (int i = 0; i < d->n_elements[m]; ++i, ++o)
        t->recipes[m].push_back({d->element_Z[o], d->element_A[o], d->element_I[o], d->element_fraction[o]});
    }
  }
  for (int m = 0; m < M; ++m)
    t->last_energy.push_back(d->data[((size_t)m * DP_T_NARRAYS + DP_T_ENERGY) * K + K - 1]);
  // one energy grid for every material (energies and log energies identical)
  t->shared_grid = true;
  for (int m = 1; m < M; ++m)
    if (std::memcmp(d->data, d->data + (size_t)m * DP_T_NARRAYS * K, 2 * (size_t)K * sizeof(double)))
      t->shared_grid = false;
  size_t nb = (size_t)M * DP_T_NARRAYS * K * sizeof(double);
  if (hipMalloc(&t->data, nb) != hipSuccess || hipMalloc(&t->scalars, (size_t)M * 4 * sizeof(double)) != hipSuccess) {
    if (t->data) (void)hipFree(t->data);
    delete t;
    return dph::fail(DP_E_HIP, "dp_upload_tables: hipMalloc failed");
  }
  if (hipMemcpy(t->data, d->data, nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(t->scalars, d->scalars, (size_t)M * 4 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(t->data);
    (void)hipFree(t->scalars);
    delete t;
    return dph::fail(DP_E_HIP, "dp_upload_tables: copy failed");
  }
  *out = t;
  return DP_OK;
}

int dp_free_tables(dp_ctx c, dp_tables t) {
  if (!t) return DP_OK;
  if (c) (void)hipSetDevice(c->device);
  if (t->data) (void)hipFree(t->data);
  if (t->scalars) (void)hipFree(t->scalars);
  delete t;
  return DP_OK;
}

}  // extern "C"

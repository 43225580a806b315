// Host preparation of transport media shared by the Monte Carlo entry points:
// runtime-mixture range tables and MixedKernel snapshots, built once per
// (components, fractions) and cached on the tables handle in device memory.
#pragma once
#include <vector>
#include "dp_internal.hpp"
#include "dp_transport.hpp"
#include "dp_del_defer.hpp"

namespace dph {
// Fills md with device pointers for the mixture (mats, fractions); the
// fractions are normalised to sum to one.  with_kernel also prepares the
// MixedKernel (needs the element recipes of every component).
int make_medium(dp_tables_s* t, std::vector<int> mats, std::vector<double> fr, dpk::Medium& md,
                bool with_kernel = true);

// Device buffers of the deferred hard-event weights of n events, carved out
// of the context's third scratch buffer.
struct DelBuffers {
  dpk::DelRec* rec = nullptr;
  int* nrec = nullptr;
  int* off = nullptr;
  uint8_t* err_pend = nullptr;
  uint8_t* err_nrec = nullptr;
  int64_t* list = nullptr;
  double* fac = nullptr;
  int* kind = nullptr;
  int64_t* count = nullptr;
  dpk::Medium* media = nullptr;  // 4 slots for callers without a media array
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
};
int del_buffers(dp_ctx c, int64_t n, DelBuffers& b);
// Evaluates every record against media[] and applies each event's factors to
// out_w (stream ordered on the context's stream).
int del_resolve(dp_ctx c, int64_t n, DelBuffers& b, const dpk::Medium* media, double mass,
                double* out_w, int* err, unsigned long long* bad, int64_t e0 = 0);
}  // namespace dph

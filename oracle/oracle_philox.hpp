// CPU oracle: counter-based Philox4x32-10 stream keyed by (seed, event).
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Philox4x32-10 as published by Salmon et al., "Parallel random numbers: as
// easy as 1, 2, 3" (SC'11).  Stream convention shared with the device code:
//   key     = (seed & 0xffffffff, seed >> 32)
//   counter = (event & 0xffffffff, event >> 32, block, 0), block = 0, 1, ...
// each block yields four 32-bit words, consumed in order; a uniform double on
// [0, 1) takes two words w0, w1: ((w0 << 32 | w1) >> 11) * 2^-53.
#pragma once
#include <stdexcept>
#include <cstdint>
#include <cmath>

namespace orc {

struct Philox4x32 {
  uint32_t k0, k1;
  uint64_t event;
  uint32_t block = 0;
  uint32_t buf[4];
  int pos = 4;

  Philox4x32(uint64_t seed, uint64_t ev) : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), event(ev) {}

  static void round(uint32_t c[4], uint32_t a, uint32_t b) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t o0 = hi1 ^ c[1] ^ a, o1 = lo1, o2 = hi0 ^ c[3] ^ b, o3 = lo0;
    c[0] = o0; c[1] = o1; c[2] = o2; c[3] = o3;
  }
  static void block10(uint32_t c[4], uint32_t a, uint32_t b) {
    for (int r = 0; r < 10; ++r) {
      round(c, a, b);
      a += 0x9E3779B9u;
      b += 0xBB67AE85u;
    }
  }
  uint32_t next_u32() {
    if (pos == 4) {
      buf[0] = (uint32_t)event; buf[1] = (uint32_t)(event >> 32); buf[2] = block++; buf[3] = 0;
      block10(buf, k0, k1);
      pos = 0;
    }
    return buf[pos++];
  }
  // Scripted draws (tests only): when set, uniform() and exponential() return
  // the listed values in order, like the scripted RNGs of the reference tests
  // (test/test_mixed_kernel.jl:29-31,66, test/test_ehs_thinning.jl:4-12).
  const double* script = nullptr;
  int script_n = 0, script_pos = 0;
  double scripted() {
    if (script_pos >= script_n) throw std::runtime_error("scripted draws exhausted");
    return script[script_pos++];
  }
  double uniform() {  // [0, 1)
    if (script) return scripted();
    uint64_t a = next_u32();
    uint64_t b = next_u32();
    return (double)(((a << 32) | b) >> 11) * 0x1.0p-53;
  }
  double uniform_open() {  // (0, 1]
    return 1.0 - uniform();
  }
  double exponential() { return script ? scripted() : -std::log(uniform_open()); }
};

}  // namespace orc

// Runs the device radix sort (diffpumas.jl_amd/csrc/dp_sort.hpp) on the host
// through the thread-per-lane stand-in in tests/emu/hip and compares it with
// std::stable_sort.  Exit status 0 = every case sorted stably.
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>
#include "dp_sort.hpp"
int main() {
  int fails = 0;
  std::mt19937 rng(7);
  for (int n : {1, 63, 1025, 2500}) {
    for (int bits : {1, 9, 14}) {
      std::vector<int32_t> k(n), v(n), ping(2 * n), pong(2 * n);
      std::vector<int32_t> hist(dps::sort_hist_ints(n));
      const int range = 1 << bits;
      std::uniform_int_distribution<int> d(0, range - 1);
      for (int i = 0; i < n; ++i) {
        k[i] = n > 100 ? d(rng) % std::max(1, range / 7 + 1) : d(rng);   // many equal keys
        v[i] = i;
      }
      const int32_t *ko, *vo;
      dps::radix_sort_pairs(k.data(), v.data(), n, bits, ping.data(), pong.data(), hist.data(),
                            nullptr, &ko, &vo);
      std::vector<int> idx(n);
      std::iota(idx.begin(), idx.end(), 0);
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return k[a] < k[b]; });
      for (int i = 0; i < n; ++i)
        if (vo[i] != idx[i] || ko[i] != k[idx[i]]) { ++fails; printf("FAIL n=%d bits=%d\n", n, bits); break; }
    }
  }
  printf("sort emulation: %d failures\n", fails);
  return fails != 0;
}

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

}  // namespace dph

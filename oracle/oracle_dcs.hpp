// CPU oracle: differential cross-sections for the hard-event weights.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// The DCS models themselves are pinned directly against the official C PUMAS
// moments and the reference's golden tables (tests/test_tables.py), so the
// oracle evaluates that pinned model code rather than a second copy of it;
// what the oracle restates independently is the transport around it.
#pragma once
#include "../diffpumas.jl_amd/csrc/dp_models.hpp"

namespace orc {
using dpm::dcs_brems;
using dpm::dcs_pair;
using dpm::dcs_photo;
using dpm::dcs_ionisation;
}  // namespace orc

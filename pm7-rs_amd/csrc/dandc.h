// SPDX-License-Identifier: GPL-3.0-or-later
//
// Host side of the divide-and-conquer SCF: the spatial partition of a system into cores with
// buffers (pm7-rs src/dandc/partition.rs).
#pragma once

#include <cstdint>
#include <vector>

namespace pm7 {

struct DcPartition {
  std::vector<int> n_core;       // per subsystem
  std::vector<int64_t> start;    // n_sub + 1
  std::vector<int> members;      // core atoms first (bisection order), then buffer atoms ascending
  int n_sub() const { return (int)n_core.size(); }
};

// Recursive median bisection of the n atoms (xyz: 3n, Bohr) into cores of at most core_size
// atoms, each grown by every other atom within `buffer` Bohr of one of its atoms.
void dc_partition(int n, const double* xyz, int core_size, double buffer, DcPartition& out);

}  // namespace pm7

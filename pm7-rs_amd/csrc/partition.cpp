// SPDX-License-Identifier: GPL-3.0-or-later
//
// Divide-and-conquer partition (pm7-rs partition, src/dandc/partition.rs:97-190), molecular case.
//
// Cores: the atom set is split at its median along the axis of largest extent (a tie between
// extents goes to the later axis), the lower half first, until a set holds at most core_size atoms
// (or one). The order along the axis is a stable sort of the coordinate, so atoms at equal
// coordinates keep their index order, which fixes which side of the median they fall on.
//
// Buffers: every atom outside the core within `buffer` Bohr (d^2 <= buffer^2) of any core atom.
// Candidates come from a uniform grid of cells `buffer` wide (the 27 cells around each core atom),
// so growing every buffer costs O(N) for a bounded density. Buffer atoms are listed in ascending
// index order; the order is immaterial to the method (every subsystem quantity is invariant under a
// permutation of its atoms).
#include "dandc.h"

#include <algorithm>
#include <cmath>
#include <unordered_map>

namespace pm7 {

namespace {

void bisect(std::vector<int>& idx, size_t lo, size_t hi, const double* xyz, size_t target,
            std::vector<std::pair<size_t, size_t>>& leaves) {
  const size_t count = hi - lo;
  if (count <= target || count < 2) {
    leaves.emplace_back(lo, hi);
    return;
  }
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t k = lo; k < hi; ++k)
    for (int a = 0; a < 3; ++a) {
      const double v = xyz[3 * idx[k] + a];
      mn[a] = std::min(mn[a], v);
      mx[a] = std::max(mx[a], v);
    }
  int axis = 0;
  for (int a = 1; a < 3; ++a)
    if (!(mx[a] - mn[a] < mx[axis] - mn[axis])) axis = a;
  std::stable_sort(idx.begin() + lo, idx.begin() + hi,
                   [&](int p, int q) { return xyz[3 * p + axis] < xyz[3 * q + axis]; });
  const size_t half = lo + count / 2;
  bisect(idx, lo, half, xyz, target, leaves);
  bisect(idx, half, hi, xyz, target, leaves);
}

}  // namespace

void dc_partition(int n, const double* xyz, int core_size, double buffer, DcPartition& out) {
  out.n_core.clear();
  out.start.assign(1, 0);
  out.members.clear();
  if (n <= 0) return;
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::vector<std::pair<size_t, size_t>> leaves;
  bisect(idx, 0, (size_t)n, xyz, (size_t)std::max(1, core_size), leaves);

  // Cell grid for the buffer search.
  const double h = std::max(buffer, 1.0e-6);
  auto cell = [&](int i, int a) { return (int64_t)std::floor(xyz[3 * i + a] / h); };
  auto key = [](int64_t x, int64_t y, int64_t z) {
    return (uint64_t)(x * 73856093LL) ^ (uint64_t)(y * 19349663LL) ^ (uint64_t)(z * 83492791LL);
  };
  std::unordered_map<uint64_t, std::vector<int>> cells;
  cells.reserve((size_t)n);
  for (int i = 0; i < n; ++i) cells[key(cell(i, 0), cell(i, 1), cell(i, 2))].push_back(i);

  const double cut2 = buffer * buffer;
  std::vector<int> stamp(n, -1);
  std::vector<int> buf;
  for (size_t s = 0; s < leaves.size(); ++s) {
    const int sid = (int)s;
    const size_t lo = leaves[s].first, hi = leaves[s].second;
    for (size_t k = lo; k < hi; ++k) stamp[idx[k]] = sid;
    buf.clear();
    for (size_t k = lo; k < hi; ++k) {
      const int c = idx[k];
      const int64_t cx = cell(c, 0), cy = cell(c, 1), cz = cell(c, 2);
      for (int64_t dx = -1; dx <= 1; ++dx)
        for (int64_t dy = -1; dy <= 1; ++dy)
          for (int64_t dz = -1; dz <= 1; ++dz) {
            auto it = cells.find(key(cx + dx, cy + dy, cz + dz));
            if (it == cells.end()) continue;
            for (int b : it->second) {
              if (stamp[b] == sid) continue;
              // A hash collision can list an atom of a far cell: the distance test decides.
              const double ex = xyz[3 * b] - xyz[3 * c], ey = xyz[3 * b + 1] - xyz[3 * c + 1],
                           ez = xyz[3 * b + 2] - xyz[3 * c + 2];
              if (ex * ex + ey * ey + ez * ez > cut2) continue;
              stamp[b] = sid;
              buf.push_back(b);
            }
          }
    }
    std::sort(buf.begin(), buf.end());
    out.n_core.push_back((int)(hi - lo));
    out.members.insert(out.members.end(), idx.begin() + lo, idx.begin() + hi);
    out.members.insert(out.members.end(), buf.begin(), buf.end());
    out.start.push_back((int64_t)out.members.size());
  }
}

}  // namespace pm7

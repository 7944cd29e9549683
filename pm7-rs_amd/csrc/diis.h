// SPDX-License-Identifier: GPL-3.0-or-later
//
// DIIS extrapolation and the level-shift stall guard, run by one thread (CDIIS, guard) or one
// wavefront (A-DIIS) of whichever kernel owns the molecule's SCF.
#pragma once

#include <hip/hip_runtime.h>

#include "pm7_device.h"

namespace pm7 {

// ---- small serial solvers (thread 0) -------------------------------------------------------

__device__ bool solve_small(double* m, double* rhs, int n, double* x) {
  for (int col = 0; col < n; ++col) {
    int piv = col;
    double best = fabs(m[col * n + col]);
    for (int row = col + 1; row < n; ++row) {
      double v = fabs(m[row * n + col]);
      if (v > best) {
        best = v;
        piv = row;
      }
    }
    if (best < 1.0e-12) return false;
    if (piv != col) {
      for (int j = 0; j < n; ++j) {
        double t = m[col * n + j];
        m[col * n + j] = m[piv * n + j];
        m[piv * n + j] = t;
      }
      double t = rhs[col];
      rhs[col] = rhs[piv];
      rhs[piv] = t;
    }
    for (int row = col + 1; row < n; ++row) {
      double factor = m[row * n + col] / m[col * n + col];
      if (factor == 0.0) continue;
      for (int j = col; j < n; ++j) m[row * n + j] -= factor * m[col * n + j];
      rhs[row] -= factor * rhs[col];
    }
  }
  for (int col = n - 1; col >= 0; --col) {
    double sum = rhs[col];
    for (int j = col + 1; j < n; ++j) sum -= m[col * n + j] * x[j];
    x[col] = sum / m[col * n + col];
  }
  return true;
}

// CDIIS over the chronological window (pm7-rs src/scf.rs:1684-1771), run by one thread with its
// <= 9x9 bordered system in the LDS scratch `wk` (b: 81, rhs: 9, x: 9). Returns the number of
// coefficients (0 on failure); coefficient i belongs to chronological entry skip + i.
__device__ int cdiis(int cnt, const int* slot, const double* G, double* coef, int* skip_out, double* wk) {
  if (cnt < 2) return 0;
  double* b = wk;
  double* rhs = wk + (HIST + 1) * (HIST + 1);
  double* x = rhs + HIST + 1;
  for (int skip = 0; skip < cnt - 1; ++skip) {
    const int m = cnt - skip, dim = m + 1;
    double scale = 0.0;
    for (int i = 0; i < m; ++i) scale = fmax(scale, fabs(G[slot[skip + i] * HIST + slot[skip + i]]));
    if (!isfinite(scale) || scale <= 0.0) continue;
    for (int i = 0; i < dim * dim; ++i) b[i] = 0.0;
    for (int i = 0; i < m; ++i) {
      for (int j = 0; j < m; ++j) b[i * dim + j] = G[slot[skip + i] * HIST + slot[skip + j]] / scale;
      b[i * dim + m] = -1.0;
      b[m * dim + i] = -1.0;
      rhs[i] = 0.0;
    }
    rhs[m] = -1.0;
    if (!solve_small(b, rhs, dim, x)) continue;
    bool finite = true;
    for (int i = 0; i < dim; ++i) finite = finite && isfinite(x[i]);
    if (!finite) continue;
    double sum = 0.0;
    for (int i = 0; i < m; ++i) sum += fabs(x[i]);
    if (sum <= 25.0) {
      for (int i = 0; i < m; ++i) coef[i] = x[i];
      *skip_out = skip;
      return m;
    }
  }
  return 0;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// A-DIIS on the simplex (pm7-rs src/scf.rs:1812-1916) from M[s][t] = <D_s, F_t>, run by one full
// wavefront: lane i owns coefficient i, its gradient row and its rank in the simplex projection,
// so the 400-step projected-gradient loop needs no memory traffic. Writes `cnt` coefficients.
__device__ void adiis_wave(int cnt, const int* slot, const double* Mm, double* coef) {
  const int lane = threadIdx.x & 63;
  const int n = cnt;
  const bool act = lane < n;
  const int sn = slot[n - 1];
  const double mnn = Mm[sn * HIST + sn];
  double d = 0.0;
  double srow[HIST];  // s[lane][j] + s[j][lane]
#pragma unroll
  for (int j = 0; j < HIST; ++j) srow[j] = 0.0;
  if (act) {
    const int si = slot[lane];
    d = Mm[si * HIST + sn] - mnn;
#pragma unroll
    for (int j = 0; j < HIST; ++j) {
      if (j < n) {
        const int sj = slot[j];
        const double sij = Mm[si * HIST + sj] - Mm[si * HIST + sn] - Mm[sn * HIST + sj] + mnn;
        const double sji = Mm[sj * HIST + si] - Mm[sj * HIST + sn] - Mm[sn * HIST + si] + mnn;
        srow[j] = sij + sji;
      }
    }
  }
  double rowabs = 0.0;
#pragma unroll
  for (int j = 0; j < HIST; ++j) rowabs += fabs(srow[j]);
  const double lr = 1.0 / fmax(1.0e-12, wave_max(act ? rowabs : 0.0));
  double c = (act && lane == n - 1) ? 1.0 : 0.0;
  for (int it = 0; it < 400; ++it) {
    double g = 2.0 * d;
#pragma unroll
    for (int j = 0; j < HIST; ++j) g += srow[j] * __shfl(c, j, 64);
    const double trial = act ? c - lr * g : -INFINITY;
    // Descending rank among the active lanes (ties by lane) and the inclusive prefix sum.
    int rank = 0;
    double css = trial;
#pragma unroll
    for (int j = 0; j < HIST; ++j) {
      const double tj = __shfl(trial, j, 64);
      const bool before = (j < n) && (tj > trial || (tj == trial && j < lane));
      rank += before ? 1 : 0;
      css += before ? tj : 0.0;
    }
    const double t = (css - 1.0) / (double)(rank + 1);
    const bool cond = act && (trial - t > 0.0);
    const double best = wave_max(cond ? (double)rank : -1.0);
    const unsigned long long win = __ballot(cond && (double)rank == best);
    const int wl = win ? (__ffsll((long long)win) - 1) : 0;
    const double theta = win ? __shfl(t, wl, 64) : 0.0;
    const double proj = act ? fmax(trial - theta, 0.0) : 0.0;
    const double delta = wave_sum(act ? fabs(proj - c) : 0.0);
    c = proj;
    if (delta < 1.0e-12) break;
  }
  if (act) coef[lane] = c;
}

// Level-shift stall guard (pm7-rs src/scf.rs:1581-1656); state lives in LDS.
struct Guard {
  double floor_, shift;
  double steps[10], energies[10];
  int n;
};

__device__ void guard_update(Guard& g, double step, double energy) {
  g.steps[g.n] = step;
  g.energies[g.n] = energy;
  g.n++;
  if (g.n < 10) return;
  double sn = INFINITY, sw = INFINITY, en = INFINITY, ew = INFINITY;
  for (int i = 0; i < 5; ++i) {
    sw = fmin(sw, g.steps[i]);
    ew = fmin(ew, g.energies[i]);
    sn = fmin(sn, g.steps[5 + i]);
    en = fmin(en, g.energies[5 + i]);
  }
  bool moved = false;
  if (sn <= sw * 0.1) {
    if (g.shift > g.floor_) {
      double relaxed = fmax(g.shift * 0.5, g.floor_);
      g.shift = relaxed < 0.25 ? g.floor_ : relaxed;
      moved = true;
    }
  } else if (sn > sw * 0.9 && en > ew - 1.0e-9 && g.shift < 64.0) {
    g.shift = g.shift < 1.0 ? 1.0 : fmin(g.shift * 2.0, 64.0);
    moved = true;
  }
  if (moved) {
    g.n = 0;
  } else {
    for (int i = 0; i < 9; ++i) {
      g.steps[i] = g.steps[i + 1];
      g.energies[i] = g.energies[i + 1];
    }
    g.n = 9;
  }
}

}  // namespace pm7

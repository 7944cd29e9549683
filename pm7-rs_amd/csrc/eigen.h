// SPDX-License-Identifier: GPL-3.0-or-later
//
// Workgroup-wide symmetric eigensolver of the SCF megakernel: parallel cyclic Jacobi on a packed
// upper triangle, warm-started from the previous SCF iteration's eigenvectors.
#pragma once

#include <hip/hip_runtime.h>

#include "mfma.h"
#include "wg.h"

namespace pm7 {

// Packed upper-triangular index of the symmetric element (i, j).
__device__ inline int pidx(int i, int j) {
  const int a = i < j ? i : j, b = i < j ? j : i;
  return b * (b + 1) / 2 + a;
}

// Scale of the Jacobi stopping test off^2 <= tol2 * scale: the matrix's total square norm, capped
// at n^2 (100 eV)^2. The cap leaves ordinary Fock matrices (elements of tens of eV) untouched; a
// matrix with a few huge diagonal entries (pseudo-atoms carry 1e4-1e7 eV) would otherwise let the
// off-diagonal residual grow with those entries and leave the ordinary orbitals mixed far beyond
// the density's tolerance.
__device__ inline double jacobi_scale(double total, int n) { return fmin(total, (double)n * (double)n * 1.0e4); }

// Parallel cyclic Jacobi on the packed symmetric n2 x n2 matrix A, accumulating the rotations
// into V (row-major, stride n2). Round-robin pairing gives m = n2/2 disjoint rotations per round;
// the 2x2 block (pair k rows, pair l columns) takes both its row and column rotation in one
// update, so a round is one pass over the packed matrix. Returns the number of sweeps.
//
// The round is bound by LDS traffic and latency, so:
//  * the m(m+1)/2 blocks are dealt out as a folded triangle (row k beside row m-1-k, m+1 blocks
//    per folded row), giving every thread the same count; a thread issues the loads of CH of its
//    blocks before any store (the blocks of a round are disjoint);
//  * the rotation accumulator loads several rows before storing them (more so when V is in L2);
//  * the rotation angle needs two square roots and two divisions:
//      d = a_qq - a_pp, x = 2 a_pq, r = sqrt(d^2 + x^2), t = sgn(d) x / (|d| + r),
//      c = 1 / sqrt(1 + t^2) = sqrt((|d| + r) / (2 r)), s = t c.
//
// TA is the storage type of A: double, or float for the reduced-precision phase of a mixed run.
// The angles are always formed in FP64 (from the stored values), and V always accumulates in
// FP64, so the rotations stay orthogonal to FP64 rounding whatever TA is. The sweeps stop when
// off-diagonal^2 <= tol2 * jacobi_scale(total^2), or after max_sweeps.
// V is split by rows: rows [0, vrows) at Vl (LDS), the rest at Vg (L2/HBM), stride n2 in both.
template <int NT, int MAXN, bool V_LDS, typename TA>
__device__ int jacobi_core(int n2, TA* A, double* Vl, double* Vg, int vrows, double2* s_cs, double* s_t, int* s_pq,
                           double* red, double tol2, int max_sweeps) {
  constexpr int NX = 32, NY = NT / NX;  // 2-D thread grid: compile-time divisors only
  constexpr int CH = 2;                 // blocks in flight per thread
  constexpr int VR = V_LDS ? 4 : 8;     // accumulator rows in flight per thread
  const int tx = threadIdx.x % NX, ty = threadIdx.x / NX;
  const int m = n2 / 2;
  // Folded block grid: (m+1)/2 folded rows of m+1 blocks; thread t takes t, t+NT, ...
  const int w1 = m + 1, rows = (m + 1) / 2, total = rows * w1;
  const int rstep = NT / w1, cstep = NT - rstep * w1;
  const int R0 = threadIdx.x / w1, C0 = threadIdx.x - R0 * w1;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    double sums[2] = {0.0, 0.0};
    for (int j = ty; j < n2; j += NY) {
      const TA* rowj = A + j * (j + 1) / 2;
      for (int i = tx; i <= j; i += NX) {
        const double v = (double)rowj[i];
        if (i == j) sums[1] += v * v;
        else sums[0] += 2.0 * v * v;
      }
    }
    wg_sum_n<NT, 2>(sums, red);
    if (sums[0] <= tol2 * jacobi_scale(sums[0] + sums[1], n2)) break;
    for (int r = 0; r < n2 - 1; ++r) {
      if (threadIdx.x < m) {
        // Round-robin pairing: position 0 stays, positions 1..n2-1 turn by one each round.
        const int k = threadIdx.x;
        int a = k - 1 + r;
        a -= (a >= n2 - 1) ? n2 - 1 : 0;
        const int p = k == 0 ? 0 : 1 + a;
        int b = n2 - 2 - k + r;
        b -= (b >= n2 - 1) ? n2 - 1 : 0;
        const int q = 1 + b;
        const double apq = (double)A[pidx(p, q)];
        double cc = 1.0, ss = 0.0, tt = 0.0;
        if (apq != 0.0) {
          const double d = (double)A[pidx(q, q)] - (double)A[pidx(p, p)];
          const double x = 2.0 * apq;
          const double rr = sqrt(d * d + x * x);
          const double den = fabs(d) + rr;
          tt = (d >= 0.0 ? x : -x) / den;
          cc = sqrt(den / (2.0 * rr));
          ss = tt * cc;
        }
        s_pq[k] = p | (q << 16);  // n2 <= 1024
        s_cs[k] = make_double2(cc, ss);
        s_t[k] = tt;
      }
      __syncthreads();
      {
        int R = R0, col = C0;
        for (int lin0 = threadIdx.x; lin0 < total; lin0 += CH * NT) {
          int i0[CH], i1[CH], i2[CH], i3[CH];
          TA x0[CH], x1[CH], x2[CH], x3[CH], ck[CH], sk[CH], cl[CH], sl[CH];
          bool act[CH], dg[CH];
#pragma unroll
          for (int u = 0; u < CH; ++u) {
            int k = -1, l = -1;
            if (lin0 + u * NT < total) {
              if (col < m - R) {
                k = R;
                l = R + col;
              } else {
                k = m - 1 - R;
                l = k + (col - (m - R));
                // odd m: the middle folded row would hold its own row twice; keep the first copy.
                if ((m & 1) && R == rows - 1) k = -1;
              }
            }
            R += rstep;
            col += cstep;
            if (col >= w1) {
              col -= w1;
              ++R;
            }
            act[u] = k >= 0;
            dg[u] = k == l;
            if (act[u]) {
              const int pqk = s_pq[k], pql = s_pq[l];
              const int pk = pqk & 0xffff, qk = pqk >> 16, pl = pql & 0xffff, ql = pql >> 16;
              const double2 csk = s_cs[k], csl = s_cs[l];
              ck[u] = (TA)csk.x;
              sk[u] = (TA)csk.y;
              cl[u] = (TA)csl.x;
              sl[u] = (TA)csl.y;
              if (dg[u]) {
                sl[u] = (TA)s_t[k];  // diagonal block: a_pp -= t a_pq, a_qq += t a_pq, a_pq = 0
                i0[u] = pidx(pk, qk);
                i1[u] = pidx(pk, pk);
                i2[u] = pidx(qk, qk);
                i3[u] = i0[u];
                act[u] = sl[u] != (TA)0;
              } else {
                i0[u] = pidx(pk, pl);
                i1[u] = pidx(pk, ql);
                i2[u] = pidx(qk, pl);
                i3[u] = pidx(qk, ql);
                act[u] = !(sk[u] == (TA)0 && sl[u] == (TA)0);
              }
              if (act[u]) {
                x0[u] = A[i0[u]];
                x1[u] = A[i1[u]];
                x2[u] = A[i2[u]];
                x3[u] = A[i3[u]];
              }
            }
          }
#pragma unroll
          for (int u = 0; u < CH; ++u) {
            if (!act[u]) continue;
            if (dg[u]) {
              const TA tk = sl[u];
              A[i1[u]] = x1[u] - tk * x0[u];
              A[i2[u]] = x2[u] + tk * x0[u];
              A[i0[u]] = (TA)0;
            } else {
              const TA y00 = ck[u] * x0[u] - sk[u] * x2[u], y01 = ck[u] * x1[u] - sk[u] * x3[u];
              const TA y10 = sk[u] * x0[u] + ck[u] * x2[u], y11 = sk[u] * x1[u] + ck[u] * x3[u];
              A[i0[u]] = cl[u] * y00 - sl[u] * y01;
              A[i1[u]] = sl[u] * y00 + cl[u] * y01;
              A[i2[u]] = cl[u] * y10 - sl[u] * y11;
              A[i3[u]] = sl[u] * y10 + cl[u] * y11;
            }
          }
        }
      }
      // Rotation accumulator: column pairs over tx, rows over ty, VR rows loaded before any store;
      // the rows held in LDS and the rest are separate passes, so each pass addresses one
      // memory space directly.
      auto rotate_rows = [&](double* __restrict__ V, int r0, int r1) {
        for (int k = tx; k < m; k += NX) {
          const double2 cs = s_cs[k];
          const double cc = cs.x, ss = cs.y;
          if (ss == 0.0) continue;
          const int pq = s_pq[k];
          const int pk = pq & 0xffff, qk = pq >> 16;
          for (int i0 = r0 + ty; i0 < r1; i0 += VR * NY) {
            double vp[VR], vq[VR];
#pragma unroll
            for (int u = 0; u < VR; ++u) {
              const int i = i0 + u * NY;
              vp[u] = i < r1 ? V[i * n2 + pk] : 0.0;
              vq[u] = i < r1 ? V[i * n2 + qk] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < VR; ++u) {
              const int i = i0 + u * NY;
              if (i < r1) {
                V[i * n2 + pk] = cc * vp[u] - ss * vq[u];
                V[i * n2 + qk] = ss * vp[u] + cc * vq[u];
              }
            }
          }
        }
      };
      if (vrows == n2) {
        rotate_rows(Vl, 0, n2);
      } else if (vrows > 0 && n2 - vrows <= NY) {
        // At most one row per thread lives in L2: its loads are issued first and complete
        // while the thread works through its LDS rows.
        const int ig = vrows + ty;
        const bool hg = ig < n2;
        for (int k = tx; k < m; k += NX) {
          const double2 cs = s_cs[k];
          const double cc = cs.x, ss = cs.y;
          if (ss == 0.0) continue;
          const int pq = s_pq[k];
          const int pk = pq & 0xffff, qk = pq >> 16;
          double gp = 0.0, gq = 0.0;
          if (hg) {
            gp = Vg[ig * n2 + pk];
            gq = Vg[ig * n2 + qk];
          }
          for (int i0 = ty; i0 < vrows; i0 += VR * NY) {
            double vp[VR], vq[VR];
#pragma unroll
            for (int u = 0; u < VR; ++u) {
              const int i = i0 + u * NY;
              vp[u] = i < vrows ? Vl[i * n2 + pk] : 0.0;
              vq[u] = i < vrows ? Vl[i * n2 + qk] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < VR; ++u) {
              const int i = i0 + u * NY;
              if (i < vrows) {
                Vl[i * n2 + pk] = cc * vp[u] - ss * vq[u];
                Vl[i * n2 + qk] = ss * vp[u] + cc * vq[u];
              }
            }
          }
          if (hg) {
            Vg[ig * n2 + pk] = cc * gp - ss * gq;
            Vg[ig * n2 + qk] = ss * gp + cc * gq;
          }
        }
      } else {
        if (vrows > 0) rotate_rows(Vl, 0, vrows);
        rotate_rows(Vg, vrows, n2);
      }
      __syncthreads();
    }
  }
  return sweep;
}

// LDS words of a size class's s_a array, and where the block solver runs. The block solver's
// dimension is ld rounded up to 16; the rows and columns past ld are decoupled zeros whose part of
// the rotation accumulator stays the identity, so the accumulator is stored ld x ld either way.
// With the accumulator in LDS the block solver covers the whole class and s_a adds one 16 x 16
// subproblem and its rotation (512 doubles) per block pair after the packed matrix (if that is in
// LDS). With the packed matrix in LDS and the accumulator in L2 it covers dimensions up to 128, its
// subproblems carved from the packed array's unused tail (128 * 129 / 2 + 8 * 512 doubles fit the
// 196 class's array).
template <int MAXN, bool A_LDS, bool V_LDS>
struct EigLds {
  static constexpr int kBMax = V_LDS ? (MAXN + 15) / 16 * 16 : ((A_LDS && MAXN >= 128) ? 128 : 0);
  static constexpr bool kBlock = kBMax > 0;
  static constexpr int kA = A_LDS ? MAXN * (MAXN + 1) / 2 : 0;
  static constexpr int kBQ = (kBlock && V_LDS) ? (kBMax / 16) * 512 : 0;
  static constexpr int kWords = kA + kBQ > 0 ? kA + kBQ : 1;
  static_assert(!(A_LDS && V_LDS) || kBMax == MAXN, "an LDS-resident packed matrix must hold the padded size");
  static_assert(V_LDS || !kBlock || 128 * 129 / 2 + 8 * 512 <= kA, "block subproblems must fit the packed array");
};

// LDS stores of this wave are visible to its own later LDS loads (a wave's LDS operations complete
// in order; the clobber keeps the compiler from moving memory accesses across).
__device__ inline void wave_lds_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

// A pointer into LDS typed as such, so accesses compile to LDS instructions with 32-bit addresses.
typedef __attribute__((address_space(3))) double lds_f64;
__device__ inline lds_f64* as_lds(double* p) { return (lds_f64*)p; }

// Blocks (I, J) of block pair k in round r of a round-robin over nb blocks (the pairing of
// jacobi_core, one level up).
__device__ inline void rr_pair(int nb, int r, int k, int& I, int& J) {
  int a = k - 1 + r;
  a -= (a >= nb - 1) ? nb - 1 : 0;
  I = k == 0 ? 0 : 1 + a;
  int b = nb - 2 - k + r;
  b -= (b >= nb - 1) ? nb - 1 : 0;
  J = 1 + b;
}

// The 2 x 2 rotation of jacobi_core that annihilates a_pq (c = 1, s = t = 0 when a_pq is 0).
__device__ inline void jacobi_angle(double app, double aqq, double apq, double& c, double& s, double& t) {
  c = 1.0;
  s = 0.0;
  t = 0.0;
  if (apq != 0.0) {
    const double d = aqq - app;
    const double x = 2.0 * apq;
    const double r = sqrt(d * d + x * x);
    const double den = fabs(d) + r;
    t = (d >= 0.0 ? x : -x) / den;
    c = sqrt(den / (2.0 * r));
    s = t * c;
  }
}

// The same rotation from the hardware reciprocal square root and reciprocal, each refined by two
// Newton steps (about 27 FP64 operations instead of the ~60 of the correctly rounded square roots
// and divisions): t is within a few ulps of jacobi_angle's, and c = (1 + t^2)^-1/2, s = t c keep
// c^2 + s^2 = 1 to rounding whatever t is, so the rotation stays orthogonal. Arguments outside the
// range where the squares neither overflow nor lose precision take the exact formula.
__device__ inline void jacobi_angle_fast(double app, double aqq, double apq, double& c, double& s, double& t) {
  c = 1.0;
  s = 0.0;
  t = 0.0;
  if (apq == 0.0) return;
  const double d = aqq - app;
  const double x = 2.0 * apq;
  const double h = fma(d, d, x * x);
  if (!(h > 1.0e-280 && h < 1.0e280)) {
    jacobi_angle(app, aqq, apq, c, s, t);
    return;
  }
  // sqrt(h): g ~ h / sqrt(h), hh ~ 1 / (2 sqrt(h)), two coupled Newton steps.
  const double y = __builtin_amdgcn_rsq(h);
  double g = h * y, hh = 0.5 * y;
  const double e0 = fma(-g, hh, 0.5);
  g = fma(g, e0, g);
  hh = fma(hh, e0, hh);
  const double d0 = fma(-g, g, h);
  g = fma(d0, hh, g);
  const double den = fabs(d) + g;
  double z = __builtin_amdgcn_rcp(den);
  z = fma(z, fma(-den, z, 1.0), z);
  z = fma(z, fma(-den, z, 1.0), z);
  t = (d >= 0.0 ? x : -x) * z;
  const double u = fma(t, t, 1.0);  // in [1, 2]
  double w = __builtin_amdgcn_rsq(u);
  const double hu = 0.5 * u;
  w = w * fma(-hu, w * w, 1.5);
  w = w * fma(-hu, w * w, 1.5);
  c = w;
  s = t * c;
}

// Block Jacobi on the packed symmetric N x N matrix A (N = 16 P; indices past the matrix are
// decoupled zero rows and stay exactly zero), accumulating into V (N x N, row-major, in LDS).
// The indices form 2P blocks of 8; a round pairs the blocks round-robin, and block pair k owns
// the 16 indices S_k (its two blocks). Per round:
//  1. wave k < P rotates its 16 x 16 diagonal block in LDS (the 2 x 2 update of jacobi_core),
//     accumulating the rotations into Q_k, and writes the rotated block back to A: the first round
//     of a sweep pairs blocks 2k, 2k+1 and turns only the pairs inside each 8-block (7 inner rounds
//     of a round-robin over 8), every later round turns only the 64 pairs across its two blocks
//     (8 inner rounds, rotation j = (j, 8 + (j + rr) mod 8)), so a sweep turns every index pair
//     once, in 7 + 8 (2P - 1) inner rounds of 8 disjoint rotations. Lane (jr, jc) = (lane / 8, lane % 8) owns
//     the 2 x 2 block of rotation pairs jr, jc (jr <= jc) and rows jr, jr + 8 of Q in pair jc's
//     columns, and forms the angles of both of its pairs itself (every lane gets the same bits),
//     so an inner round is one LDS read pass, the angles, one write pass and one wave-level wait;
//  2. every off-diagonal block becomes A[S_k, S_l] <- Q_k^T A[S_k, S_l] Q_l and every 16-row tile
//     of the accumulator V[R, S_l] <- V[R, S_l] Q_l, FP64 matrix-core products (8 and 4 per item)
//     dealt to the waves; a pair whose sweep rotated nothing has Q = I and its items are skipped.
//     The off-diagonal blocks are not touched by step 1, so each wave loads its blocks' operands
//     before step 1 and the loads complete while the inner sweeps run.
// A round is two workgroup barriers instead of one per index pair, and the bulk of its work is
// dense products on the matrix cores. As in jacobi_core, only exact zeros skip their rotation: a
// rounding-level element inside an exactly degenerate set still turns the set (by 45 degrees), and
// that decides which members an occupied count that splits the set fills.
// bq: P x 512 doubles of LDS (subproblem, then Q); s_flag: P ints. Returns the number of sweeps.
// A and V each sit in LDS or in L2 (A_LDS, V_LDS); operands that live in L2 are loaded before the
// inner sweeps. BMAX: the largest N. V is stored nv x nv (nv <= N): its rows and columns past nv
// belong to the decoupled pad indices and are the identity, read as such and never written.
template <int NT, int BMAX, bool A_LDS, bool V_LDS>
__device__ int bjacobi_core(int N, int nv, double* A, double* V, double* bq, int* s_flag, double* red, double tol2,
                            int max_sweeps) {
  constexpr int NW = NT / 64;
  constexpr int NX = 32, NY = NT / NX;
  constexpr int PMAX = BMAX / 16;
  constexpr int MAXIA = A_LDS ? 1 : (PMAX * (PMAX - 1) / 2 + NW - 1) / NW;  // prefetched off-diagonal blocks
  constexpr int MAXIV = V_LDS ? 1 : (PMAX * PMAX + NW - 1) / NW;            // prefetched accumulator tiles
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r16 = lane & 15, q4 = lane >> 4;
  const int P = N / 16, NB = N / 8;
  const int nA = P * (P - 1) / 2, nV = (N / 16) * P;
  lds_f64* bqs = as_lds(bq);
  auto vload = [&](int i, int j) -> double {
    if (i < nv && j < nv) {
      if constexpr (V_LDS) return as_lds(V)[i * nv + j];
      else return V[i * nv + j];
    }
    return i == j ? 1.0 : 0.0;
  };
  auto vstore = [&](int i, int j, double v) {
    if (i < nv && j < nv) {
      if constexpr (V_LDS) as_lds(V)[i * nv + j] = v;
      else V[i * nv + j] = v;
    }
  };
  auto aload = [&](int idx) -> double {
    if constexpr (A_LDS) return as_lds(A)[idx];
    else return A[idx];
  };
  auto astore = [&](int idx, double v) {
    if constexpr (A_LDS) as_lds(A)[idx] = v;
    else A[idx] = v;
  };
  // Off-diagonal block item -> (k, l), k < l.
  auto item_kl = [&](int it, int& k, int& l) {
    k = 0;
    int t = it;
    while (t >= P - 1 - k) {
      t -= P - 1 - k;
      ++k;
    }
    l = k + 1 + t;
  };
  const int jr = lane >> 3, jc = lane & 7;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    double sums[2] = {0.0, 0.0};
    for (int j = threadIdx.x / NX; j < N; j += NY) {
      const int rowj = j * (j + 1) / 2;
      for (int i = threadIdx.x % NX; i <= j; i += NX) {
        const double v = aload(rowj + i);
        if (i == j) sums[1] += v * v;
        else sums[0] += 2.0 * v * v;
      }
    }
    wg_sum_n<NT, 2>(sums, red);
    const double total = sums[0] + sums[1];
    if (sums[0] <= tol2 * jacobi_scale(total, N)) break;
    for (int r = 0; r < NB; ++r) {
      // Round 0 pairs blocks 2k, 2k+1 (pairs inside each block); rounds 1..NB-1 the round-robin.
      auto blocks = [&](int k, int& I, int& J) {
        if (r == 0) {
          I = 2 * k;
          J = 2 * k + 1;
        } else {
          rr_pair(NB, r - 1, k, I, J);
        }
      };
      // Operands in L2 of this wave's off-diagonal blocks and accumulator tiles (untouched by the
      // inner sweeps).
      double pav[MAXIA][4], pv[MAXIV][4];
#pragma unroll
      for (int m = 0; m < MAXIA; ++m) {
        const int it = wave + m * NW;
        if (!A_LDS && it < nA) {
          int k, l, Ik, Jk, Il, Jl;
          item_kl(it, k, l);
          blocks(k, Ik, Jk);
          blocks(l, Il, Jl);
          const int gi = r16 < 8 ? 8 * Ik + r16 : 8 * Jk + r16 - 8;
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int b = 4 * s + q4;
            pav[m][s] = aload(pidx(gi, b < 8 ? 8 * Il + b : 8 * Jl + b - 8));
          }
        }
      }
#pragma unroll
      for (int m = 0; m < MAXIV; ++m) {
        const int iv = wave + m * NW;
        if (!V_LDS && iv < nV) {
          const int rt = iv / P, l = iv - rt * P;
          int Il, Jl;
          blocks(l, Il, Jl);
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int b = 4 * s + q4;
            pv[m][s] = vload(16 * rt + r16, b < 8 ? 8 * Il + b : 8 * Jl + b - 8);
          }
        }
      }
      if (wave < P) {
        const int k = wave;
        int I, J;
        blocks(k, I, J);
        auto g = [&](int a) { return a < 8 ? 8 * I + a : 8 * J + a - 8; };
        lds_f64* B = bqs + k * 512;
        lds_f64* Q = B + 256;
        double ld4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = lane + 64 * u;
          ld4[u] = aload(pidx(g(e >> 4), g(e & 15)));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = lane + 64 * u;
          B[e] = ld4[u];
          Q[e] = (e >> 4) == (e & 15) ? 1.0 : 0.0;
        }
        wave_lds_sync();
        int any = 0;
        const int n_inner = r == 0 ? 7 : 8;
        for (int rr = 0; rr < n_inner; ++rr) {
          // Rotation pair j = (p_j, q_j) of inner round rr: round 0 runs a round-robin over the 8
          // indices of each block (j < 4 in the first, j >= 4 in the second), later rounds pair
          // index j of the first block with index (j + rr) % 8 of the second.
          auto pq16 = [&](int j, int& p, int& q) {
            if (r == 0) {
              const int h = j & 3, o = j & 4 ? 8 : 0;
              int a = h - 1 + rr;
              a -= (a >= 7) ? 7 : 0;
              p = o + (h == 0 ? 0 : 1 + a);
              int b = 6 - h + rr;
              b -= (b >= 7) ? 7 : 0;
              q = o + 1 + b;
            } else {
              p = j;
              q = 8 + ((j + rr) & 7);
            }
          };
          int pr, qr, pc, qc;
          pq16(jr, pr, qr);
          pq16(jc, pc, qc);
          const double rpp = B[pr * 17], rqq = B[qr * 17], rpq = B[pr * 16 + qr];
          const double cpp = B[pc * 17], cqq = B[qc * 17], cpq = B[pc * 16 + qc];
          double x0 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0;
          if (jr < jc) {
            x0 = B[pr * 16 + pc];
            x1 = B[pr * 16 + qc];
            x2 = B[qr * 16 + pc];
            x3 = B[qr * 16 + qc];
          }
          const double vp0 = Q[jr * 16 + pc], vq0 = Q[jr * 16 + qc];
          const double vp1 = Q[(jr + 8) * 16 + pc], vq1 = Q[(jr + 8) * 16 + qc];
          asm volatile("" ::: "memory");  // every read of the round is issued before any write
          double cr, sr, tr, cc, sc, tc;
          jacobi_angle_fast(rpp, rqq, rpq, cr, sr, tr);
          jacobi_angle_fast(cpp, cqq, cpq, cc, sc, tc);
          any |= __any(cpq != 0.0);
          if (jr == jc) {
            if (tr != 0.0) {
              B[pr * 17] = rpp - tr * rpq;
              B[qr * 17] = rqq + tr * rpq;
              B[pr * 16 + qr] = 0.0;
              B[qr * 16 + pr] = 0.0;
            }
          } else if (jr < jc) {
            if (sr != 0.0 || sc != 0.0) {
              const double y00 = cr * x0 - sr * x2, y01 = cr * x1 - sr * x3;
              const double y10 = sr * x0 + cr * x2, y11 = sr * x1 + cr * x3;
              const double z0 = cc * y00 - sc * y01, z1 = sc * y00 + cc * y01;
              const double z2 = cc * y10 - sc * y11, z3 = sc * y10 + cc * y11;
              B[pr * 16 + pc] = z0;
              B[pc * 16 + pr] = z0;
              B[pr * 16 + qc] = z1;
              B[qc * 16 + pr] = z1;
              B[qr * 16 + pc] = z2;
              B[pc * 16 + qr] = z2;
              B[qr * 16 + qc] = z3;
              B[qc * 16 + qr] = z3;
            }
          }
          if (sc != 0.0) {
            Q[jr * 16 + pc] = cc * vp0 - sc * vq0;
            Q[jr * 16 + qc] = sc * vp0 + cc * vq0;
            Q[(jr + 8) * 16 + pc] = cc * vp1 - sc * vq1;
            Q[(jr + 8) * 16 + qc] = sc * vp1 + cc * vq1;
          }
          wave_lds_sync();
        }
        if (any) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int e = lane + 64 * u, a = e >> 4, b = e & 15;
            if (a <= b) astore(pidx(g(a), g(b)), B[e]);
          }
        }
        if (lane == 0) s_flag[k] = any;
      }
      __syncthreads();
      for (int it = wave, m = 0; it < nA; it += NW, ++m) {
        int k, l;
        item_kl(it, k, l);
        if (!s_flag[k] && !s_flag[l]) continue;
        int Ik, Jk, Il, Jl;
        blocks(k, Ik, Jk);
        blocks(l, Il, Jl);
        auto gk = [&](int a) { return a < 8 ? 8 * Ik + a : 8 * Jk + a - 8; };
        auto gl = [&](int a) { return a < 8 ? 8 * Il + a : 8 * Jl + a - 8; };
        const lds_f64* Qk = bqs + k * 512 + 256;
        const lds_f64* Ql = bqs + l * 512 + 256;
        double av[4], qlv[4], qkv[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if constexpr (A_LDS) {
            av[s] = aload(pidx(gk(r16), gl(4 * s + q4)));
          } else {
            // m < MAXIA: item it = wave + m NW < nA <= MAXIA NW.
#pragma unroll
            for (int mm = 0; mm < MAXIA; ++mm)
              if (mm == m) av[s] = pav[mm][s];
          }
          qlv[s] = Ql[(4 * s + q4) * 16 + r16];
          qkv[s] = Qk[(4 * s + q4) * 16 + r16];
        }
        // X = A[S_k, S_l] Q_l, then Y = Q_k^T X (X's accumulator is the second product's B operand).
        dbl4 x = {0.0, 0.0, 0.0, 0.0}, y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; ++s) x = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], qlv[s], x, 0, 0, 0);
#pragma unroll
        for (int s = 0; s < 4; ++s) y = __builtin_amdgcn_mfma_f64_16x16x4f64(qkv[s], x[s], y, 0, 0, 0);
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) astore(pidx(gk(tile_row_f64(q4, t4)), gl(r16)), y[t4]);
      }
      for (int iv = wave, m = 0; iv < nV; iv += NW, ++m) {
        const int rt = iv / P, l = iv - rt * P;
        if (!s_flag[l]) continue;
        int Il, Jl;
        blocks(l, Il, Jl);
        auto gl = [&](int a) { return a < 8 ? 8 * Il + a : 8 * Jl + a - 8; };
        const lds_f64* Ql = bqs + l * 512 + 256;
        const int i0 = 16 * rt;
        double av[4], qlv[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if constexpr (V_LDS) {
            av[s] = vload(i0 + r16, gl(4 * s + q4));
          } else {
            // m < MAXIV: tile iv = wave + m NW < nV <= MAXIV NW.
#pragma unroll
            for (int mm = 0; mm < MAXIV; ++mm)
              if (mm == m) av[s] = pv[mm][s];
          }
          qlv[s] = Ql[(4 * s + q4) * 16 + r16];
        }
        dbl4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; ++s) x = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], qlv[s], x, 0, 0, 0);
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) vstore(i0 + tile_row_f64(q4, t4), gl(r16), x[t4]);
      }
      __syncthreads();
    }
  }
  return sweep;
}

// Relative width of an eigenvalue cluster: eigenvalues whose neighbouring gaps are all below
// TIE_REL * max(1, |eps|) eV count as one degenerate set.
constexpr double TIE_REL = 1.0e-10;

// Eigenpairs of the symmetric ld x ld matrix W (pad rows/columns zero): ascending eigenvalues
// eps[0..n) and eigenvectors as the columns of Cout. With Cprev (the previous SCF iteration's
// eigenvectors) the solve is warm-started: Jacobi diagonalises Cprev^T W Cprev, which is nearly
// diagonal once the SCF settles, and Cout = Cprev V. A and V sit in LDS when the variant allows.
// TA, tol2 and max_sweeps are passed to jacobi_core (the products around it are FP64 either way).
//
// Degenerate sets (tie_mode). 0 (RHF): eigenvalues ascending, exact ties by Jacobi slot. 1 and 2
// (UHF alpha and beta): inside a set of eigenvalues that differ by rounding only (a cluster,
// TIE_REL) the order of the vectors is decided by slot, not by the last bits of the eigenvalues:
// ascending for alpha, descending for beta. It matters only where the occupied count splits such a
// set: an open shell whose odd electrons of both spins fall into degenerate orbitals (O2 2+
// triplet: one alpha pi* and one beta pi electron). Opposite slot orders put the two into different
// members of their sets; with both in the same member the SCF ends 0.99 eV higher on O2 2+. Only a
// cold solve (no Cprev: the first SCF iteration) uses the descending order; a warm-started solve
// keeps ascending slot order, i.e. the previous iteration's order, so a set cannot swap back and
// forth. RHF keeps the plain order: a closed shell whose frontier splits a degenerate set (the
// D3h frontier of CfF3) then settles the reference's state. Eigenvalues inside a cluster may be
// out of ascending order by less than its width.
template <int NT, int MAXN, bool A_LDS, bool V_LDS, typename TA>
__device__ __attribute__((noinline)) int eigensolve(int n, int ld, const double* __restrict__ W, const double* __restrict__ Cprev,
                          double* __restrict__ Cout, double* __restrict__ eps, double* Ag, double* Vg,
                          double* __restrict__ T, double* s_a, double* s_v, double2* s_cs, double* s_t,
                          int* s_perm, int* s_pq, double* red, unsigned long long* tsub, double tol2,
                          int max_sweeps, int tie_mode, bool block) {
  const int n2 = ld;
  unsigned long long t0 = wall_clock64();
  TA* A = reinterpret_cast<TA*>(A_LDS ? s_a : Ag);
  // Block Jacobi where the size class offers it (FP64 matrix; an L2-resident packed matrix of the
  // padded size must fit the molecule's ld x ld slot): the Jacobi dimension is ld rounded up to 16.
  using Lds = EigLds<MAXN, A_LDS, V_LDS>;
  const int n16 = (n2 + 15) & ~15;
  const bool blk = Lds::kBlock && sizeof(TA) == 8 && block && n16 <= Lds::kBMax &&
                   (A_LDS || (int64_t)n16 * (n16 + 1) / 2 <= (int64_t)ld * ld);
  const int nj = blk ? n16 : n2;  // Jacobi dimension (pad rows and columns are zero)
  // The block solver's subproblems: after the packed matrix in s_a (V in LDS: after the class's
  // whole packed array; V in L2: after this matrix's).
  double* bq = V_LDS ? s_a + Lds::kA : s_a + ((n16 * (n16 + 1) / 2 + 1) & ~1);
  // Rotation accumulator: all in LDS (V_LDS), or as many leading rows as fit in the LDS the packed
  // matrix leaves free (A_LDS: the buffer holds MAXN(MAXN+1)/2 doubles), the rest in Vg.
  double* Vl = V_LDS ? s_v : (A_LDS ? s_a + (n2 * (n2 + 1) / 2 + 1) / 2 * 2 : nullptr);
  int vrows = n2;
  if (!V_LDS) {
    const int free_words = A_LDS ? MAXN * (MAXN + 1) / 2 - (n2 * (n2 + 1) / 2 + 1) / 2 * 2 : 0;
    vrows = free_words > 0 ? min(n2, free_words / n2) : 0;
  }
  if (blk) vrows = V_LDS ? n2 : 0;  // the whole n2 x n2 accumulator in LDS, or all of it at Vg
  auto vat = [&](int i, int j) -> double& { return i < vrows ? Vl[i * n2 + j] : Vg[i * n2 + j]; };
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
  const int nb = (n + 31) >> 5;  // 32 x 32 MFMA blocks (their 32 nb rows cover n2)
  if (Cprev) {
    // T = W Cprev (W symmetric: W(i, k) read as W(k, i), so both operands stream along rows).
    for (int t = wave; t < nb * nb; t += NW) {
      const int I = t / nb, J = t - I * nb, i0 = 32 * I, j0 = 32 * J;
      auto w_ik = [&](int i, int k) { return (i < n && k < n) ? W[(int64_t)k * ld + i] : 0.0; };
      auto c_kj = [&](int k, int j) { return (k < n && j < n) ? Cprev[(int64_t)k * ld + j] : 0.0; };
      dbl4 x[4];
      block32_f64(n, i0, j0, w_ik, c_kj, x);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int i = i0 + 16 * (s >> 1) + tile_row_f64(q, u), j = j0 + 16 * (s & 1) + r;
          if (i < n && j < n) T[(int64_t)i * ld + j] = x[s][u];
        }
    }
    __syncthreads();
    // A = Cprev^T T, upper triangle (pad rows and columns up to n2 are zero).
    for (int t = wave; t < nb * nb; t += NW) {
      const int I = t / nb, J = t - I * nb;
      if (I > J) continue;
      const int i0 = 32 * I, j0 = 32 * J;
      auto c_ik = [&](int i, int k) { return (i < n && k < n) ? Cprev[(int64_t)k * ld + i] : 0.0; };
      auto t_kj = [&](int k, int j) { return (k < n && j < n) ? T[(int64_t)k * ld + j] : 0.0; };
      dbl4 x[4];
      block32_f64(n, i0, j0, c_ik, t_kj, x);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int i = i0 + 16 * (s >> 1) + tile_row_f64(q, u), j = j0 + 16 * (s & 1) + r;
          if (i <= j && j < nj) A[pidx(i, j)] = (i < n && j < n) ? (TA)x[s][u] : (TA)0;
        }
    }
  } else {
    constexpr int NX = 32, NY = NT / NX;
    for (int j = threadIdx.x / NX; j < nj; j += NY)
      for (int i = threadIdx.x % NX; i <= j; i += NX)
        A[j * (j + 1) / 2 + i] = j < n2 ? (TA)W[(int64_t)i * ld + j] : (TA)0;
  }
  for (int idx = threadIdx.x; idx < n2 * n2; idx += NT) {
    const int i = idx / n2, j = idx - i * n2;
    vat(i, j) = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  unsigned long long t1 = wall_clock64();
  int sweeps;
  if constexpr (Lds::kBlock && sizeof(TA) == 8) {
    if (blk)
      sweeps = bjacobi_core<NT, Lds::kBMax, A_LDS, V_LDS>(nj, n2, reinterpret_cast<double*>(A), V_LDS ? Vl : Vg,
                                                          bq, s_perm, red, tol2, max_sweeps);
    else
      sweeps = jacobi_core<NT, MAXN, V_LDS, TA>(n2, A, Vl, Vg, vrows, s_cs, s_t, s_pq, red, tol2, max_sweeps);
  } else {
    sweeps = jacobi_core<NT, MAXN, V_LDS, TA>(n2, A, Vl, Vg, vrows, s_cs, s_t, s_pq, red, tol2, max_sweeps);
  }
  unsigned long long t2 = wall_clock64();
  for (int i = threadIdx.x; i < n; i += NT) {
    const double di = (double)A[pidx(i, i)];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = (double)A[pidx(j, j)];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    s_perm[rank] = i;
  }
  __syncthreads();
  {
    // Any neighbours closer than the cluster width? (s_pq is free until the next Jacobi round.)
    int near = 0;
    for (int k = 1 + threadIdx.x; k < n; k += NT) {
      const double lo = (double)A[pidx(s_perm[k - 1], s_perm[k - 1])], hi = (double)A[pidx(s_perm[k], s_perm[k])];
      near |= (hi - lo <= TIE_REL * fmax(1.0, fabs(hi))) ? 1 : 0;
    }
    if (threadIdx.x == 0) s_pq[0] = 0;
    __syncthreads();
    if (near && tie_mode != 0) s_pq[0] = 1;  // every writer stores the same value
    __syncthreads();
    const bool desc = tie_mode == 2 && !Cprev;
    if (s_pq[0] && threadIdx.x == 0) {
      // Order each cluster by slot (insertion sort; clusters hold a few vectors).
      int k = 0;
      while (k < n) {
        int e = k + 1;
        while (e < n) {
          const double lo = (double)A[pidx(s_perm[e - 1], s_perm[e - 1])], hi = (double)A[pidx(s_perm[e], s_perm[e])];
          if (hi - lo > TIE_REL * fmax(1.0, fabs(hi))) break;
          ++e;
        }
        for (int a = k + 1; a < e; ++a) {
          const int v = s_perm[a];
          int b = a;
          while (b > k && (desc ? s_perm[b - 1] < v : s_perm[b - 1] > v)) {
            s_perm[b] = s_perm[b - 1];
            --b;
          }
          s_perm[b] = v;
        }
        k = e;
      }
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < n; k += NT) eps[k] = (double)A[pidx(s_perm[k], s_perm[k])];
  if (Cprev) {
    // Cout = Cprev V(:, perm); pad rows and columns up to ld are written as zero.
    for (int t = wave; t < nb * nb; t += NW) {
      const int I = t / nb, J = t - I * nb, i0 = 32 * I, j0 = 32 * J;
      auto c_ik = [&](int i, int k) { return (i < n && k < n) ? Cprev[(int64_t)i * ld + k] : 0.0; };
      auto v_kj = [&](int k, int j) { return (k < n && j < n) ? vat(k, s_perm[j]) : 0.0; };
      dbl4 x[4];
      block32_f64(n, i0, j0, c_ik, v_kj, x);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int i = i0 + 16 * (s >> 1) + tile_row_f64(q, u), j = j0 + 16 * (s & 1) + r;
          if (i < ld && j < ld) Cout[(int64_t)i * ld + j] = (i < n && j < n) ? x[s][u] : 0.0;
        }
    }
  } else {
    for (int idx = threadIdx.x; idx < n2 * n2; idx += NT) {
      const int i = idx / n2, k = idx - i * n2;
      Cout[idx] = (i < n && k < n) ? vat(i, s_perm[k]) : 0.0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    tsub[0] += t1 - t0;
    tsub[1] += t2 - t1;
    tsub[2] += wall_clock64() - t2;
  }
  return sweeps;
}

// Occupied-virtual rotation of the orbitals without a full eigensolve (the reduced-precision
// phase of a mixed run): with F' = Cp^T W Cp, each occupied i / virtual a pair gets the 2x2
// Jacobi angle that annihilates F'_ia (the formula of jacobi_core), all at once as
// U = exp(K), K_ia = theta, K_ai = -theta (third-order Taylor), and Cn = Cp U is brought back to
// orthonormal columns by two Newton-Schulz steps C <- C (3 I - C^T C) / 2. Occupied orbitals stay
// the first nocc columns. Every product is an MFMA pass; scratch: Ts, Ag, Vg (ld x ld each), W
// (overwritten) and Cp (overwritten). Stationary points of the SCF are the same as with the
// full eigensolve: at self-consistency F'_ia = 0 and the step is the identity.
template <int NT>
__device__ __attribute__((noinline)) void occ_virt_rotate(int n, int ld, int nocc, double* __restrict__ W, double* __restrict__ Cp,
                                double* __restrict__ Cn, double* __restrict__ Ts, double* __restrict__ Ag,
                                double* __restrict__ Vg) {
  auto rowk = [&](const double* X) {
    return [=](int i, int k) { return (i < n && k < n) ? X[(int64_t)k * ld + i] : 0.0; };  // X(i,k), X symmetric
  };
  auto colk = [&](const double* X) {
    return [=](int k, int j) { return (k < n && j < n) ? X[(int64_t)k * ld + j] : 0.0; };  // X(k,j)
  };
  auto trans = [&](const double* X) {
    return [=](int i, int k) { return (i < n && k < n) ? X[(int64_t)k * ld + i] : 0.0; };  // X^T(i,k)
  };
  auto rowm = [&](const double* X) {
    return [=](int i, int k) { return (i < n && k < n) ? X[(int64_t)i * ld + k] : 0.0; };  // X(i,k)
  };
  auto put = [&](double* X) { return [=](int i, int j, double v) { X[(int64_t)i * ld + j] = v; }; };
  // T = W Cp; F' = Cp^T T.
  gemm_f64<NT>(n, n, rowk(W), colk(Cp), put(Ts));
  __syncthreads();
  gemm_f64<NT>(n, n, trans(Cp), colk(Ts), put(Ag));
  __syncthreads();
  // K into W.
  for (int idx = threadIdx.x; idx < n * n; idx += NT) {
    const int i = idx / n, j = idx - i * n;
    double k = 0.0;
    const bool iv = i < nocc, jv = j < nocc;
    if (iv != jv) {
      const int o = iv ? i : j, v = iv ? j : i;
      const double apq = Ag[(int64_t)o * ld + v];
      if (apq != 0.0) {
        const double d = Ag[(int64_t)v * ld + v] - Ag[(int64_t)o * ld + o];
        const double x = 2.0 * apq;
        const double rr = sqrt(d * d + x * x);
        const double th = atan((d >= 0.0 ? x : -x) / (fabs(d) + rr));
        k = iv ? th : -th;
      }
    }
    W[(int64_t)i * ld + j] = k;
  }
  __syncthreads();
  // K^2 into Vg; U = I + K + K^2/2 + K K^2/6 into Ts.
  gemm_f64<NT>(n, n, rowm(W), colk(W), put(Vg));
  __syncthreads();
  gemm_f64<NT>(n, n, rowm(W), colk(Vg), [&](int i, int j, double v) {
    const int64_t x = (int64_t)i * ld + j;
    Ts[x] = (i == j ? 1.0 : 0.0) + W[x] + 0.5 * Vg[x] + v / 6.0;
  });
  __syncthreads();
  // C1 = Cp U into Ag.
  gemm_f64<NT>(n, n, rowm(Cp), colk(Ts), put(Ag));
  __syncthreads();
  // Two Newton-Schulz steps: Ag -> Cp -> Cn.
  double* src = Ag;
  double* dst = Cp;
  for (int step = 0; step < 2; ++step) {
    gemm_f64<NT>(n, n, trans(src), colk(src), put(Vg));  // S = C^T C
    __syncthreads();
    for (int idx = threadIdx.x; idx < n * n; idx += NT) {
      const int i = idx / n, j = idx - i * n;
      const int64_t x = (int64_t)i * ld + j;
      Vg[x] = (i == j ? 1.5 : 0.0) - 0.5 * Vg[x];
    }
    __syncthreads();
    gemm_f64<NT>(n, n, rowm(src), colk(Vg), put(dst));
    __syncthreads();
    src = dst;
    dst = Cn;
  }
}

}  // namespace pm7

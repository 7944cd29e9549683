// SPDX-License-Identifier: GPL-3.0-or-later
//
// The SCF megakernel: one workgroup owns one molecule for its whole SCF, so a batch of
// molecules of uneven size is ONE launch and every molecule converges (and stops) on its own.
// Per iteration (pm7-rs src/scf.rs:1231-1295 RHF, 1380-1470 UHF):
//
//   F = H + G(P)              J as an owner-computes gather per (atom, packed orbital pair),
//                             K per atom pair, one-centre J-K on the atom's own block
//   E = 1/2 <P, H + F>        fixed-order tree reduction
//   e = FP - (FP)^T, |e|      2x2 register-tiled product, written into the DIIS ring
//   A-DIIS / CDIIS            history dots in one fused pass, the <= 9x9 bordered solve or the
//                             simplex-projected A-DIIS minimisation on thread 0
//   F' = sum c_i F_i (+ shift (1 - P/2) when the stall guard raised one)
//   (eps, C) = eig(F')        parallel cyclic Jacobi (round-robin pairing, 2x2 block updates),
//                             matrix resident in LDS when it fits, eigenvectors in L2
//   P' = 2 C_occ C_occ^T, dP  2x2 register-tiled product with the RMS change fused in
//
// No atomics anywhere: every output element has one owner and every reduction has a fixed
// order, so a molecule's result does not depend on which other molecules share the launch.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "diis.h"
#include "eigen.h"
#include "fock.h"
#include "kernels.h"
#include "mfma.h"
#include "pm7_device.h"
#include "wg.h"

namespace pm7 {


// ||FP - (FP)^T||_F^2; the commutator itself is stored into E when E != nullptr. F and P are
// symmetric, so (FP)^T = PF and E = FP - PF tile by tile on the matrix cores; both products read
// their operands along rows (X(i, k) = X(k, i)). Each wave owns the upper-triangle tiles
// wave, wave + NW, ... and writes E(i, j) and E(j, i) = -E(i, j).
template <int NT>
__device__ __attribute__((noinline)) double commutator(int n, int ld, const double* __restrict__ F, const double* __restrict__ P,
                             double* __restrict__ E, double* red) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
  const int nb = (n + 31) >> 5;  // 32 x 32 blocks (four 16 x 16 tiles each)
  double nrm = 0.0;
  for (int t = wave; t < nb * nb; t += NW) {
    const int I = t / nb, J = t - I * nb;
    if (I > J) continue;
    const int i0 = 32 * I, j0 = 32 * J;
    auto f_ik = [&](int i, int k) { return (i < n && k < n) ? F[(int64_t)k * ld + i] : 0.0; };
    auto p_kj = [&](int k, int j) { return (k < n && j < n) ? P[(int64_t)k * ld + j] : 0.0; };
    auto p_ik = [&](int i, int k) { return (i < n && k < n) ? P[(int64_t)k * ld + i] : 0.0; };
    auto f_kj = [&](int k, int j) { return (k < n && j < n) ? F[(int64_t)k * ld + j] : 0.0; };
    dbl4 fp[4], pf[4];
    block32_f64(n, i0, j0, f_ik, p_kj, fp);
    block32_f64(n, i0, j0, p_ik, f_kj, pf);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + 16 * (s >> 1) + tile_row_f64(q, u), j = j0 + 16 * (s & 1) + r;
        if (i >= n || j >= n) continue;
        const double e = fp[s][u] - pf[s][u];
        if (I == J) {
          nrm += e * e;
          if (E) E[(int64_t)i * ld + j] = e;
        } else {
          nrm += 2.0 * e * e;
          if (E) {
            E[(int64_t)i * ld + j] = e;
            E[(int64_t)j * ld + i] = -e;
          }
        }
      }
  }
  return wg_sum<NT>(nrm, red);
}

// P' = w C_occ C_occ^T (C columns sorted ascending) on the matrix cores, upper-triangle tiles
// mirrored; returns sum (P' - P)^2 over n x n (0 when Pold is null). T is the precision of the product: FP64, or FP32
// inputs with FP32 accumulation in the reduced-precision phase.
template <int NT, typename T>
__device__ __attribute__((noinline)) double density(int n, int ld, int nocc, double wgt, const double* __restrict__ C,
                          const double* __restrict__ Pold, double* __restrict__ Pn, double* red) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
  const int nb = (n + 31) >> 5;  // 32 x 32 blocks
  double dsq = 0.0;
  for (int t = wave; t < nb * nb; t += NW) {
    const int I = t / nb, J = t - I * nb;
    if (I > J) continue;
    const int i0 = 32 * I, j0 = 32 * J;
    double v[4][4];
    int rows[4][4];
    if constexpr (sizeof(T) == 8) {
      auto c_ik = [&](int i, int k) { return (i < n && k < nocc) ? C[(int64_t)i * ld + k] : 0.0; };
      auto c_kj = [&](int k, int j) { return (k < nocc && j < n) ? C[(int64_t)j * ld + k] : 0.0; };
      dbl4 x[4];
      block32_f64(nocc, i0, j0, c_ik, c_kj, x);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          v[s][u] = wgt * x[s][u];
          rows[s][u] = tile_row_f64(q, u);
        }
    } else {
      auto c_ik = [&](int i, int k) { return (i < n && k < nocc) ? (float)C[(int64_t)i * ld + k] : 0.0f; };
      auto c_kj = [&](int k, int j) { return (k < nocc && j < n) ? (float)C[(int64_t)j * ld + k] : 0.0f; };
      flt4 x[4];
      block32_f32(nocc, i0, j0, c_ik, c_kj, x);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          v[s][u] = wgt * (double)x[s][u];
          rows[s][u] = tile_row_f32(q, u);
        }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + 16 * (s >> 1) + rows[s][u], j = j0 + 16 * (s & 1) + r;
        if (i >= n || j >= n) continue;
        Pn[(int64_t)i * ld + j] = v[s][u];
        if (I != J) Pn[(int64_t)j * ld + i] = v[s][u];
        if (!Pold) continue;
        const double d = v[s][u] - Pold[(int64_t)i * ld + j];
        dsq += (I == J) ? d * d : 2.0 * d * d;
      }
  }
  return wg_sum<NT>(dsq, red);
}


template <int NT>
__device__ void level_shift(int n, int ld, double* W, const double* P, double occ, double shift) {
  if (shift <= 0.0) return;
  const double inv = 1.0 / occ;
  for (int idx = threadIdx.x; idx < n * n; idx += NT) {
    int i = idx / n, j = idx - i * n;
    double q = P[(int64_t)i * ld + j] * inv;
    W[(int64_t)i * ld + j] += shift * ((i == j) ? 1.0 - q : -q);
  }
  __syncthreads();
}

// MOPAC diagonal guess (pm7-rs src/scf.rs:1050-1114) for sp atoms, scaled by `scale`.
template <int NT>
__device__ void sad_guess(const MolCtx& c, double* P, double scale) {
  const int ld = c.ld;
  for (int idx = threadIdx.x; idx < ld * ld; idx += NT) P[idx] = 0.0;
  __syncthreads();
  for (int la = threadIdx.x; la < c.M.n_atoms; la += NT) {
    int a = c.M.atom0 + la;
    sad_atom(c.elem[c.atom_z[a]], c.M.sad_yy, scale, P, c.atom_ao[a], ld);
  }
  __syncthreads();
}

template <int NT>
__device__ void copy_mat(int ld, const double* src, double* dst) {
  for (int idx = threadIdx.x; idx < ld * ld; idx += NT) dst[idx] = src[idx];
  __syncthreads();
}

// DIIS history dots of the newest slot h against every live slot t (s_slot[0..cnt)):
// <E_h, E_t>, and with want_m also <D_h, F_t> and <D_t, F_h>, over the padded ld x ld matrices,
// into out[3 q + k] (LDS, written by thread 0). The <D, F> dots feed only A-DIIS, so iterations
// that will not run it read the error matrices alone (a third of the history bytes). A function of
// its own so its 3 x HIST partial sums do not add to the registers of the rest of the megakernel.
template <int NT>
__device__ __attribute__((noinline)) void history_dots(int ld, int cnt, const int* s_slot, const double* __restrict__ E,
                                                       const double* __restrict__ P, const double* __restrict__ F,
                                                       const double* __restrict__ hE, const double* __restrict__ hF,
                                                       const double* __restrict__ hD, bool want_m, double* red,
                                                       double* out) {
  double acc[3 * HIST];
  for (int k = 0; k < 3 * HIST; ++k) acc[k] = 0.0;
  const int64_t sq = (int64_t)ld * ld;
  if (want_m) {
    for (int idx = threadIdx.x; idx < ld * ld; idx += NT) {
      const double eh0 = E[idx], dh = P[idx], fh = F[idx];
#pragma unroll
      for (int q = 0; q < HIST; ++q) {
        if (q < cnt) {
          const int64_t off = (int64_t)s_slot[q] * sq + idx;
          acc[3 * q] += eh0 * hE[off];
          acc[3 * q + 1] += dh * hF[off];
          acc[3 * q + 2] += hD[off] * fh;
        }
      }
    }
  } else {
    for (int idx = threadIdx.x; idx < ld * ld; idx += NT) {
      const double eh0 = E[idx];
#pragma unroll
      for (int q = 0; q < HIST; ++q)
        if (q < cnt) acc[3 * q] += eh0 * hE[(int64_t)s_slot[q] * sq + idx];
    }
  }
  wg_sum_n<NT, 3 * HIST>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3 * HIST; ++k) out[k] = acc[k];
}

// The convergence test of one SCF iteration (thread 0): the caller's (|dE| < e_tol and RMS dP <
// p_tol, or a zero shift and a commutator norm below 1e-7; src/scf.rs:1274-1287). A mixed run's
// FP64 tail asks for MIXED_TAIL times both bounds; once the caller's test is met it goes on
// while it progresses (tail.extra iterations so far, tail.stall of them without a 20 % drop of
// dp below tail.best) and stops, converged, on the third stalled one or after MIXED_TAIL_EXTRA.
struct TailState {
  int extra, stall;
  double best;
};

__device__ inline bool tail_converged(const MolDev& M, int iter, double de, double dp, double err, double shift,
                                      bool low, TailState& tail) {
  const bool resid = shift == 0.0 && err < 1.0e-7;
  const bool plain = iter > 0 && de < M.e_tol && (dp < M.p_tol || resid);
  if (M.precision != 1 || low) return plain;
  const bool tight = iter > 0 && de < M.e_tol &&
                     (dp < M.p_tol * MIXED_TAIL || (shift == 0.0 && err < 1.0e-7 * MIXED_TAIL));
  if (tight) return true;
  if (!plain) return false;
  if (tail.extra++ == 0 || dp < 0.8 * tail.best) {
    tail.best = dp;
    tail.stall = 0;
  } else {
    ++tail.stall;
  }
  return tail.stall >= 3 || tail.extra > MIXED_TAIL_EXTRA;
}

// max |A - B| over the n x n block.
template <int NT>
__device__ double max_abs_diff(int n, int ld, const double* A, const double* B, double* red) {
  double m = 0.0;
  for (int idx = threadIdx.x; idx < n * n; idx += NT) {
    const int i = idx / n, j = idx - i * n;
    m = fmax(m, fabs(A[(int64_t)i * ld + j] - B[(int64_t)i * ld + j]));
  }
  return wg_max<NT>(m, red);
}

// MAXN: largest padded basis of the size class; A_LDS / V_LDS: packed Fock matrix / rotation
// accumulator of the eigensolver resident in LDS (otherwise in the molecule's slab).
template <int NT, int MAXN, bool A_LDS, bool V_LDS>
__global__ void __launch_bounds__(NT)
    k_scf(const int* __restrict__ mol_list, const MolDev* __restrict__ mols, const int* __restrict__ atom_z,
          const int* __restrict__ atom_ao, const int* __restrict__ atom_pp, const PairDev* __restrict__ pairs,
          const ElemDev* __restrict__ elem, const double* __restrict__ ocd, const double* __restrict__ w,
          double* __restrict__ slab,
          double* __restrict__ vecs, MolOut* __restrict__ out) {
  __shared__ double s_a[EigLds<MAXN, A_LDS, V_LDS>::kWords];
  __shared__ double s_v[V_LDS ? MAXN * MAXN : 1];
  __shared__ double s_red[(Wg<NT>::NW + 1) * 3 * HIST];
  __shared__ double2 s_cs[MAXN / 2];  // this round's (cos, sin) per rotation pair
  __shared__ double s_t[MAXN / 2];     // and tan
  __shared__ int s_perm[MAXN];
  __shared__ int s_pq[MAXN / 2];  // this round's rotation pairs, p | q << 16
  __shared__ double s_G[HIST * HIST], s_M[HIST * HIST], s_coef[HIST + 1];
  __shared__ int s_slot[HIST], s_cslot[HIST];
  __shared__ int s_mok[HIST];  // slot's <D, F> dots against the slots before it are in s_M
  __shared__ double s_dots[3 * HIST];
  __shared__ MolOut s_res;                   // the molecule's result record (thread 0 keeps it)
  __shared__ unsigned long long s_tk[NPHASE];  // per-phase device clocks (thread 0)
  __shared__ double s_wk[(HIST + 1) * (HIST + 3)];  // bordered DIIS system
  __shared__ int s_int[8];     // 0 head, 1 cnt, 2 ncoef, 3 done, 4 reduced-precision phase
  __shared__ double s_dbl[8];  // 0 e_old, 1 err, 2 dp
  __shared__ Guard s_guard;
  __shared__ TailState s_tail;

  const int mi = mol_list[blockIdx.x];
  MolCtx c;
  c.M = mols[mi];
  c.atom_z = atom_z;
  c.atom_ao = atom_ao;
  c.atom_pp = atom_pp;
  c.pairs = pairs;
  c.elem = elem;
  c.ocd = ocd;
  c.w = w;
  c.ld = c.M.ld;
  c.base = slab + c.M.mat0;
  c.pp = slab + c.M.pp0;
  const int n = c.M.nao, ld = c.ld;
  double* eps_a = vecs + c.M.vec0;
  double* eps_b = eps_a + ld;
  int sweeps = 0;
  bool last_rot = false;  // the last SCF step rotated the orbitals instead of diagonalising
  if (threadIdx.x == 0) {
    s_guard.floor_ = c.M.shift_floor;
    s_guard.shift = c.M.shift_floor;
    s_guard.n = 0;
    s_int[0] = 0;
    s_int[1] = 0;
    s_int[3] = 0;
    s_tail.extra = 0;  // mixed tail: iterations past the caller's convergence test
    s_tail.stall = 0;
    s_tail.best = 1.0e30;
    s_dbl[0] = 0.0;
    s_dbl[2] = 1.0e30;
    for (int k = 0; k < HIST; ++k) s_mok[k] = 0;
  }
  __syncthreads();
  double* H = c.slot(S_H);
  // The result record and the phase clocks live in LDS and are written by thread 0 only, so they
  // hold no registers through the SCF; e_el is computed by every thread (uniform) and stored last.
  MolOut& res = s_res;
  unsigned long long* tk = s_tk;
  unsigned long long t_mark = wall_clock64();
  const unsigned long long t_start = t_mark;
  // Attribute the time since the last mark to phase `ph` (thread 0 keeps the books).
  auto lap = [&](int ph) {
    const unsigned long long now = wall_clock64();
    if (threadIdx.x == 0) tk[ph] += now - t_mark;
    t_mark = now;
  };
  double e_el = 0.0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < NPHASE; ++k) tk[k] = 0;
    res.converged = 0;
    res.iterations = 0;
    res.dp_err = INFINITY;
    res.err_norm = 0.0;
    res.fp32_iters = 0;
    res.e_fp32_last = 0.0;
    res.check_de = 0.0;
    res.check_dp = 0.0;
  }
  // A mixed run starts in reduced precision only when the core Hamiltonian stays within the range
  // where FP32 rounding (6e-8 relative) is far below the 1e-5 hand-over: max |H_ij| <= 1e3 eV.
  // (Pseudo-atoms such as MOPAC's capped-bond element carry |beta| up to 1e7 eV.)
  {
    double hmax = 0.0;
    const double* H0 = c.slot(S_H);
    for (int idx = threadIdx.x; idx < c.M.nao * c.M.nao; idx += NT) {
      const int i = idx / c.M.nao, j = idx - i * c.M.nao;
      hmax = fmax(hmax, fabs(H0[(int64_t)i * c.ld + j]));
    }
    for (int off = 32; off > 0; off >>= 1) hmax = fmax(hmax, __shfl_down(hmax, off, 64));
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = hmax;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < Wg<NT>::NW; ++w) hmax = fmax(hmax, s_red[w]);
      s_int[4] = (c.M.precision == 1 && hmax <= 1.0e3) ? 1 : 0;
      s_int[5] = 0;         // reduced-precision iterations without progress
      s_dbl[3] = 1.0e30;    // best density change of that phase
    }
    __syncthreads();
  }
  // One FP64 eigensolve: to full accuracy, or, in the reduced-precision phase of a mixed run (the
  // iterations before the density change drops below rotate_below; later ones rotate occupied
  // against virtual orbitals only), to the accuracy the FP32 density supports (off-diagonal norm
  // 1e-7 of the total) in at most mixed_sweeps sweeps.
  const double tol_full = exp10((double)-c.M.eig_tol_exp);
  // tie: 0 RHF, 1 UHF alpha, 2 UHF beta; it sets the order inside degenerate clusters. In the
  // smallest class the reduced-precision cold solve (the first iteration of a mixed run, a few
  // capped sweeps) keeps the scalar rounds: where the occupation splits a degenerate set, the basis
  // inside the set that its capped sweeps leave is the one the run keeps, and the scalar rounds' is
  // the one that settles the reference's state on the small symmetric molecules of the golden set
  // (CfF3). There the packed matrix is in LDS and the scalar rounds cost little; in the larger
  // classes (packed matrix in L2) they would cost as much as the rest of the eigensolves.
  auto eig = [&](bool low, const double* Wm, const double* Cp, double* Co, double* ev, int ag, int vg, int tg,
                 int tie) -> int {
    return eigensolve<NT, MAXN, A_LDS, V_LDS, double>(n, ld, Wm, Cp, Co, ev, c.slot(ag), c.slot(vg), c.slot(tg), s_a,
                                                      s_v, s_cs, s_t, s_perm, s_pq, s_red,
                                                      tk + PH_EIG_IN, low ? 1.0e-14 : tol_full,
                                                      low ? c.M.mixed_sweeps : 40, tie,
                                                      c.M.eig_block != 0 && !(low && Cp == nullptr && MAXN <= 64));
  };

  if (!c.M.uhf) {
    // ------------------------------------------------------------------ RHF
    if (!c.M.init_p) sad_guess<NT>(c, c.slot(S_HD), 1.0);  // else the caller's density is there
    const int nocc = c.M.n_occ;
    int cbuf = 0;  // eigenvector buffer the next solve writes
    int iter = 0;
    for (; iter < c.M.max_scf; ++iter) {
      const int h = s_int[0];
      double* P = c.slot(S_HD + h);
      double* F = c.slot(S_HF + h);
      double* E = c.slot(S_HE + h);
      if (s_int[4]) fock_build<NT, float>(c, P, P, 0.5, nullptr, 0.0, F, nullptr);
      else fock_build<NT, double>(c, P, P, 0.5, nullptr, 0.0, F, nullptr);
      lap(PH_FOCK);
      double eh[2] = {0.0, 0.0};
      for (int idx = threadIdx.x; idx < n * n; idx += NT) {
        int i = idx / n, j = idx - i * n;
        int64_t x = (int64_t)i * ld + j;
        eh[0] += P[x] * H[x];
        eh[1] += P[x] * F[x];
      }
      wg_sum_n<NT, 2>(eh, s_red);
      const double e = 0.5 * (eh[0] + eh[1]);
      const double err = sqrt(commutator<NT>(n, ld, F, P, E, s_red));
      lap(PH_COMM);
      if (threadIdx.x == 0) {
        if (s_int[1] < HIST) s_int[1]++;
        s_int[0] = (h + 1) % HIST;
        const int cnt = s_int[1];
        for (int i = 0; i < cnt; ++i) s_slot[i] = (s_int[0] - cnt + i + 2 * HIST) % HIST;
      }
      __syncthreads();
      const int cnt = s_int[1];
      // History dots: <E_h,E_t> for every live slot t, and <D_h,F_t>, <D_t,F_h> when this iteration
      // runs A-DIIS. A slot added without them is marked; if A-DIIS runs again while it is live,
      // its dots against every live slot are taken then (the same products in the same order, so
      // the same values as if they had been taken on time).
      const bool use_adiis = c.M.use_diis && c.M.accel == 2 && err > c.M.adiis_switch && cnt >= 2;
      {
        history_dots<NT>(ld, cnt, s_slot, E, P, F, c.slot(S_HE), c.slot(S_HF), c.slot(S_HD), use_adiis, s_red,
                         s_dots);
        const double* acc = s_dots;  // thread 0 wrote them and reads them
        if (threadIdx.x == 0) {
          for (int q = 0; q < cnt; ++q) {
            const int t = s_slot[q];
            s_G[h * HIST + t] = acc[3 * q];
            s_G[t * HIST + h] = acc[3 * q];
            if (use_adiis) {
              s_M[h * HIST + t] = acc[3 * q + 1];
              s_M[t * HIST + h] = acc[3 * q + 2];
            }
          }
          s_mok[h] = use_adiis ? 1 : 0;
          int ncoef = 0, skip = 0;
          if (c.M.use_diis && c.M.accel != 0 && (c.M.accel == 1 || err <= c.M.adiis_switch))
            ncoef = cdiis(cnt, s_slot, s_G, s_coef, &skip, s_wk);
          for (int i = 0; i < ncoef; ++i) s_cslot[i] = s_slot[skip + i];
          s_int[2] = ncoef;
          s_dbl[1] = err;
        }
        __syncthreads();
        if (use_adiis) {
          for (int q = 0; q < cnt; ++q) {
            const int t = s_slot[q];
            if (s_mok[t]) continue;  // uniform: thread 0 wrote it before the barrier
            history_dots<NT>(ld, cnt, s_slot, c.slot(S_HE + t), c.slot(S_HD + t), c.slot(S_HF + t), c.slot(S_HE),
                             c.slot(S_HF), c.slot(S_HD), true, s_red, s_dots);
            if (threadIdx.x == 0) {
              for (int u = 0; u < cnt; ++u) {
                const int v = s_slot[u];
                s_M[t * HIST + v] = s_dots[3 * u + 1];
                s_M[v * HIST + t] = s_dots[3 * u + 2];
              }
              s_mok[t] = 1;
            }
            __syncthreads();
          }
        }
        if (use_adiis && threadIdx.x < 64) {
          adiis_wave(cnt, s_slot, s_M, s_coef);
          if (threadIdx.x == 0) {
            for (int i = 0; i < cnt; ++i) s_cslot[i] = s_slot[i];
            s_int[2] = cnt;
          }
        }
        __syncthreads();
      }
      // Extrapolated Fock.
      double* W = c.slot(S_W);
      const int ncoef = s_int[2];
      for (int idx = threadIdx.x; idx < ld * ld; idx += NT) {
        double v;
        if (ncoef == 0) {
          v = F[idx];
        } else {
          v = 0.0;
          for (int i = 0; i < ncoef; ++i) v += s_coef[i] * c.slot(S_HF)[(int64_t)s_cslot[i] * ld * ld + idx];
        }
        W[idx] = v;
      }
      __syncthreads();
      level_shift<NT>(n, ld, W, P, 2.0, s_guard.shift);
      lap(PH_DIIS);
      double* Cnew = c.slot(cbuf ? S_C1 : S_C0);
      const bool low = s_int[4] != 0;
      // Occupied-virtual rotations only once the SCF is near its fixed point (last density change
      // below ROTATE_BELOW); before that every iteration diagonalises in full, so large early steps
      // cannot carry it to another stationary state. PRECISION_MIXED keeps rotating after the
      // switch to FP64 (the rotation is FP64 GEMMs on the matrix cores); the orbitals and
      // eigenvalues it reports come from one full eigensolve of the converged Fock after the loop.
      const bool rot = iter > 0 && s_dbl[2] < c.M.rotate_below && (low || c.M.precision == 1);
      if (rot)
        occ_virt_rotate<NT>(n, ld, nocc, W, c.slot(cbuf ? S_C0 : S_C1), Cnew, c.slot(S_T), c.slot(S_AG), c.slot(S_VG));
      else
        sweeps += eig(low, W, iter > 0 ? c.slot(cbuf ? S_C0 : S_C1) : nullptr, Cnew, eps_a, S_AG, S_VG, S_T, 0);
      last_rot = rot;
      lap(PH_EIG);
      double* Pn = c.slot(S_HD + s_int[0]);
      const double dsq = low ? density<NT, float>(n, ld, nocc, 2.0, Cnew, P, Pn, s_red)
                             : density<NT, double>(n, ld, nocc, 2.0, Cnew, P, Pn, s_red);
      lap(PH_DENS);
      cbuf ^= 1;
      if (threadIdx.x == 0) {
        const double dp = sqrt(dsq / (double)((int64_t)n * n));
        const double de = fabs(e - s_dbl[0]);
        s_dbl[0] = e;
        s_dbl[2] = dp;
        const bool conv = tail_converged(c.M, iter, de, dp, err, s_guard.shift, low, s_tail);
        if (low) {
          // Reduced-precision phase: never converges here; hands over to FP64 below the switch.
          ++res.fp32_iters;
          res.e_fp32_last = e;
          // Progress: the density change beats the phase's best by 20 %; three iterations
          // without it mean FP32 noise has the upper hand, and FP64 takes over.
          if (dp < 0.8 * s_dbl[3]) {
            s_dbl[3] = dp;
            s_int[5] = 0;
          } else {
            ++s_int[5];
          }
          if (conv || dp < c.M.mixed_switch || res.fp32_iters >= MIXED_MAX_ITERS || s_int[5] >= 3) s_int[4] = 0;
          else guard_update(s_guard, dp, e);
        } else if (conv) {
          s_int[3] = 1;
        } else {
          guard_update(s_guard, dp, e);
        }
      }
      __syncthreads();
      if (s_int[3]) {
        ++iter;
        break;
      }
    }
    if (threadIdx.x == 0) {
      res.iterations = iter > c.M.max_scf ? c.M.max_scf : iter;
      res.converged = s_int[3];
      res.dp_err = s_dbl[2];
      res.err_norm = s_dbl[1];
    }
    // Final energy from the converged density.
    int pslot = S_HD + s_int[0];
    double* P = c.slot(pslot);
    double* F = c.slot(S_F);
    auto final_energy = [&](const double* Pf) {
      fock_build<NT, double>(c, Pf, Pf, 0.5, nullptr, 0.0, F, nullptr);
      double eh[2] = {0.0, 0.0};
      for (int idx = threadIdx.x; idx < n * n; idx += NT) {
        int i = idx / n, j = idx - i * n;
        int64_t x = (int64_t)i * ld + j;
        eh[0] += Pf[x] * H[x];
        eh[1] += Pf[x] * F[x];
      }
      wg_sum_n<NT, 2>(eh, s_red);
      e_el = 0.5 * (eh[0] + eh[1]);
    };
    final_energy(P);
    // A mixed run always ends with a full FP64 Roothaan step; its size is the result's stated
    // error against FP64 (err_vs_fp64_*).
    const bool check = c.M.precision == 1 && s_guard.shift == 0.0;
    if (s_guard.shift > 0.0 || last_rot || check) {
      sweeps += eig(false, F, c.slot(cbuf ? S_C0 : S_C1), c.slot(cbuf ? S_C1 : S_C0), eps_a, S_AG, S_VG, S_T, 0);
      cbuf ^= 1;
      if ((last_rot || check) && s_guard.shift == 0.0) {
        // The reported orbitals diagonalise the converged Fock; the reported density and energy
        // are theirs (one more Roothaan step from the converged density).
        pslot = S_HD + (s_int[0] + 1) % HIST;
        double* P2 = c.slot(pslot);
        (void)density<NT, double>(n, ld, nocc, 2.0, c.slot(cbuf ? S_C0 : S_C1), P, P2, s_red);
        const double e_conv = e_el;
        final_energy(P2);
        if (check) {
          const double dpm = max_abs_diff<NT>(n, ld, P2, P, s_red);
          if (threadIdx.x == 0) {
            res.check_de = fabs(e_el - e_conv);
            res.check_dp = dpm;
          }
        }
      }
    }
    if (threadIdx.x == 0) {
      res.p_slot = pslot;
      res.c_slot = cbuf ? S_C0 : S_C1;
      res.pb_slot = -1;
      res.cb_slot = -1;
    }
  } else {
    // ------------------------------------------------------------------ UHF (no DIIS)
    const int na = c.M.n_occ, nb = c.M.n_occ_b;
    const double ntot = (double)max(na + nb, 1);
    double* Pa = c.slot(U_PA);
    double* Pb = c.slot(U_PB);
    double* Pan = c.slot(U_PAN);
    double* Pbn = c.slot(U_PBN);
    double* Pt = c.slot(U_PT);
    double* Fa = c.slot(U_FA);
    double* Fb = c.slot(U_FB);
    if (!c.M.init_p) {  // else the caller's alpha and beta densities are there
      sad_guess<NT>(c, Pa, na / ntot);
      sad_guess<NT>(c, Pb, nb / ntot);
    }
    int cbuf = 0;
    int iter = 0;
    for (; iter < c.M.max_scf; ++iter) {
      for (int idx = threadIdx.x; idx < ld * ld; idx += NT) Pt[idx] = Pa[idx] + Pb[idx];
      __syncthreads();
      if (s_int[4]) fock_build<NT, float>(c, Pt, Pa, 1.0, Pb, 1.0, Fa, Fb);
      else fock_build<NT, double>(c, Pt, Pa, 1.0, Pb, 1.0, Fa, Fb);
      lap(PH_FOCK);
      double eh[3] = {0.0, 0.0, 0.0};
      for (int idx = threadIdx.x; idx < n * n; idx += NT) {
        int i = idx / n, j = idx - i * n;
        int64_t x = (int64_t)i * ld + j;
        eh[0] += Pt[x] * H[x];
        eh[1] += Pa[x] * Fa[x];
        eh[2] += Pb[x] * Fb[x];
      }
      wg_sum_n<NT, 3>(eh, s_red);
      const double e = 0.5 * (eh[0] + eh[1] + eh[2]);
      const double ea2 = commutator<NT>(n, ld, Fa, Pa, nullptr, s_red);
      const double eb2 = commutator<NT>(n, ld, Fb, Pb, nullptr, s_red);
      const double err = sqrt(ea2 + eb2);
      lap(PH_COMM);
      const double shift = s_guard.shift;
      level_shift<NT>(n, ld, Fa, Pa, 1.0, shift);
      level_shift<NT>(n, ld, Fb, Pb, 1.0, shift);
      double* Ca = c.slot(cbuf ? U_CA1 : U_CA0);
      double* Cb = c.slot(cbuf ? U_CB1 : U_CB0);
      const bool low = s_int[4] != 0;
      const bool rot = iter > 0 && s_dbl[2] < c.M.rotate_below && (low || c.M.precision == 1);
      last_rot = rot;
      if (rot) {
        occ_virt_rotate<NT>(n, ld, na, Fa, c.slot(cbuf ? U_CA0 : U_CA1), Ca, c.slot(U_T), c.slot(U_AG), c.slot(U_VG));
        occ_virt_rotate<NT>(n, ld, nb, Fb, c.slot(cbuf ? U_CB0 : U_CB1), Cb, c.slot(U_T), c.slot(U_AG), c.slot(U_VG));
      } else {
        sweeps += eig(low, Fa, iter > 0 ? c.slot(cbuf ? U_CA0 : U_CA1) : nullptr, Ca, eps_a, U_AG, U_VG, U_T, 1);
        sweeps += eig(low, Fb, iter > 0 ? c.slot(cbuf ? U_CB0 : U_CB1) : nullptr, Cb, eps_b, U_AG, U_VG, U_T, 2);
      }
      lap(PH_EIG);
      const double da = low ? density<NT, float>(n, ld, na, 1.0, Ca, Pa, Pan, s_red)
                            : density<NT, double>(n, ld, na, 1.0, Ca, Pa, Pan, s_red);
      const double db = low ? density<NT, float>(n, ld, nb, 1.0, Cb, Pb, Pbn, s_red)
                            : density<NT, double>(n, ld, nb, 1.0, Cb, Pb, Pbn, s_red);
      lap(PH_DENS);
      cbuf ^= 1;
      if (threadIdx.x == 0) {
        const double nn = (double)((int64_t)n * n);
        const double dp = sqrt(da / nn) + sqrt(db / nn);
        const double de = fabs(e - s_dbl[0]);
        s_dbl[0] = e;
        s_dbl[2] = dp;
        s_dbl[1] = err;
        const bool conv = tail_converged(c.M, iter, de, dp, err, s_guard.shift, low, s_tail);
        if (low) {
          ++res.fp32_iters;
          res.e_fp32_last = e;
          // Progress: the density change beats the phase's best by 20 %; three iterations
          // without it mean FP32 noise has the upper hand, and FP64 takes over.
          if (dp < 0.8 * s_dbl[3]) {
            s_dbl[3] = dp;
            s_int[5] = 0;
          } else {
            ++s_int[5];
          }
          if (conv || dp < c.M.mixed_switch || res.fp32_iters >= MIXED_MAX_ITERS || s_int[5] >= 3) s_int[4] = 0;
          else guard_update(s_guard, dp, e);
        } else if (conv) {
          s_int[3] = 1;
        } else {
          guard_update(s_guard, dp, e);
        }
      }
      __syncthreads();
      double* tmp = Pa;
      Pa = Pan;
      Pan = tmp;
      tmp = Pb;
      Pb = Pbn;
      Pbn = tmp;
      if (s_int[3]) {
        ++iter;
        break;
      }
    }
    if (threadIdx.x == 0) {
      res.iterations = iter > c.M.max_scf ? c.M.max_scf : iter;
      res.converged = s_int[3];
      res.dp_err = s_dbl[2];
      res.err_norm = s_dbl[1];
    }
    for (int idx = threadIdx.x; idx < ld * ld; idx += NT) Pt[idx] = Pa[idx] + Pb[idx];
    __syncthreads();
    fock_build<NT, double>(c, Pt, Pa, 1.0, Pb, 1.0, Fa, Fb);
    double eh[3] = {0.0, 0.0, 0.0};
    for (int idx = threadIdx.x; idx < n * n; idx += NT) {
      int i = idx / n, j = idx - i * n;
      int64_t x = (int64_t)i * ld + j;
      eh[0] += Pt[x] * H[x];
      eh[1] += Pa[x] * Fa[x];
      eh[2] += Pb[x] * Fb[x];
    }
    wg_sum_n<NT, 3>(eh, s_red);
    e_el = 0.5 * (eh[0] + eh[1] + eh[2]);
    // The last loop solve left alpha/beta vectors in buffer (cbuf ^ 1).
    int ca_last = cbuf ? U_CA0 : U_CA1, cb_last = cbuf ? U_CB0 : U_CB1;
    if (s_guard.shift > 0.0 || last_rot) {
      const int dst = (ca_last == U_CA0) ? U_CA1 : U_CA0;
      sweeps += eig(false, Fa, c.slot(ca_last), c.slot(dst), eps_a, U_AG, U_VG, U_T, 1);
      ca_last = dst;
    }
    {
      const int dst = (cb_last == U_CB0) ? U_CB1 : U_CB0;
      sweeps += eig(false, Fb, c.slot(cb_last), c.slot(dst), eps_b, U_AG, U_VG, U_T, 2);
      cb_last = dst;
    }
    const bool check = c.M.precision == 1 && s_guard.shift == 0.0;
    if (check && !last_rot) {  // alpha was not re-solved above
      const int dst = (ca_last == U_CA0) ? U_CA1 : U_CA0;
      sweeps += eig(false, Fa, c.slot(ca_last), c.slot(dst), eps_a, U_AG, U_VG, U_T, 1);
      ca_last = dst;
    }
    if ((last_rot || check) && s_guard.shift == 0.0) {
      // Densities and energy of the reported orbitals (see the RHF branch).
      (void)density<NT, double>(n, ld, na, 1.0, c.slot(ca_last), Pa, Pan, s_red);
      (void)density<NT, double>(n, ld, nb, 1.0, c.slot(cb_last), Pb, Pbn, s_red);
      if (check) {
        const double dpm = fmax(max_abs_diff<NT>(n, ld, Pan, Pa, s_red), max_abs_diff<NT>(n, ld, Pbn, Pb, s_red));
        if (threadIdx.x == 0) res.check_dp = dpm;
      }
      const double e_conv = e_el;
      double* tmp = Pa;
      Pa = Pan;
      Pan = tmp;
      tmp = Pb;
      Pb = Pbn;
      Pbn = tmp;
      for (int idx = threadIdx.x; idx < ld * ld; idx += NT) Pt[idx] = Pa[idx] + Pb[idx];
      __syncthreads();
      fock_build<NT, double>(c, Pt, Pa, 1.0, Pb, 1.0, Fa, Fb);
      double e3[3] = {0.0, 0.0, 0.0};
      for (int idx = threadIdx.x; idx < n * n; idx += NT) {
        int i = idx / n, j = idx - i * n;
        int64_t x = (int64_t)i * ld + j;
        e3[0] += Pt[x] * H[x];
        e3[1] += Pa[x] * Fa[x];
        e3[2] += Pb[x] * Fb[x];
      }
      wg_sum_n<NT, 3>(e3, s_red);
      e_el = 0.5 * (e3[0] + e3[1] + e3[2]);
      if (check && threadIdx.x == 0) res.check_de = fabs(e_el - e_conv);
    }
    if (threadIdx.x == 0) {
      res.p_slot = (Pa == c.slot(U_PA)) ? U_PA : U_PAN;
      res.pb_slot = (Pb == c.slot(U_PB)) ? U_PB : U_PBN;
      res.c_slot = ca_last;
      res.cb_slot = cb_last;
    }
  }
  if (threadIdx.x == 0) {
    tk[PH_TOTAL] = wall_clock64() - t_start;
    for (int k = 0; k < NPHASE; ++k) res.ticks[k] = tk[k];
    res.shift = s_guard.shift;
    res.sweeps = sweeps;
    res.e_el = e_el;
    out[mi] = res;
  }
}

// Size classes: <= 64 basis functions keep the packed matrix and the rotation accumulator in LDS,
// <= 120 the accumulator (the packed matrix in L2), <= 196 the packed matrix; larger molecules run
// on the grid-wide path.
#define PM7_SCF_INSTANCE(NT, MAXN, AL, VL)                                                                        \
  template __global__ void k_scf<NT, MAXN, AL, VL>(const int*, const MolDev*, const int*, const int*, const int*, \
                                                   const PairDev*, const ElemDev*, const double*, const double*, \
                                                   double*, double*, \
                                                   MolOut*);
PM7_SCF_INSTANCE(256, 64, true, true)
PM7_SCF_INSTANCE(512, 120, false, true)
PM7_SCF_INSTANCE(512, 196, true, false)

// One Fock build on a molecule's slab (the build_fock_spin entry point): F = H + J(Pt) - K(Ps).
__global__ void __launch_bounds__(256) k_fock_once(int mi, const MolDev* __restrict__ mols, const int* __restrict__ atom_z,
                                                    const int* __restrict__ atom_ao, const int* __restrict__ atom_pp,
                                                    const PairDev* __restrict__ pairs, const ElemDev* __restrict__ elem,
                                                    const double* __restrict__ ocd, const double* __restrict__ w,
                                                    double* __restrict__ slab, int pt_slot, int ps_slot, int f_slot) {
  MolCtx c;
  c.M = mols[mi];
  c.atom_z = atom_z;
  c.atom_ao = atom_ao;
  c.atom_pp = atom_pp;
  c.pairs = pairs;
  c.elem = elem;
  c.ocd = ocd;
  c.w = w;
  c.ld = c.M.ld;
  c.base = slab + c.M.mat0;
  c.pp = slab + c.M.pp0;
  fock_build<256, double>(c, c.slot(pt_slot), c.slot(ps_slot), 1.0, nullptr, 0.0, c.slot(f_slot), nullptr);
}

// Two-electron response Fock of many densities of one molecule against its one integral set
// (the CPHF kernel of pm7-rs cphf_ov, src/hessian.rs:1074-1081: build_fock_x(X) - H_core, with
// build_fock_x(X) = H + J(X) - K(X/2), src/fock.rs:665-684). Workgroup k takes density k:
// X_k at xs + k*nao*ld (row stride ld), G_k written at gs + k*nao*ld, packed-density scratch at
// pps + k*pp_stride. The densities are symmetric, as response densities C_v U C_o^T + transpose are.
template <int NT>
__global__ void __launch_bounds__(NT) k_fock_response(int mi, const MolDev* __restrict__ mols,
                                                        const int* __restrict__ atom_z, const int* __restrict__ atom_ao,
                                                        const int* __restrict__ atom_pp,
                                                        const PairDev* __restrict__ pairs,
                                                        const ElemDev* __restrict__ elem,
                                                        const double* __restrict__ ocd, const double* __restrict__ w,
                                                        double* __restrict__ slab, const double* __restrict__ xs,
                                                        double* __restrict__ gs, double* __restrict__ pps,
                                                        int64_t pp_stride) {
  MolCtx c;
  c.M = mols[mi];
  c.atom_z = atom_z;
  c.atom_ao = atom_ao;
  c.atom_pp = atom_pp;
  c.pairs = pairs;
  c.elem = elem;
  c.ocd = ocd;
  c.w = w;
  c.ld = c.M.ld;
  c.base = slab + c.M.mat0;
  c.pp = pps + (int64_t)blockIdx.x * pp_stride;
  c.minus_h = true;  // F(X) - H at each store, as the reference subtracts h_core from the Fock matrix
  const int64_t sq = (int64_t)c.M.nao * c.ld;  // rows below nao only: nothing reads or writes the rest
  const double* X = xs + blockIdx.x * sq;
  double* G = gs + blockIdx.x * sq;
  fock_build<NT, double>(c, X, X, 0.5, nullptr, 0.0, G, nullptr);
}

// The open-shell response kernel (pm7-rs analytic_hessian_uhf's step, src/hessian.rs:1850-1880):
// G^s = build_fock_spin_x(Xa + Xb, X^s) - H_core for s = alpha, beta. Workgroup k takes the pair
// (Xa_k, Xb_k) at xs + k*sq and xs + (n + k)*sq (sq = nao*ld), forms Xt = Xa + Xb at
// ts + k*sq, and writes Ga at gs + k*sq, Gb at gs + (n + k)*sq.
template <int NT>
__global__ void __launch_bounds__(NT) k_fock_response_spin(int mi, const MolDev* __restrict__ mols,
                                                            const int* __restrict__ atom_z,
                                                            const int* __restrict__ atom_ao,
                                                            const int* __restrict__ atom_pp,
                                                            const PairDev* __restrict__ pairs,
                                                            const ElemDev* __restrict__ elem,
                                                            const double* __restrict__ ocd,
                                                            const double* __restrict__ w, double* __restrict__ slab,
                                                            const double* __restrict__ xs, double* __restrict__ ts,
                                                            double* __restrict__ gs, double* __restrict__ pps,
                                                            int64_t pp_stride) {
  MolCtx c;
  c.M = mols[mi];
  c.atom_z = atom_z;
  c.atom_ao = atom_ao;
  c.atom_pp = atom_pp;
  c.pairs = pairs;
  c.elem = elem;
  c.ocd = ocd;
  c.w = w;
  c.ld = c.M.ld;
  c.base = slab + c.M.mat0;
  c.pp = pps + (int64_t)blockIdx.x * pp_stride;
  c.minus_h = true;
  const int n = c.M.nao;
  const int64_t sq = (int64_t)n * c.ld;
  const double* Xa = xs + blockIdx.x * sq;
  const double* Xb = xs + (gridDim.x + blockIdx.x) * sq;
  double* Xt = ts + blockIdx.x * sq;
  for (int t = threadIdx.x; t < n * n; t += NT) {
    const int64_t i = (int64_t)(t / n) * c.ld + t % n;
    Xt[i] = Xa[i] + Xb[i];
  }
  __syncthreads();
  fock_build<NT, double>(c, Xt, Xa, 1.0, Xb, 1.0, gs + blockIdx.x * sq, gs + (gridDim.x + blockIdx.x) * sq);
}

// ---------------------------------------------------------------------------------------------
// Divide and conquer (Dixon-Merz; pm7-rs src/dandc/mod.rs). One workgroup per subsystem, the
// subsystems of a system one launch; every subsystem keeps its integrals, core Hamiltonian and
// last eigenvectors in its slab across the SCF steps.
// ---------------------------------------------------------------------------------------------

// One subsystem step (solve_subsystem, src/dandc/mod.rs:520-626): F = H' + G(P) + diag(v_el)
// over the subsystem's own basis, where H' is the core Hamiltonian less the environment's nuclear
// potential (subtracted once, at set-up) and v_el the electronic environment of each atom
// (src/dandc/mod.rs:545-564); the partition-weighted energy 1/2 sum w_{mu nu} P (H' + F_alpha)
// with w = 1 core-core, 1/2 core-buffer, 0 buffer-buffer (src/dandc/mod.rs:593-613, weights
// src/dandc/partition.rs:60-67); (eps, C) = eig(F), warm-started from the previous step's vectors
// (cold at step 0); and per state the core-weighted norm sum_{mu in core} C_{mu k}^2 times the
// channel's spin scale (1 RHF, 1/2 per UHF channel), the weights of the common Fermi level
// (weighted_spectrum, src/dandc/mod.rs:667-712). Core atoms come first, so the core AOs are
// 0 .. nc_ao[mi]. UHF: F_s = H' + J(P) - K(P_s) for s = alpha, beta (src/dandc/mod.rs:566-591).
template <int NT, int MAXN, bool A_LDS, bool V_LDS>
__global__ void __launch_bounds__(NT)
    k_dc_solve(const int* __restrict__ mol_list, const MolDev* __restrict__ mols, const int* __restrict__ atom_z,
               const int* __restrict__ atom_ao, const int* __restrict__ atom_pp, const PairDev* __restrict__ pairs,
               const ElemDev* __restrict__ elem, const double* __restrict__ ocd, const double* __restrict__ w,
               double* __restrict__ slab, double* __restrict__ vecs, const int* __restrict__ nc_ao,
               const double* __restrict__ v_el, int step, double* __restrict__ wnorm,
               double* __restrict__ e_part) {
  __shared__ double s_a[EigLds<MAXN, A_LDS, V_LDS>::kWords];
  __shared__ double s_v[V_LDS ? MAXN * MAXN : 1];
  __shared__ double s_red[(Wg<NT>::NW + 1) * 3 * HIST];
  __shared__ double2 s_cs[MAXN / 2];
  __shared__ double s_t[MAXN / 2];
  __shared__ int s_perm[MAXN];
  __shared__ int s_pq[MAXN / 2];
  __shared__ unsigned long long s_tk[3];
  const int mi = mol_list[blockIdx.x];
  MolCtx c;
  c.M = mols[mi];
  c.atom_z = atom_z;
  c.atom_ao = atom_ao;
  c.atom_pp = atom_pp;
  c.pairs = pairs;
  c.elem = elem;
  c.ocd = ocd;
  c.w = w;
  c.ld = c.M.ld;
  c.base = slab + c.M.mat0;
  c.pp = slab + c.M.pp0;
  const int n = c.M.nao, ld = c.ld, nc = nc_ao[mi];
  const bool uhf = c.M.uhf;
  double* eps = vecs + c.M.vec0;
  double* wn = wnorm + c.M.vec0;
  const double* H = c.slot(S_H);
  double* Fa = c.slot(uhf ? (int)DU_FA : (int)DR_F);
  double* Fb = c.slot(DU_FB);
  const double* Pt = c.slot(uhf ? (int)DU_PT : (int)DR_P);
  if (uhf) fock_build<NT, double>(c, Pt, c.slot(DU_PA), 1.0, c.slot(DU_PB), 1.0, Fa, Fb);
  else fock_build<NT, double>(c, Pt, Pt, 0.5, nullptr, 0.0, Fa, nullptr);
  // Electronic environment on the Fock diagonal (both spins).
  for (int la = threadIdx.x; la < c.M.n_atoms; la += NT) {
    const int a = c.M.atom0 + la;
    const int o = atom_ao[a], no = elem[atom_z[a]].n_orb;
    const double v = v_el[a];
    for (int k = 0; k < no; ++k) {
      const int64_t d = (int64_t)(o + k) * ld + o + k;
      Fa[d] += v;
      if (uhf) Fb[d] += v;
    }
  }
  __syncthreads();
  double e = 0.0;
  for (int idx = threadIdx.x; idx < n * n; idx += NT) {
    const int i = idx / n, j = idx - i * n;
    const double wgt = 0.5 * ((i < nc ? 1.0 : 0.0) + (j < nc ? 1.0 : 0.0));
    if (wgt == 0.0) continue;
    const int64_t x = (int64_t)i * ld + j;
    e += 0.5 * wgt * Pt[x] * (H[x] + Fa[x]);
  }
  e = wg_sum<NT>(e, s_red);
  if (threadIdx.x == 0) {
    e_part[mi] = e;
    s_tk[0] = s_tk[1] = s_tk[2] = 0;
  }
  const double tol = exp10((double)-c.M.eig_tol_exp);
  const int nch = uhf ? 2 : 1;
  for (int ch = 0; ch < nch; ++ch) {
    const double* W = ch == 0 ? Fa : Fb;
    int c_new, c_old, ag, vg, tg;
    if (!uhf) {
      c_new = (step & 1) ? DR_C1 : DR_C0;
      c_old = (step & 1) ? DR_C0 : DR_C1;
      ag = DR_AG, vg = DR_VG, tg = DR_T;
    } else {
      c_new = ch == 0 ? ((step & 1) ? DU_CA1 : DU_CA0) : ((step & 1) ? DU_CB1 : DU_CB0);
      c_old = ch == 0 ? ((step & 1) ? DU_CA0 : DU_CA1) : ((step & 1) ? DU_CB0 : DU_CB1);
      ag = DU_AG, vg = DU_VG, tg = DU_T;
    }
    double* ev = eps + ch * ld;
    eigensolve<NT, MAXN, A_LDS, V_LDS, double>(n, ld, W, step > 0 ? c.slot(c_old) : nullptr, c.slot(c_new), ev,
                                               c.slot(ag), c.slot(vg), c.slot(tg), s_a, s_v, s_cs, s_t, s_perm,
                                               s_pq, s_red, s_tk, tol, 40, 0, c.M.eig_block != 0);
    __syncthreads();
    const double* C = c.slot(c_new);
    const double scale = uhf ? 0.5 : 1.0;
    for (int k = threadIdx.x; k < ld; k += NT) {
      double s = 0.0;
      if (k < n)
        for (int mu = 0; mu < nc; ++mu) {
          const double v = C[(int64_t)mu * ld + k];
          s += v * v;
        }
      wn[ch * ld + k] = scale * s;
      if (k >= n) ev[k] = 0.0;
    }
    __syncthreads();
  }
  if (!uhf)
    for (int k = threadIdx.x; k < ld; k += NT) {
      wn[ld + k] = 0.0;
      eps[ld + k] = 0.0;
    }
}

// Fermi-Dirac occupation of a spin-degenerate orbital (src/dandc/mod.rs:649-660).
__device__ inline double dc_occupation(double e, double fermi, double width) {
  const double x = (e - fermi) / width;
  if (x > 40.0) return 0.0;
  if (x < -40.0) return 2.0;
  return 2.0 / (1.0 + exp(x));
}

// The subsystem's occupation-weighted density at the common Fermi level (scatter_density,
// src/dandc/mod.rs:772-821): D_s = sum_k f_k s C_k C_k^T with f the Fermi-Dirac occupation
// (dropped below 1e-14) and s the channel's spin scale (1 RHF, 1/2 per UHF channel), as
// (C sqrt(f s)) (C sqrt(f s))^T on the matrix cores. The states are sorted ascending, so the
// occupied ones are a prefix. Written to DR_D (RHF) or DU_DA / DU_DB.
template <int NT>
__global__ void __launch_bounds__(NT) k_dc_density(const MolDev* __restrict__ mols, double* __restrict__ slab,
                                                     const double* __restrict__ vecs,
                                                     const double* __restrict__ scal, double width, int step) {
  __shared__ double s_red[(Wg<NT>::NW + 1) * 3 * HIST];
  __shared__ int s_cnt;
  const MolDev M = mols[blockIdx.x];
  const int n = M.nao, ld = M.ld;
  double* base = slab + M.mat0;
  const int64_t sq = (int64_t)ld * ld;
  const double fermi = scal[2];
  const int nch = M.uhf ? 2 : 1;
  for (int ch = 0; ch < nch; ++ch) {
    int cs;
    double* D;
    const double* C;
    if (!M.uhf) {
      C = base + sq * ((step & 1) ? DR_C1 : DR_C0);
      D = base + sq * DR_D;
      cs = DR_CS;
    } else {
      C = base + sq * (ch == 0 ? ((step & 1) ? DU_CA1 : DU_CA0) : ((step & 1) ? DU_CB1 : DU_CB0));
      D = base + sq * (ch == 0 ? DU_DA : DU_DB);
      cs = DU_CS;
    }
    double* Cs = base + sq * cs;
    const double* ev = vecs + M.vec0 + ch * ld;
    const double scale = M.uhf ? 0.5 : 1.0;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    // Occupied prefix length, then the scaled columns.
    for (int k = threadIdx.x; k < n; k += NT) {
      const double f = dc_occupation(ev[k], fermi, width) * scale;
      if (f >= 1.0e-14) atomicMax(&s_cnt, k + 1);
    }
    __syncthreads();
    const int nocc = s_cnt;
    for (int idx = threadIdx.x; idx < n * nocc; idx += NT) {
      const int mu = idx / nocc, k = idx - mu * nocc;
      const double f = dc_occupation(ev[k], fermi, width) * scale;
      Cs[(int64_t)mu * ld + k] = f >= 1.0e-14 ? C[(int64_t)mu * ld + k] * sqrt(f) : 0.0;
    }
    __syncthreads();
    density<NT, double>(n, ld, nocc, 1.0, Cs, nullptr, D, s_red);
    __syncthreads();
  }
}

#define PM7_DC_INSTANCE(NT, MAXN, AL, VL)                                                                          \
  template __global__ void k_dc_solve<NT, MAXN, AL, VL>(const int*, const MolDev*, const int*, const int*, const int*, \
                                                        const PairDev*, const ElemDev*, const double*, const double*, \
                                                        double*, double*, const int*, const double*, int, double*,   \
                                                        double*);
PM7_DC_INSTANCE(256, 64, true, true)
PM7_DC_INSTANCE(512, 120, false, true)
PM7_DC_INSTANCE(512, 196, true, false)

hipError_t launch_dc_solve(int variant, const BatchDev& b, int list_off, int count, const int* nc_ao,
                           const double* v_el, int step, double* wnorm, double* e_part, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int* list = b.mol_list + list_off;
#define PM7_DC_LAUNCH(NT, MAXN, AL, VL)                                                                            \
  hipLaunchKernelGGL((k_dc_solve<NT, MAXN, AL, VL>), dim3(count), dim3(NT), 0, s, list, b.mols, b.atom_z, b.atom_ao, \
                     b.atom_pp, b.pairs, b.elem, b.ocd, b.w, b.slab, b.vecs, nc_ao, v_el, step, wnorm, e_part)
  switch (variant) {
    case 0: PM7_DC_LAUNCH(256, 64, true, true); break;
    case 1: PM7_DC_LAUNCH(512, 120, false, true); break;
    case 2: PM7_DC_LAUNCH(512, 196, true, false); break;
    default: return hipErrorInvalidValue;  // subsystems above 196 basis functions are refused by the host
  }
#undef PM7_DC_LAUNCH
  return hipGetLastError();
}

hipError_t launch_dc_density(const BatchDev& b, int n_sub, const double* scal, double width, int step,
                             hipStream_t s) {
  if (n_sub <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dc_density<256>, dim3(n_sub), dim3(256), 0, s, b.mols, b.slab, b.vecs, scal, width, step);
  return hipGetLastError();
}

hipError_t launch_scf(int variant, const BatchDev& b, int list_off, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int* list = b.mol_list + list_off;
  switch (variant) {
    case 0:
      hipLaunchKernelGGL((k_scf<256, 64, true, true>), dim3(count), dim3(256), 0, s, list, b.mols, b.atom_z,
                         b.atom_ao, b.atom_pp, b.pairs, b.elem, b.ocd, b.w, b.slab, b.vecs, b.out);
      break;
    case 1:
      hipLaunchKernelGGL((k_scf<512, 120, false, true>), dim3(count), dim3(512), 0, s, list, b.mols, b.atom_z,
                         b.atom_ao, b.atom_pp, b.pairs, b.elem, b.ocd, b.w, b.slab, b.vecs, b.out);
      break;
    case 2:
      hipLaunchKernelGGL((k_scf<512, 196, true, false>), dim3(count), dim3(512), 0, s, list, b.mols, b.atom_z,
                         b.atom_ao, b.atom_pp, b.pairs, b.elem, b.ocd, b.w, b.slab, b.vecs, b.out);
      break;
    default:
      return hipErrorInvalidValue;  // the large class runs large_scf, not a megakernel
  }
  return hipGetLastError();
}

hipError_t launch_fock_once(const BatchDev& b, int mol, int pt_slot, int ps_slot, int f_slot, hipStream_t s) {
  hipLaunchKernelGGL(k_fock_once, dim3(1), dim3(256), 0, s, mol, b.mols, b.atom_z, b.atom_ao, b.atom_pp, b.pairs,
                     b.elem, b.ocd, b.w, b.slab, pt_slot, ps_slot, f_slot);
  return hipGetLastError();
}

hipError_t launch_fock_response(const BatchDev& b, int mol, const double* xs, double* gs, double* pps,
                                int64_t pp_stride, int n_dens, hipStream_t s) {
  if (n_dens <= 0) return hipSuccess;
  // Threads per density (environment PM7_AMD_RESPONSE_NT: 256, 512 or 1024), read once.
  static const int nt = [] {
    const char* e = std::getenv("PM7_AMD_RESPONSE_NT");
    const int v = e ? std::atoi(e) : RESPONSE_NT_DEFAULT;
    return (v == 256 || v == 512 || v == 1024) ? v : RESPONSE_NT_DEFAULT;
  }();
#define PM7_RESPONSE_LAUNCH(NT)                                                                                    \
  hipLaunchKernelGGL(k_fock_response<NT>, dim3(n_dens), dim3(NT), 0, s, mol, b.mols, b.atom_z, b.atom_ao,        \
                     b.atom_pp, b.pairs, b.elem, b.ocd, b.w, b.slab, xs, gs, pps, pp_stride)
  if (nt == 1024) PM7_RESPONSE_LAUNCH(1024);
  else if (nt == 512) PM7_RESPONSE_LAUNCH(512);
  else PM7_RESPONSE_LAUNCH(256);
#undef PM7_RESPONSE_LAUNCH
  return hipGetLastError();
}

hipError_t launch_fock_response_spin(const BatchDev& b, int mol, const double* xs, double* ts, double* gs,
                                     double* pps, int64_t pp_stride, int n_dens, hipStream_t s) {
  if (n_dens <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fock_response_spin<RESPONSE_NT_DEFAULT>, dim3(n_dens), dim3(RESPONSE_NT_DEFAULT), 0, s, mol,
                     b.mols, b.atom_z, b.atom_ao, b.atom_pp, b.pairs, b.elem, b.ocd, b.w, b.slab, xs, ts, gs, pps,
                     pp_stride);
  return hipGetLastError();
}

}  // namespace pm7

// SPDX-License-Identifier: GPL-3.0-or-later
//
// The large-system SCF path (basis sizes above the one-workgroup classes, up to thousands of AOs):
// one molecule owns the whole device and every step of an iteration is a grid-wide kernel or a
// library call, in the order of pm7-rs rhf_loop / uhf_loop (src/scf.rs:1231-1295, 1380-1470):
//
//   F = H + G(P)          grid kernels over (atom, orbital pair) items and atom pairs (fock.h)
//   E = 1/2 <P, H + F>    two-stage fixed-order reduction
//   e = FP - PF           rocBLAS dgemm + antisymmetrise, |e|^2 reduced in fixed order
//   DIIS                  history dots (one pass), CDIIS / A-DIIS on one wavefront (diis.h)
//   (eps, C) = eig(F')    rocSOLVER dsyevd (FP64); in the reduced-precision phase of a mixed run
//                         near convergence, occupied-virtual rotations on rocBLAS GEMMs instead
//   P = 2 C_occ C_occ^T   rocBLAS dgemm, or sgemm on FP32 copies in the reduced-precision phase
//
// Scalars (ring of DIIS slots, coefficients, stall guard, convergence) live in a control block in
// device memory, updated by one-wavefront control kernels; the host reads only the stop flag.
// Orientation: rocBLAS/rocSOLVER are column-major. Our square matrices are row-major with leading
// dimension ld; symmetric ones are the same either way. After dsyevd the eigenvector array read
// row-major is C^T (row k = orbital k), so the orbitals are kept as Ct = C^T during the SCF and
// transposed into the C slot at the end.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <memory>
#include <type_traits>
#include <vector>

#include "diis.h"
#include "fock.h"
#include "kernels.h"
#include "lockstep.h"
#include "pm7_device.h"

namespace pm7 {

namespace {

constexpr int RED_BLOCKS = 512;  // partial sums per reduction (fixed, so the order is fixed)
constexpr int NDOT = 3 * HIST + 8;

// Device control block of one large SCF.
struct LargeCtl {
  int head, cnt, ncoef, done, low, fp32_iters, stall, iter;
  int slot[HIST], cslot[HIST];
  double coef[HIST + 1];
  double G[HIST * HIST], M[HIST * HIST];
  double e_old, e_cur, err, dp, best_dp, e_low;
  double e_conv, check_de, check_dp;  // mixed precision: the final FP64 Roothaan step's size
  int tail_extra, tail_stall;          // mixed tail: iterations past the caller's convergence test
  double tail_best;
  double dots[NDOT];  // reductions land here
  Guard guard;
  double wk[(HIST + 1) * (HIST + 3)];
};

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
    if (e == hipSuccess) cap = std::max<size_t>(bytes, 256);
    return e;
  }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

__device__ MolCtx make_ctx(const MolDev& M, const BatchDev& b) {
  MolCtx c;
  c.M = M;
  c.atom_z = b.atom_z;
  c.atom_ao = b.atom_ao;
  c.atom_pp = b.atom_pp;
  c.pairs = b.pairs;
  c.elem = b.elem;
  c.ocd = b.ocd;
  c.w = b.w;
  c.ld = M.ld;
  c.base = b.slab + M.mat0;
  c.pp = b.slab + M.pp0;
  return c;
}

// ---- kernels --------------------------------------------------------------------------------

template <typename T>
__global__ void k_lg_pack(int mi, BatchDev b, const double* __restrict__ Pt) {
  const MolCtx c = make_ctx(b.mols[mi], b);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < c.M.n_pp; t += gridDim.x * blockDim.x)
    fock_pack_item<T>(c, t, Pt);
}

// Diagonal blocks: one wavefront per (atom, orbital pair) element, lanes over the partner atoms.
template <typename T>
__global__ void __launch_bounds__(256) k_lg_fock_diag(int mi, BatchDev b, const double* __restrict__ Pt,
                                                      const double* __restrict__ Psa, double xa,
                                                      const double* __restrict__ Psb, double xb,
                                                      double* __restrict__ Fa, double* __restrict__ Fb) {
  const MolCtx c = make_ctx(b.mols[mi], b);
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < c.M.n_pp; t += waves)
    fock_diag_item_wave<T>(c, t, Pt, Psa, xa, Psb, xb, Fa, Fb);
}

// Off-diagonal blocks: one thread per atom pair.
template <typename T>
__global__ void __launch_bounds__(256) k_lg_fock_pairs(int mi, BatchDev b, const double* __restrict__ Psa, double xa,
                                                       const double* __restrict__ Psb, double xb,
                                                       double* __restrict__ Fa, double* __restrict__ Fb) {
  const MolCtx c = make_ctx(b.mols[mi], b);
  const int64_t npairs = (int64_t)c.M.n_atoms * (c.M.n_atoms - 1) / 2;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += (int64_t)gridDim.x * blockDim.x)
    fock_pair_item<T>(c, (int)p, Psa, xa, Psb, xb, Fa, Fb);
}

// Block-partial sums of up to 3 dot products <X_k, Y_k> over the n x n block, into
// part[block * 3 + k]; every block visits a fixed strided set of elements.
__global__ void __launch_bounds__(256) k_lg_dots(int n, int ld, int k, const double* __restrict__ X0,
                                                 const double* __restrict__ Y0, const double* __restrict__ X1,
                                                 const double* __restrict__ Y1, const double* __restrict__ X2,
                                                 const double* __restrict__ Y2, double* __restrict__ part) {
  double acc[3] = {0.0, 0.0, 0.0};
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx / n, j = idx - i * n, x = i * ld + j;
    acc[0] += X0[x] * Y0[x];
    if (k > 1) acc[1] += X1[x] * Y1[x];
    if (k > 2) acc[2] += X2[x] * Y2[x];
  }
  __shared__ double red[3][4];
  for (int q = 0; q < 3; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0] + red[threadIdx.x][1] +
                                                             red[threadIdx.x][2] + red[threadIdx.x][3];
}

// Sum the block partials of k_lg_dots in block order into dst[0..k).
__global__ void k_lg_finish(int nblocks, int k, const double* __restrict__ part, double* __restrict__ dst) {
  if (threadIdx.x < k) {
    double s = 0.0;
    for (int bI = 0; bI < nblocks; ++bI) s += part[bI * 3 + threadIdx.x];
    dst[threadIdx.x] = s;
  }
}

// E = X - X^T (X = FP) into E (when given) and block partials of sum E^2.
__global__ void __launch_bounds__(256) k_lg_comm(int n, int ld, const double* __restrict__ X, double* __restrict__ E,
                                                 double* __restrict__ part) {
  double acc = 0.0;
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx / n, j = idx - i * n;
    const double e = X[i * ld + j] - X[j * ld + i];
    if (E) E[i * ld + j] = e;
    acc += e * e;
  }
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x * 3] = red[0] + red[1] + red[2] + red[3];
    part[blockIdx.x * 3 + 1] = 0.0;
    part[blockIdx.x * 3 + 2] = 0.0;
  }
}

// W = sum_i coef_i F_{cslot_i} (or F when no coefficients) + shift (1 - P/occ) on the virtual
// projector (the megakernel's level_shift: W += shift (I - P/occ) over the n x n block).
// F and W may be the same matrix (the in-place shift of the UHF path).
__global__ void __launch_bounds__(256) k_lg_combine(int n, int ld, const LargeCtl* __restrict__ ctl,
                                                    const double* __restrict__ hist_f, const double* F,
                                                    const double* __restrict__ P, double occ, double* W,
                                                    int use_diis) {
  const int ncoef = use_diis ? ctl->ncoef : 0;
  const double shift = ctl->guard.shift;
  const int64_t sq = (int64_t)ld * ld;
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx / n, j = idx - i * n, x = i * ld + j;
    double v;
    if (ncoef == 0) {
      v = F[x];
    } else {
      v = 0.0;
      for (int k = 0; k < ncoef; ++k) v += ctl->coef[k] * hist_f[(int64_t)ctl->cslot[k] * sq + x];
    }
    if (shift > 0.0) v += shift * ((i == j ? 1.0 : 0.0) - P[x] / occ);
    W[x] = v;
  }
}

// Pn = w Ct_occ^T Ct_occ was formed by a library product into Pn; here the change against P,
// block partials of sum (Pn - P)^2.
__global__ void __launch_bounds__(256) k_lg_diff(int n, int ld, const double* __restrict__ Pn,
                                                 const double* __restrict__ P, double* __restrict__ part) {
  double acc = 0.0;
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx / n, j = idx - i * n, x = i * ld + j;
    const double d = Pn[x] - P[x];
    acc += d * d;
  }
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x * 3] = red[0] + red[1] + red[2] + red[3];
    part[blockIdx.x * 3 + 1] = 0.0;
    part[blockIdx.x * 3 + 2] = 0.0;
  }
}

// Per-block max |Pn - P| over n x n into part[block] (the check of a mixed run's last step).
__global__ void __launch_bounds__(256) k_lg_maxdiff(int n, int ld, const double* __restrict__ Pn,
                                                    const double* __restrict__ P, double* __restrict__ part) {
  double m = 0.0;
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx / n, j = idx - i * n, x = i * ld + j;
    m = fmax(m, fabs(Pn[x] - P[x]));
  }
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// what 0: remember the converged energy; 1: the step's energy change and (with nb blocks of
// part) its max density change, folded into the control block.
__global__ void k_lg_check(LargeCtl* __restrict__ ctl, int what, int nb, const double* __restrict__ part) {
  if (threadIdx.x != 0) return;
  if (what == 0) {
    ctl->e_conv = ctl->e_cur;
    ctl->check_de = 0.0;
    ctl->check_dp = 0.0;
    return;
  }
  double m = 0.0;
  for (int k = 0; k < nb; ++k) m = fmax(m, part[k]);
  ctl->check_dp = fmax(ctl->check_dp, m);
  ctl->check_de = fabs(ctl->e_cur - ctl->e_conv);
}

__global__ void __launch_bounds__(256) k_lg_to_f32(int64_t count, const double* __restrict__ src, float* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) dst[i] = (float)src[i];
}
__global__ void __launch_bounds__(256) k_lg_to_f64(int64_t count, const float* __restrict__ src, double* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) dst[i] = (double)src[i];
}
__global__ void __launch_bounds__(256) k_lg_copy(int64_t count, const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}
// W (n x n of leading dimension ld) into an nmax x nmax slot of the batched eigensolve; the pad
// diagonal gets 1e5 + i, far above any Fock eigenvalue, so the pad eigenpairs sort last and stay
// decoupled from the molecule's own (the pad block has no coupling to it).
__global__ void __launch_bounds__(256) k_lg_pad(int n, int ld, int nmax, const double* __restrict__ W,
                                                double* __restrict__ dst) {
  const int64_t nn = (int64_t)nmax * nmax;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int j = (int)(idx / nmax), i = (int)(idx - (int64_t)j * nmax);
    dst[idx] = (i < n && j < n) ? W[(int64_t)j * ld + i] : (i == j ? 1.0e5 + i : 0.0);
  }
}
// The first n eigenpairs of a slot back into the molecule's orbitals (Ct, column j = orbital j)
// and eigenvalues; pad rows/columns of Ct are zero.
__global__ void __launch_bounds__(256) k_lg_unpad(int n, int ld, int nmax, const double* __restrict__ V,
                                                  const double* __restrict__ D, double* __restrict__ Ct,
                                                  double* __restrict__ eps) {
  const int64_t nn = (int64_t)ld * ld;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int j = (int)(idx / ld), i = (int)(idx - (int64_t)j * ld);
    Ct[idx] = (i < n && j < n) ? V[(int64_t)j * nmax + i] : 0.0;
    if (i == 0 && j < n) eps[j] = D[j];
  }
}
// dst = src^T over the ld x ld square (pad rows/columns of src are zero).
__global__ void __launch_bounds__(256) k_lg_transpose(int ld, const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t nn = (int64_t)ld * ld;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx / ld, j = idx - i * ld;
    dst[j * ld + i] = src[idx];
  }
}

// Occupied-virtual generator K from F' (symmetric, column-major = row-major): K_ia = theta,
// K_ai = -theta with the 2x2 Jacobi angle of pair (i, a) (eigen.h occ_virt_rotate), column-major.
__global__ void __launch_bounds__(256) k_lg_kgen(int n, int ld, int nocc, const double* __restrict__ Fp,
                                                 double* __restrict__ K) {
  const int64_t nn = (int64_t)ld * ld;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int j = (int)(idx / ld), i = (int)(idx - (int64_t)j * ld);  // column-major (i, j)
    double k = 0.0;
    if (i < n && j < n) {
      const bool io = i < nocc, jo = j < nocc;
      if (io != jo) {
        const int o = io ? i : j, v = io ? j : i;
        const double apq = Fp[(int64_t)v * ld + o];
        if (apq != 0.0) {
          const double d = Fp[(int64_t)v * ld + v] - Fp[(int64_t)o * ld + o];
          const double x = 2.0 * apq;
          const double rr = sqrt(d * d + x * x);
          const double th = atan((d >= 0.0 ? x : -x) / (fabs(d) + rr));
          k = io ? th : -th;
        }
      }
    }
    K[idx] = k;
  }
}

// U = I + K + K2/2 + K3/6 (all column-major, ld x ld; pad diagonal entries 1 keep U orthogonal).
__global__ void __launch_bounds__(256) k_lg_taylor(int ld, const double* __restrict__ K, const double* __restrict__ K2,
                                                   const double* __restrict__ K3, double* __restrict__ U) {
  const int64_t nn = (int64_t)ld * ld;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int64_t j = idx / ld, i = idx - j * ld;
    U[idx] = (i == j ? 1.0 : 0.0) + K[idx] + 0.5 * K2[idx] + K3[idx] / 6.0;
  }
}

// S <- 1.5 I - 0.5 S (Newton-Schulz factor), ld x ld.
__global__ void __launch_bounds__(256) k_lg_ns(int ld, double* __restrict__ S) {
  const int64_t nn = (int64_t)ld * ld;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int64_t j = idx / ld, i = idx - j * ld;
    S[idx] = (i == j ? 1.5 : 0.0) - 0.5 * S[idx];
  }
}

// Control A (one wavefront): advance the DIIS ring, fill the Gram and mixing matrices from the
// history dots, pick CDIIS or A-DIIS coefficients (src/scf.rs:1240-1266).
__global__ void __launch_bounds__(64) k_lg_ctrl_a(LargeCtl* __restrict__ ctl, int use_diis, int accel,
                                                  double adiis_switch) {
  __shared__ int s_slot[HIST];
  __shared__ double s_M[HIST * HIST];
  __shared__ double s_coef[HIST + 1];
  __shared__ int s_cnt, s_h, s_use_adiis;
  if (threadIdx.x == 0) {
    const int h = ctl->head;
    if (ctl->cnt < HIST) ctl->cnt++;
    const int cnt = ctl->cnt;
    ctl->head = (h + 1) % HIST;
    for (int i = 0; i < cnt; ++i) ctl->slot[i] = (ctl->head - cnt + i + 2 * HIST) % HIST;
    for (int q = 0; q < cnt; ++q) {
      const int t = ctl->slot[q];
      ctl->G[h * HIST + t] = ctl->dots[3 * q];
      ctl->G[t * HIST + h] = ctl->dots[3 * q];
      ctl->M[h * HIST + t] = ctl->dots[3 * q + 1];
      ctl->M[t * HIST + h] = ctl->dots[3 * q + 2];
    }
    const double err = ctl->err;
    int ncoef = 0, skip = 0;
    if (use_diis && accel != 0 && (accel == 1 || err <= adiis_switch)) {
      ncoef = cdiis(cnt, ctl->slot, ctl->G, ctl->coef, &skip, ctl->wk);
      for (int i = 0; i < ncoef; ++i) ctl->cslot[i] = ctl->slot[skip + i];
    }
    ctl->ncoef = ncoef;
    s_use_adiis = (use_diis && accel == 2 && err > adiis_switch && cnt >= 2) ? 1 : 0;
    s_cnt = cnt;
    s_h = h;
    for (int i = 0; i < cnt; ++i) s_slot[i] = ctl->slot[i];
    for (int i = 0; i < HIST * HIST; ++i) s_M[i] = ctl->M[i];
  }
  __syncthreads();
  if (s_use_adiis) {
    adiis_wave(s_cnt, s_slot, s_M, s_coef);
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 0; i < s_cnt; ++i) {
        ctl->coef[i] = s_coef[i];
        ctl->cslot[i] = s_slot[i];
      }
      ctl->ncoef = s_cnt;
    }
  }
  (void)s_h;
}

// Control B (one thread): convergence, stall guard and the precision hand-over
// (src/scf.rs:1274-1287; the megakernel's tail of an iteration).
__global__ void k_lg_ctrl_b(LargeCtl* __restrict__ ctl, int iter, int64_t nn, double e_tol, double p_tol,
                            double tail, double mixed_switch, int uhf) {
  if (threadIdx.x != 0) return;
  const double e = ctl->e_cur;
  const double dp = uhf ? sqrt(ctl->dots[0] / (double)nn) + sqrt(ctl->dots[1] / (double)nn)
                        : sqrt(ctl->dots[0] / (double)nn);
  const double de = fabs(e - ctl->e_old);
  ctl->e_old = e;
  ctl->dp = dp;
  // The caller's test; a mixed run's FP64 tail asks for `tail` times both bounds, for at most
  // MIXED_TAIL_EXTRA iterations past the first that meets the caller's (as in the megakernel).
  const bool resid_ok = ctl->guard.shift == 0.0 && ctl->err < 1.0e-7;
  bool conv = iter > 0 && de < e_tol && (dp < p_tol || resid_ok);
  if (tail < 1.0 && !ctl->low) {
    const bool tight = iter > 0 && de < e_tol && (dp < p_tol * tail || (ctl->guard.shift == 0.0 && ctl->err < 1.0e-7 * tail));
    if (!tight && conv) {  // past the caller's test: go on while the density change still drops
      if (ctl->tail_extra++ == 0 || dp < 0.8 * ctl->tail_best) {
        ctl->tail_best = dp;
        ctl->tail_stall = 0;
      } else {
        ++ctl->tail_stall;
      }
      conv = ctl->tail_stall >= 3 || ctl->tail_extra > MIXED_TAIL_EXTRA;
    } else {
      conv = tight;
    }
  }
  if (ctl->low) {
    ctl->fp32_iters++;
    ctl->e_low = e;
    if (dp < 0.8 * ctl->best_dp) {
      ctl->best_dp = dp;
      ctl->stall = 0;
    } else {
      ctl->stall++;
    }
    if (conv || dp < mixed_switch || ctl->fp32_iters >= MIXED_MAX_ITERS || ctl->stall >= 3) ctl->low = 0;
    else guard_update(ctl->guard, dp, e);
  } else if (conv) {
    ctl->done = 1;
  } else {
    guard_update(ctl->guard, dp, e);
  }
  ctl->iter = iter + 1;
}

// Energy from the dots: e = 1/2 (<P, H> + <P, F>) (RHF) or 1/2 (<Pt, H> + <Pa, Fa> + <Pb, Fb>).
__global__ void k_lg_energy(LargeCtl* __restrict__ ctl, const double* __restrict__ d3, int uhf) {
  if (threadIdx.x == 0) ctl->e_cur = 0.5 * (d3[0] + d3[1] + (uhf ? d3[2] : 0.0));
}
__global__ void k_lg_err(LargeCtl* __restrict__ ctl, const double* __restrict__ d3, int terms) {
  if (threadIdx.x == 0) ctl->err = sqrt(d3[0] + (terms > 1 ? d3[1] : 0.0));
}

__global__ void k_lg_init(LargeCtl* __restrict__ ctl, double shift_floor, int low) {
  if (threadIdx.x != 0) return;
  ctl->head = 0;
  ctl->cnt = 0;
  ctl->ncoef = 0;
  ctl->done = 0;
  ctl->low = low;
  ctl->fp32_iters = 0;
  ctl->stall = 0;
  ctl->tail_extra = 0;
  ctl->tail_stall = 0;
  ctl->tail_best = 1.0e30;
  ctl->iter = 0;
  ctl->e_old = 0.0;
  ctl->e_cur = 0.0;
  ctl->err = 0.0;
  ctl->dp = 1.0e30;
  ctl->best_dp = 1.0e30;
  ctl->e_low = 0.0;
  ctl->guard.floor_ = shift_floor;
  ctl->guard.shift = shift_floor;
  ctl->guard.n = 0;
}

// ---- s/p closed-shell Fock from per-atom integral streams ---------------------------------------
//
// For molecules whose atoms all carry 1 or 4 orbitals, each atom a gets its own stream of the pair
// integral blocks (a, b) over every partner b, built once per SCF: partners are grouped by class
// (4-orbital atoms, then 1-orbital atoms, each list in atom order and padded to a multiple of 64
// with -1) and a chunk of 64 partners is stored element-major, element k of lane t at
// [chunk][k][t], each block in the pair's own order w[e_F][kl_S] (F the pair's first atom), zero
// for a's own slot and the padding. One workgroup per atom then reads its stream once, coalesced,
// and writes its own rows of F: the Coulomb sum of the diagonal block over all partners (lanes,
// then a fixed-order reduction) and each partner's exchange block, accumulated element by element
// in the pair's own order and oriented by F and S, so the blocks (a, b) and (b, a) from two
// workgroups are the same sums bit for bit and F stays exactly symmetric.

constexpr int LGS_CHUNK = 64;
constexpr int LGS_THREADS = 256;  // threads of a k_lgs_fock workgroup

// The larger orbital index of packed pair e (pack_idx's h), at compile time.
constexpr int ce_hi(int e) {
  return e >= 36 ? 8 : e >= 28 ? 7 : e >= 21 ? 6 : e >= 15 ? 5 : e >= 10 ? 4 : e >= 6 ? 3 : e >= 3 ? 2 : e >= 1 ? 1 : 0;
}

// f(std::integral_constant<int, I>) for I in [B, E), each I a compile-time constant.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Orbital count of every atom of molecule mi, in molecule order.
__global__ void k_lgs_norb(int mi, BatchDev b, int* __restrict__ norb) {
  const MolDev M = b.mols[mi];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M.n_atoms; i += gridDim.x * blockDim.x)
    norb[i] = b.elem[b.atom_z[M.atom0 + i]].n_orb;
}

// Workgroup per (chunk c of atom a's stream): the 64 partners' blocks read one after another
// (each block contiguous), transposed through LDS, written element-major; the FP64 stream always,
// the FP32 one when st32 is set. Chunks of a: the 4-orbital partners' (cls[0, pA)), then the
// 1-orbital partners' (cls[pA, pA + pB)); a's stream starts at off[a].
__global__ void __launch_bounds__(256) k_lgs_build(int mi, BatchDev b, const int* __restrict__ cls, int pA, int pB,
                                                   const int64_t* __restrict__ off, double* __restrict__ st,
                                                   float* __restrict__ st32) {
  __shared__ double tile[100 * LGS_CHUNK];
  __shared__ int64_t s_woff[LGS_CHUNK];
  const MolDev M = b.mols[mi];
  const int la = blockIdx.y, na = M.n_atoms;
  const int npa = npair_of(b.elem[b.atom_z[M.atom0 + la]].n_orb);
  const int ca = pA / LGS_CHUNK, c = blockIdx.x;
  if (c >= ca + pB / LGS_CHUNK) return;  // the whole workgroup
  const bool inA = c < ca;
  const int K = npa * (inA ? 10 : 1);
  const int* list = inA ? cls + c * LGS_CHUNK : cls + pA + (c - ca) * LGS_CHUNK;
  const int64_t base = off[la] + (inA ? (int64_t)c * K * LGS_CHUNK
                                      : (int64_t)pA * npa * 10 + (int64_t)(c - ca) * K * LGS_CHUNK);
  if (threadIdx.x < LGS_CHUNK) {
    const int lb = list[threadIdx.x];
    s_woff[threadIdx.x] =
        (lb >= 0 && lb != la) ? b.pairs[M.pair0 + tri_index(min(la, lb), max(la, lb), na)].w_off : (int64_t)-1;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < K * LGS_CHUNK; i += 256) {  // k fastest: each block read contiguously
    const int t = i / K, k = i - t * K;
    const int64_t wo = s_woff[t];
    tile[k * LGS_CHUNK + t] = wo >= 0 ? b.w[wo + k] : 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < K * LGS_CHUNK; i += 256) {  // element-major, coalesced
    st[base + i] = tile[i];
    if (st32) st32[base + i] = (float)tile[i];
  }
}

// One chunk of partners (class with NFB orbitals) of a workgroup whose atom has NFA orbitals: lane
// `lane` takes partner cls[lane]; J gathers the Coulomb sums of a's packed pairs.
template <typename T, int NFA, int NFB>
__device__ __forceinline__ void lgs_chunk(const MolDev& M, const BatchDev& b, int la, int oa,
                                          const T* __restrict__ sp, const int* __restrict__ cls, int lane,
                                          const double* __restrict__ P, double xa, double* __restrict__ F,
                                          T (&J)[10]) {
  constexpr int NF = NFA > NFB ? NFA : NFB, NS = NFA > NFB ? NFB : NFA;
  constexpr int NPF = NF * (NF + 1) / 2, NPS = NS * (NS + 1) / 2;
  const int ld = M.ld;
  const int lb = cls[lane];
  const bool live = lb >= 0 && lb != la;
  const int bb = M.atom0 + (live ? lb : la);  // a valid atom for the loads of idle lanes
  const int ob = b.atom_ao[bb];
  const double* ppb = b.slab + M.pp0 + b.atom_pp[bb];
  // a is the pair's first atom: the larger orbital count, on a tie the lower index
  const bool af = NFA > NFB ? true : (NFA < NFB ? false : la < lb);
  // P[F_p][S_q] read at a's rows (P is symmetric): in the class-sorted basis the partners of a
  // chunk are contiguous in AO order, so these loads and the stores below are coalesced.
  T pfs[NF][NS], acc[NF][NS];
#pragma unroll
  for (int p = 0; p < NF; ++p)
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      pfs[p][q] = (T)P[af ? (int64_t)(oa + p) * ld + ob + q : (int64_t)(oa + q) * ld + ob + p];
      acc[p][q] = (T)0;
    }
  T ppv[NPF > NPS ? NPF : NPS];
  const int npb = NFB == 4 ? 10 : 1;
#pragma unroll
  for (int k = 0; k < (NPF > NPS ? NPF : NPS); ++k) ppv[k] = k < npb ? (T)ppb[k] : (T)0;
  // Every index below is a compile-time constant (static_for), so the block lives in registers.
  static_for<0, NPF>([&](auto ec) {
    constexpr int e = decltype(ec)::value, mu = ce_hi(e), p = e - mu * (mu + 1) / 2;
    static_for<0, NPS>([&](auto kc) {
      constexpr int kl = decltype(kc)::value, l = ce_hi(kl), q = kl - l * (l + 1) / 2;
      const T w = sp[(e * NPS + kl) * LGS_CHUNK];
      // Coulomb: a's packed pair is e when a is F, kl when a is S
      if constexpr (NFA > NFB) {
        J[e] += w * ppv[kl];
      } else if constexpr (NFA < NFB) {
        J[kl] += w * ppv[e];
      } else {
        const T c = w * (af ? ppv[kl] : ppv[e]);
        J[e] += af ? c : (T)0;
        J[kl] += af ? (T)0 : c;
      }
      // exchange: every (mu' nu') x (la' sg') term this packed element stands for
      acc[mu][l] += pfs[p][q] * w;
      if constexpr (mu != p) acc[p][l] += pfs[mu][q] * w;
      if constexpr (l != q) acc[mu][q] += pfs[p][l] * w;
      if constexpr (mu != p && l != q) acc[p][q] += pfs[mu][l] * w;
    });
  });
  if (!live) return;
  const double* H = b.slab + M.mat0 + (int64_t)S_H * ld * ld;
#pragma unroll
  for (int p = 0; p < NF; ++p)
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int64_t at = af ? (int64_t)(oa + p) * ld + ob + q : (int64_t)(oa + q) * ld + ob + p;  // a's row
      F[at] = H[at] - (double)((T)xa * acc[p][q]);
    }
}

// The partner chunks of segment SEG (0: both, 1: the 4-orbital partners, 2: the 1-orbital ones).
template <typename T, int NFA, int SEG>
__device__ __forceinline__ void lgs_atom(const MolDev& M, const BatchDev& b, int la, int oa, const T* __restrict__ sa,
                                         const int* __restrict__ cls, int pA, int pB, int wv, int lane,
                                         const double* __restrict__ P, double xa, double* __restrict__ F, T (&J)[10]) {
  constexpr int NPA = NFA == 4 ? 10 : 1;
  constexpr int NW = LGS_THREADS / 64;
  if constexpr (SEG != 2)
    for (int c = wv; c < pA / LGS_CHUNK; c += NW)
      lgs_chunk<T, NFA, 4>(M, b, la, oa, sa + (int64_t)c * NPA * 10 * LGS_CHUNK + lane, cls + c * LGS_CHUNK, lane,
                           P, xa, F, J);
  const T* sb = sa + (int64_t)pA * NPA * 10;
  if constexpr (SEG != 1)
    for (int c = wv; c < pB / LGS_CHUNK; c += NW)
      lgs_chunk<T, NFA, 1>(M, b, la, oa, sb + (int64_t)c * NPA * LGS_CHUNK + lane, cls + pA + c * LGS_CHUNK, lane,
                           P, xa, F, J);
}

// Workgroup per atom a with NFA orbitals: its rows of F = H + J(P) - xa K(P) (closed shell), the
// Coulomb sums from the packed density of every atom (k_lg_pack<T>), one-centre terms as
// fock_diag_item. SEG 0 takes all partners. A 4-orbital atom takes two workgroups in two launches:
// SEG 2 its 1-orbital partners, Coulomb sums into jpart[slot]; SEG 1 its 4-orbital partners, then
// jpart[slot] + its own sums and the diagonal block.
template <typename T, int NFA, int SEG>
__device__ __forceinline__ void lgs_fock_atom(const MolDev& M, const BatchDev& b, const T* __restrict__ st,
                                              const int* __restrict__ cls, int pA, int pB,
                                              const int64_t* __restrict__ off, int la, int slot,
                                              const double* __restrict__ P, double xa, double* __restrict__ F,
                                              T* __restrict__ jpart, T (&red)[10][LGS_THREADS]) {
  const int a = M.atom0 + la, ld = M.ld;
  const ElemDev ea = b.elem[b.atom_z[a]];
  constexpr int nfa = NFA;
  const int oa = b.atom_ao[a];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  T J[10];
#pragma unroll
  for (int e = 0; e < 10; ++e) J[e] = (T)0;
  lgs_atom<T, NFA, SEG>(M, b, la, oa, st + off[la], cls, pA, pB, wv, lane, P, xa, F, J);
#pragma unroll
  for (int e = 0; e < 10; ++e) red[e][threadIdx.x] = J[e];  // every thread, in thread order
  __syncthreads();
  const int npa = npair_of(nfa);
  if ((int)threadIdx.x >= npa) return;
  const int e = threadIdx.x;
  T acc = SEG == 1 ? jpart[slot * 10 + e] : (T)0;
  for (int i = 0; i < LGS_THREADS; ++i) acc += red[e][i];
  if constexpr (SEG == 2) {
    jpart[slot * 10 + e] = acc;
    return;
  }
  int mu, nu;
  unpack_e(e, mu, nu);
  T ja = (T)0, ka = (T)0;
  for (int l = 0; l < nfa; ++l)
    for (int s2 = 0; s2 < nfa; ++s2) {
      const int64_t ix = (int64_t)(oa + l) * ld + oa + s2;
      ja += (T)P[ix] * (T)oc_int(mu, nu, l, s2, ea);
      ka += (T)P[ix] * (T)oc_int(mu, l, nu, s2, ea);
    }
  const double* H = b.slab + M.mat0 + (int64_t)S_H * ld * ld;
  const int64_t i1 = (int64_t)(oa + mu) * ld + oa + nu, i2 = (int64_t)(oa + nu) * ld + oa + mu;
  const double f = H[i1] + (double)(acc + ja - (T)xa * ka);
  F[i1] = f;
  F[i2] = f;
}

// The 4-orbital atoms' 4-orbital partners (SEG 1, after k_lgs_fock_light): workgroup i takes
// atom cls[i].
template <typename T>
__global__ void __launch_bounds__(LGS_THREADS) k_lgs_fock_heavy(int mi, BatchDev b, const T* __restrict__ st,
                                                        const int* __restrict__ cls, int pA, int pB,
                                                        const int64_t* __restrict__ off,
                                                        const double* __restrict__ P, double xa,
                                                        double* __restrict__ F, T* __restrict__ jpart) {
  __shared__ T red[10][LGS_THREADS];
  const MolDev M = b.mols[mi];
  lgs_fock_atom<T, 4, 1>(M, b, st, cls, pA, pB, off, cls[blockIdx.x], blockIdx.x, P, xa, F, jpart, red);
}

// The small blocks in one launch: workgroups [0, nA) take the 4-orbital atoms' 1-orbital partners
// (SEG 2), the next nB the 1-orbital atoms (all partners); both need few registers and run at the
// same occupancy.
template <typename T>
__global__ void __launch_bounds__(LGS_THREADS) k_lgs_fock_light(int mi, BatchDev b, const T* __restrict__ st,
                                                        const int* __restrict__ cls, int pA, int pB, int nA,
                                                        const int64_t* __restrict__ off,
                                                        const double* __restrict__ P, double xa,
                                                        double* __restrict__ F, T* __restrict__ jpart) {
  __shared__ T red[10][LGS_THREADS];
  const MolDev M = b.mols[mi];
  const int i = blockIdx.x;
  if (i < nA)
    lgs_fock_atom<T, 4, 2>(M, b, st, cls, pA, pB, off, cls[i], i, P, xa, F, jpart, red);
  else
    lgs_fock_atom<T, 1, 0>(M, b, st, cls, pA, pB, off, cls[pA + i - nA], 0, P, xa, F, jpart, red);
}

// Basis permutations of n x n matrices (row stride ld): dst[i][j] = src[pi[i]][pi[j]], and for the
// orbitals (rows are orbitals, columns AOs) dst[k][j] = src[k][pi[j]].
__global__ void __launch_bounds__(256) k_lgs_perm_sym(int n, int ld, const int* __restrict__ pi,
                                                      const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int i = (int)(idx / n), j = (int)(idx - (int64_t)i * n);
    dst[(int64_t)i * ld + j] = src[(int64_t)pi[i] * ld + pi[j]];
  }
}
__global__ void __launch_bounds__(256) k_lgs_perm_cols(int n, int ld, const int* __restrict__ pi,
                                                       const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int k = (int)(idx / n), j = (int)(idx - (int64_t)k * n);
    dst[(int64_t)k * ld + j] = src[(int64_t)k * ld + pi[j]];
  }
}

// SAD guess of the large path (pm7-rs src/scf.rs:1050-1114): diagonal tore/4 per sp AO, minus the
// charge spread (sad_yy), scaled; pads and off-diagonal stay zero (the slab is zeroed).
__global__ void k_lg_sad(int mi, BatchDev b, double* __restrict__ P, double scale) {
  const MolDev M = b.mols[mi];
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < M.n_atoms; a += gridDim.x * blockDim.x) {
    sad_atom(b.elem[b.atom_z[M.atom0 + a]], M.sad_yy, scale, P, b.atom_ao[M.atom0 + a], M.ld);
  }
}

}  // namespace

// ---- host driver --------------------------------------------------------------------------------

struct LargeWork {
  rocblas_handle blas = nullptr;
  Buf ctl, part, dots, info, f32a, f32b;
  Buf lgs_st, lgs_st32, lgs_cls, lgs_off, lgs_norb;  // per-atom integral streams (s/p closed shell)
  Buf lgs_perm, lgs_ao;                               // class-sorted basis: permutations, AO offsets
  Buf lgs_jpart;                                      // partial Coulomb sums of the 4-orbital atoms
  std::vector<hipEvent_t> ev;
};

LargeWork* large_work_create() {
  auto* w = new LargeWork();
  if (rocblas_create_handle(&w->blas) != rocblas_status_success) {
    delete w;
    return nullptr;
  }
  w->ev.resize(8);
  for (auto& e : w->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      large_work_destroy(w);
      return nullptr;
    }
  return w;
}

void large_work_destroy(LargeWork* w) {
  if (!w) return;
  if (w->blas) rocblas_destroy_handle(w->blas);
  for (auto& e : w->ev)
    if (e) (void)hipEventDestroy(e);
  delete w;
}


namespace {

#define LG(expr)                                                           \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      err = std::string(#expr ": ") + hipGetErrorString(e_);              \
      return e_;                                                           \
    }                                                                      \
  } while (0)
#define LGB(expr)                                                          \
  do {                                                                     \
    rocblas_status s_ = (expr);                                            \
    if (s_ != rocblas_status_success) {                                    \
      err = std::string(#expr ": ") + rocblas_status_to_string(s_);       \
      return hipErrorUnknown;                                              \
    }                                                                      \
  } while (0)

int grid_for(int64_t items) { return (int)std::min<int64_t>(std::max<int64_t>((items + 255) / 256, 1), 4096); }

}  // namespace

// Lockstep batching of the full eigensolves of several large-system SCFs running in their own
// threads: every participant hands in its matrices for the round (or none), the last one to
// arrive pads them into one strided array and runs one rocSOLVER dsyevd_strided_batched, whose
// time for 16 matrices of a few hundred rows is close to that of one.
struct EigBatcher {
  struct Item {
    const double* W;
    int n, ld;
    double* eps;
    double* Ct;
  };
  std::unique_ptr<Lockstep<Item, hipError_t>> sync;
  std::vector<hipEvent_t> ready;
  rocblas_handle blas = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  Buf work, D, E, info;
  int nmax = 0, cap = 0;

  // One round's batched solve (run by the last participant to arrive, lock held): each matrix
  // padded into the strided batch once its owner's stream reached it, one dsyevd for all, the
  // vectors and values copied back; `done` marks the end for the owners' streams.
  hipError_t run_round(std::vector<Item>& items) {
    const int cnt = (int)items.size();
    const int64_t nn = (int64_t)nmax * nmax;
    double* wk = static_cast<double*>(work.p);
    for (int k = 0; k < cnt; ++k) {
      const hipError_t e = hipStreamWaitEvent(st, ready[k], 0);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(k_lg_pad, dim3(grid_for(nn)), dim3(256), 0, st, items[k].n, items[k].ld, nmax, items[k].W,
                         wk + k * nn);
    }
    if (rocsolver_dsyevd_strided_batched(blas, rocblas_evect_original, rocblas_fill_upper, nmax, wk, nmax, nn,
                                         static_cast<double*>(D.p), nmax, static_cast<double*>(E.p), nmax,
                                         static_cast<rocblas_int*>(info.p), cnt) != rocblas_status_success)
      return hipErrorUnknown;
    for (int k = 0; k < cnt; ++k) {
      const int64_t sq = (int64_t)items[k].ld * items[k].ld;
      hipLaunchKernelGGL(k_lg_unpad, dim3(grid_for(sq)), dim3(256), 0, st, items[k].n, items[k].ld, nmax,
                         wk + k * nn, static_cast<double*>(D.p) + (int64_t)k * nmax, items[k].Ct, items[k].eps);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(done, st);
    return e;
  }

  // This participant's matrices for the round (k = 0: none), then wait for the round; the
  // caller's stream is ordered after the batched solve.
  hipError_t submit(const Item* its, int k, hipStream_t s) {
    const hipError_t st_ = sync->submit(its, k, s);
    if (k > 0 && st_ == hipSuccess) return hipStreamWaitEvent(s, done, 0);
    return st_;
  }

  void leave() { sync->leave(); }
};

EigBatcher* eig_batcher_create(int participants, int nmax, std::string& err) {
  auto* b = new EigBatcher();
  b->nmax = nmax;
  b->cap = 2 * participants;  // UHF hands in alpha and beta
  b->sync = std::make_unique<Lockstep<EigBatcher::Item, hipError_t>>(
      participants, b->cap, hipSuccess, [b](std::vector<EigBatcher::Item>& items) { return b->run_round(items); },
      [b](int slot, void* caller) { return hipEventRecord(b->ready[slot], static_cast<hipStream_t>(caller)); });
  b->sync->set_overflow_status(hipErrorInvalidValue);
  bool ok = rocblas_create_handle(&b->blas) == rocblas_status_success && hipStreamCreate(&b->st) == hipSuccess &&
            rocblas_set_stream(b->blas, b->st) == rocblas_status_success && hipEventCreate(&b->done) == hipSuccess;
  b->ready.assign(b->cap, nullptr);
  for (auto& e : b->ready) ok = ok && hipEventCreate(&e) == hipSuccess;
  ok = ok && b->work.ensure((size_t)b->cap * nmax * nmax * sizeof(double)) == hipSuccess &&
       b->D.ensure((size_t)b->cap * nmax * sizeof(double)) == hipSuccess &&
       b->E.ensure((size_t)b->cap * nmax * sizeof(double)) == hipSuccess &&
       b->info.ensure((size_t)b->cap * sizeof(rocblas_int)) == hipSuccess;
  if (!ok) {
    err = "batched eigensolver set-up failed";
    eig_batcher_destroy(b);
    return nullptr;
  }
  return b;
}

void eig_batcher_destroy(EigBatcher* b) {
  if (!b) return;
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->blas) rocblas_destroy_handle(b->blas);
  for (auto& e : b->ready)
    if (e) (void)hipEventDestroy(e);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

hipError_t large_scf(LargeWork* lw, const BatchDev& b, const MolDev& M, int mi, hipStream_t s, MolOut& out,
                     std::string& err, EigBatcher* eb) {
  // A participant of a batched eigensolve leaves the group when its SCF loop ends (or fails).
  struct Leave {
    EigBatcher* e;
    void now() {
      if (e) e->leave();
      e = nullptr;
    }
    ~Leave() { now(); }
  } leave{eb};
  const int n = M.nao, ld = M.ld;
  const int64_t sq = (int64_t)ld * ld, nn = (int64_t)n * n;
  double* base = b.slab + M.mat0;
  auto slot = [&](int k) { return base + (int64_t)k * sq; };
  LG(lw->ctl.ensure(sizeof(LargeCtl)));
  LG(lw->part.ensure((size_t)RED_BLOCKS * 3 * sizeof(double)));
  LG(lw->dots.ensure((size_t)(NDOT + 8) * sizeof(double)));
  LG(lw->info.ensure(sizeof(rocblas_int)));
  LGB(rocblas_set_stream(lw->blas, s));
  LargeCtl* ctl = static_cast<LargeCtl*>(lw->ctl.p);
  double* ctl_dots = reinterpret_cast<double*>(static_cast<char*>(lw->ctl.p) + offsetof(LargeCtl, dots));
  double* part = static_cast<double*>(lw->part.p);
  double* d3 = static_cast<double*>(lw->dots.p);
  rocblas_int* info = static_cast<rocblas_int*>(lw->info.p);
  const bool mixed = M.precision == 1;
  const double tail = mixed ? MIXED_TAIL : 1.0;  // a mixed run's FP64 tail converges tighter
  if (mixed) {
    LG(lw->f32a.ensure((size_t)sq * sizeof(float)));
    LG(lw->f32b.ensure((size_t)sq * sizeof(float)));
  }
  float* fa = static_cast<float*>(lw->f32a.p);
  float* fb = static_cast<float*>(lw->f32b.p);
  const int RB = std::min<int>(RED_BLOCKS, grid_for(nn));
  // Reduction helpers: k dots over n x n, finished into dst (device).
  auto dots = [&](int k, const double* x0, const double* y0, const double* x1, const double* y1, const double* x2,
                  const double* y2, double* dst) -> hipError_t {
    hipLaunchKernelGGL(k_lg_dots, dim3(RB), dim3(256), 0, s, n, ld, k, x0, y0, x1, y1, x2, y2, part);
    hipLaunchKernelGGL(k_lg_finish, dim3(1), dim3(64), 0, s, RB, k, part, dst);
    return hipGetLastError();
  };
  const double one = 1.0, zero = 0.0;
  const float onef = 1.0f, zerof = 0.0f;
  // Column-major product D = op(A) op(B) on ld x ld operands.
  auto gemm = [&](rocblas_operation ta, rocblas_operation tb, int m_, int n_, int k_, double alpha, const double* A,
                  const double* B, double beta, double* D) {
    return rocblas_dgemm(lw->blas, ta, tb, m_, n_, k_, &alpha, A, ld, B, ld, &beta, D, ld);
  };
  // Orbitals in Ct (row-major, row k = orbital k) == C column-major: Ccm = Ct.
  // Density Pn = w C_occ C_occ^T (symmetric): column-major Ccm[:, :nocc] Ccm[:, :nocc]^T.
  auto density = [&](const double* Ct, int nocc, double wgt, double* Pn, bool low) -> hipError_t {
    if (nocc <= 0) return hipMemsetAsync(Pn, 0, (size_t)sq * sizeof(double), s);
    if (!low) {
      if (rocblas_dgemm(lw->blas, rocblas_operation_none, rocblas_operation_transpose, n, n, nocc, &wgt, Ct, ld, Ct,
                        ld, &zero, Pn, ld) != rocblas_status_success)
        return hipErrorUnknown;
    } else {
      // FP32 inputs, FP32 accumulation (matrix cores), result widened to FP64.
      hipLaunchKernelGGL(k_lg_to_f32, dim3(grid_for(sq)), dim3(256), 0, s, sq, Ct, fa);
      const float w32 = (float)wgt;
      if (rocblas_sgemm(lw->blas, rocblas_operation_none, rocblas_operation_transpose, n, n, nocc, &w32, fa, ld, fa,
                        ld, &zerof, fb, ld) != rocblas_status_success)
        return hipErrorUnknown;
      hipLaunchKernelGGL(k_lg_to_f64, dim3(grid_for(sq)), dim3(256), 0, s, sq, fb, Pn);
    }
    (void)onef;
    return hipGetLastError();
  };
  // Full FP64 eigensolve: W (destroyed) -> eps, orbitals into Ct.
  auto eig_full = [&](double* W, double* eps, double* Ct) -> hipError_t {
    // Pad rows/columns of W are zero: restrict the solve to the n x n block.
    if (rocsolver_dsyevd(lw->blas, rocblas_evect_original, rocblas_fill_upper, n, W, ld, eps, slot(S_AG), info) !=
        rocblas_status_success)
      return hipErrorUnknown;
    hipLaunchKernelGGL(k_lg_copy, dim3(grid_for(sq)), dim3(256), 0, s, sq, W, Ct);
    return hipGetLastError();
  };
  // One SCF iteration's full eigensolves (one RHF matrix, or alpha and beta): through the group's
  // batched solve when this SCF runs in one, else directly.
  auto eig_round = [&](int k, double* W0, double* e0, double* C0, double* W1, double* e1, double* C1) -> hipError_t {
    if (!leave.e) {
      hipError_t e_ = eig_full(W0, e0, C0);
      if (e_ == hipSuccess && k > 1) e_ = eig_full(W1, e1, C1);
      return e_;
    }
    const EigBatcher::Item its[2] = {{W0, n, ld, e0, C0}, {W1, n, ld, e1, C1}};
    return leave.e->submit(its, k, s);
  };
  // A round in which this SCF needs no full eigensolve (orbital rotations instead).
  auto eig_skip = [&]() -> hipError_t { return leave.e ? leave.e->submit(nullptr, 0, s) : hipSuccess; };
  // Occupied-virtual rotation of Ct (the reduced-precision phase near convergence).
  auto eig_rotate = [&](const double* W, int nocc, double* Ct) -> hipError_t {
    double* T1 = slot(S_T);
    double* Fp = slot(S_AG);
    double* K = slot(S_VG);
    double* K2 = slot(S_C1);
    // F' = C^T W C (column-major, C = Ccm = Ct)
    if (gemm(rocblas_operation_transpose, rocblas_operation_none, n, n, n, 1.0, Ct, W, 0.0, T1) != rocblas_status_success)
      return hipErrorUnknown;
    if (gemm(rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, T1, Ct, 0.0, Fp) != rocblas_status_success)
      return hipErrorUnknown;
    hipLaunchKernelGGL(k_lg_kgen, dim3(grid_for(sq)), dim3(256), 0, s, n, ld, nocc, Fp, K);
    if (gemm(rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, K, K, 0.0, K2) != rocblas_status_success)
      return hipErrorUnknown;
    if (gemm(rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, K, K2, 0.0, T1) != rocblas_status_success)
      return hipErrorUnknown;  // K^3
    hipLaunchKernelGGL(k_lg_taylor, dim3(grid_for(sq)), dim3(256), 0, s, ld, K, K2, T1, Fp);  // U into Fp
    if (gemm(rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, Ct, Fp, 0.0, K) != rocblas_status_success)
      return hipErrorUnknown;  // C1 = C U into K
    double* src = K;
    double* dst = K2;
    for (int step = 0; step < 2; ++step) {
      if (gemm(rocblas_operation_transpose, rocblas_operation_none, n, n, n, 1.0, src, src, 0.0, T1) !=
          rocblas_status_success)
        return hipErrorUnknown;  // S = C^T C
      hipLaunchKernelGGL(k_lg_ns, dim3(grid_for(sq)), dim3(256), 0, s, ld, T1);
      double* to = step == 1 ? Ct : dst;
      if (gemm(rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, src, T1, 0.0, to) != rocblas_status_success)
        return hipErrorUnknown;
      src = to;
    }
    return hipGetLastError();
  };

  const bool mixed_ok = mixed;  // the FP32 range guard of the megakernel is applied below
  hipLaunchKernelGGL(k_lg_init, dim3(1), dim3(64), 0, s, ctl, M.shift_floor, mixed_ok ? 1 : 0);
  std::vector<float> phase(5, 0.0f);
  // Closed-shell s/p molecules: the per-atom integral streams (k_lgs_fock), built once here, their
  // device time counted in the Fock phase. Other molecules, and PM7_AMD_LARGE_STREAMS=0, keep the
  // element and pair kernels (k_lg_fock_diag, k_lg_fock_pairs).
  int lgs_pA = 0, lgs_pB = 0, lgs_nA = 0, lgs_nB = 0;
  bool lgs_on = false;
  // With the streams the SCF runs in a class-sorted basis (the 4-orbital atoms' AOs first, then the
  // 1-orbital atoms', each in atom order): bp carries that basis's AO offsets, lgs_perm its
  // permutation (new -> old, then old -> new). H and a seeded density are permuted at the start, the
  // density, the orbitals and H back at the end; everything between is basis-independent.
  BatchDev bp = b;
  const int* lgs_n2o = nullptr;
  const int* lgs_o2n = nullptr;
  auto permute = [&](double* m, const int* pi, bool cols) -> hipError_t {
    double* tmp = slot(S_W);
    hipLaunchKernelGGL(k_lg_copy, dim3(grid_for(sq)), dim3(256), 0, s, sq, m, tmp);
    if (cols) hipLaunchKernelGGL(k_lgs_perm_cols, dim3(grid_for(nn)), dim3(256), 0, s, n, ld, pi, tmp, m);
    else hipLaunchKernelGGL(k_lgs_perm_sym, dim3(grid_for(nn)), dim3(256), 0, s, n, ld, pi, tmp, m);
    return hipGetLastError();
  };
  auto lgs_setup = [&]() -> hipError_t {
    const char* env = std::getenv("PM7_AMD_LARGE_STREAMS");  // read per SCF
    const bool enabled = !(env && std::atoi(env) == 0);
    if (!enabled || M.dpath || M.uhf) return hipSuccess;
    const int na = M.n_atoms;
    hipError_t e;
    if ((e = lw->lgs_norb.ensure((size_t)na * sizeof(int))) != hipSuccess) return e;
    hipLaunchKernelGGL(k_lgs_norb, dim3(grid_for(na)), dim3(256), 0, s, mi, b, static_cast<int*>(lw->lgs_norb.p));
    std::vector<int> norb(na);
    if ((e = hipMemcpyAsync(norb.data(), lw->lgs_norb.p, (size_t)na * sizeof(int), hipMemcpyDeviceToHost, s)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return e;
    std::vector<int> heavy, light;
    for (int i = 0; i < na; ++i) {
      if (norb[i] == 4) heavy.push_back(i);
      else if (norb[i] == 1) light.push_back(i);
      else return hipSuccess;  // an atom with another shell count: the element and pair kernels
    }
    const int pA = (int)((heavy.size() + LGS_CHUNK - 1) / LGS_CHUNK * LGS_CHUNK);
    const int pB = (int)((light.size() + LGS_CHUNK - 1) / LGS_CHUNK * LGS_CHUNK);
    std::vector<int> cls((size_t)pA + pB, -1);
    std::copy(heavy.begin(), heavy.end(), cls.begin());
    std::copy(light.begin(), light.end(), cls.begin() + pA);
    // class-sorted basis: new AO offsets and the permutation (new -> old, old -> new)
    std::vector<int> ao_new(na), n2o(n), o2n(n);
    {
      int ao_old = 0;
      std::vector<int> ao_old_of(na);
      for (int i = 0; i < na; ++i) {
        ao_old_of[i] = ao_old;
        ao_old += norb[i];
      }
      const int nh = (int)heavy.size();
      for (int r = 0; r < nh; ++r) ao_new[heavy[r]] = 4 * r;
      for (int r = 0; r < (int)light.size(); ++r) ao_new[light[r]] = 4 * nh + r;
      for (int i = 0; i < na; ++i)
        for (int k = 0; k < norb[i]; ++k) {
          n2o[ao_new[i] + k] = ao_old_of[i] + k;
          o2n[ao_old_of[i] + k] = ao_new[i] + k;
        }
    }
    std::vector<int64_t> off(na);
    int64_t total = 0;
    for (int i = 0; i < na; ++i) {
      off[i] = total;
      const int64_t npa = norb[i] == 4 ? 10 : 1;
      total += npa * ((int64_t)pA * 10 + pB);
    }
    if (lw->lgs_st.ensure((size_t)total * sizeof(double)) != hipSuccess ||
        (mixed && lw->lgs_st32.ensure((size_t)total * sizeof(float)) != hipSuccess) ||
        lw->lgs_cls.ensure(cls.size() * sizeof(int)) != hipSuccess ||
        lw->lgs_off.ensure(off.size() * sizeof(int64_t)) != hipSuccess ||
        lw->lgs_perm.ensure((size_t)2 * n * sizeof(int)) != hipSuccess ||
        lw->lgs_jpart.ensure((size_t)10 * std::max<size_t>(heavy.size(), 1) * sizeof(double)) != hipSuccess ||
        lw->lgs_ao.ensure((size_t)(M.atom0 + na) * sizeof(int)) != hipSuccess) {
      (void)hipGetLastError();  // no room for the streams: the element and pair kernels
      return hipSuccess;
    }
    if ((e = hipMemcpyAsync(lw->lgs_cls.p, cls.data(), cls.size() * sizeof(int), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(lw->lgs_off.p, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(lw->lgs_perm.p, n2o.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(static_cast<int*>(lw->lgs_perm.p) + n, o2n.data(), (size_t)n * sizeof(int),
                            hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(static_cast<int*>(lw->lgs_ao.p) + M.atom0, ao_new.data(), (size_t)na * sizeof(int),
                            hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipEventRecord(lw->ev[0], s)) != hipSuccess)
      return e;
    const int* dcls = static_cast<const int*>(lw->lgs_cls.p);
    const int64_t* doff = static_cast<const int64_t*>(lw->lgs_off.p);
    hipLaunchKernelGGL(k_lgs_build, dim3((pA + pB) / LGS_CHUNK, na), dim3(256), 0, s, mi, b, dcls, pA, pB, doff,
                       static_cast<double*>(lw->lgs_st.p), mixed ? static_cast<float*>(lw->lgs_st32.p) : nullptr);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipEventRecord(lw->ev[1], s)) != hipSuccess ||
        (e = hipEventSynchronize(lw->ev[1])) != hipSuccess)
      return e;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, lw->ev[0], lw->ev[1]) == hipSuccess) phase[0] += ms;
    lgs_pA = pA;
    lgs_pB = pB;
    lgs_nA = (int)heavy.size();
    lgs_nB = (int)light.size();
    lgs_on = true;
    bp.atom_ao = static_cast<const int*>(lw->lgs_ao.p);
    lgs_n2o = static_cast<const int*>(lw->lgs_perm.p);
    lgs_o2n = lgs_n2o + n;
    if ((e = permute(slot(S_H), lgs_n2o, false)) != hipSuccess) return e;
    if (M.init_p && (e = permute(slot(S_HD), lgs_n2o, false)) != hipSuccess) return e;  // the seeded density
    return hipSuccess;
  };
  // F = H + G(P) of a closed shell (G with K(P/2)); low: reduced-precision accumulation.
  auto fock_rhf = [&](const double* P, double* F, bool lowp) -> hipError_t {
    const BatchDev& b = bp;  // the basis the SCF runs in
    if (lowp) hipLaunchKernelGGL(k_lg_pack<float>, dim3(grid_for(M.n_pp)), dim3(256), 0, s, mi, b, P);
    else hipLaunchKernelGGL(k_lg_pack<double>, dim3(grid_for(M.n_pp)), dim3(256), 0, s, mi, b, P);
    const int* dcls = static_cast<const int*>(lw->lgs_cls.p);
    const int64_t* doff = static_cast<const int64_t*>(lw->lgs_off.p);
    auto streamed = [&](auto* st) {
      using T = std::remove_const_t<std::remove_pointer_t<decltype(st)>>;
      T* jp = static_cast<T*>(lw->lgs_jpart.p);
      if (lgs_nA + lgs_nB > 0)
        hipLaunchKernelGGL((k_lgs_fock_light<T>), dim3(lgs_nA + lgs_nB), dim3(LGS_THREADS), 0, s, mi, b, st, dcls,
                           lgs_pA, lgs_pB, lgs_nA, doff, P, 0.5, F, jp);
      if (lgs_nA > 0)
        hipLaunchKernelGGL((k_lgs_fock_heavy<T>), dim3(lgs_nA), dim3(LGS_THREADS), 0, s, mi, b, st, dcls, lgs_pA,
                           lgs_pB, doff, P, 0.5, F, jp);
    };
    if (lgs_on && lowp)
      streamed(static_cast<const float*>(lw->lgs_st32.p));
    else if (lgs_on)
      streamed(static_cast<const double*>(lw->lgs_st.p));
    else if (lowp) {
      hipLaunchKernelGGL(k_lg_fock_diag<float>, dim3(grid_for((int64_t)M.n_pp * 64)), dim3(256), 0, s, mi, b, P, P,
                         0.5, nullptr, 0.0, F, nullptr);
      hipLaunchKernelGGL(k_lg_fock_pairs<float>, dim3(grid_for((int64_t)M.n_atoms * M.n_atoms / 2)), dim3(256), 0, s,
                         mi, b, P, 0.5, nullptr, 0.0, F, nullptr);
    } else {
      hipLaunchKernelGGL(k_lg_fock_diag<double>, dim3(grid_for((int64_t)M.n_pp * 64)), dim3(256), 0, s, mi, b, P, P,
                         0.5, nullptr, 0.0, F, nullptr);
      hipLaunchKernelGGL(k_lg_fock_pairs<double>, dim3(grid_for((int64_t)M.n_atoms * M.n_atoms / 2)), dim3(256), 0,
                         s, mi, b, P, 0.5, nullptr, 0.0, F, nullptr);
    }
    return hipGetLastError();
  };
  LargeCtl host{};
  int iter = 0;
  int low = mixed_ok ? 1 : 0;
  int cslot_c = S_C0;  // Ct lives here during the SCF
  bool last_rot = false;  // the last SCF step rotated the orbitals instead of diagonalising
  double dp_prev = 1.0e30;
  if (!M.uhf) {
    const int nocc = M.n_occ;
    LG(lgs_setup());
    if (!M.init_p) hipLaunchKernelGGL(k_lg_sad, dim3(grid_for(M.n_atoms)), dim3(256), 0, s, mi, bp, slot(S_HD), 1.0);
    for (; iter < M.max_scf; ++iter) {
      const int h = iter % HIST;
      const int cnt = std::min(iter + 1, HIST);
      double* P = slot(S_HD + h);
      double* F = slot(S_HF + h);
      double* E = slot(S_HE + h);
      LG(hipEventRecord(lw->ev[0], s));
      LG(fock_rhf(P, F, low != 0));
      LG(hipEventRecord(lw->ev[1], s));
      // energy and commutator
      LG(dots(2, P, slot(S_H), P, F, nullptr, nullptr, d3));
      hipLaunchKernelGGL(k_lg_energy, dim3(1), dim3(64), 0, s, ctl, d3, 0);
      LGB(gemm(rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, P, F, 0.0, slot(S_T)));  // row-major FP
      hipLaunchKernelGGL(k_lg_comm, dim3(RB), dim3(256), 0, s, n, ld, slot(S_T), E, part);
      hipLaunchKernelGGL(k_lg_finish, dim3(1), dim3(64), 0, s, RB, 1, part, d3 + 4);
      hipLaunchKernelGGL(k_lg_err, dim3(1), dim3(64), 0, s, ctl, d3 + 4, 1);
      LG(hipEventRecord(lw->ev[2], s));
      // history dots of the new entry against every live one (ring positions as in k_lg_ctrl_a)
      const int head_next = (h + 1) % HIST;
      for (int q = 0; q < cnt; ++q) {
        const int t = (head_next - cnt + q + 2 * HIST) % HIST;
        LG(dots(3, E, slot(S_HE + t), P, slot(S_HF + t), slot(S_HD + t), F, ctl_dots + 3 * q));
      }
      hipLaunchKernelGGL(k_lg_ctrl_a, dim3(1), dim3(64), 0, s, ctl, M.use_diis, M.accel, M.adiis_switch);
      hipLaunchKernelGGL(k_lg_combine, dim3(grid_for(nn)), dim3(256), 0, s, n, ld, ctl, slot(S_HF), F, P, 2.0,
                         slot(S_W), M.use_diis);
      LG(hipGetLastError());
      LG(hipEventRecord(lw->ev[3], s));
      double* eps = b.vecs + M.vec0;
      // PRECISION_MIXED keeps rotating after the switch to FP64 (rocBLAS dgemm); the reported
      // orbitals come from one full eigensolve of the converged Fock after the loop.
      last_rot = (low || mixed) && iter > 0 && dp_prev < M.rotate_below;
      if (last_rot) {
        LG(eig_skip());
        LG(eig_rotate(slot(S_W), nocc, slot(cslot_c)));
      } else {
        LG(eig_round(1, slot(S_W), eps, slot(cslot_c), nullptr, nullptr, nullptr));
      }
      LG(hipEventRecord(lw->ev[4], s));
      double* Pn = slot(S_HD + head_next);
      LG(density(slot(cslot_c), nocc, 2.0, Pn, low != 0));
      hipLaunchKernelGGL(k_lg_diff, dim3(RB), dim3(256), 0, s, n, ld, Pn, P, part);
      hipLaunchKernelGGL(k_lg_finish, dim3(1), dim3(64), 0, s, RB, 1, part, ctl_dots);
      hipLaunchKernelGGL(k_lg_ctrl_b, dim3(1), dim3(64), 0, s, ctl, iter, nn, M.e_tol, M.p_tol, tail, M.mixed_switch,
                         0);
      LG(hipEventRecord(lw->ev[5], s));
      LG(hipMemcpyAsync(&host, ctl, sizeof host, hipMemcpyDeviceToHost, s));
      LG(hipStreamSynchronize(s));
      float ms;
      for (int k = 0; k < 5; ++k)
        if (hipEventElapsedTime(&ms, lw->ev[k], lw->ev[k + 1]) == hipSuccess) phase[k] += ms;
      (void)hipGetLastError();  // a failed timing query must not become the next launch's error
      low = host.low;
      dp_prev = host.dp;
      if (host.done) {
        ++iter;
        break;
      }
    }
    leave.now();
    // Final energy and (with a level shift) the unshifted orbitals from the converged density.
    const int hp = iter % HIST;
    double* P = slot(S_HD + hp);
    double* F = slot(S_F);
    LG(fock_rhf(P, F, false));
    LG(dots(2, P, slot(S_H), P, F, nullptr, nullptr, d3));
    hipLaunchKernelGGL(k_lg_energy, dim3(1), dim3(64), 0, s, ctl, d3, 0);
    int pslot = S_HD + hp;
    // A mixed run always ends with a full FP64 Roothaan step; its size is the stated error
    // against FP64 (as in the megakernel).
    const bool check = mixed && host.guard.shift == 0.0;
    hipLaunchKernelGGL(k_lg_check, dim3(1), dim3(64), 0, s, ctl, 0, 0, part);
    if (host.guard.shift > 0.0 || last_rot || check) {
      hipLaunchKernelGGL(k_lg_copy, dim3(grid_for(sq)), dim3(256), 0, s, sq, F, slot(S_W));
      LG(eig_full(slot(S_W), b.vecs + M.vec0, slot(cslot_c)));
      if ((last_rot || check) && host.guard.shift == 0.0) {
        // The reported orbitals diagonalise the converged Fock; the reported density and energy
        // are theirs (one more Roothaan step from the converged density).
        pslot = S_HD + (hp + 1) % HIST;
        double* P2 = slot(pslot);
        LG(density(slot(cslot_c), M.n_occ, 2.0, P2, false));
        LG(fock_rhf(P2, F, false));
        LG(dots(2, P2, slot(S_H), P2, F, nullptr, nullptr, d3));
        hipLaunchKernelGGL(k_lg_energy, dim3(1), dim3(64), 0, s, ctl, d3, 0);
        if (check) {
          hipLaunchKernelGGL(k_lg_maxdiff, dim3(RB), dim3(256), 0, s, n, ld, P2, P, part);
          hipLaunchKernelGGL(k_lg_check, dim3(1), dim3(64), 0, s, ctl, 1, RB, part);
        }
      }
    }
    if (lgs_on) {  // back to the caller's basis: the density, the orbitals (AO columns of Ct), H
      LG(permute(slot(pslot), lgs_o2n, false));
      LG(permute(slot(cslot_c), lgs_o2n, true));
      LG(permute(slot(S_H), lgs_o2n, false));
    }
    // C = Ct^T into the C1 slot for the output gather.
    hipLaunchKernelGGL(k_lg_transpose, dim3(grid_for(sq)), dim3(256), 0, s, ld, slot(cslot_c), slot(S_C1));
    LG(hipMemcpyAsync(&host, ctl, sizeof host, hipMemcpyDeviceToHost, s));
    LG(hipStreamSynchronize(s));
    out.p_slot = pslot;
    out.c_slot = S_C1;
    out.pb_slot = -1;
    out.cb_slot = -1;
  } else {
    const int na = M.n_occ, nb = M.n_occ_b;
    const double ntot = (double)std::max(na + nb, 1);
    double *Pa = slot(U_PA), *Pb = slot(U_PB), *Pan = slot(U_PAN), *Pbn = slot(U_PBN), *Pt = slot(U_PT);
    double *Fa = slot(U_FA), *Fb = slot(U_FB);
    if (!M.init_p) {  // else the caller's alpha and beta densities are there
      hipLaunchKernelGGL(k_lg_sad, dim3(grid_for(M.n_atoms)), dim3(256), 0, s, mi, b, Pa, na / ntot);
      hipLaunchKernelGGL(k_lg_sad, dim3(grid_for(M.n_atoms)), dim3(256), 0, s, mi, b, Pb, nb / ntot);
    }
    double* eps_a = b.vecs + M.vec0;
    double* eps_b = eps_a + ld;
    for (; iter < M.max_scf; ++iter) {
      LG(hipEventRecord(lw->ev[0], s));
      // Pt = Pa + Pb via two library updates: Pt = Pa; Pt += Pb
      hipLaunchKernelGGL(k_lg_copy, dim3(grid_for(sq)), dim3(256), 0, s, sq, Pa, Pt);
      LGB(rocblas_daxpy(lw->blas, (rocblas_int)sq, &one, Pb, 1, Pt, 1));
      if (low) {
        hipLaunchKernelGGL(k_lg_pack<float>, dim3(grid_for(M.n_pp)), dim3(256), 0, s, mi, b, Pt);
        hipLaunchKernelGGL(k_lg_fock_diag<float>, dim3(grid_for((int64_t)M.n_pp * 64)), dim3(256), 0, s, mi, b, Pt, Pa, 1.0, Pb, 1.0, Fa, Fb);
        hipLaunchKernelGGL(k_lg_fock_pairs<float>, dim3(grid_for((int64_t)M.n_atoms * M.n_atoms / 2)), dim3(256), 0, s, mi, b, Pa, 1.0, Pb, 1.0, Fa, Fb);
      } else {
        hipLaunchKernelGGL(k_lg_pack<double>, dim3(grid_for(M.n_pp)), dim3(256), 0, s, mi, b, Pt);
        hipLaunchKernelGGL(k_lg_fock_diag<double>, dim3(grid_for((int64_t)M.n_pp * 64)), dim3(256), 0, s, mi, b, Pt, Pa, 1.0, Pb, 1.0, Fa, Fb);
        hipLaunchKernelGGL(k_lg_fock_pairs<double>, dim3(grid_for((int64_t)M.n_atoms * M.n_atoms / 2)), dim3(256), 0, s, mi, b, Pa, 1.0, Pb, 1.0, Fa, Fb);
      }
      LG(hipGetLastError());
      LG(hipEventRecord(lw->ev[1], s));
      LG(dots(3, Pt, slot(U_H), Pa, Fa, Pb, Fb, d3));
      hipLaunchKernelGGL(k_lg_energy, dim3(1), dim3(64), 0, s, ctl, d3, 1);
      LGB(gemm(rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, Pa, Fa, 0.0, slot(U_T)));
      hipLaunchKernelGGL(k_lg_comm, dim3(RB), dim3(256), 0, s, n, ld, slot(U_T), nullptr, part);
      hipLaunchKernelGGL(k_lg_finish, dim3(1), dim3(64), 0, s, RB, 1, part, d3 + 4);
      LGB(gemm(rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, Pb, Fb, 0.0, slot(U_T)));
      hipLaunchKernelGGL(k_lg_comm, dim3(RB), dim3(256), 0, s, n, ld, slot(U_T), nullptr, part);
      hipLaunchKernelGGL(k_lg_finish, dim3(1), dim3(64), 0, s, RB, 1, part, d3 + 5);
      hipLaunchKernelGGL(k_lg_err, dim3(1), dim3(64), 0, s, ctl, d3 + 4, 2);
      LG(hipEventRecord(lw->ev[2], s));
      // level shift in place (no DIIS for UHF): F += shift (I - P)
      hipLaunchKernelGGL(k_lg_combine, dim3(grid_for(nn)), dim3(256), 0, s, n, ld, ctl, slot(S_HF), Fa, Pa, 1.0, Fa, 0);
      hipLaunchKernelGGL(k_lg_combine, dim3(grid_for(nn)), dim3(256), 0, s, n, ld, ctl, slot(S_HF), Fb, Pb, 1.0, Fb, 0);
      LG(hipEventRecord(lw->ev[3], s));
      last_rot = (low || mixed) && iter > 0 && dp_prev < M.rotate_below;
      if (last_rot) {
        LG(eig_skip());
        LG(eig_rotate(Fa, na, slot(U_CA0)));
        LG(eig_rotate(Fb, nb, slot(U_CB0)));
      } else {
        LG(eig_round(2, Fa, eps_a, slot(U_CA0), Fb, eps_b, slot(U_CB0)));
      }
      LG(hipEventRecord(lw->ev[4], s));
      LG(density(slot(U_CA0), na, 1.0, Pan, low != 0));
      LG(density(slot(U_CB0), nb, 1.0, Pbn, low != 0));
      hipLaunchKernelGGL(k_lg_diff, dim3(RB), dim3(256), 0, s, n, ld, Pan, Pa, part);
      hipLaunchKernelGGL(k_lg_finish, dim3(1), dim3(64), 0, s, RB, 1, part, ctl_dots);
      hipLaunchKernelGGL(k_lg_diff, dim3(RB), dim3(256), 0, s, n, ld, Pbn, Pb, part);
      hipLaunchKernelGGL(k_lg_finish, dim3(1), dim3(64), 0, s, RB, 1, part, ctl_dots + 1);
      hipLaunchKernelGGL(k_lg_ctrl_b, dim3(1), dim3(64), 0, s, ctl, iter, nn, M.e_tol, M.p_tol, tail, M.mixed_switch,
                         1);
      LG(hipEventRecord(lw->ev[5], s));
      LG(hipMemcpyAsync(&host, ctl, sizeof host, hipMemcpyDeviceToHost, s));
      LG(hipStreamSynchronize(s));
      float ms;
      for (int k = 0; k < 5; ++k)
        if (hipEventElapsedTime(&ms, lw->ev[k], lw->ev[k + 1]) == hipSuccess) phase[k] += ms;
      (void)hipGetLastError();  // a failed timing query must not become the next launch's error
      std::swap(Pa, Pan);
      std::swap(Pb, Pbn);
      low = host.low;
      dp_prev = host.dp;
      if (host.done) {
        ++iter;
        break;
      }
    }
    leave.now();
    hipLaunchKernelGGL(k_lg_copy, dim3(grid_for(sq)), dim3(256), 0, s, sq, Pa, Pt);
    LGB(rocblas_daxpy(lw->blas, (rocblas_int)sq, &one, Pb, 1, Pt, 1));
    hipLaunchKernelGGL(k_lg_pack<double>, dim3(grid_for(M.n_pp)), dim3(256), 0, s, mi, b, Pt);
    hipLaunchKernelGGL(k_lg_fock_diag<double>, dim3(grid_for((int64_t)M.n_pp * 64)), dim3(256), 0, s, mi, b, Pt, Pa, 1.0, Pb, 1.0, Fa, Fb);
        hipLaunchKernelGGL(k_lg_fock_pairs<double>, dim3(grid_for((int64_t)M.n_atoms * M.n_atoms / 2)), dim3(256), 0, s, mi, b, Pa, 1.0, Pb, 1.0, Fa, Fb);
    LG(dots(3, Pt, slot(U_H), Pa, Fa, Pb, Fb, d3));
    hipLaunchKernelGGL(k_lg_energy, dim3(1), dim3(64), 0, s, ctl, d3, 1);
    const bool check = mixed && host.guard.shift == 0.0;
    hipLaunchKernelGGL(k_lg_check, dim3(1), dim3(64), 0, s, ctl, 0, 0, part);
    if (host.guard.shift > 0.0 || last_rot || check) LG(eig_full(Fa, eps_a, slot(U_CA0)));
    hipLaunchKernelGGL(k_lg_copy, dim3(grid_for(sq)), dim3(256), 0, s, sq, Fb, slot(U_T));
    LG(eig_full(slot(U_T), eps_b, slot(U_CB0)));
    if ((last_rot || check) && host.guard.shift == 0.0) {
      // Densities and energy of the reported orbitals (see the RHF branch).
      LG(density(slot(U_CA0), na, 1.0, Pan, false));
      LG(density(slot(U_CB0), nb, 1.0, Pbn, false));
      if (check) {
        hipLaunchKernelGGL(k_lg_maxdiff, dim3(RB), dim3(256), 0, s, n, ld, Pan, Pa, part);
        hipLaunchKernelGGL(k_lg_check, dim3(1), dim3(64), 0, s, ctl, 1, RB, part);
        hipLaunchKernelGGL(k_lg_maxdiff, dim3(RB), dim3(256), 0, s, n, ld, Pbn, Pb, part);
        hipLaunchKernelGGL(k_lg_check, dim3(1), dim3(64), 0, s, ctl, 1, RB, part);
      }
      std::swap(Pa, Pan);
      std::swap(Pb, Pbn);
      hipLaunchKernelGGL(k_lg_copy, dim3(grid_for(sq)), dim3(256), 0, s, sq, Pa, Pt);
      LGB(rocblas_daxpy(lw->blas, (rocblas_int)sq, &one, Pb, 1, Pt, 1));
      hipLaunchKernelGGL(k_lg_pack<double>, dim3(grid_for(M.n_pp)), dim3(256), 0, s, mi, b, Pt);
      hipLaunchKernelGGL(k_lg_fock_diag<double>, dim3(grid_for((int64_t)M.n_pp * 64)), dim3(256), 0, s, mi, b, Pt, Pa, 1.0, Pb, 1.0, Fa, Fb);
      hipLaunchKernelGGL(k_lg_fock_pairs<double>, dim3(grid_for((int64_t)M.n_atoms * M.n_atoms / 2)), dim3(256), 0, s, mi, b, Pa, 1.0, Pb, 1.0, Fa, Fb);
      LG(dots(3, Pt, slot(U_H), Pa, Fa, Pb, Fb, d3));
      hipLaunchKernelGGL(k_lg_energy, dim3(1), dim3(64), 0, s, ctl, d3, 1);
      if (check) hipLaunchKernelGGL(k_lg_check, dim3(1), dim3(64), 0, s, ctl, 1, 0, part);
    }
    hipLaunchKernelGGL(k_lg_transpose, dim3(grid_for(sq)), dim3(256), 0, s, ld, slot(U_CA0), slot(U_CA1));
    hipLaunchKernelGGL(k_lg_transpose, dim3(grid_for(sq)), dim3(256), 0, s, ld, slot(U_CB0), slot(U_CB1));
    LG(hipMemcpyAsync(&host, ctl, sizeof host, hipMemcpyDeviceToHost, s));
    LG(hipStreamSynchronize(s));
    out.p_slot = (Pa == slot(U_PA)) ? U_PA : U_PAN;
    out.pb_slot = (Pb == slot(U_PB)) ? U_PB : U_PBN;
    out.c_slot = U_CA1;
    out.cb_slot = U_CB1;
  }
  out.e_el = host.e_cur;
  out.iterations = iter > M.max_scf ? M.max_scf : iter;
  out.converged = host.done;
  out.dp_err = host.dp;
  out.err_norm = host.err;
  out.shift = host.guard.shift;
  out.e_fp32_last = host.e_low;
  out.fp32_iters = host.fp32_iters;
  out.check_de = host.check_de;
  out.check_dp = host.check_dp;
  out.sweeps = 0;
  for (int k = 0; k < NPHASE; ++k) out.ticks[k] = 0;
  // phases: fock, energy+commutator, diis, eigensolve, density (ms -> 100 MHz ticks)
  for (int k = 0; k < 5; ++k) out.ticks[k] = (unsigned long long)(phase[k] * 1.0e5);
  double tot = 0.0;
  for (int k = 0; k < 5; ++k) tot += phase[k];
  out.ticks[PH_TOTAL] = (unsigned long long)(tot * 1.0e5);
  out.ticks[PH_EIG_SWEEP] = out.ticks[PH_EIG];
  (void)cslot_c;
  return hipSuccess;
}

}  // namespace pm7

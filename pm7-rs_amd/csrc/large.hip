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
// For molecules whose atoms all carry 1 or 4 orbitals. The atoms form two classes, A (4 orbitals)
// and B (1 orbital), each in atom order; an atom's position there is its rank, and its sorted index
// u is its rank (A) or nA + rank (B); the SCF runs in the AO basis of that order. Every atom pair
// is taken once, by the atom of lower sorted index, a, which is also the pair's first atom in the
// pair's own order w[e_F][kl_S] (A before B; within a class the lower index). Partners are grouped
// in chunks of 64 of one class aligned on rank (chunk c holds ranks 64c..64c+63), so partner b sits
// at lane rank_b % 64 of chunk rank_b / 64 in every stream. a's stream, built once per SCF, holds
// the chunks from the one holding a's first later partner: its class-A chunks (A atoms only), then
// its class-B chunks; element k of lane t at [chunk][k][t], zero for lanes that are not a later
// partner. Each pair's integrals are read once per Fock build.
//
// One wavefront per (a, chunk): the exchange block F(a, b) = H(a, b) - xa K_ab, stored at (a, b) and
// mirrored at (b, a), so F is exactly symmetric; a's Coulomb sum over the chunk (a fixed shuffle
// tree) into jrow; each partner's Coulomb sum from a's packed density into jcol. k_lgs_colsum adds
// the jcol entries of every partner chunk in fixed parts, and k_lgs_diag adds the parts and the jrow
// entries of each atom in a fixed order and writes its diagonal block with the one-centre terms.

constexpr int LGS_CHUNK = 64;
constexpr int LGS_NPART = 32;  // fixed parts of k_lgs_colsum's sums per partner chunk

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

struct LgsDims {
  int nA, nB, nchA, nchB;  // atoms and chunks per class
};
struct LgsAtom {           // per sorted index
  int64_t soff, joff;      // offsets of its stream and of its jcol entries
  int la, rank;            // molecule-local atom, rank in its class
};
// The chunk holding the first later partner of rank r in its own class.
__host__ __device__ inline int lgs_c0(int r) { return (r + 1) / LGS_CHUNK; }

// Orbital count of every atom of molecule mi, in molecule order.
__global__ void k_lgs_norb(int mi, BatchDev b, int* __restrict__ norb) {
  const MolDev M = b.mols[mi];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M.n_atoms; i += gridDim.x * blockDim.x)
    norb[i] = b.elem[b.atom_z[M.atom0 + i]].n_orb;
}

// Workgroup per (local chunk j, atom u): the 64 partners' blocks read one after another (each
// contiguous), transposed through LDS, written element-major; FP64 always, FP32 when st32 is set.
__global__ void __launch_bounds__(256) k_lgs_build(int mi, BatchDev b, LgsDims d, const LgsAtom* __restrict__ at,
                                                   const int* __restrict__ srt, double* __restrict__ st,
                                                   float* __restrict__ st32) {
  constexpr int TS = LGS_CHUNK + 1;  // padded row: the transposing writes spread over the LDS banks
  __shared__ double tile[100 * TS];
  __shared__ int64_t s_woff[LGS_CHUNK];
  const MolDev M = b.mols[mi];
  const int u = blockIdx.y, j = blockIdx.x, na = M.n_atoms;
  const LgsAtom A = at[u];
  const bool aA = u < d.nA;
  const int npa = aA ? 10 : 1;
  const int nsegA = aA ? d.nchA - lgs_c0(A.rank) : 0;
  const int c0B = aA ? 0 : lgs_c0(A.rank);
  if (j >= nsegA + d.nchB - c0B) return;  // the whole workgroup
  const bool inA = j < nsegA;
  const int c = inA ? lgs_c0(A.rank) + j : c0B + (j - nsegA);
  const int K = npa * (inA ? 10 : 1);
  const int64_t base = A.soff + (inA ? (int64_t)j * K * LGS_CHUNK
                                     : (int64_t)nsegA * npa * 10 * LGS_CHUNK + (int64_t)(j - nsegA) * K * LGS_CHUNK);
  if (threadIdx.x < LGS_CHUNK) {
    const int rb = c * LGS_CHUNK + threadIdx.x;
    const bool later = rb < (inA ? d.nA : d.nB) && (aA != inA || rb > A.rank);
    int64_t wo = -1;
    if (later) {
      const int lb = srt[(inA ? 0 : d.nA) + rb];
      wo = b.pairs[M.pair0 + tri_index(min(A.la, lb), max(A.la, lb), na)].w_off;
    }
    s_woff[threadIdx.x] = wo;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < K * LGS_CHUNK; i += 256) {  // k fastest: each block read contiguously
    const int t = i / K, k = i - t * K;
    const int64_t wo = s_woff[t];
    tile[k * TS + t] = wo >= 0 ? b.w[wo + k] : 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < K * LGS_CHUNK; i += 256) {  // element-major, coalesced
    const double v = tile[(i / LGS_CHUNK) * TS + i % LGS_CHUNK];
    st[base + i] = v;
    if (st32) st32[base + i] = (float)v;
  }
}

// One (a, chunk) item: a has NFA orbitals, the chunk's partners NFB; lane `lane` takes partner lb
// (live: a later partner). sp, jc: this item's stream and jcol entries at the lane; jr: a's jrow
// entry of this chunk.
template <typename T, int NFA, int NFB>
__device__ __forceinline__ void lgs_item(const MolDev& M, const BatchDev& b, int la, const T* __restrict__ sp,
                                         int lb, bool live, const double* __restrict__ P, double xa,
                                         double* __restrict__ F, T* __restrict__ jc, T* __restrict__ jr) {
  constexpr int NPA = NFA * (NFA + 1) / 2, NPB = NFB * (NFB + 1) / 2;
  const int ld = M.ld;
  const int a = M.atom0 + la;
  const int oa = b.atom_ao[a];
  const int bb = M.atom0 + (live ? lb : la);  // a valid atom for the loads of idle lanes
  const int ob = b.atom_ao[bb];
  const double* ppa = b.slab + M.pp0 + b.atom_pp[a];
  const double* ppb = b.slab + M.pp0 + b.atom_pp[bb];
  // P[a_p][b_q] at a's rows: in the class-sorted basis the partners of a chunk are contiguous in
  // AO order, so these loads and the row stores below are coalesced.
  T pab[NFA][NFB], acc[NFA][NFB];
#pragma unroll
  for (int p = 0; p < NFA; ++p)
#pragma unroll
    for (int q = 0; q < NFB; ++q) {
      pab[p][q] = (T)P[(int64_t)(oa + p) * ld + ob + q];
      acc[p][q] = (T)0;
    }
  T pa[NPA], pb[NPB], ja[NPA], jb[NPB];
#pragma unroll
  for (int k = 0; k < NPA; ++k) {
    pa[k] = (T)ppa[k];
    ja[k] = (T)0;
  }
#pragma unroll
  for (int k = 0; k < NPB; ++k) {
    pb[k] = (T)ppb[k];
    jb[k] = (T)0;
  }
  // Every index below is a compile-time constant (static_for), so the block lives in registers.
  static_for<0, NPA>([&](auto ec) {
    constexpr int e = decltype(ec)::value, mu = ce_hi(e), p = e - mu * (mu + 1) / 2;
    static_for<0, NPB>([&](auto kc) {
      constexpr int kl = decltype(kc)::value, l = ce_hi(kl), q = kl - l * (l + 1) / 2;
      const T w = sp[(e * NPB + kl) * LGS_CHUNK];
      ja[e] += w * pb[kl];
      jb[kl] += w * pa[e];
      // exchange: every (mu' nu') x (la' sg') term this packed element stands for
      acc[mu][l] += pab[p][q] * w;
      if constexpr (mu != p) acc[p][l] += pab[mu][q] * w;
      if constexpr (l != q) acc[mu][q] += pab[p][l] * w;
      if constexpr (mu != p && l != q) acc[p][q] += pab[mu][l] * w;
    });
  });
#pragma unroll
  for (int k = 0; k < NPB; ++k) jc[k * LGS_CHUNK] = jb[k];  // zero on idle lanes (zero integrals)
#pragma unroll
  for (int k = 0; k < NPA; ++k) {
    T v = ja[k];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);  // the same tree on every lane
    ja[k] = v;
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < NPA; ++k) jr[k] = ja[k];
  }
  if (!live) return;
  const double* H = b.slab + M.mat0 + (int64_t)S_H * ld * ld;
#pragma unroll
  for (int p = 0; p < NFA; ++p)
#pragma unroll
    for (int q = 0; q < NFB; ++q) {
      const int64_t at = (int64_t)(oa + p) * ld + ob + q;
      const double f = H[at] - (double)((T)xa * acc[p][q]);
      F[at] = f;
      F[(int64_t)(ob + q) * ld + oa + p] = f;
    }
}

// Class-A atoms against class-A chunks: workgroup i is (group g of four A atoms, chunk c) = wg[i];
// wavefront w takes atom u = 4g + w, so the mirrored stores of one workgroup fill 128-byte rows.
template <typename T>
__global__ void __launch_bounds__(256) k_lgs_fock_aa(int mi, BatchDev b, LgsDims d, const LgsAtom* __restrict__ at,
                                                     const int* __restrict__ srt, const int2* __restrict__ wg,
                                                     const T* __restrict__ st, const double* __restrict__ P,
                                                     double xa, double* __restrict__ F, T* __restrict__ jcol,
                                                     T* __restrict__ jrow) {
  const MolDev M = b.mols[mi];
  const int2 g = wg[blockIdx.x];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int u = 4 * g.x + w, c = g.y;
  if (u >= d.nA) return;
  const LgsAtom A = at[u];
  const int c0 = lgs_c0(A.rank);
  if (c < c0) return;
  const int rb = c * LGS_CHUNK + lane;
  const bool live = rb < d.nA && rb > A.rank;
  const int lb = live ? srt[rb] : -1;
  lgs_item<T, 4, 4>(M, b, A.la, st + A.soff + (int64_t)(c - c0) * 100 * LGS_CHUNK + lane, lb, live, P, xa, F,
                    jcol + A.joff + (int64_t)(c - c0) * 10 * LGS_CHUNK + lane,
                    jrow + ((int64_t)u * d.nchA + c) * 10);
}

// Class-B chunks: wg[i] = (g | kind << 24, chunk c); kind 0: the four A atoms 4g.., kind 1: the four
// B atoms of ranks 4g... jrow: per (u, class-B chunk), 10 words.
template <typename T>
__global__ void __launch_bounds__(256) k_lgs_fock_xb(int mi, BatchDev b, LgsDims d, const LgsAtom* __restrict__ at,
                                                     const int* __restrict__ srt, const int2* __restrict__ wg,
                                                     const T* __restrict__ st, const double* __restrict__ P,
                                                     double xa, double* __restrict__ F, T* __restrict__ jcol,
                                                     T* __restrict__ jrow) {
  const MolDev M = b.mols[mi];
  const int2 g = wg[blockIdx.x];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int kind = g.x >> 24, r = 4 * (g.x & 0xffffff) + w, c = g.y;
  const int rb = c * LGS_CHUNK + lane;
  if (kind == 0) {
    if (r >= d.nA) return;
    const LgsAtom A = at[r];
    const int nsegA = d.nchA - lgs_c0(A.rank);
    const bool live = rb < d.nB;
    const int lb = live ? srt[d.nA + rb] : -1;
    lgs_item<T, 4, 1>(M, b, A.la,
                      st + A.soff + (int64_t)nsegA * 100 * LGS_CHUNK + (int64_t)c * 10 * LGS_CHUNK + lane, lb,
                      live, P, xa, F, jcol + A.joff + (int64_t)nsegA * 10 * LGS_CHUNK + (int64_t)c * LGS_CHUNK + lane,
                      jrow + ((int64_t)r * d.nchB + c) * 10);
  } else {
    if (r >= d.nB) return;
    const int u = d.nA + r;
    const LgsAtom A = at[u];
    const int c0 = lgs_c0(r);
    if (c < c0) return;
    const bool live = rb < d.nB && rb > r;
    const int lb = live ? srt[d.nA + rb] : -1;
    lgs_item<T, 1, 1>(M, b, A.la, st + A.soff + (int64_t)(c - c0) * LGS_CHUNK + lane, lb, live, P, xa, F,
                      jcol + A.joff + (int64_t)(c - c0) * LGS_CHUNK + lane, jrow + ((int64_t)u * d.nchB + c) * 10);
  }
}

// acc[k] += jcol entries k of the contributors i0 + w, i0 + w + 4, ... below i1 at lane `lane`;
// off(i) is contributor i's entry offset for the chunk.
template <int NPB, typename T, class Off>
__device__ __forceinline__ void lgs_colsum_range(T (&acc)[10], const T* __restrict__ jcol, int i0, int i1, int w,
                                                 int lane, Off off) {
#pragma unroll 4
  for (int i = i0 + w; i < i1; i += 4) {  // unrolled: several entries' offsets and loads in flight
    const int64_t o = off(i) + lane;
#pragma unroll
    for (int k = 0; k < NPB; ++k) acc[k] += jcol[o + k * LGS_CHUNK];
  }
}

// Workgroup per (partner chunk cg: class A chunks, then class B; part p): the jcol entries of the
// chunk from the atoms of sorted index in part p of [0, nc), wavefront w taking every fourth, then
// the four wavefronts in order: part[cg][p][k][lane].
template <typename T>
__global__ void __launch_bounds__(256) k_lgs_colsum(LgsDims d, const LgsAtom* __restrict__ at,
                                                    const T* __restrict__ jcol, T* __restrict__ part) {
  __shared__ T red[4][10][LGS_CHUNK];
  const int cg = blockIdx.x, p = blockIdx.y;
  const bool xa = cg < d.nchA;
  const int C = xa ? cg : cg - d.nchA;
  const int npb = xa ? 10 : 1;
  // earlier atoms whose streams hold chunk C: ranks up to 64C + 62 of the chunk's class
  const int lim = C * LGS_CHUNK + LGS_CHUNK - 1;
  const int nc = xa ? min(d.nA, lim) : d.nA + min(d.nB, lim);
  const int i0 = (int)((int64_t)nc * p / LGS_NPART), i1 = (int)((int64_t)nc * (p + 1) / LGS_NPART);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  T acc[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) acc[k] = (T)0;
  if (xa) {
    lgs_colsum_range<10>(acc, jcol, i0, i1, w, lane, [&](int i) {
      return at[i].joff + (int64_t)(C - lgs_c0(i)) * 10 * LGS_CHUNK;
    });
  } else {
    // the A atoms' class-B entries, then the earlier B atoms' (each range strided from its start)
    const int mA = min(i1, d.nA);
    lgs_colsum_range<1>(acc, jcol, i0, mA, w, lane, [&](int i) {
      return at[i].joff + (int64_t)(d.nchA - lgs_c0(i)) * 10 * LGS_CHUNK + (int64_t)C * LGS_CHUNK;
    });
    lgs_colsum_range<1>(acc, jcol, max(i0, d.nA), i1, w, lane, [&](int i) {
      return at[i].joff + (int64_t)(C - lgs_c0(i - d.nA)) * LGS_CHUNK;
    });
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) red[w][k][lane] = acc[k];
  __syncthreads();
  if (w != 0) return;
  for (int k = 0; k < npb; ++k)
    part[(((int64_t)cg * LGS_NPART + p) * 10 + k) * LGS_CHUNK + lane] =
        ((red[0][k][lane] + red[1][k][lane]) + red[2][k][lane]) + red[3][k][lane];
}

// Diagonal element e of an atom with NFA orbitals at AO oa (one-centre parameters g_ss, g_sp, g_pp,
// g_p2, h_sp at ocp): the one-centre terms as fock_diag_item (returns xa ko, sets jo) and H, at
// element i1 (and its mirror i2) of the block.
template <typename T, int NFA>
__device__ __forceinline__ T lgs_diag_oc(const MolDev& M, const BatchDev& b, int oa, const double* __restrict__ ocp,
                                         int e, const double* __restrict__ P, double xa, double& h, int64_t& i1,
                                         int64_t& i2, T& jo) {
  const int ld = M.ld;
  ElemDev ea;  // only the one-centre fields oc_int reads
  ea.g_ss = ocp[0];
  ea.g_sp = ocp[1];
  ea.g_pp = ocp[2];
  ea.g_p2 = ocp[3];
  ea.h_sp = ocp[4];
  int mu, nu;
  unpack_e(e, mu, nu);
  double pd[NFA][NFA];
#pragma unroll
  for (int l = 0; l < NFA; ++l)
#pragma unroll
    for (int s2 = 0; s2 < NFA; ++s2) pd[l][s2] = P[(int64_t)(oa + l) * ld + oa + s2];  // all in flight
  i1 = (int64_t)(oa + mu) * ld + oa + nu;
  i2 = (int64_t)(oa + nu) * ld + oa + mu;
  h = b.slab[M.mat0 + (int64_t)S_H * ld * ld + i1];
  T ko = (T)0;
  jo = (T)0;
#pragma unroll
  for (int l = 0; l < NFA; ++l)
#pragma unroll
    for (int s2 = 0; s2 < NFA; ++s2) {
      jo += (T)pd[l][s2] * (T)oc_int(mu, nu, l, s2, ea);
      ko += (T)pd[l][s2] * (T)oc_int(mu, l, nu, s2, ea);
    }
  return (T)xa * ko;
}

// One-centre parameters (g_ss, g_sp, g_pp, g_p2, h_sp) of every atom by sorted index, once per SCF.
__global__ void k_lgs_ocpar(int mi, BatchDev b, int na, const LgsAtom* __restrict__ at, double* __restrict__ ocp) {
  const MolDev M = b.mols[mi];
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < na; u += gridDim.x * blockDim.x) {
    const ElemDev& ea = b.elem[b.atom_z[M.atom0 + at[u].la]];
    ocp[5 * u] = ea.g_ss;
    ocp[5 * u + 1] = ea.g_sp;
    ocp[5 * u + 2] = ea.g_pp;
    ocp[5 * u + 3] = ea.g_p2;
    ocp[5 * u + 4] = ea.h_sp;
  }
}

// Sum of the entries v[idx(c)] for c in [c_begin, c_end), in order, eight loads in flight.
template <typename T, class Idx>
__device__ __forceinline__ T lgs_sum_chunks(const T* __restrict__ v, int c_begin, int c_end, Idx idx) {
  T sum = (T)0;
  for (int c = c_begin; c < c_end; c += 8) {
    T x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = c + j < c_end ? v[idx(c + j)] : (T)0;
#pragma unroll
    for (int j = 0; j < 8; ++j) sum += x[j];
  }
  return sum;
}

// Wavefront per (chunk cg as in k_lgs_colsum, packed pair k), lane per atom of the chunk: the
// atom's Coulomb sum of pair k (the parts, then its jrow entries by chunk) and its diagonal
// element k with the one-centre terms. Workgroups of one wavefront: the element and jrow accesses
// are one line per lane, so they are spread over as many CUs as possible.
template <typename T>
__global__ void __launch_bounds__(64) k_lgs_diag(int mi, BatchDev b, LgsDims d, const double* __restrict__ ocp,
                                                 const T* __restrict__ part, const T* __restrict__ jrowA,
                                                 const T* __restrict__ jrowB, const double* __restrict__ P,
                                                 double xa, double* __restrict__ F) {
  const MolDev M = b.mols[mi];
  const int cg = blockIdx.x, k = blockIdx.y, lane = threadIdx.x;
  const bool inA = cg < d.nchA;
  const int r = (inA ? cg : cg - d.nchA) * LGS_CHUNK + lane;
  if (r >= (inA ? d.nA : d.nB) || k >= (inA ? 10 : 1)) return;
  const int u = inA ? r : d.nA + r;
  double h;
  int64_t i1, i2;
  T jo;
  // in the class-sorted basis the atom's AOs start at 4 r (class A) or 4 nA + r (class B)
  const double* op = ocp + 5 * u;
  const T xko = inA ? lgs_diag_oc<T, 4>(M, b, 4 * r, op, k, P, xa, h, i1, i2, jo)
                    : lgs_diag_oc<T, 1>(M, b, 4 * d.nA + r, op, k, P, xa, h, i1, i2, jo);
  T J = (T)0;
  const T* pp = part + ((int64_t)cg * LGS_NPART * 10 + k) * LGS_CHUNK + lane;
#pragma unroll
  for (int p = 0; p < LGS_NPART; ++p) J += pp[(int64_t)p * 10 * LGS_CHUNK];
  if (inA) J += lgs_sum_chunks(jrowA, lgs_c0(r), d.nchA, [&](int c) { return ((int64_t)u * d.nchA + c) * 10 + k; });
  J += lgs_sum_chunks(jrowB, inA ? 0 : lgs_c0(r), d.nchB, [&](int c) { return ((int64_t)u * d.nchB + c) * 10 + k; });
  const double f = h + (double)(J + jo - xko);
  F[i1] = f;
  F[i2] = f;
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
  Buf lgs_st, lgs_st32, lgs_at, lgs_srt, lgs_wg, lgs_norb;  // per-atom integral streams (s/p closed shell)
  Buf lgs_jcol, lgs_jrow, lgs_part;                          // partial Coulomb sums
  Buf lgs_ocp;                                               // one-centre parameters by sorted index
  LargeCtl* flag_host = nullptr;  // pinned: the stop flag's copy stays asynchronous
  Buf lgs_perm, lgs_ao;                                      // class-sorted basis: permutations, AO offsets
  std::vector<hipEvent_t> ev;
};

LargeWork* large_work_create() {
  auto* w = new LargeWork();
  if (rocblas_create_handle(&w->blas) != rocblas_status_success) {
    delete w;
    return nullptr;
  }
  w->ev.resize(10);
  for (auto& e : w->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      large_work_destroy(w);
      return nullptr;
    }
  if (hipHostMalloc(reinterpret_cast<void**>(&w->flag_host), sizeof(LargeCtl), hipHostMallocDefault) != hipSuccess) {
    w->flag_host = nullptr;
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
  if (w->flag_host) (void)hipHostFree(w->flag_host);
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
  // Closed-shell s/p molecules: the per-atom integral streams (k_lgs_fock_aa, k_lgs_fock_xb,
  // k_lgs_colsum, k_lgs_diag), built once here, their
  // device time counted in the Fock phase. Other molecules, and PM7_AMD_LARGE_STREAMS=0, keep the
  // element and pair kernels (k_lg_fock_diag, k_lg_fock_pairs).
  LgsDims lgs_d{0, 0, 0, 0};
  int lgs_n_aa = 0, lgs_n_xb = 0;  // workgroups of the class-A and class-B chunk launches
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
    const int nA = (int)heavy.size(), nB = (int)light.size();
    const LgsDims d{nA, nB, (nA + LGS_CHUNK - 1) / LGS_CHUNK, (nB + LGS_CHUNK - 1) / LGS_CHUNK};
    std::vector<int> srt(heavy);
    srt.insert(srt.end(), light.begin(), light.end());
    // class-sorted basis: new AO offsets and the permutation (new -> old, old -> new)
    std::vector<int> ao_new(na), n2o(n), o2n(n);
    {
      int ao_old = 0;
      std::vector<int> ao_old_of(na);
      for (int i = 0; i < na; ++i) {
        ao_old_of[i] = ao_old;
        ao_old += norb[i];
      }
      for (int r = 0; r < nA; ++r) ao_new[heavy[r]] = 4 * r;
      for (int r = 0; r < nB; ++r) ao_new[light[r]] = 4 * nA + r;
      for (int i = 0; i < na; ++i)
        for (int k = 0; k < norb[i]; ++k) {
          n2o[ao_new[i] + k] = ao_old_of[i] + k;
          o2n[ao_old_of[i] + k] = ao_new[i] + k;
        }
    }
    // per sorted index: stream and jcol offsets (A: its class-A chunks of 100 words and class-B
    // chunks of 10 per lane; B: its class-B chunks of 1)
    std::vector<LgsAtom> at(na);
    int64_t total = 0, jtotal = 0;
    for (int u = 0; u < na; ++u) {
      const bool aA = u < nA;
      const int r = aA ? u : u - nA;
      const int64_t segA = aA ? d.nchA - lgs_c0(r) : 0, segB = aA ? d.nchB : d.nchB - lgs_c0(r);
      at[u] = LgsAtom{total, jtotal, srt[u], r};
      total += (aA ? segA * 100 + segB * 10 : segB) * LGS_CHUNK;
      jtotal += (aA ? segA * 10 + segB : segB) * LGS_CHUNK;
    }
    // workgroups: (group of four A atoms, class-A chunk), then (group | kind << 24, class-B chunk)
    std::vector<int2> wg;
    for (int g = 0; 4 * g < nA; ++g)
      for (int c = lgs_c0(4 * g); c < d.nchA; ++c) wg.push_back(make_int2(g, c));
    const int n_aa = (int)wg.size();
    for (int g = 0; 4 * g < nA; ++g)
      for (int c = 0; c < d.nchB; ++c) wg.push_back(make_int2(g, c));
    for (int g = 0; 4 * g < nB; ++g)
      for (int c = lgs_c0(4 * g); c < d.nchB; ++c) wg.push_back(make_int2(g | 1 << 24, c));
    const size_t jrow_words = (size_t)nA * d.nchA * 10 + (size_t)na * d.nchB * 10;
    if (lw->lgs_st.ensure((size_t)total * sizeof(double)) != hipSuccess ||
        (mixed && lw->lgs_st32.ensure((size_t)total * sizeof(float)) != hipSuccess) ||
        lw->lgs_at.ensure(at.size() * sizeof(LgsAtom)) != hipSuccess ||
        lw->lgs_srt.ensure(srt.size() * sizeof(int)) != hipSuccess ||
        lw->lgs_wg.ensure(std::max<size_t>(wg.size(), 1) * sizeof(int2)) != hipSuccess ||
        lw->lgs_jcol.ensure((size_t)jtotal * sizeof(double)) != hipSuccess ||
        lw->lgs_jrow.ensure(jrow_words * sizeof(double)) != hipSuccess ||
        lw->lgs_part.ensure((size_t)(d.nchA + d.nchB) * LGS_NPART * 10 * LGS_CHUNK * sizeof(double)) !=
            hipSuccess ||
        lw->lgs_perm.ensure((size_t)2 * n * sizeof(int)) != hipSuccess ||
        lw->lgs_ocp.ensure((size_t)5 * na * sizeof(double)) != hipSuccess ||
        lw->lgs_ao.ensure((size_t)(M.atom0 + na) * sizeof(int)) != hipSuccess) {
      (void)hipGetLastError();  // no room for the streams: the element and pair kernels
      return hipSuccess;
    }
    if ((e = hipMemcpyAsync(lw->lgs_at.p, at.data(), at.size() * sizeof(LgsAtom), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(lw->lgs_srt.p, srt.data(), srt.size() * sizeof(int), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (!wg.empty() &&
         (e = hipMemcpyAsync(lw->lgs_wg.p, wg.data(), wg.size() * sizeof(int2), hipMemcpyHostToDevice, s)) !=
             hipSuccess) ||
        (e = hipMemcpyAsync(lw->lgs_perm.p, n2o.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(static_cast<int*>(lw->lgs_perm.p) + n, o2n.data(), (size_t)n * sizeof(int),
                            hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(static_cast<int*>(lw->lgs_ao.p) + M.atom0, ao_new.data(), (size_t)na * sizeof(int),
                            hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess ||  // the host vectors end with this lambda
        (e = hipEventRecord(lw->ev[0], s)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_lgs_build, dim3(d.nchA + d.nchB, na), dim3(256), 0, s, mi, b, d,
                       static_cast<const LgsAtom*>(lw->lgs_at.p), static_cast<const int*>(lw->lgs_srt.p),
                       static_cast<double*>(lw->lgs_st.p), mixed ? static_cast<float*>(lw->lgs_st32.p) : nullptr);
    hipLaunchKernelGGL(k_lgs_ocpar, dim3(grid_for(na)), dim3(256), 0, s, mi, b, na,
                       static_cast<const LgsAtom*>(lw->lgs_at.p), static_cast<double*>(lw->lgs_ocp.p));
    if ((e = hipGetLastError()) != hipSuccess || (e = hipEventRecord(lw->ev[1], s)) != hipSuccess ||
        (e = hipEventSynchronize(lw->ev[1])) != hipSuccess)
      return e;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, lw->ev[0], lw->ev[1]) == hipSuccess) phase[0] += ms;
    lgs_d = d;
    lgs_n_aa = n_aa;
    lgs_n_xb = (int)wg.size() - n_aa;
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
    auto streamed = [&](auto* st) {
      using T = std::remove_const_t<std::remove_pointer_t<decltype(st)>>;
      const LgsDims& d = lgs_d;
      const auto* at = static_cast<const LgsAtom*>(lw->lgs_at.p);
      const int* srt = static_cast<const int*>(lw->lgs_srt.p);
      const int2* wg = static_cast<const int2*>(lw->lgs_wg.p);
      T* jcol = static_cast<T*>(lw->lgs_jcol.p);
      T* jrowA = static_cast<T*>(lw->lgs_jrow.p);
      T* jrowB = jrowA + (size_t)d.nA * d.nchA * 10;
      T* part = static_cast<T*>(lw->lgs_part.p);
      const double* ocp = static_cast<const double*>(lw->lgs_ocp.p);
      if (lgs_n_aa > 0)
        hipLaunchKernelGGL((k_lgs_fock_aa<T>), dim3(lgs_n_aa), dim3(256), 0, s, mi, b, d, at, srt, wg, st, P, 0.5, F,
                           jcol, jrowA);
      if (lgs_n_xb > 0)
        hipLaunchKernelGGL((k_lgs_fock_xb<T>), dim3(lgs_n_xb), dim3(256), 0, s, mi, b, d, at, srt, wg + lgs_n_aa, st,
                           P, 0.5, F, jcol, jrowB);
      hipLaunchKernelGGL((k_lgs_colsum<T>), dim3(d.nchA + d.nchB, LGS_NPART), dim3(256), 0, s, d, at, jcol, part);
      hipLaunchKernelGGL((k_lgs_diag<T>), dim3(d.nchA + d.nchB, 10), dim3(64), 0, s, mi, b, d, ocp, part, jrowA, jrowB, P,
                         0.5, F);
      return hipSuccess;
    };
    if (lgs_on) {
      const hipError_t e = lowp ? streamed(static_cast<const float*>(lw->lgs_st32.p))
                                : streamed(static_cast<const double*>(lw->lgs_st.p));
      if (e != hipSuccess) return e;
    } else if (lowp) {
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
    // Each iteration's Fock build is enqueued at the end of the previous one, before the host waits
    // for the stop flag, so the device is not idle while the host wakes up: it is discarded when the
    // loop stops (it writes the next DIIS slot only) and redone when the precision switches.
    // Its timing events alternate between two pairs (ev 0/1 and 6/7). SCFs of a lockstep group
    // (eb set) keep the plain order.
    bool pre = false;   // this iteration's build is already enqueued
    int pre_low = 0;    // in that precision
    for (; iter < M.max_scf; ++iter) {
      const int h = iter % HIST;
      const int cnt = std::min(iter + 1, HIST);
      double* P = slot(S_HD + h);
      double* F = slot(S_HF + h);
      double* E = slot(S_HE + h);
      hipEvent_t fs = (iter & 1) ? lw->ev[6] : lw->ev[0], fe = (iter & 1) ? lw->ev[7] : lw->ev[1];
      if (!pre || pre_low != (low != 0)) {
        LG(hipEventRecord(fs, s));
        LG(fock_rhf(P, F, low != 0));
        LG(hipEventRecord(fe, s));
      }
      pre = false;
      LG(hipEventRecord(lw->ev[9], s));
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
      LG(hipMemcpyAsync(lw->flag_host, ctl, sizeof host, hipMemcpyDeviceToHost, s));
      LG(hipEventRecord(lw->ev[8], s));
      if (!eb && iter + 1 < M.max_scf) {  // the next iteration's Fock build, in the current precision
        const int hn = (iter + 1) % HIST;
        hipEvent_t fsn = (iter & 1) ? lw->ev[0] : lw->ev[6], fen = (iter & 1) ? lw->ev[1] : lw->ev[7];
        LG(hipEventRecord(fsn, s));
        LG(fock_rhf(slot(S_HD + hn), slot(S_HF + hn), low != 0));
        LG(hipEventRecord(fen, s));
        pre = true;
        pre_low = low != 0;
      }
      LG(hipEventSynchronize(lw->ev[8]));
      host = *lw->flag_host;
      float ms;
      if (hipEventElapsedTime(&ms, fs, fe) == hipSuccess) phase[0] += ms;
      if (hipEventElapsedTime(&ms, lw->ev[9], lw->ev[2]) == hipSuccess) phase[1] += ms;
      for (int k = 2; k < 5; ++k)
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

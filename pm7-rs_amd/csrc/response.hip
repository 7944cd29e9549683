// SPDX-License-Identifier: GPL-3.0-or-later
//
// The closed-shell response Fock of many densities of one molecule (the CPHF kernel of pm7-rs
// cphf_ov, src/hessian.rs:1074-1081: G(X) = build_fock_x(X) - H_core = J(X) - K(X/2),
// src/fock.rs:665-684) as contractions over the density index, so the integrals are read once per
// group of densities and X and G stream through HBM once. X and G are held density-fastest, element
// (d, i, j) at (i * nao + j) * dp + d (dp: the density count rounded up to 16, so every (i, j) run
// starts on a 128-byte line):
//
//   diagonal blocks   Fd = Wall PP: PP (n_pp x D) holds every density's packed per-atom blocks
//                     (X_mu_nu, 2 X_mu_nu off the diagonal: the Coulomb gather's packed density);
//                     Wall (n_pp x n_pp) holds the two-centre integrals w_ab[e][kl] between the
//                     packed pairs of different atoms and, on its diagonal blocks, the one-centre
//                     J - K/2 coefficients of each atom, so one FP64 MFMA product gives every
//                     diagonal-block element of every density (k_resp_wall, k_resp_pack,
//                     k_resp_gemm);
//   exchange blocks   workgroup (atom pair, 256 densities), one density per lane: the pair's X
//                     block, the pair's integrals as wave-uniform scalars, the block
//                     G[F_mu][S_l] = -xk sum_p sum_q w[(mu p)][(l q)] X[F_p][S_q] as fock_pair_item
//                     sums it, stored with its mirror G[S_l][F_mu] from the same value, so G is
//                     exactly symmetric (k_resp_pairs, which also stores the diagonal blocks from
//                     Fd). Every element is stored as (H + g) - H, as the reference subtracts H from
//                     the full Fock matrix.
//
// Each X element is read once and each G element written once, in coalesced runs over the
// densities. The result of a density depends on that density alone (every sum has a fixed order
// and never runs across densities), so a density gives the same bits alone and in any batch.
#include <hip/hip_runtime.h>

#include "fock.h"
#include "kernels.h"
#include "mfma.h"
#include "pm7_device.h"

namespace pm7 {

namespace {

// One-centre (mu nu | la si) of an element: the sp rules, or its 45 x 45 spd matrix.
__device__ inline double oc_elem(const double* __restrict__ ocd, const ElemDev& e, int mu, int nu, int la, int si) {
  if (e.n_orb == 9) return ocd[(int64_t)e.z * OC_D_WORDS + pack_idx(mu, nu) * 45 + pack_idx(la, si)];
  return oc_int(mu, nu, la, si, e);
}

}  // namespace

// Wall[t][u], t = (atom a, packed pair e), u = (atom b, packed pair kl), one thread each: between
// atoms the two-centre integral of the Coulomb gather (the fock_diag_item term), within an atom
// the coefficient of PP[u] in its one-centre J - xk K: J(mu nu | l s) and, for K, (mu l | nu s) or,
// off the packed diagonal, the mean of (mu l | nu s) and (mu s | nu l) (PP carries 2 X_ls).
__global__ void __launch_bounds__(256) k_resp_wall(int mi, const MolDev* __restrict__ mols,
                                                   const int* __restrict__ atom_z, const int* __restrict__ atom_pp,
                                                   const PairDev* __restrict__ pairs, const ElemDev* __restrict__ elem,
                                                   const double* __restrict__ ocd, const double* __restrict__ w,
                                                   double xk, double* __restrict__ wall) {
  const MolDev M = mols[mi];
  const int n = M.n_pp;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int t = (int)(idx / n), u = (int)(idx - (int64_t)t * n);
  const int la = find_atom(atom_pp, M.atom0, M.n_atoms, t), lb = find_atom(atom_pp, M.atom0, M.n_atoms, u);
  const int a = M.atom0 + la, b = M.atom0 + lb;
  const int e = t - atom_pp[a], kl = u - atom_pp[b];
  const ElemDev ea = elem[atom_z[a]];
  double v;
  if (la == lb) {
    int mu, nu, l, s;
    unpack_e(e, mu, nu);
    unpack_e(kl, l, s);
    const double j = oc_elem(ocd, ea, mu, nu, l, s);
    const double k = l == s ? oc_elem(ocd, ea, mu, l, nu, s)
                            : 0.5 * (oc_elem(ocd, ea, mu, l, nu, s) + oc_elem(ocd, ea, mu, s, nu, l));
    v = j - xk * k;
  } else {
    const int lo = la < lb ? la : lb, hi = la < lb ? lb : la;
    const PairDev P = pairs[M.pair0 + tri_index(lo, hi, M.n_atoms)];
    const int npa = npair_of(ea.n_orb), npb = npair_of(elem[atom_z[b]].n_orb);
    const double* wp = w + P.w_off;
    v = P.f == a ? wp[(int64_t)e * npb + kl] : wp[(int64_t)kl * npa + e];
  }
  wall[idx] = v;
}

// PP[t][d] = X_d[o + mu][o + nu] (x 2 off the packed diagonal), t = (atom, packed pair (mu, nu)).
__global__ void __launch_bounds__(256) k_resp_pack(int mi, const MolDev* __restrict__ mols,
                                                   const int* __restrict__ atom_ao, const int* __restrict__ atom_pp,
                                                   const double* __restrict__ xs, int n_dens, int dp,
                                                   double* __restrict__ pp) {
  const MolDev M = mols[mi];
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M.n_pp * n_dens) return;
  const int t = (int)(idx / n_dens), d = (int)(idx - (int64_t)t * n_dens);
  const int la = find_atom(atom_pp, M.atom0, M.n_atoms, t);
  const int a = M.atom0 + la;
  int mu, nu;
  unpack_e(t - atom_pp[a], mu, nu);
  const int o = atom_ao[a];
  pp[idx] = xs[((int64_t)(o + mu) * M.nao + o + nu) * dp + d] * (mu == nu ? 1.0 : 2.0);
}

// Fd (n x D) = Wall (n x n) PP (n x D) on the matrix cores: each workgroup's eight wavefronts take
// one 16 x 16 output tile over eight slices of k, summed in slice order through LDS. Each element's
// sum runs over k in the same order whatever D is.
constexpr int RESP_KSPLIT = 8;
__global__ void __launch_bounds__(64 * RESP_KSPLIT) k_resp_gemm(int n, int n_dens, const double* __restrict__ wall,
                                                                const double* __restrict__ pp,
                                                                double* __restrict__ fd) {
  __shared__ dbl4 part[RESP_KSPLIT - 1][64];
  const int td = (n_dens + 15) >> 4;
  const int I = blockIdx.x / td, J = blockIdx.x - I * td;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const int kq = ((n + 4 * RESP_KSPLIT - 1) / (4 * RESP_KSPLIT)) * 4;  // a slice of k, a multiple of 4
  const int kb = wv * kq, kl = min(kq, n - kb);
  auto fa = [&](int i, int k) { return (i < n && k < kl) ? wall[(int64_t)i * n + kb + k] : 0.0; };
  auto fb = [&](int k, int j) { return (k < kl && j < n_dens) ? pp[(int64_t)(kb + k) * n_dens + j] : 0.0; };
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  if (kl > 0) acc = tile16_f64(kl, 16 * I, 16 * J, fa, fb);
  if (wv > 0) part[wv - 1][lane] = acc;
  __syncthreads();
  if (wv != 0) return;
#pragma unroll 1
  for (int p = 0; p < RESP_KSPLIT - 1; ++p) {
    const dbl4 v = part[p][lane];
#pragma unroll
    for (int t4 = 0; t4 < 4; ++t4) acc[t4] += v[t4];
  }
#pragma unroll
  for (int t4 = 0; t4 < 4; ++t4) {
    const int i = 16 * I + tile_row_f64(q, t4), j = 16 * J + r;
    if (i < n && j < n_dens) fd[(int64_t)i * n_dens + j] = acc[t4];
  }
}

namespace {

// One exchange block of the pair (F, S) (NF >= NS orbitals) for this lane's density: the pair's X
// block in registers, each element sum_p sum_q w[(mu p)][(l q)] X[F_p][S_q] with the integrals as
// wave-uniform loads, stored as (H - xk g) - H at (F_mu, S_l) and (S_l, F_mu).
template <int NF, int NS>
__device__ __forceinline__ void resp_pair_block(const double* __restrict__ wp, const double* __restrict__ H, int ld, int nao,
                                       int oF, int oS, const double* __restrict__ xs, double* __restrict__ gs,
                                       int dp, int d, bool store, double xk) {
  constexpr int npS = NS * (NS + 1) / 2;
  constexpr int UF = NF * NS <= 16 ? NF : 1, US = NF * NS <= 16 ? NS : 1;  // whole unroll for s/p pairs
  double x[NF * NS];
#pragma unroll
  for (int p = 0; p < NF; ++p)
#pragma unroll
    for (int q = 0; q < NS; ++q) x[p * NS + q] = xs[((int64_t)(oF + p) * nao + oS + q) * dp + d];
#pragma unroll UF
  for (int mu = 0; mu < NF; ++mu)
#pragma unroll US
    for (int l = 0; l < NS; ++l) {
      double g = 0.0;
#pragma unroll
      for (int p = 0; p < NF; ++p) {
        const double* row = wp + pack_idx(mu, p) * npS;
#pragma unroll
        for (int q = 0; q < NS; ++q) g += row[pack_idx(l, q)] * x[p * NS + q];
      }
      const double h = H[(int64_t)(oF + mu) * ld + oS + l];
      const double v = (h - xk * g) - h;
      if (store) {
        gs[((int64_t)(oF + mu) * nao + oS + l) * dp + d] = v;
        gs[((int64_t)(oS + l) * nao + oF + mu) * dp + d] = v;
      }
    }
}

}  // namespace

// Workgroup (item on x, 256 densities on y): items below n_atoms (n_atoms - 1) / 2 are the atom
// pairs in the order of the pair list, the rest the atoms' diagonal blocks. SPD: atoms with up to 9
// orbitals; otherwise every atom has 1 or 4.
template <bool SPD>
__global__ void __launch_bounds__(256) k_resp_pairs(int mi, const MolDev* __restrict__ mols,
                                                    const int* __restrict__ atom_z, const int* __restrict__ atom_ao,
                                                    const int* __restrict__ atom_pp,
                                                    const PairDev* __restrict__ pairs,
                                                    const ElemDev* __restrict__ elem, const double* __restrict__ w,
                                                    const double* __restrict__ slab, const double* __restrict__ xs,
                                                    double* __restrict__ gs, int n_dens, int dp,
                                                    const double* __restrict__ fd, double xk) {
  const MolDev M = mols[mi];
  const int nao = M.nao, ld = M.ld, na = M.n_atoms;
  const int d0 = blockIdx.y * 256, d = d0 + (int)threadIdx.x;
  const bool store = d < n_dens;
  const int dd = store ? d : n_dens - 1;  // a valid density for the loads of the idle lanes
  const double* H = slab + M.mat0 + (int64_t)S_H * ld * ld;
  const int item = blockIdx.x, npair = na * (na - 1) / 2;
  if (item < npair) {
    const PairDev P = pairs[M.pair0 + item];
    const int nF = elem[atom_z[P.f]].n_orb, nS = elem[atom_z[P.s]].n_orb;
    const int oF = atom_ao[P.f], oS = atom_ao[P.s];
    const double* wp = w + P.w_off;
    const int shape = nF * 16 + nS;
    if (shape == 4 * 16 + 4) resp_pair_block<4, 4>(wp, H, ld, nao, oF, oS, xs, gs, dp, dd, store, xk);
    else if (shape == 4 * 16 + 1) resp_pair_block<4, 1>(wp, H, ld, nao, oF, oS, xs, gs, dp, dd, store, xk);
    else if (shape == 1 * 16 + 1) resp_pair_block<1, 1>(wp, H, ld, nao, oF, oS, xs, gs, dp, dd, store, xk);
    else if constexpr (SPD) {
      if (shape == 9 * 16 + 9) resp_pair_block<9, 9>(wp, H, ld, nao, oF, oS, xs, gs, dp, dd, store, xk);
      else if (shape == 9 * 16 + 4) resp_pair_block<9, 4>(wp, H, ld, nao, oF, oS, xs, gs, dp, dd, store, xk);
      else if (shape == 9 * 16 + 1) resp_pair_block<9, 1>(wp, H, ld, nao, oF, oS, xs, gs, dp, dd, store, xk);
    }
    return;
  }
  const int a = M.atom0 + (item - npair);
  const int nfa = elem[atom_z[a]].n_orb, oa = atom_ao[a], pp_a = atom_pp[a];
  for (int mu = 0; mu < nfa; ++mu)
    for (int nu = 0; nu < nfa; ++nu) {
      const double h = H[(int64_t)(oa + mu) * ld + oa + nu];
      const double v = (h + fd[(int64_t)(pp_a + pack_idx(mu, nu)) * n_dens + dd]) - h;
      if (store) gs[((int64_t)(oa + mu) * nao + oa + nu) * dp + d] = v;
    }
}

// Between the caller's layout (density d, row i, column j at (d * nao + i) * nao + j, k = i * nao + j)
// and the density-fastest layout (k * dp + d): 32 x 32 tiles through LDS, coalesced on both sides.
__global__ void __launch_bounds__(256) k_resp_transpose(const double* __restrict__ src, double* __restrict__ dst,
                                                        int64_t nk, int n_dens, int dp, int to_pairs) {
  __shared__ double t[32][33];
  const int64_t k0 = (int64_t)blockIdx.x * 32;
  const int d0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (to_pairs) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int d = d0 + ty + 8 * m;
      const int64_t k = k0 + tx;
      if (d < n_dens && k < nk) t[ty + 8 * m][tx] = src[(int64_t)d * nk + k];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int64_t k = k0 + ty + 8 * m;
      const int d = d0 + tx;
      if (d < n_dens && k < nk) dst[k * dp + d] = t[tx][ty + 8 * m];
    }
  } else {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int64_t k = k0 + ty + 8 * m;
      const int d = d0 + tx;
      if (d < n_dens && k < nk) t[tx][ty + 8 * m] = src[k * dp + d];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int d = d0 + ty + 8 * m;
      const int64_t k = k0 + tx;
      if (d < n_dens && k < nk) dst[(int64_t)d * nk + k] = t[ty + 8 * m][tx];
    }
  }
}

hipError_t launch_response_relayout(const double* src, double* dst, int nao, int n_dens, int dp, bool to_pairs,
                                    hipStream_t s) {
  const int64_t nk = (int64_t)nao * nao;
  if (nk == 0 || n_dens <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_resp_transpose, dim3((unsigned)((nk + 31) / 32), (unsigned)((n_dens + 31) / 32)), dim3(256), 0,
                     s, src, dst, nk, n_dens, dp, to_pairs ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_response_pairs(const BatchDev& b, int mol, int n_pp, int n_atoms, bool spd, const double* xs,
                                 double* gs, int n_dens, int dp, double* wall, double* pp, double* fd, hipStream_t s) {
  if (n_dens <= 0) return hipSuccess;
  const double xk = 0.5;  // build_fock_x: K(X / 2)
  const int64_t nw = (int64_t)n_pp * n_pp, npd = (int64_t)n_pp * n_dens;
  hipLaunchKernelGGL(k_resp_wall, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, mol, b.mols, b.atom_z,
                     b.atom_pp, b.pairs, b.elem, b.ocd, b.w, xk, wall);
  hipLaunchKernelGGL(k_resp_pack, dim3((unsigned)((npd + 255) / 256)), dim3(256), 0, s, mol, b.mols, b.atom_ao,
                     b.atom_pp, xs, n_dens, dp, pp);
  const int64_t tiles = (int64_t)((n_pp + 15) / 16) * ((n_dens + 15) / 16);
  hipLaunchKernelGGL(k_resp_gemm, dim3((unsigned)tiles), dim3(64 * RESP_KSPLIT), 0, s, n_pp, n_dens, wall, pp, fd);
  const int64_t items = (int64_t)n_atoms * (n_atoms - 1) / 2 + n_atoms;
  const dim3 grid((unsigned)items, (unsigned)((n_dens + 255) / 256));
  if (spd)
    hipLaunchKernelGGL(k_resp_pairs<true>, grid, dim3(256), 0, s, mol, b.mols, b.atom_z, b.atom_ao, b.atom_pp,
                       b.pairs, b.elem, b.w, b.slab, xs, gs, n_dens, dp, fd, xk);
  else
    hipLaunchKernelGGL(k_resp_pairs<false>, grid, dim3(256), 0, s, mol, b.mols, b.atom_z, b.atom_ao, b.atom_pp,
                       b.pairs, b.elem, b.w, b.slab, xs, gs, n_dens, dp, fd, xk);
  return hipGetLastError();
}

}  // namespace pm7

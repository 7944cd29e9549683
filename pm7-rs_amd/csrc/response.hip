// SPDX-License-Identifier: GPL-3.0-or-later
//
// The closed-shell response Fock of many densities of one molecule (the CPHF kernel of pm7-rs
// cphf_ov, src/hessian.rs:1074-1081: G(X) = build_fock_x(X) - H_core = J(X) - K(X/2),
// src/fock.rs:665-684) as contractions over the density index, so the integrals are read once per
// group of densities and X and G stream through HBM once, in rows:
//
//   diagonal blocks   Fd = Wall PP: PP (n_pp x D) holds every density's packed per-atom blocks
//                     (X_mu_nu, 2 X_mu_nu off the diagonal: the Coulomb gather's packed density);
//                     Wall (n_pp x n_pp) holds the two-centre integrals w_ab[e][kl] between the
//                     packed pairs of different atoms and, on its diagonal blocks, the one-centre
//                     J - K/2 coefficients of each atom, so one FP64 MFMA product gives every
//                     diagonal-block element of every density (k_resp_wall, k_resp_pack,
//                     k_resp_gemm);
//   exchange blocks   workgroup (atom a, group of DC densities): the rows a_mu of X for the group
//                     in LDS, then every element G[a_mu][j] of those rows: the exchange sum of the
//                     pair (a, atom of j) with the pair's integrals, or the diagonal-block value
//                     from Fd, stored as (H + g) - H as the reference subtracts H from the full
//                     Fock matrix (k_resp_rows). Each workgroup writes whole rows, so the stores
//                     are coalesced and nothing is written twice; the element G[b_l][a_m] of the
//                     transposed block is the same sum in the same order, so G stays exactly
//                     symmetric.
//
// The result of a density depends on that density alone (every sum has a fixed order and never
// runs across densities), so a density gives the same bits alone and in any batch.
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

// Atom (molecule-relative index) whose orbitals hold AO column j: atom_ao ascending.
__device__ inline int atom_of_ao(const int* atom_ao, int a0, int na, int j) {
  int lo = 0, hi = na - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (atom_ao[a0 + mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  return lo;
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
                                                   const double* __restrict__ xs, int n_dens,
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
  pp[idx] = xs[((int64_t)(o + mu) * n_dens + d) * M.ld + o + nu] * (mu == nu ? 1.0 : 2.0);
}

// Fd (n x D) = Wall (n x n) PP (n x D) on the matrix cores: each workgroup's four wavefronts take
// one 16 x 16 output tile over four quarters of k, summed (w0 + w1) + (w2 + w3) through LDS. Each
// element's sum runs over k in the same order whatever D is.
__global__ void __launch_bounds__(256) k_resp_gemm(int n, int n_dens, const double* __restrict__ wall,
                                                   const double* __restrict__ pp, double* __restrict__ fd) {
  __shared__ dbl4 part[3][64];
  const int td = (n_dens + 15) >> 4;
  const int I = blockIdx.x / td, J = blockIdx.x - I * td;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const int kq = ((n + 15) >> 4) << 2;  // a quarter of k, a multiple of 4
  const int kb = wv * kq, kl = min(kq, n - kb);
  auto fa = [&](int i, int k) { return (i < n && k < kl) ? wall[(int64_t)i * n + kb + k] : 0.0; };
  auto fb = [&](int k, int j) { return (k < kl && j < n_dens) ? pp[(int64_t)(kb + k) * n_dens + j] : 0.0; };
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  if (kl > 0) acc = tile16_f64(kl, 16 * I, 16 * J, fa, fb);
  if (wv > 0) part[wv - 1][lane] = acc;
  __syncthreads();
  if (wv != 0) return;
  const dbl4 p1 = part[0][lane], p2 = part[1][lane], p3 = part[2][lane];
#pragma unroll
  for (int t4 = 0; t4 < 4; ++t4) {
    const int i = 16 * I + tile_row_f64(q, t4), j = 16 * J + r;
    if (i < n && j < n_dens) fd[(int64_t)i * n_dens + j] = (acc[t4] + p1[t4]) + (p2[t4] + p3[t4]);
  }
}

// Exchange blocks on the matrix cores. For the pair (a, b) with first atom F and second S the block
// G[a_mu][b_l] over the densities d of a group is one product
//   G[(mu l)][d] = sum_k A[(mu l)][k] B[k][d],  k = (p, q) = p * nS + q over F's orbitals p and S's q,
//   A[(mu l)][k] = w[(iF p)][(iS q)],  B[k][d] = X_d[F_p][S_q],
// (iF, iS) = (mu, l) when a is F, (l, mu) otherwise, as fock_pair_item sums it. The element
// G[b_l][a_mu] of atom b's rows is the same product over the same k with the same A row, so the
// two stay equal. Rows (mu l) = mu * nfb + l; a density per MFMA column (up to 16).
//
// Workgroup (densities d0 .. d0 + dc on x, atom a on y; only atoms with one orbital when `light`,
// only the others otherwise): the rows a_mu of the group in LDS, [mu][d][ld]; wavefront v takes the
// partner atoms v, v + 4, ... and overwrites each partner's columns of those rows with its G block
// (the columns of different partners are disjoint, and a block's X is read before its G is
// written); then the workgroup writes whole rows of G, the diagonal block from Fd, each stored as
// (H + g) - H as the reference subtracts H from the full Fock matrix. Each G element is written
// once, in coalesced rows; X and G stream past the caches (non-temporal), which keep the integrals,
// H and Fd.

namespace {

typedef double nd2 __attribute__((ext_vector_type(2)));

// Lookup tables of a row workgroup: per partner atom its integral block's offset and
// AO offset | orbitals << 16 | a-is-first << 24.
__device__ inline void resp_tables(const MolDev& M, int la, const int* __restrict__ atom_z,
                                   const int* __restrict__ atom_ao, const PairDev* __restrict__ pairs,
                                   const ElemDev* __restrict__ elem, int64_t* s_woff, int* s_info) {
  const int na = M.n_atoms, a = M.atom0 + la;
  for (int lb = threadIdx.x; lb < na; lb += 256) {
    const int b = M.atom0 + lb;
    int64_t woff = 0;
    int af = 0;
    if (lb != la) {
      const PairDev P = pairs[M.pair0 + tri_index(min(la, lb), max(la, lb), na)];
      woff = P.w_off;
      af = P.f == a;
    }
    s_woff[lb] = woff;
    s_info[lb] = atom_ao[b] | (elem[atom_z[b]].n_orb << 16) | (af << 24);
  }
}

// Rows a_0 .. a_{nf-1} of the group: dc * ld contiguous words per orbital (ld even: 16-byte pairs),
// to LDS rows [mu][d][ld] of stride dstr words per orbital.
template <int U>
__device__ inline void resp_load_rows(const double* __restrict__ xs, double* sx, int oa, int nf, int n_dens,
                                      int d0, int dc, int ld, int64_t dstr) {
  const int half = dc * ld / 2;
  const nd2* src = reinterpret_cast<const nd2*>(xs);
  nd2* dst = reinterpret_cast<nd2*>(sx);
  for (int mu = 0; mu < nf; ++mu) {
    const nd2* s = src + ((int64_t)(oa + mu) * n_dens + d0) * ld / 2;
    nd2* t = dst + mu * dstr / 2;
    for (int i0 = threadIdx.x; i0 < half; i0 += 256 * U) {
      nd2 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + 256 * u < half) v[u] = __builtin_nontemporal_load(s + i0 + 256 * u);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + 256 * u < half) t[i0 + 256 * u] = v[u];
    }
  }
}

// One pair shape (NFA orbitals on this workgroup's atom, NFB on the partner, AF: this atom comes
// first): lane (r, q) of a wavefront holds, for each k-step s, the index of A[r][4 s + q] in the
// pair's integral block and the LDS offset of B[4 s + q][r] from the lane's base (sx + r * ld +
// partner AO offset), -1 where the operand is zero; they depend on the lane and the shape only, so
// they are formed once per workgroup.
template <int NFA, int NFB, bool AF>
struct RespShape {
  static constexpr int nS = AF ? NFB : NFA, nel = NFA * NFB, nps = nS * (nS + 1) / 2, S4 = (nel + 3) / 4;
  int aix[S4], bof[S4];
  __device__ void init(int r, int q, int dstr, bool dcol) {
    const int mu = r / NFB, l = r % NFB;
    const int iF = AF ? mu : l, iS = AF ? l : mu;
#pragma unroll
    for (int s4 = 0; s4 < S4; ++s4) {
      const int k = 4 * s4 + q, p = k / nS, qq = k % nS;
      aix[s4] = (r < nel && k < nel) ? pack_idx(iF, p) * nps + pack_idx(iS, qq) : -1;
      bof[s4] = (dcol && k < nel) ? (AF ? p * dstr + qq : qq * dstr + p) : -1;
    }
  }
  // Every lane loads and writes (no divergent branches): operands outside the block are read at a
  // valid address and replaced by zero, results outside it go to the lane's trash word.
  // (The loaded values are zeroed where they are used, so a fetch ahead does not wait for them.)
  // NL loads whatever the shape, so the count of loads in flight behind a fetch is the same on
  // every path and a later wait for this fetch's data leaves the loads after it in flight.
  // Index of the k-step s operand, 0 (a valid word, zeroed at use) outside the block.
  __device__ int load_index(int s4) const {
    int v = 0;
#pragma unroll
    for (int t = 0; t < S4; ++t)
      if (t == s4) v = aix[t] > 0 ? aix[t] : 0;
    return v;
  }
  // The block's product from the LDS rows, written over the same words (offsets `out`).
  __device__ void block(double* base, const double* any, double* trash, const double (&av)[4],
                        const int (&out)[4]) const {
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s4 = 0; s4 < S4; ++s4) {
      const int o = bof[s4];
      const double v = *(o < 0 ? any : base + o);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(aix[s4] < 0 ? 0.0 : av[s4], o < 0 ? 0.0 : v, acc, 0, 0, 0);
    }
#pragma unroll
    for (int t4 = 0; t4 < 4; ++t4) *(out[t4] < 0 ? trash : base + out[t4]) = acc[t4];
  }
};

// LDS offsets (from the lane's base) of the result rows q + 4 t of a block with NFB partner
// orbitals, -1 outside the block or the group.
template <int NFA, int NFB>
__device__ inline void resp_out_offsets(int q, int dstr, bool dcol, int (&out)[4]) {
#pragma unroll
  for (int t4 = 0; t4 < 4; ++t4) {
    const int i = tile_row_f64(q, t4);
    out[t4] = (i < NFA * NFB && dcol) ? (i / NFB) * dstr + i % NFB : -1;
  }
}

}  // namespace

// s/p molecules: every atom has 1 or 4 orbitals, this workgroup's atom NFA. The shape of each
// partner (its orbitals, which atom comes first) selects a specialised block, and the next
// partner's integrals are fetched while the current one is multiplied.
template <int NFA>
__global__ void __launch_bounds__(256) k_resp_rows_sp(int mi, const MolDev* __restrict__ mols,
                                                      const int* __restrict__ atom_z, const int* __restrict__ atom_ao,
                                                      const int* __restrict__ atom_pp,
                                                      const PairDev* __restrict__ pairs,
                                                      const ElemDev* __restrict__ elem, const double* __restrict__ w,
                                                      const double* __restrict__ slab, const double* __restrict__ xs,
                                                      double* __restrict__ gs, int n_dens, int dc_max, int x_words,
                                                      const double* __restrict__ fd, double xk) {
  extern __shared__ double sx[];
  const MolDev M = mols[mi];
  const int nao = M.nao, ld = M.ld, na = M.n_atoms;
  const int la = blockIdx.y, a = M.atom0 + la;
  if (elem[atom_z[a]].n_orb != NFA) return;  // the whole workgroup
  const int d0 = blockIdx.x * dc_max;
  const int dc = min(dc_max, n_dens - d0);
  const int oa = atom_ao[a];
  const int64_t dstr = (int64_t)dc_max * ld;
  int64_t* s_woff = reinterpret_cast<int64_t*>(sx + x_words);
  int* s_info = reinterpret_cast<int*>(s_woff + na);
  resp_load_rows<NFA == 1 ? 8 : 4>(xs, sx, oa, NFA, n_dens, d0, dc, ld, dstr);
  resp_tables(M, la, atom_z, atom_ao, pairs, elem, s_woff, s_info);
  __syncthreads();
  // The partner loop runs per wavefront: its index and every partner's data are wave-uniform
  // (scalar registers), so only the operands differ between lanes.
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const bool dcol = r < dc;
  double* trash = sx + x_words + na + (na + 1) / 2 + threadIdx.x;  // after the tables
  RespShape<NFA, 1, false> sh0;
  RespShape<NFA, 1, true> sh1;
  RespShape<NFA, 4, false> sh2;
  RespShape<NFA, 4, true> sh3;
  sh0.init(r, q, (int)dstr, dcol);
  sh1.init(r, q, (int)dstr, dcol);
  sh2.init(r, q, (int)dstr, dcol);
  sh3.init(r, q, (int)dstr, dcol);
  int out1[4], out4[4];
  resp_out_offsets<NFA, 1>(q, (int)dstr, dcol, out1);
  resp_out_offsets<NFA, 4>(q, (int)dstr, dcol, out4);
  struct Part {
    int ob, shape;  // shape: (orbitals == 4) << 1 | a-is-first
    double av[4];
  };
  constexpr int NL = NFA == 1 ? 1 : 4;  // k-steps of this atom's largest block
  // Every fetch issues the same NL loads on one straight path (the shape picks the indices by
  // selects), so a wait for one ring slot leaves the later slots' loads in flight; a fetch past
  // the last partner reads a valid partner whose values are never used.
  auto fetch = [&](int lb, Part& pt) {
    const int info = __builtin_amdgcn_readfirstlane(s_info[lb]);
    const int64_t wo = s_woff[lb];
    const int64_t woff = (int64_t)((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(wo >> 32)) << 32 |
                                   (uint32_t)__builtin_amdgcn_readfirstlane((int)wo));
    pt.ob = info & 0xffff;
    pt.shape = (((info >> 16) & 0xff) == 4 ? 2 : 0) | ((info >> 24) & 1);
    const double* wp = w + woff;
#pragma unroll
    for (int s4 = 0; s4 < NL; ++s4) {
      const int i = pt.shape == 0 ? sh0.load_index(s4)
                    : pt.shape == 1 ? sh1.load_index(s4)
                    : pt.shape == 2 ? sh2.load_index(s4)
                                    : sh3.load_index(s4);
      pt.av[s4] = wp[i];
    }
  };
  // Integrals fetched four partners ahead (an L2 round trip outlasts a block's products) into a
  // ring of four slots, each consumed and refilled in place.
  Part ring[4] = {};
#pragma unroll
  for (int u = 0; u < 4; ++u) fetch(min(wv + 4 * u, na - 1), ring[u]);
  for (int lb0 = wv; lb0 < na; lb0 += 16) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int lb = lb0 + 4 * u;
      if (lb >= na) break;
      Part& pt = ring[u];
      if (lb != la) {
        double* base = sx + r * ld + pt.ob;
        switch (pt.shape) {
          case 0: sh0.block(base, sx, trash, pt.av, out1); break;
          case 1: sh1.block(base, sx, trash, pt.av, out1); break;
          case 2: sh2.block(base, sx, trash, pt.av, out4); break;
          default: sh3.block(base, sx, trash, pt.av, out4); break;
        }
      }
      fetch(min(lb + 16, na - 1), pt);
    }
  }
  __syncthreads();
  // Whole rows of G: element (mu, j) per thread for every density of the group, H read first.
  const double* H = slab + M.mat0 + (int64_t)S_H * ld * ld + (int64_t)oa * ld;
  const int pp_a = atom_pp[a];
  constexpr int JP = 2;
  for (int j0 = 0; j0 < nao; j0 += 256 * JP) {
    double hv[NFA][JP];
#pragma unroll
    for (int mu = 0; mu < NFA; ++mu)
#pragma unroll
      for (int jj = 0; jj < JP; ++jj) {
        const int j = j0 + threadIdx.x + 256 * jj;
        hv[mu][jj] = j < nao ? H[mu * ld + j] : 0.0;
      }
#pragma unroll
    for (int mu = 0; mu < NFA; ++mu)
#pragma unroll
      for (int jj = 0; jj < JP; ++jj) {
        const int j = j0 + threadIdx.x + 256 * jj;
        if (j >= nao) continue;
        const double h = hv[mu][jj];
        double* g = gs + ((int64_t)(oa + mu) * n_dens + d0) * ld + j;
        if (j >= oa && j < oa + NFA) {
          const double* f = fd + (int64_t)(pp_a + pack_idx(mu, j - oa)) * n_dens + d0;
          for (int d = 0; d < dc; ++d) __builtin_nontemporal_store((h + f[d]) - h, g + d * ld);
        } else {
          const double* x = sx + mu * dstr + j;
          for (int d = 0; d < dc; ++d) __builtin_nontemporal_store((h - xk * x[d * ld]) - h, g + d * ld);
        }
      }
  }
}

// Molecules with d shells: up to 9 orbitals per atom (6 row tiles, K <= 81), shapes at run time.
__global__ void __launch_bounds__(256) k_resp_rows_spd(int mi, const MolDev* __restrict__ mols,
                                                       const int* __restrict__ atom_z, const int* __restrict__ atom_ao,
                                                       const int* __restrict__ atom_pp,
                                                       const PairDev* __restrict__ pairs,
                                                       const ElemDev* __restrict__ elem, const double* __restrict__ w,
                                                       const double* __restrict__ slab, const double* __restrict__ xs,
                                                       double* __restrict__ gs, int n_dens, int dc_max, int x_words,
                                                       int light, const double* __restrict__ fd, double xk) {
  extern __shared__ double sx[];
  const MolDev M = mols[mi];
  const int nao = M.nao, ld = M.ld, na = M.n_atoms;
  const int la = blockIdx.y, a = M.atom0 + la;
  const int nfa = elem[atom_z[a]].n_orb;
  if ((nfa == 1) != (light != 0)) return;  // the whole workgroup
  const int d0 = blockIdx.x * dc_max;
  const int dc = min(dc_max, n_dens - d0);
  const int oa = atom_ao[a];
  const int64_t dstr = (int64_t)dc_max * ld;
  int64_t* s_woff = reinterpret_cast<int64_t*>(sx + x_words);
  int* s_info = reinterpret_cast<int*>(s_woff + na);
  resp_load_rows<4>(xs, sx, oa, nfa, n_dens, d0, dc, ld, dstr);
  resp_tables(M, la, atom_z, atom_ao, pairs, elem, s_woff, s_info);
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const bool dcol = r < dc;
  for (int lb = __builtin_amdgcn_readfirstlane(wv); lb < na; lb += 4) {
    if (lb == la) continue;
    const int info = __builtin_amdgcn_readfirstlane(s_info[lb]);
    const int ob = info & 0xffff, nfb = (info >> 16) & 0xff;
    const bool af = (info >> 24) != 0;
    const int nS = af ? nfb : nfa, nel = nfa * nfb, nps = npair_of(nS);
    const int64_t wo = s_woff[lb];
    const double* wp = w + (int64_t)((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(wo >> 32)) << 32 |
                                     (uint32_t)__builtin_amdgcn_readfirstlane((int)wo));
    double* base = sx + r * ld + ob;
    dbl4 acc[6];
#pragma unroll
    for (int rt = 0; rt < 6; ++rt) acc[rt] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < nel; k0 += 4) {
      const int k = k0 + q, p = k / nS, qq = k - p * nS;
      const double bv = (dcol && k < nel) ? base[af ? p * dstr + qq : qq * dstr + p] : 0.0;
#pragma unroll
      for (int rt = 0; rt < 6; ++rt) {
        if (16 * rt >= nel) break;
        const int i = 16 * rt + r, mu = i / nfb, l = i - mu * nfb;
        const int iF = af ? mu : l, iS = af ? l : mu;
        const double av = (i < nel && k < nel) ? wp[(int64_t)pack_idx(iF, p) * nps + pack_idx(iS, qq)] : 0.0;
        acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[rt], 0, 0, 0);
      }
    }
#pragma unroll
    for (int rt = 0; rt < 6; ++rt) {
      if (16 * rt >= nel) break;
#pragma unroll
      for (int t4 = 0; t4 < 4; ++t4) {
        const int i = 16 * rt + tile_row_f64(q, t4);
        if (i < nel && dcol) {
          const int mu = i / nfb, l = i - mu * nfb;
          base[mu * dstr + l] = acc[rt][t4];
        }
      }
    }
  }
  __syncthreads();
  const double* H = slab + M.mat0 + (int64_t)S_H * ld * ld;
  const int pp_a = atom_pp[a];
  for (int mu = 0; mu < nfa; ++mu)
    for (int j = threadIdx.x; j < nao; j += 256) {
      const double h = H[(int64_t)(oa + mu) * ld + j];
      double* g = gs + ((int64_t)(oa + mu) * n_dens + d0) * ld + j;
      if (j >= oa && j < oa + nfa) {
        const double* f = fd + (int64_t)(pp_a + pack_idx(mu, j - oa)) * n_dens + d0;
        for (int d = 0; d < dc; ++d) __builtin_nontemporal_store((h + f[d]) - h, g + d * ld);
      } else {
        const double* x = sx + mu * dstr + j;
        for (int d = 0; d < dc; ++d) __builtin_nontemporal_store((h - xk * x[d * ld]) - h, g + d * ld);
      }
    }
}

// Between the caller's layout (density d, row i, column j at (d * nao + i) * nao + j) and the row
// layout of the contraction kernels ((i * n_dens + d) * ld + j), so that the rows of an atom for a
// group of densities are one contiguous run: to_rows 1 takes src in the caller's layout.
__global__ void __launch_bounds__(256) k_resp_relayout(const double* __restrict__ src, double* __restrict__ dst,
                                                       int nao, int ld, int n_dens, int to_rows) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)nao * n_dens * nao) return;
  const int64_t row = idx / nao;  // (i, d) in the row layout
  const int j = (int)(idx - row * nao);
  const int i = (int)(row / n_dens), d = (int)(row - (int64_t)i * n_dens);
  const int64_t c = ((int64_t)d * nao + i) * nao + j, r = row * ld + j;
  if (to_rows) dst[r] = src[c];
  else dst[c] = src[r];
}

hipError_t launch_response_relayout(const double* src, double* dst, int nao, int ld, int n_dens, bool to_rows,
                                    hipStream_t s) {
  const int64_t n = (int64_t)nao * n_dens * nao;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_resp_relayout, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, nao, ld, n_dens,
                     to_rows ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_response_rows(const BatchDev& b, int mol, int n_pp, int n_atoms, int nao, int ld, int nfa_max,
                                bool sp, const double* xs, double* gs, int n_dens, double* wall, double* pp,
                                double* fd, hipStream_t s) {
  if (n_dens <= 0) return hipSuccess;
  const double xk = 0.5;  // build_fock_x: K(X / 2)
  const int64_t nw = (int64_t)n_pp * n_pp, npd = (int64_t)n_pp * n_dens;
  hipLaunchKernelGGL(k_resp_wall, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, mol, b.mols, b.atom_z,
                     b.atom_pp, b.pairs, b.elem, b.ocd, b.w, xk, wall);
  hipLaunchKernelGGL(k_resp_pack, dim3((unsigned)((npd + 255) / 256)), dim3(256), 0, s, mol, b.mols, b.atom_ao,
                     b.atom_pp, xs, n_dens, pp);
  const int64_t tiles = (int64_t)((n_pp + 15) / 16) * ((n_dens + 15) / 16);
  hipLaunchKernelGGL(k_resp_gemm, dim3((unsigned)tiles), dim3(256), 0, s, n_pp, n_dens, wall, pp, fd);
  // Atoms with several orbitals, then those with one (a smaller LDS slab, more densities per group).
  for (int light = 0; light < 2; ++light) {
    const int nf = light ? 1 : nfa_max;
    if (!light && nfa_max == 1) continue;
    const int dc = response_group(nf, ld, n_atoms, light ? 16 : RESP_DC);
    const int x_words = dc * nf * ld;
    const size_t lds = (size_t)(x_words + response_table_words(n_atoms) + RESP_TRASH) * sizeof(double);
    const dim3 grid((unsigned)((n_dens + dc - 1) / dc), (unsigned)n_atoms);
    if (sp && light)
      hipLaunchKernelGGL(k_resp_rows_sp<1>, grid, dim3(256), lds, s, mol, b.mols, b.atom_z, b.atom_ao, b.atom_pp,
                         b.pairs, b.elem, b.w, b.slab, xs, gs, n_dens, dc, x_words, fd, xk);
    else if (sp)
      hipLaunchKernelGGL(k_resp_rows_sp<4>, grid, dim3(256), lds, s, mol, b.mols, b.atom_z, b.atom_ao, b.atom_pp,
                         b.pairs, b.elem, b.w, b.slab, xs, gs, n_dens, dc, x_words, fd, xk);
    else
      hipLaunchKernelGGL(k_resp_rows_spd, grid, dim3(256), lds, s, mol, b.mols, b.atom_z, b.atom_ao, b.atom_pp,
                         b.pairs, b.elem, b.w, b.slab, xs, gs, n_dens, dc, x_words, light, fd, xk);
  }
  return hipGetLastError();
}

}  // namespace pm7

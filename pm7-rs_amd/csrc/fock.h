// SPDX-License-Identifier: GPL-3.0-or-later
//
// F = H + G(P) for sp atoms (pm7-rs src/fock.rs:162-338), as independent work items so one
// workgroup (fock_build, the SCF megakernel) or a whole grid (the large-system path) can run it:
//   pack  item t: the packed per-atom density pp[t] = P[la][si] (la == si) or 2 P[la][si]
//   diag  item t: one (atom, orbital pair) element of a diagonal block: the two-centre Coulomb
//                 gather over every partner atom plus the one-centre J - K
//   pair  item p: the exchange block of atom pair p (one owner, no atomics)
// T is the type G(P) is formed in (FP64, or FP32 in the reduced-precision phase); H is added in
// FP64 either way.
#pragma once

#include <hip/hip_runtime.h>

#include "pm7_device.h"

namespace pm7 {

// Packed pair e = mu(mu+1)/2 + nu (mu >= nu) back to its orbitals, e < 45.
__device__ inline void unpack_e(int e, int& mu, int& nu) {
  mu = (e >= 36) ? 8 : (e >= 28) ? 7 : (e >= 21) ? 6 : (e >= 15) ? 5 : (e >= 10) ? 4 : (e >= 6) ? 3 : (e >= 3) ? 2
                                                                                        : (e >= 1) ? 1 : 0;
  nu = e - mu * (mu + 1) / 2;
}

// One-centre two-electron integral (mu nu | la si) of an sp atom (pm7-rs src/fock.rs:69-107).
__device__ inline double oc_int(int a, int b, int c, int d, const ElemDev& e) {
  if (a == b && c == d) {
    if (a == 0 && c == 0) return e.g_ss;
    if (a == 0 || c == 0) return e.g_sp;
    return a == c ? e.g_pp : e.g_p2;
  }
  int ba = min(a, b), bb = max(a, b), kc = min(c, d), kd = max(c, d);
  if (ba == 0 && bb != 0 && kc == 0 && kd != 0 && bb == kd) return e.h_sp;
  if (ba != 0 && bb != 0 && ba != bb && ba == kc && bb == kd) return 0.5 * (e.g_pp - e.g_p2);
  return 0.0;
}

struct MolCtx {
  MolDev M;
  const int* atom_z;
  const int* atom_ao;
  const int* atom_pp;
  const PairDev* pairs;
  const ElemDev* elem;
  const double* w;
  double* base;  // molecule slab
  double* pp;    // packed densities
  int ld;
  const double* ocd;  // one-centre spd matrices by atomic number (pm7-rs src/fock.rs:171-189)
  bool minus_h = false;  // store F - H (the response kernel G(X)) instead of F
  __device__ double* slot(int s) const { return base + (int64_t)s * ld * ld; }
};

// Stores one Fock element and its mirror; with c.minus_h, F - H, rounded as the two-step
// (H + g) - H is, so G(X) equals build_fock_x(X) - H formed from the full matrix.
__device__ inline void put_f(const MolCtx& c, double* __restrict__ F, int64_t i1, int64_t i2, double f, double h) {
  const double v = c.minus_h ? f - h : f;
  F[i1] = v;
  F[i2] = v;
}

// One-centre (mu nu | la si) of an atom: the sp rules, or the element's 45 x 45 spd matrix.
__device__ inline double oc_val(const MolCtx& c, const ElemDev& e, int mu, int nu, int la, int si) {
  if (e.n_orb == 9) return c.ocd[(int64_t)e.z * OC_D_WORDS + pack_idx(mu, nu) * 45 + pack_idx(la, si)];
  return oc_int(mu, nu, la, si, e);
}


__device__ inline int find_atom(const int* atom_pp, int a0, int na, int t) {
  int lo = 0, hi = na - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (atom_pp[a0 + mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T>
__device__ inline void fock_pack_item(const MolCtx& c, int t, const double* __restrict__ Pt) {
  const MolDev& M = c.M;
  int la = find_atom(c.atom_pp, M.atom0, M.n_atoms, t);
  int a = M.atom0 + la;
  int e = t - c.atom_pp[a];
  int mu, nu;
  unpack_e(e, mu, nu);
  int o = c.atom_ao[a];
  c.pp[t] = (double)(T)(Pt[(int64_t)(o + mu) * c.ld + o + nu] * (mu == nu ? 1.0 : 2.0));
}

template <typename T>
__device__ inline void fock_diag_item(const MolCtx& c, int t, const double* __restrict__ Pt,
                                      const double* __restrict__ Psa, double xa, const double* __restrict__ Psb,
                                      double xb, double* __restrict__ Fa, double* __restrict__ Fb) {
  const MolDev& M = c.M;
  const int ld = c.ld;
  const double* H = c.slot(S_H);
  int la = find_atom(c.atom_pp, M.atom0, M.n_atoms, t);
  int a = M.atom0 + la;
  int e = t - c.atom_pp[a];
  int mu, nu;
  unpack_e(e, mu, nu);
  const ElemDev ea = c.elem[c.atom_z[a]];
  const int npa = npair_of(ea.n_orb);
  T acc = (T)0;
  for (int lb = 0; lb < M.n_atoms; ++lb) {
    if (lb == la) continue;
    int u = la < lb ? la : lb, v = la < lb ? lb : la;
    const PairDev P = c.pairs[M.pair0 + tri_index(u, v, M.n_atoms)];
    int b = M.atom0 + lb;
    int ppb = c.atom_pp[b];
    int npb = (lb + 1 < M.n_atoms ? c.atom_pp[b + 1] : M.n_pp) - ppb;
    const double* wp = c.w + P.w_off;
    const double* pb = c.pp + ppb;
    if (P.f == a) {
      const double* row = wp + (int64_t)e * npb;
      for (int kl = 0; kl < npb; ++kl) acc += (T)row[kl] * (T)pb[kl];
    } else {
      for (int kl = 0; kl < npb; ++kl) acc += (T)wp[(int64_t)kl * npa + e] * (T)pb[kl];
    }
  }
  const int o = c.atom_ao[a];
  const int n = ea.n_orb;
  T ja = (T)0, ka = (T)0, kb = (T)0;
  for (int l = 0; l < n; ++l)
    for (int s = 0; s < n; ++s) {
      int64_t ix = (int64_t)(o + l) * ld + o + s;
      ja += (T)Pt[ix] * (T)oc_val(c, ea, mu, nu, l, s);
      const T gx = (T)oc_val(c, ea, mu, l, nu, s);
      ka += (T)Psa[ix] * gx;
      if (Fb) kb += (T)Psb[ix] * gx;
    }
  const int64_t i1 = (int64_t)(o + mu) * ld + o + nu, i2 = (int64_t)(o + nu) * ld + o + mu;
  double fa = H[i1] + (double)(acc + ja - (T)xa * ka);
  put_f(c, Fa, i1, i2, fa, H[i1]);
  if (Fb) {
    double fb = H[i1] + (double)(acc + ja - (T)xb * kb);
    put_f(c, Fb, i1, i2, fb, H[i1]);
  }
}

// The diagonal-block element t with the partner sum split over the 64 lanes of one wavefront
// (partner lb handled by lane lb % 64, then a fixed shuffle tree): for molecules with many atoms,
// where one thread per element would walk thousands of partners in sequence. All lanes of the
// wave call it with the same t; lane 0 writes.
template <typename T>
__device__ inline void fock_diag_item_wave(const MolCtx& c, int t, const double* __restrict__ Pt,
                                           const double* __restrict__ Psa, double xa,
                                           const double* __restrict__ Psb, double xb, double* __restrict__ Fa,
                                           double* __restrict__ Fb) {
  const MolDev& M = c.M;
  const int ld = c.ld;
  const int lane = threadIdx.x & 63;
  int la = find_atom(c.atom_pp, M.atom0, M.n_atoms, t);
  int a = M.atom0 + la;
  int e = t - c.atom_pp[a];
  int mu, nu;
  unpack_e(e, mu, nu);
  const ElemDev ea = c.elem[c.atom_z[a]];
  const int npa = npair_of(ea.n_orb);
  T acc = (T)0;
  for (int lb = lane; lb < M.n_atoms; lb += 64) {
    if (lb == la) continue;
    int u = la < lb ? la : lb, v = la < lb ? lb : la;
    const PairDev P = c.pairs[M.pair0 + tri_index(u, v, M.n_atoms)];
    int b = M.atom0 + lb;
    int ppb = c.atom_pp[b];
    int npb = (lb + 1 < M.n_atoms ? c.atom_pp[b + 1] : M.n_pp) - ppb;
    const double* wp = c.w + P.w_off;
    const double* pb = c.pp + ppb;
    if (P.f == a) {
      const double* row = wp + (int64_t)e * npb;
      for (int kl = 0; kl < npb; ++kl) acc += (T)row[kl] * (T)pb[kl];
    } else {
      for (int kl = 0; kl < npb; ++kl) acc += (T)wp[(int64_t)kl * npa + e] * (T)pb[kl];
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane != 0) return;
  const double* H = c.slot(S_H);
  const int o = c.atom_ao[a];
  const int n = ea.n_orb;
  T ja = (T)0, ka = (T)0, kb = (T)0;
  for (int l = 0; l < n; ++l)
    for (int s = 0; s < n; ++s) {
      int64_t ix = (int64_t)(o + l) * ld + o + s;
      ja += (T)Pt[ix] * (T)oc_val(c, ea, mu, nu, l, s);
      const T gx = (T)oc_val(c, ea, mu, l, nu, s);
      ka += (T)Psa[ix] * gx;
      if (Fb) kb += (T)Psb[ix] * gx;
    }
  const int64_t i1 = (int64_t)(o + mu) * ld + o + nu, i2 = (int64_t)(o + nu) * ld + o + mu;
  double fa = H[i1] + (double)(acc + ja - (T)xa * ka);
  put_f(c, Fa, i1, i2, fa, H[i1]);
  if (Fb) {
    double fb = H[i1] + (double)(acc + ja - (T)xb * kb);
    put_f(c, Fb, i1, i2, fb, H[i1]);
  }
}

// Exchange block of a pair with a 9-orbital atom: runtime loops, densities read from the slab.
template <typename T>
__device__ void fock_pair_item_spd(const MolCtx& c, const PairDev& P, int nf, int ns, const double* __restrict__ Psa,
                                   double xa, const double* __restrict__ Psb, double xb, double* __restrict__ Fa,
                                   double* __restrict__ Fb) {
  const int ld = c.ld, nps = npair_of(ns);
  const double* H = c.slot(S_H);
  const int of = c.atom_ao[P.f], os = c.atom_ao[P.s];
  const double* wp = c.w + P.w_off;
  for (int mu = 0; mu < nf; ++mu)
    for (int lam = 0; lam < ns; ++lam) {
      T acca = (T)0, accb = (T)0;
      for (int nu = 0; nu < nf; ++nu) {
        const double* row = wp + (int64_t)pack_idx(mu, nu) * nps;
        const double* pa = Psa + (int64_t)(of + nu) * ld + os;
        const double* pb = Fb ? Psb + (int64_t)(of + nu) * ld + os : nullptr;
        for (int sg = 0; sg < ns; ++sg) {
          const T wv = (T)row[pack_idx(lam, sg)];
          acca += (T)pa[sg] * wv;
          if (Fb) accb += (T)pb[sg] * wv;
        }
      }
      const int64_t i1 = (int64_t)(of + mu) * ld + os + lam, i2 = (int64_t)(os + lam) * ld + of + mu;
      const double fa = H[i1] - (double)((T)xa * acca);
      put_f(c, Fa, i1, i2, fa, H[i1]);
      if (Fb) {
        const double fb = H[i1] - (double)((T)xb * accb);
        put_f(c, Fb, i1, i2, fb, H[i1]);
      }
    }
}

template <typename T>
__device__ inline void fock_pair_item(const MolCtx& c, int p, const double* __restrict__ Psa, double xa,
                                      const double* __restrict__ Psb, double xb, double* __restrict__ Fa,
                                      double* __restrict__ Fb) {
  const MolDev& M = c.M;
  const int ld = c.ld;
  const double* H = c.slot(S_H);
  const PairDev P = c.pairs[M.pair0 + p];
  const ElemDev ef = c.elem[c.atom_z[P.f]];
  const ElemDev es = c.elem[c.atom_z[P.s]];
  const int nf = ef.n_orb, ns = es.n_orb, nps = npair_of(ns);
  if (nf == 9) {  // the first atom carries at least as many orbitals as the second
    fock_pair_item_spd<T>(c, P, nf, ns, Psa, xa, Psb, xb, Fa, Fb);
    return;
  }
  const int of = c.atom_ao[P.f], os = c.atom_ao[P.s];
  const double* wp = c.w + P.w_off;
  // Same-spin density block held in registers: every index below is a compile-time constant.
  T Xa[4][4], Xb[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool in = i < nf && j < ns;
      Xa[i][j] = in ? (T)Psa[(int64_t)(of + i) * ld + os + j] : (T)0;
      Xb[i][j] = (in && Fb) ? (T)Psb[(int64_t)(of + i) * ld + os + j] : (T)0;
    }
#pragma unroll
  for (int mu = 0; mu < 4; ++mu)
#pragma unroll
    for (int lam = 0; lam < 4; ++lam) {
      if (mu >= nf || lam >= ns) continue;
      T acca = (T)0, accb = (T)0;
#pragma unroll
      for (int nu = 0; nu < 4; ++nu) {
        if (nu >= nf) continue;
        const double* row = wp + (int64_t)pack_idx(mu, nu) * nps;
#pragma unroll
        for (int sg = 0; sg < 4; ++sg) {
          if (sg >= ns) continue;
          const T wv = (T)row[pack_idx(lam, sg)];
          acca += Xa[nu][sg] * wv;
          accb += Xb[nu][sg] * wv;
        }
      }
      const int64_t i1 = (int64_t)(of + mu) * ld + os + lam, i2 = (int64_t)(os + lam) * ld + of + mu;
      const double fa = H[i1] - (double)((T)xa * acca);
      put_f(c, Fa, i1, i2, fa, H[i1]);
      if (Fb) {
        const double fb = H[i1] - (double)((T)xb * accb);
        put_f(c, Fb, i1, i2, fb, H[i1]);
      }
    }
}

// F^a = H + J(Pt) - xa K(Psa) (and F^b with Psb when given), by one workgroup of NT threads.
template <int NT, typename T>
__device__ void fock_build(const MolCtx& c, const double* __restrict__ Pt, const double* __restrict__ Psa, double xa,
                           const double* __restrict__ Psb, double xb, double* __restrict__ Fa, double* __restrict__ Fb) {
  for (int t = threadIdx.x; t < c.M.n_pp; t += NT) fock_pack_item<T>(c, t, Pt);
  __syncthreads();
  for (int t = threadIdx.x; t < c.M.n_pp; t += NT) fock_diag_item<T>(c, t, Pt, Psa, xa, Psb, xb, Fa, Fb);
  const int npairs = c.M.n_atoms * (c.M.n_atoms - 1) / 2;
  for (int p = threadIdx.x; p < npairs; p += NT) fock_pair_item<T>(c, p, Psa, xa, Psb, xb, Fa, Fb);
  __syncthreads();
}

}  // namespace pm7

// SPDX-License-Identifier: GPL-3.0-or-later
//
// Device-side data layout of the PM7 engine (shared by the HIP kernels and the host driver).
//
// HBM layout of one batch (every array is one allocation, molecules concatenated):
//   atoms  : z, AO offset, packed-pair offset, xyz (Bohr)              -- per atom
//   pairs  : PairDev {mol, first, second, w_off}, triangular u<v order  -- per atom pair
//   w      : rotated two-centre blocks, row-major npair_first x npair_second, FP64
//   slab   : per molecule NSLOT square matrices of leading dimension ld (= nao rounded up to
//            even) holding H, Fock, densities, eigenvectors and the DIIS history, then the packed
//            per-atom densities used by the Coulomb contraction
//   vecs   : per molecule 2*ld eigenvalues (alpha, beta)
#pragma once

#include <stdint.h>

// The host-only sanitizer builds compile the host sources with g++: the HIP qualifiers are empty.
#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace pm7 {

constexpr double EV = 27.211386245988;        // eV per Hartree (CODATA 2018, MOPAC v23.2.5)
constexpr double A0 = 0.529177210903;         // Bohr radius in Angstrom
constexpr double EV_TO_KCAL = 23.06054783061903;
constexpr double KCAL_TO_EV = 1.0 / EV_TO_KCAL;
constexpr int MAX_Z = 103;     // atomic numbers 1..102 (the parameter tables stop at nobelium)
constexpr int NSLOT = 32;       // square matrices per molecule in the slab
constexpr int HIST = 8;         // DIIS history depth
constexpr int MIXED_MAX_ITERS = 30;  // longest reduced-precision phase of a mixed run
constexpr double ROTATE_BELOW = 1.0e-2;  // default density change below which mixed runs rotate orbitals
// A mixed run's FP64 tail converges this much tighter than p_tol (and than the 1e-7 commutator
// test): it reaches the fixed point along a different path than an FP64 run, so it has to end
// closer to it for the two densities to agree to 1e-7.
constexpr double MIXED_TAIL = 1.0e-2;
// ... for as long as it makes progress once the caller's own test (p_tol) is met: the tail stops,
// converged, after three iterations in a row without a 20 % drop of the density change below
// the tail's best (a nearly degenerate frontier settles no further), or after MIXED_TAIL_EXTRA.
constexpr int MIXED_TAIL_EXTRA = 40;

// Slab slots, RHF.
// C0/C1 double-buffer the sorted eigenvectors so each eigensolve can start from the last one;
// VG and AG hold the Jacobi rotation accumulator and the packed matrix when they do not fit LDS.
enum RhfSlot { S_H = 0, S_F = 1, S_W = 2, S_C0 = 3, S_C1 = 4, S_VG = 5, S_AG = 6, S_T = 7, S_HF = 8, S_HE = 16, S_HD = 24 };
// Slab slots, UHF (no DIIS history, so the history region is reused).
enum UhfSlot {
  U_H = 0, U_FA = 1, U_FB = 2, U_PA = 3, U_PB = 4, U_PAN = 5, U_PBN = 6, U_CA0 = 7, U_CB0 = 8,
  U_AG = 9, U_T = 10, U_CA1 = 11, U_CB1 = 12, U_PT = 13, U_VG = 14
};

// Slab slots of a divide-and-conquer subsystem (no DIIS history: the SCF's ring is free).
// RHF: gathered density, Fock, eigenvectors double-buffered across steps, eigensolver scratch,
// the occupation-weighted local density and the occupation-scaled eigenvectors.
enum DcRhfSlot { DR_P = 24, DR_F = 1, DR_C0 = 3, DR_C1 = 4, DR_VG = 5, DR_AG = 6, DR_T = 7, DR_D = 9, DR_CS = 10 };
// UHF: total, alpha and beta densities, the two Fock matrices, eigenvectors per spin (double
// buffered), scratch, the two spin channels' weighted local densities and the scaled vectors.
enum DcUhfSlot {
  DU_PT = 13, DU_PA = 3, DU_PB = 4, DU_FA = 1, DU_FB = 2, DU_CA0 = 7, DU_CA1 = 11, DU_CB0 = 8, DU_CB1 = 12,
  DU_AG = 9, DU_T = 10, DU_VG = 14, DU_DA = 15, DU_DB = 16, DU_CS = 17
};

// One subsystem-to-global contribution of the divide-and-conquer density assembly: the element
// (mu, nu) of the atom-pair block it feeds reads slab[base + mu * ld + nu] (the alpha, or only,
// channel; a UHF subsystem's beta channel sits dstep doubles further) with partition weight w.
struct DcContrib {
  int64_t base;
  int64_t dstep;
  int ld;
  int pad;
  double w;
};

// Packed local-AO descriptor of a subsystem: local atom | orbital in atom << 16 | atom's AOs << 24.
__host__ __device__ inline int dc_ao_code(int atom, int mu, int norb) { return atom | (mu << 16) | (norb << 24); }

struct ElemDev {
  double u_ss, u_pp, beta_s, beta_p, core_charge;
  double g_ss, g_sp, g_pp, g_p2, h_sp;
  double zeta_s, zeta_p, dd, qq, rho0, rho1, rho2, po9;
  double gk[4], gl[4], gm[4];
  // spd basis (MNDO/d): d-shell one-electron terms, additive terms po(1..9) and charge
  // separations ddp(2..6) of the multipole model (1-based slots)
  double u_dd, beta_d, zeta_d;
  double po[10], ddp[7];
  int n, n_orb, n_gauss, present;
  int z, iod, n_sp, n_d;  // atomic number, d occupation of the reference atom, shell quantum numbers
};

// Words of the one-centre spd matrix of a d-bearing element (45 x 45 packed orbital pairs).
constexpr int OC_D_WORDS = 45 * 45;

struct PairDev {
  int mol;       // batch-local molecule index
  int f, s;      // batch-global atom indices: `f` carries at least as many AOs as `s`
  int pad;
  int64_t w_off; // offset of the npair(f) x npair(s) block in the integral array
};

struct MolDev {
  int n_atoms, nao, ld, n_occ, n_occ_b, uhf;
  int atom0;            // batch-global index of the first atom
  int max_scf, accel, use_diis, precision, n_pp;
  int mixed_sweeps;     // Jacobi sweep cap per eigensolve in the reduced-precision phase
  int dpath;            // 1: the molecule holds a d-bearing atom, every pair takes the MNDO/d kernels
  int eig_tol_exp;      // full-accuracy eigensolves stop at off-diagonal^2 <= 10^-eig_tol_exp x total^2
  int64_t pair0;        // first pair of the molecule
  int64_t mat0;         // slab offset (doubles)
  int64_t pp0;          // slab offset of the packed densities
  int64_t vec0;         // offset into the eigenvalue array
  double e_tol, p_tol, adiis_switch, shift_floor, sad_yy, mixed_switch;
  double rotate_below;  // mixed precision: density change below which orbitals are rotated
  int init_p;           // 1: the starting density was uploaded into the first density slot(s)
  int eig_block;        // 1: block Jacobi in the size classes that offer it (0: the scalar rounds)
};

// Phases of an SCF iteration timed inside the megakernel (device wall clock, 100 MHz ticks).
// PH_EIG is split into its warm-start transform (EIG_IN), the Jacobi sweeps (EIG_SWEEP) and the
// sort + back-transform (EIG_OUT).
enum Phase {
  PH_FOCK = 0, PH_COMM = 1, PH_DIIS = 2, PH_EIG = 3, PH_DENS = 4, PH_TOTAL = 5,
  PH_EIG_IN = 6, PH_EIG_SWEEP = 7, PH_EIG_OUT = 8, NPHASE = 9
};

struct MolOut {
  double e_el, dp_err, err_norm, shift, e_fp32_last;
  // Mixed precision: energy change and max |dP| of the final full FP64 Roothaan step.
  double check_de, check_dp;
  int iterations, converged, sweeps, p_slot, c_slot, cb_slot, pb_slot, fp32_iters;
  unsigned long long ticks[NPHASE];
};

// MOPAC diagonal guess for one atom (pm7-rs src/scf.rs:1050-1114): the core charge spread over
// the s/p orbitals (a main-group d shell gets only the charge correction), a transition metal's
// filled s, then d, then p; yy spreads the molecular charge. Writes the n_orb diagonal entries.
__host__ __device__ inline void sad_atom(const ElemDev& e, double yy, double scale, double* P, int o, int ld) {
  double occ[9];
  const int n = e.n_orb;
  const double tore = e.core_charge;
  for (int k = 0; k < 9; ++k) occ[k] = 0.0;
  if (n == 1) {
    occ[0] = tore - yy;
  } else if (n == 4) {
    for (int k = 0; k < 4; ++k) occ[k] = tore * 0.25 - yy;
  } else if (n == 9) {
    const int z = e.z;
    if (z < 21 || (z > 30 && z < 39) || (z > 48 && z < 57)) {
      for (int k = 0; k < 4; ++k) occ[k] = tore * 0.25 - yy;
      for (int k = 4; k < 9; ++k) occ[k] = -yy;
    } else {
      double sum = tore - 9.0 * yy;
      occ[0] = sum < 0.0 ? 0.0 : (sum > 2.0 ? 2.0 : sum);
      sum -= 2.0;
      if (sum > 0.0) {
        for (int k = 4; k < 9; ++k) occ[k] = sum * 0.2 < 0.0 ? 0.0 : (sum * 0.2 > 2.0 ? 2.0 : sum * 0.2);
        sum -= 10.0;
        if (sum > 0.0)
          for (int k = 1; k < 4; ++k) occ[k] = sum / 3.0;
      }
    }
  }
  for (int k = 0; k < n; ++k) P[(int64_t)(o + k) * ld + o + k] = occ[k] * scale;
}

__host__ __device__ inline int npair_of(int norb) { return norb * (norb + 1) / 2; }

__host__ __device__ inline int pack_idx(int a, int b) {
  int h = a > b ? a : b, l = a > b ? b : a;
  return h * (h + 1) / 2 + l;
}

// Index of the pair (u, v), u < v, inside a molecule of n atoms (triangular, row-major).
__host__ __device__ inline int64_t tri_index(int u, int v, int n) {
  return (int64_t)u * (2 * n - u - 1) / 2 + (v - u - 1);
}

}  // namespace pm7

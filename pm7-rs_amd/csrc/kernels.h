// SPDX-License-Identifier: GPL-3.0-or-later
//
// Host-callable launchers for the engine's HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "pm7_device.h"

namespace pm7 {

// Device pointers of one prepared batch.
struct BatchDev {
  const MolDev* mols = nullptr;
  const int* mol_list = nullptr;   // molecule indices grouped by SCF kernel variant
  const int* atom_z = nullptr;
  const int* atom_ao = nullptr;    // molecule-relative AO offset
  const int* atom_pp = nullptr;    // molecule-relative packed-density offset
  const double* xyz = nullptr;     // Bohr
  const PairDev* pairs = nullptr;
  const ElemDev* elem = nullptr;
  const double2* pairpar = nullptr;
  const int2* hc_items = nullptr;  // (molecule, atom | element << 24)
  const int64_t* gath_off = nullptr;
  double* w = nullptr;
  double* slab = nullptr;
  double* vecs = nullptr;
  MolOut* out = nullptr;
  double* e_core = nullptr;        // per pair
  double* e_core_mol = nullptr;    // per molecule
  double* gath = nullptr;          // packed outputs
  double4 vpar;
  // Pairs whose overlap needs the numerical quadrature (a valence shell with n >= 4).
  const int64_t* num_pairs = nullptr;
  int n_num = 0;
  const double* gl = nullptr;      // Gauss-Legendre nodes/weights: x48, w48, x40, w40 on [-1, 1]
  double* s_num = nullptr;         // 5 local overlaps per numerical pair
  // Pairs of molecules holding a d-bearing atom (the MNDO/d kernels), and the one-centre spd
  // matrices of the d-bearing elements (OC_D_WORDS per atomic number).
  const int64_t* d_pairs = nullptr;
  int64_t n_dp = 0;
  const double* ocd = nullptr;
};

// Gauss-Legendre nodes and weights on [-1, 1] for the overlap quadrature (48 and 40 points).
constexpr int GL_WORDS = 2 * 48 + 2 * 40;

// SCF kernel size classes (padded basis size): 0 <= 64 (256 threads, eigensolver fully in LDS),
// 1 <= 120 (512 threads, packed matrix in L2, rotation accumulator in LDS), 2 <= 196 (packed
// matrix in LDS, rotations in L2), 3: larger (the grid-wide path). Classes 0 and 1, and class 2 up
// to 128 basis functions, diagonalise with the block Jacobi solver.
constexpr int SCF_VARIANTS = 4;
constexpr int VARIANT_MAX_NAO[SCF_VARIANTS] = {64, 120, 196, 1 << 30};
constexpr int LARGE_VARIANT = 3;  // above 196 basis functions: the grid-wide path (large.hip)
constexpr int GRID_FEW = 8;        // at most this many large-class-2 molecules of a batch go grid-wide
// Large-class molecules up to this basis size run in groups of up to large_group_size() SCFs whose full
// eigensolves are batched (one strided rocSOLVER call per iteration for the whole group).
constexpr int LARGE_GROUP_MAX_NAO = 512;
// 64: on a serving mix of 12-200 atom S/P/Cl chains, groups of 16/32/64 gave 134/140/146 large
// solves/s; the batcher's matrices take 2 x 64 x 512^2 doubles (268 MB) at most.
constexpr int LARGE_GROUP_DEFAULT = 64;
int large_group_size();  // LARGE_GROUP_DEFAULT, or PM7_AMD_LARGE_GROUP (2..64)

hipError_t launch_integrals(const BatchDev& b, int64_t n_pairs, int n_items, int n_mols, hipStream_t s);
hipError_t launch_scf(int variant, const BatchDev& b, int list_off, int count, hipStream_t s);
hipError_t launch_zero(double* p, int64_t words, hipStream_t s);
hipError_t launch_gather(const BatchDev& b, int n_mols, hipStream_t s);
// Nuclear gradient at the converged densities of a finished batch: dE/dd per pair into gpair
// (3 per pair), then per listed atom (items: molecule, batch-global atom) into gatom (3 per atom).
// s_numg: 60 doubles per numerical-overlap pair (its overlaps at the 12 displaced separations).
hipError_t launch_gradient(const BatchDev& b, int64_t n_pairs, const int2* items, int n_items, double* gpair,
                           double* gatom, double* s_numg, hipStream_t s);
hipError_t launch_fock_once(const BatchDev& b, int mol, int pt_slot, int ps_slot, int f_slot, hipStream_t s);
// Closed-shell response Fock of n_dens densities of molecule `mol` as contractions over the
// density index (response.hip): wall (n_pp^2 doubles), pp and fd (n_pp x n_dens each) are scratch;
// X and G density-fastest, element (d, i, j) at (i * nao + j) * dp + d. spd: some atom has d
// orbitals (else every atom has 1 or 4).
hipError_t launch_response_pairs(const BatchDev& b, int mol, int n_pp, int n_atoms, bool spd, const double* xs,
                                 double* gs, int n_dens, int dp, double* wall, double* pp, double* fd, hipStream_t s);
// Densities between the caller's layout ((d * nao + i) * nao + j) and the density-fastest one
// (to_pairs: caller's to density-fastest).
hipError_t launch_response_relayout(const double* src, double* dst, int nao, int n_dens, int dp, bool to_pairs,
                                    hipStream_t s);
// Row stride of the density-fastest layout: whole 128-byte lines per (i, j).
inline int response_dp(int n_dens) { return (n_dens + 15) / 16 * 16; }
// The contraction path's limit: the n_pp^2 integral matrix (<= 512 MB).
inline bool response_pairs_fit(int n_pp) { return n_pp <= 8192; }
// G_k = J(X_k) - K(X_k / 2) for n_dens densities of molecule `mol` (slices of nao*ld at xs / gs;
// pp_stride doubles of packed-density scratch per density at pps).
// 512 threads per density: 0.41 ms for 294 densities of a 194-AO molecule (256: 0.58, 1024: 0.55),
// measured with the subtraction of H as a separate pass (0.38 ms with it fused into the stores).
constexpr int RESPONSE_NT_DEFAULT = 512;
hipError_t launch_fock_response(const BatchDev& b, int mol, const double* xs, double* gs, double* pps,
                                int64_t pp_stride, int n_dens, hipStream_t s);
// Open shell: G^a, G^b = J(Xa + Xb) - K(X^s) for n_dens pairs (alpha slices then beta slices at xs
// and gs, sums at ts; each slice nao*ld).
hipError_t launch_fock_response_spin(const BatchDev& b, int mol, const double* xs, double* ts, double* gs,
                                     double* pps, int64_t pp_stride, int n_dens, hipStream_t s);

// Divide and conquer: device arrays of one system's partition (the subsystems are the molecules of
// a prepared batch, in subsystem order; "batch atoms" are the subsystem atoms, core atoms first).
struct DcDev {
  int n_atoms = 0, n_sub = 0, uhf = 0, n_batch_atoms = 0;
  int64_t n_slots = 0, n_el = 0, n_states = 0;
  const int64_t* diag_off = nullptr;  // per system atom: element offset of its diagonal block
  const int* norb = nullptr;          // per system atom
  const double* xyz = nullptr;        // system positions (Bohr)
  const double* core = nullptr;       // system core charges
  const int* parent = nullptr;        // per batch atom: its system atom
  const int* atom_mol = nullptr;      // per batch atom: its subsystem
  const int* nc_ao = nullptr;         // per subsystem: AOs of its core atoms
  const int* aocode = nullptr;        // per subsystem AO: dc_ao_code(local atom, orbital, atom's AOs)
  const int64_t* ao0 = nullptr;       // per subsystem: first entry in aocode
  const int64_t* amap = nullptr;      // per subsystem local atom pair (row-major): global block offset or -1
  const int64_t* amap0 = nullptr;     // per subsystem: first entry in amap
  const int64_t* slot_off = nullptr;  // per stored block: element offset
  const int* slot_dims = nullptr;     // per stored block: rows | cols << 8
  const int64_t* cstart = nullptr;    // per stored block: first contribution (n_slots + 1 entries)
  const DcContrib* contrib = nullptr;
  const double* normalizer = nullptr;  // per stored block: 1 / total partition weight
  double* P = nullptr;                 // total density (stored elements)
  double* S = nullptr;                 // spin density (UHF)
  double* pop = nullptr;               // per system atom
  double* g = nullptr;                 // per system atom: global monopole potential
  double* v_nuc = nullptr;             // per batch atom: nuclear environment
  double* v_el = nullptr;              // per batch atom: electronic environment
  double* wnorm = nullptr;             // like vecs: weighted norm per state (0 on padding)
  double* e_part = nullptr;            // per subsystem: weighted energy
  double* partial = nullptr;           // per stored block: squared change
  double* scal = nullptr;              // [rms, e_el, fermi, fermi failed]
  double* fermi_part = nullptr;        // 2 x DC_FERMI_PART_WORDS: partial counts of the Fermi passes
  double* fermi_state = nullptr;       // 2 x 4: bracket and flags of the Fermi passes
};

constexpr int DC_FERMI_PART_WORDS = 128 * 16;  // workgroups of a Fermi pass x 16 partial counts

hipError_t launch_dc_solve(int variant, const BatchDev& b, int list_off, int count, const int* nc_ao,
                           const double* v_el, int step, double* wnorm, double* e_part, hipStream_t s);
hipError_t launch_dc_density(const BatchDev& b, int n_sub, const double* scal, double width, int step,
                             hipStream_t s);
hipError_t launch_dc_pop(const DcDev& d, hipStream_t s);
// g = global monopole potential of charges q, v = each subsystem atom's environment from it.
hipError_t launch_dc_field(const BatchDev& b, const DcDev& d, const double* q, double* g, double* v,
                           hipStream_t s);
hipError_t launch_dc_shift_h(const BatchDev& b, const DcDev& d, hipStream_t s);
hipError_t launch_dc_gather(const BatchDev& b, const DcDev& d, hipStream_t s);
hipError_t launch_dc_fermi(const BatchDev& b, const DcDev& d, double n_elec, double width, hipStream_t s);
hipError_t launch_dc_assemble(const BatchDev& b, const DcDev& d, double mixing, hipStream_t s);
hipError_t launch_dc_finish(const DcDev& d, hipStream_t s);

// Large-system SCF of molecule `mi` (host copy M) of a prepared batch: grid-wide kernels and
// rocBLAS/rocSOLVER on stream s; fills `out` as the megakernel would. Blocks the calling thread.
struct LargeWork;
LargeWork* large_work_create();
void large_work_destroy(LargeWork* w);
// Several large-system SCFs in their own threads can share a batched eigensolve: created for
// `participants` SCFs of padded size up to nmax; each passes it to large_scf.
struct EigBatcher;
EigBatcher* eig_batcher_create(int participants, int nmax, std::string& err);
void eig_batcher_destroy(EigBatcher* b);
hipError_t large_scf(LargeWork* w, const BatchDev& b, const MolDev& M, int mi, hipStream_t s, MolOut& out,
                     std::string& err, EigBatcher* eb = nullptr);

}  // namespace pm7

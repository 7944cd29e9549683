// SPDX-License-Identifier: GPL-3.0-or-later
//
// Host driver and C ABI of the PM7 engine (drop-in for pm7-rs run_pm7, src/scf.rs:512-867).
//
// A call (single molecule or batch) is: validate on the host -> lay the batch out in HBM (atoms,
// atom pairs, integral stream, per-molecule matrix slab) -> one integral launch for every pair of
// every molecule -> one SCF megakernel launch per size class (molecules run concurrently, each
// in its own workgroup, each stopping when it converges) -> one gather launch -> one copy back.
// The geometry-only corrections (dispersion, EH+, MM) run on the host while the device works.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <memory>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pm7_amd.h"
#include "host.h"
#include "dandc.h"
#include "kernels.h"

using namespace pm7;

// Device scratch of pm7_fock_spin_noalloc: the batch layout of one molecule (atoms, atom pairs,
// pair integrals, matrix slab, outputs) for callers that cannot allocate device memory. One copy
// per GPU, held until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (pm7_ctx_create), never through the name.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

// Grow-only pinned host buffer (device-to-host copies at full link speed).
struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      (void)hipHostFree(p);
      p = nullptr;
      cap = 0;
    }
    size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Bump allocator over a fixed device region (the allocation-free entry point's static scratch):
// reset at the start of a call, each buffer carved at the next 256-byte boundary.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, off = 0;
};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  Arena* arena = nullptr;  // set: carve from the arena, never hipMalloc / hipFree
  hipError_t ensure(size_t bytes) {
    if (arena) {
      const size_t at = (arena->off + 255) & ~(size_t)255, want = std::max<size_t>(bytes, 256);
      if (!arena->base || at + want > arena->cap) return hipErrorOutOfMemory;
      p = arena->base + at;
      cap = want;
      arena->off = at + want;
      return hipSuccess;
    }
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return e;
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p && !arena) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

double nan_() { return std::nan(""); }

// Atomic masses (u) for the centre-of-mass dipole origin (pm7-rs src/data_tables.rs MASS: standard
// atomic weights through Rn; heavier elements have none there and count as 0 here).
const double MASS[87] = {
    0.0,     1.0079,  4.0026,  6.94,    9.01218, 10.81,   12.011,  14.0067, 15.9994, 18.9984, 20.179,  22.98977,
    24.305,  26.98154, 28.0855, 30.97376, 32.06,  35.453,  39.948,  39.098,  40.078,  44.956,  47.867,  50.942,
    51.996,  54.938,  55.845,  58.933,  58.693,  63.546,  65.38,   69.723,  72.63,   74.922,  78.971,  79.904,
    83.798,  85.468,  87.62,   88.906,  91.224,  92.906,  95.95,   97.0,    101.07,  102.91,  106.42,  107.87,
    112.41,  114.82,  118.71,  121.76,  127.6,   126.9,   131.29,  132.91,  137.33,  138.91,  140.12,  140.91,
    144.24,  145.0,   150.36,  151.96,  157.25,  158.93,  162.5,   164.93,  167.26,  168.93,  173.05,  174.97,
    178.49,  180.95,  183.84,  186.21,  190.23,  192.22,  195.08,  196.97,  200.59,  204.38,  207.2,   208.98,
    209.0,   210.0,   222.0};

std::string default_data_dir() {
  Dl_info info;
  if (dladdr((void*)&default_data_dir, &info) && info.dli_fname) {
    std::string so(info.dli_fname);
    auto slash = so.rfind('/');
    std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
    return dir + "/../data";
  }
  return "data";
}

// One validated molecule of a call.
struct Job {
  int idx = 0;
  pm7_options_v2 o{};
  std::vector<int> z;
  std::vector<double> xyz;
  int nao = 0, ld = 0, n_pp = 0, na = 0, nb = 0, uhf = 0, variant = 0;
  int dpath = 0;  // holds a d-bearing atom (or force_dpath): every pair through the MNDO/d kernels
  double sad_yy = 0.0;
  int missing_z = 0;  // MISSING_ELEMENT: the atomic number without parameters
  // Starting densities (initial_density / initial_spin_density), already split into the
  // matrices the SCF starts from: RHF P (p0), UHF P_alpha (p0) and P_beta (p1); n x n row-major.
  std::vector<double> p0, p1;
  int64_t pair0 = 0, mat0 = 0, pp0 = 0, vec0 = 0, gath0 = 0, w0 = 0;
  int atom0 = 0;
  double corr_kcal = 0.0;
};

}  // namespace

namespace {

// Device resources of one batch in flight: streams (one per SCF size class, stream 0 also
// carries uploads, integrals and the output gather), events, and grow-only device buffers.
// The synchronous batch path uses the context's lane; the request server owns several.
struct Lane {
  // One stream for everything (request-server lanes: hardware queues are few, and a stream that
  // shares a queue with another lane's long kernel would wait behind it) or one per size class.
  bool single_stream = false;
  hipStream_t st[SCF_VARIANTS] = {};
  hipEvent_t ev[12] = {};
  hipEvent_t ev_var[2 * SCF_VARIANTS] = {};
  DevBuf d_mols, d_list, d_atom_z, d_atom_ao, d_atom_pp, d_xyz, d_pairs, d_items, d_gath_off;
  DevBuf d_w, d_slab, d_vecs, d_out, d_ecore, d_ecore_mol, d_gath, d_num, d_snum, d_dpairs;
  DevBuf d_gpair, d_gatom, d_gitems, d_snumg;  // nuclear gradient (pm7_gradient_batch only)
  DevBuf d_rx, d_rg, d_rpp, d_rstage, d_rwall, d_rfd;          // response densities, their G, packed scratch, staging
  PinnedBuf h_out;  // staging for the packed results
  std::vector<int2> h_gitems;  // host source of d_gitems
  LargeWork* large = nullptr;  // rocBLAS handle and buffers of the large-system path (lazily made)
  // Several large-system SCFs in lockstep: one work set and one stream each (lazily made).
  std::vector<LargeWork*> group_work;
  std::vector<hipStream_t> group_st;
  int device = 0;
  std::string err;
  bool create(bool one_stream = false) {
    single_stream = one_stream;
    if (hipGetDevice(&device) != hipSuccess) return false;
    for (int v = 0; v < (single_stream ? 1 : SCF_VARIANTS); ++v)
      if (hipStreamCreateWithFlags(&st[v], hipStreamNonBlocking) != hipSuccess) return false;
    for (int v = 1; single_stream && v < SCF_VARIANTS; ++v) st[v] = st[0];
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) return false;
    for (auto& e : ev_var)
      if (hipEventCreate(&e) != hipSuccess) return false;
    return true;
  }
  void destroy() {
    for (auto& s : st)
      if (s) (void)hipStreamSynchronize(s);
    DevBuf* bufs[] = {&d_mols, &d_list, &d_atom_z, &d_atom_ao, &d_atom_pp, &d_xyz, &d_pairs, &d_items, &d_gath_off,
                      &d_w, &d_slab, &d_vecs, &d_out, &d_ecore, &d_ecore_mol, &d_gath, &d_num, &d_snum,
                      &d_dpairs, &d_gpair, &d_gatom, &d_gitems, &d_snumg, &d_rx, &d_rg, &d_rpp, &d_rstage,
                      &d_rwall, &d_rfd};
    for (DevBuf* b : bufs) b->release();
    h_out.release();
    if (large) large_work_destroy(large);
    large = nullptr;
    for (size_t k = 0; k < group_work.size(); ++k) {
      if (group_st[k]) (void)hipStreamSynchronize(group_st[k]);
      large_work_destroy(group_work[k]);
      if (group_st[k]) (void)hipStreamDestroy(group_st[k]);
    }
    group_work.clear();
    group_st.clear();
    for (int v = 0; v < (single_stream ? 1 : SCF_VARIANTS); ++v)
      if (st[v]) (void)hipStreamDestroy(st[v]);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : ev_var)
      if (e) (void)hipEventDestroy(e);
    for (auto& s : st) s = nullptr;
    for (auto& e : ev) e = nullptr;
    for (auto& e : ev_var) e = nullptr;
  }
  size_t held() const { return d_w.cap + d_slab.cap + d_gath.cap + d_pairs.cap + d_ecore.cap; }
  // Every device buffer of this lane is carved from `a` (the allocation-free lane).
  void use_arena(Arena* a) {
    DevBuf* bufs[] = {&d_mols, &d_list, &d_atom_z, &d_atom_ao, &d_atom_pp, &d_xyz, &d_pairs, &d_items, &d_gath_off,
                      &d_w, &d_slab, &d_vecs, &d_out, &d_ecore, &d_ecore_mol, &d_gath, &d_num, &d_snum,
                      &d_dpairs, &d_gpair, &d_gatom, &d_gitems, &d_snumg, &d_rx, &d_rg, &d_rpp, &d_rstage,
                      &d_rwall, &d_rfd};
    for (DevBuf* b : bufs) b->arena = a;
  }
};

}  // namespace

namespace {
struct AsyncBatch;
}

struct pm7_ctx {
  int device = 0;
  std::string data_dir;
  std::string err;
  ParamSet params;
  int method = -1;  // pm7_method of the loaded parameters (what options.method must equal)
  int tables = -1;  // which bundled tables are loaded (PM7 or PM7-TS)
  std::vector<ElemDev> elem_host;
  DevBuf d_elem, d_pairpar, d_gl, d_ocd;
  Lane lane;
  double timings[PM7_TIMING_SLOTS] = {};
  // Jacobi sweep cap of each full eigensolve in the reduced-precision phase of PRECISION_MIXED
  // runs (environment PM7_AMD_MIXED_SWEEPS, 1..40): three sweeps of the warm-started solve are
  // enough for the FP32 density; the DIIS-driven iterations absorb the rest.
  int mixed_sweeps = 3;
  // PRECISION_MIXED: density change below which the SCF rotates orbitals instead of
  // diagonalising (environment PM7_AMD_ROTATE_BELOW, 1e-6..1e-1).
  double rotate_below = ROTATE_BELOW;
  // Stopping test of the full-accuracy Jacobi eigensolves: off-diagonal^2 <= 10^-eig_tol_exp x
  // total^2 (environment PM7_AMD_EIG_TOL_EXP, 16..30). 1e-22 keeps every density of the
  // 256-molecule organic batch within 1e-7 of a tightly converged CPU solve (1e-20 left one at
  // 1.3e-7); against 1e-26 it saves 0.25 sweeps per solve (FP64 organic batch 7.40k -> 7.57k
  // solves/s, S/P/Cl batch 415 -> 437/s).
  int eig_tol_exp = 22;
  // Block Jacobi (16 x 16 subproblems, rotations applied as matrix-core products): bit v on means
  // size class v uses it where the class offers it (environment PM7_AMD_EIG_BLOCK: 0 the scalar
  // rounds everywhere, 1 block everywhere, larger values a mask; pm7_autotune sets it).
  int eig_block_mask = 0xF;
  // A batch holding at most GRID_FEW molecules above grid_min_nao basis functions (in the
  // largest megakernel class) runs them on the grid-wide path: one such molecule alone on one CU
  // takes 210 ms at 194 AOs against 58 ms grid-wide, while many of them side by side keep the
  // device busier as megakernels (environment PM7_AMD_GRID_MIN_NAO, 32..196, default 128).
  int grid_min_nao = 128;
  // Divide and conquer replays each step after the first as a captured HIP graph (environment
  // PM7_AMD_DC_GRAPH=0: every step's launches issued one by one).
  bool dc_graph = true;
  // Closed-shell response Fock as density-index contractions (PM7_AMD_RESPONSE_CONTRACT=0: one
  // workgroup per density, the Fock-build kernel of the SCF).
  bool response_contract = true;
  // Allocation-free entry point: its lane's buffers live in the device's static scratch array.
  Arena arena;
  Lane na_lane;
  bool na_ready = false;
  std::string na_err;
  // Asynchronous batches (pm7_batch_submit / pm7_batch_wait): up to async_max batches in flight,
  // each on its own lane; lanes are made on first use and kept.
  std::vector<Lane*> async_lanes;
  std::deque<AsyncBatch*> inflight;               // submit order
  std::unordered_map<int64_t, pm7_status> async_done;  // collected early (a submit needed the lane)
  int64_t next_ticket = 1;
  // 2: on the 256-molecule organic batch, two batches in flight raised the device throughput
  // 8.0k -> 10.2k solves/s (FP64) and 12.3k -> 15.6k (mixed) with step latency p99 47 / 32 ms;
  // 4-6 in flight reached 11-14k / 17-21k at twice the latency (16 hardware queues).
  int async_max = 2;
  bool async_one_stream = false;
  hipEvent_t epoch = nullptr;  // recorded at the first asynchronous submit: origin of timings [12], [13]
  std::mutex mu;
};

namespace {

pm7_status fail(pm7_ctx* ctx, pm7_status st, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return st;
}

void set_msg(pm7_result_v2* r, const std::string& m) { std::snprintf(r->message, sizeof r->message, "%s", m.c_str()); }

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(ctx, PM7_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// Same, recording the message on a lane (lane work may run on the server's thread).
#define LANECHK(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      L.err = std::string(#expr ": ") + hipGetErrorString(e_);                               \
      return PM7_ERR_DEVICE;                                                                 \
    }                                                                                        \
  } while (0)

const char* method_name(int m) {
  switch (m) {
    case PM7_METHOD_PM7: return "pm7";
    case PM7_METHOD_PM7_TS: return "pm7-ts";
    case PM7_METHOD_PM7_SPARKLE: return "pm7-sparkle";
    case PM7_METHOD_PM7_MINUS: return "pm7-minus";
    case PM7_METHOD_PM7_HH: return "pm7-hh";
    default: return "unknown";
  }
}

// Methods this engine runs: the four pm7-rs methods with orbital atoms only (src/method.rs:9-16).
pm7_status method_ok(int m, std::string& msg) {
  if (m == PM7_METHOD_PM7 || m == PM7_METHOD_PM7_TS || m == PM7_METHOD_PM7_MINUS || m == PM7_METHOD_PM7_HH)
    return PM7_OK;
  if (m == PM7_METHOD_PM7_SPARKLE) {
    msg = "PM7-Sparkle (lanthanide point-charge cores) is not offered by this engine";
    return PM7_ERR_UNSUPPORTED;
  }
  msg = "unknown PM7 method code " + std::to_string(m);
  return PM7_ERR_INVALID_INPUT;
}

// Input validation (pm7-rs validate_input src/scf.rs:937-992, run_pm7_once 530-627).
pm7_status validate(pm7_ctx* ctx, const pm7_molecule_v1& m, const pm7_options_v2& o, Job& j, std::string& msg) {
  if (m.n_atoms <= 0 || !m.z || !m.xyz_bohr) {
    msg = "a PM7 calculation needs at least one atom";
    return PM7_ERR_INVALID_INPUT;
  }
  if (!std::isfinite(o.charge)) return msg = "charge must be finite", PM7_ERR_INVALID_INPUT;
  if (o.max_scf <= 0) return msg = "max_scf must be greater than zero", PM7_ERR_INVALID_INPUT;
  if (!(std::isfinite(o.e_tol) && o.e_tol > 0)) return msg = "e_tol must be finite and positive", PM7_ERR_INVALID_INPUT;
  if (!(std::isfinite(o.p_tol) && o.p_tol > 0)) return msg = "p_tol must be finite and positive", PM7_ERR_INVALID_INPUT;
  if (!(std::isfinite(o.level_shift) && o.level_shift >= 0))
    return msg = "level_shift must be finite and non-negative", PM7_ERR_INVALID_INPUT;
  if (!(std::isfinite(o.adiis_switch) && o.adiis_switch >= 0))
    return msg = "adiis_switch must be finite and non-negative", PM7_ERR_INVALID_INPUT;
  if (o.multiplicity < 1) return msg = "multiplicity must be >= 1", PM7_ERR_INVALID_INPUT;
  {
    std::string m2;
    const pm7_status ms = method_ok(o.method, m2);
    if (ms != PM7_OK) return msg = m2, ms;
  }
  if (o.method != ctx->method)
    return msg = std::string("options method `") + method_name(o.method) + "` does not match parameter method `" +
                 method_name(ctx->method) + "`",
           PM7_ERR_INVALID_INPUT;
  if (o.dipole_origin < PM7_DIPOLE_CENTRE_OF_MASS || o.dipole_origin > PM7_DIPOLE_CENTRE_OF_CHARGE)
    return msg = "unknown dipole origin " + std::to_string(o.dipole_origin), PM7_ERR_INVALID_INPUT;
  const int n = m.n_atoms;
  j.z.assign(m.z, m.z + n);
  j.xyz.assign(m.xyz_bohr, m.xyz_bohr + 3 * n);
  for (int i = 0; i < 3 * n; ++i)
    if (!std::isfinite(j.xyz[i])) {
      msg = "atom " + std::to_string(i / 3 + 1) + " has non-finite coordinates";
      return PM7_ERR_INVALID_INPUT;
    }
  // Coincident atoms, O(N) bucket grid (src/scf.rs:1001-1040).
  {
    const double h = 1.0e-6;
    std::unordered_map<long long, std::vector<int>> cells;
    auto key = [](long long a, long long b, long long c) { return (a * 73856093LL) ^ (b * 19349663LL) ^ (c * 83492791LL); };
    for (int i = 0; i < n; ++i) {
      long long cx = (long long)std::floor(j.xyz[3 * i] / h), cy = (long long)std::floor(j.xyz[3 * i + 1] / h),
                cz = (long long)std::floor(j.xyz[3 * i + 2] / h);
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          for (int c = -1; c <= 1; ++c) {
            auto it = cells.find(key(cx + a, cy + b, cz + c));
            if (it == cells.end()) continue;
            for (int k : it->second) {
              double dx = j.xyz[3 * i] - j.xyz[3 * k], dy = j.xyz[3 * i + 1] - j.xyz[3 * k + 1],
                     dz = j.xyz[3 * i + 2] - j.xyz[3 * k + 2];
              if (dx * dx + dy * dy + dz * dz < 1.0e-20) {
                msg = "atoms " + std::to_string(std::min(i, k) + 1) + " and " + std::to_string(std::max(i, k) + 1) +
                      " occupy the same position";
                return PM7_ERR_INVALID_INPUT;
              }
            }
          }
      cells[key(cx, cy, cz)].push_back(i);
    }
  }
  double n_elec = 0.0, tore = 0.0;
  int nao = 0, npp = 0;
  for (int i = 0; i < n; ++i) {
    int z = j.z[i];
    if (z <= 0 || z >= MAX_Z || !ctx->params.el[z].present) {
      msg = "element " + std::to_string(z) + " has no PM7 parameters";
      j.missing_z = z;
      return PM7_ERR_MISSING_ELEMENT;
    }
    const HostElement& e = ctx->params.el[z];
    tore += e.core_charge;
    if (e.n_orb == 9) j.dpath = 1;
    nao += e.n_orb;
    npp += npair_of(e.n_orb);
  }
  n_elec = tore - o.charge;
  long ne = std::lround(n_elec);
  if (std::fabs(n_elec - (double)ne) > 1.0e-6 || ne < 0) {
    msg = "invalid electron count " + std::to_string(n_elec);
    return PM7_ERR_INVALID_INPUT;
  }
  long unp = o.multiplicity - 1;
  if (ne - unp < 0 || (ne - unp) % 2 != 0) {
    msg = "electron count " + std::to_string(ne) + " is incompatible with multiplicity " + std::to_string(o.multiplicity);
    return PM7_ERR_INVALID_INPUT;
  }
  j.na = (int)((ne + unp) / 2);
  j.nb = (int)((ne - unp) / 2);
  if (j.na > nao) {
    // More electrons of one spin than orbitals (e.g. a large negative charge): the reference's
    // occupied-column slice would fail; no density can hold them.
    msg = std::to_string(j.na) + " electrons of one spin do not fit in " + std::to_string(nao) + " orbitals";
    return PM7_ERR_INVALID_INPUT;
  }
  if (o.reference == PM7_REF_RHF && j.na != j.nb) {
    msg = "restricted (RHF) needs a closed shell; use UHF";
    return PM7_ERR_INVALID_INPUT;
  }
  j.uhf = o.reference == PM7_REF_AUTO ? (j.na != j.nb) : (o.reference == PM7_REF_UHF);
  if (o.force_dpath) j.dpath = 1;
  // Seeded densities (src/scf.rs:1201-1210 RHF, 1341-1378 UHF): a wrong shape is refused.
  if (o.initial_density || o.initial_spin_density) {
    if (o.initial_density_nao != nao) {
      msg = "the supplied initial density is " + std::to_string(o.initial_density_nao) + "x" +
            std::to_string(o.initial_density_nao) + " and this system's basis is " + std::to_string(nao) + "x" +
            std::to_string(nao);
      return PM7_ERR_INVALID_INPUT;
    }
    if (o.initial_spin_density && !o.initial_density) {
      msg = "an initial spin density needs an initial density";
      return PM7_ERR_INVALID_INPUT;
    }
    const size_t nn = (size_t)nao * nao;
    for (size_t k = 0; k < nn; ++k)
      if (!std::isfinite(o.initial_density[k]) || (o.initial_spin_density && !std::isfinite(o.initial_spin_density[k]))) {
        msg = "the supplied initial density has non-finite entries";
        return PM7_ERR_INVALID_INPUT;
      }
    if (!j.uhf) {
      j.p0.assign(o.initial_density, o.initial_density + nn);
    } else {
      const double ntot = (double)std::max(j.na + j.nb, 1);
      const double fa = j.na / ntot, fb = j.nb / ntot;
      j.p0.resize(nn);
      j.p1.resize(nn);
      for (size_t k = 0; k < nn; ++k) {
        const double sd = o.initial_spin_density ? 0.5 * o.initial_spin_density[k] : 0.0;
        j.p0[k] = o.initial_density[k] * fa + sd;
        j.p1[k] = o.initial_density[k] * fb - sd;
      }
    }
  }
  j.nao = nao;
  j.ld = nao + (nao & 1);
  j.n_pp = npp;
  j.sad_yy = (tore - (double)(j.na + j.nb)) / nao;
  j.variant = -1;
  for (int v = 0; v < SCF_VARIANTS; ++v)
    if (nao <= VARIANT_MAX_NAO[v]) {
      j.variant = v;
      break;
    }
  if (j.variant < 0) {
    msg = "no SCF path for " + std::to_string(nao) + " basis functions";
    return PM7_ERR_UNSUPPORTED;
  }
  j.o = o;
  // The seeded densities were copied into the job (a queued request outlives the caller's arrays).
  j.o.initial_density = nullptr;
  j.o.initial_spin_density = nullptr;
  return PM7_OK;
}

// Host copies of a batch's device layout (the sources of its uploads).
struct HostLayout {
  std::vector<int> atom_z, atom_ao, atom_pp, list;
  std::vector<double> xyz;
  std::vector<PairDev> pairs;
  std::vector<int2> items;
  std::vector<int64_t> gath_off, num_pairs, d_pairs;
};

struct Prepared {
  HostLayout host;
  std::vector<Job> jobs;
  std::list<std::vector<double>> staging;  // host copies of seeded densities (stable addresses)
  double need_mb = 0.0, budget_mb = 0.0;   // INSUFFICIENT_MEMORY payload
  bool fixed_classes = false;  // keep every job in its size class (no grid-wide routing)
  size_t extra_bytes = 0;      // device memory the caller adds beside the batch (budget check)
  std::vector<MolDev> mols;
  int64_t n_pairs = 0, w_words = 0, slab_words = 0, vec_words = 0, gath_words = 0, n_dpairs = 0;
  int n_items = 0;
  int list_off[SCF_VARIANTS + 1] = {};
  BatchDev b;
};

size_t bytes_needed(const Prepared& P) {
  return (size_t)(P.w_words + P.slab_words + P.vec_words + P.gath_words + P.n_dpairs) * 8 +
         (size_t)P.n_pairs * (sizeof(PairDev) + 8) +
         P.jobs.size() * (sizeof(MolDev) + sizeof(MolOut) + 64);
}

}  // namespace

int pm7::large_group_size() {
  static const int g = [] {
    const char* v = std::getenv("PM7_AMD_LARGE_GROUP");
    const int x = v ? std::atoi(v) : 0;
    return (x >= 2 && x <= 64) ? x : LARGE_GROUP_DEFAULT;
  }();
  return g;
}

namespace {

// Words a pair of a d-path molecule takes in the integral array: the npair(f) x npair(s) block and
// the packed electron-core blocks of both atoms.
int64_t pair_words_d(int nf, int ns) {
  return (int64_t)npair_of(nf) * npair_of(ns) + npair_of(nf) + npair_of(ns);
}

// Device bytes one validated job adds to a batch (the terms of bytes_needed).
size_t job_bytes(const pm7_ctx* ctx, const Job& j) {
  const int n = (int)j.z.size();
  int64_t w = 0;
  for (int u = 0; u < n; ++u)
    for (int v = u + 1; v < n; ++v) {
      const int nu = ctx->params.el[j.z[u]].n_orb, nv = ctx->params.el[j.z[v]].n_orb;
      w += j.dpath ? pair_words_d(nu, nv) + 1 : (int64_t)npair_of(nu) * npair_of(nv);
    }
  const int64_t slab = (int64_t)NSLOT * j.ld * j.ld + ((j.n_pp + 7) / 8) * 8;
  const int64_t words = w + slab + 2 * j.ld + (int64_t)(j.uhf ? 4 : 2) * j.nao * j.nao;
  return (size_t)words * 8 + (size_t)n * (n - 1) / 2 * (sizeof(PairDev) + 8) + sizeof(MolDev) + sizeof(MolOut) + 64;
}

// Lay out, upload and run the once-per-geometry kernels.
pm7_status prepare(pm7_ctx* ctx, Lane& L, Prepared& P, int64_t max_memory_mb) {
  const int nm = (int)P.jobs.size();
  // The host arrays of the uploads live in P: hipMemcpyAsync from pageable memory may read the
  // source after it returns, so a source must outlive the stream's use of it (the call's end).
  HostLayout& hl = P.host;
  std::vector<int>& atom_z = hl.atom_z;
  std::vector<int>& atom_ao = hl.atom_ao;
  std::vector<int>& atom_pp = hl.atom_pp;
  std::vector<double>& xyz = hl.xyz;
  std::vector<PairDev>& pairs = hl.pairs;
  std::vector<int2>& items = hl.items;
  std::vector<int64_t>& gath_off = hl.gath_off;
  std::vector<int64_t>& num_pairs = hl.num_pairs;  // pairs whose overlaps need the quadrature (a shell with n >= 4)
  std::vector<int64_t>& d_pairs = hl.d_pairs;      // pairs of molecules with a d-bearing atom
  hl = HostLayout();
  gath_off.assign(nm, 0);
  int64_t w_off = 0, mat = 0, vec = 0, gath = 0, pair0 = 0;
  int atom0 = 0;
  P.mols.resize(nm);
  std::vector<int> lists[SCF_VARIANTS];
  if (!P.fixed_classes) {
    int few = 0;
    for (const Job& j : P.jobs) few += (j.variant == LARGE_VARIANT - 1 && j.nao > ctx->grid_min_nao) ? 1 : 0;
    if (few > 0 && few <= GRID_FEW)
      for (Job& j : P.jobs)
        if (j.variant == LARGE_VARIANT - 1 && j.nao > ctx->grid_min_nao) j.variant = LARGE_VARIANT;
  }
  for (int mi = 0; mi < nm; ++mi) {
    Job& j = P.jobs[mi];
    const int n = (int)j.z.size();
    j.atom0 = atom0;
    j.pair0 = pair0;
    int ao = 0, pp = 0;
    for (int i = 0; i < n; ++i) {
      const HostElement& e = ctx->params.el[j.z[i]];
      atom_z.push_back(j.z[i]);
      atom_ao.push_back(ao);
      atom_pp.push_back(pp);
      for (int k = 0; k < 3; ++k) xyz.push_back(j.xyz[3 * i + k]);
      for (int el = 0; el < npair_of(e.n_orb); ++el) items.push_back(make_int2(mi, (atom0 + i) | (el << 24)));
      ao += e.n_orb;
      pp += npair_of(e.n_orb);
    }
    for (int u = 0; u < n; ++u)
      for (int v = u + 1; v < n; ++v) {
        int nu = ctx->params.el[j.z[u]].n_orb, nv = ctx->params.el[j.z[v]].n_orb;
        PairDev pd;
        pd.mol = mi;
        pd.pad = 0;
        if (nu >= nv) {
          pd.f = atom0 + u;
          pd.s = atom0 + v;
        } else {
          pd.f = atom0 + v;
          pd.s = atom0 + u;
        }
        pd.w_off = w_off;
        if (j.dpath) {  // w block, then both electron-core blocks; spd overlaps in closed form
          w_off += pair_words_d(std::max(nu, nv), std::min(nu, nv));
          d_pairs.push_back((int64_t)pairs.size());
        } else {
          w_off += (int64_t)npair_of(std::max(nu, nv)) * npair_of(std::min(nu, nv));
        }
        if (!j.dpath && (std::max(ctx->params.el[j.z[u]].n, ctx->params.el[j.z[v]].n) >= 4 ||
                         std::min(ctx->params.el[j.z[u]].n, ctx->params.el[j.z[v]].n) < 1)) {
          num_pairs.push_back((int64_t)pairs.size());
          pd.pad = (int)num_pairs.size();  // 1-based slot of its numerical overlaps
        }
        pairs.push_back(pd);
      }
    j.mat0 = mat;
    j.pp0 = mat + (int64_t)NSLOT * j.ld * j.ld;
    mat = j.pp0 + ((j.n_pp + 7) / 8) * 8;
    j.vec0 = vec;
    vec += 2 * j.ld;
    j.gath0 = gath;
    gath_off[mi] = gath;
    gath += (int64_t)(j.uhf ? 4 : 2) * j.nao * j.nao;
    MolDev& M = P.mols[mi];
    std::memset(&M, 0, sizeof M);
    M.n_atoms = n;
    M.nao = j.nao;
    M.ld = j.ld;
    M.n_occ = j.na;
    M.n_occ_b = j.nb;
    M.uhf = j.uhf;
    M.atom0 = atom0;
    M.max_scf = j.o.max_scf;
    M.accel = j.o.use_diis ? j.o.accelerator : 0;
    M.use_diis = j.o.use_diis;
    M.precision = j.o.precision;
    M.n_pp = j.n_pp;
    M.pair0 = pair0;
    M.mat0 = j.mat0;
    M.pp0 = j.pp0;
    M.vec0 = j.vec0;
    M.e_tol = j.o.e_tol;
    M.p_tol = j.o.p_tol;
    M.adiis_switch = j.o.adiis_switch;
    M.shift_floor = j.o.level_shift;
    M.sad_yy = j.sad_yy;
    M.mixed_switch = j.o.mixed_switch;
    M.mixed_sweeps = ctx->mixed_sweeps;
    M.dpath = j.dpath;
    M.rotate_below = ctx->rotate_below;
    M.eig_tol_exp = ctx->eig_tol_exp;
    M.eig_block = (ctx->eig_block_mask >> j.variant) & 1;
    M.init_p = j.p0.empty() ? 0 : 1;
    lists[j.variant].push_back(mi);
    atom0 += n;
    pair0 += (int64_t)n * (n - 1) / 2;
  }
  std::vector<int>& list = hl.list;
  for (int v = 0; v < SCF_VARIANTS; ++v) {
    P.list_off[v] = (int)list.size();
    // Largest molecules first so the longest SCFs start first.
    std::stable_sort(lists[v].begin(), lists[v].end(),
                     [&](int a, int b) { return P.jobs[a].nao > P.jobs[b].nao; });
    list.insert(list.end(), lists[v].begin(), lists[v].end());
  }
  P.list_off[SCF_VARIANTS] = (int)list.size();
  P.n_pairs = (int64_t)pairs.size();
  P.w_words = w_off;
  P.slab_words = mat;
  P.vec_words = vec;
  P.gath_words = gath;
  P.n_items = (int)items.size();
  P.n_dpairs = (int64_t)d_pairs.size();
  // HBM budget (pm7-rs memory::guard, src/memory.rs:95-118, but for device memory).
  size_t need = bytes_needed(P) + P.extra_bytes;
  size_t free_b = 0, total_b = 0;
  LANECHK(hipMemGetInfo(&free_b, &total_b));
  size_t budget = max_memory_mb > 0 ? (size_t)max_memory_mb * 1024 * 1024 : (size_t)(0.8 * (double)(free_b + L.held()));
  if (need > budget) {
    L.err = "needs " + std::to_string(need >> 20) + " MiB of device memory, budget " + std::to_string(budget >> 20) +
            " MiB (integrals + SCF working set)";
    P.need_mb = (double)need / (1024.0 * 1024.0);
    P.budget_mb = (double)budget / (1024.0 * 1024.0);
    return PM7_ERR_INSUFFICIENT_MEMORY;
  }
  auto up = [&](DevBuf& d, const void* src, size_t bytes) -> hipError_t {
    hipError_t e = d.ensure(bytes);
    if (e != hipSuccess) return e;
    if (bytes) return hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, L.st[0]);
    return hipSuccess;
  };
  LANECHK(up(L.d_mols, P.mols.data(), P.mols.size() * sizeof(MolDev)));
  LANECHK(up(L.d_list, list.data(), list.size() * sizeof(int)));
  LANECHK(up(L.d_atom_z, atom_z.data(), atom_z.size() * sizeof(int)));
  LANECHK(up(L.d_atom_ao, atom_ao.data(), atom_ao.size() * sizeof(int)));
  LANECHK(up(L.d_atom_pp, atom_pp.data(), atom_pp.size() * sizeof(int)));
  LANECHK(up(L.d_xyz, xyz.data(), xyz.size() * sizeof(double)));
  LANECHK(up(L.d_pairs, pairs.data(), pairs.size() * sizeof(PairDev)));
  LANECHK(up(L.d_items, items.data(), items.size() * sizeof(int2)));
  LANECHK(up(L.d_gath_off, gath_off.data(), gath_off.size() * sizeof(int64_t)));
  LANECHK(L.d_w.ensure((size_t)P.w_words * 8));
  LANECHK(L.d_slab.ensure((size_t)P.slab_words * 8));
  LANECHK(L.d_vecs.ensure((size_t)P.vec_words * 8));
  LANECHK(L.d_out.ensure((size_t)nm * sizeof(MolOut)));
  LANECHK(L.d_ecore.ensure((size_t)P.n_pairs * 8));
  LANECHK(L.d_ecore_mol.ensure((size_t)nm * 8));
  LANECHK(L.d_gath.ensure((size_t)P.gath_words * 8));
  LANECHK(up(L.d_num, num_pairs.data(), num_pairs.size() * sizeof(int64_t)));
  LANECHK(L.d_snum.ensure(num_pairs.size() * 5 * sizeof(double)));
  LANECHK(up(L.d_dpairs, d_pairs.data(), d_pairs.size() * sizeof(int64_t)));
  // Zero the slab: pad rows/columns of every matrix must be exact zeros.
  LANECHK(launch_zero(L.d_slab.as<double>(), P.slab_words, L.st[0]));
  // Seeded starting densities into the first density slot(s) (RHF: S_HD; UHF: U_PA, U_PB), rows
  // padded to the slab's stride. The staging copies live in P until the call ends.
  for (const Job& j : P.jobs) {
    if (j.p0.empty()) continue;
    const int nsrc = j.uhf ? 2 : 1;
    const int slots[2] = {j.uhf ? (int)U_PA : (int)S_HD, (int)U_PB};
    for (int k = 0; k < nsrc; ++k) {
      const std::vector<double>& src = k == 0 ? j.p0 : j.p1;
      P.staging.emplace_back((size_t)j.ld * j.ld, 0.0);
      std::vector<double>& pad = P.staging.back();
      for (int r = 0; r < j.nao; ++r) std::copy(src.begin() + (size_t)r * j.nao, src.begin() + (size_t)(r + 1) * j.nao, pad.begin() + (size_t)r * j.ld);
      double* dst = L.d_slab.as<double>() + j.mat0 + (int64_t)slots[k] * j.ld * j.ld;
      LANECHK(hipMemcpyAsync(dst, pad.data(), pad.size() * 8, hipMemcpyHostToDevice, L.st[0]));
    }
  }
  BatchDev& b = P.b;
  b.mols = L.d_mols.as<MolDev>();
  b.mol_list = L.d_list.as<int>();
  b.atom_z = L.d_atom_z.as<int>();
  b.atom_ao = L.d_atom_ao.as<int>();
  b.atom_pp = L.d_atom_pp.as<int>();
  b.xyz = L.d_xyz.as<double>();
  b.pairs = L.d_pairs.as<PairDev>();
  b.elem = ctx->d_elem.as<ElemDev>();
  b.pairpar = ctx->d_pairpar.as<double2>();
  b.hc_items = L.d_items.as<int2>();
  b.gath_off = L.d_gath_off.as<int64_t>();
  b.w = L.d_w.as<double>();
  b.slab = L.d_slab.as<double>();
  b.vecs = L.d_vecs.as<double>();
  b.out = L.d_out.as<MolOut>();
  b.e_core = L.d_ecore.as<double>();
  b.e_core_mol = L.d_ecore_mol.as<double>();
  b.gath = L.d_gath.as<double>();
  b.vpar = make_double4(ctx->params.vpar[1], ctx->params.vpar[2], ctx->params.vpar[3], ctx->params.vpar[4]);
  b.num_pairs = L.d_num.as<int64_t>();
  b.n_num = (int)num_pairs.size();
  b.gl = ctx->d_gl.as<double>();
  b.s_num = L.d_snum.as<double>();
  b.d_pairs = L.d_dpairs.as<int64_t>();
  b.n_dp = (int64_t)d_pairs.size();
  b.ocd = ctx->d_ocd.as<double>();
  LANECHK(hipEventRecord(L.ev[0], L.st[0]));
  LANECHK(launch_integrals(b, P.n_pairs, P.n_items, nm, L.st[0]));
  LANECHK(hipEventRecord(L.ev[1], L.st[0]));
  return PM7_OK;
}

void finish_result(pm7_ctx* ctx, const Job& j, const MolOut& R, double e_core, const double* eps_ab,
                   const double* gath, pm7_result_v2* out) {
  const int n = j.nao, nat = (int)j.z.size();
  std::memset(out, 0, sizeof *out);
  out->nao = n;
  out->n_atoms = nat;
  out->iterations = R.iterations;
  out->converged = R.converged;
  out->unrestricted = j.uhf;
  out->n_occ = j.na;
  out->n_occ_beta = j.uhf ? j.nb : -1;
  out->fp32_iterations = R.fp32_iters;
  out->err_vs_fp64_ev = R.check_de;
  out->err_vs_fp64_density = R.check_dp;
  out->jacobi_sweeps = R.sweeps;
  out->electronic_ev = R.e_el;
  out->core_ev = e_core;
  out->correction_kcal = j.corr_kcal;
  out->total_ev = R.e_el + e_core + j.corr_kcal * KCAL_TO_EV;
  double e_isol = 0.0, eheat = 0.0;
  for (int z : j.z) {
    e_isol += ctx->params.el[z].e_isol;
    eheat += ctx->params.el[z].eheat_ev;
  }
  out->heat_of_formation_kcal = (R.e_el + e_core - e_isol + eheat) * EV_TO_KCAL + j.corr_kcal;
  out->density_error = R.dp_err;
  for (int k = 0; k < 6; ++k) out->device_phase_ms[k] = (double)R.ticks[k] * 1.0e-5;  // 100 MHz clock
  for (int k = 0; k < 3; ++k) out->eig_phase_ms[k] = (double)R.ticks[PH_EIG_IN + k] * 1.0e-5;
  const size_t nn = (size_t)n * n;
  // One allocation per result, based at `density` (pm7_result_free releases it).
  const size_t words = 2 * nn + n + nat + (j.uhf ? 2 * nn + n : 0);
  double* block = (double*)std::malloc(words * sizeof(double));
  if (!block) {
    out->status = PM7_ERR_INSUFFICIENT_MEMORY;
    set_msg(out, "host allocation of the result arrays failed");
    return;
  }
  out->density = block;
  out->mo_coeff = block + nn;
  out->mo_energies = block + 2 * nn;
  out->charges = block + 2 * nn + n;
  std::memcpy(out->mo_energies, eps_ab, n * sizeof(double));
  std::memcpy(out->mo_coeff, gath + nn, nn * sizeof(double));
  if (!j.uhf) {
    std::memcpy(out->density, gath, nn * sizeof(double));
  } else {
    const double* pa = gath;
    const double* pb = gath + 2 * nn;
    out->spin_density = out->charges + nat;
    for (size_t k = 0; k < nn; ++k) {
      out->density[k] = pa[k] + pb[k];
      out->spin_density[k] = pa[k] - pb[k];
    }
    out->mo_energies_beta = out->spin_density + nn;
    out->mo_coeff_beta = out->mo_energies_beta + n;
    std::memcpy(out->mo_energies_beta, eps_ab + j.ld, n * sizeof(double));
    std::memcpy(out->mo_coeff_beta, gath + 3 * nn, nn * sizeof(double));
  }
  // Mulliken charges (src/scf.rs:777-787) and the dipole breakdown (src/dipole.rs:282-348).
  double net = 0.0;
  int ao = 0;
  std::vector<int> off(nat);
  for (int i = 0; i < nat; ++i) {
    const HostElement& e = ctx->params.el[j.z[i]];
    off[i] = ao;
    double pop = 0.0;
    for (int k = 0; k < e.n_orb; ++k) pop += out->density[(size_t)(ao + k) * n + ao + k];
    out->charges[i] = e.core_charge - pop;
    net += out->charges[i];
    ao += e.n_orb;
  }
  // Origin of the point-charge part (dipole_origin, src/dipole.rs:137-177): the coordinate origin
  // for a neutral molecule (|net| <= 0.5) whatever the option, else the centre of mass (default),
  // the coordinate origin or the centre of core charge.
  double org[3] = {0, 0, 0};
  if (std::fabs(net) > 0.5 && j.o.dipole_origin != PM7_DIPOLE_COORDINATES) {
    double tot = 0.0;
    for (int i = 0; i < nat; ++i) {
      const double m = j.o.dipole_origin == PM7_DIPOLE_CENTRE_OF_CHARGE ? ctx->params.el[j.z[i]].core_charge
                                                                         : (j.z[i] <= 86 ? MASS[j.z[i]] : 0.0);
      for (int k = 0; k < 3; ++k) org[k] += j.xyz[3 * i + k] * m;
      tot += m;
    }
    if (tot > 0)
      for (int k = 0; k < 3; ++k) org[k] /= tot;
  }
  const double D = 2.541746473;
  // p-d hybrid term of a d-bearing atom (pm7-rs src/dipole.rs:332-340); orbitals
  // s px py pz dx2-y2 dxz dz2 dyz dxy.
  double pd[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < nat; ++i) {
    const HostElement& e = ctx->params.el[j.z[i]];
    if (e.n_orb != 9) continue;
    auto P = [&](int a, int b) { return out->density[(size_t)(off[i] + a) * n + off[i] + b]; };
    const double inv3 = 1.0 / std::sqrt(3.0), hyb = -2.0 * e.ddp[5];
    pd[0] += hyb * (P(5, 3) + P(4, 1) + P(8, 2) - inv3 * P(6, 1));
    pd[1] += hyb * (P(7, 3) - P(4, 2) + P(8, 1) - inv3 * P(6, 2));
    pd[2] += hyb * (P(5, 1) + P(7, 2) + 2.0 * inv3 * P(6, 3));
  }
  for (int k = 0; k < 3; ++k) {
    double pc = 0.0, sp = 0.0;
    for (int i = 0; i < nat; ++i) {
      pc += (j.xyz[3 * i + k] - org[k]) * out->charges[i];
      const HostElement& e = ctx->params.el[j.z[i]];
      if (e.n_orb >= 4) sp += -2.0 * e.dd * out->density[(size_t)off[i] * n + off[i] + 1 + k];
    }
    out->dipole_point_charge_debye[k] = pc * D;
    out->dipole_sp_hybrid_debye[k] = sp * D;
    out->dipole_pd_hybrid_debye[k] = pd[k] * D;
    out->dipole_debye[k] = pc * D + sp * D + pd[k] * D;
  }
  out->dipole_magnitude = std::sqrt(out->dipole_debye[0] * out->dipole_debye[0] +
                                    out->dipole_debye[1] * out->dipole_debye[1] +
                                    out->dipole_debye[2] * out->dipole_debye[2]);
  out->homo_ev = j.na >= 1 ? eps_ab[j.na - 1] : nan_();
  out->lumo_ev = j.na < n ? eps_ab[j.na] : nan_();
  out->homo_ev_beta = nan_();
  out->lumo_ev_beta = nan_();
  if (j.uhf) {
    out->homo_ev_beta = j.nb >= 1 ? eps_ab[j.ld + j.nb - 1] : nan_();
    out->lumo_ev_beta = j.nb < n ? eps_ab[j.ld + j.nb] : nan_();
  }
  if (R.converged) {
    out->status = PM7_OK;
  } else {
    out->status = PM7_ERR_SCF_NOT_CONVERGED;
    set_msg(out, "SCF not converged after " + std::to_string(R.iterations) + " iterations (error " +
                     std::to_string(R.dp_err) + ")");
  }
}

// Gauss-Legendre nodes and weights on [-1, 1] for 48 and 40 points (x48, w48, x40, w40): roots of
// P_n by Newton's method from the Chebyshev guess cos(pi (i + 3/4) / (n + 1/2)), weights
// 2 / ((1 - x^2) P_n'(x)^2) (the rule of pm7-rs src/overlap_numeric.rs:137-175).
void gauss_legendre_table(std::vector<double>& out) {
  out.assign(GL_WORDS, 0.0);
  auto legendre = [](int n, double x, double& dp) {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1.0);
    return p1;
  };
  int base = 0;
  for (int n : {48, 40}) {
    double* x = out.data() + base;
    double* w = x + n;
    for (int i = 0; i < (n + 1) / 2; ++i) {
      double t = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
      for (int it = 0; it < 100; ++it) {
        const double dx = -legendre(n, t, dp) / dp;
        t += dx;
        if (std::fabs(dx) < 1.0e-15) break;
      }
      (void)legendre(n, t, dp);
      const double wt = 2.0 / ((1.0 - t * t) * dp * dp);
      x[i] = -t;
      x[n - 1 - i] = t;
      w[i] = wt;
      w[n - 1 - i] = wt;
    }
    base += 2 * n;
  }
}

// PM7, PM7- and PM7-HH share the PM7 tables; PM7-TS has its own (src/params.rs:158-170).
pm7_status ensure_params(pm7_ctx* ctx, int method) {
  const int tables = method == PM7_METHOD_PM7_TS ? PM7_METHOD_PM7_TS : PM7_METHOD_PM7;
  if (ctx->tables == tables) {
    ctx->method = method;
    return PM7_OK;
  }
  std::string err;
  if (!load_params(ctx->data_dir, tables, ctx->params, err)) return fail(ctx, PM7_ERR_MISSING_PARAMETER, err);
  to_device_table(ctx->params, ctx->elem_host);
  HIPCHK(ctx->d_elem.ensure(ctx->elem_host.size() * sizeof(ElemDev)));
  HIPCHK(hipMemcpy(ctx->d_elem.p, ctx->elem_host.data(), ctx->elem_host.size() * sizeof(ElemDev), hipMemcpyHostToDevice));
  std::vector<double2> pp(MAX_Z * MAX_Z);
  for (int i = 0; i < MAX_Z * MAX_Z; ++i) pp[i] = make_double2(ctx->params.pair[i].first, ctx->params.pair[i].second);
  HIPCHK(ctx->d_pairpar.ensure(pp.size() * sizeof(double2)));
  HIPCHK(hipMemcpy(ctx->d_pairpar.p, pp.data(), pp.size() * sizeof(double2), hipMemcpyHostToDevice));
  {
    std::vector<double> ocd((size_t)MAX_Z * OC_D_WORDS, 0.0);
    for (int z = 0; z < MAX_Z; ++z)
      if (ctx->params.el[z].oc_d.size() == (size_t)OC_D_WORDS)
        std::copy(ctx->params.el[z].oc_d.begin(), ctx->params.el[z].oc_d.end(), ocd.begin() + (size_t)z * OC_D_WORDS);
    HIPCHK(ctx->d_ocd.ensure(ocd.size() * sizeof(double)));
    HIPCHK(hipMemcpy(ctx->d_ocd.p, ocd.data(), ocd.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (!ctx->d_gl.p) {
    std::vector<double> gl;
    gauss_legendre_table(gl);
    HIPCHK(ctx->d_gl.ensure(gl.size() * sizeof(double)));
    HIPCHK(hipMemcpy(ctx->d_gl.p, gl.data(), gl.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  ctx->tables = tables;
  ctx->method = method;
  return PM7_OK;
}

// Geometric corrections of every job (host; overlaps the device work).
void host_corrections(std::vector<Job>& jobs) {
  auto work = [&](int t, int nt) {
    for (size_t k = t; k < jobs.size(); k += nt) {
      Job& j = jobs[k];
      j.corr_kcal = 0.0;
      if (j.o.method != PM7_METHOD_PM7_MINUS) {  // correction_energy, src/scf.rs:907-932
        j.corr_kcal += dispersion_kcal(j.z, j.xyz, j.o.smooth_dispersion != 0);
        j.corr_kcal += hbond_kcal(j.z, j.xyz);
        if (j.o.method == PM7_METHOD_PM7_HH) j.corr_kcal += hh_repulsion_kcal(j.z, j.xyz);
        j.corr_kcal += c_triple_kcal(j.z, j.xyz);
        j.corr_kcal += si_o_h_kcal(j.z, j.xyz);
      }
    }
  };
  const int nt = jobs.size() >= 32 ? std::min(8, std::max(1, (int)std::thread::hardware_concurrency() / 2)) : 1;
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t, nt);
  work(0, nt);
  for (auto& th : pool) th.join();
}

// SCF megakernels (one stream per size class, all after the integrals), then the output gather
// on stream 0; L.ev[5] marks the end of the batch.
pm7_status launch_batch(Lane& L, const Prepared& P, int* count) {
  for (int v = 1; v < SCF_VARIANTS && !L.single_stream; ++v) LANECHK(hipStreamWaitEvent(L.st[v], L.ev[1], 0));
  for (int v = 0; v < SCF_VARIANTS; ++v) {
    count[v] = P.list_off[v + 1] - P.list_off[v];
    if (v == LARGE_VARIANT) continue;  // run_large
    LANECHK(hipEventRecord(L.ev_var[2 * v], L.st[v]));
    LANECHK(launch_scf(v, P.b, P.list_off[v], count[v], L.st[v]));
    LANECHK(hipEventRecord(L.ev_var[2 * v + 1], L.st[v]));
  }
  return PM7_OK;
}

// The large-class molecules, one after another on the large class's stream (each fills the whole
// device while the megakernels of the other classes run beside it); their MolOut records are
// written where the megakernels would have written them.
pm7_status run_large(Lane& L, const Prepared& P) {
  const int off = P.list_off[LARGE_VARIANT], cnt = P.list_off[LARGE_VARIANT + 1] - off;
  if (cnt == 0) return PM7_OK;
  // The class's timing events bracket all of its work, one molecule at a time and grouped.
  hipStream_t ls = L.st[LARGE_VARIANT];
  LANECHK(hipEventRecord(L.ev_var[2 * LARGE_VARIANT], ls));
  std::vector<std::pair<int, int>> order;  // largest first
  for (int mi = 0; mi < (int)P.jobs.size(); ++mi)
    if (P.jobs[mi].variant == LARGE_VARIANT) order.push_back({-P.jobs[mi].nao, mi});
  std::stable_sort(order.begin(), order.end());
  // Molecules up to LARGE_GROUP_MAX_NAO run in groups, each SCF in its own thread and stream, their
  // full eigensolves batched in lockstep; larger ones one after another.
  size_t single = 0;
  while (single < order.size() && -order[single].first > LARGE_GROUP_MAX_NAO) ++single;
  if (order.size() - single < 2) single = order.size();
  if (single > 0) {
    if (!L.large) {
      L.large = large_work_create();
      if (!L.large) {
        L.err = "rocBLAS handle creation failed";
        return PM7_ERR_DEVICE;
      }
    }
    hipStream_t s = ls;
    for (size_t k = 0; k < single; ++k) {
      const int mi = order[k].second;
      MolOut out{};
      if (large_scf(L.large, P.b, P.mols[mi], mi, s, out, L.err) != hipSuccess) return PM7_ERR_DEVICE;
      LANECHK(hipMemcpyAsync(P.b.out + mi, &out, sizeof out, hipMemcpyHostToDevice, s));
      LANECHK(hipStreamSynchronize(s));
    }
  }
  const int group = large_group_size();
  for (size_t g0 = single; g0 < order.size(); g0 += group) {
    const int gn = (int)std::min<size_t>(group, order.size() - g0);
    while ((int)L.group_work.size() < gn) {
      LargeWork* w = large_work_create();
      hipStream_t st = nullptr;
      if (!w || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        large_work_destroy(w);
        L.err = "large-system work set creation failed";
        return PM7_ERR_DEVICE;
      }
      L.group_work.push_back(w);
      L.group_st.push_back(st);
    }
    int nmax = 0;
    for (int k = 0; k < gn; ++k) nmax = std::max(nmax, P.mols[order[g0 + k].second].ld);
    EigBatcher* eb = eig_batcher_create(gn, nmax, L.err);
    if (!eb) return PM7_ERR_DEVICE;
    std::vector<std::string> errs(gn);
    std::vector<int> bad(gn, 0);
    std::vector<std::thread> th;
    for (int k = 0; k < gn; ++k) {
      if (hipStreamWaitEvent(L.group_st[k], L.ev[1], 0) != hipSuccess) bad[k] = 1;  // integrals done
      th.emplace_back([&, k] {
        (void)hipSetDevice(L.device);
        const int mi = order[g0 + k].second;
        MolOut out{};
        hipStream_t st = L.group_st[k];
        hipError_t r = large_scf(L.group_work[k], P.b, P.mols[mi], mi, st, out, errs[k], eb);
        if (r == hipSuccess) r = hipMemcpyAsync(P.b.out + mi, &out, sizeof out, hipMemcpyHostToDevice, st);
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        if (r != hipSuccess) bad[k] = 1;
      });
    }
    for (auto& t : th) t.join();
    eig_batcher_destroy(eb);
    for (int k = 0; k < gn; ++k)
      if (bad[k]) {
        L.err = errs[k].empty() ? std::string("large-system SCF failed") : errs[k];
        return PM7_ERR_DEVICE;
      }
  }
  LANECHK(hipEventRecord(L.ev_var[2 * LARGE_VARIANT + 1], ls));  // every group thread synchronised its stream
  return PM7_OK;
}

// Join the class streams and gather the outputs on stream 0; L.ev[5] marks the end of the batch.
pm7_status finish_batch(Lane& L, const Prepared& P) {
  const int nm = (int)P.jobs.size();
  for (int v = 1; v < SCF_VARIANTS && !L.single_stream; ++v) {
    LANECHK(hipEventRecord(L.ev[6 + v], L.st[v]));
    LANECHK(hipStreamWaitEvent(L.st[0], L.ev[6 + v], 0));
  }
  LANECHK(hipEventRecord(L.ev[2], L.st[0]));
  LANECHK(launch_gather(P.b, nm, L.st[0]));
  LANECHK(hipEventRecord(L.ev[5], L.st[0]));
  return PM7_OK;
}

// Copy a finished batch's outputs to the host and fill results[j.idx] for every job.
pm7_status collect_batch(pm7_ctx* ctx, Lane& L, const Prepared& P, pm7_result_v2* results) {
  const int nm = (int)P.jobs.size();
  // One pinned staging area: MolOut records, core energies, eigenvalues, packed matrices.
  const size_t o_out = 0;
  const size_t o_ecore = o_out + ((size_t)nm * sizeof(MolOut) + 255) / 256 * 256;
  const size_t o_vecs = o_ecore + ((size_t)nm * 8 + 255) / 256 * 256;
  const size_t o_gath = o_vecs + ((size_t)P.vec_words * 8 + 255) / 256 * 256;
  const size_t total = o_gath + (size_t)P.gath_words * 8;
  LANECHK(L.h_out.ensure(total));
  char* h = static_cast<char*>(L.h_out.p);
  const MolOut* outs = reinterpret_cast<const MolOut*>(h + o_out);
  const double* ecore = reinterpret_cast<const double*>(h + o_ecore);
  const double* vecs = reinterpret_cast<const double*>(h + o_vecs);
  const double* gath = reinterpret_cast<const double*>(h + o_gath);
  LANECHK(hipMemcpyAsync(h + o_out, P.b.out, nm * sizeof(MolOut), hipMemcpyDeviceToHost, L.st[0]));
  LANECHK(hipMemcpyAsync(h + o_ecore, P.b.e_core_mol, nm * sizeof(double), hipMemcpyDeviceToHost, L.st[0]));
  LANECHK(hipMemcpyAsync(h + o_vecs, P.b.vecs, P.vec_words * sizeof(double), hipMemcpyDeviceToHost, L.st[0]));
  LANECHK(hipMemcpyAsync(h + o_gath, P.b.gath, P.gath_words * sizeof(double), hipMemcpyDeviceToHost, L.st[0]));
  LANECHK(hipStreamSynchronize(L.st[0]));
  // Results are independent: a few host threads assemble them (each writes only its own).
  auto work = [&](int t, int nt) {
    for (int mi = t; mi < nm; mi += nt) {
      const Job& j = P.jobs[mi];
      finish_result(ctx, j, outs[mi], ecore[mi], vecs + j.vec0, gath + j.gath0, &results[j.idx]);
    }
  };
  const int nt = nm >= 32 ? std::min(8, std::max(1, (int)std::thread::hardware_concurrency() / 2)) : 1;
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t, nt);
  work(0, nt);
  for (auto& th : pool) th.join();
  return PM7_OK;
}

// Where pm7_gradient_batch wants each molecule's gradient: flat output, offset per input index.
struct GradReq {
  double* out = nullptr;
  std::vector<int64_t> off;  // by Job::idx, in doubles
};

// Gradient of the geometric corrections host_corrections adds, kcal/mol/Bohr into g[3 n]:
// dispersion analytically; the EH+, C#C and Si-O-H terms, each only when it is non-zero at this
// geometry (zero means the molecule has no such contact or motif), by differences with
// h = 1e-5 Bohr. These terms select contacts with hard tests (distance and angle cutoffs, first
// match), so a geometry can sit exactly where a test flips and the energy jumps. A jump J shows
// as one one-sided quotient ~J/h (above 10 kcal/mol/Bohr and ten times the other): the other side
// keeps the selection made at this geometry, and its derivative is the one reported (as an
// analytic derivative of the selected terms would be); otherwise the central difference (at a
// kink, the mean of the two one-sided slopes).
void corrections_grad_kcal(const Job& j, double* g) {
  dispersion_grad_kcal(j.z, j.xyz, g, j.o.smooth_dispersion != 0);
  if (j.o.method == PM7_METHOD_PM7_HH) hh_repulsion_grad_kcal(j.z, j.xyz, g);
  using Term = double (*)(const std::vector<int>&, const std::vector<double>&);
  const Term terms[3] = {hbond_kcal, c_triple_kcal, si_o_h_kcal};
  std::vector<double> x = j.xyz;
  constexpr double h = 1.0e-5;
  for (Term f : terms) {
    const double e0 = f(j.z, j.xyz);
    if (e0 == 0.0) continue;
    for (size_t c = 0; c < x.size(); ++c) {
      x[c] = j.xyz[c] + h;
      const double ep = f(j.z, x);
      x[c] = j.xyz[c] - h;
      const double em = f(j.z, x);
      x[c] = j.xyz[c];
      const double df = (ep - e0) / h, db = (e0 - em) / h;
      const double big = std::max(std::fabs(df), std::fabs(db)), small = std::min(std::fabs(df), std::fabs(db));
      const bool jump = big > 10.0 && small < 0.1 * big;
      g[c] += jump ? (std::fabs(df) < std::fabs(db) ? df : db) : 0.5 * (df + db);
    }
  }
}

// Device gradient of every job of a finished batch (enqueued on stream 0 after the SCF and the
// gather); collect_gradient adds the corrections' gradient on the host and fills the caller's
// array. A molecule with a pair that has no device gradient (its SCF left no density) gets NaN and
// PM7_ERR_UNSUPPORTED.
pm7_status launch_gradient_batch(Lane& L, const Prepared& P, int* n_items_out) {
  std::vector<int2>& items = L.h_gitems;  // the upload's source outlives the call (see HostLayout)
  items.clear();
  int n_atoms = 0;
  for (size_t mi = 0; mi < P.jobs.size(); ++mi) {
    const Job& j = P.jobs[mi];
    for (int k = 0; k < (int)j.z.size(); ++k) items.push_back(make_int2((int)mi, j.atom0 + k));
    n_atoms = std::max(n_atoms, j.atom0 + (int)j.z.size());
  }
  *n_items_out = (int)items.size();
  LANECHK(L.d_gitems.ensure(std::max<size_t>(1, items.size()) * sizeof(int2)));
  LANECHK(L.d_gpair.ensure(std::max<size_t>(1, (size_t)P.n_pairs * 3) * sizeof(double)));
  LANECHK(L.d_gatom.ensure(std::max<size_t>(1, (size_t)n_atoms * 3) * sizeof(double)));
  if (!items.empty())
    LANECHK(hipMemcpyAsync(L.d_gitems.p, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, L.st[0]));
  LANECHK(L.d_snumg.ensure(std::max<size_t>(1, (size_t)P.b.n_num * 60) * sizeof(double)));
  LANECHK(launch_gradient(P.b, P.n_pairs, static_cast<const int2*>(L.d_gitems.p), (int)items.size(),
                          static_cast<double*>(L.d_gpair.p), static_cast<double*>(L.d_gatom.p),
                          static_cast<double*>(L.d_snumg.p), L.st[0]));
  return PM7_OK;
}

pm7_status collect_gradient(Lane& L, const Prepared& P, const GradReq& g, pm7_result_v2* results) {
  int n_atoms = 0;
  for (const Job& j : P.jobs) n_atoms = std::max(n_atoms, j.atom0 + (int)j.z.size());
  std::vector<double> ga((size_t)n_atoms * 3);
  if (n_atoms > 0) {
    LANECHK(hipMemcpyAsync(ga.data(), L.d_gatom.p, ga.size() * sizeof(double), hipMemcpyDeviceToHost, L.st[0]));
    LANECHK(hipStreamSynchronize(L.st[0]));
  }
  auto work = [&](int t, int nt) {
    for (size_t mi = t; mi < P.jobs.size(); mi += nt) {
      const Job& j = P.jobs[mi];
      const int na = (int)j.z.size();
      double* out = g.out + g.off[j.idx];
      bool ok = true;
      for (int k = 0; k < 3 * na; ++k) {
        out[k] = ga[(size_t)3 * j.atom0 + k];
        ok = ok && std::isfinite(out[k]);
      }
      pm7_result_v2& r = results[j.idx];
      if (!ok) {
        for (int k = 0; k < 3 * na; ++k) out[k] = std::nan("");
        if (r.status == PM7_OK || r.status == PM7_ERR_SCF_NOT_CONVERGED) {
          r.status = PM7_ERR_UNSUPPORTED;
          set_msg(&r, "gradient: no converged density to differentiate at");
        }
        continue;
      }
      if (j.o.method == PM7_METHOD_PM7_MINUS) continue;
      std::vector<double> gc((size_t)3 * na, 0.0);
      corrections_grad_kcal(j, gc.data());
      for (int c = 0; c < 3 * na; ++c) out[c] += gc[c] / EV_TO_KCAL;
    }
  };
  const int nt = P.jobs.size() >= 8 ? std::min(16, std::max(1, (int)std::thread::hardware_concurrency())) : 1;
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t, nt);
  work(0, nt);
  for (auto& th : pool) th.join();
  return PM7_OK;
}

pm7_status run_jobs(pm7_ctx* ctx, std::vector<Job>& jobs, pm7_result_v2* results, int64_t max_memory_mb,
                    const GradReq* grad = nullptr) {
  Lane& L = ctx->lane;
  auto t0 = std::chrono::steady_clock::now();
  Prepared P;
  P.jobs = std::move(jobs);
  pm7_status st = prepare(ctx, L, P, max_memory_mb);
  const auto t_prep = std::chrono::steady_clock::now();
  if (st != PM7_OK) {
    ctx->err = L.err;
    for (auto& j : P.jobs) {
      results[j.idx].status = st;
      results[j.idx].needed_mb = P.need_mb;
      results[j.idx].budget_mb = P.budget_mb;
      set_msg(&results[j.idx], L.err);
    }
    return st;
  }
  int count[SCF_VARIANTS];
  st = launch_batch(L, P, count);
  if (st == PM7_OK) st = run_large(L, P);
  if (st == PM7_OK) st = finish_batch(L, P);
  int n_gitems = 0;
  if (st == PM7_OK && grad) st = launch_gradient_batch(L, P, &n_gitems);
  if (st != PM7_OK) return fail(ctx, st, L.err);
  host_corrections(P.jobs);
  const auto t_launched = std::chrono::steady_clock::now();
  HIPCHK(hipEventSynchronize(L.ev[5]));
  const auto t_synced = std::chrono::steady_clock::now();
  st = collect_batch(ctx, L, P, results);
  if (st == PM7_OK && grad) st = collect_gradient(L, P, *grad, results);
  if (st != PM7_OK) return fail(ctx, st, L.err);
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, L.ev[0], L.ev[1]) == hipSuccess) ctx->timings[0] = ms;
  if (hipEventElapsedTime(&ms, L.ev[1], L.ev[2]) == hipSuccess) ctx->timings[1] = ms;
  if (hipEventElapsedTime(&ms, L.ev[2], L.ev[5]) == hipSuccess) ctx->timings[2] = ms;
  for (int v = 0; v < SCF_VARIANTS; ++v) {
    ctx->timings[4 + v] = 0.0;
    if (count[v] > 0 && hipEventElapsedTime(&ms, L.ev_var[2 * v], L.ev_var[2 * v + 1]) == hipSuccess)
      ctx->timings[4 + v] = ms;
  }
  // A timing query that failed (an event of a class that did not run) must not stay behind as
  // this thread's last error, where the next call's first launch check would report it.
  (void)hipGetLastError();
  const auto t_end = std::chrono::steady_clock::now();
  using msd = std::chrono::duration<double, std::milli>;
  ctx->timings[3] = msd(t_end - t0).count();
  ctx->timings[12] = ctx->timings[13] = 0.0;
  ctx->timings[8] = msd(t_prep - t0).count();
  ctx->timings[9] = msd(t_launched - t_prep).count();
  ctx->timings[10] = msd(t_synced - t_launched).count();
  ctx->timings[11] = msd(t_end - t_synced).count();
  return PM7_OK;
}

// ---- asynchronous batches -----------------------------------------------------------------

// One batch in flight: its lane (streams and device buffers), its layout and the caller's result
// array, which pm7_batch_wait (or a submit that needs the lane) fills.
struct AsyncBatch {
  int64_t ticket = 0;
  Lane* lane = nullptr;
  Prepared P;
  pm7_result_v2* results = nullptr;
  int count[SCF_VARIANTS] = {};
  std::chrono::steady_clock::time_point t0;
};

// Wait for batch b, fill its results, record its timings and free its lane. Lock held.
pm7_status async_collect(pm7_ctx* ctx, AsyncBatch* b) {
  Lane& L = *b->lane;
  pm7_status st = PM7_OK;
  if (hipEventSynchronize(L.ev[5]) != hipSuccess) st = fail(ctx, PM7_ERR_DEVICE, "batch wait failed");
  if (st == PM7_OK) {
    st = collect_batch(ctx, L, b->P, b->results);
    if (st != PM7_OK) ctx->err = L.err;
  }
  if (st == PM7_OK) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, L.ev[0], L.ev[1]) == hipSuccess) ctx->timings[0] = ms;
    if (hipEventElapsedTime(&ms, L.ev[1], L.ev[2]) == hipSuccess) ctx->timings[1] = ms;
    if (hipEventElapsedTime(&ms, L.ev[2], L.ev[5]) == hipSuccess) ctx->timings[2] = ms;
    for (int v = 0; v < SCF_VARIANTS; ++v) {
      ctx->timings[4 + v] = 0.0;
      if (b->count[v] > 0 && hipEventElapsedTime(&ms, L.ev_var[2 * v], L.ev_var[2 * v + 1]) == hipSuccess)
        ctx->timings[4 + v] = ms;
    }
    ctx->timings[12] = ctx->timings[13] = 0.0;
    if (ctx->epoch) {
      if (hipEventElapsedTime(&ms, ctx->epoch, L.ev[0]) == hipSuccess) ctx->timings[12] = ms;
      if (hipEventElapsedTime(&ms, ctx->epoch, L.ev[5]) == hipSuccess) ctx->timings[13] = ms;
    }
    (void)hipGetLastError();
    ctx->timings[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b->t0).count();
  }
  if (st != PM7_OK)
    for (const Job& j : b->P.jobs) {
      b->results[j.idx].status = st;
      set_msg(&b->results[j.idx], ctx->err);
    }
  return st;
}

// Collect and retire the oldest batch in flight; its status is kept for its own wait.
void async_retire_oldest(pm7_ctx* ctx) {
  AsyncBatch* b = ctx->inflight.front();
  ctx->inflight.pop_front();
  ctx->async_done[b->ticket] = async_collect(ctx, b);
  delete b;
}

// ---- request server (continuous batching) ------------------------------------------------

struct SrvReq {
  Job job;
  bool valid = false;
  int state = 0;  // 0 queued, 1 on the device, 2 finished, 3 claimed
  pm7_result_v2 res{};
  std::chrono::steady_clock::time_point t_submit, t_done;
};

// One batch in flight on its own lane.
// One batch in flight on its own lane. The batch runs in its own thread (upload, integrals, the
// megakernels, the grid-wide large-system SCFs, which block on their host control loop, and the
// result assembly), so the scheduler keeps admitting work to the other lanes meanwhile.
struct Flight {
  Lane lane;
  Prepared P;
  std::vector<int64_t> ids;
  size_t bytes = 0;
  bool busy = false;
  std::thread thr;
  std::atomic<int> done{0};  // 1: results (or the error) are ready for the scheduler
  pm7_status st = PM7_OK;
  std::vector<pm7_result_v2> res;
};

}  // namespace

struct pm7_server {
  pm7_ctx* ctx = nullptr;
  pm7_server_config_v1 cfg{};
  std::vector<std::unique_ptr<Flight>> flights;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> queue;     // admitted in arrival order
  std::deque<int64_t> finished;  // finished and not yet claimed, in completion order
  std::unordered_map<int64_t, SrvReq> reqs;
  int64_t next_id = 0;
  size_t budget = 0, inflight = 0;
  bool stopping = false;
  std::thread worker;
};

namespace {

void srv_finish(pm7_server* s, int64_t id, std::chrono::steady_clock::time_point now) {
  SrvReq& r = s->reqs[id];
  r.state = 2;
  r.t_done = now;
  s->finished.push_back(id);
}

// The scheduler: retire finished batches, admit queued requests into free lanes while the
// in-flight device memory stays within the budget, and repeat. Batches on different lanes run
// concurrently, so requests that arrive while others are on the device start without waiting
// for them to finish.
void srv_loop(pm7_server* s) {
  pm7_ctx* ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool busy = false;
    for (auto& fp : s->flights) {
      Flight& f = *fp;
      if (!f.busy) continue;
      if (!f.done.load(std::memory_order_acquire)) {
        busy = true;
        continue;
      }
      f.thr.join();
      const auto now = std::chrono::steady_clock::now();
      for (size_t k = 0; k < f.ids.size(); ++k) {
        SrvReq& r = s->reqs[f.ids[k]];
        if (f.st == PM7_OK) {
          r.res = f.res[k];
        } else {
          std::memset(&r.res, 0, sizeof r.res);
          r.res.status = f.st;
          r.res.n_atoms = (int32_t)r.job.z.size();
          set_msg(&r.res, f.lane.err.empty() ? std::string("device error") : f.lane.err);
        }
        srv_finish(s, f.ids[k], now);
      }
      f.busy = false;
      f.done.store(0);
      f.ids.clear();
      s->inflight -= f.bytes;
      s->cv_done.notify_all();
    }
    for (auto& fp : s->flights) {
      Flight& f = *fp;
      if (f.busy || s->queue.empty()) continue;
      // The oldest request and the queued requests of its size class, in arrival order: one
      // class per batch keeps a batch to a single kernel, whose time is that class's.
      std::vector<int64_t> ids;
      size_t bytes = 0;
      // Large-system groups hold a lane for many small-molecule batch times: one lane is kept
      // for the other classes, so small requests behind a run of large ones still find a lane.
      int large_flights = 0;
      for (auto& gp : s->flights)
        if (gp->busy && !gp->P.jobs.empty() && gp->P.jobs.front().variant == LARGE_VARIANT) ++large_flights;
      const bool large_ok = large_flights < std::max(1, (int)s->flights.size() - 1);
      int cls = -1;
      for (int64_t id : s->queue) {
        const int v = s->reqs[id].job.variant;
        if (v != LARGE_VARIANT || large_ok) {
          cls = v;
          break;
        }
      }
      if (cls < 0) break;  // only large requests queued and their lanes are full
      // Large-system SCFs of one batch run as a lockstep group (batched eigensolves), at most
      // large_group_size() of them, so further ones start on the other lanes.
      const int cap = cls == LARGE_VARIANT ? std::min(large_group_size(), s->cfg.max_batch) : s->cfg.max_batch;
      for (auto it = s->queue.begin(); it != s->queue.end() && (int)ids.size() < cap;) {
        SrvReq& r = s->reqs[*it];
        if (r.job.variant != cls) {
          ++it;
          continue;
        }
        const size_t b = job_bytes(ctx, r.job);
        if (s->inflight + bytes + b > s->budget && (s->inflight + bytes) > 0) break;  // wait for memory
        ids.push_back(*it);
        it = s->queue.erase(it);
        bytes += b;
      }
      if (ids.empty()) break;
      f.P = Prepared();
      for (size_t k = 0; k < ids.size(); ++k) {
        SrvReq& r = s->reqs[ids[k]];
        r.state = 1;
        f.P.jobs.push_back(r.job);
        f.P.jobs.back().idx = (int)k;
      }
      f.ids = ids;
      f.bytes = bytes;
      s->inflight += bytes;
      f.busy = true;
      f.done.store(0);
      Flight* fl = &f;
      f.thr = std::thread([ctx, s, fl] {
        (void)hipSetDevice(ctx->device);
        Flight& g = *fl;
        int count[SCF_VARIANTS];
        pm7_status st = prepare(ctx, g.lane, g.P, 0);
        if (st == PM7_OK) st = launch_batch(g.lane, g.P, count);
        if (st == PM7_OK) st = run_large(g.lane, g.P);
        if (st == PM7_OK) st = finish_batch(g.lane, g.P);
        if (st == PM7_OK) host_corrections(g.P.jobs);
        g.res.assign(g.P.jobs.size(), pm7_result_v2{});
        for (auto& r : g.res) std::memset(&r, 0, sizeof r);
        if (st == PM7_OK)
          st = hipEventSynchronize(g.lane.ev[5]) == hipSuccess ? collect_batch(ctx, g.lane, g.P, g.res.data())
                                                                : PM7_ERR_DEVICE;
        g.st = st;
        {
          std::lock_guard<std::mutex> wl(s->mu);
          g.done.store(1, std::memory_order_release);
        }
        s->cv_work.notify_all();
      });
      busy = true;
    }
    if (s->stopping && s->queue.empty() && !busy) break;
    if (busy) s->cv_work.wait_for(lk, std::chrono::microseconds(50));
    else s->cv_work.wait(lk, [&] { return s->stopping || !s->queue.empty(); });
  }
}

}  // namespace

extern "C" {

void pm7_server_default_config(pm7_server_config_v1* cfg) {
  if (!cfg) return;
  cfg->lanes = 4;
  cfg->max_batch = 256;
  cfg->max_memory_mb = 0;
}

pm7_status pm7_server_start(pm7_ctx* ctx, const pm7_server_config_v1* cfg, pm7_server** out) {
  if (!ctx || !out) return PM7_ERR_INVALID_INPUT;
  *out = nullptr;
  pm7_server_config_v1 c;
  pm7_server_default_config(&c);
  if (cfg) c = *cfg;
  if (c.lanes < 1 || c.lanes > 16 || c.max_batch < 1) return fail(ctx, PM7_ERR_INVALID_INPUT, "server: lanes 1..16, max_batch >= 1");
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  auto s = std::make_unique<pm7_server>();
  s->ctx = ctx;
  s->cfg = c;
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  s->budget = c.max_memory_mb > 0 ? (size_t)c.max_memory_mb << 20 : (size_t)(0.8 * (double)free_b);
  for (int k = 0; k < c.lanes; ++k) {
    auto f = std::make_unique<Flight>();
    if (!f->lane.create(true)) {
      f->lane.destroy();
      for (auto& g : s->flights) g->lane.destroy();
      return fail(ctx, PM7_ERR_DEVICE, "server: stream or event creation failed");
    }
    s->flights.push_back(std::move(f));
  }
  pm7_server* raw = s.release();
  raw->worker = std::thread(srv_loop, raw);
  *out = raw;
  return PM7_OK;
}

pm7_status pm7_server_submit(pm7_server* s, const pm7_molecule_v1* mol, const pm7_options_v2* opts, int64_t* request_id) {
  if (!s || !mol || !request_id) return PM7_ERR_INVALID_INPUT;
  pm7_options_v2 def;
  pm7_default_options(&def);
  SrvReq r;
  r.t_submit = std::chrono::steady_clock::now();
  std::string msg;
  pm7_status st = validate(s->ctx, *mol, opts ? *opts : def, r.job, msg);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) return PM7_ERR_INVALID_INPUT;
  const int64_t id = s->next_id++;
  *request_id = id;
  if (st == PM7_OK && job_bytes(s->ctx, r.job) > s->budget) {
    st = PM7_ERR_INSUFFICIENT_MEMORY;
    msg = "request needs " + std::to_string(job_bytes(s->ctx, r.job) >> 20) + " MiB, server budget " +
          std::to_string(s->budget >> 20) + " MiB";
    r.res.needed_mb = (double)job_bytes(s->ctx, r.job) / (1024.0 * 1024.0);
    r.res.budget_mb = (double)s->budget / (1024.0 * 1024.0);
  }
  if (st != PM7_OK) {
    r.res.status = st;
    r.res.n_atoms = mol->n_atoms;
    r.res.missing_element = r.job.missing_z;
    set_msg(&r.res, msg);
    s->reqs.emplace(id, std::move(r));
    srv_finish(s, id, std::chrono::steady_clock::now());
    s->cv_done.notify_all();
    return PM7_OK;
  }
  r.valid = true;
  s->reqs.emplace(id, std::move(r));
  s->queue.push_back(id);
  s->cv_work.notify_one();
  return PM7_OK;
}

pm7_status pm7_server_poll(pm7_server* s, int64_t request_id, int32_t timeout_ms, int64_t* done_id,
                           pm7_result_v2* out, double* latency_ms) {
  if (!s || !out) return PM7_ERR_INVALID_INPUT;
  std::unique_lock<std::mutex> lk(s->mu);
  if (request_id >= 0) {
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end() || it->second.state == 3) return PM7_ERR_INVALID_INPUT;
  }
  auto ready = [&]() -> int64_t {
    if (request_id >= 0) return s->reqs[request_id].state == 2 ? request_id : -1;
    while (!s->finished.empty()) {
      const int64_t id = s->finished.front();
      auto it = s->reqs.find(id);
      if (it != s->reqs.end() && it->second.state == 2) return id;
      s->finished.pop_front();  // claimed by id already
    }
    return -1;
  };
  int64_t id = ready();
  if (id < 0 && timeout_ms != 0) {
    if (timeout_ms < 0) s->cv_done.wait(lk, [&] { return (id = ready()) >= 0; });
    else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return (id = ready()) >= 0; });
  }
  if (id < 0) return PM7_NOT_READY;
  auto it = s->reqs.find(id);
  *out = it->second.res;
  if (done_id) *done_id = id;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(it->second.t_done - it->second.t_submit).count();
  s->reqs.erase(it);
  return PM7_OK;
}

void pm7_server_stop(pm7_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->ctx->device);
  for (auto& kv : s->reqs) pm7_result_free(&kv.second.res);
  for (auto& f : s->flights) {
    if (f->thr.joinable()) f->thr.join();
    f->lane.destroy();
  }
  delete s;
}


int32_t pm7_abi_version(void) { return PM7_ABI_VERSION; }

void pm7_default_options(pm7_options_v2* o) {
  std::memset(o, 0, sizeof *o);
  o->method = PM7_METHOD_PM7;
  o->charge = 0.0;
  o->multiplicity = 1;
  o->max_scf = 200;
  o->force_dpath = 0;
  o->smooth_dispersion = 0;
  o->dipole_origin = PM7_DIPOLE_CENTRE_OF_MASS;
  o->initial_density_nao = 0;
  o->initial_density = nullptr;
  o->initial_spin_density = nullptr;
  o->e_tol = 1.0e-8;
  o->p_tol = 1.0e-7;
  o->use_diis = 1;
  o->accelerator = PM7_ACCEL_ADIIS_CDIIS;
  o->adiis_switch = 0.1;
  o->level_shift = 0.0;
  o->reference = PM7_REF_AUTO;
  o->precision = PM7_PRECISION_FP64;
  o->mixed_switch = 1.0e-3;
  o->max_memory_mb = 0;
}

pm7_status pm7_ctx_create(int32_t device, const char* data_dir, pm7_ctx** out) {
  if (!out) return PM7_ERR_INVALID_INPUT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return PM7_ERR_DEVICE;
  if (device < 0 || device >= count) return PM7_ERR_INVALID_INPUT;
  if (hipSetDevice(device) != hipSuccess) return PM7_ERR_DEVICE;
  pm7_ctx* ctx = new pm7_ctx();
  ctx->device = device;
  ctx->data_dir = data_dir ? data_dir : default_data_dir();
  if (const char* ms = std::getenv("PM7_AMD_MIXED_SWEEPS")) {
    const int v = std::atoi(ms);
    if (v >= 1 && v <= 40) ctx->mixed_sweeps = v;
  }
  if (const char* rb = std::getenv("PM7_AMD_ROTATE_BELOW")) {
    const double v = std::strtod(rb, nullptr);
    if (v >= 1.0e-6 && v <= 1.0e-1) ctx->rotate_below = v;
  }
  if (const char* gm = std::getenv("PM7_AMD_GRID_MIN_NAO")) {
    const int v = std::atoi(gm);
    if (v >= 32 && v <= VARIANT_MAX_NAO[LARGE_VARIANT - 1]) ctx->grid_min_nao = v;
  }
  if (const char* al = std::getenv("PM7_AMD_ASYNC_LANES")) {
    const int v = std::atoi(al);
    if (v >= 1 && v <= 8) ctx->async_max = v;
  }
  if (const char* ao = std::getenv("PM7_AMD_ASYNC_ONE_STREAM")) ctx->async_one_stream = std::atoi(ao) != 0;
  if (const char* te = std::getenv("PM7_AMD_EIG_TOL_EXP")) {
    const int v = std::atoi(te);
    if (v >= 16 && v <= 30) ctx->eig_tol_exp = v;
  }
  if (const char* dg = std::getenv("PM7_AMD_DC_GRAPH")) ctx->dc_graph = std::atoi(dg) != 0;
  if (const char* rc = std::getenv("PM7_AMD_RESPONSE_CONTRACT")) ctx->response_contract = std::atoi(rc) != 0;
  if (const char* eb = std::getenv("PM7_AMD_EIG_BLOCK")) {
    const int v = std::atoi(eb);
    ctx->eig_block_mask = v == 0 ? 0 : (v == 1 ? 0xF : (v & 0xF));
  }
  if (!ctx->lane.create()) {
    ctx->lane.destroy();
    delete ctx;
    return PM7_ERR_DEVICE;
  }
  {
    // The allocation-free lane: this device's copy of the scratch array, and one stream.
    void* sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !sp) {
      ctx->na_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
    } else if (!ctx->na_lane.create(true)) {
      ctx->na_err = "stream creation for the allocation-free lane failed";
      ctx->na_lane.destroy();
    } else {
      ctx->arena.base = static_cast<char*>(sp);
      ctx->arena.cap = NOALLOC_SCRATCH_BYTES;
      ctx->na_lane.use_arena(&ctx->arena);
      ctx->na_ready = true;
    }
  }
  pm7_status st = ensure_params(ctx, PM7_METHOD_PM7);
  if (st != PM7_OK) {
    pm7_ctx_destroy(ctx);
    return st;
  }
  *out = ctx;
  return PM7_OK;
}

void pm7_ctx_destroy(pm7_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (AsyncBatch* b : ctx->inflight) {  // batches never waited for: finish, drop their results
    (void)hipEventSynchronize(b->lane->ev[5]);
    delete b;
  }
  ctx->inflight.clear();
  for (Lane* l : ctx->async_lanes) {
    l->destroy();
    delete l;
  }
  ctx->async_lanes.clear();
  if (ctx->epoch) (void)hipEventDestroy(ctx->epoch);
  ctx->lane.destroy();
  if (ctx->na_ready) ctx->na_lane.destroy();
  ctx->d_elem.release();
  ctx->d_pairpar.release();
  ctx->d_gl.release();
  ctx->d_ocd.release();
  delete ctx;
}

const char* pm7_last_error(const pm7_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

pm7_status pm7_params_load(pm7_ctx* ctx, int32_t method) {
  if (!ctx) return PM7_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  std::string msg;
  const pm7_status st = method_ok(method, msg);
  if (st != PM7_OK) return fail(ctx, st, msg);
  return ensure_params(ctx, method);
}

namespace {
pm7_status run_batch_impl(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                          int32_t n_opts, pm7_result_v2* results, double* gradient);
}  // namespace

pm7_status pm7_run_batch(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                         int32_t n_opts, pm7_result_v2* results) {
  return run_batch_impl(ctx, n, mols, opts, n_opts, results, nullptr);
}

pm7_status pm7_gradient_batch(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                              int32_t n_opts, pm7_result_v2* results, double* gradient) {
  if (n > 0 && !gradient) return PM7_ERR_INVALID_INPUT;
  return run_batch_impl(ctx, n, mols, opts, n_opts, results, gradient);
}

pm7_status pm7_gradient(pm7_ctx* ctx, const pm7_molecule_v1* mol, const pm7_options_v2* opts, pm7_result_v2* out,
                        double* gradient) {
  if (!ctx || !mol || !out || !gradient) return PM7_ERR_INVALID_INPUT;
  pm7_status st = pm7_gradient_batch(ctx, 1, mol, opts, opts ? 1 : 0, out, gradient);
  if (st != PM7_OK) return st;
  return (pm7_status)out->status;
}

namespace {
pm7_status run_batch_impl(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                          int32_t n_opts, pm7_result_v2* results, double* gradient) {
  if (!ctx || n < 0 || (n > 0 && (!mols || !results))) return PM7_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  pm7_options_v2 def;
  pm7_default_options(&def);
  std::vector<Job> jobs;
  int64_t budget = 0;
  for (int i = 0; i < n; ++i) {
    std::memset(&results[i], 0, sizeof results[i]);
    const pm7_options_v2& o = (opts && n_opts > 0) ? opts[n_opts == 1 ? 0 : i] : def;
    Job j;
    j.idx = i;
    std::string msg;
    pm7_status st = validate(ctx, mols[i], o, j, msg);
    if (st != PM7_OK) {
      results[i].status = st;
      results[i].n_atoms = mols[i].n_atoms;
      results[i].missing_element = j.missing_z;
      set_msg(&results[i], msg);
      continue;
    }
    if (o.max_memory_mb > 0) budget = budget > 0 ? std::min(budget, o.max_memory_mb) : o.max_memory_mb;
    jobs.push_back(std::move(j));
  }
  GradReq greq;
  if (gradient) {
    greq.out = gradient;
    greq.off.resize(n);
    int64_t off = 0;
    for (int i = 0; i < n; ++i) {
      greq.off[i] = off;
      const int na = std::max(0, mols[i].n_atoms);
      for (int k = 0; k < 3 * na; ++k) gradient[off + k] = std::nan("");
      off += 3 * (int64_t)na;
    }
  }
  if (jobs.empty()) return PM7_OK;
  pm7_status st = run_jobs(ctx, jobs, results, budget, gradient ? &greq : nullptr);
  return st == PM7_ERR_INSUFFICIENT_MEMORY ? PM7_OK : st;
}
}  // namespace

pm7_status pm7_batch_submit(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                            int32_t n_opts, pm7_result_v2* results, int64_t* ticket) {
  if (!ctx || n < 0 || !ticket || (n > 0 && (!mols || !results))) return PM7_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  const auto t0 = std::chrono::steady_clock::now();
  *ticket = ctx->next_ticket++;
  pm7_options_v2 def;
  pm7_default_options(&def);
  std::vector<Job> jobs;
  int64_t budget = 0;
  for (int i = 0; i < n; ++i) {
    std::memset(&results[i], 0, sizeof results[i]);
    const pm7_options_v2& o = (opts && n_opts > 0) ? opts[n_opts == 1 ? 0 : i] : def;
    Job j;
    j.idx = i;
    std::string msg;
    const pm7_status st = validate(ctx, mols[i], o, j, msg);
    if (st != PM7_OK) {
      results[i].status = st;
      results[i].n_atoms = mols[i].n_atoms;
      results[i].missing_element = j.missing_z;
      set_msg(&results[i], msg);
      continue;
    }
    if (o.max_memory_mb > 0) budget = budget > 0 ? std::min(budget, o.max_memory_mb) : o.max_memory_mb;
    jobs.push_back(std::move(j));
  }
  if (jobs.empty()) {
    ctx->async_done[*ticket] = PM7_OK;
    return PM7_OK;
  }
  // A lane: a free one, a new one while fewer than async_max exist, else the oldest batch's.
  if ((int)ctx->inflight.size() >= ctx->async_max) async_retire_oldest(ctx);
  Lane* lane = nullptr;
  for (Lane* l : ctx->async_lanes) {
    bool busy = false;
    for (const AsyncBatch* b : ctx->inflight) busy = busy || b->lane == l;
    if (!busy) {
      lane = l;
      break;
    }
  }
  if (!lane) {
    lane = new Lane();
    if (!lane->create(ctx->async_one_stream)) {
      lane->destroy();
      delete lane;
      return fail(ctx, PM7_ERR_DEVICE, "stream creation for an asynchronous batch failed");
    }
    ctx->async_lanes.push_back(lane);
  }
  if (!ctx->epoch) {
    if (hipEventCreate(&ctx->epoch) != hipSuccess || hipEventRecord(ctx->epoch, lane->st[0]) != hipSuccess)
      return fail(ctx, PM7_ERR_DEVICE, "timing event of the asynchronous batches failed");
  }
  auto* b = new AsyncBatch();
  b->ticket = *ticket;
  b->lane = lane;
  b->results = results;
  b->t0 = t0;
  b->P.jobs = std::move(jobs);
  Lane& L = *lane;
  pm7_status st = prepare(ctx, L, b->P, budget);
  if (st == PM7_OK) st = launch_batch(L, b->P, b->count);
  if (st == PM7_OK) st = run_large(L, b->P);  // molecules above 196 AOs: blocks until they finish
  if (st == PM7_OK) st = finish_batch(L, b->P);
  if (st != PM7_OK) {
    ctx->err = L.err;
    for (const Job& j : b->P.jobs) {
      results[j.idx].status = st;
      results[j.idx].needed_mb = b->P.need_mb;
      results[j.idx].budget_mb = b->P.budget_mb;
      set_msg(&results[j.idx], L.err);
    }
    if (st != PM7_ERR_INSUFFICIENT_MEMORY) (void)hipStreamSynchronize(L.st[0]);
    delete b;
    ctx->async_done[*ticket] = st == PM7_ERR_INSUFFICIENT_MEMORY ? PM7_OK : st;
    return st == PM7_ERR_INSUFFICIENT_MEMORY ? PM7_OK : st;
  }
  host_corrections(b->P.jobs);  // host work of this batch while the device runs it
  ctx->inflight.push_back(b);
  return PM7_OK;
}

pm7_status pm7_batch_wait(pm7_ctx* ctx, int64_t ticket) {
  if (!ctx) return PM7_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  auto it = ctx->async_done.find(ticket);
  if (it != ctx->async_done.end()) {
    const pm7_status st = it->second;
    ctx->async_done.erase(it);
    return st;
  }
  for (auto q = ctx->inflight.begin(); q != ctx->inflight.end(); ++q) {
    if ((*q)->ticket != ticket) continue;
    AsyncBatch* b = *q;
    ctx->inflight.erase(q);
    const pm7_status st = async_collect(ctx, b);
    delete b;
    return st;
  }
  return fail(ctx, PM7_ERR_INVALID_INPUT, "unknown batch ticket " + std::to_string(ticket));
}

pm7_status pm7_run(pm7_ctx* ctx, const pm7_molecule_v1* mol, const pm7_options_v2* opts, pm7_result_v2* out) {
  if (!ctx || !mol || !out) return PM7_ERR_INVALID_INPUT;
  pm7_status st = pm7_run_batch(ctx, 1, mol, opts, opts ? 1 : 0, out);
  if (st != PM7_OK) return st;
  return (pm7_status)out->status;
}

void pm7_result_free(pm7_result_v2* r) {
  if (!r) return;
  std::free(r->density);  // the base of the result's single allocation
  double** arrays[] = {&r->density, &r->spin_density, &r->mo_energies, &r->mo_coeff,
                       &r->mo_energies_beta, &r->mo_coeff_beta, &r->charges};
  for (double** a : arrays) *a = nullptr;
}

pm7_status pm7_core_hamiltonian(pm7_ctx* ctx, const pm7_molecule_v1* mol, double* h_out, int32_t* nao_out) {
  if (!ctx || !mol) return PM7_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  pm7_options_v2 o;
  pm7_default_options(&o);
  o.method = ctx->method;
  o.reference = PM7_REF_UHF;  // any electron count is fine for the core Hamiltonian
  o.multiplicity = 1;
  Job j;
  std::string msg;
  pm7_status st = validate(ctx, *mol, o, j, msg);
  if (st != PM7_OK && st != PM7_ERR_INVALID_INPUT) return fail(ctx, st, msg);
  if (st == PM7_ERR_INVALID_INPUT) {  // retry as a doublet when the electron count is odd
    o.multiplicity = 2;
    st = validate(ctx, *mol, o, j, msg);
    if (st != PM7_OK) return fail(ctx, st, msg);
  }
  if (nao_out) *nao_out = j.nao;
  if (!h_out) return PM7_OK;
  Prepared P;
  P.jobs.push_back(j);
  st = prepare(ctx, ctx->lane, P, 0);
  if (st != PM7_OK) return fail(ctx, st, ctx->lane.err);
  const Job& J = P.jobs[0];
  HIPCHK(hipMemcpy2DAsync(h_out, J.nao * sizeof(double), P.b.slab + J.mat0, J.ld * sizeof(double),
                          J.nao * sizeof(double), J.nao, hipMemcpyDeviceToHost, ctx->lane.st[0]));
  HIPCHK(hipStreamSynchronize(ctx->lane.st[0]));
  return PM7_OK;
}

pm7_status pm7_fock_spin(pm7_ctx* ctx, const pm7_molecule_v1* mol, const double* p_tot, const double* p_spin,
                         double* f_out) {
  if (!ctx || !mol || !p_tot || !p_spin || !f_out) return PM7_ERR_INVALID_INPUT;
  int32_t nao = 0;
  pm7_status st = pm7_core_hamiltonian(ctx, mol, nullptr, &nao);
  if (st != PM7_OK) return st;
  std::lock_guard<std::mutex> lk(ctx->mu);
  pm7_options_v2 o;
  pm7_default_options(&o);
  o.method = ctx->method;
  o.reference = PM7_REF_UHF;
  Job j;
  std::string msg;
  st = validate(ctx, *mol, o, j, msg);
  if (st != PM7_OK) {
    o.multiplicity = 2;
    st = validate(ctx, *mol, o, j, msg);
    if (st != PM7_OK) return fail(ctx, st, msg);
  }
  Prepared P;
  P.jobs.push_back(j);
  st = prepare(ctx, ctx->lane, P, 0);
  if (st != PM7_OK) return fail(ctx, st, ctx->lane.err);
  const Job& J = P.jobs[0];
  const int64_t sq = (int64_t)J.ld * J.ld;
  double* base = P.b.slab + J.mat0;
  HIPCHK(hipMemcpy2DAsync(base + 3 * sq, J.ld * sizeof(double), p_tot, J.nao * sizeof(double), J.nao * sizeof(double),
                          J.nao, hipMemcpyHostToDevice, ctx->lane.st[0]));
  HIPCHK(hipMemcpy2DAsync(base + 4 * sq, J.ld * sizeof(double), p_spin, J.nao * sizeof(double), J.nao * sizeof(double),
                          J.nao, hipMemcpyHostToDevice, ctx->lane.st[0]));
  HIPCHK(launch_fock_once(P.b, 0, 3, 4, 1, ctx->lane.st[0]));
  HIPCHK(hipMemcpy2DAsync(f_out, J.nao * sizeof(double), base + 1 * sq, J.ld * sizeof(double), J.nao * sizeof(double),
                          J.nao, hipMemcpyDeviceToHost, ctx->lane.st[0]));
  HIPCHK(hipStreamSynchronize(ctx->lane.st[0]));
  return PM7_OK;
}

pm7_status pm7_fock_spin_noalloc(pm7_ctx* ctx, const pm7_molecule_v1* mol, const double* p_tot,
                                 const double* p_spin, double* f_out) {
  if (!ctx || !mol || !p_tot || !p_spin || !f_out) return PM7_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (!ctx->na_ready) return fail(ctx, PM7_ERR_DEVICE, "allocation-free lane unavailable: " + ctx->na_err);
  pm7_options_v2 o;
  pm7_default_options(&o);
  o.method = ctx->method;
  o.reference = PM7_REF_UHF;  // any electron count: the Fock build needs no occupation
  Job j;
  std::string msg;
  pm7_status st = validate(ctx, *mol, o, j, msg);
  if (st != PM7_OK) {
    o.multiplicity = 2;
    st = validate(ctx, *mol, o, j, msg);
    if (st != PM7_OK) return fail(ctx, st, msg);
  }
  Lane& L = ctx->na_lane;
  ctx->arena.off = 0;  // every call lays the molecule out from the start of the scratch
  Prepared P;
  P.jobs.push_back(j);
  // The budget check of prepare() sees the scratch size; carving fails past it as well.
  st = prepare(ctx, L, P, (int64_t)(NOALLOC_SCRATCH_BYTES >> 20));
  if (st == PM7_ERR_DEVICE && L.err.find("out of memory") != std::string::npos) st = PM7_ERR_INSUFFICIENT_MEMORY;
  if (st != PM7_OK) {
    return fail(ctx, st, st == PM7_ERR_INSUFFICIENT_MEMORY
                             ? "molecule needs more than the " + std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                                   " MiB device scratch of pm7_fock_spin_noalloc: " + L.err
                             : L.err);
  }
  const Job& J = P.jobs[0];
  const int64_t sq = (int64_t)J.ld * J.ld;
  double* base = P.b.slab + J.mat0;
  HIPCHK(hipMemcpy2DAsync(base + 3 * sq, J.ld * sizeof(double), p_tot, J.nao * sizeof(double), J.nao * sizeof(double),
                          J.nao, hipMemcpyHostToDevice, L.st[0]));
  HIPCHK(hipMemcpy2DAsync(base + 4 * sq, J.ld * sizeof(double), p_spin, J.nao * sizeof(double), J.nao * sizeof(double),
                          J.nao, hipMemcpyHostToDevice, L.st[0]));
  HIPCHK(launch_fock_once(P.b, 0, 3, 4, 1, L.st[0]));
  HIPCHK(hipMemcpy2DAsync(f_out, J.nao * sizeof(double), base + 1 * sq, J.ld * sizeof(double), J.nao * sizeof(double),
                          J.nao, hipMemcpyDeviceToHost, L.st[0]));
  HIPCHK(hipStreamSynchronize(L.st[0]));
  return PM7_OK;
}

namespace {

// n_dens closed-shell densities (spin = false: in[0] -> out[0]) or alpha/beta pairs (spin = true:
// in[0], in[1] -> out[0], out[1]) of one molecule through the response kernel, one launch.
pm7_status response_fock_impl(pm7_ctx* ctx, const pm7_molecule_v1* mol, int32_t n_dens, bool spin,
                              const double* const* in, double* const* out) {
  const int n_sp = spin ? 2 : 1;
  if (!ctx || !mol || n_dens < 0) return PM7_ERR_INVALID_INPUT;
  for (int k = 0; k < n_sp && n_dens > 0; ++k)
    if (!in[k] || !out[k]) return PM7_ERR_INVALID_INPUT;
  int32_t nao = 0;
  pm7_status st = pm7_core_hamiltonian(ctx, mol, nullptr, &nao);
  if (st != PM7_OK) return st;
  if (n_dens == 0) return PM7_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  pm7_options_v2 o;
  pm7_default_options(&o);
  o.method = ctx->method;
  o.reference = PM7_REF_UHF;
  Job j;
  std::string msg;
  st = validate(ctx, *mol, o, j, msg);
  if (st != PM7_OK) {
    o.multiplicity = 2;
    st = validate(ctx, *mol, o, j, msg);
    if (st != PM7_OK) return fail(ctx, st, msg);
  }
  Prepared P;
  P.jobs.push_back(j);
  st = prepare(ctx, ctx->lane, P, 0);
  if (st != PM7_OK) return fail(ctx, st, ctx->lane.err);
  const Job& J = P.jobs[0];
  Lane& L = ctx->lane;
  const int64_t slice = (int64_t)J.nao * J.ld;  // densities are stored with row stride ld, nao rows each
  const int64_t pp_stride = ((int64_t)J.n_pp + 31) / 32 * 32;
  bool spd = false;  // some atom with d orbitals
  for (int z : J.z) spd = spd || ctx->params.el[z].n_orb > 4;
  // Closed shell: the density-index contractions (one integral read per 256 densities) where the
  // molecule fits them; the open-shell pairs and larger molecules take one workgroup per density.
  const bool contract = !spin && ctx->response_contract && response_pairs_fit(J.n_pp);
  const int dp = response_dp(n_dens);
  // alpha and beta slices back to back (plus the alpha + beta sums for the open-shell kernel)
  const size_t mat_bytes = contract ? (size_t)J.nao * J.nao * dp * sizeof(double)
                                : (size_t)n_sp * n_dens * (size_t)slice * sizeof(double);
  const size_t host_bytes = (size_t)n_dens * (size_t)J.nao * J.nao * sizeof(double);
  const size_t pp_bytes = (size_t)n_dens * (size_t)(contract ? J.n_pp : pp_stride) * sizeof(double);
  if (L.d_rx.ensure(mat_bytes + (spin ? mat_bytes / 2 : 0)) != hipSuccess || L.d_rg.ensure(mat_bytes) != hipSuccess ||
      L.d_rstage.ensure(n_sp * host_bytes) != hipSuccess || L.d_rpp.ensure(pp_bytes) != hipSuccess ||
      (contract && (L.d_rwall.ensure((size_t)J.n_pp * J.n_pp * sizeof(double)) != hipSuccess ||
                L.d_rfd.ensure(pp_bytes) != hipSuccess))) {
    (void)hipGetLastError();
    return fail(ctx, PM7_ERR_INSUFFICIENT_MEMORY,
                "response Fock: no device memory for " + std::to_string(n_dens) + " densities of " +
                    std::to_string(J.nao) + " AOs");
  }
  double* xs = static_cast<double*>(L.d_rx.p);
  double* gs = static_cast<double*>(L.d_rg.p);
  // One contiguous copy each way over PCIe (a pitched copy from pageable memory goes row by row);
  // the change of layout (row stride ld for the per-density kernels, density-fastest for the
  // contractions) is a device-to-device pass.
  double* stage = static_cast<double*>(L.d_rstage.p);
  const int64_t host_words = (int64_t)n_dens * J.nao * J.nao;
  for (int k = 0; k < n_sp; ++k)
    HIPCHK(hipMemcpyAsync(stage + k * host_words, in[k], host_bytes, hipMemcpyHostToDevice, L.st[0]));
  if (contract)
    HIPCHK(launch_response_relayout(stage, xs, J.nao, n_dens, dp, true, L.st[0]));
  else
    HIPCHK(hipMemcpy2DAsync(xs, J.ld * sizeof(double), stage, J.nao * sizeof(double), J.nao * sizeof(double),
                            (size_t)n_sp * n_dens * J.nao, hipMemcpyDeviceToDevice, L.st[0]));
  HIPCHK(hipEventRecord(L.ev[0], L.st[0]));
  if (spin)
    HIPCHK(launch_fock_response_spin(P.b, 0, xs, xs + 2 * n_dens * slice, gs, static_cast<double*>(L.d_rpp.p),
                                     pp_stride, n_dens, L.st[0]));
  else if (contract)
    HIPCHK(launch_response_pairs(P.b, 0, J.n_pp, (int)J.z.size(), spd, xs, gs, n_dens, dp,
                                 static_cast<double*>(L.d_rwall.p), static_cast<double*>(L.d_rpp.p),
                                 static_cast<double*>(L.d_rfd.p), L.st[0]));
  else
    HIPCHK(launch_fock_response(P.b, 0, xs, gs, static_cast<double*>(L.d_rpp.p), pp_stride, n_dens, L.st[0]));
  HIPCHK(hipEventRecord(L.ev[1], L.st[0]));
  if (contract)
    HIPCHK(launch_response_relayout(gs, stage, J.nao, n_dens, dp, false, L.st[0]));
  else
    HIPCHK(hipMemcpy2DAsync(stage, J.nao * sizeof(double), gs, J.ld * sizeof(double), J.nao * sizeof(double),
                            (size_t)n_sp * n_dens * J.nao, hipMemcpyDeviceToDevice, L.st[0]));
  for (int k = 0; k < n_sp; ++k)
    HIPCHK(hipMemcpyAsync(out[k], stage + k * host_words, host_bytes, hipMemcpyDeviceToHost, L.st[0]));
  HIPCHK(hipStreamSynchronize(L.st[0]));
  float ms = 0.0f;
  for (double& t : ctx->timings) t = 0.0;
  if (hipEventElapsedTime(&ms, L.ev[0], L.ev[1]) == hipSuccess) ctx->timings[1] = ms;
  (void)hipGetLastError();  // as in run_jobs: a failed timing query is not the next call's error
  return PM7_OK;
}

}  // namespace

pm7_status pm7_response_fock_batch(pm7_ctx* ctx, const pm7_molecule_v1* mol, int32_t n_dens, const double* dp,
                                   double* g_out) {
  const double* in[1] = {dp};
  double* out[1] = {g_out};
  return response_fock_impl(ctx, mol, n_dens, false, in, out);
}

pm7_status pm7_response_fock_spin_batch(pm7_ctx* ctx, const pm7_molecule_v1* mol, int32_t n_dens,
                                        const double* dpa, const double* dpb, double* ga_out, double* gb_out) {
  const double* in[2] = {dpa, dpb};
  double* out[2] = {ga_out, gb_out};
  return response_fock_impl(ctx, mol, n_dens, true, in, out);
}

pm7_status pm7_core_core_energy(pm7_ctx* ctx, const pm7_molecule_v1* mol, double* e_out) {
  if (!ctx || !mol || !e_out) return PM7_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  pm7_options_v2 o;
  pm7_default_options(&o);
  o.method = ctx->method;
  o.reference = PM7_REF_UHF;
  Job j;
  std::string msg;
  pm7_status st = validate(ctx, *mol, o, j, msg);
  if (st != PM7_OK) {
    o.multiplicity = 2;
    st = validate(ctx, *mol, o, j, msg);
    if (st != PM7_OK) return fail(ctx, st, msg);
  }
  Prepared P;
  P.jobs.push_back(j);
  st = prepare(ctx, ctx->lane, P, 0);
  if (st != PM7_OK) return fail(ctx, st, ctx->lane.err);
  HIPCHK(hipMemcpyAsync(e_out, P.b.e_core_mol, sizeof(double), hipMemcpyDeviceToHost, ctx->lane.st[0]));
  HIPCHK(hipStreamSynchronize(ctx->lane.st[0]));
  return PM7_OK;
}

// ---------------------------------------------------------------------------------------------
// Divide-and-conquer SCF (pm7-rs run_dandc, src/dandc/mod.rs:138-420)
// ---------------------------------------------------------------------------------------------
void pm7_dandc_default_options(pm7_dandc_options_v1* d) {
  if (!d) return;
  std::memset(d, 0, sizeof *d);
  d->buffer_bohr = 15.0;
  d->core_size = 12;
  d->max_scf = 200;
  d->fermi_width_ev = 0.1;
  d->p_tol = 1.0e-6;
  d->mixing = 0.4;
}

namespace {

// DandcOptions::validate (src/dandc/mod.rs:91-115), plus the step and tolerance limits.
bool dandc_options_ok(const pm7_dandc_options_v1& d, std::string& msg) {
  if (!std::isfinite(d.buffer_bohr) || d.buffer_bohr <= 0.0)
    return msg = "the divide-and-conquer buffer must be finite and positive", false;
  if (d.core_size < 1) return msg = "the divide-and-conquer core size must be at least 1", false;
  if (!std::isfinite(d.fermi_width_ev) || d.fermi_width_ev <= 0.0)
    return msg = "the divide-and-conquer Fermi width must be finite and positive", false;
  if (!(d.mixing > 0.0 && d.mixing <= 1.0)) return msg = "the divide-and-conquer mixing factor must lie in (0, 1]", false;
  if (d.max_scf < 1) return msg = "the divide-and-conquer max_scf must be at least 1", false;
  if (!(std::isfinite(d.p_tol) && d.p_tol > 0.0)) return msg = "the divide-and-conquer p_tol must be finite and positive", false;
  return true;
}

// Device memory of one run beside its subsystem batch.
struct DcBufs {
  DevBuf diag_off, norb, xyz, core, parent, atom_mol, nc_ao, aocode, ao0, amap, amap0, slot_off, slot_dims, cstart,
      contrib, normalizer, P, S, pop, g, v_nuc, v_el, wnorm, e_part, partial, scal, fermi_part, fermi_state;
  PinnedBuf h_scal;
  hipEvent_t ev[6] = {};  // timing (0, 1), class fan-out (2) and fan-in (3 + class)
  ~DcBufs() {
    DevBuf* all[] = {&diag_off, &norb, &xyz, &core, &parent, &atom_mol, &nc_ao, &aocode, &ao0, &amap, &amap0,
                     &slot_off, &slot_dims, &cstart, &contrib, &normalizer, &P, &S, &pop, &g, &v_nuc, &v_el,
                     &wnorm, &e_part, &partial, &scal, &fermi_part, &fermi_state};
    for (DevBuf* b : all) b->release();
    h_scal.release();
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

pm7_status dandc_impl(pm7_ctx* ctx, const pm7_molecule_v1& mol, const pm7_options_v2& o0,
                      const pm7_dandc_options_v1& d, pm7_dandc_result_v1* out) {
  auto refuse = [&](pm7_status st, const std::string& m) {
    out->status = st;
    std::snprintf(out->message, sizeof out->message, "%s", m.c_str());
    ctx->err = m;
    return st;
  };
  std::string msg;
  if (!dandc_options_ok(d, msg)) return refuse(PM7_ERR_INVALID_INPUT, msg);
  // The reference starts from the atomic guess whatever the options hold.
  pm7_options_v2 o = o0;
  o.initial_density = nullptr;
  o.initial_spin_density = nullptr;
  o.initial_density_nao = 0;
  Job whole;
  pm7_status st = validate(ctx, mol, o, whole, msg);
  if (st != PM7_OK) {
    out->n_atoms = mol.n_atoms;
    return refuse(st, msg);
  }
  const int n = mol.n_atoms;
  const double* X = mol.xyz_bohr;
  const bool uhf = whole.uhf != 0;
  std::vector<int> norb(n);
  std::vector<double> core(n);
  double tore = 0.0;
  for (int a = 0; a < n; ++a) {
    norb[a] = ctx->params.el[mol.z[a]].n_orb;
    core[a] = ctx->params.el[mol.z[a]].core_charge;
    tore += core[a];
  }
  DcPartition part;
  dc_partition(n, X, d.core_size, d.buffer_bohr, part);
  const int ns = part.n_sub();
  int largest = 0, largest_nao = 0;
  int64_t amap_words = 0, n_contrib = 0;
  for (int s = 0; s < ns; ++s) {
    const int len = (int)(part.start[s + 1] - part.start[s]), nc = part.n_core[s];
    int nao = 0;
    for (int i = 0; i < len; ++i) nao += norb[part.members[part.start[s] + i]];
    largest = std::max(largest, len);
    largest_nao = std::max(largest_nao, nao);
    amap_words += (int64_t)len * len;
    n_contrib += (int64_t)len * len - (int64_t)(len - nc) * (len - nc);
  }
  out->n_atoms = n;
  out->subsystems = ns;
  out->largest_subsystem = largest;
  out->largest_subsystem_nao = largest_nao;
  out->unrestricted = uhf;
  const double buffer_ang = d.buffer_bohr * A0;
  auto knob = [&](const std::string& what) {
    char b[512];
    std::snprintf(b, sizeof b,
                  "%s\nThis is a divide-and-conquer subsystem, not the whole system: a %.2f A buffer around a "
                  "%d-atom core gives a largest subsystem of %d atoms and %d orbitals. Reduce the buffer, or drop "
                  "divide and conquer and use the exact SCF (pm7_run): below roughly 350 atoms the exact SCF is both "
                  "faster and exact.",
                  what.c_str(), buffer_ang, d.core_size, largest, largest_nao);
    return std::string(b);
  };
  if (largest_nao > VARIANT_MAX_NAO[LARGE_VARIANT - 1])
    return refuse(PM7_ERR_UNSUPPORTED, knob("one subsystem workgroup holds at most " +
                                            std::to_string(VARIANT_MAX_NAO[LARGE_VARIANT - 1]) +
                                            " basis functions"));

  // Stored pattern: the union of the subsystems' weighted atom pairs, CSR over atoms with sorted
  // columns (src/dandc/mod.rs:218-230, src/dandc/sparse.rs with_pattern).
  std::vector<int64_t> row_start(n + 1, 0);
  std::vector<int> cols;
  std::vector<int64_t> offs;
  {
    std::vector<std::vector<int>> nbr(n);
    for (int s = 0; s < ns; ++s) {
      const int* m = &part.members[part.start[s]];
      const int len = (int)(part.start[s + 1] - part.start[s]), nc = part.n_core[s];
      for (int i = 0; i < len; ++i)
        for (int j = 0; j < len; ++j)
          if (i < nc || j < nc) nbr[m[i]].push_back(m[j]);
    }
    int64_t total = 0;
    for (int a = 0; a < n; ++a) {
      std::vector<int>& l = nbr[a];
      std::sort(l.begin(), l.end());
      l.erase(std::unique(l.begin(), l.end()), l.end());
      for (int b : l) {
        cols.push_back(b);
        offs.push_back(total);
        total += (int64_t)norb[a] * norb[b];
      }
      row_start[a + 1] = (int64_t)cols.size();
      std::vector<int>().swap(l);
    }
    offs.push_back(total);
  }
  const int64_t n_slots = (int64_t)cols.size(), n_el = offs.back();
  auto slot_of = [&](int a, int b) -> int64_t {
    auto lo = cols.begin() + row_start[a], hi = cols.begin() + row_start[a + 1];
    auto it = std::lower_bound(lo, hi, b);
    return (it != hi && *it == b) ? (int64_t)(it - cols.begin()) : -1;
  };
  out->stored_pairs = n_slots;
  out->stored_elements = n_el;
  // Partition weights normalised by their realised totals (src/dandc/mod.rs:232-258).
  std::vector<double> normalizer(n_slots, 0.0);
  for (int s = 0; s < ns; ++s) {
    const int* m = &part.members[part.start[s]];
    const int len = (int)(part.start[s + 1] - part.start[s]), nc = part.n_core[s];
    for (int i = 0; i < len; ++i)
      for (int j = 0; j < len; ++j) {
        const double w = 0.5 * ((i < nc ? 1.0 : 0.0) + (j < nc ? 1.0 : 0.0));
        if (w > 0.0) normalizer[slot_of(m[i], m[j])] += w;
      }
  }
  for (double& w : normalizer) w = w > 1.0e-12 ? 1.0 / w : 0.0;
  // Starting densities: MOPAC's atomic guess of the whole system on the diagonal blocks, and for
  // an open shell a spin density scaled from its diagonal (src/dandc/mod.rs:261-306).
  const double n_elec = tore - o.charge;
  const double yy = whole.nao > 0 ? (tore - n_elec) / whole.nao : 0.0;
  const double spin_scale = tore > 0.0 ? 2.0 * (0.5 * (o.multiplicity - 1)) / tore : 0.0;
  std::vector<double> P0(n_el, 0.0), S0(uhf ? n_el : 0, 0.0);
  std::vector<int64_t> diag_off(n);
  for (int a = 0; a < n; ++a) {
    const int64_t k = slot_of(a, a);
    diag_off[a] = offs[k];
    const int na = norb[a];
    sad_atom(ctx->elem_host[mol.z[a]], yy, 1.0, P0.data() + offs[k], 0, na);
    if (uhf)
      for (int i = 0; i < na; ++i) S0[offs[k] + i * na + i] = spin_scale * P0[offs[k] + i * na + i];
  }

  // The subsystems as the molecules of one batch: integrals and core Hamiltonians once.
  Lane& L = ctx->lane;
  Prepared P;
  P.fixed_classes = true;
  P.jobs.resize(ns);
  for (int s = 0; s < ns; ++s) {
    Job& j = P.jobs[s];
    j.idx = s;
    j.o = o;
    const int len = (int)(part.start[s + 1] - part.start[s]);
    int nao = 0, npp = 0;
    for (int i = 0; i < len; ++i) {
      const int a = part.members[part.start[s] + i];
      j.z.push_back(mol.z[a]);
      for (int k = 0; k < 3; ++k) j.xyz.push_back(X[3 * a + k]);
      nao += norb[a];
      npp += npair_of(norb[a]);
      if (norb[a] == 9) j.dpath = 1;  // the cluster's own atoms decide (build_core of the cluster)
    }
    j.nao = nao;
    j.ld = nao + (nao & 1);
    j.n_pp = npp;
    j.uhf = uhf;
    j.variant = -1;
    for (int v = 0; v < LARGE_VARIANT; ++v)
      if (nao <= VARIANT_MAX_NAO[v]) {
        j.variant = v;
        break;
      }
  }
  P.extra_bytes = (size_t)n_el * 8 * (uhf ? 3 : 2) + (size_t)n_slots * (8 + 4 + 8 + 8 + 8) +
                  (size_t)n_contrib * sizeof(DcContrib) + (size_t)amap_words * 8 + (size_t)n * 64 +
                  (size_t)part.members.size() * 40;
  st = prepare(ctx, L, P, o.max_memory_mb);
  if (st != PM7_OK) {
    out->needed_mb = P.need_mb;
    out->budget_mb = P.budget_mb;
    return refuse(st, st == PM7_ERR_INSUFFICIENT_MEMORY ? knob(L.err) : L.err);
  }

  // Host tables of the device side.
  const int nb_atoms = (int)part.members.size();
  std::vector<int> atom_mol(nb_atoms), nc_ao(ns), aocode;
  std::vector<int64_t> ao0(ns), amap, amap0(ns), cstart(n_slots + 1, 0);
  std::vector<int> slot_dims(n_slots);
  for (int a = 0; a < n; ++a)
    for (int64_t k = row_start[a]; k < row_start[a + 1]; ++k) slot_dims[k] = norb[a] | (norb[cols[k]] << 8);
  amap.reserve(amap_words);
  for (int s = 0; s < ns; ++s) {
    const int* m = &part.members[part.start[s]];
    const int len = (int)(part.start[s + 1] - part.start[s]), nc = part.n_core[s];
    ao0[s] = (int64_t)aocode.size();
    amap0[s] = (int64_t)amap.size();
    int core_ao = 0;
    for (int i = 0; i < len; ++i) {
      atom_mol[part.start[s] + i] = s;
      for (int k = 0; k < norb[m[i]]; ++k) aocode.push_back(dc_ao_code(i, k, norb[m[i]]));
      if (i < nc) core_ao += norb[m[i]];
    }
    nc_ao[s] = core_ao;
    for (int i = 0; i < len; ++i)
      for (int j = 0; j < len; ++j) {
        const int64_t k = slot_of(m[i], m[j]);
        amap.push_back(k < 0 ? -1 : offs[k]);
        if (k >= 0 && (i < nc || j < nc)) ++cstart[k + 1];
      }
  }
  for (int64_t k = 0; k < n_slots; ++k) cstart[k + 1] += cstart[k];
  std::vector<DcContrib> contrib(n_contrib);
  {
    std::vector<int64_t> fill(cstart.begin(), cstart.end() - 1);
    for (int s = 0; s < ns; ++s) {
      const Job& j = P.jobs[s];
      const int* m = &part.members[part.start[s]];
      const int len = (int)(part.start[s + 1] - part.start[s]), nc = part.n_core[s];
      const int64_t sq = (int64_t)j.ld * j.ld;
      int oi = 0;
      for (int i = 0; i < len; ++i) {
        int oj = 0;
        for (int jj = 0; jj < len; ++jj) {
          if (i < nc || jj < nc) {
            DcContrib c;
            c.base = j.mat0 + (uhf ? (int64_t)DU_DA : (int64_t)DR_D) * sq + (int64_t)oi * j.ld + oj;
            c.dstep = (int64_t)(DU_DB - DU_DA) * sq;
            c.ld = j.ld;
            c.pad = 0;
            c.w = 0.5 * ((i < nc ? 1.0 : 0.0) + (jj < nc ? 1.0 : 0.0));
            contrib[fill[slot_of(m[i], m[jj])]++] = c;
          }
          oj += norb[m[jj]];
        }
        oi += norb[m[i]];
      }
    }
  }
  std::vector<double> xyz(X, X + 3 * n);

  DcBufs B;
  hipStream_t s0 = L.st[0];
#define DCCHK(expr)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return refuse(PM7_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)
  auto up = [&](DevBuf& buf, const void* src, size_t bytes) -> hipError_t {
    hipError_t e = buf.ensure(std::max<size_t>(bytes, 8));
    if (e != hipSuccess || !bytes) return e;
    return hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, s0);
  };
  DCCHK(up(B.diag_off, diag_off.data(), diag_off.size() * 8));
  DCCHK(up(B.norb, norb.data(), norb.size() * 4));
  DCCHK(up(B.xyz, xyz.data(), xyz.size() * 8));
  DCCHK(up(B.core, core.data(), core.size() * 8));
  DCCHK(up(B.parent, part.members.data(), part.members.size() * 4));
  DCCHK(up(B.atom_mol, atom_mol.data(), atom_mol.size() * 4));
  DCCHK(up(B.nc_ao, nc_ao.data(), nc_ao.size() * 4));
  DCCHK(up(B.aocode, aocode.data(), aocode.size() * 4));
  DCCHK(up(B.ao0, ao0.data(), ao0.size() * 8));
  DCCHK(up(B.amap, amap.data(), amap.size() * 8));
  DCCHK(up(B.amap0, amap0.data(), amap0.size() * 8));
  DCCHK(up(B.slot_off, offs.data(), offs.size() * 8));
  DCCHK(up(B.slot_dims, slot_dims.data(), slot_dims.size() * 4));
  DCCHK(up(B.cstart, cstart.data(), cstart.size() * 8));
  DCCHK(up(B.contrib, contrib.data(), contrib.size() * sizeof(DcContrib)));
  DCCHK(up(B.normalizer, normalizer.data(), normalizer.size() * 8));
  DCCHK(up(B.P, P0.data(), P0.size() * 8));
  DCCHK(up(B.S, S0.data(), S0.size() * 8));
  DCCHK(B.pop.ensure((size_t)n * 8));
  DCCHK(B.g.ensure((size_t)n * 8));
  DCCHK(B.v_nuc.ensure((size_t)nb_atoms * 8));
  DCCHK(B.v_el.ensure((size_t)nb_atoms * 8));
  DCCHK(B.wnorm.ensure((size_t)P.vec_words * 8));
  DCCHK(B.e_part.ensure((size_t)ns * 8));
  DCCHK(B.partial.ensure((size_t)std::max<int64_t>(n_slots, 1) * 8));
  DCCHK(B.scal.ensure(8 * 8));
  DCCHK(B.fermi_part.ensure((size_t)2 * DC_FERMI_PART_WORDS * 8));
  DCCHK(B.fermi_state.ensure(8 * 8));
  DCCHK(B.h_scal.ensure(8 * 8));
  for (int k = 0; k < 6; ++k) DCCHK(hipEventCreateWithFlags(&B.ev[k], k < 2 ? hipEventDefault : hipEventDisableTiming));
  DcDev dd;
  dd.n_atoms = n;
  dd.n_sub = ns;
  dd.uhf = uhf;
  dd.n_batch_atoms = nb_atoms;
  dd.n_slots = n_slots;
  dd.n_el = n_el;
  dd.n_states = P.vec_words;
  dd.diag_off = B.diag_off.as<int64_t>();
  dd.norb = B.norb.as<int>();
  dd.xyz = B.xyz.as<double>();
  dd.core = B.core.as<double>();
  dd.parent = B.parent.as<int>();
  dd.atom_mol = B.atom_mol.as<int>();
  dd.nc_ao = B.nc_ao.as<int>();
  dd.aocode = B.aocode.as<int>();
  dd.ao0 = B.ao0.as<int64_t>();
  dd.amap = B.amap.as<int64_t>();
  dd.amap0 = B.amap0.as<int64_t>();
  dd.slot_off = B.slot_off.as<int64_t>();
  dd.slot_dims = B.slot_dims.as<int>();
  dd.cstart = B.cstart.as<int64_t>();
  dd.contrib = B.contrib.as<DcContrib>();
  dd.normalizer = B.normalizer.as<double>();
  dd.P = B.P.as<double>();
  dd.S = uhf ? B.S.as<double>() : nullptr;
  dd.pop = B.pop.as<double>();
  dd.g = B.g.as<double>();
  dd.v_nuc = B.v_nuc.as<double>();
  dd.v_el = B.v_el.as<double>();
  dd.wnorm = B.wnorm.as<double>();
  dd.e_part = B.e_part.as<double>();
  dd.partial = B.partial.as<double>();
  dd.scal = B.scal.as<double>();
  dd.fermi_part = B.fermi_part.as<double>();
  dd.fermi_state = B.fermi_state.as<double>();
  const BatchDev& b = P.b;
  // The nuclear environment, once per geometry, into the subsystems' core Hamiltonians.
  DCCHK(launch_dc_field(b, dd, dd.core, dd.g, dd.v_nuc, s0));
  DCCHK(launch_dc_shift_h(b, dd, s0));
  DCCHK(hipEventRecord(B.ev[0], s0));
  const double* hs = static_cast<const double*>(B.h_scal.p);
  int it = 0;
  bool converged = false;
  double rms = INFINITY, e_el = 0.0, fermi = 0.0;
  // One step's launches (about 22: the environment, the gather, the size classes' solves forked
  // onto their streams and joined by events, the Fermi passes, the densities, the assembly, the
  // finish and the 32-byte read-back). They are the same every step but for the parity of the
  // eigenvector buffers (and step 0's cold solve), so steps >= 1 replay one of two HIP graphs, each
  // captured the first time its parity comes up (no memset nodes: every buffer the step clears is
  // written by a kernel). PM7_AMD_DC_GRAPH=0 issues every step eagerly.
  auto issue_step = [&](int step) -> hipError_t {
    hipError_t e;
    if ((e = launch_dc_pop(dd, s0)) != hipSuccess) return e;
    if ((e = launch_dc_field(b, dd, dd.pop, dd.g, dd.v_el, s0)) != hipSuccess) return e;
    if ((e = launch_dc_gather(b, dd, s0)) != hipSuccess) return e;
    if ((e = hipEventRecord(B.ev[2], s0)) != hipSuccess) return e;
    for (int v = 0; v < LARGE_VARIANT; ++v) {
      const int cnt = P.list_off[v + 1] - P.list_off[v];
      if (cnt == 0) continue;
      hipStream_t sv = v == 0 ? s0 : L.st[v];
      if (sv != s0 && (e = hipStreamWaitEvent(sv, B.ev[2], 0)) != hipSuccess) return e;
      if ((e = launch_dc_solve(v, b, P.list_off[v], cnt, dd.nc_ao, dd.v_el, step, dd.wnorm, dd.e_part, sv)) !=
          hipSuccess)
        return e;
      if (sv != s0) {
        if ((e = hipEventRecord(B.ev[3 + v], sv)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s0, B.ev[3 + v], 0)) != hipSuccess) return e;
      }
    }
    if ((e = launch_dc_fermi(b, dd, n_elec, d.fermi_width_ev, s0)) != hipSuccess) return e;
    if ((e = launch_dc_density(b, ns, dd.scal, d.fermi_width_ev, step, s0)) != hipSuccess) return e;
    if ((e = launch_dc_assemble(b, dd, d.mixing, s0)) != hipSuccess) return e;
    if ((e = launch_dc_finish(dd, s0)) != hipSuccess) return e;
    return hipMemcpyAsync(B.h_scal.p, dd.scal, 4 * 8, hipMemcpyDeviceToHost, s0);
  };
  struct StepGraphs {
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    ~StepGraphs() {
      for (hipGraphExec_t& g : exec)
        if (g) (void)hipGraphExecDestroy(g);
    }
  } graphs;
  const bool use_graph = ctx->dc_graph;
  for (; it < d.max_scf;) {
    if (use_graph && it >= 1) {
      hipGraphExec_t& gx = graphs.exec[it & 1];
      if (!gx) {
        hipGraph_t g = nullptr;
        DCCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
        const hipError_t e_issue = issue_step(it);
        const hipError_t e_end = hipStreamEndCapture(s0, &g);
        if (e_issue != hipSuccess || e_end != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          return refuse(PM7_ERR_DEVICE, std::string("capturing a divide-and-conquer step: ") +
                                            hipGetErrorString(e_issue != hipSuccess ? e_issue : e_end));
        }
        const hipError_t e_inst = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DCCHK(e_inst);
      }
      DCCHK(hipGraphLaunch(gx, s0));
    } else {
      DCCHK(issue_step(it));
    }
    DCCHK(hipStreamSynchronize(s0));
    ++it;
    if (hs[3] != 0.0) {
      // No chemical potential brackets the electron count (src/dandc/mod.rs:744-758).
      std::vector<double> ev(P.vec_words);
      DCCHK(hipMemcpy(ev.data(), b.vecs, ev.size() * 8, hipMemcpyDeviceToHost));
      double lo = INFINITY, hi = -INFINITY;
      for (const Job& j : P.jobs) {
        lo = std::min(lo, ev[j.vec0]);
        hi = std::max(hi, ev[j.vec0 + j.nao - 1]);
      }
      char bmsg[256];
      std::snprintf(bmsg, sizeof bmsg,
                    "no chemical potential in [-500, 500] eV holds %g electrons; the subsystem spectra span %.3f "
                    "to %.3f eV",
                    n_elec, lo, hi);
      out->iterations = it;
      return refuse(PM7_ERR_INVALID_INPUT, bmsg);
    }
    rms = hs[0];
    e_el = hs[1];
    fermi = hs[2];
    if (rms < d.p_tol) {
      converged = true;
      break;
    }
  }
  DCCHK(hipEventRecord(B.ev[1], s0));
  DCCHK(hipEventSynchronize(B.ev[1]));
  float ms = 0.0f;
  DCCHK(hipEventElapsedTime(&ms, B.ev[0], B.ev[1]));
  out->device_ms = ms;
  out->iterations = it;
  out->converged = converged;
  out->density_error = rms;
  out->electronic_ev = e_el;
  out->fermi_ev = fermi;

  // Results: one allocation based at row_start.
  const size_t bytes = (size_t)(n + 1) * 8 + (size_t)(n_slots + 1) * 8 + (size_t)n_el * 8 * (uhf ? 2 : 1) +
                       (size_t)n * 8 + (size_t)n_slots * 4 + (size_t)n * 4;
  char* blk = static_cast<char*>(std::malloc(std::max<size_t>(bytes, 8)));
  if (!blk) return refuse(PM7_ERR_INSUFFICIENT_MEMORY, "host allocation of the divide-and-conquer result failed");
  char* q = blk;
  auto take = [&](size_t nbytes) {
    char* r = q;
    q += nbytes;
    return r;
  };
  out->row_start = reinterpret_cast<int64_t*>(take((size_t)(n + 1) * 8));
  out->offsets = reinterpret_cast<int64_t*>(take((size_t)(n_slots + 1) * 8));
  out->density = reinterpret_cast<double*>(take((size_t)n_el * 8));
  out->spin_density = uhf ? reinterpret_cast<double*>(take((size_t)n_el * 8)) : nullptr;
  out->charges = reinterpret_cast<double*>(take((size_t)n * 8));
  out->columns = reinterpret_cast<int32_t*>(take((size_t)n_slots * 4));
  out->norb = reinterpret_cast<int32_t*>(take((size_t)n * 4));
  std::copy(row_start.begin(), row_start.end(), out->row_start);
  std::copy(offs.begin(), offs.end(), out->offsets);
  std::copy(cols.begin(), cols.end(), out->columns);
  std::copy(norb.begin(), norb.end(), out->norb);
  DCCHK(hipMemcpy(out->density, dd.P, (size_t)n_el * 8, hipMemcpyDeviceToHost));
  if (uhf) DCCHK(hipMemcpy(out->spin_density, dd.S, (size_t)n_el * 8, hipMemcpyDeviceToHost));
  for (int a = 0; a < n; ++a) {
    double pop = 0.0;
    for (int i = 0; i < norb[a]; ++i) pop += out->density[diag_off[a] + i * norb[a] + i];
    out->charges[a] = core[a] - pop;
  }
#undef DCCHK
  out->status = PM7_OK;
  return PM7_OK;
}

}  // namespace

pm7_status pm7_run_dandc(pm7_ctx* ctx, const pm7_molecule_v1* mol, const pm7_options_v2* opts,
                         const pm7_dandc_options_v1* dopts, pm7_dandc_result_v1* out) {
  if (!ctx || !mol || !out) return PM7_ERR_INVALID_INPUT;
  std::memset(out, 0, sizeof *out);
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  pm7_options_v2 o;
  if (opts) o = *opts;
  else pm7_default_options(&o);
  pm7_dandc_options_v1 d;
  if (dopts) d = *dopts;
  else pm7_dandc_default_options(&d);
  if (!opts) o.method = ctx->method;
  const pm7_status st = dandc_impl(ctx, *mol, o, d, out);
  if (st != PM7_OK) {
    pm7_dandc_result_free(out);
    out->status = st;
  }
  return st;
}

void pm7_dandc_result_free(pm7_dandc_result_v1* r) {
  if (!r) return;
  std::free(r->row_start);  // the base of the result's single allocation
  r->row_start = nullptr;
  r->offsets = nullptr;
  r->columns = nullptr;
  r->norb = nullptr;
  r->density = nullptr;
  r->spin_density = nullptr;
  r->charges = nullptr;
}

pm7_status pm7_dandc_partition(const pm7_molecule_v1* mol, int32_t core_size, double buffer_bohr,
                               pm7_partition_v1* out) {
  if (!out) return PM7_ERR_INVALID_INPUT;
  std::memset(out, 0, sizeof *out);
  if (!mol || mol->n_atoms <= 0 || !mol->xyz_bohr || core_size < 1 || !std::isfinite(buffer_bohr) ||
      buffer_bohr <= 0.0)
    return PM7_ERR_INVALID_INPUT;
  for (int i = 0; i < 3 * mol->n_atoms; ++i)
    if (!std::isfinite(mol->xyz_bohr[i])) return PM7_ERR_INVALID_INPUT;
  DcPartition part;
  dc_partition(mol->n_atoms, mol->xyz_bohr, core_size, buffer_bohr, part);
  const int ns = part.n_sub();
  const size_t bytes = (size_t)(ns + 1) * 8 + (size_t)ns * 4 + part.members.size() * 4;
  char* blk = static_cast<char*>(std::malloc(bytes));
  if (!blk) return PM7_ERR_INSUFFICIENT_MEMORY;
  out->n_sub = ns;
  out->start = reinterpret_cast<int64_t*>(blk);
  out->n_core = reinterpret_cast<int32_t*>(blk + (size_t)(ns + 1) * 8);
  out->members = out->n_core + ns;
  std::copy(part.start.begin(), part.start.end(), out->start);
  std::copy(part.n_core.begin(), part.n_core.end(), out->n_core);
  std::copy(part.members.begin(), part.members.end(), out->members);
  return PM7_OK;
}

void pm7_partition_free(pm7_partition_v1* p) {
  if (!p) return;
  std::free(p->start);  // the base of the single allocation
  p->start = nullptr;
  p->n_core = nullptr;
  p->members = nullptr;
  p->n_sub = 0;
}

void pm7_last_timings(const pm7_ctx* ctx, double* ms, int32_t n) {
  if (!ctx || !ms) return;
  for (int k = 0; k < n && k < PM7_TIMING_SLOTS; ++k) ms[k] = ctx->timings[k];
}

}  // extern "C"

namespace pm7 {

TuneKnobs tune_knobs(const pm7_ctx* ctx) {
  return TuneKnobs{ctx->eig_block_mask, ctx->mixed_sweeps, ctx->async_max};
}

void set_tune_knobs(pm7_ctx* ctx, const TuneKnobs& k) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->eig_block_mask = k.eig_block_mask & 0xF;
  if (k.mixed_sweeps >= 1 && k.mixed_sweeps <= 40) ctx->mixed_sweeps = k.mixed_sweeps;
  if (k.batches_in_flight >= 1 && k.batches_in_flight <= 8) ctx->async_max = k.batches_in_flight;
}

int ctx_device(const pm7_ctx* ctx) { return ctx->device; }

int ctx_batches_in_flight(pm7_ctx* ctx) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  return (int)ctx->inflight.size();
}

void ctx_set_error(pm7_ctx* ctx, const std::string& msg) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->err = msg;
}

}  // namespace pm7

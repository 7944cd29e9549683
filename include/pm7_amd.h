/* SPDX-License-Identifier: GPL-3.0-or-later
 *
 * pm7_amd: PM7 NDDO SCF single points on one MI355X (gfx950), behind a plain C ABI.
 *
 * The boundary replaces pm7-rs's
 *     pub fn run_pm7(molecule: &Molecule, params: &Pm7Parameters, options: &Pm7Options)
 *         -> Result<Pm7Result>                                     (pm7-rs src/scf.rs:512-516)
 * Units follow the reference: coordinates in Bohr (src/system.rs:23), energies in eV, heats of
 * formation in kcal/mol, matrices row-major with eigenvectors as COLUMNS in ascending eigenvalue
 * order (src/linalg.rs:229-230). Status codes map one-to-one onto pm7-rs's Pm7Error variants
 * (src/error.rs:13-48) plus PM7_ERR_DEVICE.
 *
 * Threading: a context owns one device and its streams; calls on one context are serialised
 * internally. Separate contexts may be used from separate threads.
 */
#ifndef PM7_AMD_H
#define PM7_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM7_ABI_VERSION 2

typedef enum {
  PM7_OK = 0,
  PM7_ERR_INVALID_INPUT = 1,      /* Pm7Error::InvalidInput */
  PM7_ERR_MISSING_ELEMENT = 2,    /* Pm7Error::MissingElement */
  PM7_ERR_MISSING_PARAMETER = 3,  /* Pm7Error::MissingParameter */
  PM7_ERR_UNSUPPORTED = 4,        /* Pm7Error::UnsupportedFeature */
  PM7_ERR_LINEAR_ALGEBRA = 5,     /* Pm7Error::LinearAlgebra */
  PM7_ERR_SCF_NOT_CONVERGED = 6,  /* Pm7Error::ScfNotConverged (energies are still filled) */
  PM7_ERR_INSUFFICIENT_MEMORY = 7,/* Pm7Error::InsufficientMemory */
  PM7_ERR_DEVICE = 8,             /* HIP runtime failure */
  PM7_NOT_READY = 9               /* request server: the request has not finished yet */
} pm7_status;

/* pm7-rs Pm7Method (src/method.rs:9-16) with its parameter-cache slots (src/params.rs:145-150).
 * PM7-Sparkle (lanthanide +3 point cores, src/params.rs:171-172) is refused with
 * PM7_ERR_UNSUPPORTED; any other code is PM7_ERR_INVALID_INPUT. PM7-HH is PM7 plus the H-H
 * repulsion correction (src/hh_rep.rs:39-60, src/scf.rs:922-923). */
typedef enum {
  PM7_METHOD_PM7 = 0,
  PM7_METHOD_PM7_TS = 1,
  PM7_METHOD_PM7_SPARKLE = 2,
  PM7_METHOD_PM7_MINUS = 3,
  PM7_METHOD_PM7_HH = 4
} pm7_method;
typedef enum { PM7_ACCEL_NONE = 0, PM7_ACCEL_CDIIS = 1, PM7_ACCEL_ADIIS_CDIIS = 2 } pm7_accelerator;
typedef enum { PM7_REF_AUTO = 0, PM7_REF_RHF = 1, PM7_REF_UHF = 2 } pm7_reference;
typedef enum { PM7_PRECISION_FP64 = 0, PM7_PRECISION_MIXED = 1 } pm7_precision;
/* pm7-rs DipoleOrigin (src/dipole.rs:71-96): origin of the point-charge dipole of an ion
 * (a molecule with |net charge| <= 0.5 always uses the coordinate origin, src/dipole.rs:143-146). */
typedef enum {
  PM7_DIPOLE_CENTRE_OF_MASS = 0,
  PM7_DIPOLE_COORDINATES = 1,
  PM7_DIPOLE_CENTRE_OF_CHARGE = 2
} pm7_dipole_origin;

typedef struct pm7_ctx pm7_ctx;

/* pm7-rs Molecule (src/system.rs:20-39): atomic numbers and Bohr coordinates, xyz[3*i + k]. */
typedef struct {
  int32_t n_atoms;
  const int32_t* z;
  const double* xyz_bohr;
} pm7_molecule_v1;

/* pm7-rs Pm7Options (src/scf.rs:73-234), defaults from pm7_default_options (src/scf.rs:207-234).
 * Not offered: periodic cells, external fields, stability analysis and the Hessian's exchange cutoff
 * and CPHF budget (outside the molecular single point). */
typedef struct {
  int32_t method;        /* pm7_method; must match the loaded tables (src/scf.rs:530-535) */
  double charge;         /* options.charge is the charge used (src/scf.rs:592) */
  int32_t multiplicity;  /* >= 1 */
  int32_t max_scf;       /* default 200 */
  double e_tol;          /* eV, default 1e-8 */
  double p_tol;          /* RMS density change, default 1e-7 */
  int32_t use_diis;      /* default 1; 0 forces plain iteration */
  int32_t accelerator;   /* pm7_accelerator, default ADIIS->CDIIS */
  double adiis_switch;   /* commutator norm of the A-DIIS -> CDIIS switch, default 0.1 */
  double level_shift;    /* static level-shift floor in eV, default 0 */
  int32_t reference;     /* pm7_reference, default AUTO */
  int32_t precision;     /* pm7_precision, default FP64 (engine extension) */
  double mixed_switch;   /* FP32 -> FP64 hand-over density change for PRECISION_MIXED (1e-3) */
  int64_t max_memory_mb; /* device-memory budget for this call; <= 0 means 80% of free HBM */
  int32_t force_dpath;   /* every pair through the MNDO/d kernels, sp molecules too (src/scf.rs:97) */
  int32_t smooth_dispersion; /* continuous carbon C6 (src/scf.rs:120, src/dispersion.rs:133-175) */
  int32_t dipole_origin; /* pm7_dipole_origin (src/scf.rs:187) */
  int32_t initial_density_nao; /* rows = columns of the two arrays below; must equal the basis size */
  /* Start the SCF from this density instead of the atomic guess (src/scf.rs:161-163, 1201-1210),
   * nao*nao row-major, or NULL. UHF: P_alpha = n_alpha/n P, P_beta = n_beta/n P (src/scf.rs:1341-1360). */
  const double* initial_density;
  /* UHF only: add this spin density P_alpha - P_beta (P_alpha += s/2, P_beta -= s/2;
   * src/scf.rs:170, 1361-1378), nao*nao row-major, or NULL. */
  const double* initial_spin_density;
} pm7_options_v2;

/* pm7-rs Pm7Result (src/scf.rs:262-342), molecular fields. The arrays of one result are allocated
 * by the library (one allocation) and released together by pm7_result_free. */
typedef struct {
  int32_t status;
  int32_t iterations;
  int32_t converged;
  int32_t unrestricted;
  int32_t nao;
  int32_t n_atoms;
  int32_t n_occ;
  int32_t n_occ_beta;        /* -1 for RHF */
  int32_t fp32_iterations;   /* iterations run in FP32 under PRECISION_MIXED */
  int32_t jacobi_sweeps;     /* total eigensolver sweeps (diagnostic) */
  double electronic_ev;
  double core_ev;
  double total_ev;
  double heat_of_formation_kcal;
  double correction_kcal;    /* dispersion + EH+ + MM terms included in total/HOF */
  double density_error;      /* last RMS density change */
  double homo_ev, lumo_ev, homo_ev_beta, lumo_ev_beta; /* NaN when absent */
  double dipole_debye[3];
  double dipole_point_charge_debye[3];
  double dipole_sp_hybrid_debye[3];
  double dipole_pd_hybrid_debye[3];
  double dipole_magnitude;
  double device_phase_ms[6]; /* time inside this molecule's SCF workgroup: Fock build, energy +
                                commutator, DIIS extrapolation, eigensolve, density, whole SCF */
  double eig_phase_ms[3];    /* eigensolve split: warm-start transform, Jacobi sweeps, sort and
                                back-transform */
  /* PRECISION_MIXED: how far the result can be from the FP64 one. The reported density and energy
   * come from one full FP64 eigensolve of the Fock matrix of the converged density (a Roothaan
   * step); err_vs_fp64_ev = |E(step) - E(converged)| and err_vs_fp64_density = max |P(step) -
   * P(converged)|, both 0 at an exact fixed point. An FP64 run reports 0 for both (it is the
   * reference). */
  double err_vs_fp64_ev;
  double err_vs_fp64_density;
  int32_t missing_element;   /* PM7_ERR_MISSING_ELEMENT: the atomic number (Pm7Error::MissingElement) */
  int32_t reserved_;
  double needed_mb, budget_mb; /* PM7_ERR_INSUFFICIENT_MEMORY (Pm7Error::InsufficientMemory); what
                                  ran out is named in message */
  double* density;           /* nao*nao, total density */
  double* spin_density;      /* nao*nao, P_alpha - P_beta, UHF only (else NULL) */
  double* mo_energies;       /* nao */
  double* mo_coeff;          /* nao*nao, columns = MOs */
  double* mo_energies_beta;  /* UHF only */
  double* mo_coeff_beta;     /* UHF only */
  double* charges;           /* n_atoms Mulliken net charges */
  char message[256];
} pm7_result_v2;

/* Context: device + streams + parameter tables + pooled device memory.
 * data_dir: directory of the bundled parameter tables (NULL: next to the library). */
pm7_status pm7_ctx_create(int32_t device, const char* data_dir, pm7_ctx** out);
void pm7_ctx_destroy(pm7_ctx* ctx);
const char* pm7_last_error(const pm7_ctx* ctx);
int32_t pm7_abi_version(void);

/* Replaces Pm7Parameters::method (src/params.rs:130-155): (re)load the bundled tables. */
pm7_status pm7_params_load(pm7_ctx* ctx, int32_t method);

void pm7_default_options(pm7_options_v2* opts);

/* Replaces run_pm7 (src/scf.rs:512-867) for one molecule. */
pm7_status pm7_run(pm7_ctx* ctx, const pm7_molecule_v1* mol, const pm7_options_v2* opts,
                   pm7_result_v2* out);

/* Many molecules of uneven size in one persistent launch per size class; results[i] is
 * exactly what pm7_run would return for mols[i] (fixed-order reductions, no atomics).
 * opts may hold one entry (shared) or n entries. Returns PM7_OK when every call was dispatched;
 * per-molecule outcomes are in results[i].status. */
pm7_status pm7_run_batch(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols,
                         const pm7_options_v2* opts, int32_t n_opts, pm7_result_v2* results);

void pm7_result_free(pm7_result_v2* res);

/* pm7_run_batch split in two, so that consecutive batches overlap on the device (no pm7-rs
 * counterpart; it serves callers that issue batch after batch, as rayon callers of run_pm7 do).
 * pm7_batch_submit validates, lays out, uploads and launches the batch on one of the context's
 * asynchronous lanes (own streams and device buffers; PM7_AMD_ASYNC_LANES, default 2, batches in
 * flight) and returns with a ticket; a batch with molecules above 196 AOs runs those before it
 * returns. results[] must stay valid until pm7_batch_wait(ticket) has returned: the wait fills
 * them, exactly as pm7_run_batch would, and returns what pm7_run_batch would have returned. A
 * submit while every lane is busy first completes the oldest batch in flight (its wait then returns
 * at once). Per-molecule outcomes are in results[i].status. */
pm7_status pm7_batch_submit(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                            int32_t n_opts, pm7_result_v2* results, int64_t* ticket);
pm7_status pm7_batch_wait(pm7_ctx* ctx, int64_t ticket);

/* Replaces closed_form_gradient (src/gradient.rs:500-555) for many molecules: the SCF of each
 * (results[i] exactly as pm7_run_batch fills it), then dE_total/dR at the converged density on the
 * device, in eV/Bohr, plus the gradient of the geometric corrections (host). Molecule i's atom a,
 * axis k lands at gradient[3 * (n_atoms of molecules 0..i-1) + 3 * a + k]; the caller sizes the
 * array to 3 x the total atom count. A molecule whose SCF left no density gets NaN and
 * results[i].status = PM7_ERR_UNSUPPORTED (its energies, if any, are still filled). */
pm7_status pm7_gradient_batch(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols,
                              const pm7_options_v2* opts, int32_t n_opts, pm7_result_v2* results,
                              double* gradient);

/* One molecule (gradient: 3 * n_atoms doubles, eV/Bohr); returns the result's status. */
pm7_status pm7_gradient(pm7_ctx* ctx, const pm7_molecule_v1* mol, const pm7_options_v2* opts,
                        pm7_result_v2* out, double* gradient);

/* Replaces build_core (src/hamiltonian.rs:224-231): the core Hamiltonian H (nao*nao). */
pm7_status pm7_core_hamiltonian(pm7_ctx* ctx, const pm7_molecule_v1* mol, double* h_out,
                                int32_t* nao_out);

/* Replaces build_fock_spin (src/fock.rs:133-147): F = H + J(p_tot) - K(p_spin). */
pm7_status pm7_fock_spin(pm7_ctx* ctx, const pm7_molecule_v1* mol, const double* p_tot,
                         const double* p_spin, double* f_out);

/* No counterpart in the reference: pm7_fock_spin for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same inputs,
 * output and result bits as pm7_fock_spin; the molecule's device layout (atoms, pairs, pair
 * integrals, matrix slab) lives in a static per-device scratch of 32 MiB, so the call makes no
 * hipMalloc/hipFree and calls no allocating library routine. A molecule whose layout exceeds the
 * scratch gets PM7_ERR_INSUFFICIENT_MEMORY. Every call on a device shares that device's scratch:
 * calls on one context are serialised, but calls that overlap on one GPU from different contexts
 * race. */
pm7_status pm7_fock_spin_noalloc(pm7_ctx* ctx, const pm7_molecule_v1* mol, const double* p_tot,
                                 const double* p_spin, double* f_out);

/* Replaces the CPHF response kernel of cphf_ov (src/hessian.rs:1074-1081), which calls
 * build_fock_x (src/fock.rs:665-684) once per response density and subtracts H_core:
 * g_out[k] = build_fock_x(dp[k]) - H = J(dp[k]) - K(dp[k] / 2), for n_dens symmetric densities of one
 * molecule (each nao*nao, row-major, back to back; g_out the same shape). The pair integrals are
 * built once and every density is contracted against them on the device (one workgroup each).
 * n_dens = 0 does nothing; PM7_ERR_INSUFFICIENT_MEMORY when the densities do not fit. */
pm7_status pm7_response_fock_batch(pm7_ctx* ctx, const pm7_molecule_v1* mol, int32_t n_dens,
                                   const double* dp, double* g_out);

/* The open-shell response kernel of analytic_hessian_uhf (src/hessian.rs:1850-1880):
 * ga_out[k] = build_fock_spin_x(dpa[k] + dpb[k], dpa[k]) - H and gb_out[k] likewise with dpb[k],
 * for n_dens pairs of symmetric alpha/beta response densities (each array n_dens nao*nao matrices). */
pm7_status pm7_response_fock_spin_batch(pm7_ctx* ctx, const pm7_molecule_v1* mol, int32_t n_dens,
                                        const double* dpa, const double* dpb, double* ga_out,
                                        double* gb_out);

/* Replaces core_core_energy (src/repulsion.rs:16-36), eV. */
pm7_status pm7_core_core_energy(pm7_ctx* ctx, const pm7_molecule_v1* mol, double* e_out);

/* Divide-and-conquer SCF (pm7-rs DandcOptions, src/dandc/mod.rs:57-116; defaults from
 * pm7_dandc_default_options, src/dandc/mod.rs:77-89). */
typedef struct {
  double buffer_bohr;    /* buffer radius around each core, Bohr (default 15) */
  int32_t core_size;     /* target atoms per core (default 12) */
  int32_t max_scf;       /* default 200 */
  double fermi_width_ev; /* Fermi smearing of the common chemical potential (default 0.1 eV) */
  double p_tol;          /* RMS change over the stored density elements (default 1e-6) */
  double mixing;         /* linear density mixing, in (0, 1] (default 0.4) */
} pm7_dandc_options_v1;

/* pm7-rs DandcResult (src/dandc/mod.rs:117-135). The density is stored as atom-pair blocks, CSR
 * over atoms (pm7-rs SparseDensity, src/dandc/sparse.rs): the blocks of atom a are slots
 * row_start[a] .. row_start[a+1]-1, slot k is the pair (a, columns[k]) and holds the
 * norb[a] x norb[columns[k]] block row-major at density + offsets[k]. The arrays are allocated by
 * the library (one allocation) and released by pm7_dandc_result_free. */
typedef struct {
  int32_t status;
  int32_t converged;
  int32_t iterations;
  int32_t unrestricted;
  int32_t subsystems;
  int32_t largest_subsystem;      /* atoms */
  int32_t largest_subsystem_nao;  /* basis functions */
  int32_t n_atoms;
  int64_t stored_pairs;
  int64_t stored_elements;
  double fermi_ev;                /* common chemical potential */
  double electronic_ev;           /* sum of the subsystems' weighted energies */
  double density_error;           /* last RMS density change */
  double device_ms;               /* device time of the SCF steps (HIP events) */
  double needed_mb, budget_mb;    /* PM7_ERR_INSUFFICIENT_MEMORY */
  int64_t* row_start;             /* n_atoms + 1 */
  int32_t* columns;               /* stored_pairs */
  int64_t* offsets;               /* stored_pairs + 1 */
  int32_t* norb;                  /* n_atoms */
  double* density;                /* stored_elements */
  double* spin_density;           /* stored_elements, UHF only (else NULL) */
  double* charges;                /* n_atoms Mulliken net charges (mulliken_charges, src/dandc/mod.rs:823-835) */
  char message[512];
} pm7_dandc_result_v1;

void pm7_dandc_default_options(pm7_dandc_options_v1* dopts);

/* Replaces run_dandc (src/dandc/mod.rs:138-420). Every subsystem's Fock build, eigensolve and
 * weighted density of a step run as one launch each (one workgroup per subsystem, integrals and
 * eigenvectors resident across steps); the global density, the environment field, the common
 * Fermi level and the mixing stay on the device. Subsystems above 196 basis functions are refused
 * (PM7_ERR_UNSUPPORTED) with the subsystem's atoms and orbitals named, as is a partition whose
 * device working set exceeds the budget (PM7_ERR_INSUFFICIENT_MEMORY). Like the reference, an
 * unconverged run returns PM7_OK with converged = 0. */
pm7_status pm7_run_dandc(pm7_ctx* ctx, const pm7_molecule_v1* mol, const pm7_options_v2* opts,
                         const pm7_dandc_options_v1* dopts, pm7_dandc_result_v1* out);
void pm7_dandc_result_free(pm7_dandc_result_v1* res);

/* Replaces partition (src/dandc/partition.rs:97-133), host only (no context, no device): recursive
 * median bisection into cores of at most core_size atoms, each grown by every atom within
 * buffer_bohr of one of its atoms. Subsystem s holds members[start[s] .. start[s+1]-1], its
 * n_core[s] core atoms first, then its buffer atoms in ascending index order. Released by
 * pm7_partition_free. */
typedef struct {
  int32_t n_sub;
  int32_t* n_core;   /* n_sub */
  int64_t* start;    /* n_sub + 1 */
  int32_t* members;  /* start[n_sub] */
} pm7_partition_v1;
pm7_status pm7_dandc_partition(const pm7_molecule_v1* mol, int32_t core_size, double buffer_bohr,
                               pm7_partition_v1* out);
void pm7_partition_free(pm7_partition_v1* part);

/* Request server: continuous batching of SCF requests (engine extension; no pm7-rs counterpart,
 * it serves the callers that would otherwise call run_pm7 once per request). A scheduler thread
 * admits queued requests, in arrival order, into up to `lanes` batches in flight at once (each on
 * its own streams), as long as the device memory of the in-flight batches stays within the
 * budget; a request that arrives while others run starts as soon as a lane is free. Each result
 * equals what pm7_run returns for that request alone (no atomics, fixed reduction orders).
 * pm7_params_load must not be called while a server runs on the context. */
typedef struct pm7_server pm7_server;
typedef struct {
  int32_t lanes;         /* batches in flight at once, 1..16 (default 4) */
  int32_t max_batch;     /* molecules per batch (default 256) */
  int64_t max_memory_mb; /* device memory for the in-flight batches; <= 0: 80% of free HBM */
} pm7_server_config_v1;

void pm7_server_default_config(pm7_server_config_v1* cfg);
pm7_status pm7_server_start(pm7_ctx* ctx, const pm7_server_config_v1* cfg, pm7_server** out);
/* Queue one request (opts may be NULL for the defaults). Invalid input, unsupported elements and
 * requests larger than the whole budget finish at once with that status. */
pm7_status pm7_server_submit(pm7_server* srv, const pm7_molecule_v1* mol, const pm7_options_v2* opts,
                             int64_t* request_id);
/* Claim a finished request: request_id >= 0 waits for that request, < 0 for any unclaimed one.
 * timeout_ms: 0 = do not wait, < 0 = wait without limit. Fills *out (release with
 * pm7_result_free), the id and the latency from submit to completion in ms; PM7_NOT_READY when
 * nothing finished in time. */
pm7_status pm7_server_poll(pm7_server* srv, int64_t request_id, int32_t timeout_ms, int64_t* done_id,
                           pm7_result_v2* out, double* latency_ms);
/* Finish the queued and in-flight requests, release unclaimed results, stop the thread. */
void pm7_server_stop(pm7_server* srv);

/* Autotuner (no pm7-rs counterpart; it serves callers that run many batches of one kind of
 * workload). Times candidate launch shapes on the caller's batch and keeps the fastest in the
 * context: per size class (<= 64, <= 120, <= 196 basis functions) the block Jacobi solver or the
 * scalar rounds, the Jacobi sweep cap of the reduced-precision phase (mixed runs only), and the
 * number of batches pm7_batch_submit keeps in flight, the one with the highest device throughput
 * whose median step latency stays within latency_budget (>= 1, 0: 2) times the one-in-flight
 * latency. Each candidate runs `steps` (1..64, 0: 4) batches after one untimed batch. Every
 * candidate gives results within the parity bars of the defaults; the number in flight does not
 * change results at all. cache_path (may be NULL): a text file of earlier choices keyed by device
 * and workload signature; a matching line is applied without timing (candidates = 0), a new choice
 * is written back. No batch may be in flight on the context during the call. */
typedef struct {
  int32_t eig_block_mask;       /* bit v: size class v uses the block Jacobi solver */
  int32_t mixed_sweeps;         /* Jacobi sweep cap of the reduced-precision phase */
  int32_t batches_in_flight;    /* pm7_batch_submit depth */
  int32_t candidates;           /* settings timed (0: taken from the cache file) */
  double solves_per_s;          /* the chosen setting: device throughput of `steps` batches */
  double step_ms;               /* its median submit-to-result latency per batch */
  double default_solves_per_s;  /* the context's settings before the call, measured the same way */
  double tune_ms;               /* wall time of the call */
} pm7_tune_result_v1;
pm7_status pm7_autotune(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                        int32_t n_opts, int32_t steps, double latency_budget, const char* cache_path,
                        pm7_tune_result_v1* out);

/* Timing of the last pm7_run / pm7_run_batch call in milliseconds, measured with HIP events on
 * the stream each kernel runs on: [0] integrals + H_core, [1] all SCF megakernels, [2] output
 * gather, [3] whole call (host wall clock), [4..7] SCF megakernel of size class 0/1/2/3 (<= 64,
 * <= 120, <= 196, larger basis functions; 0 when the class was empty), then host wall clock:
 * [8] validation, layout and upload, [9] kernel launches, host-side corrections and the copy of
 * the packed results to the host (which waits for the kernels), [10] remaining device wait,
 * [11] result assembly. After pm7_response_fock_batch: [1] its response kernel, the rest 0.
 * After pm7_batch_wait: [0]-[2] and [4]-[7] of that batch, [3] submit to wait, and [12], [13] the
 * device times at which its integrals kernel started and its output gather ended, in ms from the
 * context's first asynchronous submit (so the device span of several batches in flight is
 * max [13] - min [12]); 0 after the synchronous calls.
 * Writes min(n, PM7_TIMING_SLOTS) values. */
#define PM7_TIMING_SLOTS 14
void pm7_last_timings(const pm7_ctx* ctx, double* ms, int32_t n);

#ifdef __cplusplus
}
#endif
#endif

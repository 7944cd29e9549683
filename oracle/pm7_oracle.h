/* SPDX-License-Identifier: GPL-3.0-or-later
 *
 * CPU FP64 oracle for the PM7 NDDO SCF single point.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP engine is compared against; it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * the engine ships. It restates the pm7-rs algorithm (file:line citations in pm7_oracle.cpp) in
 * plain serial/threaded C++ with no GPU code.
 */
#ifndef PM7_ORACLE_H
#define PM7_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int method;          /* 0 = PM7, 1 = PM7-TS, 3 = PM7-minus (no post-SCF corrections), 4 = PM7-HH */
  int multiplicity;    /* >= 1 */
  int max_scf;         /* default 200 */
  int use_diis;        /* default 1 */
  int accelerator;     /* 0 none, 1 CDIIS, 2 A-DIIS->CDIIS (default) */
  int reference;       /* 0 auto, 1 RHF, 2 UHF */
  int threads;         /* worker threads for the pair/row loops; <= 0 means all cores */
  double charge;
  double e_tol;        /* default 1e-8 eV */
  double p_tol;        /* default 1e-7 */
  double adiis_switch; /* default 0.1 */
  double level_shift;  /* default 0 eV */
  int force_dpath;     /* every pair through the MNDO/d integrals (src/scf.rs:97) */
  int smooth_dispersion; /* continuous carbon C6 (src/dispersion.rs:72-175) */
  int dipole_origin;   /* 0 centre of mass, 1 coordinates, 2 centre of core charge (src/dipole.rs:71-96) */
  int pad_;
  const double* initial_density;      /* nao*nao or NULL (src/scf.rs:1201-1210, 1341-1360) */
  const double* initial_spin_density; /* nao*nao or NULL, UHF only (src/scf.rs:1361-1378) */
} oracle_options;

typedef struct {
  int status; /* 0 ok, 1 invalid input, 2 missing element, 3 unsupported, 5 not converged */
  int iterations;
  int converged;
  int unrestricted;
  int nao;
  int n_occ;
  int n_occ_beta;
  int n_atoms;
  double electronic_ev;
  double core_ev;
  double total_ev;
  double hof_kcal;
  double correction_kcal;
  double density_error;
  double homo_ev, lumo_ev, homo_ev_beta, lumo_ev_beta;
  double dipole[3];
  double dipole_point[3];
  double dipole_sp[3];
  double dipole_pd[3];
  char message[256];
} oracle_result;

void oracle_default_options(oracle_options* o);

/* Full single point. z[n], xyz_bohr[3n]. Optional outputs (may be NULL): density[nao*nao],
 * spin_density[nao*nao], eps[nao], charges[n]. nao can be queried with oracle_nao first. */
int oracle_run(const char* data_dir, int n_atoms, const int* z, const double* xyz_bohr,
               const oracle_options* opts, oracle_result* out, double* density,
               double* spin_density, double* eps, double* charges);

int oracle_nao(const char* data_dir, int n_atoms, const int* z);

/* Pieces, exposed so tests can pin them one by one. */
int oracle_hcore(const char* data_dir, int n_atoms, const int* z, const double* xyz_bohr,
                 double* h /* nao*nao */);
/* F = H + G(P_tot, P_spin) for one spin; RHF uses p_spin = P/2. */
int oracle_fock(const char* data_dir, int n_atoms, const int* z, const double* xyz_bohr,
                const double* p_tot, const double* p_spin, double* f);
/* Rotated two-centre block for the ordered pair (a heavier-or-equal first): w[npa*npb],
 * e1b[81], e2a[81] (9x9 row-major), s[81] overlap in the molecular frame. */
int oracle_pair_block(const char* data_dir, int za, int zb, const double* d_bohr, double* w,
                      double* e1b, double* e2a, double* s);
double oracle_core_core(const char* data_dir, int n_atoms, const int* z, const double* xyz_bohr);
/* Post-SCF geometric corrections in kcal/mol: [dispersion, hbond, c_triple, si_o_h]. */
void oracle_corrections(int n_atoms, const int* z, const double* xyz_bohr, double* out4);
/* [smooth dispersion, PM7-HH H-H repulsion] in kcal/mol. */
void oracle_corrections2(int n_atoms, const int* z, const double* xyz_bohr, double* out2);
/* Symmetric eigensolve (Householder + implicit QL), ascending, vectors as columns. */
int oracle_eigh(int n, const double* a, double* w, double* v);
/* Local-frame Slater overlaps [s111, s211, s121, s221, s222] (atom a has the higher n);
 * numeric != 0 forces the prolate-spheroidal quadrature used for n >= 4. */
int oracle_overlap_locals(int na, double zsa, double zpa, int nb, double zsb, double zpb, double r, int numeric,
                          double* out5);
/* Parameter derived quantities for tests: [n_orb, dd, qq, rho0, rho1, rho2, e_isol, eheat_ev, po9]. */
int oracle_element(const char* data_dir, int method, int z, double* out9);

/* Divide-and-conquer SCF (src/dandc/mod.rs:57-420), molecular case. */
typedef struct {
  double buffer_bohr;    /* default 15 */
  int core_size;         /* default 12 */
  int max_scf;           /* default 200 */
  double fermi_width_ev; /* default 0.1 */
  double p_tol;          /* default 1e-6 */
  double mixing;         /* default 0.4 */
} oracle_dandc_options;

typedef struct {
  int status;            /* 0 ok (also when not converged), 1 invalid input / no Fermi level, 2 missing element */
  int converged;
  int iterations;
  int unrestricted;
  int subsystems;
  int largest_subsystem;
  int largest_subsystem_nao;
  int nao;
  long long stored_pairs;
  long long stored_elements;
  double fermi_ev;
  double electronic_ev;
  double density_error;
  char message[256];
} oracle_dandc_result;

void oracle_default_dandc_options(oracle_dandc_options* d);
/* dense_density / dense_spin: nao*nao (the stored blocks, zero elsewhere) or NULL; charges[n] or NULL. */
int oracle_run_dandc(const char* data_dir, int n_atoms, const int* z, const double* xyz_bohr,
                     const oracle_options* opts, const oracle_dandc_options* dopts, oracle_dandc_result* out,
                     double* dense_density, double* dense_spin, double* charges);
/* Partition (src/dandc/partition.rs:97-133): returns the total member count; with the arrays NULL
 * only counts (n_sub and the return value), else fills n_core[n_sub], start[n_sub+1], members[cap]. */
int oracle_dandc_partition(int n_atoms, const double* xyz_bohr, int core_size, double buffer, int* n_sub,
                           int* n_core, int* start, int* members, int cap);

#ifdef __cplusplus
}
#endif
#endif

// SPDX-License-Identifier: GPL-3.0-or-later
//
// Host-side pieces of the engine (parameters, geometric post-SCF corrections, batch layout).
#pragma once

#include <array>
#include <string>
#include <utility>
#include <vector>

#include "pm7_device.h"

struct pm7_ctx;

namespace pm7 {

// Launch-shape settings of a context that pm7_autotune chooses (engine.cpp keeps them).
struct TuneKnobs {
  int eig_block_mask;     // bit v: size class v diagonalises with the block Jacobi solver
  int mixed_sweeps;       // Jacobi sweep cap of the reduced-precision phase
  int batches_in_flight;  // pm7_batch_submit depth
};
TuneKnobs tune_knobs(const pm7_ctx* ctx);
void set_tune_knobs(pm7_ctx* ctx, const TuneKnobs& k);
int ctx_device(const pm7_ctx* ctx);
int ctx_batches_in_flight(pm7_ctx* ctx);
void ctx_set_error(pm7_ctx* ctx, const std::string& msg);

struct HostElement {
  bool present = false;
  int z = 0, n = 0, ios = 0, iop = 0, iod = 0, n_orb = 0;
  double core_charge = 0, eheat_ev = 0, e_isol = 0;
  double u_ss = 0, u_pp = 0, u_dd = 0, zeta_s = 0, zeta_p = 0, zeta_d = 0;
  double beta_s = 0, beta_p = 0, beta_d = 0;
  double g_ss = 0, g_sp = 0, g_pp = 0, g_p2 = 0, h_sp = 0, poc = 0;
  double dd = 0, qq = 0, rho0 = 0, rho1 = 0, rho2 = 0;
  double zsn = 0, zpn = 0, zdn = 0, f0sd = 0, g2sd = 0;
  std::vector<std::array<double, 3>> gauss;
  // MNDO/d data (1-based slots): additive terms po(1..9), charge separations ddp(2..6), the
  // one-centre d integrals repd(1..52), the isolated-atom d-shell correction, and for a d-bearing
  // element the 45 x 45 one-centre matrix over packed orbital pairs.
  std::array<double, 10> po{};
  std::array<double, 7> ddp{};
  std::array<double, 53> repd{};
  double eiscor = 0;
  std::vector<double> oc_d;
};

struct ParamSet {
  int method = 0;
  HostElement el[MAX_Z];
  std::vector<std::pair<double, double>> pair;  // dense MAX_Z x MAX_Z (alpb, xfac), fallbacks resolved
  double vpar[61] = {0};
};

bool load_params(const std::string& dir, int method, ParamSet& P, std::string& err);
void derive_mndod(HostElement& e);
int shell_principal(int z, bool d_shell);
void to_device_table(const ParamSet& P, std::vector<ElemDev>& out);

// Geometric post-SCF corrections in kcal/mol (coordinates in Bohr).
// D2 dispersion; smooth: the carbon C6 from a continuous neighbour count (smooth_dispersion).
double dispersion_kcal(const std::vector<int>& z, const std::vector<double>& xyz, bool smooth = false);
// Analytic gradient of dispersion_kcal, kcal/mol/Bohr, accumulated into g[3 n].
void dispersion_grad_kcal(const std::vector<int>& z, const std::vector<double>& xyz, double* g, bool smooth = false);
// PM7-HH hydrogen-hydrogen repulsion and its analytic gradient (accumulated into g[3 n]).
double hh_repulsion_kcal(const std::vector<int>& z, const std::vector<double>& xyz);
void hh_repulsion_grad_kcal(const std::vector<int>& z, const std::vector<double>& xyz, double* g);
double hbond_kcal(const std::vector<int>& z, const std::vector<double>& xyz);
double c_triple_kcal(const std::vector<int>& z, const std::vector<double>& xyz);
double si_o_h_kcal(const std::vector<int>& z, const std::vector<double>& xyz);

}  // namespace pm7

// SPDX-License-Identifier: GPL-3.0-or-later
//
// PM7 parameter tables: read the bundled MOPAC v23.2.5 data (data/pm7_*.tsv) and derive the
// NDDO multipole quantities the integral kernels need (pm7-rs src/params.rs:205-432, 602-652).
#include "host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace pm7 {

namespace {

std::vector<std::vector<std::string>> tsv(const std::string& path) {
  std::vector<std::vector<std::string>> rows;
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) return rows;
  char buf[8192];
  while (std::fgets(buf, sizeof buf, fp)) {
    if (buf[0] == '#' || buf[0] == '\n') continue;
    std::vector<std::string> f;
    char* save = nullptr;
    for (char* t = strtok_r(buf, "\t\n", &save); t; t = strtok_r(nullptr, "\t\n", &save)) f.emplace_back(t);
    rows.push_back(std::move(f));
  }
  std::fclose(fp);
  return rows;
}

// Charge separations (MOPAC ddpo) and the fixed 5-step calpar secant for the additive terms.
double dd_sep(int n, double zs, double zp) {
  return (2.0 * n + 1.0) * std::pow(4.0 * zs * zp, n + 0.5) / std::pow(zs + zp, 2.0 * n + 2.0) / std::sqrt(3.0);
}
double qq_sep(int n, double zp) { return std::sqrt((4.0 * n * n + 6.0 * n + 2.0) / 20.0) / zp; }

template <class Fn>
double secant5(double target, double x0, Fn f) {
  double a = x0, b = x0 + 0.04;
  for (int i = 0; i < 5; ++i) {
    double fa = f(a), fb = f(b);
    if (std::fabs(fb - fa) < 1.0e-25) break;
    double c = a + (b - a) * (target - fa) / (fb - fa);
    a = b;
    b = c;
  }
  return b;
}
double sgn(double x) { return std::signbit(x) ? -1.0 : 1.0; }

double rho1_of(double hsp_ev, double dd) {
  double t = hsp_ev / EV;
  auto f = [dd](double d) { return 0.5 * d - 0.5 / std::sqrt(4.0 * dd * dd + 1.0 / (d * d)); };
  return 0.5 / secant5(t, std::pow(std::fabs(t) / (dd * dd), 1.0 / 3.0) * sgn(t), f);
}
double rho2_of(double hpp_ev, double qq) {
  double t = hpp_ev / EV;
  auto f = [qq](double q) {
    return 0.25 * q - 0.5 / std::sqrt(4.0 * qq * qq + 1.0 / (q * q)) + 0.25 / std::sqrt(8.0 * qq * qq + 1.0 / (q * q));
  };
  return 0.5 / secant5(t, std::pow(std::fabs(t) / (3.0 * std::pow(qq, 4)), 0.2) * sgn(t), f);
}

}  // namespace

bool load_params(const std::string& dir, int method, ParamSet& P, std::string& err) {
  const std::string stem = dir + (method == 1 ? "/pm7ts" : "/pm7");
  auto rows = tsv(stem + "_elements.tsv");
  if (rows.size() < 2) {
    err = "parameter table " + stem + "_elements.tsv not found";
    return false;
  }
  P = ParamSet();
  P.method = method;
  const auto& hdr = rows[0];
  auto col = [&](const char* name) -> int {
    for (size_t i = 0; i < hdr.size(); ++i)
      if (hdr[i] == name) return (int)i;
    return -1;
  };
  for (size_t r = 1; r < rows.size(); ++r) {
    auto v = [&](const char* name) {
      int c = col(name);
      return (c >= 0 && c < (int)rows[r].size()) ? std::strtod(rows[r][c].c_str(), nullptr) : 0.0;
    };
    int z = (int)v("z");
    if (z <= 0 || z >= MAX_Z) continue;
    HostElement& e = P.el[z];
    e.present = true;
    e.z = z;
    e.n = (int)v("n");
    e.ios = (int)v("ios");
    e.iop = (int)v("iop");
    e.iod = (int)v("iod");
    e.core_charge = v("core_charge");
    e.eheat_ev = v("eheat_kcal") * KCAL_TO_EV;
    e.u_ss = v("u_ss");
    e.u_pp = v("u_pp");
    e.u_dd = v("u_dd");
    e.zeta_s = v("zeta_s");
    e.zeta_p = v("zeta_p");
    e.zeta_d = v("zeta_d");
    e.beta_s = v("beta_s");
    e.beta_p = v("beta_p");
    e.beta_d = v("beta_d");
    e.g_ss = v("g_ss");
    e.g_sp = v("g_sp");
    e.g_pp = v("g_pp");
    e.g_p2 = v("g_p2");
    e.h_sp = v("h_sp");
    e.poc = v("poc");
    e.zsn = v("zsn");
    e.zpn = v("zpn");
    e.zdn = v("zdn");
    e.f0sd = v("f0sd");
    e.g2sd = v("g2sd");
    e.n_orb = e.zeta_d > 1.0e-20 ? 9 : (e.zeta_p > 1.0e-20 ? 4 : 1);
    if (e.n_orb >= 4) {
      double zp = std::max(e.zeta_p, 0.3);
      e.dd = dd_sep(e.n, e.zeta_s, zp);
      e.qq = qq_sep(e.n, zp);
      e.rho1 = rho1_of(std::max(e.h_sp, 1.0e-7), e.dd);
      e.rho2 = rho2_of(std::max(0.5 * (e.g_pp - e.g_p2), 0.1), e.qq);
    }
    e.rho0 = e.g_ss != 0.0 ? EV / (2.0 * e.g_ss) : 0.0;
    for (int s = 1; s <= 4; ++s) {
      char a[8], b[8], c[8];
      std::snprintf(a, sizeof a, "g1_%d", s);
      std::snprintf(b, sizeof b, "g2_%d", s);
      std::snprintf(c, sizeof c, "g3_%d", s);
      double k = v(a), l = v(b), m = v(c);
      if (k != 0.0 || l != 0.0 || m != 0.0) e.gauss.push_back({k, l, m});
    }
    // Isolated-atom reference energy (MOPAC moldat): occupation-weighted one-centre terms.
    const double p = e.iop, l = std::min(p, 6.0 - p);
    e.e_isol = e.u_ss * e.ios + e.u_pp * p + e.u_dd * e.iod + e.g_ss * std::max(e.ios - 1, 0) +
               e.g_pp * (-0.5 * l * (l - 1.0) / 2.0) + e.g_sp * (e.ios * p) +
               e.g_p2 * (p * (p - 1.0) / 2.0 + 0.5 * l * (l - 1.0) / 2.0) + e.h_sp * (-p * e.ios * 0.5);
    // MNDO/d data for every element (a pair in a molecule with a d shell needs both atoms'
    // po/ddp); the partly-filled d-shell term joins the isolated-atom energy.
    derive_mndod(e);
    e.e_isol += e.eiscor;
  }
  std::vector<double> alpb(MAX_Z * MAX_Z, 0.0), xfac(MAX_Z * MAX_Z, 0.0);
  std::vector<char> has(MAX_Z * MAX_Z, 0);
  for (auto& row : tsv(stem + "_pairs.tsv")) {
    if (row.size() < 4 || row[0] == "z_hi") continue;
    int a = std::atoi(row[0].c_str()), b = std::atoi(row[1].c_str());
    if (a <= 0 || b <= 0 || a >= MAX_Z || b >= MAX_Z) continue;
    double al = std::strtod(row[2].c_str(), nullptr), xf = std::strtod(row[3].c_str(), nullptr);
    for (int k = 0; k < 2; ++k) {
      int i = k ? a : b, j = k ? b : a;
      alpb[i * MAX_Z + j] = al;
      xfac[i * MAX_Z + j] = xf;
      has[i * MAX_Z + j] = 1;
    }
  }
  // A missing pair takes the average of the two homonuclear pairs (pm7-rs src/params.rs:172-197).
  P.pair.assign(MAX_Z * MAX_Z, {0.0, 0.0});
  for (int i = 1; i < MAX_Z; ++i)
    for (int j = 1; j < MAX_Z; ++j) {
      int ij = i * MAX_Z + j;
      if (has[ij]) {
        P.pair[ij] = {alpb[ij], xfac[ij]};
      } else {
        int ii = i * MAX_Z + i, jj = j * MAX_Z + j;
        P.pair[ij] = {0.5 * (alpb[ii] + alpb[jj]), 0.5 * (xfac[ii] + xfac[jj])};
      }
    }
  for (auto& row : tsv(stem + "_vpar.tsv")) {
    if (row.size() < 2 || row[0] == "index") continue;
    int i = std::atoi(row[0].c_str());
    if (i >= 1 && i <= 60) P.vpar[i] = std::strtod(row[1].c_str(), nullptr);
  }
  return true;
}

void to_device_table(const ParamSet& P, std::vector<ElemDev>& out) {
  out.assign(MAX_Z, ElemDev{});
  for (int z = 0; z < MAX_Z; ++z) {
    const HostElement& e = P.el[z];
    ElemDev& d = out[z];
    std::memset(&d, 0, sizeof d);
    if (!e.present) continue;
    d.present = 1;
    d.u_ss = e.u_ss;
    d.u_pp = e.u_pp;
    d.beta_s = e.beta_s;
    d.beta_p = e.beta_p;
    d.core_charge = e.core_charge;
    d.g_ss = e.g_ss;
    d.g_sp = e.g_sp;
    d.g_pp = e.g_pp;
    d.g_p2 = e.g_p2;
    d.h_sp = e.h_sp;
    d.zeta_s = e.zeta_s;
    d.zeta_p = e.zeta_p;
    d.dd = e.dd;
    d.qq = e.qq;
    d.rho0 = e.rho0;
    d.rho1 = e.rho1;
    d.rho2 = e.rho2;
    d.po9 = e.poc > 1.0e-5 ? e.poc : e.rho0;
    d.n = e.n;
    d.n_orb = e.n_orb;
    d.z = z;
    d.iod = e.iod;
    d.u_dd = e.u_dd;
    d.beta_d = e.beta_d;
    d.zeta_d = e.zeta_d;
    d.n_sp = shell_principal(z, false);
    d.n_d = shell_principal(z, true);
    for (int k = 0; k < 10; ++k) d.po[k] = e.po[k];
    for (int k = 0; k < 7; ++k) d.ddp[k] = e.ddp[k];
    d.n_gauss = (int)std::min<size_t>(4, e.gauss.size());
    for (int k = 0; k < d.n_gauss; ++k) {
      d.gk[k] = e.gauss[k][0];
      d.gl[k] = e.gauss[k][1];
      d.gm[k] = e.gauss[k][2];
    }
  }
}

}  // namespace pm7

// SPDX-License-Identifier: GPL-3.0-or-later
//
// Per-element MNDO/d data for the spd basis (pm7-rs src/mndod.rs:188-735, MOPAC inid): the 52
// one-centre d integrals, the Klopman-Ohno additive terms po(1..9), the multipole charge
// separations ddp(2..6), the isolated-atom d-shell correction and, for a d-bearing element, the
// 45 x 45 one-centre two-electron matrix the Fock build reads. Computed once per parameter load,
// on the host, in FP64.
#include <algorithm>
#include <cmath>

#include "host.h"

namespace pm7 {
namespace {

inline int pack2(int a, int b) {
  const int h = std::max(a, b), l = std::min(a, b);
  return h * (h + 1) / 2 + l;
}

int shell_n_sp(int z) {  // principal quantum number of the s/p shell (src/mndod.rs:206-216)
  if (z <= 2) return 1;
  if (z <= 10) return 2;
  if (z <= 18) return 3;
  if (z <= 36) return 4;
  if (z <= 54) return 5;
  if (z <= 86) return 6;
  return 0;
}
int shell_n_d(int z) {  // principal quantum number of the d shell (src/mndod.rs:219-227)
  if (z <= 30) return 3;
  if (z <= 48) return 4;
  if (z <= 80) return 5;
  if (z <= 86) return 6;
  return 0;
}
bool is_main_group(int z) {  // false only for the transition metals (src/mndod.rs:230-232)
  return !((z >= 21 && z <= 29) || (z >= 39 && z <= 47) || (z >= 57 && z <= 79));
}

// fx[i] = (i-1)!, bin[i][j] = C(i-1, j-1) (src/mndod.rs:235-258).
struct FactBinom {
  double fx[31] = {0};
  double bin[31][31] = {{0}};
  FactBinom() {
    fx[1] = 1.0;
    for (int i = 2; i <= 30; ++i) fx[i] = fx[i - 1] * (i - 1.0);
    for (int i = 0; i <= 30; ++i) bin[i][1] = 1.0;
    for (int i = 2; i <= 30; ++i)
      for (int j = 2; j <= i; ++j) bin[i][j] = bin[i - 1][j - 1] + bin[i - 1][j];
  }
};
const FactBinom& fbx() {
  static const FactBinom t;
  return t;
}

// Radial Slater-Condon integral R^k (na ea nb eb | nc ec nd ed) in eV (src/mndod.rs:262-331).
double slater_condon(int k, int na, double ea, int nb, double eb, int nc, double ec, int nd, double ed) {
  const FactBinom& T = fbx();
  const double la = std::log(ea), lb = std::log(eb), lc = std::log(ec), ld = std::log(ed);
  const int nab = na + nb, ncd = nc + nd, n = nab + ncd;
  const double ecd = ec + ed, eab = ea + eb, e = ecd + eab;
  const double le = std::log(e), l2 = std::log(2.0), lcd = std::log(ecd), lab = std::log(eab);
  const double ff = T.fx[n] / std::sqrt(T.fx[2 * na + 1] * T.fx[2 * nb + 1] * T.fx[2 * nc + 1] * T.fx[2 * nd + 1]);
  const double c = EV * ff *
                   std::exp(na * la + nb * lb + nc * lc + nd * ld + 0.5 * (la + lb + lc + ld) + l2 * (n + 2.0) -
                            le * n);
  double s0 = 1.0 / e, s1 = 0.0, s2 = 0.0;
  const int m = ncd - k;
  for (int i = 1; i <= m; ++i) {
    s0 *= e / ecd;
    s1 += s0 * (T.bin[ncd - k][i] - T.bin[ncd + k + 1][i]) / T.bin[n][i];
  }
  const int m2 = ncd + k + 1;
  for (int i = m + 1; i <= m2; ++i) {
    s0 *= e / ecd;
    s2 += s0 * T.bin[m2][i] / T.bin[n][i];
  }
  const double s3 = std::exp(le * n - lcd * m2 - lab * (nab - k)) / T.bin[n][m2];
  return c * (s1 - s2 + s3);
}

// Isolated-atom d-shell correction: reference-configuration weights of the radial integrals
// F0(dd) G2(sd) ... per transition metal (src/mndod.rs:480-520), then the weighted sum.
double d_shell_correction(int z, double r016, double r066, double r244, double r266, double r466) {
  struct W { int z, a016, a066, a244, a266, a466; };
  static const W tab[] = {
      {21, 2, 0, 1, 0, 0},    {22, 4, 1, 2, 8, 1},    {23, 6, 3, 3, 15, 8},   {24, 5, 10, 5, 35, 35},
      {25, 10, 10, 5, 35, 35}, {26, 12, 15, 6, 35, 35}, {27, 14, 21, 7, 43, 36}, {28, 16, 28, 8, 50, 43},
      {29, 10, 45, 5, 70, 70}, {39, 2, 0, 1, 0, 0},    {40, 4, 1, 2, 8, 1},    {41, 4, 6, 4, 21, 21},
      {42, 5, 10, 5, 35, 35},  {43, 10, 10, 5, 35, 35}, {44, 7, 21, 5, 43, 36}, {45, 8, 28, 5, 50, 43},
      {46, 0, 45, 0, 70, 70},  {47, 10, 45, 5, 70, 70}, {57, 2, 0, 1, 0, 0},    {71, 2, 0, 1, 0, 0},
      {72, 4, 1, 2, 8, 1},     {73, 6, 3, 3, 15, 8},   {74, 5, 10, 5, 35, 35}, {75, 10, 10, 5, 35, 35},
      {76, 12, 15, 6, 35, 35}, {77, 14, 21, 7, 43, 36}, {78, 9, 36, 5, 56, 56}, {79, 10, 45, 5, 70, 70}};
  for (const W& w : tab)
    if (w.z == z)
      return w.a016 * r016 + w.a066 * r066 - w.a244 * r244 / 5.0 - w.a266 * r266 / 49.0 - w.a466 * r466 / 49.0;
  return 0.0;
}

// The 52 one-centre d integrals from the twelve radial parameters (src/mndod.rs:334-471). The
// square roots are MOPAC's own truncated constants.
void one_centre_d(int z, double zsn, double zpn, double zdn, double f0sd, double g2sd, double repd[53],
                  double& eiscor) {
  const int ns = shell_n_sp(z), nd = shell_n_d(z);
  const double es = zsn, ep = zpn, ed = zdn;
  double r016 = slater_condon(0, ns, es, ns, es, nd, ed, nd, ed);
  const double r036 = slater_condon(0, ns, ep, ns, ep, nd, ed, nd, ed);
  const double r066 = slater_condon(0, nd, ed, nd, ed, nd, ed, nd, ed);
  const double r155 = slater_condon(1, ns, ep, nd, ed, ns, ep, nd, ed);
  const double r125 = slater_condon(1, ns, es, ns, ep, ns, ep, nd, ed);
  double r244 = slater_condon(2, ns, es, nd, ed, ns, es, nd, ed);
  const double r236 = slater_condon(2, ns, ep, ns, ep, nd, ed, nd, ed);
  const double r266 = slater_condon(2, nd, ed, nd, ed, nd, ed, nd, ed);
  const double r234 = slater_condon(2, ns, ep, ns, ep, ns, es, nd, ed);
  const double r246 = slater_condon(2, ns, es, nd, ed, nd, ed, nd, ed);
  const double r355 = slater_condon(3, ns, ep, nd, ed, ns, ep, nd, ed);
  const double r466 = slater_condon(4, nd, ed, nd, ed, nd, ed, nd, ed);
  if (f0sd > 0.001) r016 = f0sd;
  if (g2sd > 0.001) r244 = g2sd;
  eiscor = d_shell_correction(z, r016, r066, r244, r266, r466);
  const double s3 = 1.7320508, s5 = 2.23606797, s15 = 3.87298334;
  double* d = repd;
  for (int i = 0; i < 53; ++i) d[i] = 0.0;
  d[1] = r016;
  d[2] = 2.0 / (3.0 * s5) * r125;
  d[3] = 1.0 / s15 * r125;
  d[4] = 2.0 / (5.0 * s5) * r234;
  d[5] = r036 + 4.0 / 35.0 * r236;
  d[6] = r036 + 2.0 / 35.0 * r236;
  d[7] = r036 - 4.0 / 35.0 * r236;
  d[8] = -1.0 / (3.0 * s5) * r125;
  d[9] = std::sqrt(3.0 / 125.0) * r234;
  d[10] = s3 / 35.0 * r236;
  d[11] = 3.0 / 35.0 * r236;
  d[12] = -1.0 / (5.0 * s5) * r234;
  d[13] = r036 - 2.0 / 35.0 * r236;
  d[14] = -2.0 * s3 / 35.0 * r236;
  d[15] = -d[3];
  d[16] = -d[11];
  d[17] = -d[9];
  d[18] = -d[14];
  d[19] = 1.0 / 5.0 * r244;
  d[20] = 2.0 / (7.0 * s5) * r246;
  d[21] = d[20] / 2.0;
  d[22] = -d[20];
  d[23] = 4.0 / 15.0 * r155 + 27.0 / 245.0 * r355;
  d[24] = 2.0 * s3 / 15.0 * r155 - 9.0 * s3 / 245.0 * r355;
  d[25] = 1.0 / 15.0 * r155 + 18.0 / 245.0 * r355;
  d[26] = -s3 / 15.0 * r155 + 12.0 * s3 / 245.0 * r355;
  d[27] = -s3 / 15.0 * r155 - 3.0 * s3 / 245.0 * r355;
  d[28] = -d[27];
  d[29] = r066 + 4.0 / 49.0 * r266 + 4.0 / 49.0 * r466;
  d[30] = r066 + 2.0 / 49.0 * r266 - 24.0 / 441.0 * r466;
  d[31] = r066 - 4.0 / 49.0 * r266 + 6.0 / 441.0 * r466;
  d[32] = std::sqrt(3.0 / 245.0) * r246;
  d[33] = 1.0 / 5.0 * r155 + 24.0 / 245.0 * r355;
  d[34] = 1.0 / 5.0 * r155 - 6.0 / 245.0 * r355;
  d[35] = 3.0 / 49.0 * r355;
  d[36] = 1.0 / 49.0 * r266 + 30.0 / 441.0 * r466;
  d[37] = s3 / 49.0 * r266 - 5.0 * s3 / 441.0 * r466;
  d[38] = r066 - 2.0 / 49.0 * r266 - 4.0 / 441.0 * r466;
  d[39] = -2.0 * s3 / 49.0 * r266 + 10.0 * s3 / 441.0 * r466;
  d[40] = -d[32];
  d[41] = -d[34];
  d[42] = -d[35];
  d[43] = -d[37];
  d[44] = 3.0 / 49.0 * r266 + 20.0 / 441.0 * r466;
  d[45] = -d[39];
  d[46] = 1.0 / 5.0 * r155 - 3.0 / 35.0 * r355;
  d[47] = -d[46];
  d[48] = 4.0 / 49.0 * r266 + 15.0 / 441.0 * r466;
  d[49] = 3.0 / 49.0 * r266 - 5.0 / 147.0 * r466;
  d[50] = -d[49];
  d[51] = r066 + 4.0 / 49.0 * r266 - 34.0 / 441.0 * r466;
  d[52] = 35.0 / 441.0 * r466;
}

// Multipole coefficient a_l of the charge distribution (n1 z1)(n2 z2) (Thiel-Voityuk Eq. 7;
// src/mndod.rs:523-535).
double multipole_coef(double z1, double z2, int n1, int n2, int l) {
  const FactBinom& T = fbx();
  const double zz = z1 + z2 + 1.0e-20;
  const double a = 2.0 * z1 / zz, b = 2.0 * z2 / zz;
  return T.fx[n1 + n2 + l + 1] / std::sqrt(T.fx[2 * n1 + 1] * T.fx[2 * n2 + 1]) * std::pow(a, n1) * std::sqrt(a) *
         std::pow(b, n2) * std::sqrt(b) * std::pow(2.0, l) / std::pow(zz, l);
}

// Additive term rho of multipole order l whose self-interaction at zero separation equals fg
// (src/mndod.rs:557-609): closed form for l = 0, else a 100-step golden-section search on
// [0.1, 5] Bohr with the squared residual.
double additive_term(int l, double d, double fg) {
  if (l == 0) return 0.5 * EV / fg;
  const double dsq = d * d, ev4 = EV * 0.25, ev8 = EV / 8.0;
  auto resid = [&](double y) {
    double v = l == 1 ? ev4 * (1.0 / y - 1.0 / std::sqrt(y * y + dsq))
                      : ev8 * (1.0 / y - 2.0 / std::sqrt(y * y + dsq * 0.5) + 1.0 / std::sqrt(y * y + dsq));
    return (v - fg) * (v - fg);
  };
  double a1 = 0.1, a2 = 5.0, f1 = 0.0, f2 = 0.0;
  for (int it = 0; it < 100; ++it) {
    const double delta = a2 - a1;
    if (delta < 1.0e-8) break;
    const double y1 = a1 + delta * 0.382, y2 = a1 + delta * 0.618;
    f1 = resid(y1);
    f2 = resid(y2);
    if (f1 < f2) a2 = y2;
    else a1 = y1;
  }
  return f1 >= f2 ? a2 : a1;
}

// The one-centre spd matrix of a d-bearing element (src/mndod.rs:693-735): sp block from
// gss gsp gpp gp2 hsp, d block from repd. Each entry (ij, kl, n) below sets the packed pairs
// (ij, kl) and (kl, ij) to repd[n]; pairs are packed h(h+1)/2 + l over the orbital order
// s px py pz dx2-y2 dxz dz2 dyz dxy.
const short ONE_CENTRE_D[] = {
    0,14,1, 0,20,1, 0,27,1, 0,35,1, 0,44,1, 6,24,2, 1,11,3, 1,18,3,
    1,38,3, 3,31,3, 3,37,3, 6,16,3, 6,30,3, 9,21,4, 9,27,5, 2,14,6,
    2,20,6, 2,44,6, 5,14,6, 5,35,6, 5,44,6, 9,20,6, 9,35,6, 2,35,7,
    5,20,7, 9,14,7, 9,44,7, 1,22,8, 3,23,8, 2,10,9, 4,36,9, 7,15,9,
    8,28,9, 7,26,10, 8,34,10, 4,33,11, 7,19,11, 7,43,11, 8,41,11, 2,21,12,
    5,21,12, 2,27,13, 5,27,13, 2,25,14, 4,42,14, 3,12,15, 8,32,16, 5,10,17,
    5,25,18, 10,10,19, 15,15,19, 21,21,19, 28,28,19, 36,36,19, 21,27,20, 15,26,21,
    20,21,21, 21,35,21, 28,34,21, 10,25,22, 14,21,22, 21,44,22, 36,42,22, 24,24,23,
    16,24,24, 24,30,24, 22,22,25, 23,23,25, 18,22,26, 23,31,26, 11,22,27, 22,38,27,
    23,37,27, 12,23,28, 14,14,29, 20,20,29, 27,27,29, 35,35,29, 44,44,29, 20,27,30,
    27,35,30, 14,27,31, 27,44,31, 10,20,32, 15,19,32, 15,43,32, 28,41,32, 33,36,32,
    11,11,33, 12,12,33, 16,16,33, 18,18,33, 30,30,33, 31,31,33, 37,37,33, 38,38,33,
    11,18,34, 16,30,34, 18,38,34, 31,37,34, 13,13,35, 13,16,35, 17,17,35, 17,29,35,
    17,39,35, 29,29,35, 29,39,35, 39,39,35, 26,26,36, 34,34,36, 19,26,37, 26,43,37,
    34,41,37, 14,20,38, 14,35,38, 20,35,38, 20,44,38, 35,44,38, 20,25,39, 33,42,39,
    10,35,40, 28,32,40, 12,31,41, 13,30,42, 32,34,43, 19,19,44, 32,32,44, 33,33,44,
    41,41,44, 43,43,44, 25,35,45, 11,38,46, 12,37,47, 25,25,48, 42,42,48, 19,43,49,
    32,41,50, 14,44,51, 40,40,52,
};


}  // namespace

int shell_principal(int z, bool d_shell) { return d_shell ? shell_n_d(z) : shell_n_sp(z); }

void derive_mndod(HostElement& e) {
  e.po.fill(0.0);
  e.ddp.fill(0.0);
  e.repd.fill(0.0);
  e.eiscor = 0.0;
  e.oc_d.clear();
  const bool has_d = e.zeta_d > 1.0e-20;
  const int z = e.z;
  double aij[7] = {0};
  if (z >= 3 && !(e.zeta_s * e.zeta_p < 0.01)) {
    const int nsp = shell_n_sp(z);
    aij[2] = multipole_coef(e.zeta_s, e.zeta_p, nsp, nsp, 1);
    aij[3] = multipole_coef(e.zeta_p, e.zeta_p, nsp, nsp, 2);
    if (has_d) {
      const int nd = shell_n_d(z);
      aij[4] = multipole_coef(e.zeta_s, e.zeta_d, nsp, nd, 2);
      aij[5] = multipole_coef(e.zeta_p, e.zeta_d, nsp, nd, 1);
      aij[6] = multipole_coef(e.zeta_d, e.zeta_d, nd, nd, 2);
    }
  }
  double* rp = e.repd.data();
  if (has_d && e.zdn > 1.0e-4) one_centre_d(z, e.zsn, e.zpn, e.zdn, e.f0sd, e.g2sd, rp, e.eiscor);
  double* po = e.po.data();
  double* ddp = e.ddp.data();
  if (e.g_ss > 0.1) po[1] = additive_term(0, 1.0, e.g_ss);
  if (z >= 3) {
    ddp[2] = aij[2] / std::sqrt(12.0);
    po[2] = additive_term(1, ddp[2], e.h_sp);
    po[7] = po[1];
    ddp[3] = std::sqrt(aij[3] * 0.1);
    po[3] = additive_term(2, ddp[3], 0.5 * (e.g_pp - e.g_p2));
    if (has_d) {
      ddp[4] = std::sqrt(aij[4] * std::sqrt(1.0 / 60.0));
      po[4] = additive_term(2, ddp[4], rp[19]);
      ddp[5] = aij[5] / std::sqrt(20.0);
      po[5] = additive_term(1, ddp[5], rp[23] - 1.8 * rp[35]);
      const double fg = 0.2 * (rp[29] + 2.0 * rp[30] + 2.0 * rp[31]);
      po[8] = fg > 1.0e-5 ? additive_term(0, 1.0, fg) : 1.0e5;
      ddp[6] = std::sqrt(aij[6] / 14.0);
      po[6] = additive_term(2, ddp[6], rp[44] - (20.0 / 35.0) * rp[52]);
    }
  }
  if (is_main_group(z) || !has_d) {  // main-group elements keep their validated sp terms
    po[1] = e.rho0;
    if (e.rho1 > 1.0e-5) po[2] = e.rho1;
    if (e.rho2 > 1.0e-5) po[3] = e.rho2;
    po[7] = po[1];
    ddp[2] = e.dd;
    ddp[3] = e.qq * std::sqrt(2.0);
  }
  po[9] = e.poc > 1.0e-5 ? e.poc : po[1];
  if (!has_d) return;
  e.oc_d.assign(45 * 45, 0.0);
  auto set = [&](int a, int b, double v) {
    e.oc_d[a * 45 + b] = v;
    e.oc_d[b * 45 + a] = v;
  };
  set(0, 0, e.g_ss);
  const int pd[3] = {2, 5, 9};  // packed (px px), (py py), (pz pz)
  for (int a = 0; a < 3; ++a) {
    set(pd[a], 0, e.g_sp);
    set(pd[a], pd[a], e.g_pp);
    for (int b = 0; b < a; ++b) set(pd[a], pd[b], e.g_p2);
  }
  for (int p : {1, 3, 6}) set(p, p, e.h_sp);
  for (int p : {4, 7, 8}) set(p, p, 0.5 * (e.g_pp - e.g_p2));
  for (size_t i = 0; i < sizeof(ONE_CENTRE_D) / sizeof(short); i += 3)
    set(ONE_CENTRE_D[i], ONE_CENTRE_D[i + 1], rp[ONE_CENTRE_D[i + 2]]);
}

}  // namespace pm7

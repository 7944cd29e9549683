// SPDX-License-Identifier: GPL-3.0-or-later
//
// CPU FP64 oracle for the PM7 NDDO SCF single point behind pm7-rs `run_pm7`.
//
// TEST INFRASTRUCTURE ONLY: this is the checker, never the product. Only tests/, the smoke entry
// and bench.py's cpu_baseline leg load it. It restates, function by function, the reference's
// algorithm; every routine cites the pm7-rs source it follows (paths relative to the pm7-rs
// tree, e.g. src/scf.rs:1189). The reference itself is Rust and cannot be built in this image,
// so this restatement is pinned instead by the reference's golden vectors
// (tools/oracle/baselines/reference.json), by the reference's unit-test constants and by closed
// forms; see tests/test_oracle_*.py.
//
// Scope of this restatement: the sp Dewar-Sabelli-Klopman path (closed-form Slater overlaps for
// n <= 3, the prolate-spheroidal quadrature for heavier sp shells), the MNDO/d path that every
// molecule holding a d-bearing element takes (one-centre spd integrals, multipole two-centre
// integrals, the spd rotation, the PM7 d-balance, spd overlaps), RHF and UHF, A-DIIS -> CDIIS,
// the level-shift stall guard, core-core repulsion, D2 dispersion, the EH+ hydrogen-bond term, the
// MM corrections, Mulliken charges, the dipole breakdown (point charge, sp and pd hybrid) and the
// heat of formation. Two tables are data that the formulas need as given: the multipole content
// of each local charge distribution (MULTIPOLES) and the one-centre spd integral map
// (ONE_CENTRE_D).

#include "pm7_oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------------
// Constants (src/constants.rs:14-27): CODATA-2018 values MOPAC v23.2.5 uses.
// ---------------------------------------------------------------------------------------------
constexpr double EV = 27.211386245988;          // eV per Hartree
constexpr double A0 = 0.529177210903;           // Bohr radius, Angstrom
constexpr double EV_TO_KCAL = 23.06054783061903;
constexpr double KCAL_TO_EV = 1.0 / EV_TO_KCAL;
constexpr double AU_DIPOLE_TO_DEBYE = 2.541746473;
constexpr double PI = 3.14159265358979323846;

// Atomic masses (u) for the centre-of-mass dipole origin (src/data_tables.rs MASS: standard
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

// ---------------------------------------------------------------------------------------------
// Tiny deterministic parallel-for: static contiguous chunks, each output written by one thread.
// ---------------------------------------------------------------------------------------------
int g_threads = 1;

template <class F>
void parallel_for(long n, F&& body) {
  int t = std::max(1, std::min<int>(g_threads, (int)std::max<long>(1, n / 8)));
  if (t <= 1 || n < 64) {
    for (long i = 0; i < n; ++i) body(i);
    return;
  }
  std::vector<std::thread> pool;
  long chunk = (n + t - 1) / t;
  for (int k = 0; k < t; ++k) {
    long lo = k * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    pool.emplace_back([lo, hi, &body]() {
      for (long i = lo; i < hi; ++i) body(i);
    });
  }
  for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------------------------
// Parameters (src/params.rs:205-432, 604-652)
// ---------------------------------------------------------------------------------------------
// MNDO/d one-centre data of an element (src/mndod.rs:188-204), 1-based slots.
struct DShell {
  double repd[53] = {0};  // one-centre d integrals (eV)
  double po[10] = {0};    // additive terms (Bohr): 1 ss, 2 sp, 3 pp, 4 sd, 5 pd, 6 dd, 7 pp mono, 8 dd mono, 9 core
  double ddp[7] = {0};    // charge separations (Bohr): 2 sp, 3 pp, 4 sd, 5 pd, 6 dd
  double aij[7] = {0};    // multipole coefficients
  double eiscor = 0.0;    // isolated-atom d-shell correction (eV)
  bool has_d = false;
  std::vector<double> oc;  // 45 x 45 one-centre matrix of a d-bearing element
};

struct Element {
  int z = 0, n = 0, ios = 0, iop = 0, iod = 0, n_orb = 0;
  double core_charge = 0, eheat_ev = 0, u_ss = 0, u_pp = 0, u_dd = 0;
  double zeta_s = 0, zeta_p = 0, zeta_d = 0, beta_s = 0, beta_p = 0, beta_d = 0;
  double g_ss = 0, g_sp = 0, g_pp = 0, g_p2 = 0, h_sp = 0;
  double zsn = 0, zpn = 0, zdn = 0, f0sd = 0, g2sd = 0, poc = 0;
  std::vector<std::array<double, 3>> gauss;  // (K, L, M)
  double e_isol = 0, dd = 0, qq = 0, rho0 = 0, rho1 = 0, rho2 = 0;
  DShell ds;
  bool has_p() const { return n_orb >= 4; }
  double po9() const { return poc > 1.0e-5 ? poc : rho0; }  // src/params.rs:80-87
};

struct Params {
  int method = 0;
  std::map<int, Element> el;
  std::map<std::pair<int, int>, std::pair<double, double>> pairs;  // (hi,lo) -> (alpb,xfac)
  double vpar[61] = {0};
  const Element* get(int z) const {
    auto it = el.find(z);
    return it == el.end() ? nullptr : &it->second;
  }
  // src/params.rs:172-197: explicit pair, else average of the two diagonal pairs.
  std::pair<double, double> pair(int zi, int zj) const {
    auto key = std::make_pair(std::max(zi, zj), std::min(zi, zj));
    auto it = pairs.find(key);
    if (it != pairs.end()) return it->second;
    auto ii = pairs.find({zi, zi});
    auto jj = pairs.find({zj, zj});
    std::pair<double, double> a = ii == pairs.end() ? std::make_pair(0.0, 0.0) : ii->second;
    std::pair<double, double> b = jj == pairs.end() ? std::make_pair(0.0, 0.0) : jj->second;
    return {0.5 * (a.first + b.first), 0.5 * (a.second + b.second)};
  }
};

// MOPAC ddpo.F90 charge separations (src/params.rs:602-612).
double dd_charge_sep(int n, double zs, double zp) {
  double nn = n;
  return (2.0 * nn + 1.0) * std::pow(4.0 * zs * zp, nn + 0.5) / std::pow(zs + zp, 2.0 * nn + 2.0) /
         std::sqrt(3.0);
}
double qq_charge_sep(int n, double zp) {
  double nn = n;
  return std::sqrt((4.0 * nn * nn + 6.0 * nn + 2.0) / 20.0) / zp;
}
// MOPAC calpar.F90 fixed 5-step secant (src/params.rs:621-636).
template <class F>
double calpar_secant(double target, double guess, F f) {
  double d1 = guess, d2 = guess + 0.04;
  for (int it = 0; it < 5; ++it) {
    double f1 = f(d1), f2 = f(d2);
    if (std::fabs(f2 - f1) < 1.0e-25) break;
    double d3 = d1 + (d2 - d1) * (target - f1) / (f2 - f1);
    d1 = d2;
    d2 = d3;
  }
  return d2;
}
double signum(double x) { return std::signbit(x) ? -1.0 : 1.0; }
double additive_rho1(double hsp_ev, double dd) {  // src/params.rs:638-643
  double hsp = hsp_ev / EV;
  auto f = [dd](double d) { return 0.5 * d - 0.5 / std::sqrt(4.0 * dd * dd + 1.0 / (d * d)); };
  double guess = std::pow(std::fabs(hsp) / (dd * dd), 1.0 / 3.0) * signum(hsp);
  return 0.5 / calpar_secant(hsp, guess, f);
}
double additive_rho2(double hpp_ev, double qq) {  // src/params.rs:645-652
  double hpp = hpp_ev / EV;
  auto f = [qq](double q) {
    return 0.25 * q - 0.5 / std::sqrt(4.0 * qq * qq + 1.0 / (q * q)) +
           0.25 / std::sqrt(8.0 * qq * qq + 1.0 / (q * q));
  };
  double guess = std::pow(std::fabs(hpp) / (3.0 * std::pow(qq, 4)), 0.2) * signum(hpp);
  return 0.5 / calpar_secant(hpp, guess, f);
}

std::vector<std::vector<std::string>> read_tsv(const std::string& path) {
  std::vector<std::vector<std::string>> rows;
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) return rows;
  char line[4096];
  while (std::fgets(line, sizeof line, fp)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    std::vector<std::string> fields;
    char* save = nullptr;
    for (char* tok = strtok_r(line, "\t\n", &save); tok; tok = strtok_r(nullptr, "\t\n", &save))
      fields.emplace_back(tok);
    rows.push_back(fields);
  }
  std::fclose(fp);
  return rows;
}

std::map<std::string, Params> g_params_cache;

void derive_dshell(Element& e);

const Params* load_params(const char* data_dir, int method) {
  std::string key = std::string(data_dir ? data_dir : ".") + "#" + std::to_string(method);
  auto it = g_params_cache.find(key);
  if (it != g_params_cache.end()) return &it->second;
  std::string prefix = std::string(data_dir ? data_dir : ".") + (method == 1 ? "/pm7ts" : "/pm7");
  auto er = read_tsv(prefix + "_elements.tsv");
  if (er.size() < 2) return nullptr;
  Params P;
  P.method = method;
  const auto& hdr = er[0];
  auto col = [&](const char* name) {
    for (size_t i = 0; i < hdr.size(); ++i)
      if (hdr[i] == name) return (int)i;
    return -1;
  };
  for (size_t r = 1; r < er.size(); ++r) {
    auto v = [&](const char* name) { return std::strtod(er[r][col(name)].c_str(), nullptr); };
    Element e;
    e.z = (int)v("z");
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
    e.zsn = v("zsn");
    e.zpn = v("zpn");
    e.zdn = v("zdn");
    e.f0sd = v("f0sd");
    e.g2sd = v("g2sd");
    e.poc = v("poc");
    e.n_orb = e.zeta_d > 1.0e-20 ? 9 : (e.zeta_p > 1.0e-20 ? 4 : 1);
    if (e.n_orb >= 4) {
      e.dd = dd_charge_sep(e.n, e.zeta_s, std::max(e.zeta_p, 0.3));
      e.qq = qq_charge_sep(e.n, std::max(e.zeta_p, 0.3));
      e.rho1 = additive_rho1(std::max(e.h_sp, 1.0e-7), e.dd);
      e.rho2 = additive_rho2(std::max(0.5 * (e.g_pp - e.g_p2), 0.1), e.qq);
    }
    for (int s = 1; s <= 4; ++s) {
      char a[8], b[8], c[8];
      std::snprintf(a, sizeof a, "g1_%d", s);
      std::snprintf(b, sizeof b, "g2_%d", s);
      std::snprintf(c, sizeof c, "g3_%d", s);
      double k = v(a), l = v(b), m = v(c);
      if (k != 0.0 || l != 0.0 || m != 0.0) e.gauss.push_back({k, l, m});
    }
    // Isolated-atom energy (src/params.rs:268-286); the d-shell eiscor term is added below
    // (src/params.rs:316-340).
    double gssc = std::max(e.ios - 1, 0);
    double p = e.iop;
    double l = std::min(p, 6.0 - p);
    double gspc = e.ios * p;
    double gp2c = p * (p - 1.0) / 2.0 + 0.5 * l * (l - 1.0) / 2.0;
    double gppc = -0.5 * l * (l - 1.0) / 2.0;
    double hspc = -p * e.ios * 0.5;
    e.e_isol = e.u_ss * e.ios + e.u_pp * p + e.u_dd * e.iod + e.g_ss * gssc + e.g_pp * gppc +
               e.g_sp * gspc + e.g_p2 * gp2c + e.h_sp * hspc;
    e.rho0 = e.g_ss != 0.0 ? EV / (2.0 * e.g_ss) : 0.0;
    derive_dshell(e);  // every element: the MNDO/d kernel needs po/ddp of both atoms
    e.e_isol += e.ds.eiscor;
    P.el[e.z] = e;
  }
  for (auto& row : read_tsv(prefix + "_pairs.tsv")) {
    if (row.size() < 4 || row[0] == "z_hi") continue;
    int hi = std::atoi(row[0].c_str()), lo = std::atoi(row[1].c_str());
    P.pairs[{std::max(hi, lo), std::min(hi, lo)}] = {std::strtod(row[2].c_str(), nullptr),
                                                     std::strtod(row[3].c_str(), nullptr)};
  }
  for (auto& row : read_tsv(prefix + "_vpar.tsv")) {
    if (row.size() < 2 || row[0] == "index") continue;
    int idx = std::atoi(row[0].c_str());
    if (idx >= 1 && idx <= 60) P.vpar[idx] = std::strtod(row[1].c_str(), nullptr);
  }
  auto res = g_params_cache.emplace(key, P);
  return &res.first->second;
}

// ---------------------------------------------------------------------------------------------
// Geometry and basis (src/basis.rs:37-62, src/pbc/images.rs:180-199)
// ---------------------------------------------------------------------------------------------
struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

struct Mol {
  std::vector<int> z;
  std::vector<V3> pos;  // Bohr
  std::vector<int> off, norb;
  std::vector<const Element*> el;
  int nao = 0;
  int n() const { return (int)z.size(); }
};

// Matrix helpers, row-major n x n.
using Mat = std::vector<double>;

inline int pack(int a, int b) {
  int h = std::max(a, b), l = std::min(a, b);
  return h * (h + 1) / 2 + l;
}

// ---------------------------------------------------------------------------------------------
// sp two-centre integrals (src/integrals.rs:16-538)
// ---------------------------------------------------------------------------------------------
// PM7 feathering toward the point charge (src/integrals.rs:31-41).
void feather_to_point(double r, double& cfrac, double& point) {
  point = EV / r;
  double r_ang = r * A0;
  cfrac = r_ang < 7.0 ? 1.0 - std::exp(-(r_ang - 7.0) * (r_ang - 7.0) * 0.22) : 0.0;
}
void feather_ri(double* ri, int n, double r, const int* mono, int nmono) {  // :47-56
  double cfrac, point;
  feather_to_point(r, cfrac, point);
  double add = point * (1.0 - cfrac);
  for (int i = 0; i < n; ++i) ri[i] *= cfrac;
  for (int k = 0; k < nmono; ++k) ri[mono[k]] += add;
}
const int XX_MONO[10] = {0, 2, 3, 10, 11, 15, 16, 17, 18, 20};  // :61
const int XH_MONO[3] = {0, 2, 3};                               // :63

// Shortest-arc quaternion rotation taking v onto +x (src/integrals.rs:74-110).
void rotation_to_x(double vx, double vy, double vz, double R[3][3]) {
  double qx = 0.0, qy = vz, qz = -vy, qw = vx + 1.0;
  if (std::fabs(qw) < 1.0e-13) {
    qx = 0.0;
    qy = 0.0;
    qz = 1.0;
    qw = 0.0;
  }
  double inv = 1.0 / std::sqrt(qx * qx + qy * qy + qz * qz + qw * qw);
  qy *= inv;
  qz *= inv;
  qw *= inv;
  R[0][0] = (qy * qy + qz * qz) * -2.0 + 1.0;
  R[0][1] = qz * qw * -2.0;
  R[0][2] = qy * qw * 2.0;
  R[1][0] = qz * qw * 2.0;
  R[1][1] = qz * qz * -2.0 + 1.0;
  R[1][2] = qy * qz * 2.0;
  R[2][0] = qy * qw * -2.0;
  R[2][1] = qy * qz * 2.0;
  R[2][2] = qy * qy * -2.0 + 1.0;
}

inline double isq(double x, double c) { return 1.0 / std::sqrt(x + c); }  // 1/sqrt(x + c)
inline double isq2(double a, double c) { return 1.0 / std::sqrt(a * a + c); }

void local_xh(const Element& ei, const Element& ej, double r, double ri[4]) {  // :273-294
  const double ev1 = EV / 2.0, ev2 = EV / 4.0;
  double da = ei.dd, qa = ei.qq * 2.0;
  double aee = (ei.rho0 + ej.rho0) * (ei.rho0 + ej.rho0);
  double ade = (ei.rho1 + ej.rho0) * (ei.rho1 + ej.rho0);
  double aqe = (ei.rho2 + ej.rho0) * (ei.rho2 + ej.rho0);
  double ee = EV / std::sqrt(r * r + aee);
  double ev1dsqr6 = ev1 / std::sqrt(r * r + aqe);
  ri[0] = ee;
  ri[1] = ev1 * isq2(r + da, ade) - ev1 * isq2(r - da, ade);
  ri[2] = ee + ev2 * isq2(r + qa, aqe) + ev2 * isq2(r - qa, aqe) - ev1dsqr6;
  ri[3] = ee + ev1 / std::sqrt(r * r + qa * qa + aqe) - ev1dsqr6;
  feather_ri(ri, 4, r, XH_MONO, 3);
}

void local_xx(const Element& ei, const Element& ej, double r, double ri[22]) {  // :296-450
  const double ev1 = EV / 2.0, ev2 = EV / 4.0, ev3 = EV / 8.0, ev4 = EV / 16.0;
  double da = ei.dd, db = ej.dd, qa = ei.qq * 2.0, qb = ej.qq * 2.0, qa1 = ei.qq, qb1 = ej.qq;
  auto sq = [](double x) { return x * x; };
  double aee = sq(ei.rho0 + ej.rho0), ade = sq(ei.rho1 + ej.rho0), aqe = sq(ei.rho2 + ej.rho0);
  double aed = sq(ei.rho0 + ej.rho1), aeq = sq(ei.rho0 + ej.rho2), axx = sq(ei.rho1 + ej.rho1);
  double adq = sq(ei.rho1 + ej.rho2), aqd = sq(ei.rho2 + ej.rho1), aqq = sq(ei.rho2 + ej.rho2);
  double r2 = r * r;
  double ee = EV / std::sqrt(r2 + aee);
  double dze = ev1 * isq2(r - da, ade) - ev1 * isq2(r + da, ade);
  double ev1dsqr6 = ev1 / std::sqrt(r2 + aqe);
  double qzze = ev2 * isq2(r - qa, aqe) + ev2 * isq2(r + qa, aqe) - ev1dsqr6;
  double qxxe = ev1 * isq(r2 + qa * qa, aqe) - ev1dsqr6;
  double edz = ev1 * isq2(r + db, aed) - ev1 * isq2(r - db, aed);
  double ev1dsqr12 = ev1 / std::sqrt(r2 + aeq);
  double eqzz = ev2 * isq2(r - qb, aeq) + ev2 * isq2(r + qb, aeq) - ev1dsqr12;
  double eqxx = ev1 * isq(r2 + qb * qb, aeq) - ev1dsqr12;
  double e20 = ev2 * isq2(r + da, adq), e22 = ev2 * isq2(r - da, adq);
  double e24 = ev2 * isq2(r - db, aqd), e26 = ev2 * isq2(r + db, aqd);
  double e36 = ev2 / std::sqrt(r2 + aqq);
  double e39 = ev2 * isq(r2 + qa * qa, aqq), e40 = ev2 * isq(r2 + qb * qb, aqq);
  double e42 = ev3 * isq2(r - qb, aqq), e44 = ev3 * isq2(r + qb, aqq);
  double e46 = ev3 * isq2(r + qa, aqq), e48 = ev3 * isq2(r - qa, aqq);
  ri[0] = ee;
  ri[1] = -dze;
  ri[2] = ee + qzze;
  ri[3] = ee + qxxe;
  ri[4] = -edz;
  ri[5] = ev2 * isq2(r + da - db, axx) + ev2 * isq2(r - da + db, axx) -
          ev2 * isq2(r - da - db, axx) - ev2 * isq2(r + da + db, axx);
  ri[6] = ev1 * isq(r2 + sq(da - db), axx) - ev1 * isq(r2 + sq(da + db), axx);
  ri[7] = -edz + ev3 * isq2(r + qa - db, aqd) - ev3 * isq2(r + qa + db, aqd) +
          ev3 * isq2(r - qa - db, aqd) - ev3 * isq2(r - qa + db, aqd) - e24 + e26;
  ri[8] = -edz - e24 + ev2 * isq(sq(r - db) + qa * qa, aqd) + e26 -
          ev2 * isq(sq(r + db) + qa * qa, aqd);
  ri[9] = ev2 * isq(sq(qa1 - db) + sq(r + qa1), aqd) - ev2 * isq(sq(qa1 - db) + sq(r - qa1), aqd) -
          ev2 * isq(sq(qa1 + db) + sq(r + qa1), aqd) + ev2 * isq(sq(qa1 + db) + sq(r - qa1), aqd);
  ri[10] = ee + eqzz;
  ri[11] = ee + eqxx;
  ri[12] = -dze + ev3 * isq2(r + da - qb, adq) - ev3 * isq2(r - da - qb, adq) +
           ev3 * isq2(r + da + qb, adq) - ev3 * isq2(r - da + qb, adq) + e22 - e20;
  ri[13] = -dze - e20 + ev2 * isq(sq(r + da) + qb * qb, adq) + e22 -
           ev2 * isq(sq(r - da) + qb * qb, adq);
  ri[14] = ev2 * isq(sq(da - qb1) + sq(r - qb1), adq) - ev2 * isq(sq(da - qb1) + sq(r + qb1), adq) -
           ev2 * isq(sq(da + qb1) + sq(r - qb1), adq) + ev2 * isq(sq(da + qb1) + sq(r + qb1), adq);
  ri[15] = ee + eqzz + qzze + ev4 * isq2(r + qa - qb, aqq) + ev4 * isq2(r + qa + qb, aqq) +
           ev4 * isq2(r - qa - qb, aqq) + ev4 * isq2(r - qa + qb, aqq) - e48 - e46 - e42 - e44 +
           e36;
  ri[16] = ee + eqzz + qxxe + ev3 * isq(sq(r - qb) + qa * qa, aqq) +
           ev3 * isq(sq(r + qb) + qa * qa, aqq) - e42 - e44 - e39 + e36;
  ri[17] = ee + eqxx + qzze + ev3 * isq(sq(r + qa) + qb * qb, aqq) +
           ev3 * isq(sq(r - qa) + qb * qb, aqq) - e46 - e48 - e40 + e36;
  double qxxqxx = ev3 * isq(r2 + sq(qa - qb), aqq) + ev3 * isq(r2 + sq(qa + qb), aqq) - e39 - e40 + e36;
  ri[18] = ee + eqxx + qxxe + qxxqxx;
  double m2 = sq(qa1 - qb1), p2 = sq(qa1 + qb1);
  ri[19] = ev3 * isq(sq(r + qa1 - qb1) + m2, aqq) - ev3 * isq(sq(r + qa1 + qb1) + m2, aqq) -
           ev3 * isq(sq(r - qa1 - qb1) + m2, aqq) + ev3 * isq(sq(r - qa1 + qb1) + m2, aqq) -
           ev3 * isq(sq(r + qa1 - qb1) + p2, aqq) + ev3 * isq(sq(r + qa1 + qb1) + p2, aqq) +
           ev3 * isq(sq(r - qa1 - qb1) + p2, aqq) - ev3 * isq(sq(r - qa1 + qb1) + p2, aqq);
  double qxxqyy = ev2 * isq(r2 + qa * qa + qb * qb, aqq) - e39 - e40 + e36;
  ri[20] = ee + eqxx + qxxe + qxxqyy;
  ri[21] = (qxxqxx - qxxqyy) * 0.5;
  feather_ri(ri, 22, r, XX_MONO, 10);
}

// src/integrals.rs:452-464
void build_wxh(const double ri[4], const double* r0, const double* r1, const double* r2,
               double w[10]) {
  w[pack(0, 0)] = ri[0];
  for (int k = 0; k < 3; ++k) w[pack(k + 1, 0)] = ri[1] * r0[k];
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l <= k; ++l)
      w[pack(k + 1, l + 1)] = ri[2] * (r0[k] * r0[l]) + ri[3] * (r1[k] * r1[l] + r2[k] * r2[l]);
}

// src/integrals.rs:466-538: the 10x10 block (row = pack on atom i, col = pack on atom j).
void rotate_xx(const double ri[22], const double* r0, const double* r1, const double* r2,
               double w[100]) {
  int idx = 0;
  for (int kk = 0; kk < 4; ++kk)
    for (int ll = 0; ll <= kk; ++ll)
      for (int mm = 0; mm < 4; ++mm)
        for (int nn = 0; nn <= mm; ++nn) {
          int k = kk - 1, l = ll - 1, m = mm - 1, n = nn - 1;
          double val;
          if (kk == 0) {
            if (mm == 0)
              val = ri[0];
            else if (nn == 0)
              val = ri[4] * r0[m];
            else
              val = ri[10] * (r0[m] * r0[n]) + ri[11] * (r1[m] * r1[n] + r2[m] * r2[n]);
          } else if (ll == 0) {
            if (mm == 0)
              val = ri[1] * r0[k];
            else if (nn == 0)
              val = ri[5] * (r0[k] * r0[m]) + ri[6] * (r1[k] * r1[m] + r2[k] * r2[m]);
            else {
              double t0 = r0[k] * r0[m] * r0[n];
              double t1 = (r1[m] * r1[n] + r2[m] * r2[n]) * r0[k];
              double mix = r1[k] * (r1[n] * r0[m] + r1[m] * r0[n]) + r2[k] * (r2[m] * r0[n] + r2[n] * r0[m]);
              val = ri[12] * t0 + ri[13] * t1 + ri[14] * mix;
            }
          } else if (mm == 0) {
            val = ri[2] * (r0[k] * r0[l]) + ri[3] * (r1[k] * r1[l] + r2[k] * r2[l]);
          } else if (nn == 0) {
            double t0 = r0[k] * r0[l] * r0[m];
            double t1 = (r1[k] * r1[l] + r2[k] * r2[l]) * r0[m];
            double t2 = r1[l] * r1[m] + r2[l] * r2[m];
            val = ri[7] * t0 + ri[8] * t1 + ri[9] * (r0[k] * t2 + r0[l] * (r1[k] * r1[m] + r2[k] * r2[m]));
          } else {
            double t0 = r0[k] * r0[l] * r0[m] * r0[n];
            double v = ri[15] * t0;
            v += ri[16] * ((r1[k] * r1[l] + r2[k] * r2[l]) * r0[m] * r0[n]);
            v += ri[17] * ((r1[m] * r1[n] + r2[m] * r2[n]) * (r0[k] * r0[l]));
            v += ri[18] * (r1[k] * r1[l] * r1[m] * r1[n] + r2[k] * r2[l] * r2[m] * r2[n]);
            double mix1 = r0[m] * (r1[l] * r1[n] + r2[l] * r2[n]);
            double mix2 = r0[n] * (r1[l] * r1[m] + r2[l] * r2[m]);
            double val5 = r0[k] * (mix1 + mix2) +
                          r0[l] * (r0[m] * (r1[k] * r1[n] + r2[k] * r2[n]) + r0[n] * (r1[k] * r1[m] + r2[k] * r2[m]));
            v += ri[19] * val5;
            v += ri[20] * (r1[k] * r1[l] * r2[m] * r2[n] + r2[k] * r2[l] * r1[m] * r1[n]);
            double cross = (r1[k] * r2[l] + r2[k] * r1[l]) * (r1[m] * r2[n] + r2[m] * r1[n]);
            v += ri[21] * cross;
            val = v;
          }
          w[idx++] = val;
        }
}

struct PairBlock {
  int a = 0, b = 0;      // ordered: a has at least as many AOs as b
  int npa = 1, npb = 1;  // packed pair counts (1, 10 or 45)
  std::vector<double> w;  // npa x npb
  double e1b[9][9] = {{0}}, e2a[9][9] = {{0}}, s[9][9] = {{0}};
};

// src/integrals.rs:186-271 for the sp path.
void pair_two_electron(const Element& ei, const Element& ej, V3 d, PairBlock& out) {
  double r = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
  double inv = 1.0 / r;
  double x[3] = {d.x * inv, d.y * inv, d.z * inv};
  bool hi = ei.has_p(), hj = ej.has_p();
  double R[3][3];
  rotation_to_x(-x[0], -x[1], -x[2], R);
  if (!hi && !hj) {
    double aee = (ei.rho0 + ej.rho0) * (ei.rho0 + ej.rho0);
    double ee_nddo = EV / std::sqrt(r * r + aee);
    double cfrac, point;
    feather_to_point(r, cfrac, point);
    double ee = ee_nddo * cfrac + point * (1.0 - cfrac);
    out.npa = out.npb = 1;
    out.w.assign(1, ee);
    out.e1b[0][0] = ee * -ej.core_charge;
    out.e2a[0][0] = ee * -ei.core_charge;
    return;
  }
  if (hi && !hj) {
    double ri[4];
    local_xh(ei, ej, r, ri);
    double w[10];
    build_wxh(ri, R[0], R[1], R[2], w);
    out.npa = 10;
    out.npb = 1;
    out.w.assign(w, w + 10);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) out.e1b[a][b] = w[pack(a, b)] * -ej.core_charge;
    out.e2a[0][0] = w[0] * -ei.core_charge;
    return;
  }
  double ri[22];
  local_xx(ei, ej, r, ri);
  double w[100];
  rotate_xx(ri, R[0], R[1], R[2], w);
  out.npa = out.npb = 10;
  out.w.assign(w, w + 100);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      out.e1b[a][b] = w[pack(a, b) * 10] * -ej.core_charge;
      out.e2a[a][b] = w[pack(a, b)] * -ei.core_charge;
    }
}

// ---------------------------------------------------------------------------------------------
// Slater overlaps (src/overlap.rs:16-40, 120-328, 403-462)
// ---------------------------------------------------------------------------------------------
void aintgs(double x, double a[13]) {
  for (int i = 0; i < 13; ++i) a[i] = 0.0;
  if (std::fabs(x) < 1.0e-12) return;
  double inv = 1.0 / x;
  a[0] = std::exp(-x) * inv;
  for (int i = 1; i < 13; ++i) a[i] = a[0] + a[i - 1] * inv * i;
}
void bintgs(double x, double b[13]) {
  double absx = std::fabs(x);
  if (absx > 0.5) {
    double inv = 1.0 / x;
    double tx = std::exp(x) * inv;
    double tmx = std::exp(-x) * inv * -1.0;
    b[0] = tx + tmx;
    for (int i = 1; i < 13; ++i) {
      double sign = (i % 2 == 0) ? 1.0 : -1.0;
      b[i] = tx * sign + tmx + b[i - 1] * inv * i;
    }
  } else if (absx > 1.0e-6) {
    double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x4 * x, x6 = x4 * x2;
    b[0] = x2 * (1.0 / 3.0) + x4 * (1.0 / 60.0) + x6 * (1.0 / 2520.0) + 2.0;
    b[2] = x2 * (1.0 / 5.0) + x4 * (1.0 / 84.0) + x6 * (1.0 / 3240.0) + 2.0 / 3.0;
    b[4] = x2 * (1.0 / 7.0) + x4 * (1.0 / 108.0) + x6 * (1.0 / 3960.0) + 2.0 / 5.0;
    b[6] = x2 * (1.0 / 9.0) + x4 * (1.0 / 132.0) + x6 * (1.0 / 4680.0) + 2.0 / 7.0;
    b[8] = x2 * (1.0 / 11.0) + x4 * (1.0 / 156.0) + x6 * (1.0 / 5400.0) + 2.0 / 9.0;
    b[10] = x2 * (1.0 / 13.0) + x4 * (1.0 / 180.0) + x6 * (1.0 / 6120.0) + 2.0 / 11.0;
    b[12] = x2 * (1.0 / 15.0) + x4 * (1.0 / 204.0) + x6 * (1.0 / 6840.0) + 2.0 / 13.0;
    b[1] = x * (-2.0 / 3.0) - x3 * (1.0 / 15.0) - x5 * (1.0 / 420.0);
    b[3] = x * (-2.0 / 5.0) - x3 * (1.0 / 21.0) - x5 * (1.0 / 540.0);
    b[5] = x * (-2.0 / 7.0) - x3 * (1.0 / 27.0) - x5 * (1.0 / 660.0);
    b[7] = x * (-2.0 / 9.0) - x3 * (1.0 / 33.0) - x5 * (1.0 / 780.0);
    b[9] = x * (-2.0 / 11.0) - x3 * (1.0 / 39.0) - x5 * (1.0 / 900.0);
    b[11] = x * (-2.0 / 13.0) - x3 * (1.0 / 45.0) - x5 * (1.0 / 1020.0);
  } else {
    for (int i = 0; i < 13; ++i) b[i] = (i % 2 == 0) ? 2.0 / (i + 1.0) : 0.0;
  }
}

// Gauss-Legendre nodes and weights on [-1, 1] by Newton iteration on P_n (src/overlap_numeric.rs:137-175).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double t = std::cos(PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      double dx = -p1 / dp;
      t += dx;
      if (std::fabs(dx) < 1.0e-15) break;
    }
    {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
    }
    const double wt = 2.0 / ((1.0 - t * t) * dp * dp);
    x[i] = -t;
    x[n - 1 - i] = t;
    w[i] = wt;
    w[n - 1 - i] = wt;
  }
}

// Overlap of Slater orbitals (na, la, m; za) on a and (nb, lb, m; zb) on b (b on +z at distance r)
// by quadrature in prolate spheroidal coordinates xi = (ra + rb)/r, eta = (ra - rb)/r:
// dV = (r/2)^3 (xi^2 - eta^2) dxi deta dphi, 48 Gauss-Legendre nodes on xi in [1, 1 + 60/p],
// p = (za + zb) r / 2, and 40 on eta (src/overlap_numeric.rs:45-115).
double sto_overlap(int na, double za, int la, int m, int nb, double zb, int lb, double r) {
  double c;
  if (la == 0 && lb == 0) c = 0.5;
  else if ((la == 1) != (lb == 1)) c = (m == 0) ? std::sqrt(3.0) / 2.0 : 0.0;
  else c = (m == 0) ? 1.5 : 0.75;
  if (c == 0.0) return 0.0;
  auto fact = [](int k) {
    double f = 1.0;
    for (int i = 2; i <= k; ++i) f *= i;
    return f;
  };
  const double norm_a = std::pow(2.0 * za, na + 0.5) / std::sqrt(fact(2 * na));
  const double norm_b = std::pow(2.0 * zb, nb + 0.5) / std::sqrt(fact(2 * nb));
  const double hr = 0.5 * r;
  const double p = (za + zb) * hr;
  const double xi_max = 1.0 + 60.0 / std::max(p, 0.05);
  static std::vector<double> gx48, gw48, gx40, gw40;
  if (gx48.empty()) {
    gauss_legendre(48, gx48, gw48);
    gauss_legendre(40, gx40, gw40);
  }
  const double xm = 0.5 * (1.0 + xi_max), xh = 0.5 * (xi_max - 1.0);
  double sum = 0.0;
  for (int i = 0; i < 48; ++i) {
    const double xi = xm + xh * gx48[i], wx = xh * gw48[i];
    for (int j = 0; j < 40; ++j) {
      const double eta = gx40[j], we = gw40[j];
      const double ra = hr * (xi + eta), rb = hr * (xi - eta);
      if (ra <= 0.0 || rb <= 0.0) continue;
      const double ca = (1.0 + xi * eta) / (xi + eta), cb = (xi * eta - 1.0) / (xi - eta);
      const double ang_a = la == 0 ? 1.0 : (m == 0 ? ca : std::sqrt(std::max(1.0 - ca * ca, 0.0)));
      const double ang_b = lb == 0 ? 1.0 : (m == 0 ? cb : std::sqrt(std::max(1.0 - cb * cb, 0.0)));
      const double radial = std::pow(ra, na - 1) * std::pow(rb, nb - 1) * std::exp(-za * ra - zb * rb);
      sum += radial * (wx * we * ang_a * ang_b * (xi * xi - eta * eta));
    }
  }
  return hr * hr * hr * sum * (norm_a * norm_b * c);
}

// Local-frame overlaps of the numerical path, in the analytic kernel's sign convention
// (src/overlap_numeric.rs:24-43).
void overlap_locals_numeric(const Element& ea, const Element& eb, double r, double s[5]) {
  s[0] = sto_overlap(ea.n, ea.zeta_s, 0, 0, eb.n, eb.zeta_s, 0, r);
  s[1] = sto_overlap(ea.n, ea.zeta_p, 1, 0, eb.n, eb.zeta_s, 0, r);
  s[2] = -sto_overlap(ea.n, ea.zeta_s, 0, 0, eb.n, eb.zeta_p, 1, r);
  s[3] = -sto_overlap(ea.n, ea.zeta_p, 1, 0, eb.n, eb.zeta_p, 1, r);
  s[4] = sto_overlap(ea.n, ea.zeta_p, 1, 1, eb.n, eb.zeta_p, 1, r);
}

// Local-frame sigma/pi overlaps s111 s211 s121 s221 s222 for ea.n >= eb.n: closed forms for
// n <= 3 (src/overlap.rs:120-298), the quadrature above for n >= 4.
bool overlap_locals(const Element& ea, const Element& eb, double r, double s[5]) {
  int jcall;
  if (ea.n == 1 && eb.n == 1) jcall = 2;
  else if (ea.n == 2 && eb.n == 1) jcall = 3;
  else if (ea.n == 2 && eb.n == 2) jcall = 4;
  else if (ea.n == 3 && eb.n == 1) jcall = 431;
  else if (ea.n == 3 && eb.n == 2) jcall = 5;
  else if (ea.n == 3 && eb.n == 3) jcall = 6;
  else {
    overlap_locals_numeric(ea, eb, r, s);
    return true;
  }
  double za[2] = {ea.zeta_s, ea.zeta_p}, zb[2] = {eb.zeta_s, eb.zeta_p};
  double a111[13], b111[13], a211[13], b211[13], a121[13], b121[13], a22[13], b22[13];
  aintgs(r * (0.5 * (za[0] + zb[0])), a111);
  bintgs(r * (0.5 * (za[0] - zb[0])), b111);
  aintgs(r * (0.5 * (za[1] + zb[0])), a211);
  bintgs(r * (0.5 * (za[1] - zb[0])), b211);
  aintgs(r * (0.5 * (za[0] + zb[1])), a121);
  bintgs(r * (0.5 * (za[0] - zb[1])), b121);
  aintgs(r * (0.5 * (za[1] + zb[1])), a22);
  bintgs(r * (0.5 * (za[1] - zb[1])), b22);
  double s111 = 0, s211 = 0, s121 = 0, s221 = 0, s222 = 0;
  const double sq3 = std::sqrt(3.0);
  auto p = [](double x, double e) { return std::pow(x, e); };
  auto pi = [](double x, int k) {
    double v = 1.0;
    for (int i = 0; i < k; ++i) v *= x;
    return v;
  };
  switch (jcall) {
    case 2:
      s111 = p(r * r * (za[0] * zb[0]), 1.5) * (a111[2] * b111[0] - b111[2] * a111[0]) / 4.0;
      break;
    case 3:
      s111 = pi(r, 4) * (p(zb[0], 1.5) * p(za[0], 2.5)) *
             (a111[3] * b111[0] - b111[3] * a111[0] + a111[2] * b111[1] - b111[2] * a111[1]) / (sq3 * 8.0);
      s211 = pi(r, 4) * (p(zb[0], 1.5) * p(za[1], 2.5)) *
             (a211[2] * b211[0] - b211[2] * a211[0] + a211[3] * b211[1] - b211[3] * a211[1]) / 8.0;
      break;
    case 4: {
      s111 = pi(r, 5) * p(zb[0] * za[0], 2.5) * (a111[4] * b111[0] + b111[4] * a111[0] - a111[2] * b111[2] * 2.0) / 48.0;
      s211 = pi(r, 5) * p(zb[0] * za[1], 2.5) *
             (a211[3] * (b211[0] - b211[2]) - a211[1] * (b211[2] - b211[4]) + b211[3] * (a211[0] - a211[2]) -
              b211[1] * (a211[2] - a211[4])) / (16.0 * sq3);
      s121 = pi(r, 5) * p(zb[1] * za[0], 2.5) *
             (a121[3] * (b121[0] - b121[2]) - a121[1] * (b121[2] - b121[4]) - b121[3] * (a121[0] - a121[2]) +
              b121[1] * (a121[2] - a121[4])) / (16.0 * sq3);
      double w = pi(r, 5) * p(zb[1] * za[1], 2.5) / 16.0;
      s221 = -(w * (b22[2] * (a22[4] + a22[0]) - a22[2] * (b22[4] + b22[0])));
      s222 = w * (a22[4] * (b22[0] - b22[2]) - b22[4] * (a22[0] - a22[2]) - a22[2] * b22[0] + b22[2] * a22[0]) * 0.5;
      break;
    }
    case 431:
      s111 = pi(r, 5) * (p(zb[0], 1.5) * p(za[0], 3.5)) *
             (a111[4] * b111[0] + b111[1] * a111[3] * 2.0 - a111[1] * b111[3] * 2.0 - b111[4] * a111[0]) /
             (std::sqrt(10.0) * 24.0);
      s211 = pi(r, 5) * (p(zb[0], 1.5) * p(za[1], 3.5)) *
             (a211[3] * (b211[0] + b211[2]) - a211[1] * (b211[4] + b211[2]) + b211[1] * (a211[2] + a211[4]) -
              b211[3] * (a211[2] + a211[0])) / (8.0 * std::sqrt(30.0));
      break;
    case 5:
      s111 = pi(r, 6) * (p(zb[0], 2.5) * p(za[0], 3.5)) *
             (a111[5] * b111[0] + b111[1] * a111[4] - b111[2] * a111[3] * 2.0 - a111[2] * b111[3] * 2.0 +
              b111[4] * a111[1] + b111[5] * a111[0]) / (std::sqrt(30.0) * 48.0);
      s211 = pi(r, 6) * (p(zb[0], 2.5) * p(za[1], 3.5)) *
             (a211[4] * b211[0] + b211[1] * a211[5] - b211[3] * a211[3] * 2.0 - a211[2] * b211[2] * 2.0 +
              a211[1] * b211[5] + a211[0] * b211[4]) / (48.0 * std::sqrt(10.0));
      s121 = pi(r, 6) * (p(zb[1], 2.5) * p(za[0], 3.5)) *
             ((a121[4] * b121[0] - a121[5] * b121[1]) + (a121[3] * b121[1] - a121[4] * b121[2]) * 2.0 -
              (a121[1] * b121[3] - a121[2] * b121[4]) * 2.0 - (a121[0] * b121[4] - a121[1] * b121[5])) /
             (48.0 * std::sqrt(10.0));
      s221 = pi(r, 6) * (p(zb[1], 2.5) * p(za[1], 3.5)) *
             ((a22[3] * b22[0] - a22[5] * b22[2]) + (a22[2] * b22[1] - a22[4] * b22[3]) -
              (a22[1] * b22[2] - a22[3] * b22[4]) - (a22[0] * b22[3] - a22[2] * b22[5])) / (16.0 * std::sqrt(30.0));
      s222 = pi(r, 6) * (p(zb[1], 2.5) * p(za[1], 3.5)) *
             ((a22[5] - a22[3]) * (b22[0] - b22[2]) + (a22[4] - a22[2]) * (b22[1] - b22[3]) -
              (a22[3] - a22[1]) * (b22[2] - b22[4]) - (a22[2] - a22[0]) * (b22[3] - b22[5])) /
             (32.0 * std::sqrt(30.0));
      break;
    case 6:
      s111 = pi(r, 7) * p(zb[0] * za[0], 3.5) *
             (a111[6] * b111[0] - b111[2] * a111[4] * 3.0 + a111[2] * b111[4] * 3.0 - a111[0] * b111[6]) / 1440.0;
      s211 = pi(r, 7) * p(zb[0] * za[1], 3.5) *
             ((a211[5] * b211[0] + a211[6] * b211[1]) + (-a211[4] * b211[1] - a211[5] * b211[2]) -
              (a211[3] * b211[2] + a211[4] * b211[3]) * 2.0 - (-a211[2] * b211[3] - a211[3] * b211[4]) * 2.0 +
              (a211[1] * b211[4] + a211[2] * b211[5]) + (-a211[0] * b211[5] - a211[1] * b211[6])) /
             (480.0 * sq3);
      s121 = pi(r, 7) * p(zb[1] * za[0], 3.5) *
             ((a121[5] * b121[0] - a121[6] * b121[1]) + (a121[4] * b121[1] - a121[5] * b121[2]) -
              (a121[3] * b121[2] - a121[4] * b121[3]) * 2.0 - (a121[2] * b121[3] - a121[3] * b121[4]) * 2.0 +
              (a121[1] * b121[4] - a121[2] * b121[5]) + (a121[0] * b121[5] - a121[1] * b121[6])) /
             (480.0 * sq3);
      s221 = pi(r, 7) * (p(zb[1], 3.5) * p(za[1], 3.5)) *
             ((a22[4] * b22[0] - a22[6] * b22[2]) - (a22[2] * b22[2] - a22[4] * b22[4]) * 2.0 +
              (a22[0] * b22[4] - a22[2] * b22[6])) / 480.0;
      s222 = pi(r, 7) * (p(zb[1], 3.5) * p(za[1], 3.5)) *
             ((a22[6] - a22[4]) * (b22[0] - b22[2]) - (a22[4] - a22[2]) * (b22[2] - b22[4]) * 2.0 +
              (a22[2] - a22[0]) * (b22[4] - b22[6])) / 960.0;
      break;
  }
  s[0] = s111;
  s[1] = s211;
  s[2] = s121;
  s[3] = s221;
  s[4] = s222;
  return true;
}

// src/overlap.rs:290-328
void build_di(const double s[5], const double dir[3], double di[4][4]) {
  double rr[3][3];
  rotation_to_x(dir[0], dir[1], dir[2], rr);
  double rot[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) rot[a][b] = rr[b][a];
  di[0][0] = s[0];
  for (int a = 0; a < 3; ++a) {
    di[1 + a][0] = s[1] * rot[a][0];
    di[0][1 + a] = -s[2] * rot[a][0];
  }
  double mm[3] = {-s[3], s[4], s[4]};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double v = 0.0;
      for (int k = 0; k < 3; ++k) v += rot[a][k] * mm[k] * rot[b][k];
      di[1 + a][1 + b] = v;
    }
}

// src/overlap.rs:16-40 with pos_i at the origin and pos_j = d.
bool diatom_overlap(const Element& ei, const Element& ej, V3 d, double out[4][4]) {
  double r = norm(d);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) out[a][b] = (a == b) ? 1.0 : 0.0;
  if (r < 1.0e-8) return true;
  double x[3] = {d.x / r, d.y / r, d.z / r};
  double s[5];
  double di[4][4];
  if (ei.n >= ej.n) {
    if (!overlap_locals(ei, ej, r, s)) return false;
    build_di(s, x, di);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) out[a][b] = di[a][b];
  } else {
    if (!overlap_locals(ej, ei, r, s)) return false;
    double mx[3] = {-x[0], -x[1], -x[2]};
    build_di(s, mx, di);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) out[a][b] = di[b][a];
  }
  return true;
}

// ---------------------------------------------------------------------------------------------
// MNDO/d one-centre machinery (src/mndod.rs:188-735): per element, the 52 one-centre d integrals
// repd, the Klopman-Ohno additive terms po(1..9), the multipole charge separations ddp(2..6), the
// isolated-atom d-shell correction eiscor and, for a d-bearing element, the 45 x 45 one-centre
// matrix. Arrays keep the 1-based slots of the formulas (slot 0 unused).
// ---------------------------------------------------------------------------------------------
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

void derive_dshell(Element& e) {  // src/mndod.rs:612-690 (MOPAC inid for one element)
  DShell& D = e.ds;
  D = DShell();
  D.has_d = e.zeta_d > 1.0e-20;
  const int z = e.z;
  if (z >= 3 && !(e.zeta_s * e.zeta_p < 0.01)) {  // src/mndod.rs:538-555
    const int nsp = shell_n_sp(z);
    D.aij[2] = multipole_coef(e.zeta_s, e.zeta_p, nsp, nsp, 1);
    D.aij[3] = multipole_coef(e.zeta_p, e.zeta_p, nsp, nsp, 2);
    if (D.has_d) {
      const int nd = shell_n_d(z);
      D.aij[4] = multipole_coef(e.zeta_s, e.zeta_d, nsp, nd, 2);
      D.aij[5] = multipole_coef(e.zeta_p, e.zeta_d, nsp, nd, 1);
      D.aij[6] = multipole_coef(e.zeta_d, e.zeta_d, nd, nd, 2);
    }
  }
  if (D.has_d && e.zdn > 1.0e-4) one_centre_d(z, e.zsn, e.zpn, e.zdn, e.f0sd, e.g2sd, D.repd, D.eiscor);
  const double* rp = D.repd;
  if (e.g_ss > 0.1) D.po[1] = additive_term(0, 1.0, e.g_ss);
  if (z >= 3) {
    D.ddp[2] = D.aij[2] / std::sqrt(12.0);
    D.po[2] = additive_term(1, D.ddp[2], e.h_sp);
    D.po[7] = D.po[1];
    D.ddp[3] = std::sqrt(D.aij[3] * 0.1);
    D.po[3] = additive_term(2, D.ddp[3], 0.5 * (e.g_pp - e.g_p2));
    if (D.has_d) {
      D.ddp[4] = std::sqrt(D.aij[4] * std::sqrt(1.0 / 60.0));
      D.po[4] = additive_term(2, D.ddp[4], rp[19]);
      D.ddp[5] = D.aij[5] / std::sqrt(20.0);
      D.po[5] = additive_term(1, D.ddp[5], rp[23] - 1.8 * rp[35]);
      const double fg = 0.2 * (rp[29] + 2.0 * rp[30] + 2.0 * rp[31]);
      D.po[8] = fg > 1.0e-5 ? additive_term(0, 1.0, fg) : 1.0e5;
      D.ddp[6] = std::sqrt(D.aij[6] / 14.0);
      D.po[6] = additive_term(2, D.ddp[6], rp[44] - (20.0 / 35.0) * rp[52]);
    }
  }
  if (is_main_group(z) || !D.has_d) {  // the validated sp terms win for main-group elements
    D.po[1] = e.rho0;
    if (e.rho1 > 1.0e-5) D.po[2] = e.rho1;
    if (e.rho2 > 1.0e-5) D.po[3] = e.rho2;
    D.po[7] = D.po[1];
    D.ddp[2] = e.dd;
    D.ddp[3] = e.qq * std::sqrt(2.0);
  }
  D.po[9] = e.poc > 1.0e-5 ? e.poc : D.po[1];
  if (!D.has_d) return;
  D.oc.assign(45 * 45, 0.0);
  auto set = [&](int a, int b, double v) { D.oc[a * 45 + b] = v; D.oc[b * 45 + a] = v; };
  set(0, 0, e.g_ss);
  const int pd[3] = {2, 5, 9};  // packed (px px), (py py), (pz pz)
  for (int a = 0; a < 3; ++a) {
    set(pd[a], 0, e.g_sp);
    set(pd[a], pd[a], e.g_pp);
    for (int b = 0; b < a; ++b) set(pd[a], pd[b], e.g_p2);
  }
  for (int p : {1, 3, 6}) set(p, p, e.h_sp);                   // (s p | s p)
  for (int p : {4, 7, 8}) set(p, p, 0.5 * (e.g_pp - e.g_p2));  // (p p' | p p')
  for (size_t i = 0; i < sizeof(ONE_CENTRE_D) / sizeof(short); i += 3)
    set(ONE_CENTRE_D[i], ONE_CENTRE_D[i + 1], rp[ONE_CENTRE_D[i + 2]]);
}
// ---------------------------------------------------------------------------------------------
// MNDO/d two-centre integrals (src/mndod_twocenter.rs). Every pair of a molecule holding a
// d-bearing atom goes through here (src/hamiltonian.rs:347-354, 424-428). Local frame: first
// p orbital along the bond (sigma), then the two pi; d orbitals d1 = d-sigma, d2 d3 = d-pi (with
// the first and second pi), d4 d5 = d-delta. Local and molecular charge distributions are both
// packed h(h+1)/2 + l over (s, p1, p2, p3, d1..d5).
// ---------------------------------------------------------------------------------------------
// Interaction of two point-charge multipole configurations in Hartree (src/mndod.rs:30-186):
// order l1 on the first atom, l2 on the second, azimuthal |m|, separations da db, additive add.
double multipole_pair(double r, int l1, int l2, int m, double da, double db, double add) {
  // q(x, s2) = 1 / sqrt(x^2 + (s2 + add))
  auto q = [add](double x, double s2) { return 1.0 / std::sqrt(x * x + (s2 + add)); };
  const double dm2 = (da - db) * (da - db), dp2 = (da + db) * (da + db);
  if (l1 == 0 && l2 == 0) return q(r, 0.0);
  if (l1 == 1 && l2 == 0) return (q(r - da, 0.0) - q(r + da, 0.0)) * 0.5;
  if (l1 == 0 && l2 == 1) return (q(r + db, 0.0) - q(r - db, 0.0)) * 0.5;
  if (l1 == 1 && l2 == 1) {
    if (m == 0) return (q(r + da - db, 0.0) + q(r - da + db, 0.0) - q(r - da - db, 0.0) - q(r + da + db, 0.0)) * 0.25;
    return (q(r, dm2) - q(r, dp2)) * 0.5;
  }
  if (l1 == 0 && l2 == 2) return (q(r - db, 0.0) - q(r, db * db) * 2.0 + q(r + db, 0.0)) * 0.25;
  if (l1 == 2 && l2 == 0) return (q(r - da, 0.0) - q(r, da * da) * 2.0 + q(r + da, 0.0)) * 0.25;
  if (l1 == 1 && l2 == 2) {
    if (m == 0)
      return (q(r - da - db, 0.0) - q(r - da, db * db) * 2.0 + q(r + db - da, 0.0) - q(r - db + da, 0.0) +
              q(r + da, db * db) * 2.0 - q(r + da + db, 0.0)) * 0.125;
    const double ab = db / std::sqrt(2.0), u = (da - ab) * (da - ab), v = (da + ab) * (da + ab);
    return (q(r - ab, u) * -2.0 + q(r + ab, u) * 2.0 + q(r - ab, v) * 2.0 - q(r + ab, v) * 2.0) * 0.125;
  }
  if (l1 == 2 && l2 == 1) {
    if (m == 0)
      return (-q(r - da - db, 0.0) + q(r - db, da * da) * 2.0 - q(r + da - db, 0.0) + q(r - da + db, 0.0) -
              q(r + db, da * da) * 2.0 + q(r + da + db, 0.0)) * 0.125;
    const double aa = da / std::sqrt(2.0), u = (aa - db) * (aa - db), v = (aa + db) * (aa + db);
    return (q(r + aa, u) * -2.0 + q(r - aa, u) * 2.0 + q(r + aa, v) * 2.0 - q(r - aa, v) * 2.0) * 0.125;
  }
  // l1 == l2 == 2
  const double xyxy = q(r, dm2) * 4.0 + q(r, dp2) * 4.0 - q(r, da * da + db * db) * 8.0;
  if (m == 0) {
    const double zzzz = q(r - da - db, 0.0) + q(r + da + db, 0.0) + q(r - da + db, 0.0) + q(r + da - db, 0.0) -
                        q(r - da, db * db) * 2.0 - q(r - db, da * da) * 2.0 - q(r + da, db * db) * 2.0 -
                        q(r + db, da * da) * 2.0 + q(r, dm2) * 2.0 + q(r, dp2) * 2.0;
    return zzzz * (1.0 / 16.0) - xyxy * (1.0 / 64.0);
  }
  if (m == 1) {
    const double aa = da / std::sqrt(2.0), ab = db / std::sqrt(2.0);
    const double u = (aa - ab) * (aa - ab), v = (aa + ab) * (aa + ab);
    return (q(r + aa - ab, u) * 2.0 - q(r + aa + ab, u) * 2.0 - q(r - aa - ab, u) * 2.0 + q(r - aa + ab, u) * 2.0 -
            q(r + aa - ab, v) * 2.0 + q(r + aa + ab, v) * 2.0 + q(r - aa - ab, v) * 2.0 - q(r - aa + ab, v) * 2.0) *
           (1.0 / 16.0);
  }
  return xyxy * (1.0 / 16.0);
}

// Multipole content of each local charge distribution (a >= b): coefficient c of the (l, m)
// multipole, truncated at l = 2 as in MNDO/d (src/mndod_tables.rs CH; the decimals are MOPAC's).
// Local orbitals 0 s, 1 p-sigma, 2 p-pi, 3 p-pi', 4 d-sigma, 5 d-pi, 6 d-pi', 7 d-delta, 8 d-delta'.
struct MultipoleTerm {
  signed char a, b, l, m;
  double c;
};
const MultipoleTerm MULTIPOLES[] = {
    {0, 0, 0, 0, 1.0},         {1, 0, 1, 0, 1.0},        {2, 0, 1, 1, 1.0},         {3, 0, 1, -1, 1.0},
    {4, 0, 2, 0, 1.1547005},   {5, 0, 2, 1, 1.0},        {6, 0, 2, -1, 1.0},        {7, 0, 2, 2, 1.0},
    {8, 0, 2, -2, 1.0},        {1, 1, 0, 0, 1.0},        {1, 1, 2, 0, 1.3333333},   {2, 1, 2, 1, 1.0},
    {3, 1, 2, -1, 1.0},        {4, 1, 1, 0, 1.1547005},  {5, 1, 1, 1, 1.0},         {6, 1, 1, -1, 1.0},
    {2, 2, 0, 0, 1.0},         {2, 2, 2, 0, -0.66666667}, {2, 2, 2, 2, 1.0},        {3, 2, 2, -2, 1.0},
    {4, 2, 1, 1, -0.57735027}, {5, 2, 1, 0, 1.0},        {7, 2, 1, 1, 1.0},         {8, 2, 1, -1, 1.0},
    {3, 3, 0, 0, 1.0},         {3, 3, 2, 0, -0.66666667}, {3, 3, 2, 2, -1.0},       {4, 3, 1, -1, -0.57735027},
    {6, 3, 1, 0, 1.0},         {7, 3, 1, -1, -1.0},      {8, 3, 1, 1, 1.0},         {4, 4, 0, 0, 1.0},
    {4, 4, 2, 0, 1.3333333},   {5, 4, 2, 1, 0.57735027}, {6, 4, 2, -1, 0.57735027}, {7, 4, 2, 2, -1.1547005},
    {8, 4, 2, -2, -1.1547005}, {5, 5, 0, 0, 1.0},        {5, 5, 2, 0, 0.66666667},  {5, 5, 2, 2, 1.0},
    {6, 5, 2, -2, 1.0},        {7, 5, 2, 1, 1.0},        {8, 5, 2, -1, 1.0},        {6, 6, 0, 0, 1.0},
    {6, 6, 2, 0, 0.66666667},  {6, 6, 2, 2, -1.0},       {7, 6, 2, -1, -1.0},       {8, 6, 2, 1, 1.0},
    {7, 7, 0, 0, 1.0},         {7, 7, 2, 0, -1.3333333}, {8, 8, 0, 0, 1.0},         {8, 8, 2, 0, -1.3333333}};

struct MultipoleTable {
  double c[45][3][5] = {{{0}}};  // [packed pair][l][m + 2]
  MultipoleTable() {
    for (const MultipoleTerm& t : MULTIPOLES) c[pack(t.a, t.b)][t.l][t.m + 2] = t.c;
  }
};
const MultipoleTable& multipoles() {
  static const MultipoleTable t;
  return t;
}

inline int ang_l(int orb) { return orb == 0 ? 0 : (orb <= 3 ? 1 : 2); }
inline int pair_kind(int la, int lb) {  // 1 ss, 2 sp, 3 pp, 4 sd, 5 pd, 6 dd
  int h = std::max(la, lb) + 1, l = std::min(la, lb) + 1;
  return h * (h - 1) / 2 + l;
}

// One local two-centre integral (ab | cd) in Hartree by the multipole expansion
// (src/mndod_twocenter.rs:582-654). core_side 1 (2): the monopole of an (ss) distribution on the
// first (second) atom takes the core additive term po(9).
double local_multipole(const DShell& A, const DShell& B, int a, int b, int c, int d, int core_side, double r) {
  const MultipoleTable& T = multipoles();
  const int la = ang_l(a), lb = ang_l(b), lc = ang_l(c), ld = ang_l(d);
  const int kij = pair_kind(la, lb), kkl = pair_kind(lc, ld);
  const int pij = pack(a, b), pkl = pack(c, d);
  auto mono = [](const DShell& D, int kind, bool core) {
    if (kind == 1) return core ? D.po[9] : D.po[1];
    if (kind == 3) return D.po[7];
    if (kind == 6) return D.po[8];
    return D.po[1];
  };
  double sum = 0.0;
  for (int l1 = std::min(std::abs(la - lb), 2); l1 <= std::min(la + lb, 2); ++l1) {
    const double dij = l1 == 0 ? 0.0 : A.ddp[kij];
    const double aij = l1 == 0 ? mono(A, kij, core_side == 1) : A.po[kij];
    for (int l2 = std::min(std::abs(lc - ld), 2); l2 <= std::min(lc + ld, 2); ++l2) {
      const double dkl = l2 == 0 ? 0.0 : B.ddp[kkl];
      const double akl = l2 == 0 ? mono(B, kkl, core_side == 2) : B.po[kkl];
      const double add = (aij + akl) * (aij + akl);
      const int mm = std::min(l1, l2);
      for (int m = -mm; m <= mm; ++m) {
        const double ccc = T.c[pij][l1][m + 2] * T.c[pkl][l2][m + 2];
        if (ccc == 0.0) continue;
        sum += multipole_pair(r, l1, l2, std::abs(m), dij, dkl, add) * ccc;
      }
    }
  }
  return sum;
}

// Local (ab | cd) over the s/p distributions (packed 0..9) as an index into the 22 sp integrals
// ri[1..22] (0: zero by symmetry); the standard MNDO layout (src/mndod_tables.rs IND2 with
// src/mndod_twocenter.rs:31-37 IPOS).
const signed char SP_LOCAL[10][10] = {
    {1, 5, 11, 0, 0, 12, 0, 0, 0, 12},   {2, 6, 13, 0, 0, 14, 0, 0, 0, 14},
    {3, 8, 16, 0, 0, 18, 0, 0, 0, 18},   {0, 0, 0, 7, 15, 0, 0, 0, 0, 0},
    {0, 0, 0, 10, 20, 0, 0, 0, 0, 0},    {4, 9, 17, 0, 0, 19, 0, 0, 0, 21},
    {0, 0, 0, 0, 0, 0, 7, 15, 0, 0},     {0, 0, 0, 0, 0, 0, 10, 20, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 22, 0},     {4, 9, 17, 0, 0, 21, 0, 0, 0, 19}};

// Sign of each sp integral in the MNDO/d local frame relative to the sp kernel's ri
// (src/mndod_twocenter.rs:39-43 NRI).
const double SP_SIGN[23] = {0, 1, -1, 1, 1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1};

// The 22 sp local integrals of the MNDO/d path (src/mndod_twocenter.rs:244-470): the sp kernel's
// formulas with the po(1..3) additive terms, feathered, in the MNDO/d sign convention (1-based).
void local_sp_ri(const Element& ei, const Element& ej, double r, double ri[23]) {
  Element a = ei, b = ej;
  a.rho0 = ei.ds.po[1], a.rho1 = ei.ds.po[2], a.rho2 = ei.ds.po[3];
  b.rho0 = ej.ds.po[1], b.rho1 = ej.ds.po[2], b.rho2 = ej.ds.po[3];
  for (int k = 0; k < 23; ++k) ri[k] = 0.0;
  const bool si = ei.has_p(), sj = ej.has_p();
  if (si && sj) {
    double t[22];
    local_xx(a, b, r, t);
    for (int k = 0; k < 22; ++k) ri[k + 1] = t[k];
  } else if (si) {
    double t[4];
    local_xh(a, b, r, t);
    for (int k = 0; k < 4; ++k) ri[k + 1] = t[k];
  } else if (sj) {
    double t[4];
    local_xh(b, a, r, t);  // the mirrored pair: (ss|ss) (ss|ps) (ss|pp) (ss|pi pi)
    ri[1] = t[0];
    ri[5] = -t[1];
    ri[11] = t[2];
    ri[12] = t[3];
  } else {
    double aee = (a.rho0 + b.rho0) * (a.rho0 + b.rho0);
    double cfrac, point;
    feather_to_point(r, cfrac, point);
    ri[1] = EV / std::sqrt(r * r + aee) * cfrac + point * (1.0 - cfrac);
  }
  for (int k = 1; k <= 22; ++k) ri[k] *= SP_SIGN[k];
}

// Orbital rotation R[local][molecular] for the pair direction v = R_j - R_i
// (src/mndod_twocenter.rs:73-147): the p block takes the local (sigma, pi, pi') onto x y z, the d
// block the local (d-sigma, d-pi, d-pi', d-delta, d-delta') onto dx2-y2 dxz dz2 dyz dxy.
double orbital_rotation(V3 v, bool with_d, double R[9][9]) {
  const double pt5sq3 = 0.8660254037841;
  const double b = v.x * v.x + v.y * v.y;
  const double r = std::sqrt(b + v.z * v.z);
  const double sqb = std::sqrt(b);
  double sb = sqb / r, ca, sa, cb;
  if (sb > 1.0e-7) {
    ca = v.x / sqb;
    sa = v.y / sqb;
    cb = v.z / r;
  } else {
    sa = 0.0;
    sb = 0.0;
    ca = cb = v.z < 0.0 ? -1.0 : (v.z > 0.0 ? 1.0 : 0.0);
  }
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) R[i][j] = 0.0;
  R[0][0] = 1.0;
  double p[3][3] = {{ca * sb, sa * sb, cb}, {ca * cb, sa * cb, -sb}, {-sa, ca, 0.0}};
  for (int k = 0; k < 3; ++k)
    for (int K = 0; K < 3; ++K) R[1 + k][1 + K] = p[k][K];
  if (with_d) {
    const double c2a = ca * ca * 2.0 - 1.0, c2b = cb * cb * 2.0 - 1.0, s2a = sa * ca * 2.0, s2b = sb * cb * 2.0;
    const double d[5][5] = {
        {c2a * sb * sb * pt5sq3, ca * s2b * pt5sq3, cb * cb - sb * sb * 0.5, sa * s2b * pt5sq3, s2a * sb * sb * pt5sq3},
        {c2a * s2b * 0.5, ca * c2b, -(s2b * pt5sq3), sa * c2b, s2a * s2b * 0.5},
        {-(s2a * sb), -(sa * cb), 0.0, ca * cb, c2a * sb},
        {c2a * (cb * cb + sb * sb * 0.5), ca * s2b * -0.5, sb * sb * pt5sq3, sa * s2b * -0.5, s2a * (cb * cb + sb * sb * 0.5)},
        {-(s2a * cb), sa * sb, 0.0, -(ca * sb), c2a * cb}};
    for (int k = 0; k < 5; ++k)
      for (int K = 0; K < 5; ++K) R[4 + k][4 + K] = d[k][K];
  }
  return r;
}

// Charge-distribution transform: T[IJ][ab] takes local distribution ab (a >= b) into molecular
// IJ (I >= J) for an atom with n orbitals (src/mndod_twocenter.rs:742-915, tx + rotate_left).
void distribution_transform(int n, const double R[9][9], std::vector<double>& T) {
  const int np = n * (n + 1) / 2;
  T.assign((size_t)np * np, 0.0);
  for (int I = 0; I < n; ++I)
    for (int J = 0; J <= I; ++J)
      for (int a = 0; a < n; ++a)
        for (int b = 0; b <= a; ++b)
          T[(size_t)pack(I, J) * np + pack(a, b)] =
              a == b ? R[a][I] * R[a][J] : R[a][I] * R[b][J] + R[b][I] * R[a][J];
}

// PM7 d-imbalance correction (src/mndod_twocenter.rs:918-1000): the mean Coulomb integral
// between a transition metal's d shell and the partner's shells is set to (ss|ss), and the
// matching electron-core terms likewise. W is npa x npb, core[k][side] with k 1..10.
void d_balance(const Element& ei, const Element& ej, std::vector<double>& W, int npb, double core[11][3]) {
  auto w = [&](int ij, int kl) -> double& { return W[(size_t)ij * npb + kl]; };
  const double w00 = W[0];
  const int ni = ei.n_orb, nj = ej.n_orb;
  const int P[3] = {pack(1, 1), pack(2, 2), pack(3, 3)};
  int Dd[5];
  for (int k = 0; k < 5; ++k) Dd[k] = pack(4 + k, 4 + k);
  if (ei.iod > 0) {
    if (nj > 1) {
      double sum = 0.0;
      for (int a : Dd) sum += w(a, P[0]) + w(a, P[1]) + w(a, P[2]);
      sum = w00 - sum * (1.0 / 15.0);
      for (int a : Dd)
        for (int b : P) w(a, b) += sum;
      const double s = core[1][1] - (core[3][1] + core[4][1] * 2.0) * (1.0 / 3.0);
      core[3][1] += s;
      core[4][1] += s;
    }
    double sum = 0.0;
    for (int a : Dd) sum += w(a, 0);
    sum = w00 - sum * (1.0 / 5.0);
    for (int a : Dd) w(a, 0) += sum;
    const double s = core[1][1] - (core[7][1] + core[9][1] * 2.0 + core[10][1] * 2.0) * (1.0 / 5.0);
    core[7][1] += s;
    core[9][1] += s;
    core[10][1] += s;
  }
  if (ej.iod > 0) {
    if (ei.iod > 0) {
      double sum = 0.0;
      for (int a : Dd)
        for (int b : Dd) sum += w(a, b);
      sum = w00 - sum * (1.0 / 25.0);
      for (int a : Dd)
        for (int b : Dd) w(a, b) += sum;
    }
    if (ni > 1) {
      double sum = 0.0;
      for (int a : P)
        for (int b : Dd) sum += w(a, b);
      sum = w00 - sum * (1.0 / 15.0);
      for (int a : P)
        for (int b : Dd) w(a, b) += sum;
      const double s = core[1][2] - (core[3][2] + core[4][2] * 2.0) * (1.0 / 3.0);
      core[3][2] += s;
      core[4][2] += s;
    }
    double sum = 0.0;
    for (int b : Dd) sum += w(0, b);
    sum = w00 - sum * (1.0 / 5.0);
    for (int b : Dd) w(0, b) += sum;
    const double s = core[1][2] - (core[7][2] + core[9][2] * 2.0 + core[10][2] * 2.0) * (1.0 / 5.0);
    core[7][2] += s;
    core[9][2] += s;
    core[10][2] += s;
  }
}

// Electron-core attraction block of one atom in the molecular frame from the ten local terms
// core[k][side] (src/mndod_twocenter.rs:1003-1045): E = R^T E_local R.
void electron_core_block(const double core[11][3], int side, int n, const double R[9][9], double E[9][9]) {
  double L[9][9] = {{0}};
  auto c = [&](int k) { return core[k][side]; };
  L[0][0] = c(1);
  L[0][1] = L[1][0] = c(2);
  L[1][1] = c(3);
  L[2][2] = L[3][3] = c(4);
  L[0][4] = L[4][0] = c(5);
  L[1][4] = L[4][1] = c(6);
  L[4][4] = c(7);
  L[2][5] = L[5][2] = c(8);
  L[3][6] = L[6][3] = c(8);
  L[5][5] = L[6][6] = c(9);
  L[7][7] = L[8][8] = c(10);
  for (int K = 0; K < 9; ++K)
    for (int M = 0; M < 9; ++M) E[K][M] = 0.0;
  for (int K = 0; K < n; ++K)
    for (int M = 0; M < n; ++M) {
      double acc = 0.0;
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) acc += R[a][K] * L[a][b] * R[b][M];
      E[K][M] = acc;
    }
}

// Local electron-core terms core[1..4][side] of the s/p shells (src/mndod_twocenter.rs:482-572).
void local_core_sp(const Element& ei, const Element& ej, double r, double core[11][3]) {
  const DShell &A = ei.ds, &B = ej.ds;
  const double r2 = r * r, aci = A.po[9], acj = B.po[9];
  for (int k = 0; k < 11; ++k) core[k][0] = core[k][1] = core[k][2] = 0.0;
  core[1][1] = -(1.0 / std::sqrt(r2 + (acj + A.po[1]) * (acj + A.po[1]))) * (ej.core_charge * EV);
  core[1][2] = -(1.0 / std::sqrt(r2 + (aci + B.po[1]) * (aci + B.po[1]))) * (ei.core_charge * EV);
  // terms of a p shell on atom X in the field of the other core (additive ac)
  auto shell = [&](const DShell& X, double ac, double& t2, double& t3, double& t4) {
    const double pp = (ac + X.po[7]) * (ac + X.po[7]);
    const double da = X.ddp[2], qa = X.ddp[3] / std::sqrt(2.0), q2 = qa + qa;
    const double ad = (X.po[2] + ac) * (X.po[2] + ac), aq = (X.po[3] + ac) * (X.po[3] + ac);
    const double x1 = 1.0 / std::sqrt(r2 + pp);
    const double x2 = 1.0 / std::sqrt(r2 + aq) * -0.5;
    const double x3 = 1.0 / std::sqrt((r + da) * (r + da) + ad) * -0.5;
    const double x4 = 1.0 / std::sqrt((r - da) * (r - da) + ad) * 0.5;
    const double x5 = 1.0 / std::sqrt((r - q2) * (r - q2) + aq) * 0.25;
    const double x6 = 1.0 / std::sqrt((r + q2) * (r + q2) + aq) * 0.25;
    const double x7 = 1.0 / std::sqrt(r2 + (q2 * q2 + aq)) * 0.5;
    t2 = (x3 + x4) * EV;
    t3 = (x1 + x2 + x5 + x6) * EV;
    t4 = (x1 + x2 + x7) * EV;
  };
  if (ei.has_p()) {
    double t2, t3, t4;
    shell(A, acj, t2, t3, t4);
    core[2][1] = -t2 * ej.core_charge;
    core[3][1] = -t3 * ej.core_charge;
    core[4][1] = -t4 * ej.core_charge;
  }
  if (ej.has_p()) {
    double t2, t3, t4;
    shell(B, aci, t2, t3, t4);
    core[2][2] = t2 * ei.core_charge;  // the dipole term changes sign seen from the second atom
    core[3][2] = -t3 * ei.core_charge;
    core[4][2] = -t4 * ei.core_charge;
  }
}

// Two-centre two-electron and electron-core integrals of the ordered pair (ei first), d = R_j -
// R_i (src/mndod_twocenter.rs:1050-1112). Fills out.w (npa x npb), e1b and e2a.
void pair_two_electron_d(const Element& ei, const Element& ej, V3 d, PairBlock& out) {
  const DShell &A = ei.ds, &B = ej.ds;
  const int ni = ei.n_orb, nj = ej.n_orb;
  const bool with_d = ni == 9 || nj == 9;
  double R[9][9];
  const double r = orbital_rotation(d, with_d, R);
  double ri[23];
  local_sp_ri(ei, ej, r, ri);
  double core[11][3];
  local_core_sp(ei, ej, r, core);
  const int npa = ni * (ni + 1) / 2, npb = nj * (nj + 1) / 2;
  // Local integrals: s/p distributions from ri, the rest by the multipole expansion, feathered
  // (src/mndod_twocenter.rs:657-708).
  double cfrac, point;
  feather_to_point(r, cfrac, point);
  const double point_add = point * (1.0 - cfrac);
  std::vector<double> L((size_t)npa * npb, 0.0);
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b <= a; ++b)
      for (int c = 0; c < nj; ++c)
        for (int e = 0; e <= c; ++e) {
          const int ij = pack(a, b), kl = pack(c, e);
          double v;
          if (a < 4 && c < 4) {
            const int k = SP_LOCAL[ij][kl];
            v = k ? ri[k] : 0.0;
          } else {
            v = local_multipole(A, B, a, b, c, e, 0, r) * EV * cfrac;
            if (a == b && c == e) v += point_add;
          }
          L[(size_t)ij * npb + kl] = v;
        }
  if (with_d) {  // d electron-core terms (src/mndod_twocenter.rs:709-734)
    // distributions (d-sigma s) (d-sigma p-sigma) (d-sigma d-sigma) (d-pi p-pi) (d-pi d-pi) (d-delta d-delta)
    const int da[6] = {4, 4, 4, 5, 5, 7}, db[6] = {0, 1, 4, 2, 5, 7};
    for (int k = 0; k < 6; ++k) {
      if (nj == 9) core[5 + k][2] = -local_multipole(A, B, 0, 0, da[k], db[k], 1, r) * (EV * ei.core_charge);
      if (ni == 9) core[5 + k][1] = -local_multipole(A, B, da[k], db[k], 0, 0, 2, r) * (EV * ej.core_charge);
    }
  }
  // Feathering of the electron-core terms toward the point charge (src/mndod_twocenter.rs:1064-1081).
  {
    const double p1 = (1.0 / r) * (-EV * ej.core_charge), p2 = (1.0 / r) * (-EV * ei.core_charge);
    for (int k = 1; k <= 10; ++k) {
      core[k][1] *= cfrac;
      core[k][2] *= cfrac;
    }
    for (int k : {1, 3, 4, 7, 9, 10}) {
      core[k][1] += p1 * (1.0 - cfrac);
      core[k][2] += p2 * (1.0 - cfrac);
    }
  }
  // Rotate both distribution indices into the molecular frame: W = Ti L Tj^T.
  std::vector<double> Ti, Tj;
  distribution_transform(ni, R, Ti);
  distribution_transform(nj, R, Tj);
  std::vector<double> LT((size_t)npa * npb, 0.0);
  for (int ij = 0; ij < npa; ++ij)
    for (int KL = 0; KL < npb; ++KL) {
      double acc = 0.0;
      for (int kl = 0; kl < npb; ++kl) acc += L[(size_t)ij * npb + kl] * Tj[(size_t)KL * npb + kl];
      LT[(size_t)ij * npb + KL] = acc;
    }
  out.npa = npa;
  out.npb = npb;
  out.w.assign((size_t)npa * npb, 0.0);
  for (int IJ = 0; IJ < npa; ++IJ)
    for (int KL = 0; KL < npb; ++KL) {
      double acc = 0.0;
      for (int ij = 0; ij < npa; ++ij) acc += Ti[(size_t)IJ * npa + ij] * LT[(size_t)ij * npb + KL];
      out.w[(size_t)IJ * npb + KL] = acc;
    }
  d_balance(ei, ej, out.w, npb, core);
  electron_core_block(core, 1, ni, R, out.e1b);
  electron_core_block(core, 2, nj, R, out.e2a);
}

// ---- Overlaps of the spd basis (src/overlap_d.rs) ----
double fact_d(int n) {
  static double f[25];
  static bool init = false;
  if (!init) {
    f[0] = 1.0;
    for (int i = 1; i < 25; ++i) f[i] = f[i - 1] * i;
    init = true;
  }
  return f[n];
}

// B integrals bf[0..12] of argument x (src/overlap_d.rs:39-82).
void b_integrals(double x, double bf[13]) {
  const double ax = std::fabs(x);
  if (ax > 3.0) {
    const double ex = std::exp(x), emx = 1.0 / ex, inv = 1.0 / x;
    bf[0] = (ex - emx) * inv;
    for (int i = 1; i <= 12; ++i) bf[i] = (bf[i - 1] * i + ex * (i % 2 == 0 ? 1.0 : -1.0) - emx) * inv;
  } else if (ax <= 1.0e-6) {
    for (int i = 0; i <= 12; ++i) bf[i] = (2 * ((i + 1) % 2)) / (i + 1.0);
  } else {
    const int last = ax > 2.0 ? 15 : (ax > 1.0 ? 12 : (ax > 0.5 ? 7 : 6));
    for (int i = 0; i <= 12; ++i) {
      double y = 0.0;
      for (int m = 0; m <= last; ++m) {
        const double xf = m != 0 ? fact_d(m) : 1.0;
        const double coeff = (2 * ((m + i + 1) % 2)) / (xf * (m + i + 1));
        double xm = 1.0;
        for (int t = 0; t < m; ++t) xm *= -x;
        y += xm * coeff;
      }
      bf[i] = y;
    }
  }
}

// Overlap of Slater shells (na, la) on a and (nb, lb) on b with azimuthal component m, exponents
// ua ub, distance r, in the local frame (src/overlap_d.rs:84-176); l and m are 0-based here.
double shell_overlap(int na, int nb, int la, int lb, int m, double ua, double ub, double r) {
  auto aff = [](int l, int mm, int i) {
    if (l == 0 && mm == 0 && i == 0) return 1.0;
    if (l == 1 && mm == 0 && i == 0) return 1.0;
    if (l == 1 && mm == 1 && i == 0) return std::sqrt(0.5);
    if (l == 2 && mm == 0 && i == 0) return 1.5;
    if (l == 2 && mm == 1 && i == 0) return std::sqrt(1.5);
    if (l == 2 && mm == 2 && i == 0) return std::sqrt(0.375);
    if (l == 2 && mm == 0 && i == 2) return -0.5;
    return 0.0;
  };
  auto bi = [](int i, int j) { return (j < 0 || j > i) ? 0.0 : fact_d(i) / (fact_d(j) * fact_d(i - j)); };
  const double p = r * ((ua + ub) * 0.5), bb = r * ((ua - ub) * 0.5);
  double af[20];
  const double quo = 1.0 / p;
  af[0] = quo * std::exp(-p);
  for (int n = 1; n < 20; ++n) af[n] = af[n - 1] * n * quo + af[0];
  double bf[13];
  b_integrals(bb, bf);
  double sum = 0.0;
  for (int i = 0; i <= la - m; i += 2) {
    const int ia = na + i - la, ic = la - i - m;
    for (int j = 0; j <= lb - m; j += 2) {
      const int ib = nb + j - lb, id = lb - j - m;
      double sum1 = 0.0;
      const int iab = ia + ib;
      for (int k1 = 0; k1 <= ia; ++k1)
        for (int k2 = 0; k2 <= ib; ++k2)
          for (int k3 = 0; k3 <= ic; ++k3)
            for (int k4 = 0; k4 <= id; ++k4)
              for (int k5 = 0; k5 <= m; ++k5) {
                const int iaf = iab - k1 - k2 + k3 + k4 + 2 * k5;
                for (int k6 = 0; k6 <= m; ++k6) {
                  const int ibf = k1 + k2 + k3 + k4 + 2 * k6;
                  const double sign = 1 - 2 * ((m + k2 + k4 + k5 + k6) % 2);
                  const double coeff = bi(id, k4) * bi(ic, k3) * bi(ib, k2) * bi(ia, k1) * bi(m, k5) * bi(m, k6) * sign;
                  sum1 += af[iaf] * bf[ibf] * coeff;
                }
              }
      sum += sum1 * (aff(la, m, i) * aff(lb, m, j));
    }
  }
  double rpow = 1.0;
  for (int t = 0; t < na + nb + 1; ++t) rpow *= r;
  const double norm = std::sqrt(ua * ub / (fact_d(na + na) * fact_d(nb + nb)) * ((la + la + 1) * (lb + lb + 1)));
  return sum * rpow * (std::pow(ua, na) * std::pow(ub, nb) / 2.0 * norm);
}

// Molecular-frame overlap block S (9 x 9) of the ordered pair, d = R_j - R_i
// (src/overlap_d.rs:178-330): local sigma/pi/delta shell overlaps combined with the
// real-harmonic direction coefficients c(shell, component, local axis).
void diat_overlap_d(const Element& ei, const Element& ej, V3 d, double S[9][9]) {
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) S[i][j] = 0.0;
  const int ni = ei.n_orb, nj = ej.n_orb;
  if (ni == 0 || nj == 0) return;
  const int nij = std::max(ni, nj);
  const double xy2 = d.x * d.x + d.y * d.y;
  const double r = std::sqrt(xy2 + d.z * d.z), xy = std::sqrt(xy2);
  double ca, cb, sa, sb;
  if (xy >= 1.0e-10) {
    ca = d.x / xy;
    cb = d.z / r;
    sa = d.y / xy;
    sb = xy / r;
  } else if (d.z < 0.0) {
    ca = cb = -1.0;
    sa = sb = 0.0;
  } else if (d.z == 0.0) {
    ca = cb = sa = sb = 0.0;
  } else {
    ca = cb = 1.0;
    sa = sb = 0.0;
  }
  // c[shell 1..3][component 1..5][local axis 1..5]; axis 3 sigma, 2 and 4 pi, 1 and 5 delta.
  double c[4][6][6] = {{{0}}};
  c[1][3][3] = 1.0;
  if (nij >= 2) {
    c[2][4][4] = ca * cb;
    c[2][4][3] = ca * sb;
    c[2][4][2] = -sa;
    c[2][2][4] = sa * cb;
    c[2][2][3] = sa * sb;
    c[2][2][2] = ca;
    c[2][3][4] = -sb;
    c[2][3][3] = cb;
    c[2][3][2] = 0.0;
    if (nij >= 5) {
      const double rt34 = 0.86602540378444, rt13 = 0.57735026918963;
      const double c2a = ca * ca * 2.0 - 1.0, c2b = cb * cb * 2.0 - 1.0, s2a = sa * ca * 2.0, s2b = sb * cb * 2.0;
      c[3][5][5] = c2a * cb * cb + c2a * sb * sb * 0.5;
      c[3][5][4] = c2a * s2b * 0.5;
      c[3][5][3] = c2a * sb * sb * rt34;
      c[3][5][2] = -(s2a * sb);
      c[3][5][1] = -(s2a * cb);
      c[3][4][5] = ca * s2b * -0.5;
      c[3][4][4] = ca * c2b;
      c[3][4][3] = ca * s2b * rt34;
      c[3][4][2] = -(sa * cb);
      c[3][4][1] = sa * sb;
      c[3][3][5] = sb * sb * (rt13 * 1.5);
      c[3][3][4] = -(s2b * rt34);
      c[3][3][3] = cb * cb - sb * sb * 0.5;
      c[3][2][5] = sa * s2b * -0.5;
      c[3][2][4] = sa * c2b;
      c[3][2][3] = sa * s2b * rt34;
      c[3][2][2] = ca * cb;
      c[3][2][1] = -(ca * sb);
      c[3][1][5] = s2a * cb * cb + s2a * sb * sb * 0.5;
      c[3][1][4] = s2a * s2b * 0.5;
      c[3][1][3] = s2a * sb * sb * rt34;
      c[3][1][2] = c2a * sb;
      c[3][1][1] = c2a * cb;
    }
  }
  // orbital of (shell, component): p components 2 3 4 are py pz px, d components 1..5 are
  // dxy dyz dz2 dxz dx2-y2 (0-based orbital indices).
  const int orbital[4][6] = {{-1, -1, -1, -1, -1, -1}, {-1, 0, 0, 0, 0, -1}, {-1, -1, 2, 3, 1, -1}, {-1, 8, 7, 6, 5, 4}};
  auto shells = [](int n) { return n == 1 ? 1 : (n == 4 ? 2 : 3); };
  const int sa_n = shells(ni), sb_n = shells(nj);
  const double ei_exp[3] = {ei.zeta_s, ei.zeta_p, std::max(ei.zeta_d, 0.3)};
  const double ej_exp[3] = {ej.zeta_s, ej.zeta_p, std::max(ej.zeta_d, 0.3)};
  double s[4][4][4] = {{{0}}};
  const int newk = std::min(sa_n - 1, sb_n - 1);
  for (int i = 1; i <= sa_n; ++i) {
    const int pq1 = i == 3 ? shell_n_d(ei.z) : shell_n_sp(ei.z);
    for (int j = 1; j <= sb_n; ++j) {
      const int pq2 = j == 3 ? shell_n_d(ej.z) : shell_n_sp(ej.z);
      for (int k = 1; k <= newk + 1; ++k) {
        if (k > i || k > j) continue;
        s[i][j][k] = shell_overlap(std::max(pq1, i), std::max(pq2, j), i - 1, j - 1, k - 1, ei_exp[i - 1],
                                   ej_exp[j - 1], r);
      }
    }
  }
  for (int i = 1; i <= sa_n; ++i)
    for (int j = 1; j <= sb_n; ++j) {
      const double aa = j == 2 ? -1.0 : 1.0, bb = j == 3 ? -1.0 : 1.0;
      for (int k = 4 - i; k <= 2 + i; ++k)
        for (int l = 4 - j; l <= 2 + j; ++l) {
          const int oi = orbital[i][k], oj = orbital[j][l];
          if (oi < 0 || oj < 0) continue;
          S[oi][oj] = s[i][j][1] * (c[i][k][3] * c[j][l][3]) * aa +
                      s[i][j][2] * (c[i][k][4] * c[j][l][4] + c[i][k][2] * c[j][l][2]) * bb +
                      s[i][j][3] * (c[i][k][5] * c[j][l][5] + c[i][k][1] * c[j][l][1]);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Core Hamiltonian (src/hamiltonian.rs:310-638, molecular branch)
// ---------------------------------------------------------------------------------------------
double beta_of(const Element& e, int orb) { return orb == 0 ? e.beta_s : (orb <= 3 ? e.beta_p : e.beta_d); }

struct Core {
  Mat h;
  std::vector<PairBlock> pairs;
  bool d_path = false;  // every pair through the MNDO/d kernels
};

int build_mol(const Params& P, int n, const int* z, const double* xyz, Mol& m, std::string& err) {
  m.z.assign(z, z + n);
  m.pos.resize(n);
  m.off.resize(n);
  m.norb.resize(n);
  m.el.resize(n);
  int nao = 0;
  for (int i = 0; i < n; ++i) {
    m.pos[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    const Element* e = P.get(z[i]);
    if (!e) {
      err = "missing element " + std::to_string(z[i]);
      return 2;
    }
    m.el[i] = e;
    m.off[i] = nao;
    m.norb[i] = e->n_orb;
    nao += e->n_orb;
  }
  m.nao = nao;
  return 0;
}

Core build_core(const Mol& m, bool force_dpath = false) {
  int nao = m.nao, n = m.n();
  Core c;
  c.h.assign((size_t)nao * nao, 0.0);
  for (int i = 0; i < n; ++i) {
    const Element& e = *m.el[i];
    for (int k = 0; k < m.norb[i]; ++k)
      c.h[(size_t)(m.off[i] + k) * nao + m.off[i] + k] = k == 0 ? e.u_ss : (k <= 3 ? e.u_pp : e.u_dd);
  }
  // Molecular pair list u < v, heavier (more AOs) first (src/hamiltonian.rs:386-392).
  std::vector<std::pair<int, int>> list;
  for (int u = 0; u < n; ++u)
    for (int v = u + 1; v < n; ++v) list.push_back({u, v});
  c.pairs.resize(list.size());
  c.d_path = force_dpath;  // force_dpath (src/scf.rs:97, 544-582)
  for (int i = 0; i < n; ++i) c.d_path = c.d_path || m.el[i]->n_orb == 9;  // src/hamiltonian.rs:347-354
  parallel_for((long)list.size(), [&](long p) {
    int u = list[p].first, v = list[p].second;
    V3 d = sub(m.pos[v], m.pos[u]);
    int a = u, b = v;
    if (m.el[u]->n_orb < m.el[v]->n_orb) {
      a = v;
      b = u;
      d = {-d.x, -d.y, -d.z};
    }
    PairBlock& pb = c.pairs[p];
    pb.a = a;
    pb.b = b;
    if (c.d_path) {
      pair_two_electron_d(*m.el[a], *m.el[b], d, pb);
      diat_overlap_d(*m.el[a], *m.el[b], d, pb.s);
      return;
    }
    double r = norm(d);
    V3 dn = {d.x / r * r, d.y / r * r, d.z / r * r};  // pair_two_electron(ea, eb, d/r, r)
    pair_two_electron(*m.el[a], *m.el[b], dn, pb);
    double s4[4][4];
    diatom_overlap(*m.el[a], *m.el[b], d, s4);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) pb.s[i][j] = s4[i][j];
  });
  for (const auto& pb : c.pairs) {
    const Element& ea = *m.el[pb.a];
    const Element& eb = *m.el[pb.b];
    int oa = m.off[pb.a], ob = m.off[pb.b], na = m.norb[pb.a], nb = m.norb[pb.b];
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < na; ++j) c.h[(size_t)(oa + i) * nao + oa + j] += pb.e1b[i][j];
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) c.h[(size_t)(ob + i) * nao + ob + j] += pb.e2a[i][j];
    for (int i = 0; i < na; ++i) {
      double bi = beta_of(ea, i);
      for (int j = 0; j < nb; ++j) {
        double bj = beta_of(eb, j);
        double v = 0.5 * (bi + bj) * pb.s[i][j];
        c.h[(size_t)(oa + i) * nao + ob + j] += v;
        c.h[(size_t)(ob + j) * nao + oa + i] += v;
      }
    }
  }
  return c;
}

// ---------------------------------------------------------------------------------------------
// Fock build (src/fock.rs:69-107 one-centre integrals, 149-338 build_fock_spin_x)
// ---------------------------------------------------------------------------------------------
double oc_two_electron(int a, int b, int c, int d, const Element& e) {
  if (a == b && c == d) {
    if (a == 0 && c == 0) return e.g_ss;
    if (a == 0 || c == 0) return e.g_sp;
    return a == c ? e.g_pp : e.g_p2;
  }
  int ba = std::min(a, b), bb = std::max(a, b), kc = std::min(c, d), kd = std::max(c, d);
  if (ba == 0 && bb != 0 && kc == 0 && kd != 0 && bb == kd) return e.h_sp;
  if (ba != 0 && bb != 0 && ba != bb && ba == kc && bb == kd) return 0.5 * (e.g_pp - e.g_p2);
  return 0.0;
}

Mat build_fock_spin(const Mol& m, const Core& c, const Mat& pt, const Mat& ps) {
  int nao = m.nao, n = m.n();
  Mat f = c.h;
  auto P = [&](const Mat& x, int i, int j) { return x[(size_t)i * nao + j]; };
  // One-centre blocks.
  std::vector<std::vector<double>> oc(n);
  parallel_for(n, [&](long ia) {
    int off = m.off[ia], nn = m.norb[ia];
    const Element& e = *m.el[ia];
    oc[ia].assign(nn * nn, 0.0);
    // spd atoms read the 45 x 45 one-centre matrix (src/fock.rs:171-189)
    auto g = [&](int a, int b, int cc, int d) {
      return nn == 9 ? e.ds.oc[pack(a, b) * 45 + pack(cc, d)] : oc_two_electron(a, b, cc, d, e);
    };
    for (int mu = 0; mu < nn; ++mu)
      for (int nu = 0; nu < nn; ++nu) {
        double acc = 0.0;
        for (int la = 0; la < nn; ++la)
          for (int si = 0; si < nn; ++si) {
            acc += P(pt, off + la, off + si) * g(mu, nu, la, si);
            acc -= P(ps, off + la, off + si) * g(mu, la, nu, si);
          }
        oc[ia][mu * nn + nu] = acc;
      }
  });
  for (int ia = 0; ia < n; ++ia) {
    int off = m.off[ia], nn = m.norb[ia];
    for (int mu = 0; mu < nn; ++mu)
      for (int nu = 0; nu < nn; ++nu) f[(size_t)(off + mu) * nao + off + nu] += oc[ia][mu * nn + nu];
  }
  // Two-centre blocks.
  struct TC {
    double jaa[81], jbb[81], kab[81];
  };
  std::vector<TC> tc(c.pairs.size());
  parallel_for((long)c.pairs.size(), [&](long p) {
    const PairBlock& pb = c.pairs[p];
    int oa = m.off[pb.a], ob = m.off[pb.b], na = m.norb[pb.a], nb = m.norb[pb.b];
    int stride = pb.npb;
    const double* w = pb.w.data();
    TC& t = tc[p];
    for (int mu = 0; mu < na; ++mu)
      for (int nu = 0; nu < na; ++nu) {
        int row = pack(mu, nu) * stride;
        double acc = 0.0;
        for (int la = 0; la < nb; ++la)
          for (int si = 0; si < nb; ++si) acc += P(pt, ob + la, ob + si) * w[row + pack(la, si)];
        t.jaa[mu * na + nu] = acc;
      }
    for (int la = 0; la < nb; ++la)
      for (int si = 0; si < nb; ++si) {
        int col = pack(la, si);
        double acc = 0.0;
        for (int mu = 0; mu < na; ++mu)
          for (int nu = 0; nu < na; ++nu) acc += P(pt, oa + mu, oa + nu) * w[pack(mu, nu) * stride + col];
        t.jbb[la * nb + si] = acc;
      }
    for (int mu = 0; mu < na; ++mu)
      for (int la = 0; la < nb; ++la) {
        double acc = 0.0;
        for (int nu = 0; nu < na; ++nu) {
          int row = pack(mu, nu) * stride;
          for (int si = 0; si < nb; ++si) acc += P(ps, oa + nu, ob + si) * w[row + pack(la, si)];
        }
        t.kab[mu * nb + la] = -acc;
      }
  });
  for (size_t p = 0; p < c.pairs.size(); ++p) {
    const PairBlock& pb = c.pairs[p];
    int oa = m.off[pb.a], ob = m.off[pb.b], na = m.norb[pb.a], nb = m.norb[pb.b];
    const TC& t = tc[p];
    for (int mu = 0; mu < na; ++mu)
      for (int nu = 0; nu < na; ++nu) f[(size_t)(oa + mu) * nao + oa + nu] += t.jaa[mu * na + nu];
    for (int la = 0; la < nb; ++la)
      for (int si = 0; si < nb; ++si) f[(size_t)(ob + la) * nao + ob + si] += t.jbb[la * nb + si];
    for (int mu = 0; mu < na; ++mu)
      for (int la = 0; la < nb; ++la) {
        double v = t.kab[mu * nb + la];
        f[(size_t)(oa + mu) * nao + ob + la] += v;
        f[(size_t)(ob + la) * nao + oa + mu] += v;
      }
  }
  return f;
}

Mat build_fock_rhf(const Mol& m, const Core& c, const Mat& p) {  // src/fock.rs:648-686
  Mat half = p;
  for (auto& v : half) v *= 0.5;
  return build_fock_spin(m, c, p, half);
}

// ---------------------------------------------------------------------------------------------
// Dense linear algebra
// ---------------------------------------------------------------------------------------------
double frob(const Mat& a, const Mat& b) {  // src/linalg.rs:138-141 (in-order fold)
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// Symmetric eigensolve: Householder tridiagonalisation with accumulated transforms followed by
// the implicit-shift QL iteration (the EISPACK tred2/tql2 pair; Golub & Van Loan 8.3). Stands
// in for faer's self_adjoint_eigen (src/linalg.rs:231-278): ascending eigenvalues, eigenvectors
// as columns. Eigenvector signs are not pinned by the reference, only P and eps are.
void eigh_tql(int n, const Mat& a, std::vector<double>& w, Mat& vout) {
  std::vector<double> V(a), d(n), e(n);
  auto v = [&](int i, int j) -> double& { return V[(size_t)i * n + j]; };
  if (n == 1) {
    w.assign(1, a[0]);
    vout.assign(1, 1.0);
    return;
  }
  for (int j = 0; j < n; ++j) d[j] = v(n - 1, j);
  for (int i = n - 1; i > 0; --i) {
    double scale = 0.0, h = 0.0;
    for (int k = 0; k < i; ++k) scale += std::fabs(d[k]);
    if (scale == 0.0) {
      e[i] = d[i - 1];
      for (int j = 0; j < i; ++j) {
        d[j] = v(i - 1, j);
        v(i, j) = 0.0;
        v(j, i) = 0.0;
      }
    } else {
      for (int k = 0; k < i; ++k) {
        d[k] /= scale;
        h += d[k] * d[k];
      }
      double f = d[i - 1];
      double g = std::sqrt(h);
      if (f > 0) g = -g;
      e[i] = scale * g;
      h -= f * g;
      d[i - 1] = f - g;
      for (int j = 0; j < i; ++j) e[j] = 0.0;
      for (int j = 0; j < i; ++j) {
        f = d[j];
        v(j, i) = f;
        g = e[j] + v(j, j) * f;
        for (int k = j + 1; k <= i - 1; ++k) {
          g += v(k, j) * d[k];
          e[k] += v(k, j) * f;
        }
        e[j] = g;
      }
      f = 0.0;
      for (int j = 0; j < i; ++j) {
        e[j] /= h;
        f += e[j] * d[j];
      }
      double hh = f / (h + h);
      for (int j = 0; j < i; ++j) e[j] -= hh * d[j];
      for (int j = 0; j < i; ++j) {
        f = d[j];
        g = e[j];
        for (int k = j; k <= i - 1; ++k) v(k, j) -= (f * e[k] + g * d[k]);
        d[j] = v(i - 1, j);
        v(i, j) = 0.0;
      }
    }
    d[i] = h;
  }
  for (int i = 0; i < n - 1; ++i) {
    v(n - 1, i) = v(i, i);
    v(i, i) = 1.0;
    double h = d[i + 1];
    if (h != 0.0) {
      for (int k = 0; k <= i; ++k) d[k] = v(k, i + 1) / h;
      for (int j = 0; j <= i; ++j) {
        double g = 0.0;
        for (int k = 0; k <= i; ++k) g += v(k, i + 1) * v(k, j);
        for (int k = 0; k <= i; ++k) v(k, j) -= g * d[k];
      }
    }
    for (int k = 0; k <= i; ++k) v(k, i + 1) = 0.0;
  }
  for (int j = 0; j < n; ++j) {
    d[j] = v(n - 1, j);
    v(n - 1, j) = 0.0;
  }
  v(n - 1, n - 1) = 1.0;
  e[0] = 0.0;
  // tql2
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  double f = 0.0, tst1 = 0.0;
  const double eps = std::ldexp(1.0, -52);
  for (int l = 0; l < n; ++l) {
    tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
    int mm = l;
    while (mm < n) {
      if (std::fabs(e[mm]) <= eps * tst1) break;
      ++mm;
    }
    if (mm >= n) mm = n - 1;
    if (mm > l) {
      int iter = 0;
      do {
        ++iter;
        double g = d[l];
        double p = (d[l + 1] - g) / (2.0 * e[l]);
        double r = std::hypot(p, 1.0);
        if (p < 0) r = -r;
        d[l] = e[l] / (p + r);
        d[l + 1] = e[l] * (p + r);
        double dl1 = d[l + 1];
        double h = g - d[l];
        for (int i = l + 2; i < n; ++i) d[i] -= h;
        f += h;
        p = d[mm];
        double c = 1.0, c2 = c, c3 = c, el1 = e[l + 1], s = 0.0, s2 = 0.0;
        for (int i = mm - 1; i >= l; --i) {
          c3 = c2;
          c2 = c;
          s2 = s;
          g = c * e[i];
          h = c * p;
          r = std::hypot(p, e[i]);
          e[i + 1] = s * r;
          s = e[i] / r;
          c = p / r;
          p = c * d[i] - s * g;
          d[i + 1] = h + s * (c * g + s * d[i]);
          for (int k = 0; k < n; ++k) {
            h = v(k, i + 1);
            v(k, i + 1) = s * v(k, i) + c * h;
            v(k, i) = c * v(k, i) - s * h;
          }
        }
        p = -s * s2 * c3 * el1 * e[l] / dl1;
        e[l] = s * p;
        d[l] = c * p;
      } while (std::fabs(e[l]) > eps * tst1 && iter < 200);
    }
    d[l] += f;
    e[l] = 0.0;
  }
  // ascending sort with vectors
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  w.resize(n);
  vout.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    w[j] = d[idx[j]];
    for (int i = 0; i < n; ++i) vout[(size_t)i * n + j] = v(i, idx[j]);
  }
}


// Parallel-order cyclic Jacobi (round-robin pairing), the same rotation schedule as the engine's
// device eigensolver. Selected with PM7_ORACLE_EIGEN=jacobi to study eigenvector-basis effects.
void eigh_jacobi(int n, const Mat& a, std::vector<double>& w, Mat& vout) {
  int n2 = n + (n & 1), m = n2 / 2;
  std::vector<double> A((size_t)n2 * n2, 0.0), V((size_t)n2 * n2, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) A[(size_t)i * n2 + j] = a[(size_t)i * n + j];
  for (int i = 0; i < n2; ++i) V[(size_t)i * n2 + i] = 1.0;
  auto player = [n2](int pos, int r) { return pos == 0 ? 0 : 1 + (pos - 1 + r) % (n2 - 1); };
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int i = 0; i < n2; ++i)
      for (int j = 0; j < n2; ++j) (i == j ? dg : off) += A[(size_t)i * n2 + j] * A[(size_t)i * n2 + j];
    if (off <= 1.0e-26 * (off + dg)) break;
    for (int r = 0; r < n2 - 1; ++r) {
      std::vector<double> c(m), s(m), t(m);
      for (int k = 0; k < m; ++k) {
        int p = player(k, r), q = player(n2 - 1 - k, r);
        double apq = A[(size_t)p * n2 + q];
        c[k] = 1.0; s[k] = 0.0; t[k] = 0.0;
        if (apq != 0.0) {
          // t = sgn(d) x / (|d| + r), c = sqrt((|d| + r) / 2r): the smaller root of t^2 + 2 t d/x - 1 = 0
          double d = A[(size_t)q * n2 + q] - A[(size_t)p * n2 + p], x = 2.0 * apq;
          double r = std::sqrt(d * d + x * x), den = std::fabs(d) + r;
          t[k] = (d >= 0.0 ? x : -x) / den;
          c[k] = std::sqrt(den / (2.0 * r));
          s[k] = t[k] * c[k];
        }
      }
      std::vector<double> B(A);
      for (int k = 0; k < m; ++k)
        for (int l = 0; l < m; ++l) {
          int pk = player(k, r), qk = player(n2 - 1 - k, r), pl = player(l, r), ql = player(n2 - 1 - l, r);
          if (k == l) {
            double apq = A[(size_t)pk * n2 + qk];
            B[(size_t)pk * n2 + pk] = A[(size_t)pk * n2 + pk] - t[k] * apq;
            B[(size_t)qk * n2 + qk] = A[(size_t)qk * n2 + qk] + t[k] * apq;
            B[(size_t)pk * n2 + qk] = B[(size_t)qk * n2 + pk] = 0.0;
            continue;
          }
          double x00 = A[(size_t)pk * n2 + pl], x01 = A[(size_t)pk * n2 + ql], x10 = A[(size_t)qk * n2 + pl], x11 = A[(size_t)qk * n2 + ql];
          double y00 = c[k] * x00 - s[k] * x10, y01 = c[k] * x01 - s[k] * x11, y10 = s[k] * x00 + c[k] * x10, y11 = s[k] * x01 + c[k] * x11;
          B[(size_t)pk * n2 + pl] = c[l] * y00 - s[l] * y01;
          B[(size_t)pk * n2 + ql] = s[l] * y00 + c[l] * y01;
          B[(size_t)qk * n2 + pl] = c[l] * y10 - s[l] * y11;
          B[(size_t)qk * n2 + ql] = s[l] * y10 + c[l] * y11;
        }
      A.swap(B);
      for (int i = 0; i < n2; ++i)
        for (int k = 0; k < m; ++k) {
          if (s[k] == 0.0) continue;
          int pk = player(k, r), qk = player(n2 - 1 - k, r);
          double vp = V[(size_t)i * n2 + pk], vq = V[(size_t)i * n2 + qk];
          V[(size_t)i * n2 + pk] = c[k] * vp - s[k] * vq;
          V[(size_t)i * n2 + qk] = s[k] * vp + c[k] * vq;
        }
    }
  }
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) {
    int rank = 0;
    double di = A[(size_t)i * n2 + i];
    for (int j = 0; j < n; ++j) {
      double dj = A[(size_t)j * n2 + j];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    perm[rank] = i;
  }
  w.resize(n);
  vout.assign((size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k) {
    w[k] = A[(size_t)perm[k] * n2 + perm[k]];
    for (int i = 0; i < n; ++i) vout[(size_t)i * n + k] = V[(size_t)i * n2 + perm[k]];
  }
}

void eigh(int n, const Mat& a, std::vector<double>& w, Mat& vout) {
  static const bool use_jacobi = [] {
    const char* e = std::getenv("PM7_ORACLE_EIGEN");
    return e && std::strcmp(e, "jacobi") == 0;
  }();
  if (use_jacobi) eigh_jacobi(n, a, w, vout);
  else eigh_tql(n, a, w, vout);
}

// P = w * C_occ C_occ^T (src/scf.rs:1121-1141)
Mat density_from_coeff(const Mat& c, int n, int nocc, double weight) {
  Mat p((size_t)n * n, 0.0);
  parallel_for(n, [&](long mu) {
    for (int nu = 0; nu < n; ++nu) {
      double dotv = 0.0;
      for (int k = 0; k < nocc; ++k) dotv += c[(size_t)mu * n + k] * c[(size_t)nu * n + k];
      p[(size_t)mu * n + nu] = weight * dotv;
    }
  });
  return p;
}

// [F,P] = FP - (FP)^T (src/scf.rs:1148-1156)
void commutator(const Mat& f, const Mat& p, int n, Mat& out) {
  Mat fp((size_t)n * n, 0.0);
  parallel_for(n, [&](long i) {
    for (int k = 0; k < n; ++k) {
      double a = f[(size_t)i * n + k];
      for (int j = 0; j < n; ++j) fp[(size_t)i * n + j] += a * p[(size_t)k * n + j];
    }
  });
  out.resize((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[(size_t)i * n + j] = fp[(size_t)i * n + j] - fp[(size_t)j * n + i];
}

double rms_diff(const Mat& a, const Mat& b) {  // src/scf.rs:1178-1187
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
  return std::sqrt(s / (double)std::max<size_t>(1, a.size()));
}

// ---------------------------------------------------------------------------------------------
// SCF accelerators (src/scf.rs:1534-1916)
// ---------------------------------------------------------------------------------------------
const double MAX_LEVEL_SHIFT = 64.0;
const int STALL_WINDOW = 5;
const double ENERGY_PROGRESS_EV = 1.0e-9;
const double DIIS_MAX_WEIGHT = 25.0;

void apply_level_shift(Mat& f, const Mat& p, int n, double occ, double shift) {
  if (shift <= 0.0) return;
  double inv = 1.0 / occ;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double q = p[(size_t)i * n + j] * inv;
      f[(size_t)i * n + j] += shift * ((i == j) ? 1.0 - q : -q);
    }
}

struct StallGuard {  // src/scf.rs:1581-1656
  double floor_, shift;
  std::vector<double> steps, energies;
  explicit StallGuard(double fl) : floor_(fl), shift(fl) {}
  void update(double step, double energy) {
    steps.push_back(step);
    energies.push_back(energy);
    size_t n = steps.size();
    if (n < 2 * STALL_WINDOW) return;
    auto smallest = [](const std::vector<double>& v, size_t lo, size_t hi) {
      double m = INFINITY;
      for (size_t i = lo; i < hi; ++i) m = std::min(m, v[i]);
      return m;
    };
    double step_now = smallest(steps, n - STALL_WINDOW, n);
    double step_was = smallest(steps, n - 2 * STALL_WINDOW, n - STALL_WINDOW);
    double e_now = smallest(energies, n - STALL_WINDOW, n);
    double e_was = smallest(energies, n - 2 * STALL_WINDOW, n - STALL_WINDOW);
    bool moved;
    if (step_now <= step_was * 0.1) {
      if (shift <= floor_) {
        moved = false;
      } else {
        double relaxed = std::max(shift * 0.5, floor_);
        shift = relaxed < 0.25 ? floor_ : relaxed;
        moved = true;
      }
    } else if (step_now > step_was * 0.9 && e_now > e_was - ENERGY_PROGRESS_EV && shift < MAX_LEVEL_SHIFT) {
      shift = shift < 1.0 ? 1.0 : std::min(shift * 2.0, MAX_LEVEL_SHIFT);
      moved = true;
    } else {
      moved = false;
    }
    if (moved) {
      steps.clear();
      energies.clear();
    } else {
      steps.erase(steps.begin());
      energies.erase(energies.begin());
    }
  }
};

bool solve_bordered_small(std::vector<double> m, std::vector<double> rhs, int n, std::vector<double>& x) {
  for (int col = 0; col < n; ++col) {
    int piv = col;
    double best = std::fabs(m[col * n + col]);
    for (int row = col + 1; row < n; ++row) {
      double v = std::fabs(m[row * n + col]);
      if (v > best) {
        best = v;
        piv = row;
      }
    }
    if (best < 1.0e-12) return false;
    if (piv != col) {
      for (int j = 0; j < n; ++j) std::swap(m[col * n + j], m[piv * n + j]);
      std::swap(rhs[col], rhs[piv]);
    }
    for (int row = col + 1; row < n; ++row) {
      double factor = m[row * n + col] / m[col * n + col];
      if (factor == 0.0) continue;
      for (int j = col; j < n; ++j) m[row * n + j] -= factor * m[col * n + j];
      rhs[row] -= factor * rhs[col];
    }
  }
  x.assign(n, 0.0);
  for (int col = n - 1; col >= 0; --col) {
    double sum = rhs[col];
    for (int j = col + 1; j < n; ++j) sum -= m[col * n + j] * x[j];
    x[col] = sum / m[col * n + col];
  }
  return true;
}

bool diis_window_coeffs(const std::vector<double>& gram, int n, int skip, std::vector<double>& c) {
  int m = n - skip;
  if (m < 2) return false;
  double scale = 0.0;
  for (int i = 0; i < m; ++i) scale = std::max(scale, std::fabs(gram[(skip + i) * n + skip + i]));
  if (!std::isfinite(scale) || scale <= 0.0) return false;
  int dim = m + 1;
  std::vector<double> b((size_t)dim * dim, 0.0), rhs(dim, 0.0);
  for (int i = 0; i < m; ++i) {
    for (int j = 0; j < m; ++j) b[i * dim + j] = gram[(skip + i) * n + skip + j] / scale;
    b[i * dim + m] = -1.0;
    b[m * dim + i] = -1.0;
  }
  rhs[m] = -1.0;
  if (!solve_bordered_small(b, rhs, dim, c)) return false;
  for (double v : c)
    if (!std::isfinite(v)) return false;
  return true;
}

bool diis_coeffs(const std::vector<Mat>& es, int& skip_out, std::vector<double>& coeffs) {
  int n = (int)es.size();
  if (n < 2) return false;
  std::vector<double> gram((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) gram[i * n + j] = gram[j * n + i] = frob(es[i], es[j]);
  for (int skip = 0; skip < n - 1; ++skip) {
    std::vector<double> c;
    if (diis_window_coeffs(gram, n, skip, c)) {
      double sum = 0.0;
      for (int i = 0; i < n - skip; ++i) sum += std::fabs(c[i]);
      if (sum <= DIIS_MAX_WEIGHT) {
        skip_out = skip;
        coeffs = c;
        return true;
      }
    }
  }
  return false;
}

Mat combine(const std::vector<Mat>& fs, int from, const std::vector<double>& coeffs) {
  Mat out(fs[from].size(), 0.0);
  for (size_t i = from; i < fs.size(); ++i) {
    double ci = coeffs[i - from];
    for (size_t k = 0; k < out.size(); ++k) out[k] += ci * fs[i][k];
  }
  return out;
}

std::vector<double> simplex_project(const std::vector<double>& v) {
  std::vector<double> u(v);
  std::sort(u.begin(), u.end(), [](double a, double b) { return a > b; });
  double css = 0.0, theta = 0.0;
  for (size_t j = 0; j < u.size(); ++j) {
    css += u[j];
    double t = (css - 1.0) / (double)(j + 1);
    if (u[j] - t > 0.0) theta = t;
  }
  std::vector<double> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[i] = std::max(v[i] - theta, 0.0);
  return out;
}

std::vector<double> solve_adiis_simplex(const std::vector<double>& d, const std::vector<std::vector<double>>& s) {
  int n = (int)d.size();
  double l = 1.0e-12;
  for (int i = 0; i < n; ++i) {
    double row = 0.0;
    for (int j = 0; j < n; ++j) row += std::fabs(s[i][j] + s[j][i]);
    l = std::max(l, row);
  }
  double lr = 1.0 / l;
  std::vector<double> c(n, 0.0);
  c[n - 1] = 1.0;
  for (int it = 0; it < 400; ++it) {
    std::vector<double> g(n), trial(n);
    for (int k = 0; k < n; ++k) {
      double acc = 2.0 * d[k];
      for (int j = 0; j < n; ++j) acc += (s[k][j] + s[j][k]) * c[j];
      g[k] = acc;
    }
    for (int i = 0; i < n; ++i) trial[i] = c[i] - lr * g[i];
    auto proj = simplex_project(trial);
    double delta = 0.0;
    for (int i = 0; i < n; ++i) delta += std::fabs(proj[i] - c[i]);
    c = proj;
    if (delta < 1.0e-12) break;
  }
  return c;
}

bool adiis_extrapolate(const std::vector<Mat>& D, const std::vector<Mat>& F, Mat& out) {
  int n = (int)D.size();
  if (n < 2) return false;
  const Mat& dn = D[n - 1];
  const Mat& fn = F[n - 1];
  std::vector<double> d(n);
  std::vector<std::vector<double>> s(n, std::vector<double>(n));
  for (int i = 0; i < n; ++i) {
    double acc = 0.0;
    for (size_t k = 0; k < dn.size(); ++k) acc += (D[i][k] - dn[k]) * fn[k];
    d[i] = acc;
    for (int j = 0; j < n; ++j) {
      double a2 = 0.0;
      for (size_t k = 0; k < dn.size(); ++k) a2 += (D[i][k] - dn[k]) * (F[j][k] - fn[k]);
      s[i][j] = a2;
    }
  }
  auto c = solve_adiis_simplex(d, s);
  out = combine(F, 0, c);
  return true;
}

void push_hist(std::vector<Mat>& h, const Mat& m, size_t max) {
  if (h.size() >= max) h.erase(h.begin());
  h.push_back(m);
}

// MOPAC diagonal guess (src/scf.rs:1050-1114).
Mat sad_density(const Mol& m, double n_electrons) {
  int nao = m.nao;
  Mat p((size_t)nao * nao, 0.0);
  double tore_sum = 0.0;
  for (int i = 0; i < m.n(); ++i) tore_sum += m.el[i]->core_charge;
  double yy = nao > 0 ? (tore_sum - n_electrons) / nao : 0.0;
  for (int i = 0; i < m.n(); ++i) {
    int off = m.off[i];
    double tore = m.el[i]->core_charge;
    if (m.norb[i] == 1) {
      p[(size_t)off * nao + off] = tore - yy;
    } else if (m.norb[i] == 4) {
      double w = tore * 0.25 - yy;
      for (int k = 0; k < 4; ++k) p[(size_t)(off + k) * nao + off + k] = w;
    } else if (m.norb[i] == 9) {
      const int zi = m.z[i];
      auto at = [&](int k) -> double& { return p[(size_t)(off + k) * nao + off + k]; };
      if (zi < 21 || (zi > 30 && zi < 39) || (zi > 48 && zi < 57)) {
        double w = tore * 0.25 - yy;
        for (int k = 0; k < 4; ++k) at(k) = w;
        for (int k = 4; k < 9; ++k) at(k) = -yy;
      } else {  // transition metal: s (<= 2), then d, then p
        double sum = tore - 9.0 * yy;
        at(0) = std::min(std::max(sum, 0.0), 2.0);
        sum -= 2.0;
        if (sum > 0.0) {
          for (int k = 4; k < 9; ++k) at(k) = std::min(std::max(sum * 0.2, 0.0), 2.0);
          sum -= 10.0;
          if (sum > 0.0)
            for (int k = 1; k < 4; ++k) at(k) = sum / 3.0;
        }
      }
    }
  }
  return p;
}

struct ScfState {
  Mat density, spin;
  std::vector<double> eps, eps_b;
  int n_occ = 0, n_occ_b = 0;
  double e_el = 0, dp_err = INFINITY;
  bool converged = false, uhf = false;
  int iterations = 0;
};

ScfState rhf_loop(const Mol& m, const Core& core, int n_occ, const oracle_options& o) {  // :1189-1322
  int nao = m.nao;
  ScfState st;
  st.n_occ = n_occ;
  // The caller's density when there is one (src/scf.rs:1201-1210).
  Mat density = o.initial_density ? Mat(o.initial_density, o.initial_density + (size_t)m.nao * m.nao)
                                  : sad_density(m, 2.0 * n_occ);
  double e_old = 0.0;
  std::vector<Mat> hf, he, hd;
  int accel = o.use_diis ? o.accelerator : 0;
  StallGuard guard(o.level_shift);
  Mat mo;
  std::vector<double> eps;
  for (int iter = 0; iter < o.max_scf; ++iter) {
    st.iterations = iter + 1;
    Mat f = build_fock_rhf(m, core, density);
    double e = 0.5 * (frob(density, core.h) + frob(density, f));
    push_hist(hf, f, 8);
    Mat err;
    commutator(f, density, nao, err);
    push_hist(he, err, 8);
    push_hist(hd, density, 8);
    double err_norm = std::sqrt(frob(err, err));
    Mat fuse;
    bool ok = false;
    if (accel == 1 || (accel == 2 && err_norm <= o.adiis_switch)) {
      int skip;
      std::vector<double> c;
      if (diis_coeffs(he, skip, c)) {
        fuse = combine(hf, skip, c);
        ok = true;
      }
    } else if (accel == 2) {
      ok = adiis_extrapolate(hd, hf, fuse);
    }
    if (!ok) fuse = f;
    apply_level_shift(fuse, density, nao, 2.0, guard.shift);
    Mat c;
    eigh(nao, fuse, eps, c);
    Mat pn = density_from_coeff(c, nao, n_occ, 2.0);
    double dp = rms_diff(pn, density);
    st.dp_err = dp;
    double de = std::fabs(e - e_old);
    density = pn;
    mo = c;
    e_old = e;
    bool residual_ok = guard.shift == 0.0 && err_norm < 1.0e-7;
    if (iter > 0 && de < o.e_tol && (dp < o.p_tol || residual_ok)) {
      st.converged = true;
      break;
    }
    guard.update(dp, e);
  }
  Mat ff = build_fock_rhf(m, core, density);
  st.e_el = 0.5 * (frob(density, core.h) + frob(density, ff));
  if (guard.shift > 0.0) eigh(nao, ff, eps, mo);
  st.density = density;
  st.eps = eps;
  return st;
}

ScfState uhf_loop(const Mol& m, const Core& core, int na, int nb, const oracle_options& o) {  // :1325-1532
  int nao = m.nao;
  ScfState st;
  st.uhf = true;
  st.n_occ = na;
  st.n_occ_b = nb;
  Mat total = o.initial_density ? Mat(o.initial_density, o.initial_density + (size_t)m.nao * m.nao)
                                : sad_density(m, (double)(na + nb));  // src/scf.rs:1341-1360
  double ntot = std::max(na + nb, 1);
  double fa = na / ntot, fb = nb / ntot;
  Mat pa = total, pb = total;
  for (auto& v : pa) v *= fa;
  for (auto& v : pb) v *= fb;
  if (o.initial_spin_density)  // P_alpha += s/2, P_beta -= s/2 (src/scf.rs:1361-1378)
    for (size_t k = 0; k < pa.size(); ++k) {
      pa[k] += 0.5 * o.initial_spin_density[k];
      pb[k] -= 0.5 * o.initial_spin_density[k];
    }
  double e_old = 0.0;
  StallGuard guard(o.level_shift);
  std::vector<double> eps_a;
  Mat ca;
  for (int iter = 0; iter < o.max_scf; ++iter) {
    st.iterations = iter + 1;
    Mat pt(pa);
    for (size_t k = 0; k < pt.size(); ++k) pt[k] += pb[k];
    Mat Fa = build_fock_spin(m, core, pt, pa);
    Mat Fb = build_fock_spin(m, core, pt, pb);
    double e = 0.5 * (frob(pt, core.h) + frob(pa, Fa) + frob(pb, Fb));
    Mat ea, eb;
    commutator(Fa, pa, nao, ea);
    commutator(Fb, pb, nao, eb);
    double s = 0.0;
    for (double v : ea) s += v * v;
    for (double v : eb) s += v * v;
    double err_norm = std::sqrt(s);
    apply_level_shift(Fa, pa, nao, 1.0, guard.shift);
    apply_level_shift(Fb, pb, nao, 1.0, guard.shift);
    std::vector<double> wa, wb;
    Mat cA, cB;
    eigh(nao, Fa, wa, cA);
    eigh(nao, Fb, wb, cB);
    Mat pan = density_from_coeff(cA, nao, na, 1.0);
    Mat pbn = density_from_coeff(cB, nao, nb, 1.0);
    double dp = rms_diff(pan, pa) + rms_diff(pbn, pb);
    st.dp_err = dp;
    double de = std::fabs(e - e_old);
    eps_a = wa;
    ca = cA;
    pa = pan;
    pb = pbn;
    e_old = e;
    bool residual_ok = guard.shift == 0.0 && err_norm < 1.0e-7;
    if (iter > 0 && de < o.e_tol && (dp < o.p_tol || residual_ok)) {
      st.converged = true;
      break;
    }
    guard.update(dp, e);
  }
  Mat dens(pa), spin(pa);
  for (size_t k = 0; k < dens.size(); ++k) {
    dens[k] += pb[k];
    spin[k] -= pb[k];
  }
  Mat Fa = build_fock_spin(m, core, dens, pa);
  Mat Fb = build_fock_spin(m, core, dens, pb);
  st.e_el = 0.5 * (frob(dens, core.h) + frob(pa, Fa) + frob(pb, Fb));
  if (guard.shift > 0.0) eigh(nao, Fa, eps_a, ca);
  Mat cb;
  eigh(nao, Fb, st.eps_b, cb);
  st.density = dens;
  st.spin = spin;
  st.eps = eps_a;
  return st;
}

// ---------------------------------------------------------------------------------------------
// Core-core repulsion (src/repulsion.rs:16-36, 164-263)
// ---------------------------------------------------------------------------------------------
double core_gaussian_slot(const Element& src, double zz, double r_ang, size_t slot) {
  if (slot >= src.gauss.size()) return 0.0;
  double k = src.gauss[slot][0], l = src.gauss[slot][1], mm = src.gauss[slot][2];
  double ax = (r_ang - mm) * (r_ang - mm) * l;
  if (ax >= 25.0) return 0.0;
  return std::exp(-ax) * (zz * k) / r_ang;
}

double pair_core_energy(const Params& P, const Element& ei, const Element& ej, double r) {
  int zi = ei.z, zj = ej.z;
  double r_ang = r * A0;
  double zz = ei.core_charge * ej.core_charge;
  double gab_nddo = EV / std::sqrt(r * r + (ei.po9() + ej.po9()) * (ei.po9() + ej.po9()));
  double cfrac, point;
  feather_to_point(r, cfrac, point);
  double gab = gab_nddo * cfrac + point * (1.0 - cfrac);
  double enuc = gab * zz;
  auto pr = P.pair(zi, zj);
  double alpb0 = pr.first, xfac = pr.second;
  bool defined = std::fabs(xfac) > 1.0e-5;
  if (defined) {
    double alpb = alpb0 < 1.0e-6 ? 1.2 : alpb0;
    double scale = 1.0 + std::exp(-(r_ang + std::pow(r_ang, 6) * 0.0003) * alpb) * (2.0 * xfac);
    int hi = std::max(zi, zj), lo = std::min(zi, zj);
    if (lo == 1 && (hi == 6 || hi == 7)) {
      scale = 1.0 + std::exp(-(r_ang * r_ang) * alpb) * (2.0 * xfac);
    } else if (lo == 1 && hi == 8) {
      scale = 1.0 + std::exp(-(r_ang * r_ang) * alpb) * (2.0 * xfac) -
              std::exp(r_ang * (-2.0 * P.vpar[4])) * P.vpar[3];
    } else if (lo == 6 && hi == 6) {
      scale = scale + std::exp(r_ang * (-P.vpar[2])) * P.vpar[1];
    } else if (lo == 8 && hi == 14) {
      scale = scale - std::exp(-(r_ang - 2.9) * (r_ang - 2.9)) * 0.0007;
    }
    enuc *= scale;
  } else {
    bool sparkle = (zi >= 57 && zi <= 71) || (zj >= 57 && zj <= 71);
    double ex = sparkle ? -3.0 : -2.18;
    enuc *= 1.0 + std::exp(r_ang * ex) * 10.0;
  }
  double corr = core_gaussian_slot(ei, zz, r_ang, 0);
  corr += core_gaussian_slot(ej, zz, r_ang, 0);
  if (!defined)
    for (size_t s = 0; s < 4; ++s) {
      corr += core_gaussian_slot(ei, zz, r_ang, s);
      corr += core_gaussian_slot(ej, zz, r_ang, s);
    }
  double reduced = r_ang / (std::pow((double)zi, 0.3333) + std::pow((double)zj, 0.3333));
  if (reduced < 3.0) {
    double guard = 1.0e-8 / std::pow(reduced, 12);
    corr += guard < 1.0e5 ? guard : 1.0e5;
  }
  return enuc + corr;
}

double core_core_energy(const Params& P, const Mol& m) {
  double sum = 0.0;
  for (int i = 0; i < m.n(); ++i)
    for (int j = i + 1; j < m.n(); ++j) sum += pair_core_energy(P, *m.el[i], *m.el[j], norm(sub(m.pos[j], m.pos[i])));
  return sum;
}

// ---------------------------------------------------------------------------------------------
// Post-SCF geometric corrections (kcal/mol)
// ---------------------------------------------------------------------------------------------
// Covalent radii (src/constants.rs:64-74 covalent_radius_angstrom), Z = 0..86, else 1.5.
double covalent_radius(int z) {
  static const double R[87] = {
      0.0,  0.31, 0.28, 1.28, 0.96, 0.84, 0.76, 0.71, 0.66, 0.57, 0.58, 1.66, 1.41, 1.21, 1.11,
      1.07, 1.05, 1.02, 1.06, 2.03, 1.76, 1.70, 1.60, 1.53, 1.39, 1.39, 1.32, 1.26, 1.24, 1.32,
      1.22, 1.22, 1.20, 1.19, 1.20, 1.20, 1.16, 2.20, 1.95, 1.90, 1.75, 1.64, 1.54, 1.47, 1.46,
      1.42, 1.39, 1.45, 1.44, 1.42, 1.39, 1.39, 1.38, 1.39, 1.40, 2.44, 2.15, 2.07, 2.04, 2.03,
      2.01, 1.99, 1.98, 1.98, 1.96, 1.94, 1.92, 1.92, 1.89, 1.90, 1.87, 1.87, 1.75, 1.70, 1.62,
      1.51, 1.44, 1.41, 1.36, 1.36, 1.32, 1.45, 1.46, 1.48, 1.40, 1.50, 1.50};
  return (z >= 0 && z <= 86) ? R[z] : 1.5;
}

// D2 dispersion (src/dispersion.rs:19-264); smooth: the continuous carbon C6 of
// dispersion_energy_smooth_cut (src/dispersion.rs:72-175, 383-412).
double dispersion_energy(const std::vector<int>& z, const std::vector<V3>& pos, bool smooth = false) {
  static const double ALPHA = 15.450118, S = 1.226593, CSCALE = 2.286419;
  auto c6tab = [](int zz) {
    switch (zz) {
      case 1: return 0.16; case 2: return 0.084; case 5: return 5.79; case 6: return 1.65;
      case 7: return 1.11; case 8: return 0.70; case 9: return 0.57; case 10: return 0.45;
      case 15: return 3.25; case 16: return 5.79; case 17: return 5.97; case 18: return 3.71;
      case 27: return 0.04; case 35: return 11.60; case 36: return 4.47; case 53: return 25.80;
      case 54: return 16.50; default: return 0.0;
    }
  };
  auto r0tab = [](int zz) {
    switch (zz) {
      case 1: return 156.0; case 2: return 140.0; case 5: return 180.0; case 6: return 170.0;
      case 7: return 155.0; case 8: return 152.0; case 9: return 147.0; case 10: return 154.0;
      case 15: return 180.0; case 16: return 180.0; case 17: return 175.0; case 18: return 188.0;
      case 28: return 140.0; case 35: return 185.0; case 36: return 202.0; case 53: return 198.0;
      case 54: return 216.0; default: return 0.0;
    }
  };
  auto nefftab = [](int zz) {
    switch (zz) {
      case 1: return 0.80; case 2: return 1.42; case 5: return 2.16; case 6: return 2.50;
      case 7: return 2.82; case 8: return 3.15; case 9: return 3.48; case 10: return 3.81;
      case 15: return 4.50; case 16: return 4.80; case 17: return 5.10; case 18: return 5.40;
      case 27: return 2.90; case 35: return 6.00; case 36: return 6.30; case 53: return 6.95;
      case 54: return 7.25; default: return 0.0;
    }
  };
  int n = (int)z.size();
  std::vector<int> nb(n, 0);
  std::vector<double> cn(n, 0.0);  // smooth neighbour counts (count_weight, src/dispersion.rs:106-120)
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double d = norm(sub(pos[j], pos[i]));
      if (d * A0 < 1.3 * (covalent_radius(z[i]) + covalent_radius(z[j]))) {
        nb[i]++;
        nb[j]++;
      }
      const double rcut = 1.3 * (covalent_radius(z[i]) + covalent_radius(z[j])) / A0;
      const double ex = std::exp(-26.0 * (rcut / d - 1.0));
      const double w = std::isfinite(ex) ? 1.0 / (1.0 + ex) : 0.0;
      cn[i] += w;
      cn[j] += w;
    }
  // c6_atom_smooth (src/dispersion.rs:133-140): 1.65 + (0.95 - 1.65) s, s = 1/(1 + e^{-40 (cn - 3.5)}).
  auto c6c = [&](int i) {
    if (!smooth) return nb[i] == 4 ? 0.95 : 1.65;
    const double sw = 1.0 / (1.0 + std::exp(-(cn[i] - 3.5) * 40.0));
    return 1.65 + (0.95 - 1.65) * sw;
  };
  double sum = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      int zi = z[i], zj = z[j];
      bool ok = c6tab(zi) != 0 && r0tab(zi) != 0 && nefftab(zi) != 0 && c6tab(zj) != 0 &&
                r0tab(zj) != 0 && nefftab(zj) != 0;
      if (!ok) continue;
      double ni = nefftab(zi), nj = nefftab(zj);
      double c6i = zi == 6 ? c6c(i) : c6tab(zi);
      double c6j = zj == 6 ? c6c(j) : c6tab(zj);
      double c6 = 2.0 * std::pow(c6i * c6i * c6j * c6j * ni * nj, 1.0 / 3.0) /
                  (std::pow(c6i * nj * nj, 1.0 / 3.0) + std::pow(c6j * ni * ni, 1.0 / 3.0));
      double ri = r0tab(zi), rj = r0tab(zj);
      double r0 = (ri * ri * ri + rj * rj * rj) / (ri * ri + rj * rj) / 1000.0 * 2.0;
      double r = norm(sub(pos[j], pos[i]));
      double rij_nm = r * (A0 * 0.1);
      double damp_denom = std::exp((rij_nm * (1.0 / (S * r0)) - 1.0) * (-ALPHA)) + 1.0;
      double e = (1.0 / std::pow(rij_nm, 6)) * (c6 / (1000.0 * 4.184)) / damp_denom;
      sum += e * (-CSCALE);
    }
  return sum;
}

// PM7-HH H-H repulsion (src/hh_rep.rs:21-60): every H-H pair adds poly(r / Angstrom), kcal/mol.
double hh_repulsion_energy(const std::vector<int>& z, const std::vector<V3>& pos) {
  double sum = 0.0;
  const int n = (int)z.size();
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (z[i] != 1 || z[j] != 1) continue;
      const double r = norm(sub(pos[j], pos[i])) * A0;
      double e;
      if (r <= 1.0) {
        e = 25.46293603147693;
      } else if (r < 1.5) {
        e = std::pow(r, 5) * -2714.9523516034697 + std::pow(r, 4) * 17103.650110591705 +
            std::pow(r, 3) * -42511.85798221796 + r * r * 52063.19679913834 + r * -31430.65833597229 +
            7516.08469609514;
      } else {
        e = std::exp(std::pow(r, 1.72905) * -1.53965) * 118.7326;
      }
      sum += e;
    }
  return sum;
}

// EH+ hydrogen-bond correction (src/hbond.rs:53-626, 1294-1322), brute-force scans in ascending
// index order (the reference's spatial grid only pre-filters the same scans).
namespace hb {
const double MOPAC_TWO_PI = 6.28318530717959;
const double SHORTCUT = 2.4, LONGCUT = 7.0, COVCUT = 1.2;
double deg(double d) { return PI / (180.0 / d); }
double dist(const std::vector<V3>& c, int a, int b) { return norm(sub(c[a], c[b])); }
double bangle(const std::vector<V3>& c, int i, int j, int k) {
  V3 a = sub(c[i], c[j]), b = sub(c[j], c[k]), d = sub(c[i], c[k]);
  double d2ij = dot(a, a), d2jk = dot(b, b), d2ik = dot(d, d);
  double xy = std::sqrt(d2ij * d2jk);
  if (xy < 1.0e-20) return 0.0;
  double t = std::min(1.0, std::max(-1.0, 0.5 * (d2ij + d2jk - d2ik) / xy));
  return std::acos(t);
}
double dang(double a1, double a2, double b1, double b2) {
  const double zero = 1.0e-6;
  if ((std::fabs(a1) >= zero || std::fabs(a2) >= zero) && (std::fabs(b1) >= zero || std::fabs(b2) >= zero)) {
    double an = 1.0 / std::sqrt(a1 * a1 + a2 * a2), bn = 1.0 / std::sqrt(b1 * b1 + b2 * b2);
    a1 *= an;
    a2 *= an;
    b1 *= bn;
    b2 *= bn;
    double sinth = a1 * b2 - a2 * b1;
    double costh = std::min(1.0, std::max(-1.0, a1 * b1 + a2 * b2));
    double rcos = std::acos(costh);
    if (std::fabs(rcos) >= 4.0e-5) {
      if (sinth > 0.0) rcos = MOPAC_TWO_PI - rcos;
      return -rcos;
    }
  }
  return 0.0;
}
double dihed(const std::vector<V3>& c, int i, int j, int k, int l) {
  double xi1 = c[i].x - c[k].x, xj1 = c[j].x - c[k].x, xl1 = c[l].x - c[k].x;
  double yi1 = c[i].y - c[k].y, yj1 = c[j].y - c[k].y, yl1 = c[l].y - c[k].y;
  double zi1 = c[i].z - c[k].z, zj1 = c[j].z - c[k].z, zl1 = c[l].z - c[k].z;
  double dd = std::sqrt(xj1 * xj1 + yj1 * yj1 + zj1 * zj1);
  double cosa = dd > 0.0 ? std::min(1.0, std::max(-1.0, zj1 / dd)) : 0.0;
  double ddd = 1.0 - cosa * cosa;
  double xi2, xl2, yi2, yl2, costh, sinth;
  double yxdist = ddd > 0.0 ? dd * std::sqrt(ddd) : 0.0;
  if (ddd > 0.0 && yxdist > 1.0e-6) {
    double cosph = yj1 / yxdist, sinph = xj1 / yxdist;
    xi2 = xi1 * cosph - yi1 * sinph;
    xl2 = xl1 * cosph - yl1 * sinph;
    yi2 = xi1 * sinph + yi1 * cosph;
    double yj2 = xj1 * sinph + yj1 * cosph;
    yl2 = xl1 * sinph + yl1 * cosph;
    costh = cosa;
    sinth = yj2 / dd;
  } else {
    xi2 = xi1;
    xl2 = xl1;
    yi2 = yi1;
    yl2 = yl1;
    costh = cosa;
    sinth = 0.0;
  }
  double yi3 = yi2 * costh - zi1 * sinth;
  double yl3 = yl2 * costh - zl1 * sinth;
  double angle = dang(xl2, yl3, xi2, yi3);
  if (angle < 0.0) angle += 2.0 * PI;
  if (angle >= MOPAC_TWO_PI) angle = 0.0;
  return angle;
}
double covrad_scaled(int z) {
  static const double COVRAD[94] = {
      0.32, 0.46, 1.20, 0.94, 0.77, 0.75, 0.71, 0.63, 0.64, 0.67, 1.40, 1.25, 1.13, 1.04, 1.10,
      1.02, 0.99, 0.96, 1.76, 1.54, 1.33, 1.22, 1.21, 1.10, 1.07, 1.04, 1.00, 0.99, 1.01, 1.09,
      1.12, 1.09, 1.15, 1.10, 1.14, 1.17, 1.89, 1.67, 1.47, 1.39, 1.32, 1.24, 1.15, 1.13, 1.13,
      1.08, 1.15, 1.23, 1.28, 1.26, 1.26, 1.23, 1.32, 1.31, 2.09, 1.76, 1.62, 1.47, 1.58, 1.57,
      1.56, 1.55, 1.51, 1.52, 1.51, 1.50, 1.49, 1.49, 1.48, 1.53, 1.46, 1.37, 1.31, 1.23, 1.18,
      1.16, 1.11, 1.12, 1.13, 1.32, 1.30, 1.30, 1.36, 1.31, 1.38, 1.42, 2.01, 1.81, 1.67, 1.58,
      1.52, 1.53, 1.54, 1.55};
  if (z <= 0 || z > 94) return 1.5 * 4.0 / 3.0;
  return COVRAD[z - 1] * 4.0 / 3.0;
}
double bonding(const std::vector<int>& nat, int x, int y) { return covrad_scaled(nat[x]) + covrad_scaled(nat[y]); }

struct HBond {
  int s[9];
  int nra, nrb;
  bool disabled;
};

std::vector<int> neighbours_capped(const std::vector<V3>& c, const std::vector<int>& nat, int atom) {
  std::vector<int> list;
  for (int j = 0; j < (int)c.size(); ++j) {
    if (j != atom && dist(c, j, atom) < bonding(nat, j, atom)) {
      list.push_back(j);
      if (list.size() == 5) {
        int worst = 0;
        double wd = -1.0;
        for (int k = 0; k < 5; ++k) {
          double d = dist(c, atom, list[k]);
          if (d > wd) {
            wd = d;
            worst = k;
          }
        }
        list.erase(list.begin() + worst);
      }
    }
  }
  return list;
}
std::vector<int> neighbours_all(const std::vector<V3>& c, const std::vector<int>& nat, int atom) {
  std::vector<int> list;
  for (int k = 0; k < (int)c.size(); ++k)
    if (k != atom && dist(c, k, atom) < bonding(nat, k, atom)) list.push_back(k);
  return list;
}
void assign_refs(const std::vector<V3>& c, const std::vector<int>& nat, const std::vector<int>& nb, int x, int h,
                 int* s, int base) {
  int i2 = base, i3 = base + 1, i4 = base + 2;
  auto dh = [&](int a) { return dist(c, a, h); };
  int n = (int)nb.size();
  if (n == 3 || n == 4) {
    std::vector<int> order(nb);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return dh(b) < dh(a); });
    s[i2] = order[0];
    s[i3] = order[1];
    s[i4] = order[2];
  } else if (n == 2) {
    int far, other;
    if (dh(nb[0]) >= dh(nb[1])) {
      far = nb[0];
      other = nb[1];
    } else {
      far = nb[1];
      other = nb[0];
    }
    s[i2] = far;
    s[i3] = other;
    s[i4] = dist(c, x, h) < bonding(nat, x, h) ? h : x;
  } else if (n == 1) {
    s[i2] = nb[0];
    auto nc = neighbours_all(c, nat, nb[0]);
    int best = nb[0];
    double bd = -1.0;
    for (int k : nc) {
      double d = dh(k);
      if (d > bd) {
        bd = d;
        best = k;
      }
    }
    s[i3] = best;
    s[i4] = dist(c, x, h) < bonding(nat, x, h) ? h : x;
  } else {
    int fill = dist(c, x, h) < bonding(nat, x, h) ? h : x;
    s[i2] = s[i3] = s[i4] = fill;
  }
}
std::vector<HBond> build_hbonds(const std::vector<V3>& c, const std::vector<int>& nat) {
  int numat = (int)c.size();
  bool any = false;
  for (int zz : nat) any |= (zz == 7 || zz == 8);
  if (!any) return {};
  const double rah2 = 1.4 * 1.4;
  std::vector<int> acceptors, bonding_h;
  std::vector<char> used(numat, 0);
  for (int i = 0; i < numat; ++i) {
    if (nat[i] == 7 || nat[i] == 8) {
      acceptors.push_back(i);
      for (int j = 0; j < numat; ++j) {
        if (nat[j] != 1) continue;
        V3 d = sub(c[i], c[j]);
        if (!used[j] && dot(d, d) < rah2) {
          bonding_h.push_back(j);
          used[j] = 1;
        }
      }
    }
  }
  const double cutoff2 = 49.0;
  std::vector<std::tuple<int, int, int>> pairs;
  std::set<std::tuple<int, int, int>> seen;
  for (int i : acceptors) {
    for (int j : bonding_h) {
      V3 dij = sub(c[i], c[j]);
      if (dot(dij, dij) >= rah2) continue;
      for (int k : acceptors) {
        if (k == i) continue;
        V3 dki = sub(c[k], c[i]);
        if (dot(dki, dki) >= cutoff2) continue;
        if (bangle(c, k, j, i) <= PI * 0.5) continue;
        auto key = std::make_tuple(j, std::min(i, k), std::max(i, k));
        if (seen.count(key)) continue;
        seen.insert(key);
        pairs.push_back({i, j, k});
      }
    }
  }
  std::vector<HBond> out;
  for (auto& t : pairs) {
    int a1 = std::get<0>(t), h = std::get<1>(t), a5 = std::get<2>(t);
    HBond b;
    for (int k = 0; k < 9; ++k) b.s[k] = 0;
    b.s[0] = a1;
    b.s[4] = a5;
    b.s[8] = h;
    auto na = neighbours_capped(c, nat, a1);
    auto nb = neighbours_capped(c, nat, a5);
    b.nra = (int)na.size();
    b.nrb = (int)nb.size();
    bool dis = false;
    for (int j : na) {
      if (j == a5) dis = true;
      for (int kk : nb)
        if (j == kk && kk != h) dis = true;
    }
    b.disabled = dis;
    if (!dis) {
      assign_refs(c, nat, na, a1, h, b.s, 1);
      assign_refs(c, nat, nb, a5, h, b.s, 5);
    }
    out.push_back(b);
  }
  return out;
}
double wrap_pi(double v) {
  if (v <= -PI) v += 2.0 * PI;
  if (v > PI) v -= 2.0 * PI;
  return v;
}
bool side_terms(const std::vector<V3>& c, const std::vector<int>& nat, int nrbonds, int heavy, int r1, int r2, int r3,
                int h, double& a2c, double& tcos) {
  double angle2_shift = 0.0, angle2_shift_2 = 0.0, torsion_shift = 0.0;
  bool tcs = false, tcs2 = false;
  if (nat[heavy] == 8) {
    if (nrbonds == 1) {
      angle2_shift = PI;
      angle2_shift_2 = deg(120.0);
      torsion_shift = 0.0;
      tcs2 = true;
    } else {
      angle2_shift = deg(109.48);
      angle2_shift_2 = angle2_shift;
      torsion_shift = deg(54.74);
    }
  } else if (nat[heavy] == 7) {
    if (nrbonds == 2) {
      angle2_shift = deg(120.0);
      angle2_shift_2 = angle2_shift;
      torsion_shift = 0.0;
    } else {
      angle2_shift = deg(109.48);
      angle2_shift_2 = angle2_shift;
      torsion_shift = deg(54.74);
      tcs = true;
    }
  }
  double torsion_check = 0.0;
  if (tcs) {
    double tc = dihed(c, r2, r1, heavy, r3);
    if (tc <= -PI) tc += 2.0 * PI;
    if (tc > PI) tc -= 2.0 * PI;
    tc = tc < 0.0 ? -PI - tc : PI - tc;
    double bac = tc;
    double tcp = tc < 0.0 ? -tc : tc;
    tcp *= 180.0 / PI;
    torsion_shift += deg((54.74 - tcp) / 54.74 * 35.26);
    angle2_shift -= deg((54.74 - tcp) / 54.74 * 19.48);
    angle2_shift_2 = angle2_shift;
    torsion_check = bac;
  }
  double angle2 = bangle(c, r1, heavy, h);
  double ac = std::cos(angle2_shift - angle2);
  double ac2 = std::cos(angle2_shift_2 - angle2);
  if (ac2 > ac) ac = ac2;
  if (ac <= 0.0) return false;
  double tcorr = dihed(c, r2, r1, heavy, h);
  if (tcorr <= -PI) tcorr += 2.0 * PI;
  if (tcorr > PI) tcorr -= 2.0 * PI;
  if (!tcs2 || std::fabs(tcorr * 180.0 / PI) > 90.0) tcorr = tcorr < 0.0 ? -PI - tcorr : PI - tcorr;
  double tc;
  if (torsion_check < 0.0) {
    tc = std::cos(wrap_pi(torsion_shift - tcorr));
  } else if (torsion_check > 0.0) {
    tc = std::cos(wrap_pi(-torsion_shift - tcorr));
  } else {
    double c1 = std::cos(wrap_pi(torsion_shift - tcorr));
    double c2 = std::cos(wrap_pi(-torsion_shift - tcorr));
    tc = c2 > c1 ? c2 : c1;
  }
  if (dist(c, h, heavy) > dist(c, h, r1) && tcs2) tc = 0.0;
  if (r2 == r3 || h == r3) tc = 1.0;
  a2c = ac;
  tcos = tc;
  return true;
}
double eh_plus(const std::vector<V3>& c, const HBond& b, const std::vector<int>& nat) {
  if (b.disabled) return 0.0;
  double a0 = A0;
  double scale_nsp3 = -0.171271 * a0 * a0, scale_osp3 = -0.098822 * a0 * a0;
  double scale_nsp2 = -0.171271 * a0 * a0, scale_osp2 = scale_osp3;
  double hartree2kcal = EV * EV_TO_KCAL;
  int s1 = b.s[0], s2 = b.s[1], s3 = b.s[2], s4 = b.s[3], s5 = b.s[4], s6 = b.s[5], s7 = b.s[6], s8 = b.s[7],
      h = b.s[8];
  double angle_cos = -std::cos(bangle(c, s1, h, s5));
  if (angle_cos <= 0.0) return 0.0;
  double a2c, tcos, a2cn, tcn;
  if (!side_terms(c, nat, b.nra, s1, s2, s3, s4, h, a2c, tcos)) return 0.0;
  if (!side_terms(c, nat, b.nrb, s5, s6, s7, s8, h, a2cn, tcn)) return 0.0;
  tcn = std::fabs(tcn);
  double sa = nat[s1] == 7 ? (b.nra >= 3 ? scale_nsp3 : scale_nsp2) : (b.nra >= 2 ? scale_osp3 : scale_osp2);
  double sb = nat[s5] == 7 ? (b.nrb >= 3 ? scale_nsp3 : scale_nsp2) : (b.nrb >= 2 ? scale_osp3 : scale_osp2);
  double sc = (sa + sb) / 2.0;
  double had = dist(c, h, s1), hbd = dist(c, h, s5);
  double xc_min = std::min(had, hbd);
  double gap = std::max(had, hbd) - xc_min;
  double damping = gap > 0.5 ? 1.0 - 1.0 / (1.0 + std::exp(-60.0 * (xc_min / COVCUT - 1.0))) : 1.0;
  double oo = dist(c, s1, s5);
  damping /= 1.0 + std::exp(-100.0 * (oo / SHORTCUT - 1.0));
  damping *= 1.0 - 1.0 / (1.0 + std::exp(-10.0 * (oo / LONGCUT - 1.0)));
  double xy = dist(c, s1, s5);
  double inner = 1.0 - a2c * tcos * a2cn * tcn;
  double e = sc / std::pow(xy, 2.0) * (angle_cos * angle_cos) * (1.0 - inner * inner) * hartree2kcal * damping;
  if (nat[s1] == 8 && nat[s5] == 8) {
    double t = std::max(xy - 2.67, 0.0);
    e += -2.5 * std::exp(-80.0 * t * t) * std::pow(angle_cos, 4);
  }
  return e;
}
}  // namespace hb

double hbond_energy(const std::vector<int>& z, const std::vector<V3>& pos) {
  std::vector<V3> c(pos.size());
  for (size_t i = 0; i < pos.size(); ++i) c[i] = {pos[i].x * A0, pos[i].y * A0, pos[i].z * A0};
  double sum = 0.0;
  for (auto& b : hb::build_hbonds(c, z)) sum += hb::eh_plus(c, b, z);
  return sum;
}

// MM corrections (src/mm_corrections.rs:46-240).
double c_triple_bond_energy(const std::vector<int>& z, const std::vector<V3>& pos) {
  double sum = 0.0;
  for (size_t a = 0; a < z.size(); ++a)
    for (size_t b = a + 1; b < z.size(); ++b) {
      if (z[a] != 6 || z[b] != 6) continue;
      double r = norm(sub(pos[b], pos[a])) * A0;
      if (r >= 1.33) continue;
      double w;
      if (r < 1.21) {
        w = 1.0;
      } else {
        double x = (r - 1.21) / (1.33 - 1.21);
        double x3 = x * x * x, x4 = std::pow(x, 4), x5 = std::pow(x, 5), x6 = std::pow(x, 6);
        double smooth = 1.0 - 10.0 * x3 + 15.0 * x4 - 6.0 * x5;
        double tail = x3 - 3.0 * x4 + 3.0 * x5 - x6;
        w = smooth + (-5.0 + 25.0 * x) * tail;
      }
      sum += w;
    }
  return sum * 12.0;
}
double si_o_h_energy(const std::vector<int>& z, const std::vector<V3>& pos) {
  int n = (int)z.size();
  double sum = 0.0;
  for (int o = 0; o < n; ++o) {
    if (z[o] != 8) continue;
    auto nearest = [&](int zz, double cut) {
      int best = -1;
      double bd = 0.0;
      for (int k = 0; k < n; ++k) {
        if (k == o || z[k] != zz) continue;
        double r = norm(sub(pos[k], pos[o])) * A0;
        if (r < cut && (best < 0 || r < bd)) {
          best = k;
          bd = r;
        }
      }
      return best;
    };
    int si = nearest(14, 2.4), h = nearest(1, 1.4);
    if (si < 0 || h < 0) continue;
    double rsio = norm(sub(pos[o], pos[si])) * A0, roh = norm(sub(pos[o], pos[h])) * A0;
    double damp = std::exp(-33.0 * std::max(rsio * rsio - 1.7 * 1.7, 0.0)) * std::exp(-68.0 * std::max(roh * roh - 1.0, 0.0));
    double ref = 125.0 * PI / 180.0;
    V3 u = sub(pos[si], pos[o]), v = sub(pos[h], pos[o]);
    double cs = std::min(1.0, std::max(-1.0, dot(u, v) / (norm(u) * norm(v))));
    double ang = std::acos(cs);
    sum += 15.0 * damp * (ang - ref) * (ang - ref);
  }
  return sum;
}

void set_msg(oracle_result* out, const std::string& s) {
  if (!out) return;
  std::snprintf(out->message, sizeof out->message, "%s", s.c_str());
}

// ---------------------------------------------------------------------------------------------
// Divide-and-conquer SCF (src/dandc/mod.rs, src/dandc/partition.rs, src/dandc/sparse.rs),
// molecular case. Written for clarity: brute-force buffer search, dense per-subsystem algebra.
// ---------------------------------------------------------------------------------------------
struct DcSub {
  std::vector<int> parent;  // core atoms (bisection order), then buffer atoms (ascending)
  int n_core = 0;
  int size() const { return (int)parent.size(); }
  double weight(int i, int j) const {  // src/dandc/partition.rs:60-67
    const bool a = i < n_core, b = j < n_core;
    return (a && b) ? 1.0 : ((a || b) ? 0.5 : 0.0);
  }
};

inline double axis_of(const V3& p, int a) { return a == 0 ? p.x : (a == 1 ? p.y : p.z); }

// src/dandc/partition.rs:136-173: split at the median of a stable sort along the axis of largest
// extent (Iterator::max_by keeps the last of equal maxima), lower half first.
void dc_bisect(std::vector<int> idx, const std::vector<V3>& pos, size_t target, std::vector<std::vector<int>>& out) {
  if (idx.size() <= target || idx.size() < 2) {
    out.push_back(std::move(idx));
    return;
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i : idx)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], axis_of(pos[i], a));
      hi[a] = std::max(hi[a], axis_of(pos[i], a));
    }
  int axis = 0;
  for (int a = 1; a < 3; ++a)
    if (hi[a] - lo[a] >= hi[axis] - lo[axis]) axis = a;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return axis_of(pos[a], axis) < axis_of(pos[b], axis); });
  const size_t half = idx.size() / 2;
  std::vector<int> right(idx.begin() + half, idx.end());
  idx.resize(half);
  dc_bisect(std::move(idx), pos, target, out);
  dc_bisect(std::move(right), pos, target, out);
}

// src/dandc/partition.rs:97-133, 175-222 (molecule: the one image t = 0).
std::vector<DcSub> dc_partition(const std::vector<V3>& pos, int core_size, double buffer) {
  const int n = (int)pos.size();
  std::vector<int> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  std::vector<std::vector<int>> cores;
  dc_bisect(all, pos, (size_t)std::max(1, core_size), cores);
  std::vector<DcSub> subs;
  const double cut2 = buffer * buffer;
  for (auto& c : cores) {
    DcSub s;
    s.parent = c;
    s.n_core = (int)c.size();
    std::vector<char> in_core(n, 0);
    for (int i : c) in_core[i] = 1;
    for (int b = 0; b < n; ++b) {
      if (in_core[b]) continue;
      for (int i : c) {
        V3 d = sub(pos[b], pos[i]);
        if (dot(d, d) <= cut2) {
          s.parent.push_back(b);
          break;
        }
      }
    }
    subs.push_back(std::move(s));
  }
  return subs;
}

// Atom-pair block storage, CSR over atoms with sorted columns (src/dandc/sparse.rs).
struct DcSparse {
  std::vector<int> norb, cols;
  std::vector<long> row_start, offs;
  std::vector<double> data;
  long slot(int a, int b) const {
    auto lo = cols.begin() + row_start[a], hi = cols.begin() + row_start[a + 1];
    auto it = std::lower_bound(lo, hi, b);
    return (it != hi && *it == b) ? (long)(it - cols.begin()) : -1;
  }
  double* block(int a, int b) {
    long k = slot(a, b);
    return k < 0 ? nullptr : data.data() + offs[k];
  }
  const double* block(int a, int b) const {
    long k = slot(a, b);
    return k < 0 ? nullptr : data.data() + offs[k];
  }
  double population(int a) const {
    const double* d = block(a, a);
    double s = 0.0;
    if (d)
      for (int i = 0; i < norb[a]; ++i) s += d[i * norb[a] + i];
    return s;
  }
};

struct DcSolution {
  std::vector<double> eps, eps_b;
  Mat c, c_b;
  bool beta = false;
  double energy = 0.0;
};

struct DcOut {
  int status = 0, converged = 0, iterations = 0, unrestricted = 0, subsystems = 0, largest = 0, largest_nao = 0;
  long stored_pairs = 0, stored_elements = 0;
  double fermi = 0.0, e_el = 0.0, dp = INFINITY;
  std::string msg;
  DcSparse density, spin;
};

inline double dc_occupation(double e, double fermi, double width) {  // src/dandc/mod.rs:649-660
  const double x = (e - fermi) / width;
  if (x > 40.0) return 0.0;
  if (x < -40.0) return 2.0;
  return 2.0 / (1.0 + std::exp(x));
}

DcOut dandc_run(const Params& P, const Mol& m, const oracle_options& o, double buffer, int core_size, double width,
                int max_scf, double p_tol, double mixing) {
  DcOut R;
  const int n = m.n();
  std::vector<DcSub> subs = dc_partition(m.pos, core_size, buffer);
  R.subsystems = (int)subs.size();
  for (auto& s : subs) {
    int nao = 0;
    for (int a : s.parent) nao += m.norb[a];
    R.largest = std::max(R.largest, s.size());
    R.largest_nao = std::max(R.largest_nao, nao);
  }
  // Pattern and normaliser (src/dandc/mod.rs:218-258).
  std::vector<std::vector<int>> nbr(n);
  for (auto& s : subs)
    for (int i = 0; i < s.size(); ++i)
      for (int j = 0; j < s.size(); ++j)
        if (s.weight(i, j) > 0.0) nbr[s.parent[i]].push_back(s.parent[j]);
  DcSparse pat;
  pat.norb = m.norb;
  pat.row_start.assign(n + 1, 0);
  long total = 0;
  for (int a = 0; a < n; ++a) {
    auto l = nbr[a];
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
    for (int b : l) {
      pat.cols.push_back(b);
      pat.offs.push_back(total);
      total += (long)m.norb[a] * m.norb[b];
    }
    pat.row_start[a + 1] = (long)pat.cols.size();
  }
  pat.offs.push_back(total);
  pat.data.assign(total, 0.0);
  R.stored_pairs = (long)pat.cols.size();
  R.stored_elements = total;
  std::vector<double> wsum(pat.cols.size(), 0.0);
  for (auto& s : subs)
    for (int i = 0; i < s.size(); ++i)
      for (int j = 0; j < s.size(); ++j) {
        const double w = s.weight(i, j);
        if (w > 0.0) wsum[pat.slot(s.parent[i], s.parent[j])] += w;
      }
  std::vector<double> normalizer(wsum.size());
  for (size_t k = 0; k < wsum.size(); ++k) normalizer[k] = wsum[k] > 1.0e-12 ? 1.0 / wsum[k] : 0.0;
  // Electron count, reference, starting densities (src/dandc/mod.rs:260-306).
  double tore = 0.0;
  for (int a = 0; a < n; ++a) tore += m.el[a]->core_charge;
  const double n_elec = tore - o.charge;
  const bool uhf = o.reference == 2 || (o.reference == 0 && o.multiplicity > 1);
  R.unrestricted = uhf;
  const double nbe = (o.multiplicity - 1.0) * 0.5;
  DcSparse dens = pat, spin = pat;
  {
    Mat guess = sad_density(m, n_elec);
    for (int a = 0; a < n; ++a) {
      double* d = dens.block(a, a);
      const int na = m.norb[a], off = m.off[a];
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < na; ++j) d[i * na + j] += guess[(size_t)(off + i) * m.nao + off + j];
    }
    if (uhf) {
      const double scale = tore > 0.0 ? 2.0 * nbe / tore : 0.0;
      for (int a = 0; a < n; ++a) {
        const int na = m.norb[a];
        const double* d = dens.block(a, a);
        double* sp = spin.block(a, a);
        for (int i = 0; i < na; ++i) sp[i * na + i] += scale * d[i * na + i];
      }
    }
  }
  // Clusters and their core Hamiltonians (geometry only, built once); the nuclear far field.
  std::vector<Mol> cl(subs.size());
  std::vector<Core> cores(subs.size());
  for (size_t s = 0; s < subs.size(); ++s) {
    std::vector<int> zz;
    std::vector<double> xx;
    for (int a : subs[s].parent) {
      zz.push_back(m.z[a]);
      xx.push_back(m.pos[a].x);
      xx.push_back(m.pos[a].y);
      xx.push_back(m.pos[a].z);
    }
    std::string err;
    build_mol(P, (int)zz.size(), zz.data(), xx.data(), cl[s], err);
  }
  std::vector<double> core_q(n);
  for (int a = 0; a < n; ++a) core_q[a] = m.el[a]->core_charge;
  auto far = [&](const std::vector<double>& q) {  // src/dandc/mod.rs:903-958 (molecule: direct sum)
    std::vector<double> v(n, 0.0);
    for (int a = 0; a < n; ++a) {
      double t = 0.0;
      for (int b = 0; b < n; ++b)
        if (b != a) t += EV * q[b] / norm(sub(m.pos[b], m.pos[a]));
      v[a] = t;
    }
    return v;
  };
  // Environment of each subsystem atom: the global potential less its own cluster
  // (src/dandc/mod.rs:961-1011).
  auto local = [&](const std::vector<double>& g, const std::vector<double>& q, size_t s) {
    const DcSub& S = subs[s];
    std::vector<double> v(S.size());
    for (int i = 0; i < S.size(); ++i) {
      double t = 0.0;
      for (int j = 0; j < S.size(); ++j) {
        if (i == j) continue;
        const double r = norm(sub(m.pos[S.parent[j]], m.pos[S.parent[i]]));
        if (r < 1.0e-8) continue;
        t += EV * q[S.parent[j]] / r;
      }
      v[i] = g[S.parent[i]] - t;
    }
    return v;
  };
  const std::vector<double> g_nuc = far(core_q);
  const int threads = g_threads;
  g_threads = 1;  // the subsystem loop is the parallel one
  auto par_subs = [&](auto&& body) {
    const long ns = (long)subs.size();
    const int t = std::max(1, std::min<int>(threads, (int)ns));
    std::vector<std::thread> pool;
    const long chunk = (ns + t - 1) / t;
    for (int k = 0; k < t; ++k) {
      const long lo = k * chunk, hi = std::min(ns, lo + chunk);
      if (lo >= hi) break;
      pool.emplace_back([lo, hi, &body]() {
        for (long s = lo; s < hi; ++s) body((size_t)s);
      });
    }
    for (auto& th : pool) th.join();
  };
  par_subs([&](size_t s) {
    cores[s] = build_core(cl[s]);
    const std::vector<double> vn = local(g_nuc, core_q, s);
    for (int i = 0; i < subs[s].size(); ++i)
      for (int k = 0; k < cl[s].norb[i]; ++k) {
        const int d = cl[s].off[i] + k;
        cores[s].h[(size_t)d * cl[s].nao + d] -= vn[i];
      }
  });
  std::vector<DcSolution> sol(subs.size());
  for (int it = 0; it < max_scf; ++it) {
    R.iterations = it + 1;
    std::vector<double> pops(n);
    for (int a = 0; a < n; ++a) pops[a] = dens.population(a);
    const std::vector<double> g_el = far(pops);
    // solve_subsystem (src/dandc/mod.rs:520-626)
    par_subs([&](size_t s) {
      const DcSub& S = subs[s];
      const Mol& cm = cl[s];
      const int nn = cm.nao;
      const std::vector<double> ve = local(g_el, pops, s);
      auto local_of = [&](const DcSparse& src) {
        Mat p((size_t)nn * nn, 0.0);
        for (int i = 0; i < S.size(); ++i)
          for (int j = 0; j < S.size(); ++j) {
            const double* b = src.block(S.parent[i], S.parent[j]);
            if (!b) continue;
            const int ni = cm.norb[i], nj = cm.norb[j];
            for (int mu = 0; mu < ni; ++mu)
              for (int nu = 0; nu < nj; ++nu) p[(size_t)(cm.off[i] + mu) * nn + cm.off[j] + nu] = b[mu * nj + nu];
          }
        return p;
      };
      auto add_env = [&](Mat& f) {
        for (int i = 0; i < S.size(); ++i)
          for (int k = 0; k < cm.norb[i]; ++k) f[(size_t)(cm.off[i] + k) * nn + cm.off[i] + k] += ve[i];
      };
      const Mat pl = local_of(dens);
      Mat f;
      Mat fb;
      if (!uhf) {
        f = build_fock_rhf(cm, cores[s], pl);
      } else {
        const Mat sl = local_of(spin);
        Mat pa(pl.size()), pb(pl.size());
        for (size_t k = 0; k < pl.size(); ++k) {
          pa[k] = 0.5 * (pl[k] + sl[k]);
          pb[k] = 0.5 * (pl[k] - sl[k]);
        }
        f = build_fock_spin(cm, cores[s], pl, pa);
        fb = build_fock_spin(cm, cores[s], pl, pb);
        add_env(fb);
      }
      add_env(f);
      DcSolution& out = sol[s];
      eigh(nn, f, out.eps, out.c);
      out.beta = uhf;
      if (uhf) eigh(nn, fb, out.eps_b, out.c_b);
      double e = 0.0;
      for (int i = 0; i < S.size(); ++i)
        for (int j = 0; j < S.size(); ++j) {
          const double w = S.weight(i, j);
          if (w == 0.0) continue;
          for (int mu = 0; mu < cm.norb[i]; ++mu)
            for (int nu = 0; nu < cm.norb[j]; ++nu) {
              const size_t x = (size_t)(cm.off[i] + mu) * nn + cm.off[j] + nu;
              e += 0.5 * w * pl[x] * (cores[s].h[x] + f[x]);
            }
        }
      out.energy = e;
    });
    // Weighted spectrum and the common Fermi level (src/dandc/mod.rs:667-770).
    std::vector<std::vector<std::pair<double, double>>> spec(subs.size());
    for (size_t s = 0; s < subs.size(); ++s) {
      const DcSub& S = subs[s];
      const Mol& cm = cl[s];
      const int nn = cm.nao;
      for (int ch = 0; ch < (uhf ? 2 : 1); ++ch) {
        const std::vector<double>& ev = ch == 0 ? sol[s].eps : sol[s].eps_b;
        const Mat& c = ch == 0 ? sol[s].c : sol[s].c_b;
        const double scale = uhf ? 0.5 : 1.0;
        for (int k = 0; k < nn; ++k) {
          double nrm = 0.0;
          for (int i = 0; i < S.n_core; ++i)
            for (int mu = 0; mu < cm.norb[i]; ++mu) {
              const double v = c[(size_t)(cm.off[i] + mu) * nn + k];
              nrm += v * v;
            }
          spec[s].push_back({ev[k], scale * nrm});
        }
      }
    }
    auto count = [&](double mu) {
      double t = 0.0;
      for (auto& st : spec) {
        double part = 0.0;
        for (auto& ew : st) {
          const double f = dc_occupation(ew.first, mu, width);
          part += f < 1.0e-14 ? 0.0 : f * ew.second;
        }
        t += part;
      }
      return t;
    };
    double lo = -500.0, hi = 500.0;
    if (count(lo) > n_elec || count(hi) < n_elec) {
      R.status = 1;
      R.msg = "no chemical potential in [-500, 500] eV holds the electron count";
      break;
    }
    for (int k = 0; k < 200; ++k) {
      const double mid = 0.5 * (lo + hi);
      if (count(mid) < n_elec) lo = mid;
      else hi = mid;
      if (hi - lo < 1.0e-12) break;
    }
    R.fermi = 0.5 * (lo + hi);
    // scatter_density, normalisation, RMS change, mixing (src/dandc/mod.rs:310-407, 772-821).
    DcSparse next = pat, next_s = pat;
    for (size_t s = 0; s < subs.size(); ++s) {
      const DcSub& S = subs[s];
      const Mol& cm = cl[s];
      const int nn = cm.nao;
      for (int spin_pass = 0; spin_pass < (uhf ? 2 : 1); ++spin_pass) {
        DcSparse& dst = spin_pass ? next_s : next;
        for (int i = 0; i < S.size(); ++i)
          for (int j = 0; j < S.size(); ++j) {
            const double w = S.weight(i, j);
            if (w == 0.0) continue;
            const int na = cm.norb[i], nb = cm.norb[j], oi = cm.off[i], oj = cm.off[j];
            std::vector<double> blk((size_t)na * nb, 0.0);
            for (int ch = 0; ch < (uhf ? 2 : 1); ++ch) {
              const double scale = !uhf ? 1.0 : (ch == 0 ? 0.5 : (spin_pass ? -0.5 : 0.5));
              const std::vector<double>& ev = ch == 0 ? sol[s].eps : sol[s].eps_b;
              const Mat& c = ch == 0 ? sol[s].c : sol[s].c_b;
              for (int k = 0; k < nn; ++k) {
                const double f = dc_occupation(ev[k], R.fermi, width) * scale;
                if (std::fabs(f) < 1.0e-14) continue;
                for (int mu = 0; mu < na; ++mu) {
                  const double cm_ = c[(size_t)(oi + mu) * nn + k];
                  for (int nu = 0; nu < nb; ++nu) blk[(size_t)mu * nb + nu] += f * cm_ * c[(size_t)(oj + nu) * nn + k];
                }
              }
            }
            double* d = dst.block(S.parent[i], S.parent[j]);
            for (size_t x = 0; x < blk.size(); ++x) d[x] += w * blk[x];
          }
      }
    }
    for (size_t k = 0; k + 1 < pat.offs.size(); ++k)
      for (long x = pat.offs[k]; x < pat.offs[k + 1]; ++x) {
        next.data[x] *= normalizer[k];
        next_s.data[x] *= normalizer[k];
      }
    double sq = 0.0;
    for (size_t x = 0; x < dens.data.size(); ++x) sq += (dens.data[x] - next.data[x]) * (dens.data[x] - next.data[x]);
    R.dp = std::sqrt(sq / (double)std::max<size_t>(dens.data.size(), 1));
    for (size_t x = 0; x < dens.data.size(); ++x) dens.data[x] = (1.0 - mixing) * dens.data[x] + mixing * next.data[x];
    if (uhf)
      for (size_t x = 0; x < spin.data.size(); ++x)
        spin.data[x] = (1.0 - mixing) * spin.data[x] + mixing * next_s.data[x];
    double e = 0.0;
    for (auto& so : sol) e += so.energy;
    R.e_el = e;
    if (R.dp < p_tol) {
      R.converged = 1;
      break;
    }
  }
  g_threads = threads;
  R.density = std::move(dens);
  if (uhf) R.spin = std::move(spin);
  return R;
}

}  // namespace

// =============================================================================================
// C API
// =============================================================================================
extern "C" {

void oracle_default_options(oracle_options* o) {
  o->method = 0;
  o->multiplicity = 1;
  o->max_scf = 200;
  o->use_diis = 1;
  o->accelerator = 2;
  o->reference = 0;
  o->threads = 0;
  o->charge = 0.0;
  o->e_tol = 1.0e-8;
  o->p_tol = 1.0e-7;
  o->adiis_switch = 0.1;
  o->level_shift = 0.0;
  o->force_dpath = 0;
  o->smooth_dispersion = 0;
  o->dipole_origin = 0;
  o->pad_ = 0;
  o->initial_density = nullptr;
  o->initial_spin_density = nullptr;
}

int oracle_nao(const char* data_dir, int n_atoms, const int* z) {
  const Params* P = load_params(data_dir, 0);
  if (!P) return -1;
  int nao = 0;
  for (int i = 0; i < n_atoms; ++i) {
    const Element* e = P->get(z[i]);
    if (!e) return -1;
    nao += e->n_orb;
  }
  return nao;
}

int oracle_element(const char* data_dir, int method, int z, double* out9) {
  const Params* P = load_params(data_dir, method);
  if (!P) return 1;
  const Element* e = P->get(z);
  if (!e) return 2;
  double v[9] = {(double)e->n_orb, e->dd, e->qq, e->rho0, e->rho1, e->rho2, e->e_isol, e->eheat_ev, e->po9()};
  std::memcpy(out9, v, sizeof v);
  return 0;
}

// Local-frame overlaps (s111, s211, s121, s221, s222) of two Slater shells sets at distance r:
// numeric != 0 forces the quadrature (also for n <= 3, to check it against the closed forms).
int oracle_overlap_locals(int na, double zsa, double zpa, int nb, double zsb, double zpb, double r, int numeric,
                          double* out5) {
  Element ea{}, eb{};
  ea.n = na;
  ea.zeta_s = zsa;
  ea.zeta_p = zpa;
  eb.n = nb;
  eb.zeta_s = zsb;
  eb.zeta_p = zpb;
  if (numeric) {
    overlap_locals_numeric(ea, eb, r, out5);
    return 0;
  }
  return overlap_locals(ea, eb, r, out5) ? 0 : 1;
}

int oracle_pair_block(const char* data_dir, int za, int zb, const double* d, double* w, double* e1b,
                      double* e2a, double* s) {
  const Params* P = load_params(data_dir, 0);
  if (!P) return 1;
  const Element* ea = P->get(za);
  const Element* eb = P->get(zb);
  if (!ea || !eb) return 2;
  if (ea->n_orb == 9 || eb->n_orb == 9) return 3;
  PairBlock pb;
  V3 dv = {d[0], d[1], d[2]};
  pair_two_electron(*ea, *eb, dv, pb);
  double ss[4][4];
  diatom_overlap(*ea, *eb, dv, ss);
  for (size_t i = 0; i < pb.w.size(); ++i) w[i] = pb.w[i];
  for (int i = 0; i < 81; ++i) e1b[i] = e2a[i] = s[i] = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      e1b[i * 9 + j] = pb.e1b[i][j];
      e2a[i * 9 + j] = pb.e2a[i][j];
      s[i * 9 + j] = ss[i][j];
    }
  return 0;
}

int oracle_hcore(const char* data_dir, int n_atoms, const int* z, const double* xyz, double* h) {
  const Params* P = load_params(data_dir, 0);
  if (!P) return 1;
  Mol m;
  std::string err;
  int st = build_mol(*P, n_atoms, z, xyz, m, err);
  if (st) return st;
  g_threads = (int)std::max(1u, std::thread::hardware_concurrency());
  Core c = build_core(m);
  std::memcpy(h, c.h.data(), c.h.size() * sizeof(double));
  return 0;
}

int oracle_fock(const char* data_dir, int n_atoms, const int* z, const double* xyz, const double* pt,
                const double* ps, double* f) {
  const Params* P = load_params(data_dir, 0);
  if (!P) return 1;
  Mol m;
  std::string err;
  int st = build_mol(*P, n_atoms, z, xyz, m, err);
  if (st) return st;
  g_threads = (int)std::max(1u, std::thread::hardware_concurrency());
  Core c = build_core(m);
  size_t nn = (size_t)m.nao * m.nao;
  Mat PT(pt, pt + nn), PS(ps, ps + nn);
  Mat F = build_fock_spin(m, c, PT, PS);
  std::memcpy(f, F.data(), nn * sizeof(double));
  return 0;
}

double oracle_core_core(const char* data_dir, int n_atoms, const int* z, const double* xyz) {
  const Params* P = load_params(data_dir, 0);
  if (!P) return NAN;
  Mol m;
  std::string err;
  if (build_mol(*P, n_atoms, z, xyz, m, err)) return NAN;
  return core_core_energy(*P, m);
}

void oracle_corrections2(int n_atoms, const int* z, const double* xyz, double* out2) {
  std::vector<int> zz(z, z + n_atoms);
  std::vector<V3> pos(n_atoms);
  for (int i = 0; i < n_atoms; ++i) pos[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  out2[0] = dispersion_energy(zz, pos, true);
  out2[1] = hh_repulsion_energy(zz, pos);
}

void oracle_corrections(int n_atoms, const int* z, const double* xyz, double* out4) {
  std::vector<int> zz(z, z + n_atoms);
  std::vector<V3> pos(n_atoms);
  for (int i = 0; i < n_atoms; ++i) pos[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  out4[0] = dispersion_energy(zz, pos);
  out4[1] = hbond_energy(zz, pos);
  out4[2] = c_triple_bond_energy(zz, pos);
  out4[3] = si_o_h_energy(zz, pos);
}

int oracle_eigh(int n, const double* a, double* w, double* v) {
  Mat A(a, a + (size_t)n * n), V;
  std::vector<double> W;
  eigh(n, A, W, V);
  std::memcpy(w, W.data(), n * sizeof(double));
  std::memcpy(v, V.data(), (size_t)n * n * sizeof(double));
  return 0;
}

// run_pm7 / run_pm7_once, molecular branch (src/scf.rs:512-867).
int oracle_run(const char* data_dir, int n_atoms, const int* z, const double* xyz,
               const oracle_options* opts, oracle_result* out, double* density, double* spin_density,
               double* eps, double* charges) {
  oracle_options o;
  if (opts)
    o = *opts;
  else
    oracle_default_options(&o);
  std::memset(out, 0, sizeof *out);
  g_threads = o.threads > 0 ? o.threads : (int)std::max(1u, std::thread::hardware_concurrency());
  if (n_atoms <= 0 || o.max_scf <= 0 || !(o.e_tol > 0) || !(o.p_tol > 0) || o.multiplicity < 1 ||
      !std::isfinite(o.charge) || o.level_shift < 0 || o.adiis_switch < 0) {
    set_msg(out, "invalid options or empty molecule");
    return out->status = 1;
  }
  for (int i = 0; i < n_atoms; ++i)
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(xyz[3 * i + k])) {
        set_msg(out, "non-finite coordinates");
        return out->status = 1;
      }
  for (int i = 0; i < n_atoms; ++i)
    for (int j = 0; j < i; ++j) {
      double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
      if (dx * dx + dy * dy + dz * dz < 1.0e-20) {
        set_msg(out, "coincident atoms");
        return out->status = 1;
      }
    }
  const Params* P = load_params(data_dir, o.method == 1 ? 1 : 0);
  if (!P) {
    set_msg(out, "parameter tables not found");
    return out->status = 1;
  }
  Mol m;
  std::string err;
  int st = build_mol(*P, n_atoms, z, xyz, m, err);
  if (st) {
    set_msg(out, err);
    return out->status = st;
  }
  out->nao = m.nao;
  out->n_atoms = n_atoms;
  double n_elec = 0.0;
  for (int i = 0; i < n_atoms; ++i) n_elec += m.el[i]->core_charge;
  n_elec -= o.charge;
  long ne = std::lround(n_elec);
  if (std::fabs(n_elec - ne) > 1.0e-6 || ne < 0) {
    set_msg(out, "invalid electron count");
    return out->status = 1;
  }
  long unp = o.multiplicity - 1;
  if (ne - unp < 0 || (ne - unp) % 2 != 0) {
    set_msg(out, "electron count incompatible with multiplicity");
    return out->status = 1;
  }
  int na = (int)((ne + unp) / 2), nb = (int)((ne - unp) / 2);
  bool uhf = o.reference == 0 ? (na != nb) : (o.reference == 2);
  if (o.reference == 1 && na != nb) {
    set_msg(out, "restricted reference needs a closed shell");
    return out->status = 1;
  }
  Core core = build_core(m, o.force_dpath != 0);
  ScfState s = uhf ? uhf_loop(m, core, na, nb, o) : rhf_loop(m, core, na, o);
  double core_ev = core_core_energy(*P, m);
  double corr = 0.0;
  if (o.method != 3) {  // correction_energy (src/scf.rs:907-932)
    corr += dispersion_energy(m.z, m.pos, o.smooth_dispersion != 0);
    corr += hbond_energy(m.z, m.pos);
    if (o.method == 4) corr += hh_repulsion_energy(m.z, m.pos);
    corr += c_triple_bond_energy(m.z, m.pos);
    corr += si_o_h_energy(m.z, m.pos);
  }
  double e_isol = 0.0, eheat = 0.0;
  for (int i = 0; i < n_atoms; ++i) {
    e_isol += m.el[i]->e_isol;
    eheat += m.el[i]->eheat_ev;
  }
  out->electronic_ev = s.e_el;
  out->core_ev = core_ev;
  out->correction_kcal = corr;
  out->total_ev = s.e_el + core_ev + corr * KCAL_TO_EV;
  out->hof_kcal = (s.e_el + core_ev - e_isol + eheat) * EV_TO_KCAL + corr;
  out->iterations = s.iterations;
  out->converged = s.converged;
  out->unrestricted = s.uhf;
  out->density_error = s.dp_err;
  out->n_occ = s.n_occ;
  out->n_occ_beta = s.uhf ? s.n_occ_b : -1;
  int nao = m.nao;
  // Charges (src/scf.rs:777-787) and dipole (src/dipole.rs:282-348, centre-of-mass origin).
  std::vector<double> q(n_atoms);
  for (int i = 0; i < n_atoms; ++i) {
    double pop = 0.0;
    for (int k = 0; k < m.norb[i]; ++k) pop += s.density[(size_t)(m.off[i] + k) * nao + m.off[i] + k];
    q[i] = m.el[i]->core_charge - pop;
  }
  double net = 0.0;
  for (double v : q) net += v;
  V3 origin = {0, 0, 0};  // dipole_origin (src/dipole.rs:137-177)
  if (std::fabs(net) > 0.5 && o.dipole_origin != 1) {
    double tot = 0.0;
    for (int i = 0; i < n_atoms; ++i) {
      double mass = o.dipole_origin == 2 ? m.el[i]->core_charge : ((z[i] >= 0 && z[i] <= 86) ? MASS[z[i]] : 0.0);
      origin.x += m.pos[i].x * mass;
      origin.y += m.pos[i].y * mass;
      origin.z += m.pos[i].z * mass;
      tot += mass;
    }
    if (tot > 0) origin = {origin.x / tot, origin.y / tot, origin.z / tot};
  }
  double pc[3] = {0, 0, 0}, sp[3] = {0, 0, 0}, pd[3] = {0, 0, 0};
  for (int i = 0; i < n_atoms; ++i) {
    pc[0] += (m.pos[i].x - origin.x) * q[i];
    pc[1] += (m.pos[i].y - origin.y) * q[i];
    pc[2] += (m.pos[i].z - origin.z) * q[i];
    if (m.el[i]->has_p()) {
      int off = m.off[i];
      double hyb = -2.0 * m.el[i]->dd;
      for (int k = 0; k < 3; ++k) sp[k] += hyb * s.density[(size_t)off * nao + off + 1 + k];
    }
    if (m.norb[i] == 9) {  // p-d hybrid term (src/dipole.rs:332-340)
      int off = m.off[i];
      auto P = [&](int a, int b) { return s.density[(size_t)(off + a) * nao + off + b]; };
      const double inv3 = 1.0 / std::sqrt(3.0);
      double hyb = -2.0 * m.el[i]->ds.ddp[5];
      pd[0] += hyb * (P(5, 3) + P(4, 1) + P(8, 2) - inv3 * P(6, 1));
      pd[1] += hyb * (P(7, 3) - P(4, 2) + P(8, 1) - inv3 * P(6, 2));
      pd[2] += hyb * (P(5, 1) + P(7, 2) + 2.0 * inv3 * P(6, 3));
    }
  }
  for (int k = 0; k < 3; ++k) {
    out->dipole_point[k] = pc[k] * AU_DIPOLE_TO_DEBYE;
    out->dipole_sp[k] = sp[k] * AU_DIPOLE_TO_DEBYE;
    out->dipole_pd[k] = pd[k] * AU_DIPOLE_TO_DEBYE;
    out->dipole[k] = out->dipole_point[k] + out->dipole_sp[k] + out->dipole_pd[k];
  }
  out->homo_ev = s.n_occ >= 1 ? s.eps[s.n_occ - 1] : NAN;
  out->lumo_ev = s.n_occ < nao ? s.eps[s.n_occ] : NAN;
  out->homo_ev_beta = NAN;
  out->lumo_ev_beta = NAN;
  if (s.uhf) {
    out->homo_ev_beta = s.n_occ_b >= 1 ? s.eps_b[s.n_occ_b - 1] : NAN;
    out->lumo_ev_beta = s.n_occ_b < nao ? s.eps_b[s.n_occ_b] : NAN;
  }
  if (density) std::memcpy(density, s.density.data(), (size_t)nao * nao * sizeof(double));
  if (spin_density) {
    if (s.uhf)
      std::memcpy(spin_density, s.spin.data(), (size_t)nao * nao * sizeof(double));
    else
      std::memset(spin_density, 0, (size_t)nao * nao * sizeof(double));
  }
  if (eps) std::memcpy(eps, s.eps.data(), nao * sizeof(double));
  if (charges) std::memcpy(charges, q.data(), n_atoms * sizeof(double));
  if (!s.converged) {
    set_msg(out, "SCF not converged");
    return out->status = 5;
  }
  return out->status = 0;
}

int oracle_dandc_partition(int n_atoms, const double* xyz, int core_size, double buffer, int* n_sub,
                           int* n_core, int* start, int* members, int cap) {
  std::vector<V3> pos(n_atoms);
  for (int i = 0; i < n_atoms; ++i) pos[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  std::vector<DcSub> subs = dc_partition(pos, core_size, buffer);
  *n_sub = (int)subs.size();
  long total = 0;
  for (auto& s : subs) total += s.size();
  if (!n_core || !start || !members) return (int)total;
  if (total > cap) return -1;
  int at = 0;
  for (size_t k = 0; k < subs.size(); ++k) {
    n_core[k] = subs[k].n_core;
    start[k] = at;
    for (int a : subs[k].parent) members[at++] = a;
  }
  start[subs.size()] = at;
  return (int)total;
}

void oracle_default_dandc_options(oracle_dandc_options* d) {  // src/dandc/mod.rs:77-89
  d->buffer_bohr = 15.0;
  d->core_size = 12;
  d->max_scf = 200;
  d->fermi_width_ev = 0.1;
  d->p_tol = 1.0e-6;
  d->mixing = 0.4;
}

int oracle_run_dandc(const char* data_dir, int n_atoms, const int* z, const double* xyz, const oracle_options* opts,
                     const oracle_dandc_options* dopts, oracle_dandc_result* out, double* dense_density,
                     double* dense_spin, double* charges) {
  oracle_options o;
  if (opts) o = *opts;
  else oracle_default_options(&o);
  oracle_dandc_options d;
  if (dopts) d = *dopts;
  else oracle_default_dandc_options(&d);
  std::memset(out, 0, sizeof *out);
  g_threads = o.threads > 0 ? o.threads : (int)std::max(1u, std::thread::hardware_concurrency());
  auto bad = [&](const char* m) {
    std::snprintf(out->message, sizeof out->message, "%s", m);
    return out->status = 1;
  };
  if (n_atoms <= 0) return bad("empty molecule");
  if (!(d.buffer_bohr > 0.0) || !std::isfinite(d.buffer_bohr) || d.core_size < 1 || !(d.fermi_width_ev > 0.0) ||
      !(d.mixing > 0.0 && d.mixing <= 1.0) || d.max_scf < 1)
    return bad("invalid divide-and-conquer options");
  const Params* P = load_params(data_dir, o.method == 1 ? 1 : 0);
  if (!P) return bad("parameter tables not found");
  Mol m;
  std::string err;
  int st = build_mol(*P, n_atoms, z, xyz, m, err);
  if (st) {
    std::snprintf(out->message, sizeof out->message, "%s", err.c_str());
    return out->status = st;
  }
  DcOut R = dandc_run(*P, m, o, d.buffer_bohr, d.core_size, d.fermi_width_ev, d.max_scf, d.p_tol, d.mixing);
  out->status = R.status;
  out->converged = R.converged;
  out->iterations = R.iterations;
  out->unrestricted = R.unrestricted;
  out->subsystems = R.subsystems;
  out->largest_subsystem = R.largest;
  out->largest_subsystem_nao = R.largest_nao;
  out->stored_pairs = R.stored_pairs;
  out->stored_elements = R.stored_elements;
  out->fermi_ev = R.fermi;
  out->electronic_ev = R.e_el;
  out->density_error = R.dp;
  out->nao = m.nao;
  std::snprintf(out->message, sizeof out->message, "%s", R.msg.c_str());
  if (R.status) return R.status;
  auto dense = [&](const DcSparse& sp, double* dst) {
    std::memset(dst, 0, sizeof(double) * (size_t)m.nao * m.nao);
    for (int a = 0; a < n_atoms; ++a)
      for (long k = sp.row_start[a]; k < sp.row_start[a + 1]; ++k) {
        const int b = sp.cols[k], na = m.norb[a], nb = m.norb[b];
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j)
            dst[(size_t)(m.off[a] + i) * m.nao + m.off[b] + j] = sp.data[sp.offs[k] + i * nb + j];
      }
  };
  if (dense_density) dense(R.density, dense_density);
  if (dense_spin) {
    if (R.unrestricted) dense(R.spin, dense_spin);
    else std::memset(dense_spin, 0, sizeof(double) * (size_t)m.nao * m.nao);
  }
  if (charges)
    for (int a = 0; a < n_atoms; ++a) charges[a] = m.el[a]->core_charge - R.density.population(a);
  return 0;
}

}  // extern "C"

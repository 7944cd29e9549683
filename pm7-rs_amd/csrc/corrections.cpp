// SPDX-License-Identifier: GPL-3.0-or-later
//
// Geometry-only post-SCF corrections added to every PM7 heat of formation (kcal/mol):
//   D2 dispersion with bond-count dependent carbon C6   (pm7-rs src/dispersion.rs:19-264)
//   EH+ hydrogen-bond term                               (pm7-rs src/hbond.rs:53-626, 1294-1322)
//   C#C and Si-O-H molecular-mechanics terms              (pm7-rs src/mm_corrections.rs:46-240)
// They depend on the geometry only, so they run once per request on the host while the SCF runs
// on the device. The hydrogen-bond topology search uses a uniform cell grid whose candidate lists
// are visited in ascending atom order, which is the order the EH+ perception rules depend on.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <set>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "host.h"

namespace pm7 {

namespace {

constexpr double PI = 3.14159265358979323846;

struct P3 {
  double x, y, z;
};
inline P3 at(const std::vector<double>& xyz, int i, double s = 1.0) {
  return {xyz[3 * i] * s, xyz[3 * i + 1] * s, xyz[3 * i + 2] * s};
}
inline double d2(P3 a, P3 b) {
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}

double covalent_radius(int z) {  // Cordero/Pyykko radii (Angstrom), geometric bond perception only
  static const double R[87] = {
      0.0,  0.31, 0.28, 1.28, 0.96, 0.84, 0.76, 0.71, 0.66, 0.57, 0.58, 1.66, 1.41, 1.21, 1.11, 1.07, 1.05, 1.02,
      1.06, 2.03, 1.76, 1.70, 1.60, 1.53, 1.39, 1.39, 1.32, 1.26, 1.24, 1.32, 1.22, 1.22, 1.20, 1.19, 1.20, 1.20,
      1.16, 2.20, 1.95, 1.90, 1.75, 1.64, 1.54, 1.47, 1.46, 1.42, 1.39, 1.45, 1.44, 1.42, 1.39, 1.39, 1.38, 1.39,
      1.40, 2.44, 2.15, 2.07, 2.04, 2.03, 2.01, 1.99, 1.98, 1.98, 1.96, 1.94, 1.92, 1.92, 1.89, 1.90, 1.87, 1.87,
      1.75, 1.70, 1.62, 1.51, 1.44, 1.41, 1.36, 1.36, 1.32, 1.45, 1.46, 1.48, 1.40, 1.50, 1.50};
  return (z >= 0 && z < 87) ? R[z] : 1.5;
}

// Spatial hash over a subset of atoms; `near` returns members of the 27 surrounding cells sorted.
struct Grid {
  double h;
  std::unordered_map<int64_t, std::vector<int>> cells;
  static int64_t key(int64_t i, int64_t j, int64_t k) {
    return ((i & 0x1FFFFF) << 42) | ((j & 0x1FFFFF) << 21) | (k & 0x1FFFFF);
  }
  Grid(const std::vector<P3>& c, const std::vector<int>& members, double spacing) : h(std::max(spacing, 1e-6)) {
    for (int m : members)
      cells[key((int64_t)std::floor(c[m].x / h), (int64_t)std::floor(c[m].y / h), (int64_t)std::floor(c[m].z / h))]
          .push_back(m);
  }
  void near(P3 p, std::vector<int>& out) const {
    out.clear();
    int64_t i = (int64_t)std::floor(p.x / h), j = (int64_t)std::floor(p.y / h), k = (int64_t)std::floor(p.z / h);
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          auto it = cells.find(key(i + a, j + b, k + c));
          if (it != cells.end()) out.insert(out.end(), it->second.begin(), it->second.end());
        }
    std::sort(out.begin(), out.end());
  }
};

// ---- EH+ geometry primitives (Angstrom) ----
const double TWO_PI_MOPAC = 6.28318530717959;
inline double dist(const std::vector<P3>& c, int a, int b) { return std::sqrt(d2(c[a], c[b])); }
double bangle(const std::vector<P3>& c, int i, int j, int k) {
  double a = d2(c[i], c[j]), b = d2(c[j], c[k]), d = d2(c[i], c[k]);
  double xy = std::sqrt(a * b);
  if (xy < 1.0e-20) return 0.0;
  return std::acos(std::clamp(0.5 * (a + b - d) / xy, -1.0, 1.0));
}
double dang(double a1, double a2, double b1, double b2) {
  const double z = 1.0e-6;
  if ((std::fabs(a1) >= z || std::fabs(a2) >= z) && (std::fabs(b1) >= z || std::fabs(b2) >= z)) {
    double an = 1.0 / std::sqrt(a1 * a1 + a2 * a2), bn = 1.0 / std::sqrt(b1 * b1 + b2 * b2);
    a1 *= an;
    a2 *= an;
    b1 *= bn;
    b2 *= bn;
    double sn = a1 * b2 - a2 * b1;
    double r = std::acos(std::clamp(a1 * b1 + a2 * b2, -1.0, 1.0));
    if (std::fabs(r) >= 4.0e-5) {
      if (sn > 0.0) r = TWO_PI_MOPAC - r;
      return -r;
    }
  }
  return 0.0;
}
double dihed(const std::vector<P3>& c, int i, int j, int k, int l) {
  const P3 K = c[k];
  double xi1 = c[i].x - K.x, xj1 = c[j].x - K.x, xl1 = c[l].x - K.x;
  double yi1 = c[i].y - K.y, yj1 = c[j].y - K.y, yl1 = c[l].y - K.y;
  double zi1 = c[i].z - K.z, zj1 = c[j].z - K.z, zl1 = c[l].z - K.z;
  double dd = std::sqrt(xj1 * xj1 + yj1 * yj1 + zj1 * zj1);
  double cosa = dd > 0.0 ? std::clamp(zj1 / dd, -1.0, 1.0) : 0.0;
  double ddd = 1.0 - cosa * cosa;
  double yx = ddd > 0.0 ? dd * std::sqrt(ddd) : 0.0;
  double xi2 = xi1, xl2 = xl1, yi2 = yi1, yl2 = yl1, sinth = 0.0;
  if (ddd > 0.0 && yx > 1.0e-6) {
    double cph = yj1 / yx, sph = xj1 / yx;
    xi2 = xi1 * cph - yi1 * sph;
    xl2 = xl1 * cph - yl1 * sph;
    yi2 = xi1 * sph + yi1 * cph;
    yl2 = xl1 * sph + yl1 * cph;
    sinth = (xj1 * sph + yj1 * cph) / dd;
  }
  double yi3 = yi2 * cosa - zi1 * sinth, yl3 = yl2 * cosa - zl1 * sinth;
  double ang = dang(xl2, yl3, xi2, yi3);
  if (ang < 0.0) ang += 2.0 * PI;
  if (ang >= TWO_PI_MOPAC) ang = 0.0;
  return ang;
}
double covrad43(int z) {  // MOPAC covrad table scaled by 4/3
  static const double C[94] = {
      0.32, 0.46, 1.20, 0.94, 0.77, 0.75, 0.71, 0.63, 0.64, 0.67, 1.40, 1.25, 1.13, 1.04, 1.10, 1.02, 0.99, 0.96, 1.76,
      1.54, 1.33, 1.22, 1.21, 1.10, 1.07, 1.04, 1.00, 0.99, 1.01, 1.09, 1.12, 1.09, 1.15, 1.10, 1.14, 1.17, 1.89, 1.67,
      1.47, 1.39, 1.32, 1.24, 1.15, 1.13, 1.13, 1.08, 1.15, 1.23, 1.28, 1.26, 1.26, 1.23, 1.32, 1.31, 2.09, 1.76, 1.62,
      1.47, 1.58, 1.57, 1.56, 1.55, 1.51, 1.52, 1.51, 1.50, 1.49, 1.49, 1.48, 1.53, 1.46, 1.37, 1.31, 1.23, 1.18, 1.16,
      1.11, 1.12, 1.13, 1.32, 1.30, 1.30, 1.36, 1.31, 1.38, 1.42, 2.01, 1.81, 1.67, 1.58, 1.52, 1.53, 1.54, 1.55};
  return (z >= 1 && z <= 94) ? C[z - 1] * 4.0 / 3.0 : 1.5 * 4.0 / 3.0;
}

struct HB {
  int s[9];
  int nra, nrb;
  bool off;
};

struct HbTopo {
  const std::vector<P3>& c;
  const std::vector<int>& z;
  Grid all;
  std::vector<int> cand;
  HbTopo(const std::vector<P3>& c_, const std::vector<int>& z_, const std::vector<int>& everyone, double span)
      : c(c_), z(z_), all(c_, everyone, span) {}
  double bond(int x, int y) const { return covrad43(z[x]) + covrad43(z[y]); }
  std::vector<int> capped(int atom) {
    std::vector<int> list;
    all.near(c[atom], cand);
    for (int j : cand) {
      if (j == atom || dist(c, j, atom) >= bond(j, atom)) continue;
      list.push_back(j);
      if (list.size() == 5) {
        int worst = 0;
        double wd = -1.0;
        for (int k = 0; k < 5; ++k)
          if (dist(c, atom, list[k]) > wd) {
            wd = dist(c, atom, list[k]);
            worst = k;
          }
        list.erase(list.begin() + worst);
      }
    }
    return list;
  }
  std::vector<int> uncapped(int atom) {
    std::vector<int> list;
    all.near(c[atom], cand);
    for (int j : cand)
      if (j != atom && dist(c, j, atom) < bond(j, atom)) list.push_back(j);
    return list;
  }
  void refs(const std::vector<int>& nb, int x, int h, int* s, int base) {
    auto dh = [&](int a) { return dist(c, a, h); };
    int fill_h = dist(c, x, h) < bond(x, h) ? h : x;
    if (nb.size() == 3 || nb.size() == 4) {
      std::vector<int> o(nb);
      std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return dh(b) < dh(a); });
      s[base] = o[0];
      s[base + 1] = o[1];
      s[base + 2] = o[2];
    } else if (nb.size() == 2) {
      bool first_far = dh(nb[0]) >= dh(nb[1]);
      s[base] = first_far ? nb[0] : nb[1];
      s[base + 1] = first_far ? nb[1] : nb[0];
      s[base + 2] = fill_h;
    } else if (nb.size() == 1) {
      s[base] = nb[0];
      int best = nb[0];
      double bd = -1.0;
      for (int k : uncapped(nb[0]))
        if (dh(k) > bd) {
          bd = dh(k);
          best = k;
        }
      s[base + 1] = best;
      s[base + 2] = fill_h;
    } else {
      s[base] = s[base + 1] = s[base + 2] = fill_h;
    }
  }
};

bool side(const std::vector<P3>& c, const std::vector<int>& z, int nrb, int heavy, int r1, int r2, int r3, int h,
          double& a2c, double& tcos) {
  auto deg = [](double d) { return PI / (180.0 / d); };
  auto wrap = [](double v) {
    if (v <= -PI) v += 2.0 * PI;
    if (v > PI) v -= 2.0 * PI;
    return v;
  };
  double sh = 0.0, sh2 = 0.0, tsh = 0.0;
  bool check = false, check2 = false;
  if (z[heavy] == 8) {
    if (nrb == 1) {
      sh = PI;
      sh2 = deg(120.0);
      check2 = true;
    } else {
      sh = sh2 = deg(109.48);
      tsh = deg(54.74);
    }
  } else if (z[heavy] == 7) {
    if (nrb == 2) {
      sh = sh2 = deg(120.0);
    } else {
      sh = sh2 = deg(109.48);
      tsh = deg(54.74);
      check = true;
    }
  }
  double tc_sign = 0.0;
  if (check) {
    double tc = wrap(dihed(c, r2, r1, heavy, r3));
    tc = tc < 0.0 ? -PI - tc : PI - tc;
    double tcp = std::fabs(tc) * 180.0 / PI;
    tsh += deg((54.74 - tcp) / 54.74 * 35.26);
    sh -= deg((54.74 - tcp) / 54.74 * 19.48);
    sh2 = sh;
    tc_sign = tc;
  }
  double ang = bangle(c, r1, heavy, h);
  double ac = std::max(std::cos(sh - ang), std::cos(sh2 - ang));
  if (ac <= 0.0) return false;
  double tcorr = wrap(dihed(c, r2, r1, heavy, h));
  if (!check2 || std::fabs(tcorr * 180.0 / PI) > 90.0) tcorr = tcorr < 0.0 ? -PI - tcorr : PI - tcorr;
  double t;
  if (tc_sign < 0.0) t = std::cos(wrap(tsh - tcorr));
  else if (tc_sign > 0.0) t = std::cos(wrap(-tsh - tcorr));
  else t = std::max(std::cos(wrap(tsh - tcorr)), std::cos(wrap(-tsh - tcorr)));
  if (dist(c, h, heavy) > dist(c, h, r1) && check2) t = 0.0;
  if (r2 == r3 || h == r3) t = 1.0;
  a2c = ac;
  tcos = t;
  return true;
}

double eh_plus(const std::vector<P3>& c, const std::vector<int>& z, const HB& b) {
  if (b.off) return 0.0;
  const double a2 = A0 * A0;
  const double nsp = -0.171271 * a2, osp = -0.098822 * a2;
  const double h2k = EV * EV_TO_KCAL;
  int s1 = b.s[0], s5 = b.s[4], h = b.s[8];
  double acos_ = -std::cos(bangle(c, s1, h, s5));
  if (acos_ <= 0.0) return 0.0;
  double a2c, tc, a2cn, tcn;
  if (!side(c, z, b.nra, s1, b.s[1], b.s[2], b.s[3], h, a2c, tc)) return 0.0;
  if (!side(c, z, b.nrb, s5, b.s[5], b.s[6], b.s[7], h, a2cn, tcn)) return 0.0;
  tcn = std::fabs(tcn);
  double sa = z[s1] == 7 ? nsp : osp, sb = z[s5] == 7 ? nsp : osp;
  double sc = (sa + sb) / 2.0;
  double ha = dist(c, h, s1), hb = dist(c, h, s5);
  double xmin = std::min(ha, hb), gap = std::max(ha, hb) - xmin;
  double damp = gap > 0.5 ? 1.0 - 1.0 / (1.0 + std::exp(-60.0 * (xmin / 1.2 - 1.0))) : 1.0;
  double oo = dist(c, s1, s5);
  damp /= 1.0 + std::exp(-100.0 * (oo / 2.4 - 1.0));
  damp *= 1.0 - 1.0 / (1.0 + std::exp(-10.0 * (oo / 7.0 - 1.0)));
  double inner = 1.0 - a2c * tc * a2cn * tcn;
  double e = sc / (oo * oo) * (acos_ * acos_) * (1.0 - inner * inner) * h2k * damp;
  if (z[s1] == 8 && z[s5] == 8) {
    double t = std::max(oo - 2.67, 0.0);
    e += -2.5 * std::exp(-80.0 * t * t) * std::pow(acos_, 4);
  }
  return e;
}

}  // namespace

namespace {
// D2 dispersion energy (src/dispersion.rs:210-283); with g, also its gradient (kcal/mol/Bohr,
// accumulated into g[3n]).
//  * Discrete path (MOPAC's): a carbon with four covalent neighbours takes C6 = 0.95, any other
//    1.65. The counts are step functions of the distances and contribute nothing to the gradient,
//    so each pair's term differentiates through its distance only (at fixed bond topology).
//  * Smooth path (smooth_dispersion, src/dispersion.rs:72-175, 416-500): each neighbour counts
//    w = 1 / (1 + exp(-26 (r_cut / r - 1))), and the carbon C6 is 1.65 - 0.70 s with
//    s = 1 / (1 + exp(-40 (count - 3.5))); the gradient then has a second part, dE/dcount_i
//    (from each pair's Slater-Kirkwood combination) times dcount_i/dR over the counting pairs.
double dispersion_impl(const std::vector<int>& z, const std::vector<double>& xyz, double* g, bool smooth) {
  auto c6t = [](int q) {
    switch (q) {
      case 1: return 0.16; case 2: return 0.084; case 5: return 5.79; case 6: return 1.65; case 7: return 1.11;
      case 8: return 0.70; case 9: return 0.57; case 10: return 0.45; case 15: return 3.25; case 16: return 5.79;
      case 17: return 5.97; case 18: return 3.71; case 27: return 0.04; case 35: return 11.60; case 36: return 4.47;
      case 53: return 25.80; case 54: return 16.50; default: return 0.0;
    }
  };
  auto r0t = [](int q) {
    switch (q) {
      case 1: return 156.0; case 2: return 140.0; case 5: return 180.0; case 6: return 170.0; case 7: return 155.0;
      case 8: return 152.0; case 9: return 147.0; case 10: return 154.0; case 15: return 180.0; case 16: return 180.0;
      case 17: return 175.0; case 18: return 188.0; case 28: return 140.0; case 35: return 185.0; case 36: return 202.0;
      case 53: return 198.0; case 54: return 216.0; default: return 0.0;
    }
  };
  auto nft = [](int q) {
    switch (q) {
      case 1: return 0.80; case 2: return 1.42; case 5: return 2.16; case 6: return 2.50; case 7: return 2.82;
      case 8: return 3.15; case 9: return 3.48; case 10: return 3.81; case 15: return 4.50; case 16: return 4.80;
      case 17: return 5.10; case 18: return 5.40; case 27: return 2.90; case 35: return 6.00; case 36: return 6.30;
      case 53: return 6.95; case 54: return 7.25; default: return 0.0;
    }
  };
  constexpr double COUNT_K = 26.0, SWITCH_K = 40.0;
  const int n = (int)z.size();
  auto rcut = [&](int i, int j) { return 1.3 * (covalent_radius(z[i]) + covalent_radius(z[j])) / A0; };  // Bohr
  // Neighbour weight 1/(1 + exp(-k (r_cut/r - 1))) and its r-derivative (0 where exp overflows).
  auto weight = [&](double r, double rc, double* dw) {
    const double e = std::exp(-COUNT_K * (rc / r - 1.0));
    if (!std::isfinite(e)) {
      *dw = 0.0;
      return 0.0;
    }
    const double w = 1.0 / (1.0 + e);
    *dw = -w * (1.0 - w) * COUNT_K * rc / (r * r);
    return w;
  };
  std::vector<double> c6a(n, 0.0), dc6(n, 0.0);
  {
    std::vector<double> cnt(n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        const double r = std::sqrt(d2(at(xyz, i), at(xyz, j)));
        double w;
        if (smooth) {
          double dw;
          w = weight(r, rcut(i, j), &dw);
        } else {
          w = (r * A0 < 1.3 * (covalent_radius(z[i]) + covalent_radius(z[j]))) ? 1.0 : 0.0;
        }
        cnt[i] += w;
        cnt[j] += w;
      }
    for (int i = 0; i < n; ++i) {
      if (z[i] != 6) {
        c6a[i] = c6t(z[i]);
      } else if (!smooth) {
        c6a[i] = cnt[i] == 4.0 ? 0.95 : 1.65;
      } else {
        const double sw = 1.0 / (1.0 + std::exp(-(cnt[i] - 3.5) * SWITCH_K));
        c6a[i] = 1.65 + (0.95 - 1.65) * sw;
        dc6[i] = (0.95 - 1.65) * sw * (1.0 - sw) * SWITCH_K;
      }
    }
  }
  const double ALPHA = 15.450118, S = 1.226593, CSCALE = 2.286419;
  std::vector<double> dedcn(smooth && g ? n : 0, 0.0);
  double sum = 0.0;
  for (int i = 0; i < n; ++i) {
    int zi = z[i];
    if (c6t(zi) == 0.0 || r0t(zi) == 0.0 || nft(zi) == 0.0) continue;
    for (int j = i + 1; j < n; ++j) {
      int zj = z[j];
      if (c6t(zj) == 0.0 || r0t(zj) == 0.0 || nft(zj) == 0.0) continue;
      double ni = nft(zi), nj = nft(zj);
      double ci = c6a[i], cj = c6a[j];
      const double bi = std::cbrt(ci * nj * nj), bj = std::cbrt(cj * ni * ni);
      double c6 = 2.0 * std::cbrt(ci * ci * cj * cj * ni * nj) / (bi + bj);
      double ri = r0t(zi), rj = r0t(zj);
      double r0 = (ri * ri * ri + rj * rj * rj) / (ri * ri + rj * rj) / 1000.0 * 2.0;
      double rnm = std::sqrt(d2(at(xyz, i), at(xyz, j))) * (A0 * 0.1);
      double denom = std::exp((rnm / (S * r0) - 1.0) * -ALPHA) + 1.0;
      const double e = -CSCALE * (c6 / (1000.0 * 4.184)) / (std::pow(rnm, 6) * denom);
      sum += e;
      if (g) {
        // E = -K / (rnm^6 D), D = 1 + exp(-ALPHA (rnm / (S r0) - 1)):
        // dE/drnm = K (6 / rnm + D' / D) / (rnm^6 D), D' = -(ALPHA / (S r0)) (D - 1).
        const double k = CSCALE * (c6 / (1000.0 * 4.184));
        const double dd = -(ALPHA / (S * r0)) * (denom - 1.0);
        const double de_drnm = k * (6.0 / rnm + dd / denom) / (std::pow(rnm, 6) * denom);
        const double r_bohr = rnm / (A0 * 0.1);
        const double f = de_drnm * (A0 * 0.1) / r_bohr;
        for (int c = 0; c < 3; ++c) {
          const double dx = xyz[3 * i + c] - xyz[3 * j + c];
          g[3 * i + c] += f * dx;
          g[3 * j + c] -= f * dx;
        }
        if (smooth) {
          // E is linear in the combined C6, so dE/dc6_i = E d ln C6 / dc6_i, with
          // d ln C6 / dc6_i = (2/3 - b_i / (3 (b_i + b_j))) / c6_i (Slater-Kirkwood rule).
          const double sb = bi + bj;
          if (sb > 0.0 && ci > 0.0 && cj > 0.0) {
            dedcn[i] += e * (2.0 / 3.0 - bi / (3.0 * sb)) / ci * dc6[i];
            dedcn[j] += e * (2.0 / 3.0 - bj / (3.0 * sb)) / cj * dc6[j];
          }
        }
      }
    }
  }
  if (g && smooth) {
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        if (dedcn[i] == 0.0 && dedcn[j] == 0.0) continue;
        const double r = std::sqrt(d2(at(xyz, i), at(xyz, j)));
        double dw;
        (void)weight(r, rcut(i, j), &dw);
        if (dw == 0.0) continue;
        const double f = (dedcn[i] + dedcn[j]) * dw / r;
        for (int c = 0; c < 3; ++c) {
          const double dx = xyz[3 * i + c] - xyz[3 * j + c];
          g[3 * i + c] += f * dx;
          g[3 * j + c] -= f * dx;
        }
      }
  }
  return sum;
}

// MOPAC's H-H repulsion (src/hh_rep.rs:21-36) for r in Angstrom, kcal/mol, and dE/dr.
double hh_poly(double r, double* d) {
  if (r <= 1.0) {
    *d = 0.0;
    return 25.46293603147693;
  }
  if (r < 1.5) {
    const double c5 = -2714.9523516034697, c4 = 17103.650110591705, c3 = -42511.85798221796,
                 c2 = 52063.19679913834, c1 = -31430.65833597229, c0 = 7516.08469609514;
    *d = (((5.0 * c5 * r + 4.0 * c4) * r + 3.0 * c3) * r + 2.0 * c2) * r + c1;
    return ((((c5 * r + c4) * r + c3) * r + c2) * r + c1) * r + c0;
  }
  const double p = std::pow(r, 1.72905), e = 118.7326 * std::exp(-1.53965 * p);
  *d = e * (-1.53965 * 1.72905 * p / r);
  return e;
}
}  // namespace

double dispersion_kcal(const std::vector<int>& z, const std::vector<double>& xyz, bool smooth) {
  return dispersion_impl(z, xyz, nullptr, smooth);
}

void dispersion_grad_kcal(const std::vector<int>& z, const std::vector<double>& xyz, double* g, bool smooth) {
  (void)dispersion_impl(z, xyz, g, smooth);
}

// PM7-HH: sum over H-H pairs of hh_poly(r) (src/hh_rep.rs:48-60).
double hh_repulsion_kcal(const std::vector<int>& z, const std::vector<double>& xyz) {
  double e = 0.0, d;
  const int n = (int)z.size();
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (z[i] == 1 && z[j] == 1) e += hh_poly(std::sqrt(d2(at(xyz, i), at(xyz, j))) * A0, &d);
  return e;
}

void hh_repulsion_grad_kcal(const std::vector<int>& z, const std::vector<double>& xyz, double* g) {
  const int n = (int)z.size();
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (z[i] != 1 || z[j] != 1) continue;
      const double r = std::sqrt(d2(at(xyz, i), at(xyz, j)));
      double d;
      (void)hh_poly(r * A0, &d);
      const double f = d * A0 / r;  // dE/dr_bohr / r
      for (int c = 0; c < 3; ++c) {
        const double dx = xyz[3 * i + c] - xyz[3 * j + c];
        g[3 * i + c] += f * dx;
        g[3 * j + c] -= f * dx;
      }
    }
}

double hbond_kcal(const std::vector<int>& z, const std::vector<double>& xyz) {
  const int n = (int)z.size();
  std::vector<P3> c(n);
  for (int i = 0; i < n; ++i) c[i] = at(xyz, i, A0);
  std::vector<int> acc, hyd, everyone(n);
  double maxr = 0.0;
  for (int i = 0; i < n; ++i) {
    everyone[i] = i;
    if (z[i] == 7 || z[i] == 8) acc.push_back(i);
    if (z[i] == 1) hyd.push_back(i);
    maxr = std::max(maxr, covrad43(z[i]));
  }
  if (acc.empty()) return 0.0;
  const double rah2 = 1.4 * 1.4;
  Grid hgrid(c, hyd, 1.4), agrid(c, acc, 7.0);
  std::vector<int> bonding_h, cand, nearacc;
  std::vector<int> slot_of(n, -1);
  std::vector<char> used(n, 0);
  for (int i : acc) {
    hgrid.near(c[i], cand);
    for (int j : cand)
      if (!used[j] && d2(c[i], c[j]) < rah2) {
        slot_of[j] = (int)bonding_h.size();
        bonding_h.push_back(j);
        used[j] = 1;
      }
  }
  std::vector<std::tuple<int, int, int>> pairs;
  std::set<std::tuple<int, int, int>> seen;
  std::vector<int> slots;
  for (int i : acc) {
    hgrid.near(c[i], cand);
    slots.clear();
    for (int j : cand)
      if (slot_of[j] >= 0) slots.push_back(slot_of[j]);
    std::sort(slots.begin(), slots.end());
    agrid.near(c[i], nearacc);
    for (int sl : slots) {
      int j = bonding_h[sl];
      if (d2(c[i], c[j]) >= rah2) continue;
      for (int k : nearacc) {
        if (k == i || d2(c[k], c[i]) >= 49.0) continue;
        if (bangle(c, k, j, i) <= PI * 0.5) continue;
        if (!seen.insert({j, std::min(i, k), std::max(i, k)}).second) continue;
        pairs.push_back({i, j, k});
      }
    }
  }
  HbTopo topo(c, z, everyone, 2.0 * maxr);
  double sum = 0.0;
  for (auto& t : pairs) {
    HB b{};
    int a1 = std::get<0>(t), h = std::get<1>(t), a5 = std::get<2>(t);
    b.s[0] = a1;
    b.s[4] = a5;
    b.s[8] = h;
    auto na = topo.capped(a1), nb = topo.capped(a5);
    b.nra = (int)na.size();
    b.nrb = (int)nb.size();
    b.off = false;
    for (int j : na) {
      if (j == a5) b.off = true;
      for (int k : nb)
        if (j == k && k != h) b.off = true;
    }
    if (!b.off) {
      topo.refs(na, a1, h, b.s, 1);
      topo.refs(nb, a5, h, b.s, 5);
    }
    sum += eh_plus(c, z, b);
  }
  return sum;
}

double c_triple_kcal(const std::vector<int>& z, const std::vector<double>& xyz) {
  double sum = 0.0;
  const int n = (int)z.size();
  for (int a = 0; a < n; ++a) {
    if (z[a] != 6) continue;
    for (int b = a + 1; b < n; ++b) {
      if (z[b] != 6) continue;
      double r = std::sqrt(d2(at(xyz, a), at(xyz, b))) * A0;
      if (r >= 1.33) continue;
      if (r < 1.21) {
        sum += 1.0;
        continue;
      }
      double x = (r - 1.21) / 0.12, x3 = x * x * x, x4 = x3 * x, x5 = x4 * x, x6 = x5 * x;
      sum += (1.0 - 10.0 * x3 + 15.0 * x4 - 6.0 * x5) + (-5.0 + 25.0 * x) * (x3 - 3.0 * x4 + 3.0 * x5 - x6);
    }
  }
  return 12.0 * sum;
}

double si_o_h_kcal(const std::vector<int>& z, const std::vector<double>& xyz) {
  const int n = (int)z.size();
  double sum = 0.0;
  for (int o = 0; o < n; ++o) {
    if (z[o] != 8) continue;
    auto nearest = [&](int want, double cut) {
      int best = -1;
      double bd = 0.0;
      for (int k = 0; k < n; ++k) {
        if (k == o || z[k] != want) continue;
        double r = std::sqrt(d2(at(xyz, k), at(xyz, o))) * A0;
        if (r < cut && (best < 0 || r < bd)) {
          best = k;
          bd = r;
        }
      }
      return best;
    };
    int si = nearest(14, 2.4), h = nearest(1, 1.4);
    if (si < 0 || h < 0) continue;
    P3 O = at(xyz, o), Si = at(xyz, si), Hh = at(xyz, h);
    double rs = std::sqrt(d2(O, Si)) * A0, rh = std::sqrt(d2(O, Hh)) * A0;
    double damp = std::exp(-33.0 * std::max(rs * rs - 1.7 * 1.7, 0.0)) * std::exp(-68.0 * std::max(rh * rh - 1.0, 0.0));
    P3 u = {Si.x - O.x, Si.y - O.y, Si.z - O.z}, v = {Hh.x - O.x, Hh.y - O.y, Hh.z - O.z};
    double nu = std::sqrt(u.x * u.x + u.y * u.y + u.z * u.z), nv = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    double ang = std::acos(std::clamp((u.x * v.x + u.y * v.y + u.z * v.z) / (nu * nv), -1.0, 1.0));
    double ref = 125.0 * PI / 180.0;
    sum += 15.0 * damp * (ang - ref) * (ang - ref);
  }
  return sum;
}

}  // namespace pm7

// SPDX-License-Identifier: GPL-3.0-or-later
//
// MNDO/d two-centre kernels for molecules holding a d-bearing atom (pm7-rs
// src/mndod_twocenter.rs, src/overlap_d.rs; every pair of such a molecule takes them,
// src/hamiltonian.rs:347-354). Included by integrals.hip, whose sp helpers (local_xx, local_xh,
// feather, pair_core) it reuses.
//
// One 64-thread workgroup per atom pair: the orbital rotation R (9 x 9), the local integrals L
// (npair_f x npair_s, up to 45 x 45) and the half-rotated block live in LDS; the lanes share the
// multipole integrals, the two rotation passes, the electron-core blocks and the overlaps.
// Local frame: first p orbital along the bond (sigma), then the two pi; local d orbitals
// d-sigma, d-pi, d-pi', d-delta, d-delta'. Charge distributions are packed h(h+1)/2 + l in both
// frames over (s, p1, p2, p3, d1..d5).
#pragma once

namespace pm7 {

// ---- tables ---------------------------------------------------------------------------------
// Multipole content of each local charge distribution: coefficient of the (l, m) multipole,
// truncated at l = 2 (MNDO/d; the decimals are MOPAC's).
struct MpTerm {
  signed char a, b, l, m;
  double c;
};
constexpr MpTerm MP_TERMS[] = {
    {0, 0, 0, 0, 1.0},         {1, 0, 1, 0, 1.0},         {2, 0, 1, 1, 1.0},         {3, 0, 1, -1, 1.0},
    {4, 0, 2, 0, 1.1547005},   {5, 0, 2, 1, 1.0},         {6, 0, 2, -1, 1.0},        {7, 0, 2, 2, 1.0},
    {8, 0, 2, -2, 1.0},        {1, 1, 0, 0, 1.0},         {1, 1, 2, 0, 1.3333333},   {2, 1, 2, 1, 1.0},
    {3, 1, 2, -1, 1.0},        {4, 1, 1, 0, 1.1547005},   {5, 1, 1, 1, 1.0},         {6, 1, 1, -1, 1.0},
    {2, 2, 0, 0, 1.0},         {2, 2, 2, 0, -0.66666667}, {2, 2, 2, 2, 1.0},         {3, 2, 2, -2, 1.0},
    {4, 2, 1, 1, -0.57735027}, {5, 2, 1, 0, 1.0},         {7, 2, 1, 1, 1.0},         {8, 2, 1, -1, 1.0},
    {3, 3, 0, 0, 1.0},         {3, 3, 2, 0, -0.66666667}, {3, 3, 2, 2, -1.0},        {4, 3, 1, -1, -0.57735027},
    {6, 3, 1, 0, 1.0},         {7, 3, 1, -1, -1.0},       {8, 3, 1, 1, 1.0},         {4, 4, 0, 0, 1.0},
    {4, 4, 2, 0, 1.3333333},   {5, 4, 2, 1, 0.57735027},  {6, 4, 2, -1, 0.57735027}, {7, 4, 2, 2, -1.1547005},
    {8, 4, 2, -2, -1.1547005}, {5, 5, 0, 0, 1.0},         {5, 5, 2, 0, 0.66666667},  {5, 5, 2, 2, 1.0},
    {6, 5, 2, -2, 1.0},        {7, 5, 2, 1, 1.0},         {8, 5, 2, -1, 1.0},        {6, 6, 0, 0, 1.0},
    {6, 6, 2, 0, 0.66666667},  {6, 6, 2, 2, -1.0},        {7, 6, 2, -1, -1.0},       {8, 6, 2, 1, 1.0},
    {7, 7, 0, 0, 1.0},         {7, 7, 2, 0, -1.3333333},  {8, 8, 0, 0, 1.0},         {8, 8, 2, 0, -1.3333333}};

struct DTables {
  double ch[45][3][5];  // [packed pair][l][m + 2]
  signed char pa[45], pb[45], cls[45];  // orbitals (a >= b) of a packed pair, its shell class 1..6
  signed char sp_local[10][10];          // local (ab|cd) over s/p distributions -> ri index
  double sp_sign[23];
  double fact[25];
};

constexpr int shell_of(int orb) { return orb == 0 ? 0 : (orb <= 3 ? 1 : 2); }

constexpr DTables make_dtables() {
  DTables t{};
  for (const MpTerm& m : MP_TERMS) t.ch[m.a * (m.a + 1) / 2 + m.b][m.l][m.m + 2] = m.c;
  for (int a = 0; a < 9; ++a)
    for (int b = 0; b <= a; ++b) {
      const int p = a * (a + 1) / 2 + b;
      t.pa[p] = (signed char)a;
      t.pb[p] = (signed char)b;
      const int h = shell_of(a) + 1, l = shell_of(b) + 1;
      t.cls[p] = (signed char)(h * (h - 1) / 2 + l);  // 1 ss, 2 sp, 3 pp, 4 sd, 5 pd, 6 dd
    }
  // The standard MNDO layout of the 22 sp local integrals over (s, p-sigma, p-pi, p-pi').
  const signed char sp[10][10] = {{1, 5, 11, 0, 0, 12, 0, 0, 0, 12}, {2, 6, 13, 0, 0, 14, 0, 0, 0, 14},
                                  {3, 8, 16, 0, 0, 18, 0, 0, 0, 18}, {0, 0, 0, 7, 15, 0, 0, 0, 0, 0},
                                  {0, 0, 0, 10, 20, 0, 0, 0, 0, 0},  {4, 9, 17, 0, 0, 19, 0, 0, 0, 21},
                                  {0, 0, 0, 0, 0, 0, 7, 15, 0, 0},   {0, 0, 0, 0, 0, 0, 10, 20, 0, 0},
                                  {0, 0, 0, 0, 0, 0, 0, 0, 22, 0},   {4, 9, 17, 0, 0, 21, 0, 0, 0, 19}};
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 10; ++j) t.sp_local[i][j] = sp[i][j];
  // Sign of each sp integral in this local frame relative to the sp kernel's ri.
  const double sg[23] = {0, 1, -1, 1, 1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1};
  for (int i = 0; i < 23; ++i) t.sp_sign[i] = sg[i];
  t.fact[0] = 1.0;
  for (int i = 1; i < 25; ++i) t.fact[i] = t.fact[i - 1] * i;
  return t;
}

__constant__ DTables c_dt = make_dtables();

// ---- local two-centre integrals ---------------------------------------------------------------
// Interaction of two point-charge multipoles in Hartree (pm7-rs src/mndod.rs:30-186): order l1
// on the first atom, l2 on the second, azimuthal |m|, separations da db, additive term add.
__device__ double multipole_pair(double r, int l1, int l2, int m, double da, double db, double add) {
  auto q = [add](double x, double s2) { return 1.0 / sqrt(x * x + (s2 + add)); };
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
    const double ab = db / sqrt(2.0), u = (da - ab) * (da - ab), v = (da + ab) * (da + ab);
    return (q(r - ab, u) * -2.0 + q(r + ab, u) * 2.0 + q(r - ab, v) * 2.0 - q(r + ab, v) * 2.0) * 0.125;
  }
  if (l1 == 2 && l2 == 1) {
    if (m == 0)
      return (-q(r - da - db, 0.0) + q(r - db, da * da) * 2.0 - q(r + da - db, 0.0) + q(r - da + db, 0.0) -
              q(r + db, da * da) * 2.0 + q(r + da + db, 0.0)) * 0.125;
    const double aa = da / sqrt(2.0), u = (aa - db) * (aa - db), v = (aa + db) * (aa + db);
    return (q(r + aa, u) * -2.0 + q(r - aa, u) * 2.0 + q(r + aa, v) * 2.0 - q(r - aa, v) * 2.0) * 0.125;
  }
  const double xyxy = q(r, dm2) * 4.0 + q(r, dp2) * 4.0 - q(r, da * da + db * db) * 8.0;
  if (m == 0) {
    const double zzzz = q(r - da - db, 0.0) + q(r + da + db, 0.0) + q(r - da + db, 0.0) + q(r + da - db, 0.0) -
                        q(r - da, db * db) * 2.0 - q(r - db, da * da) * 2.0 - q(r + da, db * db) * 2.0 -
                        q(r + db, da * da) * 2.0 + q(r, dm2) * 2.0 + q(r, dp2) * 2.0;
    return zzzz * (1.0 / 16.0) - xyxy * (1.0 / 64.0);
  }
  if (m == 1) {
    const double aa = da / sqrt(2.0), ab = db / sqrt(2.0);
    const double u = (aa - ab) * (aa - ab), v = (aa + ab) * (aa + ab);
    return (q(r + aa - ab, u) * 2.0 - q(r + aa + ab, u) * 2.0 - q(r - aa - ab, u) * 2.0 + q(r - aa + ab, u) * 2.0 -
            q(r + aa - ab, v) * 2.0 + q(r + aa + ab, v) * 2.0 + q(r - aa - ab, v) * 2.0 - q(r - aa + ab, v) * 2.0) *
           (1.0 / 16.0);
  }
  return xyxy * (1.0 / 16.0);
}

// Local (ab | cd) in Hartree by the multipole expansion (pm7-rs src/mndod_twocenter.rs:582-654).
// core_side 1 (2): the monopole of an (ss) distribution on the first (second) atom takes po(9).
__device__ double local_multipole(const ElemDev& A, const ElemDev& B, int pij, int pkl, int core_side, double r) {
  const int a = c_dt.pa[pij], b = c_dt.pb[pij], c = c_dt.pa[pkl], d = c_dt.pb[pkl];
  const int la = shell_of(a), lb = shell_of(b), lc = shell_of(c), ld = shell_of(d);
  const int kij = c_dt.cls[pij], kkl = c_dt.cls[pkl];
  auto mono = [](const ElemDev& D, int kind, bool core) {
    if (kind == 1) return core ? D.po[9] : D.po[1];
    if (kind == 3) return D.po[7];
    if (kind == 6) return D.po[8];
    return D.po[1];
  };
  double sum = 0.0;
  for (int l1 = min(abs(la - lb), 2); l1 <= min(la + lb, 2); ++l1) {
    const double dij = l1 == 0 ? 0.0 : A.ddp[kij];
    const double aij = l1 == 0 ? mono(A, kij, core_side == 1) : A.po[kij];
    for (int l2 = min(abs(lc - ld), 2); l2 <= min(lc + ld, 2); ++l2) {
      const double dkl = l2 == 0 ? 0.0 : B.ddp[kkl];
      const double akl = l2 == 0 ? mono(B, kkl, core_side == 2) : B.po[kkl];
      const double add = (aij + akl) * (aij + akl);
      const int mm = min(l1, l2);
      for (int m = -mm; m <= mm; ++m) {
        const double ccc = c_dt.ch[pij][l1][m + 2] * c_dt.ch[pkl][l2][m + 2];
        if (ccc == 0.0) continue;
        sum += multipole_pair(r, l1, l2, abs(m), dij, dkl, add) * ccc;
      }
    }
  }
  return sum;
}

// The 22 sp local integrals with the po(1..3) additive terms, feathered, in the MNDO/d sign
// convention (1-based; pm7-rs src/mndod_twocenter.rs:244-470).
__device__ void local_sp_ri(const ElemDev& ei, const ElemDev& ej, double r, double ri[23]) {
  ElemDev a = ei, b = ej;
  a.rho0 = ei.po[1], a.rho1 = ei.po[2], a.rho2 = ei.po[3];
  b.rho0 = ej.po[1], b.rho1 = ej.po[2], b.rho2 = ej.po[3];
  for (int k = 0; k < 23; ++k) ri[k] = 0.0;
  const bool si = ei.n_orb >= 4, sj = ej.n_orb >= 4;
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
    local_xh(b, a, r, t);
    ri[1] = t[0];
    ri[5] = -t[1];
    ri[11] = t[2];
    ri[12] = t[3];
  } else {
    double cf, pt;
    feather(r, cf, pt);
    ri[1] = EV / sqrt(r * r + sq(a.rho0 + b.rho0)) * cf + pt * (1.0 - cf);
  }
  for (int k = 1; k <= 22; ++k) ri[k] *= c_dt.sp_sign[k];
}

// Local electron-core terms core[k][side], k 1..4, of the s/p shells (src/mndod_twocenter.rs:482-572).
__device__ void local_core_sp(const ElemDev& ei, const ElemDev& ej, double r, double core[11][3]) {
  const double r2 = r * r, aci = ei.po[9], acj = ej.po[9];
  core[1][1] = -(1.0 / sqrt(r2 + (acj + ei.po[1]) * (acj + ei.po[1]))) * (ej.core_charge * EV);
  core[1][2] = -(1.0 / sqrt(r2 + (aci + ej.po[1]) * (aci + ej.po[1]))) * (ei.core_charge * EV);
  auto shell = [&](const ElemDev& X, double ac, double& t2, double& t3, double& t4) {
    const double pp = (ac + X.po[7]) * (ac + X.po[7]);
    const double da = X.ddp[2], qa = X.ddp[3] / sqrt(2.0), q2 = qa + qa;
    const double ad = (X.po[2] + ac) * (X.po[2] + ac), aq = (X.po[3] + ac) * (X.po[3] + ac);
    const double x1 = 1.0 / sqrt(r2 + pp);
    const double x2 = 1.0 / sqrt(r2 + aq) * -0.5;
    const double x3 = 1.0 / sqrt((r + da) * (r + da) + ad) * -0.5;
    const double x4 = 1.0 / sqrt((r - da) * (r - da) + ad) * 0.5;
    const double x5 = 1.0 / sqrt((r - q2) * (r - q2) + aq) * 0.25;
    const double x6 = 1.0 / sqrt((r + q2) * (r + q2) + aq) * 0.25;
    const double x7 = 1.0 / sqrt(r2 + (q2 * q2 + aq)) * 0.5;
    t2 = (x3 + x4) * EV;
    t3 = (x1 + x2 + x5 + x6) * EV;
    t4 = (x1 + x2 + x7) * EV;
  };
  if (ei.n_orb >= 4) {
    double t2, t3, t4;
    shell(ei, acj, t2, t3, t4);
    core[2][1] = -t2 * ej.core_charge;
    core[3][1] = -t3 * ej.core_charge;
    core[4][1] = -t4 * ej.core_charge;
  }
  if (ej.n_orb >= 4) {
    double t2, t3, t4;
    shell(ej, aci, t2, t3, t4);
    core[2][2] = t2 * ei.core_charge;  // the dipole term changes sign seen from the second atom
    core[3][2] = -t3 * ei.core_charge;
    core[4][2] = -t4 * ei.core_charge;
  }
}

// Orbital rotation R[local][molecular] (row-major 9 x 9) for v = R_s - R_f
// (src/mndod_twocenter.rs:73-147).
__device__ void orbital_rotation(const double v[3], double r, bool with_d, double* R) {
  const double pt5sq3 = 0.8660254037841;
  const double b = v[0] * v[0] + v[1] * v[1];
  const double sqb = sqrt(b);
  double sb = sqb / r, ca, sa, cb;
  if (sb > 1.0e-7) {
    ca = v[0] / sqb;
    sa = v[1] / sqb;
    cb = v[2] / r;
  } else {
    sa = 0.0;
    sb = 0.0;
    ca = cb = v[2] < 0.0 ? -1.0 : (v[2] > 0.0 ? 1.0 : 0.0);
  }
  for (int i = 0; i < 81; ++i) R[i] = 0.0;
  R[0] = 1.0;
  const double p[3][3] = {{ca * sb, sa * sb, cb}, {ca * cb, sa * cb, -sb}, {-sa, ca, 0.0}};
  for (int k = 0; k < 3; ++k)
    for (int K = 0; K < 3; ++K) R[(1 + k) * 9 + 1 + K] = p[k][K];
  if (with_d) {
    const double c2a = ca * ca * 2.0 - 1.0, c2b = cb * cb * 2.0 - 1.0, s2a = sa * ca * 2.0, s2b = sb * cb * 2.0;
    const double d[5][5] = {
        {c2a * sb * sb * pt5sq3, ca * s2b * pt5sq3, cb * cb - sb * sb * 0.5, sa * s2b * pt5sq3, s2a * sb * sb * pt5sq3},
        {c2a * s2b * 0.5, ca * c2b, -(s2b * pt5sq3), sa * c2b, s2a * s2b * 0.5},
        {-(s2a * sb), -(sa * cb), 0.0, ca * cb, c2a * sb},
        {c2a * (cb * cb + sb * sb * 0.5), ca * s2b * -0.5, sb * sb * pt5sq3, sa * s2b * -0.5,
         s2a * (cb * cb + sb * sb * 0.5)},
        {-(s2a * cb), sa * sb, 0.0, -(ca * sb), c2a * cb}};
    for (int k = 0; k < 5; ++k)
      for (int K = 0; K < 5; ++K) R[(4 + k) * 9 + 4 + K] = d[k][K];
  }
}

// Coefficient taking local distribution ab into molecular IJ (packed indices).
__device__ inline double dist_coef(const double* R, int IJ, int ab) {
  const int I = c_dt.pa[IJ], J = c_dt.pb[IJ], a = c_dt.pa[ab], b = c_dt.pb[ab];
  return a == b ? R[a * 9 + I] * R[a * 9 + J] : R[a * 9 + I] * R[b * 9 + J] + R[b * 9 + I] * R[a * 9 + J];
}

// ---- spd overlaps (src/overlap_d.rs) ----------------------------------------------------------
__device__ void b_integrals(double x, double bf[13]) {
  const double ax = fabs(x);
  if (ax > 3.0) {
    const double ex = exp(x), emx = 1.0 / ex, inv = 1.0 / x;
    bf[0] = (ex - emx) * inv;
    for (int i = 1; i <= 12; ++i) bf[i] = (bf[i - 1] * i + ex * (i % 2 == 0 ? 1.0 : -1.0) - emx) * inv;
  } else if (ax <= 1.0e-6) {
    for (int i = 0; i <= 12; ++i) bf[i] = (2 * ((i + 1) % 2)) / (i + 1.0);
  } else {
    const int last = ax > 2.0 ? 15 : (ax > 1.0 ? 12 : (ax > 0.5 ? 7 : 6));
    for (int i = 0; i <= 12; ++i) {
      double y = 0.0;
      for (int m = 0; m <= last; ++m) {
        const double xf = m != 0 ? c_dt.fact[m] : 1.0;
        const double coeff = (2 * ((m + i + 1) % 2)) / (xf * (m + i + 1));
        double xm = 1.0;
        for (int t = 0; t < m; ++t) xm *= -x;
        y += xm * coeff;
      }
      bf[i] = y;
    }
  }
}

// Overlap of Slater shells (na, la) and (nb, lb), azimuthal m, exponents ua ub, distance r, local
// frame (src/overlap_d.rs:84-176); l and m 0-based.
__device__ double shell_overlap(int na, int nb, int la, int lb, int m, double ua, double ub, double r) {
  auto aff = [](int l, int mm, int i) {
    if (i == 0) {
      if (l == 0 || (l == 1 && mm == 0)) return mm == 0 ? 1.0 : 0.0;
      if (l == 1) return sqrt(0.5);
      if (mm == 0) return 1.5;
      return mm == 1 ? sqrt(1.5) : sqrt(0.375);
    }
    return (l == 2 && mm == 0 && i == 2) ? -0.5 : 0.0;
  };
  auto bi = [](int i, int j) { return (j < 0 || j > i) ? 0.0 : c_dt.fact[i] / (c_dt.fact[j] * c_dt.fact[i - j]); };
  const double p = r * ((ua + ub) * 0.5), bb = r * ((ua - ub) * 0.5);
  double af[20];
  const double quo = 1.0 / p;
  af[0] = quo * exp(-p);
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
  const double norm = sqrt(ua * ub / (c_dt.fact[na + na] * c_dt.fact[nb + nb]) * ((la + la + 1) * (lb + lb + 1)));
  return sum * rpow * (pow(ua, (double)na) * pow(ub, (double)nb) / 2.0 * norm);
}

// Direction coefficients c[shell 1..3][component 1..5][local axis 1..5] of the overlap rotation
// (src/overlap_d.rs:178-238); axis 3 sigma, 2 and 4 pi, 1 and 5 delta.
__device__ void overlap_directions(const double d[3], int nij, double c[4][6][6]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 6; ++j)
      for (int k = 0; k < 6; ++k) c[i][j][k] = 0.0;
  const double xy2 = d[0] * d[0] + d[1] * d[1];
  const double r = sqrt(xy2 + d[2] * d[2]), xy = sqrt(xy2);
  double ca, cb, sa, sb;
  if (xy >= 1.0e-10) {
    ca = d[0] / xy;
    cb = d[2] / r;
    sa = d[1] / xy;
    sb = xy / r;
  } else if (d[2] < 0.0) {
    ca = cb = -1.0;
    sa = sb = 0.0;
  } else if (d[2] == 0.0) {
    ca = cb = sa = sb = 0.0;
  } else {
    ca = cb = 1.0;
    sa = sb = 0.0;
  }
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
}

// (shell 1..3, component 1..5) of each orbital s px py pz dx2-y2 dxz dz2 dyz dxy.
__device__ inline void orbital_shell(int o, int& sh, int& comp) {
  const int s[9] = {1, 2, 2, 2, 3, 3, 3, 3, 3};
  const int c[9] = {3, 4, 2, 3, 5, 4, 3, 2, 1};
  sh = s[o];
  comp = c[o];
}

// ---- the pair kernel ---------------------------------------------------------------------------
// Words stored per pair of a d-path molecule: the molecular block w (npf x nps), then the packed
// electron-core blocks of the first atom (npf) and of the second (nps).
__host__ __device__ inline int64_t dpair_words(int nf, int ns) {
  return (int64_t)npair_of(nf) * npair_of(ns) + npair_of(nf) + npair_of(ns);
}

// LDS working set of one MNDO/d pair evaluation by a 64-lane workgroup.
struct DPairShared {
  double L[45 * 45];
  double T[45 * 45];
  double R[81];
  double ri[23];
  double core[11][3];
  double S[4][4][4];
  double C[4][6][6];
  double ecb[90];   // electron-core blocks E = R^T E_local R: first atom's packed pairs, then the second's
  double hres[81];  // resonance block (beta_i + beta_j) S_ij / 2, nf x ns row-major
  double e_core;    // core-core energy
};

// Everything one MNDO/d pair contributes at separation d (second atom minus first, Bohr): the
// rotated two-electron block (s.L, npf x nps), the electron-core blocks, the resonance block and
// the core-core energy. All 64 lanes call it; it ends with the results visible to all of them.
__device__ void pair_d_eval(const ElemDev& ef, const ElemDev& es, int zf, int zs, const double d[3],
                            const double2* __restrict__ pairpar, double4 vpar, DPairShared& s) {
  const int lane = threadIdx.x;
  const int nf = ef.n_orb, ns = es.n_orb, npf = npair_of(nf), nps = npair_of(ns);
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const bool with_d = nf == 9 || ns == 9;
  double cf, pt;
  feather(r, cf, pt);
  // Stage 1: rotation, sp local integrals, electron-core terms, shell overlaps (lanes in parallel).
  if (lane == 0) orbital_rotation(d, r, with_d, s.R);
  if (lane == 1) local_sp_ri(ef, es, r, s.ri);
  if (lane == 2) {
    double core[11][3];
    for (int k = 0; k < 11; ++k) core[k][0] = core[k][1] = core[k][2] = 0.0;
    local_core_sp(ef, es, r, core);
    for (int k = 1; k <= 4; ++k) {
      s.core[k][1] = core[k][1];
      s.core[k][2] = core[k][2];
    }
  }
  if (lane >= 4 && lane < 16) {  // d electron-core terms: 6 distributions x 2 sides
    // (d-sigma s) (d-sigma p-sigma) (d-sigma d-sigma) (d-pi p-pi) (d-pi d-pi) (d-delta d-delta)
    const int dist[6] = {10, 11, 14, 17, 20, 35};
    const int k = (lane - 4) >> 1, side = 1 + ((lane - 4) & 1);
    double v = 0.0;
    if (side == 2 && ns == 9) v = -local_multipole(ef, es, 0, dist[k], 1, r) * (EV * ef.core_charge);
    if (side == 1 && nf == 9) v = -local_multipole(ef, es, dist[k], 0, 2, r) * (EV * es.core_charge);
    s.core[5 + k][side] = v;
  }
  if (lane == 20) overlap_directions(d, max(nf, ns), s.C);
  {
    // shell overlaps s[i][j][k]: lanes 32.. take (i, j, k) in 1..3
    const int t = lane - 32;
    if (t >= 0 && t < 27) {
      const int i = 1 + t / 9, j = 1 + (t / 3) % 3, k = 1 + t % 3;
      const int sa = nf == 1 ? 1 : (nf == 4 ? 2 : 3), sb = ns == 1 ? 1 : (ns == 4 ? 2 : 3);
      const int newk = min(sa - 1, sb - 1);
      double v = 0.0;
      if (i <= sa && j <= sb && k <= newk + 1 && k <= i && k <= j) {
        const double ei[3] = {ef.zeta_s, ef.zeta_p, fmax(ef.zeta_d, 0.3)};
        const double ej[3] = {es.zeta_s, es.zeta_p, fmax(es.zeta_d, 0.3)};
        const int pq1 = i == 3 ? ef.n_d : ef.n_sp, pq2 = j == 3 ? es.n_d : es.n_sp;
        v = shell_overlap(max(pq1, i), max(pq2, j), i - 1, j - 1, k - 1, ei[i - 1], ej[j - 1], r);
      }
      s.S[i][j][k] = v;
    }
  }
  __syncthreads();
  // Stage 2: feather the electron-core terms (src/mndod_twocenter.rs:1064-1081); local integrals.
  if (lane < 20) {
    const int k = 1 + lane / 2, side = 1 + (lane & 1);
    if (k <= 10) {
      double v = s.core[k][side] * cf;
      if (k == 1 || k == 3 || k == 4 || k == 7 || k == 9 || k == 10)
        v += (1.0 / r) * (-EV * (side == 1 ? es.core_charge : ef.core_charge)) * (1.0 - cf);
      s.core[k][side] = v;
    }
  }
  const double point_add = pt * (1.0 - cf);
  for (int idx = lane; idx < npf * nps; idx += 64) {
    const int ij = idx / nps, kl = idx - ij * nps;
    double v;
    if (ij < 10 && kl < 10) {
      const int k = c_dt.sp_local[ij][kl];
      v = k ? s.ri[k] : 0.0;
    } else {
      v = local_multipole(ef, es, ij, kl, 0, r) * EV * cf;
      if (c_dt.pa[ij] == c_dt.pb[ij] && c_dt.pa[kl] == c_dt.pb[kl]) v += point_add;
    }
    s.L[idx] = v;
  }
  __syncthreads();
  // Stage 3: rotate the second index, then the first (only same-class distributions mix).
  for (int idx = lane; idx < npf * nps; idx += 64) {
    const int ij = idx / nps, KL = idx - ij * nps;
    const int cl = c_dt.cls[KL];
    double acc = 0.0;
    for (int kl = 0; kl < nps; ++kl)
      if (c_dt.cls[kl] == cl) acc += s.L[ij * nps + kl] * dist_coef(s.R, KL, kl);
    s.T[idx] = acc;
  }
  __syncthreads();
  for (int idx = lane; idx < npf * nps; idx += 64) {
    const int IJ = idx / nps, KL = idx - IJ * nps;
    const int cl = c_dt.cls[IJ];
    double acc = 0.0;
    for (int ij = 0; ij < npf; ++ij)
      if (c_dt.cls[ij] == cl) acc += dist_coef(s.R, IJ, ij) * s.T[ij * nps + KL];
    s.L[idx] = acc;
  }
  __syncthreads();
  // Stage 4: PM7 d-imbalance correction (src/mndod_twocenter.rs:918-1000), one lane.
  if (lane == 0 && (ef.iod > 0 || es.iod > 0)) {
    auto W = [&](int ij, int kl) -> double& { return s.L[ij * nps + kl]; };
    const double w00 = s.L[0];
    const int Pp[3] = {2, 5, 9}, Dd[5] = {14, 20, 27, 35, 44};
    if (ef.iod > 0) {
      if (ns > 1) {
        double sum = 0.0;
        for (int a : Dd) sum += W(a, Pp[0]) + W(a, Pp[1]) + W(a, Pp[2]);
        sum = w00 - sum * (1.0 / 15.0);
        for (int a : Dd)
          for (int b : Pp) W(a, b) += sum;
        const double sh = s.core[1][1] - (s.core[3][1] + s.core[4][1] * 2.0) * (1.0 / 3.0);
        s.core[3][1] += sh;
        s.core[4][1] += sh;
      }
      double sum = 0.0;
      for (int a : Dd) sum += W(a, 0);
      sum = w00 - sum * (1.0 / 5.0);
      for (int a : Dd) W(a, 0) += sum;
      const double sh = s.core[1][1] - (s.core[7][1] + s.core[9][1] * 2.0 + s.core[10][1] * 2.0) * (1.0 / 5.0);
      s.core[7][1] += sh;
      s.core[9][1] += sh;
      s.core[10][1] += sh;
    }
    if (es.iod > 0) {
      if (ef.iod > 0) {
        double sum = 0.0;
        for (int a : Dd)
          for (int b : Dd) sum += W(a, b);
        sum = w00 - sum * (1.0 / 25.0);
        for (int a : Dd)
          for (int b : Dd) W(a, b) += sum;
      }
      if (nf > 1) {
        double sum = 0.0;
        for (int a : Pp)
          for (int b : Dd) sum += W(a, b);
        sum = w00 - sum * (1.0 / 15.0);
        for (int a : Pp)
          for (int b : Dd) W(a, b) += sum;
        const double sh = s.core[1][2] - (s.core[3][2] + s.core[4][2] * 2.0) * (1.0 / 3.0);
        s.core[3][2] += sh;
        s.core[4][2] += sh;
      }
      double sum = 0.0;
      for (int b : Dd) sum += W(0, b);
      sum = w00 - sum * (1.0 / 5.0);
      for (int b : Dd) W(0, b) += sum;
      const double sh = s.core[1][2] - (s.core[7][2] + s.core[9][2] * 2.0 + s.core[10][2] * 2.0) * (1.0 / 5.0);
      s.core[7][2] += sh;
      s.core[9][2] += sh;
      s.core[10][2] += sh;
    }
  }
  __syncthreads();
  // Stage 5: the electron-core blocks E = R^T E_local R (src/mndod_twocenter.rs:1003-1045), the
  // resonance blocks beta*S and the core-core energy.
  for (int t = lane; t < npf + nps; t += 64) {
    const int side = t < npf ? 1 : 2;
    const int p = side == 1 ? t : t - npf, n = side == 1 ? nf : ns;
    const int K = c_dt.pa[p], Mo = c_dt.pb[p];
    auto Lc = [&](int a, int b) -> double {  // local electron-core matrix
      if (a > b) {
        const int x = a;
        a = b;
        b = x;
      }
      if (a == 0 && b == 0) return s.core[1][side];
      if (a == 0 && b == 1) return s.core[2][side];
      if (a == 1 && b == 1) return s.core[3][side];
      if ((a == 2 && b == 2) || (a == 3 && b == 3)) return s.core[4][side];
      if (a == 0 && b == 4) return s.core[5][side];
      if (a == 1 && b == 4) return s.core[6][side];
      if (a == 4 && b == 4) return s.core[7][side];
      if ((a == 2 && b == 5) || (a == 3 && b == 6)) return s.core[8][side];
      if ((a == 5 && b == 5) || (a == 6 && b == 6)) return s.core[9][side];
      if ((a == 7 && b == 7) || (a == 8 && b == 8)) return s.core[10][side];
      return 0.0;
    };
    double acc = 0.0;
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        const double l = Lc(a, b);
        if (l != 0.0) acc += s.R[a * 9 + K] * l * s.R[b * 9 + Mo];
      }
    s.ecb[t] = acc;
  }
  for (int t = lane; t < nf * ns; t += 64) {
    const int oi = t / ns, oj = t - oi * ns;
    int i, k, j, l;
    orbital_shell(oi, i, k);
    orbital_shell(oj, j, l);
    const double aa = j == 2 ? -1.0 : 1.0, bb = j == 3 ? -1.0 : 1.0;
    const double sv = s.S[i][j][1] * (s.C[i][k][3] * s.C[j][l][3]) * aa +
                      s.S[i][j][2] * (s.C[i][k][4] * s.C[j][l][4] + s.C[i][k][2] * s.C[j][l][2]) * bb +
                      s.S[i][j][3] * (s.C[i][k][5] * s.C[j][l][5] + s.C[i][k][1] * s.C[j][l][1]);
    const double bi = oi == 0 ? ef.beta_s : (oi <= 3 ? ef.beta_p : ef.beta_d);
    const double bj = oj == 0 ? es.beta_s : (oj <= 3 ? es.beta_p : es.beta_d);
    s.hres[t] = 0.5 * (bi + bj) * sv;
  }
  if (lane == 0) s.e_core = pair_core(ef, es, zf, zs, r, pairpar, vpar);
  __syncthreads();
}

__global__ void __launch_bounds__(64) k_pairs_d(const int64_t* __restrict__ dpairs, const PairDev* __restrict__ pairs,
                                                const MolDev* __restrict__ mols, const int* __restrict__ atom_z,
                                                const int* __restrict__ atom_ao, const double* __restrict__ xyz,
                                                const ElemDev* __restrict__ elem, const double2* __restrict__ pairpar,
                                                double4 vpar, double* __restrict__ w, double* __restrict__ slab,
                                                double* __restrict__ e_core) {
  __shared__ DPairShared s;
  const int lane = threadIdx.x;
  const int64_t g = dpairs[blockIdx.x];
  const PairDev P = pairs[g];
  const MolDev& M = mols[P.mol];
  const int zf = atom_z[P.f], zs = atom_z[P.s];
  const ElemDev ef = elem[zf];
  const ElemDev es = elem[zs];
  const int nf = ef.n_orb, ns = es.n_orb, npf = npair_of(nf), nps = npair_of(ns);
  const double d[3] = {xyz[3 * P.s] - xyz[3 * P.f], xyz[3 * P.s + 1] - xyz[3 * P.f + 1],
                       xyz[3 * P.s + 2] - xyz[3 * P.f + 2]};
  pair_d_eval(ef, es, zf, zs, d, pairpar, vpar, s);
  double* wp = w + P.w_off;
  for (int idx = lane; idx < npf * nps; idx += 64) wp[idx] = s.L[idx];
  for (int t = lane; t < npf + nps; t += 64) wp[(int64_t)npf * nps + t] = s.ecb[t];
  const int ld = M.ld;
  double* H = slab + M.mat0 + (int64_t)S_H * ld * ld;
  const int of = atom_ao[P.f], os = atom_ao[P.s];
  for (int t = lane; t < nf * ns; t += 64) {
    const int oi = t / ns, oj = t - oi * ns;
    H[(int64_t)(of + oi) * ld + os + oj] = s.hres[t];
    H[(int64_t)(os + oj) * ld + of + oi] = s.hres[t];
  }
  if (lane == 0) e_core[g] = s.e_core;
}

}  // namespace pm7

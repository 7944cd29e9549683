// SPDX-License-Identifier: GPL-3.0-or-later
//
// Once-per-geometry kernels: the sp two-centre integrals, overlaps, resonance and core-core
// terms of every atom pair (pm7-rs src/integrals.rs:16-538, src/overlap.rs:16-462,
// src/hamiltonian.rs:310-582, src/repulsion.rs:164-263), and the owner-computes assembly of the
// one-centre diagonal blocks of H_core.
//
// k_pairs       one thread per atom pair (every pair of every molecule in the batch). It writes
//               the rotated w block into the pair's slot of the integral stream, the two
//               off-diagonal resonance blocks of H (sole owner of those elements, so no atomics),
//               and the pair's core-core energy into a per-pair slot that a later fixed-order
//               reduction sums.
// k_hcore_diag  one thread per (atom, packed orbital pair): U + sum over partner atoms of the
//               electron-core attraction, read back out of the w blocks (e1b = -Z_b (mu nu|s s)).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "pm7_device.h"

namespace pm7 {

__device__ inline void feather(double r, double& cfrac, double& point) {
  point = EV / r;
  double ra = r * A0;
  cfrac = ra < 7.0 ? 1.0 - exp(-(ra - 7.0) * (ra - 7.0) * 0.22) : 0.0;
}

__device__ inline void rot_to_x(double vx, double vy, double vz, double R[3][3]) {
  double qy = vz, qz = -vy, qw = vx + 1.0;
  if (fabs(qw) < 1.0e-13) {
    qy = 0.0;
    qz = 1.0;
    qw = 0.0;
  }
  double inv = 1.0 / sqrt(qy * qy + qz * qz + qw * qw);
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

__device__ inline double rs(double a, double c) { return 1.0 / sqrt(a * a + c); }
__device__ inline double rq(double x, double c) { return 1.0 / sqrt(x + c); }
__device__ inline double sq(double x) { return x * x; }

// Local-frame X-H integrals (4) with feathering.
__device__ void local_xh(const ElemDev& ei, const ElemDev& ej, double r, double ri[4]) {
  const double ev1 = EV / 2.0, ev2 = EV / 4.0;
  double da = ei.dd, qa = ei.qq * 2.0;
  double aee = sq(ei.rho0 + ej.rho0), ade = sq(ei.rho1 + ej.rho0), aqe = sq(ei.rho2 + ej.rho0);
  double ee = EV / sqrt(r * r + aee);
  double e6 = ev1 / sqrt(r * r + aqe);
  ri[0] = ee;
  ri[1] = ev1 * rs(r + da, ade) - ev1 * rs(r - da, ade);
  ri[2] = ee + ev2 * rs(r + qa, aqe) + ev2 * rs(r - qa, aqe) - e6;
  ri[3] = ee + ev1 / sqrt(r * r + qa * qa + aqe) - e6;
  double cf, pt;
  feather(r, cf, pt);
  double add = pt * (1.0 - cf);
  for (int k = 0; k < 4; ++k) ri[k] *= cf;
  ri[0] += add;
  ri[2] += add;
  ri[3] += add;
}

// Local-frame X-X integrals (22) with feathering.
__device__ void local_xx(const ElemDev& ei, const ElemDev& ej, double r, double ri[22]) {
  const double ev1 = EV / 2.0, ev2 = EV / 4.0, ev3 = EV / 8.0, ev4 = EV / 16.0;
  double da = ei.dd, db = ej.dd, qa = ei.qq * 2.0, qb = ej.qq * 2.0, qa1 = ei.qq, qb1 = ej.qq;
  double aee = sq(ei.rho0 + ej.rho0), ade = sq(ei.rho1 + ej.rho0), aqe = sq(ei.rho2 + ej.rho0);
  double aed = sq(ei.rho0 + ej.rho1), aeq = sq(ei.rho0 + ej.rho2), axx = sq(ei.rho1 + ej.rho1);
  double adq = sq(ei.rho1 + ej.rho2), aqd = sq(ei.rho2 + ej.rho1), aqq = sq(ei.rho2 + ej.rho2);
  double r2 = r * r;
  double ee = EV / sqrt(r2 + aee);
  double dze = ev1 * rs(r - da, ade) - ev1 * rs(r + da, ade);
  double e6 = ev1 / sqrt(r2 + aqe);
  double qzze = ev2 * rs(r - qa, aqe) + ev2 * rs(r + qa, aqe) - e6;
  double qxxe = ev1 * rq(r2 + qa * qa, aqe) - e6;
  double edz = ev1 * rs(r + db, aed) - ev1 * rs(r - db, aed);
  double e12 = ev1 / sqrt(r2 + aeq);
  double eqzz = ev2 * rs(r - qb, aeq) + ev2 * rs(r + qb, aeq) - e12;
  double eqxx = ev1 * rq(r2 + qb * qb, aeq) - e12;
  double e20 = ev2 * rs(r + da, adq), e22 = ev2 * rs(r - da, adq);
  double e24 = ev2 * rs(r - db, aqd), e26 = ev2 * rs(r + db, aqd);
  double e36 = ev2 / sqrt(r2 + aqq);
  double e39 = ev2 * rq(r2 + qa * qa, aqq), e40 = ev2 * rq(r2 + qb * qb, aqq);
  double e42 = ev3 * rs(r - qb, aqq), e44 = ev3 * rs(r + qb, aqq);
  double e46 = ev3 * rs(r + qa, aqq), e48 = ev3 * rs(r - qa, aqq);
  ri[0] = ee;
  ri[1] = -dze;
  ri[2] = ee + qzze;
  ri[3] = ee + qxxe;
  ri[4] = -edz;
  ri[5] = ev2 * rs(r + da - db, axx) + ev2 * rs(r - da + db, axx) - ev2 * rs(r - da - db, axx) -
          ev2 * rs(r + da + db, axx);
  ri[6] = ev1 * rq(r2 + sq(da - db), axx) - ev1 * rq(r2 + sq(da + db), axx);
  ri[7] = -edz + ev3 * rs(r + qa - db, aqd) - ev3 * rs(r + qa + db, aqd) + ev3 * rs(r - qa - db, aqd) -
          ev3 * rs(r - qa + db, aqd) - e24 + e26;
  ri[8] = -edz - e24 + ev2 * rq(sq(r - db) + qa * qa, aqd) + e26 - ev2 * rq(sq(r + db) + qa * qa, aqd);
  ri[9] = ev2 * rq(sq(qa1 - db) + sq(r + qa1), aqd) - ev2 * rq(sq(qa1 - db) + sq(r - qa1), aqd) -
          ev2 * rq(sq(qa1 + db) + sq(r + qa1), aqd) + ev2 * rq(sq(qa1 + db) + sq(r - qa1), aqd);
  ri[10] = ee + eqzz;
  ri[11] = ee + eqxx;
  ri[12] = -dze + ev3 * rs(r + da - qb, adq) - ev3 * rs(r - da - qb, adq) + ev3 * rs(r + da + qb, adq) -
           ev3 * rs(r - da + qb, adq) + e22 - e20;
  ri[13] = -dze - e20 + ev2 * rq(sq(r + da) + qb * qb, adq) + e22 - ev2 * rq(sq(r - da) + qb * qb, adq);
  ri[14] = ev2 * rq(sq(da - qb1) + sq(r - qb1), adq) - ev2 * rq(sq(da - qb1) + sq(r + qb1), adq) -
           ev2 * rq(sq(da + qb1) + sq(r - qb1), adq) + ev2 * rq(sq(da + qb1) + sq(r + qb1), adq);
  ri[15] = ee + eqzz + qzze + ev4 * rs(r + qa - qb, aqq) + ev4 * rs(r + qa + qb, aqq) +
           ev4 * rs(r - qa - qb, aqq) + ev4 * rs(r - qa + qb, aqq) - e48 - e46 - e42 - e44 + e36;
  ri[16] = ee + eqzz + qxxe + ev3 * rq(sq(r - qb) + qa * qa, aqq) + ev3 * rq(sq(r + qb) + qa * qa, aqq) -
           e42 - e44 - e39 + e36;
  ri[17] = ee + eqxx + qzze + ev3 * rq(sq(r + qa) + qb * qb, aqq) + ev3 * rq(sq(r - qa) + qb * qb, aqq) -
           e46 - e48 - e40 + e36;
  double qxxqxx = ev3 * rq(r2 + sq(qa - qb), aqq) + ev3 * rq(r2 + sq(qa + qb), aqq) - e39 - e40 + e36;
  ri[18] = ee + eqxx + qxxe + qxxqxx;
  double m2 = sq(qa1 - qb1), p2 = sq(qa1 + qb1);
  ri[19] = ev3 * rq(sq(r + qa1 - qb1) + m2, aqq) - ev3 * rq(sq(r + qa1 + qb1) + m2, aqq) -
           ev3 * rq(sq(r - qa1 - qb1) + m2, aqq) + ev3 * rq(sq(r - qa1 + qb1) + m2, aqq) -
           ev3 * rq(sq(r + qa1 - qb1) + p2, aqq) + ev3 * rq(sq(r + qa1 + qb1) + p2, aqq) +
           ev3 * rq(sq(r - qa1 - qb1) + p2, aqq) - ev3 * rq(sq(r - qa1 + qb1) + p2, aqq);
  double qxxqyy = ev2 * rq(r2 + qa * qa + qb * qb, aqq) - e39 - e40 + e36;
  ri[20] = ee + eqxx + qxxe + qxxqyy;
  ri[21] = (qxxqxx - qxxqyy) * 0.5;
  double cf, pt;
  feather(r, cf, pt);
  double add = pt * (1.0 - cf);
  for (int k = 0; k < 22; ++k) ri[k] *= cf;
  ri[0] += add;
  ri[2] += add;
  ri[3] += add;
  ri[10] += add;
  ri[11] += add;
  ri[15] += add;
  ri[16] += add;
  ri[17] += add;
  ri[18] += add;
  ri[20] += add;
}

// Rotate the 22 local X-X integrals into the molecular frame, streaming the 100 outputs.
__device__ void rotate_xx_store(const double ri[22], const double R[3][3], double* __restrict__ w) {
  const double* r0 = R[0];
  const double* r1 = R[1];
  const double* r2 = R[2];
  int idx = 0;
#pragma unroll 1
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll 1
    for (int ll = 0; ll <= kk; ++ll)
      for (int mm = 0; mm < 4; ++mm)
        for (int nn = 0; nn <= mm; ++nn) {
          int k = kk - 1, l = ll - 1, m = mm - 1, n = nn - 1;
          double v;
          if (kk == 0) {
            if (mm == 0)
              v = ri[0];
            else if (nn == 0)
              v = ri[4] * r0[m];
            else
              v = ri[10] * (r0[m] * r0[n]) + ri[11] * (r1[m] * r1[n] + r2[m] * r2[n]);
          } else if (ll == 0) {
            if (mm == 0)
              v = ri[1] * r0[k];
            else if (nn == 0)
              v = ri[5] * (r0[k] * r0[m]) + ri[6] * (r1[k] * r1[m] + r2[k] * r2[m]);
            else {
              double t0 = r0[k] * r0[m] * r0[n];
              double t1 = (r1[m] * r1[n] + r2[m] * r2[n]) * r0[k];
              double mix = r1[k] * (r1[n] * r0[m] + r1[m] * r0[n]) + r2[k] * (r2[m] * r0[n] + r2[n] * r0[m]);
              v = ri[12] * t0 + ri[13] * t1 + ri[14] * mix;
            }
          } else if (mm == 0) {
            v = ri[2] * (r0[k] * r0[l]) + ri[3] * (r1[k] * r1[l] + r2[k] * r2[l]);
          } else if (nn == 0) {
            double t0 = r0[k] * r0[l] * r0[m];
            double t1 = (r1[k] * r1[l] + r2[k] * r2[l]) * r0[m];
            double t2 = r1[l] * r1[m] + r2[l] * r2[m];
            v = ri[7] * t0 + ri[8] * t1 + ri[9] * (r0[k] * t2 + r0[l] * (r1[k] * r1[m] + r2[k] * r2[m]));
          } else {
            double acc = ri[15] * (r0[k] * r0[l] * r0[m] * r0[n]);
            acc += ri[16] * ((r1[k] * r1[l] + r2[k] * r2[l]) * r0[m] * r0[n]);
            acc += ri[17] * ((r1[m] * r1[n] + r2[m] * r2[n]) * (r0[k] * r0[l]));
            acc += ri[18] * (r1[k] * r1[l] * r1[m] * r1[n] + r2[k] * r2[l] * r2[m] * r2[n]);
            double mix1 = r0[m] * (r1[l] * r1[n] + r2[l] * r2[n]);
            double mix2 = r0[n] * (r1[l] * r1[m] + r2[l] * r2[m]);
            double v5 = r0[k] * (mix1 + mix2) +
                        r0[l] * (r0[m] * (r1[k] * r1[n] + r2[k] * r2[n]) + r0[n] * (r1[k] * r1[m] + r2[k] * r2[m]));
            acc += ri[19] * v5;
            acc += ri[20] * (r1[k] * r1[l] * r2[m] * r2[n] + r2[k] * r2[l] * r1[m] * r1[n]);
            acc += ri[21] * ((r1[k] * r2[l] + r2[k] * r1[l]) * (r1[m] * r2[n] + r2[m] * r1[n]));
            v = acc;
          }
          w[idx++] = v;
        }
}

// Auxiliary A_k(x) and B_k(x) integrals, k = 0..6 (all the n <= 3 Slater overlaps need).
__device__ void aint7(double x, double a[7]) {
  for (int i = 0; i < 7; ++i) a[i] = 0.0;
  if (fabs(x) < 1.0e-12) return;
  double inv = 1.0 / x;
  a[0] = exp(-x) * inv;
  for (int i = 1; i < 7; ++i) a[i] = a[0] + a[i - 1] * inv * i;
}
__device__ void bint7(double x, double b[7]) {
  double ax = fabs(x);
  if (ax > 0.5) {
    double inv = 1.0 / x;
    double tx = exp(x) * inv, tmx = -exp(-x) * inv;
    b[0] = tx + tmx;
    for (int i = 1; i < 7; ++i) b[i] = ((i & 1) ? -tx : tx) + tmx + b[i - 1] * inv * i;
  } else if (ax > 1.0e-6) {
    double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x4 * x, x6 = x4 * x2;
    b[0] = x2 * (1.0 / 3.0) + x4 * (1.0 / 60.0) + x6 * (1.0 / 2520.0) + 2.0;
    b[2] = x2 * (1.0 / 5.0) + x4 * (1.0 / 84.0) + x6 * (1.0 / 3240.0) + 2.0 / 3.0;
    b[4] = x2 * (1.0 / 7.0) + x4 * (1.0 / 108.0) + x6 * (1.0 / 3960.0) + 2.0 / 5.0;
    b[6] = x2 * (1.0 / 9.0) + x4 * (1.0 / 132.0) + x6 * (1.0 / 4680.0) + 2.0 / 7.0;
    b[1] = x * (-2.0 / 3.0) - x3 * (1.0 / 15.0) - x5 * (1.0 / 420.0);
    b[3] = x * (-2.0 / 5.0) - x3 * (1.0 / 21.0) - x5 * (1.0 / 540.0);
    b[5] = x * (-2.0 / 7.0) - x3 * (1.0 / 27.0) - x5 * (1.0 / 660.0);
  } else {
    for (int i = 0; i < 7; ++i) b[i] = (i & 1) ? 0.0 : 2.0 / (i + 1.0);
  }
}

// sigma/pi local overlaps (s111, s211, s121, s221, s222) for ea.n >= eb.n, n <= 3.
__device__ void overlap_locals(const ElemDev& ea, const ElemDev& eb, double r, double s[5]) {
  double za0 = ea.zeta_s, za1 = ea.zeta_p, zb0 = eb.zeta_s, zb1 = eb.zeta_p;
  double A1[7], B1[7], A2[7], B2[7], A3[7], B3[7], A4[7], B4[7];
  aint7(r * (0.5 * (za0 + zb0)), A1);
  bint7(r * (0.5 * (za0 - zb0)), B1);
  aint7(r * (0.5 * (za1 + zb0)), A2);
  bint7(r * (0.5 * (za1 - zb0)), B2);
  aint7(r * (0.5 * (za0 + zb1)), A3);
  bint7(r * (0.5 * (za0 - zb1)), B3);
  aint7(r * (0.5 * (za1 + zb1)), A4);
  bint7(r * (0.5 * (za1 - zb1)), B4);
  const double sq3 = 1.7320508075688772;
  double r2 = r * r, r4 = r2 * r2, r5 = r4 * r, r6 = r5 * r, r7 = r6 * r;
  for (int k = 0; k < 5; ++k) s[k] = 0.0;
  int na = ea.n, nb = eb.n;
  if (na == 1 && nb == 1) {
    s[0] = pow(r2 * (za0 * zb0), 1.5) * (A1[2] * B1[0] - B1[2] * A1[0]) / 4.0;
  } else if (na == 2 && nb == 1) {
    s[0] = r4 * (pow(zb0, 1.5) * pow(za0, 2.5)) * (A1[3] * B1[0] - B1[3] * A1[0] + A1[2] * B1[1] - B1[2] * A1[1]) /
           (sq3 * 8.0);
    s[1] = r4 * (pow(zb0, 1.5) * pow(za1, 2.5)) * (A2[2] * B2[0] - B2[2] * A2[0] + A2[3] * B2[1] - B2[3] * A2[1]) / 8.0;
  } else if (na == 2 && nb == 2) {
    s[0] = r5 * pow(zb0 * za0, 2.5) * (A1[4] * B1[0] + B1[4] * A1[0] - A1[2] * B1[2] * 2.0) / 48.0;
    s[1] = r5 * pow(zb0 * za1, 2.5) *
           (A2[3] * (B2[0] - B2[2]) - A2[1] * (B2[2] - B2[4]) + B2[3] * (A2[0] - A2[2]) - B2[1] * (A2[2] - A2[4])) /
           (16.0 * sq3);
    s[2] = r5 * pow(zb1 * za0, 2.5) *
           (A3[3] * (B3[0] - B3[2]) - A3[1] * (B3[2] - B3[4]) - B3[3] * (A3[0] - A3[2]) + B3[1] * (A3[2] - A3[4])) /
           (16.0 * sq3);
    double w = r5 * pow(zb1 * za1, 2.5) / 16.0;
    s[3] = -(w * (B4[2] * (A4[4] + A4[0]) - A4[2] * (B4[4] + B4[0])));
    s[4] = w * (A4[4] * (B4[0] - B4[2]) - B4[4] * (A4[0] - A4[2]) - A4[2] * B4[0] + B4[2] * A4[0]) * 0.5;
  } else if (na == 3 && nb == 1) {
    s[0] = r5 * (pow(zb0, 1.5) * pow(za0, 3.5)) *
           (A1[4] * B1[0] + B1[1] * A1[3] * 2.0 - A1[1] * B1[3] * 2.0 - B1[4] * A1[0]) / (sqrt(10.0) * 24.0);
    s[1] = r5 * (pow(zb0, 1.5) * pow(za1, 3.5)) *
           (A2[3] * (B2[0] + B2[2]) - A2[1] * (B2[4] + B2[2]) + B2[1] * (A2[2] + A2[4]) - B2[3] * (A2[2] + A2[0])) /
           (8.0 * sqrt(30.0));
  } else if (na == 3 && nb == 2) {
    s[0] = r6 * (pow(zb0, 2.5) * pow(za0, 3.5)) *
           (A1[5] * B1[0] + B1[1] * A1[4] - B1[2] * A1[3] * 2.0 - A1[2] * B1[3] * 2.0 + B1[4] * A1[1] + B1[5] * A1[0]) /
           (sqrt(30.0) * 48.0);
    s[1] = r6 * (pow(zb0, 2.5) * pow(za1, 3.5)) *
           (A2[4] * B2[0] + B2[1] * A2[5] - B2[3] * A2[3] * 2.0 - A2[2] * B2[2] * 2.0 + A2[1] * B2[5] + A2[0] * B2[4]) /
           (48.0 * sqrt(10.0));
    s[2] = r6 * (pow(zb1, 2.5) * pow(za0, 3.5)) *
           ((A3[4] * B3[0] - A3[5] * B3[1]) + (A3[3] * B3[1] - A3[4] * B3[2]) * 2.0 -
            (A3[1] * B3[3] - A3[2] * B3[4]) * 2.0 - (A3[0] * B3[4] - A3[1] * B3[5])) /
           (48.0 * sqrt(10.0));
    s[3] = r6 * (pow(zb1, 2.5) * pow(za1, 3.5)) *
           ((A4[3] * B4[0] - A4[5] * B4[2]) + (A4[2] * B4[1] - A4[4] * B4[3]) - (A4[1] * B4[2] - A4[3] * B4[4]) -
            (A4[0] * B4[3] - A4[2] * B4[5])) /
           (16.0 * sqrt(30.0));
    s[4] = r6 * (pow(zb1, 2.5) * pow(za1, 3.5)) *
           ((A4[5] - A4[3]) * (B4[0] - B4[2]) + (A4[4] - A4[2]) * (B4[1] - B4[3]) - (A4[3] - A4[1]) * (B4[2] - B4[4]) -
            (A4[2] - A4[0]) * (B4[3] - B4[5])) /
           (32.0 * sqrt(30.0));
  } else if (na == 3 && nb == 3) {
    s[0] = r7 * pow(zb0 * za0, 3.5) * (A1[6] * B1[0] - B1[2] * A1[4] * 3.0 + A1[2] * B1[4] * 3.0 - A1[0] * B1[6]) /
           1440.0;
    s[1] = r7 * pow(zb0 * za1, 3.5) *
           ((A2[5] * B2[0] + A2[6] * B2[1]) + (-A2[4] * B2[1] - A2[5] * B2[2]) - (A2[3] * B2[2] + A2[4] * B2[3]) * 2.0 -
            (-A2[2] * B2[3] - A2[3] * B2[4]) * 2.0 + (A2[1] * B2[4] + A2[2] * B2[5]) + (-A2[0] * B2[5] - A2[1] * B2[6])) /
           (480.0 * sq3);
    s[2] = r7 * pow(zb1 * za0, 3.5) *
           ((A3[5] * B3[0] - A3[6] * B3[1]) + (A3[4] * B3[1] - A3[5] * B3[2]) - (A3[3] * B3[2] - A3[4] * B3[3]) * 2.0 -
            (A3[2] * B3[3] - A3[3] * B3[4]) * 2.0 + (A3[1] * B3[4] - A3[2] * B3[5]) + (A3[0] * B3[5] - A3[1] * B3[6])) /
           (480.0 * sq3);
    s[3] = r7 * (pow(zb1, 3.5) * pow(za1, 3.5)) *
           ((A4[4] * B4[0] - A4[6] * B4[2]) - (A4[2] * B4[2] - A4[4] * B4[4]) * 2.0 + (A4[0] * B4[4] - A4[2] * B4[6])) /
           480.0;
    s[4] = r7 * (pow(zb1, 3.5) * pow(za1, 3.5)) *
           ((A4[6] - A4[4]) * (B4[0] - B4[2]) - (A4[4] - A4[2]) * (B4[2] - B4[4]) * 2.0 + (A4[2] - A4[0]) * (B4[4] - B4[6])) /
           960.0;
  }
}

// Numerical local overlaps for pairs with a valence shell n >= 4 (pm7-rs src/overlap_numeric.rs:24-115):
// the five sigma/pi overlaps of Slater shells by quadrature in prolate spheroidal coordinates,
// xi in [1, 1 + 60/p] (48 Gauss-Legendre nodes, p = (za + zb) r / 2), eta in [-1, 1] (40 nodes),
// dV = (r/2)^3 (xi^2 - eta^2). One wave per pair: lanes split the 1920 nodes of each integral,
// then a fixed shuffle tree sums them. Atom a is the one with the higher n.
// With ndisp = GRAD_DISP (gradient calls) item = pair * GRAD_DISP + axis * 4 + step evaluates the
// pair at the displaced separation d + GRAD_STEPS[step] along that axis, as k_pair_grad needs.
constexpr int GRAD_DISP = 12;
constexpr double GRAD_H = 1.0e-3;
__device__ inline double grad_step(int t) { return (t & 1 ? -1.0 : 1.0) * (t < 2 ? GRAD_H : 2.0 * GRAD_H); }

__global__ void __launch_bounds__(64) k_overlap_numeric(int n_num, const int64_t* __restrict__ num_pairs,
                                                         const PairDev* __restrict__ pairs,
                                                         const int* __restrict__ atom_z, const double* __restrict__ xyz,
                                                         const ElemDev* __restrict__ elem,
                                                         const double* __restrict__ gl, double* __restrict__ s_num,
                                                         int ndisp) {
  const int item = blockIdx.x;
  if (item >= n_num * (ndisp > 0 ? ndisp : 1)) return;
  const PairDev P = pairs[num_pairs[ndisp > 0 ? item / ndisp : item]];
  const ElemDev ef = elem[atom_z[P.f]], es = elem[atom_z[P.s]];
  const bool swap = ef.n < es.n;
  const ElemDev& ea = swap ? es : ef;
  const ElemDev& eb = swap ? ef : es;
  double d[3] = {xyz[3 * P.s] - xyz[3 * P.f], xyz[3 * P.s + 1] - xyz[3 * P.f + 1], xyz[3 * P.s + 2] - xyz[3 * P.f + 2]};
  if (ndisp > 0) {
    const int t = item % ndisp;
    d[t / 4] += grad_step(t % 4);
  }
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double hr = 0.5 * r;
  const int lane = threadIdx.x;
  // (zeta_a, zeta_b, l_a, l_b, m, sign): ss, p_sigma s, s p_sigma, p_sigma p_sigma, p_pi p_pi.
  const double zas[5] = {ea.zeta_s, ea.zeta_p, ea.zeta_s, ea.zeta_p, ea.zeta_p};
  const double zbs[5] = {eb.zeta_s, eb.zeta_s, eb.zeta_p, eb.zeta_p, eb.zeta_p};
  const int las[5] = {0, 1, 0, 1, 1}, lbs[5] = {0, 0, 1, 1, 1}, ms[5] = {0, 0, 0, 0, 1};
  const double sgn[5] = {1.0, 1.0, -1.0, -1.0, 1.0};
  const double cst[5] = {0.5, 0.86602540378443864676, 0.86602540378443864676, 1.5, 0.75};
  for (int q = 0; q < 5; ++q) {
    const double za = zas[q], zb = zbs[q];
    const int la = las[q], lb = lbs[q], m = ms[q];
    const double p = (za + zb) * hr;
    const double xi_max = 1.0 + 60.0 / fmax(p, 0.05);
    const double xm = 0.5 * (1.0 + xi_max), xh = 0.5 * (xi_max - 1.0);
    double acc = 0.0;
    for (int t = lane; t < 48 * 40; t += 64) {
      const int i = t / 40, j = t - 40 * i;
      const double xi = xm + xh * gl[i], wx = xh * gl[48 + i];
      const double eta = gl[96 + j], we = gl[136 + j];
      const double ra = hr * (xi + eta), rb = hr * (xi - eta);
      if (ra <= 0.0 || rb <= 0.0) continue;
      const double ca = (1.0 + xi * eta) / (xi + eta), cb = (xi * eta - 1.0) / (xi - eta);
      const double ang_a = la == 0 ? 1.0 : (m == 0 ? ca : sqrt(fmax(1.0 - ca * ca, 0.0)));
      const double ang_b = lb == 0 ? 1.0 : (m == 0 ? cb : sqrt(fmax(1.0 - cb * cb, 0.0)));
      double radial = exp(-za * ra - zb * rb);
      for (int k = 1; k < ea.n; ++k) radial *= ra;
      for (int k = 1; k < eb.n; ++k) radial *= rb;
      if (ea.n < 1) radial /= ra;  // n = 0 shells (a few sparkle-like entries): r^(n-1) = 1/r
      if (eb.n < 1) radial /= rb;
      acc += radial * (wx * we * ang_a * ang_b * (xi * xi - eta * eta));
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) {
      double fa = 1.0, fb = 1.0;
      for (int k = 2; k <= 2 * ea.n; ++k) fa *= k;
      for (int k = 2; k <= 2 * eb.n; ++k) fb *= k;
      const double norm = pow(2.0 * za, ea.n + 0.5) / sqrt(fa) * pow(2.0 * zb, eb.n + 0.5) / sqrt(fb);
      s_num[5 * (int64_t)item + q] = sgn[q] * hr * hr * hr * acc * norm * cst[q];
    }
  }
}

// Molecular-frame 4x4 overlap block S[i][j] (atom f rows, atom s cols); d = R_s - R_f. snum: the
// pair's numerical local overlaps (valence n >= 4), else null.
__device__ void overlap_block(const ElemDev& ef, const ElemDev& es, const double d[3], double r, double S[4][4],
                              const double* snum) {
  double sl[5];
  double dir[3];
  bool swap = ef.n < es.n;
  if (snum) {
    for (int k = 0; k < 5; ++k) sl[k] = snum[k];
  }
  if (!swap) {
    if (!snum) overlap_locals(ef, es, r, sl);
    dir[0] = d[0] / r;
    dir[1] = d[1] / r;
    dir[2] = d[2] / r;
  } else {
    if (!snum) overlap_locals(es, ef, r, sl);
    dir[0] = -d[0] / r;
    dir[1] = -d[1] / r;
    dir[2] = -d[2] / r;
  }
  double rr[3][3];
  rot_to_x(dir[0], dir[1], dir[2], rr);
  double di[4][4];
  di[0][0] = sl[0];
  for (int a = 0; a < 3; ++a) {
    di[1 + a][0] = sl[1] * rr[0][a];
    di[0][1 + a] = -sl[2] * rr[0][a];
  }
  double mm[3] = {-sl[3], sl[4], sl[4]};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double v = 0.0;
      for (int k = 0; k < 3; ++k) v += rr[k][a] * mm[k] * rr[k][b];
      di[1 + a][1 + b] = v;
    }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) S[a][b] = swap ? di[b][a] : di[a][b];
}

__device__ double gauss_slot(const ElemDev& e, double zz, double ra, int slot) {
  if (slot >= e.n_gauss) return 0.0;
  double ax = (ra - e.gm[slot]) * (ra - e.gm[slot]) * e.gl[slot];
  if (ax >= 25.0) return 0.0;
  return exp(-ax) * (zz * e.gk[slot]) / ra;
}

// PM7 core-core repulsion of one pair (MOPAC ccrep), FP64.
__device__ double pair_core(const ElemDev& ei, const ElemDev& ej, int zi, int zj, double r,
                            const double2* __restrict__ pairpar, double4 vpar) {
  double ra = r * A0;
  double zz = ei.core_charge * ej.core_charge;
  double po = ei.po9 + ej.po9;
  double gab_nddo = EV / sqrt(r * r + po * po);
  double cf, pt;
  feather(r, cf, pt);
  double enuc = (gab_nddo * cf + pt * (1.0 - cf)) * zz;
  double2 pp = pairpar[zi * MAX_Z + zj];
  bool defined = fabs(pp.y) > 1.0e-5;
  if (defined) {
    double alpb = pp.x < 1.0e-6 ? 1.2 : pp.x;
    double scale = 1.0 + exp(-(ra + pow(ra, 6.0) * 0.0003) * alpb) * (2.0 * pp.y);
    int hi = zi > zj ? zi : zj, lo = zi > zj ? zj : zi;
    if (lo == 1 && (hi == 6 || hi == 7)) {
      scale = 1.0 + exp(-(ra * ra) * alpb) * (2.0 * pp.y);
    } else if (lo == 1 && hi == 8) {
      scale = 1.0 + exp(-(ra * ra) * alpb) * (2.0 * pp.y) - exp(ra * (-2.0 * vpar.w)) * vpar.z;
    } else if (lo == 6 && hi == 6) {
      scale = scale + exp(ra * (-vpar.y)) * vpar.x;
    } else if (lo == 8 && hi == 14) {
      scale = scale - exp(-(ra - 2.9) * (ra - 2.9)) * 0.0007;
    }
    enuc *= scale;
  } else {
    bool sparkle = (zi >= 57 && zi <= 71) || (zj >= 57 && zj <= 71);
    enuc *= 1.0 + exp(ra * (sparkle ? -3.0 : -2.18)) * 10.0;
  }
  double corr = gauss_slot(ei, zz, ra, 0) + gauss_slot(ej, zz, ra, 0);
  if (!defined)
    for (int k = 0; k < 4; ++k) corr += gauss_slot(ei, zz, ra, k) + gauss_slot(ej, zz, ra, k);
  double red = ra / (pow((double)zi, 0.3333) + pow((double)zj, 0.3333));
  if (red < 3.0) {
    double g = 1.0e-8 / pow(red, 12.0);
    corr += g < 1.0e5 ? g : 1.0e5;
  }
  return enuc + corr;
}

__global__ void __launch_bounds__(128) k_pairs(int64_t n_pairs, const PairDev* __restrict__ pairs,
                                                const MolDev* __restrict__ mols, const int* __restrict__ atom_z,
                                                const int* __restrict__ atom_ao, const double* __restrict__ xyz,
                                                const ElemDev* __restrict__ elem, const double2* __restrict__ pairpar,
                                                double4 vpar, double* __restrict__ w, double* __restrict__ slab,
                                                double* __restrict__ e_core, const double* __restrict__ s_num) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pairs) return;
  const PairDev P = pairs[g];
  const MolDev& M = mols[P.mol];
  if (M.dpath) return;  // k_pairs_d owns the pairs of molecules with a d shell
  int zf = atom_z[P.f], zs = atom_z[P.s];
  const ElemDev ef = elem[zf];
  const ElemDev es = elem[zs];
  double d[3] = {xyz[3 * P.s] - xyz[3 * P.f], xyz[3 * P.s + 1] - xyz[3 * P.f + 1], xyz[3 * P.s + 2] - xyz[3 * P.f + 2]};
  double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double* wp = w + P.w_off;
  // Two-electron frame points along -d (the reference's convention).
  double R[3][3];
  rot_to_x(-d[0] / r, -d[1] / r, -d[2] / r, R);
  bool hf = ef.n_orb >= 4, hs = es.n_orb >= 4;
  if (!hf && !hs) {
    double cf, pt;
    feather(r, cf, pt);
    double ee = EV / sqrt(r * r + sq(ef.rho0 + es.rho0));
    wp[0] = ee * cf + pt * (1.0 - cf);
  } else if (hf && !hs) {
    double ri[4];
    local_xh(ef, es, r, ri);
    wp[0] = ri[0];
    for (int k = 0; k < 3; ++k) wp[pack_idx(k + 1, 0)] = ri[1] * R[0][k];
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l <= k; ++l)
        wp[pack_idx(k + 1, l + 1)] = ri[2] * (R[0][k] * R[0][l]) + ri[3] * (R[1][k] * R[1][l] + R[2][k] * R[2][l]);
  } else {
    double ri[22];
    local_xx(ef, es, r, ri);
    rotate_xx_store(ri, R, wp);
  }
  // Resonance beta*S into the two off-diagonal blocks of H.
  double S[4][4];
  overlap_block(ef, es, d, r, S, P.pad > 0 ? s_num + 5 * (int64_t)(P.pad - 1) : nullptr);
  const int ld = M.ld;
  double* H = slab + M.mat0 + (int64_t)S_H * ld * ld;
  int of = atom_ao[P.f], os = atom_ao[P.s];
  for (int i = 0; i < ef.n_orb; ++i) {
    double bi = i == 0 ? ef.beta_s : ef.beta_p;
    for (int j = 0; j < es.n_orb; ++j) {
      double bj = j == 0 ? es.beta_s : es.beta_p;
      double v = 0.5 * (bi + bj) * S[i][j];
      H[(int64_t)(of + i) * ld + os + j] = v;
      H[(int64_t)(os + j) * ld + of + i] = v;
    }
  }
  e_core[g] = pair_core(ef, es, zf, zs, r, pairpar, vpar);
}

// H_aa[mu nu] = U delta + sum_b -Z_b (mu nu | s_b s_b), summed over partners in atom order.
__global__ void __launch_bounds__(128) k_hcore_diag(int n_items, const int2* __restrict__ items,
                                                     const PairDev* __restrict__ pairs, const MolDev* __restrict__ mols,
                                                     const int* __restrict__ atom_z, const int* __restrict__ atom_ao,
                                                     const ElemDev* __restrict__ elem, const double* __restrict__ w,
                                                     double* __restrict__ slab) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_items) return;
  int2 it = items[g];  // (molecule, global atom) packed element derived below
  int mol = it.x;
  int a = it.y & 0xFFFFFF, e = it.y >> 24;
  const MolDev& M = mols[mol];
  const ElemDev ea = elem[atom_z[a]];
  int mu = 0, nu = 0;
  for (int h = 0, c = 0; h < 9; ++h)
    for (int l = 0; l <= h; ++l, ++c)
      if (c == e) {
        mu = h;
        nu = l;
      }
  double acc = (mu == nu) ? (mu == 0 ? ea.u_ss : (mu <= 3 ? ea.u_pp : ea.u_dd)) : 0.0;
  int la = a - M.atom0;
  int npa = npair_of(ea.n_orb);
  for (int lb = 0; lb < M.n_atoms; ++lb) {
    if (lb == la) continue;
    int u = la < lb ? la : lb, v = la < lb ? lb : la;
    const PairDev P = pairs[M.pair0 + tri_index(u, v, M.n_atoms)];
    int b = M.atom0 + lb;
    const ElemDev eb = elem[atom_z[b]];
    const double* wp = w + P.w_off;
    const int npb = npair_of(eb.n_orb);
    double val;
    if (M.dpath) {
      // electron-core blocks stored after w: first atom's, then the second's
      val = P.f == a ? wp[(int64_t)npa * npb + e] : wp[(int64_t)npa * npb + npb + e];
      acc += val;
      continue;
    }
    if (P.f == a) {
      val = wp[(int64_t)e * npb];      // (mu nu | s s) with a first
    } else {
      val = wp[e];                      // row 0 of the (b, a) block
    }
    acc += -eb.core_charge * val;
  }
  const int ld = M.ld;
  double* H = slab + M.mat0;
  int o = atom_ao[a];
  H[(int64_t)(o + mu) * ld + o + nu] = acc;
  H[(int64_t)(o + nu) * ld + o + mu] = acc;
}

// Core-core energy of each molecule: its pairs summed in triangular order by one workgroup.
__global__ void __launch_bounds__(256) k_core_reduce(const MolDev* __restrict__ mols, const double* __restrict__ e_core,
                                                      double* __restrict__ e_mol) {
  __shared__ double red[256];
  const MolDev M = mols[blockIdx.x];
  const int64_t np = (int64_t)M.n_atoms * (M.n_atoms - 1) / 2;
  double acc = 0.0;
  for (int64_t p = threadIdx.x; p < np; p += 256) acc += e_core[M.pair0 + p];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) e_mol[blockIdx.x] = red[0];
}

// Pack each molecule's outputs (density, eigenvectors; beta copies for UHF) without padding.
__global__ void __launch_bounds__(256) k_gather(const MolDev* __restrict__ mols, const MolOut* __restrict__ out,
                                                 const double* __restrict__ slab, const int64_t* __restrict__ gath_off,
                                                 double* __restrict__ gath) {
  const MolDev M = mols[blockIdx.x];
  const MolOut R = out[blockIdx.x];
  const int n = M.nao, ld = M.ld;
  const int slots[4] = {R.p_slot, R.c_slot, R.pb_slot, R.cb_slot};
  const int nslots = M.uhf ? 4 : 2;
  double* dst = gath + gath_off[blockIdx.x];
  for (int k = 0; k < nslots; ++k) {
    const double* src = slab + M.mat0 + (int64_t)slots[k] * ld * ld;
    for (int idx = threadIdx.x; idx < n * n; idx += 256) {
      int i = idx / n, j = idx - i * n;
      dst[(int64_t)k * n * n + idx] = src[(int64_t)i * ld + j];
    }
  }
}

}  // namespace pm7

#include "integrals_d.h"

namespace pm7 {

// ---- nuclear gradient at the converged density (sp pairs) ----------------------------------
//
// At fixed density the NDDO energy is a sum of one-centre terms, which do not move, and one term
// per atom pair that depends only on the pair's separation d = R_s - R_f (pm7-rs src/gradient.rs:
// fixed-density Hellmann-Feynman gradient, 300-420; no Pulay term: the basis is orthonormal by
// construction). pair_energy is that term:
//   resonance       2 sum_{i in f, j in s} P_ij (beta_i + beta_j) S_ij / 2
//   electron-core   sum_{mu nu in f} P_mu nu (-Z_s)(mu nu|s s) + sum_{la si in s} P_la si (-Z_f)(s s|la si)
//   Coulomb         sum_{mu nu in f} sum_{la si in s} P_mu nu P_la si (mu nu|la si)
//   exchange        -sum_spin sum P^spin_mu la P^spin_nu si (mu nu|la si)    (P^a = P^b = P/2 for RHF)
//   core-core       ccrep(r)
// with the integrals built exactly as k_pairs builds them. Its derivative along d comes from the
// fourth-order central difference (8 [E(h) - E(-h)] - [E(2h) - E(-2h)]) / 12h, h = 1e-3 Bohr
// (truncation ~h^4, rounding ~1e-11 eV/Bohr).

// Densities of one pair, gathered once: packed one-centre blocks (off-diagonal doubled) and the
// f x s blocks of the total density and of each spin (nspin = 1 for RHF: P/2 stands for both).
struct PairDens {
  double pf[10], ps[10];
  double pfs[4][4];
  double xs[2][4][4];
  int nspin;
};

__device__ double pair_energy(const ElemDev& ef, const ElemDev& es, int zf, int zs, const double d[3],
                              const PairDens& D, const double2* __restrict__ pairpar, double4 vpar,
                              const double* snum) {
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const int nf = ef.n_orb, ns = es.n_orb;
  const int npf = npair_of(nf), nps = npair_of(ns);
  double w[100];
  double R[3][3];
  rot_to_x(-d[0] / r, -d[1] / r, -d[2] / r, R);
  if (nf < 4 && ns < 4) {
    double cf, pt;
    feather(r, cf, pt);
    w[0] = EV / sqrt(r * r + sq(ef.rho0 + es.rho0)) * cf + pt * (1.0 - cf);
  } else if (ns < 4) {
    double ri[4];
    local_xh(ef, es, r, ri);
    w[0] = ri[0];
    for (int k = 0; k < 3; ++k) w[pack_idx(k + 1, 0)] = ri[1] * R[0][k];
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l <= k; ++l)
        w[pack_idx(k + 1, l + 1)] = ri[2] * (R[0][k] * R[0][l]) + ri[3] * (R[1][k] * R[1][l] + R[2][k] * R[2][l]);
  } else {
    double ri[22];
    local_xx(ef, es, r, ri);
    rotate_xx_store(ri, R, w);
  }
  double S[4][4];
  overlap_block(ef, es, d, r, S, snum);
  double e = 0.0;
  for (int i = 0; i < nf; ++i) {
    const double bi = i == 0 ? ef.beta_s : ef.beta_p;
    for (int j = 0; j < ns; ++j) e += D.pfs[i][j] * (bi + (j == 0 ? es.beta_s : es.beta_p)) * S[i][j];
  }
  for (int a = 0; a < npf; ++a) {
    double row = 0.0;
    for (int b = 0; b < nps; ++b) row += w[a * nps + b] * D.ps[b];
    e += D.pf[a] * (row - es.core_charge * w[a * nps]);
  }
  for (int b = 0; b < nps; ++b) e -= ef.core_charge * D.ps[b] * w[b];
  const double kx = D.nspin == 1 ? -2.0 : -1.0;
  for (int sp = 0; sp < D.nspin; ++sp) {
    double ex = 0.0;
    for (int mu = 0; mu < nf; ++mu)
      for (int nu = 0; nu < nf; ++nu)
        for (int la = 0; la < ns; ++la)
          for (int si = 0; si < ns; ++si)
            ex += D.xs[sp][mu][la] * D.xs[sp][nu][si] * w[pack_idx(mu, nu) * nps + pack_idx(la, si)];
    e += kx * ex;
  }
  return e + pair_core(ef, es, zf, zs, r, pairpar, vpar);
}

// dE/dd of every sp-path pair of the molecules whose result holds a converged density
// (out[mol].p_slot); pairs whose overlap needs the numerical quadrature get NaN, which the host
// reports as an unsupported gradient. Pairs of MNDO/d molecules are k_pair_grad_d's.
__global__ void __launch_bounds__(128) k_pair_grad(int64_t n_pairs, const PairDev* __restrict__ pairs,
                                                    const MolDev* __restrict__ mols, const MolOut* __restrict__ out,
                                                    const int* __restrict__ atom_z, const int* __restrict__ atom_ao,
                                                    const double* __restrict__ xyz, const ElemDev* __restrict__ elem,
                                                    const double2* __restrict__ pairpar, double4 vpar,
                                                    const double* __restrict__ slab, const double* __restrict__ s_numg,
                                                    double* __restrict__ gpair) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pairs) return;
  const PairDev P = pairs[g];
  const MolDev& M = mols[P.mol];
  const MolOut& O = out[P.mol];
  double* go = gpair + 3 * g;
  if (M.dpath) return;  // k_pair_grad_d owns the pairs of molecules with a d shell
  if ((P.pad > 0 && !s_numg) || O.p_slot < 0) {
    go[0] = go[1] = go[2] = NAN;
    return;
  }
  const int zf = atom_z[P.f], zs = atom_z[P.s];
  const ElemDev ef = elem[zf];
  const ElemDev es = elem[zs];
  const int ld = M.ld, of = atom_ao[P.f], os = atom_ao[P.s];
  const int nf = ef.n_orb, ns = es.n_orb;
  const double* Pa = slab + M.mat0 + (int64_t)O.p_slot * ld * ld;
  const double* Pb = M.uhf ? slab + M.mat0 + (int64_t)O.pb_slot * ld * ld : nullptr;
  auto tot = [&](int i, int j) { return Pb ? Pa[(int64_t)i * ld + j] + Pb[(int64_t)i * ld + j] : Pa[(int64_t)i * ld + j]; };
  PairDens D;
  D.nspin = Pb ? 2 : 1;
  for (int mu = 0; mu < nf; ++mu)
    for (int nu = 0; nu <= mu; ++nu) D.pf[pack_idx(mu, nu)] = (mu == nu ? 1.0 : 2.0) * tot(of + mu, of + nu);
  for (int la = 0; la < ns; ++la)
    for (int si = 0; si <= la; ++si) D.ps[pack_idx(la, si)] = (la == si ? 1.0 : 2.0) * tot(os + la, os + si);
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j < ns; ++j) {
      D.pfs[i][j] = tot(of + i, os + j);
      // RHF: P/2 for both spins (the exchange weight carries the factor 2); UHF: P^a and P^b.
      D.xs[0][i][j] = Pb ? Pa[(int64_t)(of + i) * ld + os + j] : 0.5 * Pa[(int64_t)(of + i) * ld + os + j];
      D.xs[1][i][j] = Pb ? Pb[(int64_t)(of + i) * ld + os + j] : 0.0;
    }
  const double d0[3] = {xyz[3 * P.s] - xyz[3 * P.f], xyz[3 * P.s + 1] - xyz[3 * P.f + 1],
                        xyz[3 * P.s + 2] - xyz[3 * P.f + 2]};
  // numerical overlaps at the displaced separations, from k_overlap_numeric(ndisp = GRAD_DISP)
  const double* sn = P.pad > 0 ? s_numg + 5 * (int64_t)GRAD_DISP * (P.pad - 1) : nullptr;
  for (int k = 0; k < 3; ++k) {
    double ev[4];
    for (int t = 0; t < 4; ++t) {
      double d[3] = {d0[0], d0[1], d0[2]};
      d[k] += grad_step(t);
      ev[t] = pair_energy(ef, es, zf, zs, d, D, pairpar, vpar, sn ? sn + 5 * (k * 4 + t) : nullptr);
    }
    go[k] = (8.0 * (ev[0] - ev[1]) - (ev[2] - ev[3])) / (12.0 * GRAD_H);
  }
}

// dE/dd of every pair of an MNDO/d molecule, one 64-lane workgroup per pair: the pair's energy at
// fixed density (the terms of pair_energy, with the spd blocks of pair_d_eval: w, the electron-core
// blocks, the resonance block, core-core) at d +- h, d +- 2h along each axis, summed over the lanes
// in a fixed order, and the same fourth-order difference.
__global__ void __launch_bounds__(64) k_pair_grad_d(const int64_t* __restrict__ dpairs, const PairDev* __restrict__ pairs,
                                                     const MolDev* __restrict__ mols, const MolOut* __restrict__ out,
                                                     const int* __restrict__ atom_z, const int* __restrict__ atom_ao,
                                                     const double* __restrict__ xyz, const ElemDev* __restrict__ elem,
                                                     const double2* __restrict__ pairpar, double4 vpar,
                                                     const double* __restrict__ slab, double* __restrict__ gpair) {
  __shared__ DPairShared s;
  __shared__ double pf[45], ps[45], pfs[81], xs[2][81], red[64], ev[4];
  const int lane = threadIdx.x;
  const int64_t g = dpairs[blockIdx.x];
  const PairDev P = pairs[g];
  const MolDev& M = mols[P.mol];
  const MolOut& O = out[P.mol];
  double* go = gpair + 3 * g;
  if (O.p_slot < 0) {
    if (lane == 0) go[0] = go[1] = go[2] = NAN;
    return;
  }
  const int zf = atom_z[P.f], zs = atom_z[P.s];
  const ElemDev ef = elem[zf];
  const ElemDev es = elem[zs];
  const int nf = ef.n_orb, ns = es.n_orb, npf = npair_of(nf), nps = npair_of(ns);
  const int ld = M.ld, of = atom_ao[P.f], os = atom_ao[P.s];
  const double* Pa = slab + M.mat0 + (int64_t)O.p_slot * ld * ld;
  const double* Pb = M.uhf ? slab + M.mat0 + (int64_t)O.pb_slot * ld * ld : nullptr;
  auto tot = [&](int i, int j) { return Pb ? Pa[(int64_t)i * ld + j] + Pb[(int64_t)i * ld + j] : Pa[(int64_t)i * ld + j]; };
  for (int e = lane; e < npf; e += 64) {
    const int mu = c_dt.pa[e], nu = c_dt.pb[e];
    pf[e] = (mu == nu ? 1.0 : 2.0) * tot(of + mu, of + nu);
  }
  for (int e = lane; e < nps; e += 64) {
    const int la = c_dt.pa[e], si = c_dt.pb[e];
    ps[e] = (la == si ? 1.0 : 2.0) * tot(os + la, os + si);
  }
  for (int t = lane; t < nf * ns; t += 64) {
    const int i = t / ns, j = t - i * ns;
    pfs[t] = tot(of + i, os + j);
    const double a = Pa[(int64_t)(of + i) * ld + os + j];
    xs[0][t] = Pb ? a : 0.5 * a;  // RHF: P/2 stands for both spins
    xs[1][t] = Pb ? Pb[(int64_t)(of + i) * ld + os + j] : 0.0;
  }
  const int nspin = Pb ? 2 : 1;
  const double kx = Pb ? -1.0 : -2.0;
  const double d0[3] = {xyz[3 * P.s] - xyz[3 * P.f], xyz[3 * P.s + 1] - xyz[3 * P.f + 1],
                        xyz[3 * P.s + 2] - xyz[3 * P.f + 2]};
  __syncthreads();
  for (int k = 0; k < 3; ++k) {
    for (int t = 0; t < 4; ++t) {
      double d[3] = {d0[0], d0[1], d0[2]};
      d[k] += grad_step(t);
      pair_d_eval(ef, es, zf, zs, d, pairpar, vpar, s);
      double acc = 0.0;
      for (int idx = lane; idx < npf * nps; idx += 64) {
        const int ij = idx / nps, kl = idx - ij * nps;
        acc += pf[ij] * s.L[idx] * ps[kl];
      }
      for (int e = lane; e < npf + nps; e += 64) acc += (e < npf ? pf[e] : ps[e - npf]) * s.ecb[e];
      for (int t2 = lane; t2 < nf * ns; t2 += 64) acc += 2.0 * pfs[t2] * s.hres[t2];
      for (int q = lane; q < nf * nf; q += 64) {
        const int mu = q / nf, nu = q - mu * nf;
        const double* row = s.L + pack_idx(mu, nu) * nps;
        double ex = 0.0;
        for (int sp = 0; sp < nspin; ++sp)
          for (int la = 0; la < ns; ++la)
            for (int si = 0; si < ns; ++si)
              ex += xs[sp][mu * ns + la] * xs[sp][nu * ns + si] * row[pack_idx(la, si)];
        acc += kx * ex;
      }
      red[lane] = acc;
      __syncthreads();
      if (lane == 0) {
        double e = s.e_core;
        for (int i = 0; i < 64; ++i) e += red[i];
        ev[t] = e;
      }
      __syncthreads();
    }
    if (lane == 0) go[k] = (8.0 * (ev[0] - ev[1]) - (ev[2] - ev[3])) / (12.0 * GRAD_H);
    __syncthreads();
  }
}

// Gradient of each listed atom (items: molecule, batch-global atom): its pairs' dE/dd in partner
// order, + for the pair's second atom, - for its first.
__global__ void __launch_bounds__(128) k_atom_grad(int n_items, const int2* __restrict__ items,
                                                    const PairDev* __restrict__ pairs, const MolDev* __restrict__ mols,
                                                    const double* __restrict__ gpair, double* __restrict__ gatom) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_items) return;
  const int2 it = items[t];
  const MolDev& M = mols[it.x];
  const int a = it.y, la = a - M.atom0;
  double g[3] = {0.0, 0.0, 0.0};
  for (int lb = 0; lb < M.n_atoms; ++lb) {
    if (lb == la) continue;
    const int u = la < lb ? la : lb, v = la < lb ? lb : la;
    const int64_t pi = M.pair0 + tri_index(u, v, M.n_atoms);
    const double sgn = pairs[pi].s == a ? 1.0 : -1.0;
    for (int k = 0; k < 3; ++k) g[k] += sgn * gpair[3 * pi + k];
  }
  for (int k = 0; k < 3; ++k) gatom[3 * a + k] = g[k];
}

hipError_t launch_gradient(const BatchDev& b, int64_t n_pairs, const int2* items, int n_items, double* gpair,
                           double* gatom, double* s_numg, hipStream_t s) {
  if (b.n_num > 0)
    hipLaunchKernelGGL(k_overlap_numeric, dim3(b.n_num * GRAD_DISP), dim3(64), 0, s, b.n_num, b.num_pairs, b.pairs,
                       b.atom_z, b.xyz, b.elem, b.gl, s_numg, GRAD_DISP);
  if (n_pairs > 0)
    hipLaunchKernelGGL(k_pair_grad, dim3((unsigned)((n_pairs + 127) / 128)), dim3(128), 0, s, n_pairs, b.pairs, b.mols,
                       b.out, b.atom_z, b.atom_ao, b.xyz, b.elem, b.pairpar, b.vpar, b.slab,
                       b.n_num > 0 ? s_numg : nullptr, gpair);
  if (b.n_dp > 0)
    hipLaunchKernelGGL(k_pair_grad_d, dim3((unsigned)b.n_dp), dim3(64), 0, s, b.d_pairs, b.pairs, b.mols, b.out,
                       b.atom_z, b.atom_ao, b.xyz, b.elem, b.pairpar, b.vpar, b.slab, gpair);
  if (n_items > 0)
    hipLaunchKernelGGL(k_atom_grad, dim3((n_items + 127) / 128), dim3(128), 0, s, n_items, items, b.pairs, b.mols, gpair,
                       gatom);
  return hipGetLastError();
}

hipError_t launch_integrals(const BatchDev& b, int64_t n_pairs, int n_items, int n_mols, hipStream_t s) {
  if (b.n_dp > 0)
    hipLaunchKernelGGL(k_pairs_d, dim3((unsigned)b.n_dp), dim3(64), 0, s, b.d_pairs, b.pairs, b.mols, b.atom_z,
                       b.atom_ao, b.xyz, b.elem, b.pairpar, b.vpar, b.w, b.slab, b.e_core);
  if (b.n_num > 0)
    hipLaunchKernelGGL(k_overlap_numeric, dim3(b.n_num), dim3(64), 0, s, b.n_num, b.num_pairs, b.pairs, b.atom_z, b.xyz,
                       b.elem, b.gl, b.s_num, 0);
  if (n_pairs > 0) {
    int64_t blocks = (n_pairs + 127) / 128;
    hipLaunchKernelGGL(k_pairs, dim3((unsigned)blocks), dim3(128), 0, s, n_pairs, b.pairs, b.mols, b.atom_z, b.atom_ao,
                       b.xyz, b.elem, b.pairpar, b.vpar, b.w, b.slab, b.e_core, b.s_num);
  }
  if (n_items > 0) {
    hipLaunchKernelGGL(k_hcore_diag, dim3((n_items + 127) / 128), dim3(128), 0, s, n_items, b.hc_items, b.pairs, b.mols,
                       b.atom_z, b.atom_ao, b.elem, b.w, b.slab);
  }
  if (n_mols > 0) hipLaunchKernelGGL(k_core_reduce, dim3(n_mols), dim3(256), 0, s, b.mols, b.e_core, b.e_core_mol);
  return hipGetLastError();
}

// Zero `words` doubles (a fill kernel rather than a memset, so the clearing is an ordinary kernel
// node wherever the stream is captured).
__global__ void __launch_bounds__(256) k_zero(double* __restrict__ p, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (int64_t)gridDim.x * 256) p[i] = 0.0;
}

hipError_t launch_zero(double* p, int64_t words, hipStream_t s) {
  if (words <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((words + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, words);
  return hipGetLastError();
}

hipError_t launch_gather(const BatchDev& b, int n_mols, hipStream_t s) {
  if (n_mols > 0) hipLaunchKernelGGL(k_gather, dim3(n_mols), dim3(256), 0, s, b.mols, b.out, b.slab, b.gath_off, b.gath);
  return hipGetLastError();
}

}  // namespace pm7

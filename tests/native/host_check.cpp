// SPDX-License-Identifier: GPL-3.0-or-later
//
// Host-code check of the engine under AddressSanitizer + UndefinedBehaviorSanitizer (built by
// `make -C pm7-rs_amd sanitize`, run by tests/test_sanitizers.py): parameter loading and the
// derived MNDO/d data for every method's tables, the geometric corrections and their gradients,
// and the divide-and-conquer partition, checked against brute force. No device code.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dandc.h"
#include "host.h"

using namespace pm7;

static int failures = 0;
#define CHECK(c, ...)                     \
  do {                                    \
    if (!(c)) {                           \
      std::fprintf(stderr, __VA_ARGS__);  \
      std::fprintf(stderr, "\n");         \
      ++failures;                         \
    }                                     \
  } while (0)

// xorshift for reproducible geometries
static uint64_t rs = 88172645463325252ull;
static double urand() {
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return (double)(rs >> 11) / 9007199254740992.0;
}

static void corrections_on(const std::vector<int>& z, const std::vector<double>& xyz) {
  const int n = (int)z.size();
  for (int smooth = 0; smooth < 2; ++smooth) {
    const double e = dispersion_kcal(z, xyz, smooth != 0);
    std::vector<double> g(3 * n, 0.0);
    dispersion_grad_kcal(z, xyz, g.data(), smooth != 0);
    // central difference of one coordinate
    std::vector<double> x = xyz;
    const double h = 1e-5;
    x[0] += h;
    const double ep = dispersion_kcal(z, x, smooth != 0);
    x[0] -= 2 * h;
    const double em = dispersion_kcal(z, x, smooth != 0);
    CHECK(std::isfinite(e), "dispersion not finite");
    CHECK(std::fabs((ep - em) / (2 * h) - g[0]) < 1e-5 * (1 + std::fabs(g[0])), "dispersion gradient %g vs %g",
          g[0], (ep - em) / (2 * h));
  }
  std::vector<double> g(3 * n, 0.0);
  const double hh = hh_repulsion_kcal(z, xyz);
  hh_repulsion_grad_kcal(z, xyz, g.data());
  CHECK(std::isfinite(hh), "hh not finite");
  CHECK(std::isfinite(hbond_kcal(z, xyz)), "hbond not finite");
  CHECK(std::isfinite(c_triple_kcal(z, xyz)), "c_triple not finite");
  CHECK(std::isfinite(si_o_h_kcal(z, xyz)), "si_o_h not finite");
}

static void partition_on(const std::vector<double>& xyz, int core, double buffer) {
  const int n = (int)xyz.size() / 3;
  DcPartition p;
  dc_partition(n, xyz.data(), core, buffer, p);
  std::vector<int> owned(n, 0);
  for (int s = 0; s < p.n_sub(); ++s) {
    const int* m = p.members.data() + p.start[s];
    const int len = (int)(p.start[s + 1] - p.start[s]), nc = p.n_core[s];
    CHECK(nc >= 1 && nc <= std::max(core, 1), "core of %d atoms", nc);
    std::vector<char> in(n, 0);
    for (int i = 0; i < len; ++i) {
      CHECK(!in[m[i]], "atom %d twice in subsystem %d", m[i], s);
      in[m[i]] = 1;
    }
    for (int i = 0; i < nc; ++i) ++owned[m[i]];
    for (int b = 0; b < n; ++b) {
      double d2 = INFINITY;
      for (int i = 0; i < nc; ++i) {
        const double dx = xyz[3 * b] - xyz[3 * m[i]], dy = xyz[3 * b + 1] - xyz[3 * m[i] + 1],
                     dz = xyz[3 * b + 2] - xyz[3 * m[i] + 2];
        d2 = std::min(d2, dx * dx + dy * dy + dz * dz);
      }
      bool core_atom = false;
      for (int i = 0; i < nc; ++i) core_atom = core_atom || m[i] == b;
      if (!core_atom) CHECK((bool)in[b] == (d2 <= buffer * buffer), "buffer membership of atom %d in %d", b, s);
    }
  }
  for (int a = 0; a < n; ++a) CHECK(owned[a] == 1, "atom %d owned %d times", a, owned[a]);
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "pm7-rs_amd/data";
  for (int method : {0, 1}) {
    ParamSet P;
    std::string err;
    CHECK(load_params(dir, method, P, err), "load_params(%d): %s", method, err.c_str());
    std::vector<ElemDev> tab;
    to_device_table(P, tab);
    CHECK(tab.size() >= (size_t)MAX_Z, "device table of %zu entries", tab.size());
  }
  // Molecules: water, a C/H/O/N/S/P/Cl/Si mix in random positions (no two atoms within 0.9 A).
  const double A = 1.0 / 0.529177210903;
  corrections_on({8, 1, 1}, {0, 0, 0, 0.9584 * A, 0, 0, -0.24 * A, 0.9278 * A, 0});
  for (int trial = 0; trial < 4; ++trial) {
    const int zs[] = {1, 6, 7, 8, 14, 15, 16, 17};
    std::vector<int> z;
    std::vector<double> x;
    while ((int)z.size() < 40) {
      const double p[3] = {12 * urand() * A, 12 * urand() * A, 12 * urand() * A};
      bool ok = true;
      for (size_t k = 0; k < z.size(); ++k) {
        const double dx = p[0] - x[3 * k], dy = p[1] - x[3 * k + 1], dz = p[2] - x[3 * k + 2];
        ok = ok && dx * dx + dy * dy + dz * dz > 0.9 * 0.9 * A * A;
      }
      if (!ok) continue;
      z.push_back(zs[(int)(urand() * 8) % 8]);
      x.insert(x.end(), p, p + 3);
    }
    corrections_on(z, x);
  }
  // Partitions: a chain, a random cloud, coincident coordinates along one axis (ties).
  std::vector<double> chain;
  for (int i = 0; i < 200; ++i) chain.insert(chain.end(), {2.0 * i, 0.0, 0.0});
  partition_on(chain, 5, 6.0);
  partition_on(chain, 1, 0.5);
  std::vector<double> cloud;
  for (int i = 0; i < 500; ++i) cloud.insert(cloud.end(), {40 * urand(), 40 * urand(), 40 * urand()});
  partition_on(cloud, 12, 15.0);
  partition_on(cloud, 7, 3.0);
  std::vector<double> ties;
  for (int i = 0; i < 90; ++i) ties.insert(ties.end(), {1.26 * (i / 3), (double)(i % 3), 0.0});
  partition_on(ties, 8, 10.0);
  partition_on({0.0, 0.0, 0.0}, 4, 1.0);
  if (failures) std::fprintf(stderr, "host_check: %d failures\n", failures);
  else std::printf("host_check: ok\n");
  return failures ? 1 : 0;
}

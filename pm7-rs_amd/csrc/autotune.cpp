// SPDX-License-Identifier: GPL-3.0-or-later
//
// pm7_autotune: chooses a context's launch shapes by timing them on the caller's own batch.
//
// Knobs (TuneKnobs, kept by the engine): the eigensolver of each megakernel size class (block
// Jacobi or scalar rounds), the Jacobi sweep cap of the reduced-precision phase, and the number of
// batches in flight. The kernel-speed knobs are timed one batch at a time (device time of each
// batch, median over the steps), by coordinate descent from the current setting; the depth is then
// timed with `steps` batches through pm7_batch_submit / pm7_batch_wait (device span of all of
// them, and each batch's submit-to-result latency), and the fastest depth whose median latency
// stays within the budget wins. Nothing here touches the device directly: every measurement is a
// call of the public entry points plus their own HIP-event timings (pm7_last_timings).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pm7_amd.h"
#include "host.h"

namespace {

using pm7::TuneKnobs;

struct Measure {
  pm7_status st = PM7_OK;
  double solves_per_s = 0.0;
  double step_ms = 0.0;  // median submit-to-result latency
};

double median(std::vector<double> v) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t m = v.size() / 2;
  return v.size() % 2 ? v[m] : 0.5 * (v[m - 1] + v[m]);
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Runs `steps` batches with `depth` in flight and measures device throughput (span from the first
// batch's integrals kernel to the last batch's output gather) and the median submit-to-result
// latency. One untimed batch first.
Measure measure(pm7_ctx* ctx, int n, const pm7_molecule_v1* mols, const pm7_options_v2* opts, int n_opts,
                int steps, int depth) {
  Measure m;
  std::vector<pm7_result_v2> res((size_t)n * (steps + 1));
  {
    m.st = pm7_run_batch(ctx, n, mols, opts, n_opts, res.data() + (size_t)n * steps);
    for (int i = 0; i < n; ++i) pm7_result_free(&res[(size_t)n * steps + i]);
    if (m.st != PM7_OK) return m;
  }
  struct Flight {
    int64_t ticket;
    double t0;
  };
  std::deque<Flight> q;
  std::vector<double> lat;
  double first = 1e300, last = -1e300;
  auto finish_oldest = [&]() -> pm7_status {
    const Flight f = q.front();
    q.pop_front();
    const pm7_status st = pm7_batch_wait(ctx, f.ticket);
    lat.push_back(now_ms() - f.t0);
    double t[PM7_TIMING_SLOTS] = {};
    pm7_last_timings(ctx, t, PM7_TIMING_SLOTS);
    first = std::min(first, t[12]);
    last = std::max(last, t[13]);
    return st;
  };
  for (int s = 0; s < steps && m.st == PM7_OK; ++s) {
    const double t0 = now_ms();
    int64_t ticket = 0;
    m.st = pm7_batch_submit(ctx, n, mols, opts, n_opts, res.data() + (size_t)n * s, &ticket);
    if (m.st != PM7_OK) break;
    q.push_back({ticket, t0});
    if ((int)q.size() >= depth) m.st = finish_oldest();
  }
  while (!q.empty()) {
    const pm7_status st = finish_oldest();
    if (m.st == PM7_OK) m.st = st;
  }
  for (int s = 0; s < steps; ++s)
    for (int i = 0; i < n; ++i) pm7_result_free(&res[(size_t)n * s + i]);
  if (m.st != PM7_OK) return m;
  const double span = last - first;
  m.solves_per_s = span > 0.0 ? 1000.0 * (double)n * steps / span : 0.0;
  m.step_ms = median(lat);
  return m;
}

// Workload signature of a cache line: molecule count, total atoms and an FNV-1a hash of the atomic
// numbers and the options that change the work (precision, reference, multiplicity).
std::string signature(const pm7_ctx* ctx, int n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                      int n_opts) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    for (int b = 0; b < 8; ++b) {
      h ^= (v >> (8 * b)) & 0xff;
      h *= 1099511628211ull;
    }
  };
  int64_t atoms = 0;
  for (int i = 0; i < n; ++i) {
    atoms += mols[i].n_atoms;
    for (int a = 0; a < mols[i].n_atoms; ++a) mix((uint64_t)mols[i].z[a]);
    const pm7_options_v2* o = n_opts == 1 ? opts : (n_opts == n ? opts + i : nullptr);
    if (o) {
      mix((uint64_t)o->precision);
      mix((uint64_t)o->reference);
      mix((uint64_t)o->multiplicity);
    }
  }
  hipDeviceProp_t prop{};
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, pm7::ctx_device(ctx)) == hipSuccess) arch = prop.gcnArchName;
  char buf[160];
  std::snprintf(buf, sizeof buf, "%s/n%d/a%lld/h%016llx", arch.c_str(), n, (long long)atoms, (unsigned long long)h);
  for (char* c = buf; *c; ++c)
    if (*c == ' ') *c = '_';
  return buf;
}

bool cache_lookup(const char* path, const std::string& key, TuneKnobs& k, double& sps, double& ms) {
  std::ifstream f(path);
  std::string line;
  bool found = false;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    std::string kk;
    TuneKnobs t{};
    double a = 0.0, b = 0.0;
    if (!(in >> kk >> t.eig_block_mask >> t.mixed_sweeps >> t.batches_in_flight >> a >> b)) continue;
    if (kk != key) continue;
    k = t;
    sps = a;
    ms = b;
    found = true;  // the last matching line wins
  }
  return found;
}

void cache_store(const char* path, const std::string& key, const TuneKnobs& k, double sps, double ms) {
  std::vector<std::string> keep;
  {
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream in(line);
      std::string kk;
      if (in >> kk && kk != key) keep.push_back(line);
    }
  }
  std::ofstream f(path, std::ios::trunc);
  for (const std::string& l : keep) f << l << "\n";
  f << key << " " << k.eig_block_mask << " " << k.mixed_sweeps << " " << k.batches_in_flight << " " << sps << " "
    << ms << "\n";
}

}  // namespace

extern "C" pm7_status pm7_autotune(pm7_ctx* ctx, int32_t n, const pm7_molecule_v1* mols, const pm7_options_v2* opts,
                                   int32_t n_opts, int32_t steps, double latency_budget, const char* cache_path,
                                   pm7_tune_result_v1* out) {
  if (!ctx || !mols || n <= 0 || !out) return PM7_ERR_INVALID_INPUT;
  if (n_opts != 0 && n_opts != 1 && n_opts != n) return PM7_ERR_INVALID_INPUT;
  if (n_opts > 0 && !opts) return PM7_ERR_INVALID_INPUT;
  if (steps == 0) steps = 4;
  if (latency_budget == 0.0) latency_budget = 2.0;
  if (steps < 1 || steps > 64 || !(latency_budget >= 1.0)) return PM7_ERR_INVALID_INPUT;
  if (pm7::ctx_batches_in_flight(ctx) != 0) {
    pm7::ctx_set_error(ctx, "pm7_autotune: batches are in flight on this context");
    return PM7_ERR_INVALID_INPUT;
  }
  const double t_start = now_ms();
  *out = pm7_tune_result_v1{};
  const TuneKnobs start = pm7::tune_knobs(ctx);
  const std::string key = signature(ctx, n, mols, opts, n_opts);
  if (cache_path) {
    TuneKnobs k{};
    double sps = 0.0, ms = 0.0;
    if (cache_lookup(cache_path, key, k, sps, ms)) {
      pm7::set_tune_knobs(ctx, k);
      const TuneKnobs got = pm7::tune_knobs(ctx);
      *out = pm7_tune_result_v1{got.eig_block_mask, got.mixed_sweeps, got.batches_in_flight, 0, sps, ms, 0.0,
                                now_ms() - t_start};
      return PM7_OK;
    }
  }
  bool mixed = false;
  for (int i = 0; i < n_opts; ++i) mixed = mixed || opts[i].precision == PM7_PRECISION_MIXED;
  int candidates = 0;
  auto run = [&](const TuneKnobs& k, int depth, Measure& m) -> bool {
    TuneKnobs kk = k;
    kk.batches_in_flight = depth;
    pm7::set_tune_knobs(ctx, kk);
    m = measure(ctx, n, mols, opts, n_opts, steps, depth);
    ++candidates;
    return m.st == PM7_OK;
  };
  auto restore_fail = [&](pm7_status st) {
    pm7::set_tune_knobs(ctx, start);
    return st;
  };
  Measure base;
  if (!run(start, start.batches_in_flight, base)) return restore_fail(base.st);
  // Kernel-speed knobs, one batch at a time: coordinate descent from the current setting.
  TuneKnobs best = start;
  Measure m1;
  if (!run(best, 1, m1)) return restore_fail(m1.st);
  double best_sps = m1.solves_per_s;
  for (int bit = 0; bit < 3; ++bit) {
    TuneKnobs t = best;
    t.eig_block_mask ^= 1 << bit;
    Measure m;
    if (!run(t, 1, m)) return restore_fail(m.st);
    if (m.solves_per_s > best_sps * 1.01) {
      best = t;
      best_sps = m.solves_per_s;
    }
  }
  if (mixed) {
    const int caps[] = {3, 5, 8};
    for (int cap : caps) {
      if (cap == best.mixed_sweeps) continue;
      TuneKnobs t = best;
      t.mixed_sweeps = cap;
      Measure m;
      if (!run(t, 1, m)) return restore_fail(m.st);
      if (m.solves_per_s > best_sps * 1.01) {
        best = t;
        best_sps = m.solves_per_s;
      }
    }
  }
  // Depth: the highest throughput within the latency budget.
  Measure one;
  if (!run(best, 1, one)) return restore_fail(one.st);
  int best_depth = 1;
  Measure chosen = one;
  for (int depth = 2; depth <= 4; ++depth) {
    Measure m;
    if (!run(best, depth, m)) return restore_fail(m.st);
    if (m.step_ms <= latency_budget * one.step_ms && m.solves_per_s > chosen.solves_per_s * 1.01) {
      best_depth = depth;
      chosen = m;
    }
  }
  best.batches_in_flight = best_depth;
  pm7::set_tune_knobs(ctx, best);
  if (cache_path) cache_store(cache_path, key, best, chosen.solves_per_s, chosen.step_ms);
  *out = pm7_tune_result_v1{best.eig_block_mask, best.mixed_sweeps, best.batches_in_flight, candidates,
                            chosen.solves_per_s, chosen.step_ms, base.solves_per_s, now_ms() - t_start};
  return PM7_OK;
}

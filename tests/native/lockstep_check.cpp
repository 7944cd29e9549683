// SPDX-License-Identifier: GPL-3.0-or-later
//
// ThreadSanitizer check of the lockstep rounds the grouped large-system SCFs use to batch their
// eigensolves (pm7-rs_amd/csrc/lockstep.h), with the device work stubbed: participants run
// different numbers of rounds, hand in 0-2 items per round and leave at different times, some
// while the others wait. Every item must be processed exactly once, before its submit returns,
// in a round holding only items handed in while it was open. Built by
// `make -C pm7-rs_amd sanitize`, run by tests/test_sanitizers.py.
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

#include "lockstep.h"

struct Item {
  int owner, round, serial;
};

int main() {
  int failures = 0;
  for (int trial = 0; trial < 40; ++trial) {
    const int P = 2 + trial % 7;
    std::vector<std::vector<std::atomic<int>>> seen(P);
    std::vector<int> rounds_of(P);
    for (int p = 0; p < P; ++p) {
      rounds_of[p] = 1 + (p * 7 + trial * 3) % 23;
      seen[p] = std::vector<std::atomic<int>>(2 * rounds_of[p]);
      for (auto& s : seen[p]) s = 0;
    }
    std::atomic<int> bad{0};
    pm7::Lockstep<Item, int> ls(
        P, 2 * P, 0,
        [&](std::vector<Item>& items) {
          // every item of a round belongs to a different (owner, serial) and is seen once
          for (const Item& it : items) seen[it.owner][it.serial].fetch_add(1);
          return 0;
        },
        [&](int slot, void* caller) { return (slot >= 0 && caller != nullptr) ? 0 : 1; });
    std::vector<std::thread> th;
    for (int p = 0; p < P; ++p)
      th.emplace_back([&, p] {
        int serial = 0;
        for (int r = 0; r < rounds_of[p]; ++r) {
          const int k = (p + r + trial) % 3;  // 0, 1 or 2 items (2 = an open shell's alpha and beta)
          Item its[2];
          for (int q = 0; q < k && q < 2; ++q) its[q] = Item{p, r, serial + q};
          const int kk = k > 2 ? 2 : k;
          int token = p;
          if (ls.submit(its, kk, &token) != 0) bad.fetch_add(1);
          for (int q = 0; q < kk; ++q)
            if (seen[p][serial + q].load() != 1) bad.fetch_add(1);  // processed before submit returned
          serial += kk;
        }
        ls.leave();
      });
    for (auto& t : th) t.join();
    for (int p = 0; p < P; ++p)
      for (auto& s : seen[p])
        if (s.load() > 1) ++failures;
    failures += bad.load();
  }
  if (failures) std::fprintf(stderr, "lockstep_check: %d failures\n", failures);
  else std::printf("lockstep_check: ok\n");
  return failures ? 1 : 0;
}

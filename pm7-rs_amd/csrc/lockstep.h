// SPDX-License-Identifier: GPL-3.0-or-later
//
// Lockstep rounds of a fixed group of host threads (the grouped large-system SCFs, which batch
// their full eigensolves into one library call per SCF iteration). Host-only (std:: threading), so
// the synchronisation can be checked under ThreadSanitizer with the device work stubbed.
//
// Each participant calls submit() once per round with its items (possibly none) and blocks until
// the round has run; the last of the active participants to arrive runs the round, with the lock
// held, over every item handed in. A participant whose work is finished calls leave(); if the
// others are all waiting at that moment, the leaver runs their round. The first failing status
// sticks: later rounds run nothing and every submit returns it.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <vector>

namespace pm7 {

template <class Item, class Status>
class Lockstep {
 public:
  // run(items): the work of one full round; add(slot, caller): called for each item as it is
  // handed in (slot = its index in the round, caller = submit's token), e.g. to record an event.
  Lockstep(int participants, int cap, Status ok, std::function<Status(std::vector<Item>&)> run,
           std::function<Status(int, void*)> add)
      : active_(participants), cap_(cap), ok_(ok), status_(ok), run_(std::move(run)), add_(std::move(add)) {
    items_.reserve(cap);
  }

  // Hands in k items and waits for the round; returns the round's status.
  Status submit(const Item* its, int k, void* caller) {
    std::unique_lock<std::mutex> lk(mu_);
    for (int q = 0; q < k && status_ == ok_; ++q) {
      if ((int)items_.size() >= cap_) {
        status_ = overflow_;
        break;
      }
      const int slot = (int)items_.size();
      items_.push_back(its[q]);
      const Status s = add_(slot, caller);
      if (s != ok_) status_ = s;
    }
    const uint64_t mine = round_;
    if (++arrived_ == active_) run_round();
    else cv_.wait(lk, [&] { return round_ != mine; });
    return status_;
  }

  void leave() {
    std::lock_guard<std::mutex> lk(mu_);
    --active_;
    if (active_ > 0 && arrived_ == active_) run_round();
  }

  void set_overflow_status(Status s) { overflow_ = s; }
  uint64_t rounds() const { return round_; }  // rounds completed (call after every participant left)

 private:
  // With the lock held, by the last arrival.
  void run_round() {
    if (!items_.empty() && status_ == ok_) {
      const Status s = run_(items_);
      if (s != ok_) status_ = s;
    }
    items_.clear();
    arrived_ = 0;
    ++round_;
    cv_.notify_all();
  }

  std::mutex mu_;
  std::condition_variable cv_;
  int active_, arrived_ = 0, cap_;
  uint64_t round_ = 0;
  Status ok_, status_, overflow_{};
  std::vector<Item> items_;
  std::function<Status(std::vector<Item>&)> run_;
  std::function<Status(int, void*)> add_;
};

}  // namespace pm7

// c3d_wait.h -- the serving worker's wait (plain C++, no HIP).
#pragma once

#include <chrono>
#include <condition_variable>
#include <mutex>

namespace c3d {

// Waits on `cv` (with `lk` held) until pred() holds.  Before every wait,
// deadline(until) says whether pred() can become true by the passage of
// time alone (a request's batching window running out) and by when; the
// wait is then bounded by that time.  Deciding this again after every
// wake-up matters: a request that arrives during an unbounded wait makes
// the wait bounded from then on, so it starts when its window ends even if
// nothing else ever notifies `cv`.
template <class Pred, class Deadline>
void wait_work(std::unique_lock<std::mutex>& lk, std::condition_variable& cv, Pred pred, Deadline deadline) {
  while (!pred()) {
    std::chrono::steady_clock::time_point until;
    if (deadline(until))
      cv.wait_until(lk, until);
    else
      cv.wait(lk);
  }
}

}  // namespace c3d

// CPU check of c3d::wait_work (contour3d_amd/csrc/c3d_wait.h), the serving
// worker's wait: a request that arrives while the idle worker waits with no
// time bound must start when its batching window ends, with no further
// notification, and not before.  Built with -DDECIDE_ONCE it uses a wait
// that picks the bounded or unbounded form only once, before waiting (the
// mistake this guards against); that build must fail.
#include "c3d_wait.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <thread>

namespace {

double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct Shared {
  std::mutex m;
  std::condition_variable cv;
  std::deque<double> queue;  // submit times
  bool started = false;
  double t_start = 0;
};

void worker(Shared& S, double wait_s) {
  std::unique_lock<std::mutex> lk(S.m);
  auto pred = [&] { return !S.queue.empty() && now_s() >= S.queue.front() + wait_s; };
  auto deadline = [&](std::chrono::steady_clock::time_point& until) {
    if (S.queue.empty()) return false;
    until = std::chrono::steady_clock::now() + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                                   std::chrono::duration<double>(S.queue.front() + wait_s - now_s()));
    return true;
  };
#if defined(DECIDE_ONCE)
  std::chrono::steady_clock::time_point until;
  if (deadline(until))
    S.cv.wait_until(lk, until, pred);
  else
    S.cv.wait(lk, pred);
#else
  c3d::wait_work(lk, S.cv, pred, deadline);
#endif
  S.started = pred();
  S.t_start = now_s();
  S.cv.notify_all();
}

// one scenario: the worker is already waiting when `n` requests arrive
// together with one notification; returns 0 when the first starts in time
int scenario(int n, double wait_s) {
  static Shared pool[4];
  static int used = 0;
  Shared& S = pool[used++];
  std::thread t(worker, std::ref(S), wait_s);
  t.detach();  // a worker that never wakes must not keep the check from ending
  std::this_thread::sleep_for(std::chrono::milliseconds(50));
  double t_submit;
  {
    std::lock_guard<std::mutex> g(S.m);
    t_submit = now_s();
    for (int i = 0; i < n; i++) S.queue.push_back(t_submit);
  }
  S.cv.notify_all();
  std::unique_lock<std::mutex> lk(S.m);
  S.cv.wait_for(lk, std::chrono::seconds(2), [&] { return S.started; });
  if (!S.started) {
    std::printf("MISS n=%d window=%.3f s: not started 2 s after submit\n", n, wait_s);
    return 1;
  }
  const double dt = S.t_start - t_submit;
  std::printf("n=%d window=%.3f s: started after %.4f s\n", n, wait_s, dt);
  if (dt < wait_s || dt > wait_s + 0.5) {
    std::printf("MISS n=%d: start outside [window, window + 0.5 s]\n", n);
    return 1;
  }
  return 0;
}

}  // namespace

int main() {
  int bad = 0;
  bad += scenario(1, 0.02);
  bad += scenario(3, 0.05);
  std::fflush(stdout);
  std::_Exit(bad ? 1 : 0);  // detached workers may still be blocked
}

// The pm_pipeline hand-off protocol under ThreadSanitizer (built with g++ by
// tests/test_host_sanitizers.py): holders circulate released -> producer ->
// assigned -> consumer -> released through two BoundedQueues, as in
// vsgm_amd/csrc/host/pipeline.cc. Checks that every bin is consumed once
// with the contents its producer wrote, that a failing consumer's close()
// ends every thread, and that close() wakes a push blocked on a full queue.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <thread>
#include <vector>

#include "bounded_queue.h"

using vsgm::BoundedQueue;

namespace {

int failures = 0;

void Expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "check failed: %s\n", what);
    ++failures;
  }
}

// fail_at < 0: no failure; otherwise the consumer that takes that bin fails.
void Pipeline(int holders_n, int producers, int consumers, int bins, int fail_at) {
  std::vector<std::vector<uint64_t>> holders(holders_n, std::vector<uint64_t>(257));
  std::vector<int> holder_bin(holders_n, -1);
  BoundedQueue assigned(holders_n), released(holders_n);
  for (int i = 0; i < holders_n; ++i) released.push(i);
  std::atomic<int> next_bin{0}, producers_done{0};
  std::atomic<bool> failed{false};
  std::vector<std::atomic<int>> seen(bins);
  for (auto& s : seen) s = 0;
  std::atomic<int> bad{0};
  std::vector<std::thread> threads;
  for (int p = 0; p < producers; ++p)
    threads.emplace_back([&] {
      for (int bin; (bin = next_bin++) < bins && !failed;) {
        int h = -1;
        if (!released.pop(h)) break;
        for (size_t i = 0; i < holders[h].size(); ++i) holders[h][i] = (uint64_t)bin * 1000003u + i;
        holder_bin[h] = bin;
        assigned.push(h);
      }
      if (++producers_done == producers) assigned.close();
    });
  for (int c = 0; c < consumers; ++c)
    threads.emplace_back([&] {
      int h;
      while (assigned.pop(h)) {
        const int bin = holder_bin[h];
        for (size_t i = 0; i < holders[h].size(); ++i)
          if (holders[h][i] != (uint64_t)bin * 1000003u + i) ++bad;
        ++seen[bin];
        if (bin == fail_at) {
          failed = true;
          released.close();
          assigned.close();
          return;
        }
        released.push(h);
      }
    });
  for (auto& t : threads) t.join();
  Expect(bad == 0, "holder contents as written");
  int once = 0, more = 0;
  for (auto& s : seen) once += s == 1, more += s > 1;
  Expect(more == 0, "no bin consumed twice");
  if (fail_at < 0) Expect(once == bins, "every bin consumed");
}

void BlockedPushWakesOnClose() {
  BoundedQueue q(1);
  q.push(1);
  std::atomic<int> result{-1};
  std::thread t([&] { result = q.push(2) ? 1 : 0; });
  q.close();
  t.join();
  int v = 0;
  Expect(result == 0, "push after close is refused");
  Expect(q.pop(v) && v == 1, "queued item still drains after close");
  Expect(!q.pop(v), "drained closed queue reports the end");
}

}  // namespace

int main() {
  for (int rep = 0; rep < 20; ++rep) {
    Pipeline(1, 1, 1, 50, -1);
    Pipeline(2, 3, 1, 200, -1);
    Pipeline(4, 2, 3, 300, -1);
    Pipeline(3, 4, 2, 300, 17 + rep);
  }
  BlockedPushWakesOnClose();
  if (failures) return 1;
  std::printf("queue checks passed\n");
  return 0;
}

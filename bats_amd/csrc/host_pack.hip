// Host-side packing of inference-server batches (no GPU code): a batch's event lists, held in
// the server's host arena, are gathered back to back into the batch's pinned staging buffer
// with their row offsets, on a small pool of worker threads that lives as long as the process.
#include "common.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace {

// Runs a job split into parts on the caller and up to kWorkers threads, and waits for all.
class PackPool {
 public:
  static constexpr int kWorkers = 7;

  PackPool() {
    for (int k = 1; k <= kWorkers; ++k) {
      std::thread([this, k] { loop(k); }).detach();
    }
  }

  void run(int parts, const std::function<void(int, int)> &fn) {
    std::lock_guard<std::mutex> serial(callers_);
    parts = std::max(1, std::min(parts, kWorkers + 1));
    if (parts > 1) {
      std::lock_guard<std::mutex> l(m_);
      job_ = &fn;
      parts_ = parts;
      remaining_ = parts - 1;
      ++generation_;
    }
    if (parts > 1) go_.notify_all();
    fn(0, parts);
    if (parts > 1) {
      std::unique_lock<std::mutex> l(m_);
      done_.wait(l, [this] { return remaining_ == 0; });
    }
  }

 private:
  void loop(int k) {
    long seen = 0;
    std::unique_lock<std::mutex> l(m_);
    for (;;) {
      go_.wait(l, [&] { return generation_ != seen; });
      seen = generation_;
      if (k >= parts_) continue;
      const std::function<void(int, int)> *fn = job_;
      const int parts = parts_;
      l.unlock();
      (*fn)(k, parts);
      l.lock();
      if (--remaining_ == 0) done_.notify_one();
    }
  }

  std::mutex callers_, m_;
  std::condition_variable go_, done_;
  const std::function<void(int, int)> *job_ = nullptr;
  int parts_ = 0;
  int remaining_ = 0;
  long generation_ = 0;
};

PackPool &pack_pool() {
  static PackPool *pool = new PackPool();  // never destroyed: its threads run until exit
  return *pool;
}

}  // namespace

extern "C" int bats_pack_event_lists(int n, const int64_t *src_offsets, const int32_t *lengths,
                                     const int32_t *src_indices, const float *src_times,
                                     int32_t *row_offsets, int32_t *dst_indices,
                                     float *dst_times, int64_t dst_capacity) {
  using namespace bats;
  BATS_REQUIRE(n >= 0, "bad request count");
  BATS_REQUIRE(n == 0 || (src_offsets && lengths && src_indices && src_times && row_offsets),
               "null pointer");
  int64_t total = 0;
  for (int r = 0; r < n; ++r) {
    BATS_REQUIRE(lengths[r] >= 0 && src_offsets[r] >= 0, "negative offset or length");
    row_offsets[r] = (int32_t)total;
    total += lengths[r];
    BATS_REQUIRE(total <= dst_capacity && total <= INT32_MAX, "batch larger than its buffer");
  }
  if (row_offsets) row_offsets[n] = (int32_t)total;
  if (total == 0) return 0;
  BATS_REQUIRE(dst_indices && dst_times, "null destination");
  // parts of about 64 K events each, at most one per thread
  const int parts = (int)std::min<int64_t>(PackPool::kWorkers + 1, (total + 65535) / 65536);
  const std::function<void(int, int)> job = [&](int p, int P) {
    const int r0 = (int)((int64_t)n * p / P), r1 = (int)((int64_t)n * (p + 1) / P);
    for (int r = r0; r < r1; ++r) {
      const size_t len = (size_t)lengths[r];
      std::memcpy(dst_indices + row_offsets[r], src_indices + src_offsets[r], len * 4);
      std::memcpy(dst_times + row_offsets[r], src_times + src_offsets[r], len * 4);
    }
  };
  pack_pool().run(parts, job);
  return 0;
}

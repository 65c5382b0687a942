// Error state and small host helpers of the BATS HIP library.
#include "common.hpp"

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace bats {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(const std::string &msg) {
  set_error(msg);
  return 1;
}

int check_launch(const char *what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(err));
  return 0;
}

struct TimedLaunch {
  std::string name;
  hipEvent_t e0, e1;
};
static std::mutex g_timing_mu;
static std::atomic<bool> g_timing_on{false};
static std::vector<TimedLaunch> g_pending;
static std::map<std::string, std::pair<double, long long>> g_totals;

bool timing_enabled() { return g_timing_on.load(std::memory_order_relaxed); }

void timing_begin(const char *, hipStream_t st, hipEvent_t *e0, hipEvent_t *e1) {
  if (hipEventCreate(e0) != hipSuccess || hipEventCreate(e1) != hipSuccess) {
    *e0 = *e1 = nullptr;
    return;
  }
  (void)hipEventRecord(*e0, st);
}

void timing_end(const char *name, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  if (!e0 || !e1) return;
  (void)hipEventRecord(e1, st);
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_pending.push_back({name, e0, e1});
}

static void drain_pending() {
  for (auto &p : g_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.e1) == hipSuccess && hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
      auto &tot = g_totals[p.name];
      tot.first += ms;
      tot.second += 1;
    }
    (void)hipEventDestroy(p.e0);
    (void)hipEventDestroy(p.e1);
  }
  g_pending.clear();
}

}  // namespace bats

extern "C" void bats_timing_enable(int on) { bats::g_timing_on.store(on != 0); }

extern "C" void bats_timing_reset(void) {
  std::lock_guard<std::mutex> lk(bats::g_timing_mu);
  bats::drain_pending();
  bats::g_totals.clear();
}

extern "C" int bats_timing_query(const char *name, double *total_ms, long long *count) {
  std::lock_guard<std::mutex> lk(bats::g_timing_mu);
  bats::drain_pending();
  auto it = bats::g_totals.find(name ? name : "");
  if (total_ms) *total_ms = it == bats::g_totals.end() ? 0.0 : it->second.first;
  if (count) *count = it == bats::g_totals.end() ? 0 : it->second.second;
  return 0;
}

extern "C" const char *bats_last_error(void) { return bats::g_last_error.c_str(); }

extern "C" const char *bats_version(void) { return "bats_amd 0.1 (gfx950)"; }

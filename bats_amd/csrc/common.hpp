// Shared helpers of the BATS HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

namespace bats {

// ---------------------------------------------------------------- error reporting
void set_error(const std::string &msg);
int fail(const std::string &msg);
int check_launch(const char *what);

#define BATS_REQUIRE(cond, msg)                                   \
  do {                                                            \
    if (!(cond)) return ::bats::fail(std::string(__func__) + ": " + (msg)); \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Optional per-kernel timing with HIP events (off by default; not used under graph capture).
bool timing_enabled();
void timing_begin(const char *name, hipStream_t st, hipEvent_t *e0, hipEvent_t *e1);
void timing_end(const char *name, hipStream_t st, hipEvent_t e0, hipEvent_t e1);

// Launch wrapper: records start/stop events around the launch when timing is enabled (and
// `allow`, for launches that must not create events, is true).
#define BATS_TIMED_IF(allow, name, st, ...)                     \
  do {                                                          \
    hipEvent_t _e0 = nullptr, _e1 = nullptr;                    \
    const bool _t = (allow) && ::bats::timing_enabled();        \
    if (_t) ::bats::timing_begin(name, st, &_e0, &_e1);         \
    __VA_ARGS__;                                                \
    if (_t) ::bats::timing_end(name, st, _e0, _e1);             \
  } while (0)
#define BATS_TIMED(name, st, ...) BATS_TIMED_IF(true, name, st, __VA_ARGS__)

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// ---------------------------------------------------------------- device helpers
// Monotone map of a float's bits to an unsigned key (ascending float order = ascending key).
__device__ __forceinline__ uint32_t float_key(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

__device__ __forceinline__ bool is_inf_f(float v) { return isinf(v); }

constexpr float kInfF = __builtin_huge_valf();
constexpr double kInfD = __builtin_huge_val();

// Descriptor word 3 of a raw (untyped, unswizzled) buffer resource on gfx9-family targets.
constexpr int kBufferWord3 = 0x00020000;

// A float32 through a buffer resource: per-lane byte offset plus a wave-uniform byte offset
// (held in an SGPR, so the lane's address needs no per-load arithmetic).
__device__ __forceinline__ float buffer_load_f32(__amdgpu_buffer_rsrc_t r, int lane_off,
                                                 int uniform_off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, lane_off, uniform_off, 0));
}

// Read-only memory the same for every lane of a wave, read through scalar loads into SGPRs
// (the constant address space): the memory must not be written while the kernel runs.
typedef const __attribute__((address_space(4))) float uniform_f32;
typedef const __attribute__((address_space(4))) int uniform_i32;

__device__ __forceinline__ float4 uniform_load4(uniform_f32 *p, int k) {
  return make_float4(p[4 * k], p[4 * k + 1], p[4 * k + 2], p[4 * k + 3]);
}

// One time-sorted input event of one sample, read uniformly by every lane of a wave.
// es/e are the float32 exps exactly as stored for the backward, widened to float64. The
// crossing filter's bounds for the interval after the event are precomputed in float64 with
// zq = exp(t/tau) and zl = exp(min(t_next, t_max)/tau) (capped at 1e300):
//   czl2 = c zl^2, zl1 = zl (1 + eps), blo = 2c zq (1 - eps), bhi = 2c zl (1 + eps).
// tl = min(t_next, t_max). Padding events (after the sample's last event) have es = e = 0
// and bounds that no state passes.
struct alignas(16) Event {
  double es;
  double e;
  double czl2;
  double zl1;
  double blo;
  double bhi;
  double zq;  // exp(t / tau)
  double zl;  // exp(tl / tau), at most 1e300
  int idx;
  float t;
  float tl;
  float pad;
};
static_assert(sizeof(Event) == 80, "Event layout");

// Relative margin of the crossing filter (far above float64 rounding of the quantities it
// compares, far below the spacing of distinct float32 spike times).
constexpr double kFilterEps = 1e-10;
// Events are padded to a multiple of this many per sample.
constexpr int kEventPad = 8;

// Per-sample time-ordered event lists built by sort_events_kernel (lif_forward.hip).
// The fields of an event the lane-per-neuron solver reads for every event: the two exps and
// the interval ends, widened by kFilterEps (the exact ends are in the Event record).
struct alignas(32) FastEvent {
  double es;
  double e;
  double zq_lo;  // exp(t/tau) * (1 - kFilterEps)
  double zl_hi;  // exp(tl/tau) * (1 + kFilterEps)
};
static_assert(sizeof(FastEvent) == 32, "FastEvent layout");

struct EventLists {
  Event *events;  // [B, Pcap]
  FastEvent *fast;  // [B, Pcap]
  int *idx;       // [B, Pcap] pre-neuron index of each event (also in Event), for wide loads
  int *counts;    // [B] real events per sample (padding excluded)
  int Pcap;       // multiple of kEventPad
};
size_t event_lists_bytes(int B, int S);
int build_event_lists(hipStream_t st, int B, int S, int K_pre, const float *pre_times,
                      float tau_s, float tau, float c, float tmax, float *pre_es, float *pre_e,
                      void *ws, EventLists *out, bool timed = true);

// Workgroups are dealt to the 8 XCDs round robin; this maps dispatch index p of nblk to a
// logical index such that consecutive logical blocks run on one XCD and share its L2.
__device__ __forceinline__ int xcd_logical_block(int p, int nblk) {
  constexpr int kXcd = 8;
  const int q = nblk / kXcd, r = nblk % kXcd;
  const int xcd = p % kXcd, k = p / kXcd;
  return xcd * q + min(xcd, r) + k;
}

// LDS ordering inside one wave: its DS instructions execute in order, so a fence at
// wavefront scope (which also keeps the compiler from moving LDS accesses across it) is all
// that a wave-private array needs.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Fill kernel (used instead of hipMemsetAsync so that every call stays graph-capturable).
template <typename T>
__global__ void fill_kernel(T *__restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace bats

// Event-pair timer behind gm1hip_kernel_timer_arm / _collect.
#include <vector>

#include "timing.h"

namespace gm1 {

int g_timed_kernel = TK_NONE;

namespace {
struct Pair {
  hipEvent_t start = nullptr, stop = nullptr;
};
std::vector<Pair> g_pairs;
size_t g_used = 0;
bool g_overflow = false;

void release() {
  for (Pair& p : g_pairs) {
    if (p.start) (void)hipEventDestroy(p.start);
    if (p.stop) (void)hipEventDestroy(p.stop);
  }
  g_pairs.clear();
  g_used = 0;
}
}  // namespace

void timer_begin(hipStream_t s) {
  if (g_used < g_pairs.size()) (void)hipEventRecord(g_pairs[g_used].start, s);
  else g_overflow = true;
}

void timer_end(hipStream_t s) {
  if (g_used < g_pairs.size()) (void)hipEventRecord(g_pairs[g_used++].stop, s);
}

}  // namespace gm1

extern "C" {

// Arm timing of one kernel id for up to `capacity` launches (events are
// created here, before the timed region).  kernel 0 disarms.
int gm1hip_kernel_timer_arm(int kernel, int capacity) {
  using namespace gm1;
  g_timed_kernel = TK_NONE;
  release();
  g_overflow = false;
  if (kernel == TK_NONE) return 0;
  if (kernel < TK_ACOUSTIC_COLUMN || kernel > TK_ACOUSTIC_SMALL_STEP || capacity < 1) return (int)hipErrorInvalidValue;
  g_pairs.resize((size_t)capacity);
  for (Pair& p : g_pairs) {
    hipError_t e = hipEventCreate(&p.start);
    if (e == hipSuccess) e = hipEventCreate(&p.stop);
    if (e != hipSuccess) {
      release();
      return (int)e;
    }
  }
  g_timed_kernel = kernel;
  return 0;
}

// Wait for the recorded launches, sum their durations and disarm.
// Returns hipErrorInvalidValue if more launches happened than were armed.
int gm1hip_kernel_timer_collect(double* total_ms, int* launches) {
  using namespace gm1;
  g_timed_kernel = TK_NONE;
  double sum = 0.0;
  for (size_t q = 0; q < g_used; ++q) {
    hipError_t e = hipEventSynchronize(g_pairs[q].stop);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, g_pairs[q].start, g_pairs[q].stop);
    if (e != hipSuccess) {
      release();
      return (int)e;
    }
    sum += ms;
  }
  *total_ms = sum;
  *launches = (int)g_used;
  const bool overflow = g_overflow;
  release();
  return overflow ? (int)hipErrorInvalidValue : 0;
}

}  // extern "C"

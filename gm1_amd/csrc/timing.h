// Live per-kernel timing for throughput reports.
//
// A caller arms one kernel id with a capacity; every launch of that kernel
// is then bracketed by a pair of HIP events recorded on the launch stream
// (same-stream events add no synchronisation).  Collecting waits for the
// last event and returns the summed GPU time.  Disarmed (the default), the
// hooks are one integer compare per launch.
#pragma once
#include <hip/hip_runtime.h>

namespace gm1 {

enum TimedKernel : int {
  TK_NONE = 0,
  TK_ACOUSTIC_COLUMN = 1,  // k_snd_column: one small step's vertical pass
  TK_ACOUSTIC_PGRAD = 2,   // k_snd_pgrad: one small step's horizontal pass
  TK_SCALAR_ADV5 = 3,      // k_scalar_adv5: one scalar's fluxes + tendency
  TK_MOM_ADVU = 4,
  TK_MOM_ADVV = 5,
  TK_MOM_ADVW = 6,
  TK_ACOUSTIC_SMALL_STEP = 7,  // one small step's pressure-gradient and column work (one or two passes)
};

extern int g_timed_kernel;
void timer_begin(hipStream_t s);
void timer_end(hipStream_t s);

}  // namespace gm1

#define GM1_TIMED(id, stream, launch)                      \
  do {                                                     \
    const bool gm1_t_ = ::gm1::g_timed_kernel == (id);     \
    if (gm1_t_) ::gm1::timer_begin(stream);                \
    launch;                                                \
    if (gm1_t_) ::gm1::timer_end(stream);                  \
  } while (0)

// FP32 instances of the fast-path sweeps (FP32 or FP64 field I/O).
#include "force_fast_impl.h"

namespace gfmd {

int run_fast_f32(const ForceArgs &a, bool io_double, hipStream_t st,
                 hipEvent_t *ev) {
  if (io_double) return run_fast<double, float>(a, st, ev);
  return run_fast<float, float>(a, st, ev);
}

}  // namespace gfmd

// FP64 instances of the fast-path sweeps.
#include "force_fast_impl.h"

namespace gfmd {

int run_fast_f64(const ForceArgs &a, hipStream_t st, hipEvent_t *ev) {
  return run_fast<double, double>(a, st, ev);
}

}  // namespace gfmd

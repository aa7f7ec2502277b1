// culamp_pass_fp32.hip -- mixed-precision (FP32 pair arithmetic) instantiations
// of the pair-pass kernel.
#include "culamp_pair_kernel.h"

namespace culamp {
hipError_t launch_pass_fp32(const PassDesc& d, int kind, const LaunchCfg& c,
                            const Work& w, hipStream_t s) {
  if (kind == KIND_LSE) return launch_cost<1, KIND_LSE>(d, c, w, s);
  return launch_cost<1, KIND_SUM>(d, c, w, s);
}
}  // namespace culamp

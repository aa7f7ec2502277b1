// culamp_pass_fp64.hip -- FP64 instantiations of the pair-pass kernel.
#include "culamp_pair_kernel.h"

namespace culamp {
hipError_t launch_pass_fp64(const PassDesc& d, int kind, const LaunchCfg& c,
                            const Work& w, hipStream_t s) {
  switch (kind) {
    case KIND_LSE: return launch_cost<0, KIND_LSE>(d, c, w, s);
    case KIND_SUM: return launch_cost<0, KIND_SUM>(d, c, w, s);
    case KIND_BARY: return launch_cost<0, KIND_BARY>(d, c, w, s);
    case KIND_MAX: return launch_cost<0, KIND_MAX>(d, c, w, s);
    default: return launch_cost<0, KIND_MIN>(d, c, w, s);
  }
}
hipError_t launch_feature_fp64(const PassDesc& d, int kind, const LaunchCfg& c,
                               const Work& w, hipStream_t s) {
  switch (kind) {
    case KIND_LSE: return launch_feature_metric<0, KIND_LSE>(d, c, w, s);
    case KIND_SUM: return launch_feature_metric<0, KIND_SUM>(d, c, w, s);
    case KIND_MAX: return launch_feature_metric<0, KIND_MAX>(d, c, w, s);
    case KIND_MIN: return launch_feature_metric<0, KIND_MIN>(d, c, w, s);
    default: return hipErrorInvalidValue;
  }
}
hipError_t launch_batch_fp64(const BatchLaunch& b, int kind, const Work& w, hipStream_t s) {
  if (b.family == FAMILY_FEATURES) {
    switch (kind) {
      case KIND_LSE: return launch_batch_features<0, KIND_LSE>(b, w, s);
      case KIND_SUM: return launch_batch_features<0, KIND_SUM>(b, w, s);
      case KIND_MAX: return launch_batch_features<0, KIND_MAX>(b, w, s);
      case KIND_MIN: return launch_batch_features<0, KIND_MIN>(b, w, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (b.family != FAMILY_POINTS) return hipErrorInvalidValue;
  switch (kind) {
    case KIND_LSE: return launch_batch_points<0, KIND_LSE>(b, w, s);
    case KIND_SUM: return launch_batch_points<0, KIND_SUM>(b, w, s);
    case KIND_MAX: return launch_batch_points<0, KIND_MAX>(b, w, s);
    case KIND_MIN: return launch_batch_points<0, KIND_MIN>(b, w, s);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace culamp

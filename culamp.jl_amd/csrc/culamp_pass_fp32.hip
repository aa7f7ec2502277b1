// culamp_pass_fp32.hip -- mixed-precision (FP32 pair arithmetic) instantiations
// of the pair-pass kernel.
#include "culamp_pair_kernel.h"

namespace culamp {
hipError_t launch_pass_fp32(const PassDesc& d, int kind, const LaunchCfg& c,
                            const Work& w, hipStream_t s) {
  if (kind == KIND_LSE) return launch_cost<1, KIND_LSE>(d, c, w, s);
  return launch_cost<1, KIND_SUM>(d, c, w, s);
}
hipError_t launch_feature_fp32(const PassDesc& d, int kind, const LaunchCfg& c,
                               const Work& w, hipStream_t s) {
  if (kind == KIND_LSE) return launch_feature_metric<1, KIND_LSE>(d, c, w, s);
  return launch_feature_metric<1, KIND_SUM>(d, c, w, s);
}
hipError_t launch_feature_mfma(const PassDesc& d, int kind, int nsplit, const LaunchCfg& c,
                               const Work& w, hipStream_t s) {
  if (kind == KIND_LSE) {
    return nsplit == 3 ? launch_mfma_ks<KIND_LSE, 3>(d, c, w, s) : launch_mfma_ks<KIND_LSE, 1>(d, c, w, s);
  }
  return nsplit == 3 ? launch_mfma_ks<KIND_SUM, 3>(d, c, w, s) : launch_mfma_ks<KIND_SUM, 1>(d, c, w, s);
}
hipError_t launch_batch_fp32(const BatchLaunch& b, int kind, const Work& w, hipStream_t s) {
  if (kind != KIND_LSE && kind != KIND_SUM) return hipErrorInvalidValue;
  if (b.family == FAMILY_FEATURES)
    return kind == KIND_LSE ? launch_batch_features<1, KIND_LSE>(b, w, s) : launch_batch_features<1, KIND_SUM>(b, w, s);
  if (b.family == FAMILY_POINTS)
    return kind == KIND_LSE ? launch_batch_points<1, KIND_LSE>(b, w, s) : launch_batch_points<1, KIND_SUM>(b, w, s);
  return hipErrorInvalidValue;
}
hipError_t launch_batch_mfma(const BatchLaunch& b, int kind, const Work& w, hipStream_t s) {
  if (kind == KIND_LSE)
    return b.nsplit == 3 ? launch_batch_mfma_ks<KIND_LSE, 3>(b, w, s) : launch_batch_mfma_ks<KIND_LSE, 1>(b, w, s);
  if (kind == KIND_SUM)
    return b.nsplit == 3 ? launch_batch_mfma_ks<KIND_SUM, 3>(b, w, s) : launch_batch_mfma_ks<KIND_SUM, 1>(b, w, s);
  return hipErrorInvalidValue;
}
}  // namespace culamp

// culamp_pairs.hip -- launch-shape selection and dispatch of the pair-pass
// kernel (culamp_pair_kernel.h), plus the small reduction and elementwise
// kernels the host loops use.
#include "culamp_internal.h"

#include <math.h>

namespace culamp {
namespace {

constexpr int kSub = 8;  // must match the pair kernel's sub-tile

// one workgroup per reduction of the spec
__global__ __launch_bounds__(kReduceThreads) void reduce_kernel(const ReduceSpec spec, double* out) {
  reduce_op(spec, blockIdx.x, out);
}

__global__ void axpby_kernel(double* out, const double* a, double sa,
                             const double* b, double sb, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = axpby_elem(sa, a[i], sb, b[i]);
}
__global__ void copy3_kernel(double* d1, const double* s1, double* d2, const double* s2,
                             int64_t n, double* d3, const double* s3, int64_t n3) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    d1[i] = s1[i];
    if (d2 != nullptr) d2[i] = s2[i];
  }
  if (i < n3) d3[i] = s3[i];
}
__global__ void log_kernel(double* out, const double* a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = log(a[i]);
}
__global__ void fill_kernel(double* out, double v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}
__global__ void etamu_kernel(double* out, const double* c, double add, double div, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (c[i] + add) / div;
}
__global__ void log_minus_kernel(double* out, const double* r, const double* a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = log(r[i]) - a[i];
}
__global__ void theta_update_kernel(double* theta, const double* theta0,
                                    const double* res, const double* c,
                                    const double* etamu, int64_t m,
                                    double eta_mu, int adjust, double minv,
                                    double maxv, double winf) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m)
    theta[j] = theta_update(res[j], c[j], etamu[j], theta0[j], eta_mu, adjust,
                            minv, maxv, winf);
}

inline unsigned lin_blocks(int64_t n) { return (unsigned)((n + 255) / 256); }


}  // namespace

// FP64 fields per target of a split partial (the sums travel with their
// compensation terms): BARY 6, SUM 4, LSE 3, MAX / MIN 1.
int pass_fields(int kind) {
  return kind == KIND_BARY ? 6 : (kind == KIND_MAX || kind == KIND_MIN) ? 1 : kind == KIND_SUM ? 4 : 3;
}

// Launch shape of the point kernel: kLanesPerTarget lanes share a target, so
// a workgroup owns (256 / kLanesPerTarget) * rpt targets.  Take the largest
// rpt (at most 2) that still gives ~8 workgroups per CU once the sources are
// split into chunks of >= 256 points (one LDS tile); fewer, longer splits
// keep the merge traffic small.
LaunchCfg choose_launch(int64_t nt, int64_t ns, int precision, int kind,
                        int num_cu, int force_rpt, int force_splits, int layout) {
  LaunchCfg c{};
  const int64_t want = (int64_t)8 * (num_cu > 0 ? num_cu : 256);
  const int64_t tpb = kThreads / kLanesPerTarget;
  const bool fp64_only = kind == KIND_BARY || kind == KIND_MAX || kind == KIND_MIN;
  // (4 targets per thread measured slower than 2 in mixed precision at
  // 8192-16384 targets: the extra registers cost more than the saved LDS reads)
  const int max_rpt = 2;
  int rpt = 1;
  int64_t splits = 1;
  const int64_t cap = (ns + kTile - 1) / kTile;
  for (int cand = max_rpt; cand >= 1; cand /= 2) {
    const int64_t tb = (nt + tpb * cand - 1) / (tpb * cand);
    int64_t sp = (want + tb - 1) / tb;
    if (sp > cap) sp = cap;
    if (sp < 1) sp = 1;
    rpt = cand;
    splits = sp;
    if (tb * sp >= want) break;
  }
  const bool allow4 = precision == 1 && !fp64_only;  // (forced shapes only)
  if (force_rpt == 1 || force_rpt == 2 || (force_rpt == 4 && allow4)) rpt = force_rpt;
  if (force_splits > 0) splits = force_splits;
  if (splits > 65535) splits = 65535;
  int64_t len = (ns + splits - 1) / splits;
  len = (len + kSub - 1) / kSub * kSub;
  if (len < kSub) len = kSub;
  splits = ns > 0 ? (ns + len - 1) / len : 1;
  c.rpt = rpt;
  c.splits = (int)splits;
  c.split_len = len;
  c.tblocks = (nt + tpb * rpt - 1) / (tpb * rpt);
  if (c.tblocks < 1) c.tblocks = 1;
  // Workgroup-wide splits when one target per lane group leaves many splits
  // to merge and the targets alone fill every CU (kWideTargets per
  // workgroup): the split merge then needs no second launch.  A forced
  // launch shape keeps the split layout unless the wide one is asked for.
  // (Only about one workgroup per CU: a 16-wave workgroup holds a CU's
  // registers alone, so a second round of workgroups runs as a tail; at
  // 8192 targets the split layout measured 15 % faster.)  Both precisions:
  // at 4096 x 4096 the FP64 row LSE takes 24.0 against 24.3 us and the FP64
  // column pass 20.0 against 20.7 us.
  const bool forced = force_rpt != 0 || force_splits != 0;
  const int64_t ncu = num_cu > 0 ? num_cu : 256;
  const int64_t wide_blocks = (nt + kWideTargets - 1) / kWideTargets;
  const bool auto_wide = !forced && layout == LAYOUT_AUTO && !fp64_only &&
                         c.rpt == 1 && c.splits >= 8 && wide_blocks >= ncu &&
                         wide_blocks <= ncu + ncu / 4;
  if ((layout == LAYOUT_WIDE || auto_wide) && kind != KIND_BARY) {
    int64_t want_splits = force_splits > 0 ? force_splits : kWideWaves;
    if (want_splits > kWideWaves) want_splits = kWideWaves;
    int64_t wl = (ns + want_splits - 1) / want_splits;
    wl = (wl + kSub - 1) / kSub * kSub;
    if (wl < kSub) wl = kSub;
    c.wide = 1;
    c.rpt = 1;
    c.split_len = wl;
    c.splits = ns > 0 ? (int)((ns + wl - 1) / wl) : 1;
    c.tblocks = (nt + kWideTargets - 1) / kWideTargets;
    if (c.tblocks < 1) c.tblocks = 1;
  }
  return c;
}

hipError_t launch_pass(const PassDesc& d, int precision, int kind,
                       const LaunchCfg& c, const Work& w, hipStream_t s) {
  const bool fp64_only = kind == KIND_BARY || kind == KIND_MAX || kind == KIND_MIN;
  if (precision == 1 && !fp64_only) return launch_pass_fp32(d, kind, c, w, s);
  return launch_pass_fp64(d, kind, c, w, s);
}

LaunchCfg choose_feature_launch(int64_t nt, int64_t ns, int num_cu, int force_splits) {
  LaunchCfg c{};
  const int64_t want = (int64_t)8 * (num_cu > 0 ? num_cu : 256);
  c.rpt = 1;
  c.tblocks = (nt + kFeatTargets - 1) / kFeatTargets;
  if (c.tblocks < 1) c.tblocks = 1;
  int64_t splits = (want + c.tblocks - 1) / c.tblocks;
  const int64_t cap = (ns + 63) / 64;
  if (splits > cap) splits = cap;
  if (force_splits > 0) splits = force_splits;
  if (splits < 1) splits = 1;
  if (splits > 65535) splits = 65535;
  int64_t len = (ns + splits - 1) / splits;
  if (len < 1) len = 1;
  c.split_len = len;
  c.splits = ns > 0 ? (int)((ns + len - 1) / len) : 1;
  return c;
}

LaunchCfg choose_mfma_launch(int64_t nt, int64_t ns, int num_cu, int force_splits) {
  LaunchCfg c{};
  const int64_t want = (int64_t)8 * (num_cu > 0 ? num_cu : 256);
  c.rpt = 1;
  c.tblocks = (nt + 127) / 128;
  if (c.tblocks < 1) c.tblocks = 1;
  int64_t splits = (want + c.tblocks - 1) / c.tblocks;
  const int64_t cap = (ns + 63) / 64;
  if (splits > cap) splits = cap;
  if (force_splits > 0) splits = force_splits;
  if (splits < 1) splits = 1;
  if (splits > 65535) splits = 65535;
  int64_t len = (ns + splits - 1) / splits;
  len = (len + 31) / 32 * 32;
  if (len < 32) len = 32;
  c.split_len = len;
  c.splits = ns > 0 ? (int)((ns + len - 1) / len) : 1;
  return c;
}

hipError_t launch_reduce(const ReduceSpec& spec, double* out, hipStream_t s) {
  if (spec.nops <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)spec.nops), dim3(kReduceThreads), 0, s, spec,
                     out);
  return hipGetLastError();
}

hipError_t launch_axpby(double* out, const double* a, double sa, const double* b,
                        double sb, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpby_kernel, dim3(lin_blocks(n)), dim3(256), 0, s, out, a, sa, b, sb, n);
  return hipGetLastError();
}
hipError_t launch_copy3(double* d1, const double* s1, double* d2, const double* s2, int64_t n,
                        double* d3, const double* s3, int64_t n3, hipStream_t s) {
  const int64_t len = n > n3 ? n : n3;
  if (len <= 0) return hipSuccess;
  hipLaunchKernelGGL(copy3_kernel, dim3(lin_blocks(len)), dim3(256), 0, s, d1, s1, d2, s2, n, d3,
                     s3, n3);
  return hipGetLastError();
}
hipError_t launch_log(double* out, const double* a, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(log_kernel, dim3(lin_blocks(n)), dim3(256), 0, s, out, a, n);
  return hipGetLastError();
}
hipError_t launch_fill(double* out, double v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3(lin_blocks(n)), dim3(256), 0, s, out, v, n);
  return hipGetLastError();
}
hipError_t launch_etamu(double* out, const double* c, double add, double div,
                        int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(etamu_kernel, dim3(lin_blocks(n)), dim3(256), 0, s, out, c, add, div, n);
  return hipGetLastError();
}
hipError_t launch_log_minus(double* out, const double* r, const double* a,
                            int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(log_minus_kernel, dim3(lin_blocks(n)), dim3(256), 0, s, out, r, a, n);
  return hipGetLastError();
}
hipError_t launch_theta_update(double* theta, const double* theta0,
                               const double* res, const double* c,
                               const double* etamu, int64_t m, double eta_mu,
                               int adjust, double minv, double maxv,
                               double winf, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(theta_update_kernel, dim3(lin_blocks(m)), dim3(256), 0, s,
                     theta, theta0, res, c, etamu, m, eta_mu, adjust, minv, maxv, winf);
  return hipGetLastError();
}

}  // namespace culamp

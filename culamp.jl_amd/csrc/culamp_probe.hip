// Compute-rate probes of the device (no reference counterpart): the
// throughput of v_exp_f32 (the one transcendental per pair of every pass)
// and of the non-packed FP32 fused multiply-add, measured with HIP events on
// the context's stream.  They give the compute side of a pass's roofline:
// pairs per second against exponentials per second.
#include "culamp_host.h"

#include <algorithm>

namespace {

constexpr int kProbeThreads = 256;
constexpr int kProbeBlocks = 8192;
constexpr int kProbeChains = 8;  // independent chains per thread

// x <- 2^(-x) stays in (0.5, 1]; one v_exp_f32 (with a negated source) per step.
__global__ __launch_bounds__(kProbeThreads) void exp_probe(float* out, int iters) {
  float x[kProbeChains];
#pragma unroll
  for (int k = 0; k < kProbeChains; ++k) x[k] = 0.5f + 0.01f * k + 1e-6f * threadIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) x[k] = __builtin_amdgcn_exp2f(-x[k]);
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < kProbeChains; ++k) s += x[k];
  out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

// x <- 0.999 x + 0.001 converges to 1; one v_fma_f32 per step.
__global__ __launch_bounds__(kProbeThreads) void fma_probe(float* out, int iters, float a, float b) {
  float x[kProbeChains];
#pragma unroll
  for (int k = 0; k < kProbeChains; ++k) x[k] = 0.5f + 0.01f * k + 1e-6f * threadIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) x[k] = fmaf(x[k], a, b);
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < kProbeChains; ++k) s += x[k];
  out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

}  // namespace

using namespace culamp_host;

extern "C" {

int culamp_probe_rates(culamp_ctx* ctx, double* exp_per_s, double* fma_per_s) {
  TRY(check_ctx(ctx));
  if (!exp_per_s || !fma_per_s) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  DevBufs B;
  double* buf;
  TRY(B.alloc(&buf, (int64_t)kProbeBlocks * kProbeThreads / 2));
  float* out = reinterpret_cast<float*>(buf);
  const int iters = 2048;
  const double lanes = (double)kProbeBlocks * kProbeThreads * kProbeChains * iters;
  double best_exp = 0.0, best_fma = 0.0;
  for (int rep = 0; rep < 3; ++rep) {
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(exp_probe, dim3(kProbeBlocks), dim3(kProbeThreads), 0, ctx->stream, out, iters);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms > 0.f) best_exp = std::max(best_exp, lanes / (ms * 1e-3));
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(fma_probe, dim3(kProbeBlocks), dim3(kProbeThreads), 0, ctx->stream, out, iters,
                       0.999f, 0.001f);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms > 0.f) best_fma = std::max(best_fma, lanes / (ms * 1e-3));
  }
  *exp_per_s = best_exp;
  *fma_per_s = best_fma;
  return CULAMP_OK;
}

}  // extern "C"

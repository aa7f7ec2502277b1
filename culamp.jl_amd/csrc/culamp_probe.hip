// Compute-rate probes of the device (no reference counterpart): the
// throughput of v_exp_f32 (the one transcendental per pair of every pass),
// of the non-packed FP32 fused multiply-add and of the bf16 MFMA, measured
// with HIP events on the context's stream.  They give the compute side of a
// pass's roofline: pairs per second against exponentials per second, and the
// matrix-core rate the device sustains against its published peak.
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

// bf16 MFMA throughput: every wave runs kMfmaChains independent chains of
// v_mfma_f32_32x32x16_bf16 (32 x 32 x 16 x 2 = 32768 flops each), three
// waves per SIMD as in the pair pass.  The operands stay in registers.
typedef __bf16 probe_bf16x8 __attribute__((ext_vector_type(8)));
typedef float probe_f32x16 __attribute__((ext_vector_type(16)));
constexpr int kMfmaChains = 4;

__global__ __launch_bounds__(kProbeThreads) __attribute__((amdgpu_waves_per_eu(3))) void mfma_probe(float* out,
                                                                                                     int iters) {
  const int lane = threadIdx.x & 63;
  probe_bf16x8 a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    a[e] = (__bf16)(0.01f * (float)((lane + e) & 7));
    b[e] = (__bf16)(-0.01f * (float)((lane + 3 * e) & 7));
  }
  probe_f32x16 acc[kMfmaChains];
#pragma unroll
  for (int k = 0; k < kMfmaChains; ++k)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[k][r] = 0.001f * (float)(k + r);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kMfmaChains; ++k) acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[k], 0, 0, 0);
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < kMfmaChains; ++k)
#pragma unroll
    for (int r = 0; r < 16; ++r) s += acc[k][r];
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

int culamp_probe_mfma(culamp_ctx* ctx, double* bf16_flops_per_s) {
  TRY(check_ctx(ctx));
  if (!bf16_flops_per_s) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  int dev = 0, ncu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // three workgroups of four waves per CU: three waves on every SIMD
  const int blocks = 3 * std::max(ncu, 1);
  DevBufs B;
  double* buf;
  TRY(B.alloc(&buf, ((int64_t)blocks * kProbeThreads + 1) / 2));
  float* out = reinterpret_cast<float*>(buf);
  const int iters = 4096;
  const double flops = (double)blocks * (kProbeThreads / 64) * kMfmaChains * iters * 32768.0;
  double best = 0.0;
  for (int rep = 0; rep < 3; ++rep) {
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(mfma_probe, dim3(blocks), dim3(kProbeThreads), 0, ctx->stream, out, iters);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms > 0.f) best = std::max(best, flops / (ms * 1e-3));
  }
  *bf16_flops_per_s = best;
  return CULAMP_OK;
}

}  // extern "C"

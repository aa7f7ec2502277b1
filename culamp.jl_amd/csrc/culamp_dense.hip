// culamp_dense.hip -- passes over a stored cost matrix W: the reference's
// plain dense GPU path (warp_logsumexp_fused! and residual_c!,
// src/solvers/MirrorProx.jl:236-333), kept as the HBM-streaming comparator of
// the on-the-fly passes.  The same row LSE and column marginal are computed
// from n*m stored costs instead of from the points, so every pass streams
// the whole matrix from HBM.
//
// Layout: a dense problem holds W (n x m, row-major) and W^T (m x n,
// row-major), FP64 or FP32 (mixed precision), as the reference keeps W and
// Wt (MirrorProx.jl:423-424) so that both passes read rows contiguously.  A
// pass is one wave per output row: lane l reads columns l, l + 64, ...,
// kDenseUnroll of them per step (each wave load is one contiguous 256- or
// 512-byte segment), folds them into its state with the pair kernels'
// arithmetic, and the 64 lanes combine by butterfly.
#include "culamp_host.h"
#include "culamp_pair_kernel.h"

#include <algorithm>
#include <vector>

struct culamp_dense {
  culamp_ctx* ctx = nullptr;
  int64_t n = 0, m = 0;
  int precision = 0;   // CULAMP_FP64: double W; CULAMP_MIXED: float W
  void* w = nullptr;   // n x m row-major
  void* wt = nullptr;  // m x n row-major
  void* terms = nullptr;  // per-column (exponent term, weight) of the current pass
};

namespace culamp {
namespace {

constexpr int kDenseWaves = 16;   // rows (one per wave) per workgroup
constexpr int kDenseUnroll = 8;   // columns per lane per step: 8 loads in flight per lane

// Per-column exponent term and weight of a pass, formed once per pass (the
// pair kernels' source_terms: FP32 head and folded remainder in mixed
// precision) so that the row kernel's work per element is one fma, one
// exponential and one accumulation.
template <int PREC>
__global__ void dense_terms(const PassDesc d, void* terms) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.ns) return;
  R a, w;
  source_terms_v<PREC, true, R>(d, d.srcA != nullptr ? d.srcA[c] : 0.0,
                                d.srcB != nullptr ? d.srcB[c] : 1.0, &a, &w);
  ((R*)terms)[2 * c] = a;
  ((R*)terms)[2 * c + 1] = w;
}

// v_rc = k * M_rc + alpha * srcA[c] + gamma * tgtA[r]; KIND_LSE: out_r =
// LSE_c v_rc; KIND_SUM: out_r = sum_c srcB[c] exp(v_rc).  M is rows x cols
// row-major; the PassDesc carries k, gamma/tgtA (as target_offset reads it),
// nt = rows, ns = cols, out0 and epi; the column terms come from dense_terms.
// A workgroup of kDenseWaves waves owns kDenseWaves consecutive rows (one per
// wave); the column terms of each chunk of kDenseChunk columns are staged in
// LDS once for all of them, and each wave's matrix loads for the next chunk
// are issued before the current chunk is folded.
template <int PREC, int KIND>
__global__ __launch_bounds__(64 * kDenseWaves) void dense_rows(const PassDesc d, const void* mat,
                                                               const void* terms) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  constexpr int U = kDenseUnroll, CH = 64 * U;
  __shared__ R s_a[CH], s_w[CH];
  const R* T = (const R*)terms;
  const R* M = (const R*)mat;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row = (int64_t)blockIdx.x * kDenseWaves + wave;
  const bool live = row < d.nt;
  const int64_t rw = live ? row : d.nt - 1;
  const int64_t ncols = d.ns;
  const R k = PREC == 0 ? (R)d.k : (R)(d.k * kLog2e);
  TState<R> st;
  R tq;
  {
    const double off = target_offset<PREC, KIND>(d, rw);
    if (PREC == 1) {
      const float q2 = (float)(off * kLog2e);
      tq = (R)q2;
      st.qd = (double)q2;
    } else {
      tq = (R)off;
      st.qd = off;
    }
    st.m = empty_max<PREC, R>();
    st.S = st.Sc = st.Sy0 = st.Sy1 = st.Sy2 = 0.0;
    st.acc = (R)0;
  }
  const R* Mrow = M + rw * ncols;
  R mv[U];
  auto fetch = [&](const int64_t c0) {
#pragma unroll
    for (int u = 0; u < U; ++u) mv[u] = Mrow[min(c0 + 64 * u + lane, ncols - 1)];
  };
  fetch(0);
  for (int64_t c0 = 0; c0 < ncols; c0 += CH) {
    __syncthreads();
    for (int i = threadIdx.x; i < CH; i += 64 * kDenseWaves) {
      const int64_t c = c0 + i;
      const bool ok = c < ncols;
      s_a[i] = ok ? T[2 * c] : (R)-INFINITY;
      s_w[i] = ok ? T[2 * c + 1] : (R)0;
    }
    __syncthreads();
    R cur[U];
#pragma unroll
    for (int u = 0; u < U; ++u) cur[u] = mv[u];
    if (c0 + CH < ncols) fetch(c0 + CH);
    R v[U], wv[U], dz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      v[u] = fma(k, cur[u], s_a[64 * u + lane] + tq);
      wv[u] = s_w[64 * u + lane];
      dz[u] = (R)0;
    }
    // (the running max moves every step, not lazily: a lane's columns of one
    // row span the row's whole exponent range)
    accumulate<PREC, KIND, U, R, false>(st, v, wv, dz, dz, dz, dz);
  }
  combine_lanes<PREC, KIND, 64, R>(st);
  if (lane == 0 && live) epilogue_one<PREC, KIND, R>(d, row, st);
}

template <typename T>
__global__ void transpose_kernel(const T* __restrict__ a, T* __restrict__ at, int64_t rows,
                                 int64_t cols) {
  __shared__ T tile[32][33];
  const int64_t bx = (int64_t)blockIdx.x * 32, by = (int64_t)blockIdx.y * 32;
  for (int i = threadIdx.y; i < 32; i += blockDim.y) {
    const int64_t r = by + i, c = bx + threadIdx.x;
    if (r < rows && c < cols) tile[i][threadIdx.x] = a[r * cols + c];
  }
  __syncthreads();
  for (int i = threadIdx.y; i < 32; i += blockDim.y) {
    const int64_t c = bx + i, r = by + threadIdx.x;
    if (r < rows && c < cols) at[c * rows + r] = tile[threadIdx.x][i];
  }
}

__global__ void narrow_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

}  // namespace
}  // namespace culamp

using namespace culamp;
using namespace culamp_host;

namespace {

int run_dense(culamp_ctx* ctx, const culamp_dense* dm, bool transposed, const PassDesc& d,
              int kind) {
  const unsigned blocks = (unsigned)((d.nt + kDenseWaves - 1) / kDenseWaves);
  const void* mat = transposed ? dm->wt : dm->w;
  const unsigned tblocks = (unsigned)((d.ns + 255) / 256);
  void* T = dm->terms;
  if (dm->precision == CULAMP_FP64) {
    hipLaunchKernelGGL(dense_terms<0>, dim3(tblocks), dim3(256), 0, ctx->stream, d, T);
    if (kind == KIND_LSE) hipLaunchKernelGGL((dense_rows<0, KIND_LSE>), dim3(blocks), dim3(64 * kDenseWaves), 0, ctx->stream, d, mat, T);
    else hipLaunchKernelGGL((dense_rows<0, KIND_SUM>), dim3(blocks), dim3(64 * kDenseWaves), 0, ctx->stream, d, mat, T);
  } else {
    hipLaunchKernelGGL(dense_terms<1>, dim3(tblocks), dim3(256), 0, ctx->stream, d, T);
    if (kind == KIND_LSE) hipLaunchKernelGGL((dense_rows<1, KIND_LSE>), dim3(blocks), dim3(64 * kDenseWaves), 0, ctx->stream, d, mat, T);
    else hipLaunchKernelGGL((dense_rows<1, KIND_SUM>), dim3(blocks), dim3(64 * kDenseWaves), 0, ctx->stream, d, mat, T);
  }
  HIP_TRY(hipGetLastError());
  ctx->pairs += (double)d.nt * (double)d.ns;
  ctx->passes += 1;
  return CULAMP_OK;
}

int check_dense(culamp_ctx* ctx, const culamp_dense* dm) {
  TRY(check_ctx(ctx));
  if (!dm || dm->ctx != ctx) return fail(CULAMP_ERR_INVALID_ARG, "null dense problem or one of another context");
  return CULAMP_OK;
}

}  // namespace

extern "C" {

int culamp_dense_create(culamp_ctx* ctx, const double* w, int64_t n, int64_t m, int precision,
                        culamp_dense** out) {
  TRY(check_ctx(ctx));
  if (!out) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  if (!w || n <= 0 || m <= 0) return fail(CULAMP_ERR_INVALID_ARG, "dense matrix needs n, m > 0 and data");
  if (precision != CULAMP_FP64 && precision != CULAMP_MIXED)
    return fail(CULAMP_ERR_INVALID_ARG, "dense precision must be CULAMP_FP64 or CULAMP_MIXED");
  culamp_dense* dm = new culamp_dense();
  dm->ctx = ctx; dm->n = n; dm->m = m; dm->precision = precision;
  const size_t count = (size_t)n * (size_t)m;
  const size_t esz = precision == CULAMP_FP64 ? sizeof(double) : sizeof(float);
  double* stage = nullptr;
  hipError_t e = hipMalloc(&dm->w, count * esz);
  if (e == hipSuccess) e = hipMalloc(&dm->wt, count * esz);
  if (e == hipSuccess) e = hipMalloc(&dm->terms, 2 * esz * (size_t)std::max(n, m));
  if (e == hipSuccess) e = hipMalloc(&stage, count * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(stage, w, count * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    const dim3 tb(32, 8), tg((unsigned)((m + 31) / 32), (unsigned)((n + 31) / 32));
    if (precision == CULAMP_FP64) {
      e = hipMemcpyAsync(dm->w, stage, count * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
      if (e == hipSuccess) hipLaunchKernelGGL(transpose_kernel<double>, tg, tb, 0, ctx->stream, stage, (double*)dm->wt, n, m);
    } else {
      hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream, stage, (float*)dm->w, (int64_t)count);
      hipLaunchKernelGGL(transpose_kernel<float>, tg, tb, 0, ctx->stream, (const float*)dm->w, (float*)dm->wt, n, m);
    }
    if (e == hipSuccess) e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (stage) (void)hipFree(stage);
  if (e != hipSuccess) {
    culamp_dense_destroy(dm);
    return fail(CULAMP_ERR_HIP, "dense problem upload failed: %s", hipGetErrorString(e));
  }
  ctx->denses.insert(dm);
  *out = dm;
  return CULAMP_OK;
}

int culamp_dense_destroy(culamp_dense* dm) {
  if (!dm) return CULAMP_OK;
  dm->ctx->denses.erase(dm);
  (void)hipSetDevice(dm->ctx->device);
  (void)hipStreamSynchronize(dm->ctx->stream);
  if (dm->w) (void)hipFree(dm->w);
  if (dm->wt) (void)hipFree(dm->wt);
  if (dm->terms) (void)hipFree(dm->terms);
  delete dm;
  return CULAMP_OK;
}

// warp_logsumexp_fused! (src/solvers/MirrorProx.jl:236-296):
// out_i = LSE_j(-(st W_ij / (2 W_inf) + theta_j) / reg)
int culamp_dense_lse_rows(culamp_ctx* ctx, const culamp_dense* dm, const double* theta,
                          double reg, double st, double winf, double* out) {
  TRY(check_dense(ctx, dm));
  if (!theta || !out) return fail(CULAMP_ERR_INVALID_ARG, "null device vector");
  PassDesc d = blank();
  d.nt = dm->n; d.ns = dm->m;
  d.k = -(st / (2.0 * winf)) / reg;
  d.srcA = theta; d.alpha = -1.0 / reg;
  d.epi = EPI_LSE; d.out0 = out;
  return run_dense(ctx, dm, false, d, KIND_LSE);
}

// residual_c! (src/solvers/MirrorProx.jl:299-333):
// out_j = sum_i r_i exp(-(st W_ij / (2 W_inf) + theta_j) / reg - logZ_i)
int culamp_dense_col_marginal(culamp_ctx* ctx, const culamp_dense* dm, const double* r,
                              const double* theta, const double* logZ, double reg, double st,
                              double winf, double* out) {
  TRY(check_dense(ctx, dm));
  if (!r || !theta || !logZ || !out) return fail(CULAMP_ERR_INVALID_ARG, "null device vector");
  PassDesc d = blank();
  d.nt = dm->m; d.ns = dm->n;
  d.k = -(st / (2.0 * winf)) / reg;
  d.srcA = logZ; d.alpha = -1.0;
  d.srcB = r;
  d.tgtA = theta; d.gamma = -1.0 / reg;
  d.epi = EPI_SUM; d.out0 = out;
  return run_dense(ctx, dm, true, d, KIND_SUM);
}

}  // extern "C"

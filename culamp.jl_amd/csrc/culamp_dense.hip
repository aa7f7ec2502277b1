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
// pass is one wave per output row: the lanes read the row in contiguous
// 16-byte pieces (see dense_rows), fold them into their states with the pair
// kernels' arithmetic, and the 64 lanes combine by butterfly.
#include "culamp_host.h"
#include "culamp_pair_kernel.h"

#include <algorithm>
#include <cmath>
#include <vector>

struct culamp_dense {
  culamp_ctx* ctx = nullptr;
  int64_t n = 0, m = 0;
  int precision = 0;   // CULAMP_FP64: double W; CULAMP_MIXED: float W
  void* w = nullptr;   // n x m row-major
  void* wt = nullptr;  // m x n row-major
  void* terms = nullptr;  // per-column (exponent term, weight) of a pass, when formed apart
  double winf = 0.0;      // max |W_ij| (norm(W, Inf) of the reference)
};

namespace culamp {
namespace {

constexpr int kDenseWaves = 16;   // rows (one per wave) per workgroup

// Per-column exponent term and weight of a pass (the pair kernels' source
// terms: FP32 head and folded remainder in mixed precision), formed once per
// pass when many row workgroups would otherwise each form all of them.
template <int PREC, bool EXPKIND>
__global__ void dense_terms(const PassDesc d, void* terms) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.ns) return;
  R a, w;
  source_terms_v<PREC, EXPKIND, R>(d, d.srcA != nullptr ? d.srcA[c] : 0.0,
                                   d.srcB != nullptr ? d.srcB[c] : 1.0, &a, &w);
  ((R*)terms)[2 * c] = a;
  ((R*)terms)[2 * c + 1] = w;
}

// v_rc = k * M_rc + alpha * srcA[c] + gamma * tgtA[r]; KIND_LSE: out_r =
// LSE_c v_rc; KIND_SUM: out_r = sum_c srcB[c] exp(v_rc) (and, with out1,
// out1_r = sum_c srcB[c] exp(v_rc) M_rc); KIND_MIN: out_r = min_c v_rc with
// gamma = 0 (the min-plus of the eta = 0 dual).  M is rows x cols
// row-major; the PassDesc carries k, gamma/tgtA (as target_offset reads it),
// nt = rows, ns = cols, out0 and epi.
// A workgroup of kDenseWaves waves owns kDenseWaves consecutive rows (one per
// wave).  Each lane loads E columns per step, VW per load (16-byte loads when
// the row length allows it), so a wave's step is one contiguous span of
// 64 E elements, and the loads for the next chunk are issued before the
// current chunk is folded.  The launch sizes E to the waves a CU holds: with
// at most one workgroup per CU a step is 8 KB (FP32, E = 32) or 8 KB (FP64,
// E = 16) per wave, with more workgroups 2 KB or 4 KB (E = 8), about 64 to
// 128 KB in flight per CU either way (measured on 4096 x 4096 and 16384 x
// 16384 FP32 matrices: 17.2 against 20.0 us and 198 against 236 us; 16 KB
// steps at 4096 x 4096 took 22.6 us).  The column
// terms of each chunk are staged in LDS once for the workgroup's kDenseWaves
// rows: computed by the workgroup itself (FUSED: the pass is one launch; used
// when there are few row workgroups) or read from dense_terms' output.  The
// fold adds 8 terms at a time, as the pair kernels' sub-tiles do.
template <int PREC, int KIND, int VW, int E, bool FUSED>
__global__ __launch_bounds__(64 * kDenseWaves) void dense_rows(const PassDesc d, const void* mat,
                                                               const void* terms) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  typedef R RV __attribute__((ext_vector_type(VW)));
  constexpr bool kExpKind = KIND == KIND_LSE || KIND == KIND_SUM;
  constexpr int U = E / VW, CH = 64 * U * VW;
  constexpr int G = 8;  // terms per fold
  static_assert((U * VW) % G == 0, "whole folds per step");
  __shared__ __align__(16) R s_a[CH];
  __shared__ __align__(16) R s_w[CH];
  const R* M = (const R*)mat;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row = (int64_t)blockIdx.x * kDenseWaves + wave;
  const bool live = row < d.nt;
  const int64_t rw = live ? row : d.nt - 1;
  const int64_t ncols = d.ns;
  const int64_t nvec = ncols / VW;  // (the launch picks VW dividing ncols)
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
    st.acc = KIND == KIND_MIN ? (R)INFINITY : (R)0;
  }
  const bool needc = KIND == KIND_SUM && d.out1 != nullptr;
  const RV* Mrow = (const RV*)(M + rw * ncols);
  RV mv[U];
  auto fetch = [&](const int64_t v0) {
#pragma unroll
    for (int u = 0; u < U; ++u) mv[u] = Mrow[min(v0 + 64 * u + lane, nvec - 1)];
  };
  fetch(0);
  for (int64_t c0 = 0; c0 < ncols; c0 += CH) {
    flush_blocks<PREC, KIND, R>(st);  // the previous chunk's block sums
    __syncthreads();
    for (int i = threadIdx.x; i < CH; i += 64 * kDenseWaves) {
      const int64_t c = c0 + i;
      R a = (R)-INFINITY, w = (R)0;
      if (c < ncols) {
        if constexpr (FUSED) {
          source_terms_v<PREC, kExpKind, R>(d, d.srcA != nullptr ? d.srcA[c] : 0.0,
                                            d.srcB != nullptr ? d.srcB[c] : 1.0, &a, &w);
        } else {
          a = ((const R*)terms)[2 * c];
          w = ((const R*)terms)[2 * c + 1];
        }
      }
      s_a[i] = a;
      s_w[i] = w;
    }
    __syncthreads();
    RV cur[U];
#pragma unroll
    for (int u = 0; u < U; ++u) cur[u] = mv[u];
    if (c0 + CH < ncols) fetch((c0 + CH) / VW);
    // element e of load u is column c0 + (64 u + lane) VW + e
#pragma unroll
    for (int g0 = 0; g0 < U * VW; g0 += G) {
      R v[G], wv[G], cc[G], dz[G];
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const int u = (g0 + q) / VW, e = (g0 + q) % VW;
        const int idx = (64 * u + lane) * VW + e;
        cc[q] = cur[u][e];
        v[q] = fma(k, cc[q], s_a[idx] + tq);
        wv[q] = s_w[idx];
        dz[q] = (R)0;
        if constexpr (KIND == KIND_MIN)
          if (c0 + idx < ncols) st.acc = st.acc < v[q] ? st.acc : v[q];
      }
      if constexpr (KIND != KIND_MIN) {
        if (needc) {
          accumulate<PREC, KIND, G, R, true, true>(st, v, wv, cc, dz, dz, dz);
        } else {
          // (the running max moves every fold, not lazily: a lane's columns
          // of one row span the row's whole exponent range)
          accumulate<PREC, KIND, G, R, false, true>(st, v, wv, dz, dz, dz, dz);
        }
      }
    }
  }
  combine_lanes<PREC, KIND, 64, R>(st);  // (flushes the last chunk's block sums)
  if (lane == 0 && live) epilogue_one<PREC, KIND, R>(d, row, st);
}

// out_r = sum_c M_rc x_c (FP64 accumulation), one wave per row.
template <int PREC>
__global__ __launch_bounds__(kThreads) void dense_gemv(const void* mat, const double* x,
                                                       double* out, int64_t rows, int64_t cols) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  const R* M = (const R*)mat;
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (r >= rows) return;
  double acc = 0.0;
  for (int64_t c = lane; c < cols; c += 64) acc = fma((double)M[r * cols + c], x[c], acc);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) out[r] = acc;
}

// The plan of a dense LAMP solve: out_rc = r_r exp(k M_rc + alpha x_c - L_r)
// in FP64 (the reference's r .* softmax(...), MirrorProx.jl:501).
template <int PREC>
__global__ void dense_plan(const void* mat, const double* x, const double* L, const double* r,
                           double k, double alpha, double* out, int64_t rows, int64_t cols) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  const R* M = (const R*)mat;
  const int64_t total = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / cols, col = i - row * cols;
    out[i] = r[row] * exp(k * (double)M[i] + alpha * x[col] - L[row]);
  }
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
  const dim3 g(blocks), b(64 * kDenseWaves);
  // 16-byte loads when every row starts 16-byte aligned (the row length is a
  // multiple of the vector width; the allocation itself is 256-byte aligned)
  // Terms formed by each row workgroup while there are at most about two
  // workgroups per CU (one launch per pass); beyond that the workgroups
  // would repeat the term work and its reads many times over, so a separate
  // launch forms them once.
  const int64_t ncu = ctx->num_cu > 0 ? ctx->num_cu : 256;
  const bool fused = (int64_t)blocks <= 2 * ncu;
  const bool few = (int64_t)blocks <= ncu;
  void* T = dm->terms;
  auto go = [&](auto prec_c, auto vw_c) {
    constexpr int P = decltype(prec_c)::value, V = decltype(vw_c)::value;
    if (!fused) {
      const unsigned tblocks = (unsigned)((d.ns + 255) / 256);
      if (kind == KIND_MIN) hipLaunchKernelGGL((dense_terms<P, false>), dim3(tblocks), dim3(256), 0, ctx->stream, d, T);
      else hipLaunchKernelGGL((dense_terms<P, true>), dim3(tblocks), dim3(256), 0, ctx->stream, d, T);
    }
    auto rows = [&](auto kind_c, auto fused_c) {
      constexpr int K = decltype(kind_c)::value;
      constexpr bool F = decltype(fused_c)::value;
      constexpr int EW = V == 1 ? 8 : (P == 0 ? 16 : 32);  // E at most one workgroup per CU
      if (few) hipLaunchKernelGGL((dense_rows<P, K, V, EW, F>), g, b, 0, ctx->stream, d, mat, T);
      else hipLaunchKernelGGL((dense_rows<P, K, V, 8, F>), g, b, 0, ctx->stream, d, mat, T);
    };
    auto by_kind = [&](auto fused_c) {
      if (kind == KIND_LSE) rows(std::integral_constant<int, KIND_LSE>{}, fused_c);
      else if (kind == KIND_SUM) rows(std::integral_constant<int, KIND_SUM>{}, fused_c);
      else rows(std::integral_constant<int, KIND_MIN>{}, fused_c);
    };
    if (fused) by_kind(std::true_type{});
    else by_kind(std::false_type{});
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  if (dm->precision == CULAMP_FP64) {
    if (d.ns % 2 == 0) go(I0{}, std::integral_constant<int, 2>{});
    else go(I0{}, I1{});
  } else {
    if (d.ns % 4 == 0) go(I1{}, std::integral_constant<int, 4>{});
    else go(I1{}, I1{});
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
  for (size_t i = 0; i < count; ++i) dm->winf = std::max(dm->winf, std::fabs(w[i]));
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
  if (dm->terms) (void)hipFree(dm->terms);
  if (dm->w) (void)hipFree(dm->w);
  if (dm->wt) (void)hipFree(dm->wt);
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

// LAMP on a stored cost matrix (src/solvers/MirrorProx.jl:397-503): the
// dense GPU solver of the reference's `run` command.  W is n x m (rows: r,
// columns: c), W_inf = max |W_ij|, eta_p' = eta_p / (2 W_inf), s = 1; every
// `frequency` iterations a check over the plan P = r .* softmax(-(W / 2W_inf
// + nu') / eta_t) (rows) logs infeas = |c - colsum P|_1, the cost of the
// rounded plan, primalv and dualv, and stops when infeas and primal - dual are
// below epsilon / 2.  Returns theta and, optionally, the final plan.
int culamp_dense_lamp_solve(culamp_ctx* ctx, const culamp_dense* dm, const double* r,
                            const double* c, const culamp_args* a, int64_t frequency,
                            const double* theta0, culamp_log_fn log_fn, void* user,
                            double* theta_out, double* plan_out, culamp_result* out) {
  TRY(check_dense(ctx, dm));
  if (!r || !c || !a || !theta_out) return fail(CULAMP_ERR_INVALID_ARG, "null argument");
  if (frequency <= 0) return fail(CULAMP_ERR_INVALID_ARG, "frequency must be positive");
  const int64_t n = dm->n, m = dm->m;
  const double W = dm->winf;
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  DevBufs B;
  double *dr, *dc, *theta, *thetab, *nu, *nub, *res, *etamu, *L, *colsum, *colcost, *bvec,
      *rowb, *rnu, *u, *lse2, *tmp;
  for (double** v : {&dr, &L, &rowb, &u, &lse2}) TRY(B.alloc(v, n));
  for (double** v : {&dc, &theta, &thetab, &nu, &nub, &res, &etamu, &colsum, &colcost, &bvec, &rnu})
    TRY(B.alloc(v, m));
  TRY(B.alloc(&tmp, std::max(n, m)));
  HIP_TRY(hipMemcpyAsync(dr, r, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(dc, c, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
  if (theta0) HIP_TRY(hipMemcpyAsync(theta, theta0, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
  else HIP_TRY(launch_fill(theta, 0.0, m, ctx->stream));
  for (double* v : {nu, nub, thetab}) HIP_TRY(hipMemcpyAsync(v, theta, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
  const double etap = a->eta_p / 2.0 / W;                                      // 409
  HIP_TRY(launch_etamu(etamu, dc, a->alpha / (double)m, a->tau_mu, m, ctx->stream));  // 417
  double rlogr = 0.0;
  for (int64_t i = 0; i < n; ++i) rlogr += r[i] > 0.0 ? r[i] * std::log(r[i]) : 0.0;
  const double taup = a->tau_p;
  const double minv = std::tanh(-a->B / 2.0), maxv = std::tanh(a->B / 2.0);
  const double t0 = now_s();

  // row LSE of -(s W / (2 W_inf) + x_j) / reg, and the column pass over it
  auto row_lse = [&](const double* x, double reg, double* dst) {
    PassDesc d = blank();
    d.nt = n; d.ns = m;
    d.k = -(1.0 / (2.0 * W)) / reg;
    d.srcA = x; d.alpha = -1.0 / reg;
    d.epi = EPI_LSE; d.out0 = dst;
    return run_dense(ctx, dm, false, d, KIND_LSE);
  };
  auto col_pass = [&](const double* x, const double* logZ, const double* wr, double reg,
                      double* dst, double* dcost) {
    PassDesc d = blank();
    d.nt = m; d.ns = n;
    d.k = -(1.0 / (2.0 * W)) / reg;
    d.srcA = logZ; d.alpha = -1.0;
    d.srcB = wr;
    d.tgtA = x; d.gamma = -1.0 / reg;
    d.epi = EPI_SUM; d.out0 = dst; d.out1 = dcost;
    return run_dense(ctx, dm, true, d, KIND_SUM);
  };
  auto infeas = [&](const double* x, double reg) {  // 425-428
    TRY(row_lse(x, reg, L));
    return col_pass(x, L, dr, reg, res, nullptr);
  };

  double etat = INFINITY;
  double primal = NAN, dual = NAN, feas = NAN, objective = NAN;
  int converged = 0;
  int64_t steps = 0;
  if (a->verbose) emit(log_fn, user, "time(s),iter,infeas,ot_objective,primal,dual,pdgap,solver\n");
  TRY(infeas(nu, etat));                                                       // 436
  for (int64_t it = 1; it <= a->itermax; ++it) {
    const double elapsed = now_s() - t0;
    if (elapsed > a->tmax) break;
    if ((it - 1) % frequency == 0) {                                           // 444-461
      TRY(row_lse(nu, etat, L));
      TRY(col_pass(nu, L, dr, etat, colsum, colcost));
      ReduceSpec rs{};
      rs.nops = 4;
      rs.op[0] = 1; rs.a[0] = dc; rs.b[0] = colsum; rs.len[0] = m;   // |c - colsum|_1
      rs.op[1] = 0; rs.a[1] = colcost; rs.len[1] = m;                // <W, r .* p>
      rs.op[2] = 2; rs.a[2] = nu; rs.b[2] = colsum; rs.len[2] = m;   // <nu, colsum>
      rs.op[3] = 2; rs.a[3] = dr; rs.b[3] = L; rs.len[3] = n;        // <r, L>
      double v[4];
      TRY(reduce_to_host(ctx, rs, v));
      feas = v[0];
      const double wrp = v[1];
      // sum_i r_i sum_j p_ij log p_ij with log p_ij = v_ij - L_i and rows of p
      // summing to one
      const double rnegent = -(wrp / (2.0 * W) + v[2]) / etat - v[3];
      primal = 2.0 * (0.5 * wrp + W * feas + etap * rnegent + etap * rlogr);  // primalv, 28-32
      // dualv (7-13)
      if (etap == 0.0) {
        PassDesc d = blank();
        d.nt = n; d.ns = m;
        d.k = 1.0;
        d.srcA = theta; d.alpha = 2.0 * W;
        d.epi = EPI_MIN; d.out0 = lse2;
        TRY(run_dense(ctx, dm, false, d, KIND_MIN));
      } else {
        TRY(row_lse(theta, etap / W, lse2));
      }
      ReduceSpec rd{};
      rd.nops = 2;
      rd.op[0] = 2; rd.a[0] = dc; rd.b[0] = theta; rd.len[0] = m;
      rd.op[1] = 2; rd.a[1] = dr; rd.b[1] = lse2; rd.len[1] = n;
      double w2[2];
      TRY(reduce_to_host(ctx, rd, w2));
      if (etap == 0.0) dual = -2.0 * W * w2[0] + w2[1];
      else dual = 2.0 * ((-W * w2[0] - etap * w2[1]) + etap * rlogr);
      // cost of round(r .* p, r, c) (Utilities.jl:158-165): the rows of r .* p
      // already sum to r (row scaling 1), b_j = min(c_j / colsum_j, 1),
      // rmu = r - rowsum(P b), rnu = c - b .* colsum, and
      // <W, round> = <b, colcost> + rmu' W rnu / |rmu|_1
      {
        std::vector<double> hs(m), hb(m), hrn(m);
        TRY(copy_out(ctx, hs.data(), colsum, m));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (int64_t j = 0; j < m; ++j) {
          hb[j] = std::min(c[j] / hs[j], 1.0);
          hrn[j] = c[j] - hb[j] * hs[j];
        }
        HIP_TRY(hipMemcpyAsync(bvec, hb.data(), sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(rnu, hrn.data(), sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
        // rowsum(P b)_i = r_i sum_j p_ij b_j
        PassDesc d = blank();
        d.nt = n; d.ns = m;
        d.k = -(1.0 / (2.0 * W)) / etat;
        d.srcA = nu; d.alpha = -1.0 / etat;
        d.srcB = bvec;
        d.tgtA = L; d.gamma = -1.0;
        d.epi = EPI_SUM; d.out0 = rowb;
        TRY(run_dense(ctx, dm, false, d, KIND_SUM));
        const unsigned gb = (unsigned)((n + kThreads / 64 - 1) / (kThreads / 64));
        if (dm->precision == CULAMP_FP64) hipLaunchKernelGGL(dense_gemv<0>, dim3(gb), dim3(kThreads), 0, ctx->stream, dm->w, rnu, u, n, m);
        else hipLaunchKernelGGL(dense_gemv<1>, dim3(gb), dim3(kThreads), 0, ctx->stream, dm->w, rnu, u, n, m);
        HIP_TRY(hipGetLastError());
        std::vector<double> hrb(n), hu(n);
        TRY(copy_out(ctx, hrb.data(), rowb, n));
        TRY(copy_out(ctx, hu.data(), u, n));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<double> hcc(m);
        TRY(copy_out(ctx, hcc.data(), colcost, m));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        double bc = 0.0, num = 0.0, l1 = 0.0;
        for (int64_t j = 0; j < m; ++j) bc += hb[j] * hcc[j];
        for (int64_t i = 0; i < n; ++i) {
          const double rmu = r[i] - r[i] * hrb[i];
          num += rmu * hu[i];
          l1 += std::fabs(rmu);
        }
        objective = bc + num / l1;
      }
      if (a->verbose)
        emit(log_fn, user, fmt("%.6e", elapsed) + "," + std::to_string(it) + "," +
                               fmt("%.14e", feas) + "," + fmt("%.14e", objective) + "," +
                               fmt("%.14e", primal) + "," + fmt("%.14e", dual) + "," +
                               fmt("%.14e", primal - dual) + ",lamp_cuda\n");
      if (feas < a->epsilon / 2.0 && primal - dual < a->epsilon / 2.0) {
        converged = 1;
        break;
      }
    }
    HIP_TRY(launch_theta_update(thetab, theta, res, dc, etamu, m, a->tau_mu * a->eta_mu, 0, minv,
                                maxv, 1.0, ctx->stream));                       // 463
    etat = 1.0 / (taup + (1.0 / etat) * (1.0 - etap));                         // 465
    HIP_TRY(launch_axpby(nub, theta, taup * etat, nu, 1.0 - taup * etat, m, ctx->stream));  // 467
    TRY(infeas(nub, etat));                                                    // 475
    // the reference's update (476), then the clamp (484), in one pass
    HIP_TRY(launch_theta_update(theta, theta, res, dc, etamu, m, a->eta_mu, 1, minv, maxv, 1.0,
                                ctx->stream));
    HIP_TRY(launch_axpby(nu, nu, 1.0 - taup * etat, thetab, taup * etat, m, ctx->stream));  // 477
    TRY(infeas(nu, etat));                                                     // 487
    steps += 1;
  }
  TRY(copy_out(ctx, theta_out, theta, m));
  if (plan_out) {                                                              // 501
    TRY(row_lse(nu, etat, L));
    DevBufs P;
    double* dplan;
    TRY(P.alloc(&dplan, n * m));
    const double k = -(1.0 / (2.0 * W)) / etat, alpha = -1.0 / etat;
    if (dm->precision == CULAMP_FP64) hipLaunchKernelGGL(dense_plan<0>, dim3(4096), dim3(256), 0, ctx->stream, dm->w, nu, L, dr, k, alpha, dplan, n, m);
    else hipLaunchKernelGGL(dense_plan<1>, dim3(4096), dim3(256), 0, ctx->stream, dm->w, nu, L, dr, k, alpha, dplan, n, m);
    HIP_TRY(hipGetLastError());
    TRY(copy_out(ctx, plan_out, dplan, n * m));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (out) {
    out->iterations = steps;
    out->objective = objective;
    out->residual = feas;
    out->primal = primal;
    out->dual = dual;
    out->winf = W;
    out->eta_t = etat;
    out->elapsed_s = now_s() - t0;
    out->converged = converged;
    out->pairs = ctx->pairs - pairs0;
    out->passes = ctx->passes - passes0;
    out->nonfinite = !(std::isfinite(feas) && std::isfinite(objective));
    out->fp64_from = 0;
  }
  return CULAMP_OK;
}

}  // extern "C"

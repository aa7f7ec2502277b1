// culamp_internal.h -- types shared by the pair-pass kernels and the host side
// of libculamp.  Not part of the public ABI (see include/culamp.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace culamp {

// Cost kinds of the reference kernels' `l2dist`
// (src/solvers/MirrorProxKernel.jl:869-878).
enum Cost : int { COST_L1 = 0, COST_SQL2 = 1, COST_LINF = 2, COST_GENP = 3 };

inline int cost_of_p(double p) {
  if (p == 1.0) return COST_L1;
  if (p == 2.0) return COST_SQL2;
  if (p == __builtin_inf()) return COST_LINF;
  return COST_GENP;
}

// What a pass reduces over the sources s for each target t, with
//   v_ts = k * C(t, s) + alpha * srcA[s] + gamma * tgtA[t]
//   KIND_LSE : log sum_s exp(v_ts)                       (max-shifted)
//   KIND_SUM : sum_s srcB[s] exp(v_ts), sum_s srcB[s] exp(v_ts) C(t,s)
//   KIND_BARY: sum_s srcB[s] exp(v_ts) pos_s / sum_s srcB[s] exp(v_ts)
//   KIND_MAX : max_s C(t, s)
//   KIND_MIN : min_s C(t, s) + alpha * srcA[s]
enum Kind : int { KIND_LSE = 0, KIND_SUM = 1, KIND_BARY = 2, KIND_MAX = 3, KIND_MIN = 4 };

// Feature metrics of src/CellUtilities.jl:5 ("pearson" = correlation).
enum Metric : int { METRIC_L1 = 0, METRIC_L2 = 1, METRIC_COSINE = 2, METRIC_CORRELATION = 3 };

// Per-target epilogue applied once the split partials are merged.
enum Epi : int {
  EPI_LSE = 0,       // out0[t] = LSE
  EPI_SK_HALF = 1,   // out0[t] = log(marg[t]) - LSE
  EPI_SUM = 2,       // out0[t] = S, out1[t] = Sc
  EPI_SK_RESID = 3,  // out0[t] = marg[t] - S, out1[t] = Sc
  EPI_LAMP_COL = 4,  // out0[t] = S, out1[t] = Sc, then the fused theta / nu update
  EPI_BARY = 5,      // out0[3t + k] = weighted mean of source coordinate k
  EPI_MAX = 6,       // out0[t] = max
  EPI_MIN = 7,       // out0[t] = min
  EPI_ADJ = 8        // KIND_BARY state, adjoint of an LSE pass (see PassDesc::adj_*)
};

struct PassDesc {
  const double* tgt;  // 3 x nt, column-major (point-major)
  const double* src;  // 3 x ns
  int64_t nt, ns;
  double p;           // exponent for COST_GENP
  double k;           // cost scale in the exponent (natural-log units)
  const double* srcA; double alpha;   // source exponent term (may be null)
  const double* srcB;                 // source multiplicative weight (may be null)
  const double* tgtA; double gamma;   // target exponent term (may be null)
  int prev_offset;    // LSE kinds, mixed precision: centre each target on the
                      // previous value of its output (see pair_pass)
  int epi;
  double* out0; double* out1;
  const double* marg;
  // EPI_LAMP_COL fused update (update_theta_residual, MirrorProx.jl:159-180,
  // and the nu mixing of MirrorProxKernel.jl:1109/1115).
  int fuse_update; int adjust;
  const double* th0; const double* cvec; const double* etamu;
  double eta_mu, minv, maxv, upd_winf;
  double* th_out;
  const double* nu_in; const double* mix_src; double* nu_out; double mix_t;
  // Feature (cell-similarity) passes: metric >= 0 selects the feature kernel.
  // Rows are column-major (row index fastest, leading dimension = row
  // count): divided by their sums for l1, raw for the dot metrics; rc =
  // [n~ | mu~ | c~ | s] per row (s^-2 * squared norm, s^-1 * mean,
  // s^-2 * centred squared norm, row sum).
  int metric;
  int64_t dim;
  double fscale;
  const double* tf64; const double* sf64;
  const float* tf32; const float* sf32;
  const double* trc; const double* src_rc;
  // MFMA feature path: unit rows in bf16 (hi, and lo = remainder), row-major
  // with width 16 * zks.
  const __bf16* tz_hi; const __bf16* tz_lo;
  const __bf16* sz_hi; const __bf16* sz_lo;
  int zks;
  // Device-resident iteration scalars for graph-captured loops: with
  // dyn_kind != DYN_NONE the kernel reads [reg, st, mix] from dyn and derives
  // k = -(st / (2 dyn_winf)) / reg and alpha = -1/reg (DYN_LSE) or
  // gamma = -1/reg and mix_t = mix (DYN_COL), as the host would.
  const double* dyn;
  int dyn_kind;
  double dyn_winf;
  // EPI_ADJ (squared-Euclidean gradients of an LSE pass), with the pass's
  // S_t = sum_s w_s e^{v_ts} and Y_t = sum_s w_s e^{v_ts} pos_s:
  //   adj_norm == 0: out1[t] = adj_beta out1[t] + adj_c1 S_t
  //                  out0[3t+k] += adj_c0 (tgt_k S_t - Y_tk)
  //   adj_norm == 1: out0[3t+k] += adj_c0 tgtW[t] (tgt_k - Y_tk / S_t)
  // (out0 / out1 may be null).
  const double* tgtW;
  double adj_c0, adj_c1, adj_beta;
  int adj_norm;
  // Staged source records (mixed-precision LSE / SUM passes of the solver
  // loops).  With `rec` set the pass takes each source's FP32 record
  // {x, y, z, a} and weight from rec / recw instead of converting src / srcA /
  // srcB itself; the records hold exactly what that conversion gives.  An
  // epilogue with rec_out set writes the records of the next pass, whose
  // sources are this pass's targets, with srcA = the value the epilogue
  // writes (out0 of an LSE pass, nu_out of EPI_LAMP_COL), exponent factor
  // rec_alpha, weights rec_wsrc (null: 1) and, when rec_scaled (a mixed
  // squared-Euclidean LSE consumer), coordinates scaled by sqrt(-rec_k log2 e)
  // where rec_k is the consumer's k.
  const float4* rec;
  const float* recw;
  float4* rec_out;
  float* recw_out;
  double rec_alpha, rec_k;
  int rec_scaled;
  const double* rec_wsrc;
};

enum DynKind : int { DYN_NONE = 0, DYN_LSE = 1, DYN_COL = 2 };

// Batched launches: one workgroup per item; instances carry their own launch
// shape and partial-buffer base.
struct BatchItem {
  int32_t inst;  // < 0: padding, the workgroup does nothing
  int32_t split;
  int64_t tb;
};
struct BatchInst {
  int64_t pbase;      // offset of the instance's partials in Work.partial
  int64_t ntp;        // padded target count of the instance
  int64_t split_len;
  int32_t splits;
  int32_t pad;
};

struct Work {
  double* partial;     // split partials, [field][split][target]
};

// One batched pass: a PassDesc and a BatchInst per instance, and the item
// tables of the pass and of its split merge.  Every instance of a launch
// uses the same kernel family and template arguments.
enum BatchFamily : int { FAMILY_POINTS = 0, FAMILY_FEATURES = 1, FAMILY_MFMA = 2 };
struct BatchLaunch {
  const PassDesc* descs;
  const BatchInst* insts;
  const BatchItem* items;
  const BatchItem* mitems;
  int64_t nitems, nmitems;
  int family;
  int rpt;     // FAMILY_POINTS: targets per thread
  int cost;    // FAMILY_POINTS: Cost
  int metric;  // FAMILY_FEATURES
  int zks;     // FAMILY_MFMA: row width / 16
  int nsplit;  // FAMILY_MFMA: 1 or 3 products
};

// Apply the dynamic scalars of a pass (see PassDesc::dyn).
__device__ __forceinline__ PassDesc resolve_desc(const PassDesc& din) {
#pragma clang fp contract(off)
  // (selects rather than a field chosen at run time, so that the copy stays
  // in registers instead of private memory)
  PassDesc d = din;
  if (d.dyn_kind != DYN_NONE) {
    const double reg = d.dyn[0], st = d.dyn[1], mix = d.dyn[2];
    const double c1 = st / (2.0 * d.dyn_winf);
    const double invreg = 1.0 / reg;
    const bool lse = d.dyn_kind == DYN_LSE;
    d.k = -c1 * invreg;
    d.alpha = lse ? -invreg : d.alpha;
    d.gamma = lse ? d.gamma : -invreg;
    d.mix_t = lse ? d.mix_t : mix;
  }
  return d;
}

struct LaunchCfg {
  int rpt;          // targets per thread
  int splits;       // source splits (grid.y; inside the workgroup when wide)
  int64_t split_len;
  int64_t tblocks;  // grid.x
  int wide;         // point passes: splits merged inside one workgroup (pair_pass_wide)
};

// Layout of a point pass's source splits: automatic, split over workgroups
// (partials in HBM, merged by a second launch), or inside one workgroup.
enum Layout : int { LAYOUT_AUTO = 0, LAYOUT_SPLIT = 1, LAYOUT_WIDE = 2 };

constexpr int kThreads = 256;
constexpr int kTile = 256;      // sources staged in LDS per step
constexpr int kLanesPerTarget = 4;  // point kernel: adjacent lanes sharing a target
constexpr int kMergeLanes = 4;      // split merge: adjacent lanes sharing a target
// Workgroup-wide split layout (pair_pass_wide): one wave per source split,
// one wave's worth of targets per workgroup, merged by one wave.
constexpr int kWideWaves = 16;
constexpr int kWideThreads = 64 * kWideWaves;
constexpr int kWideTargets = 64 / kLanesPerTarget;
constexpr int kMaxFields = 6;   // partial fields per target (BARY)

int pass_fields(int kind);
LaunchCfg choose_launch(int64_t nt, int64_t ns, int precision, int kind,
                        int num_cu, int force_rpt, int force_splits, int layout = LAYOUT_AUTO);
hipError_t launch_pass(const PassDesc& d, int precision, int kind,
                       const LaunchCfg& cfg, const Work& w, hipStream_t s);
hipError_t launch_pass_fp32(const PassDesc& d, int kind, const LaunchCfg& c,
                            const Work& w, hipStream_t s);
hipError_t launch_pass_fp64(const PassDesc& d, int kind, const LaunchCfg& c,
                            const Work& w, hipStream_t s);
// Feature passes (64 targets per workgroup; LaunchCfg.rpt is unused).
constexpr int kFeatTargets = 64;
// The bf16 unit rows of the MFMA feature path are scaled by this factor
// before rounding, so that the dot of two rounded rows (each of norm at most
// kUnitRowScale (1 + 2^-9) < 1) can never leave [-kUnitRowScale^2,
// kUnitRowScale^2] and needs no clamp to [-1, 1]; the kernel divides the
// factor back out in its exponent scale.  (1 - 2^-8)^2 is exact in FP32.
constexpr double kUnitRowScale = 1.0 - 1.0 / 256.0;
constexpr double kUnitRowScale2 = kUnitRowScale * kUnitRowScale;
LaunchCfg choose_feature_launch(int64_t nt, int64_t ns, int num_cu, int force_splits);
hipError_t launch_feature_fp32(const PassDesc& d, int kind, const LaunchCfg& c,
                               const Work& w, hipStream_t s);
hipError_t launch_feature_fp64(const PassDesc& d, int kind, const LaunchCfg& c,
                               const Work& w, hipStream_t s);
// MFMA feature passes (128 targets per workgroup): nsplit 1 (bf16) or 3
// (bf16 head + remainder).
LaunchCfg choose_mfma_launch(int64_t nt, int64_t ns, int num_cu, int force_splits);
hipError_t launch_feature_mfma(const PassDesc& d, int kind, int nsplit, const LaunchCfg& c,
                               const Work& w, hipStream_t s);

// Small device reductions used by the host loops (deterministic, one block).
// ops[k] selects: 0 sum a, 1 sum |a - b|, 2 sum a*b, 3 max a, 4 sum a log a
// over a > 0, 5 sum |a|.  Results go to out[k].
struct ReduceSpec {
  int nops;
  int op[8];
  const double* a[8];
  const double* b[8];
  int64_t len[8];
};
hipError_t launch_reduce(const ReduceSpec& spec, double* out, hipStream_t s);

// Batched passes (FAMILY_POINTS / FAMILY_FEATURES in FP32 or FP64, MFMA).
hipError_t launch_batch_fp32(const BatchLaunch& b, int kind, const Work& w, hipStream_t s);
hipError_t launch_batch_fp64(const BatchLaunch& b, int kind, const Work& w, hipStream_t s);
hipError_t launch_batch_mfma(const BatchLaunch& b, int kind, const Work& w, hipStream_t s);

// Threads of a reduction workgroup.
constexpr int kReduceThreads = 1024;

// Compensated FP64 summation (Knuth's TwoSum): s + x is rounded into s and
// its exact rounding error is added to the running compensation c, so that
// s + c carries the sum with an error independent of the number of terms
// (to first order).  Contraction is off: every operation must round as
// written.
__device__ __forceinline__ void two_sum_acc(double& s, double& c, const double x) {
#pragma clang fp contract(off)
  const double t = s + x;
  const double bp = t - s;
  c += (s - (t - bp)) + (x - bp);
  s = t;
}

// One element's contribution to reduction `op` (see ReduceSpec), folded into
// the compensated pair (s, c).  Products are split exactly with an FMA
// (x * y = p + e), so a dot product is as accurate as in twice the working
// precision.
__device__ __forceinline__ void reduce_fold(const int op, double& s, double& c, const double x,
                                            const double y) {
#pragma clang fp contract(off)
  if (op == 3) {
    s = fmax(s, x);
  } else if (op == 2) {
    const double p = x * y;
    c += fma(x, y, -p);
    two_sum_acc(s, c, p);
  } else if (op == 0) {
    two_sum_acc(s, c, x);
  } else if (op == 1) {
    two_sum_acc(s, c, fabs(x - y));
  } else if (op == 4) {
    if (x > 0.0) two_sum_acc(s, c, x * log(x));
  } else {
    two_sum_acc(s, c, fabs(x));
  }
}

// Reduction k of `spec` by one workgroup (the reductions of a spec run as
// separate workgroups, so their memory round trips overlap): each thread folds
// four strided elements per round with their loads in flight together into a
// compensated pair (sum, error), the pairs meet by butterfly within each wave
// and across the waves through LDS, each meeting adding the two sums with
// TwoSum; the result is sum + error.  A fixed order, so the result depends
// only on the input.
__device__ __forceinline__ void reduce_op(const ReduceSpec& spec, const int k, double* out) {
#pragma clang fp contract(off)
  constexpr int W = kReduceThreads / 64;
  constexpr int U = 4;
  __shared__ double sh[2][W];
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int op = spec.op[k];
  const double* a = spec.a[k];
  const double* b = spec.b[k];
  const int64_t n = spec.len[k];
  const bool two = op == 1 || op == 2;
  const bool mx = op == 3;
  double v = mx ? -INFINITY : 0.0, ve = 0.0;
  for (int64_t i0 = tid; i0 < n; i0 += U * kReduceThreads) {
    double xa[U], xb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + (int64_t)u * kReduceThreads;
      xa[u] = 0.0;
      xb[u] = 0.0;
      if (i < n) {
        xa[u] = a[i];
        if (two) xb[u] = b[i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + (int64_t)u * kReduceThreads < n) reduce_fold(op, v, ve, xa[u], xb[u]);
  }
  auto meet = [&](const double o, const double oe) {
    if (mx) {
      v = fmax(v, o);
    } else {
      ve += oe;
      two_sum_acc(v, ve, o);
    }
  };
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off);
    const double oe = __shfl_xor(ve, off);
    meet(o, oe);
  }
  if (lane == 0) {
    sh[0][wave] = v;
    sh[1][wave] = ve;
  }
  __syncthreads();
  if (wave != 0) return;
  v = lane < W ? sh[0][lane] : (mx ? -INFINITY : 0.0);
  ve = lane < W ? sh[1][lane] : 0.0;
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off);
    const double oe = __shfl_xor(ve, off);
    meet(o, oe);
  }
  // (v itself is the plain sum in the same order: an infinite or NaN term
  // leaves it as plain summation would, while its error term is NaN)
  if (lane == 0) out[k] = (mx || !isfinite(v)) ? v : v + ve;
}

// out = sa * a + sb * b with both products rounded (no fused multiply-add),
// shared by every kernel that mixes potentials so they agree bit for bit.
__device__ __forceinline__ double axpby_elem(double sa, double a, double sb, double b) {
#pragma clang fp contract(off)
  return sa * a + sb * b;
}

// Elementwise helpers.
// d1 = s1 and d2 = s2 (n doubles each; d2 may be null) and d3 = s3 (n3
// doubles) in one launch.
hipError_t launch_copy3(double* d1, const double* s1, double* d2, const double* s2, int64_t n,
                        double* d3, const double* s3, int64_t n3, hipStream_t s);
hipError_t launch_axpby(double* out, const double* a, double sa, const double* b,
                        double sb, int64_t n, hipStream_t s);  // out = sa*a + sb*b
hipError_t launch_log(double* out, const double* a, int64_t n, hipStream_t s);
hipError_t launch_fill(double* out, double v, int64_t n, hipStream_t s);
hipError_t launch_etamu(double* out, const double* c, double add, double div,
                        int64_t n, hipStream_t s);  // out = (c + add) / div
hipError_t launch_log_minus(double* out, const double* r, const double* a,
                            int64_t n, hipStream_t s);  // out = log(r) - a
hipError_t launch_theta_update(double* theta, const double* theta0,
                               const double* res, const double* c,
                               const double* etamu, int64_t m, double eta_mu,
                               int adjust, double minv, double maxv,
                               double winf, hipStream_t s);

// update_theta_residual (src/solvers/MirrorProx.jl:159-180), overflow-safe form.
__device__ __forceinline__ double theta_update(double res, double c,
                                               double etamu, double th0,
                                               double eta_mu, int adjust,
                                               double minv, double maxv,
                                               double winf) {
  double dif = 2.0 * winf * (res - c) / etamu;
  double mx = fmax(dif, -dif);
  double base = (th0 + 1.0) / (1.0 - th0);
  if (eta_mu != 0.0) {
    asm volatile("" ::: "memory");  // (keeps pow behind the branch: x^1 = x)
    base = pow(base, 1.0 - eta_mu);
  }
  double e = exp(dif - mx) * base;
  double em = exp(-mx);
  double t = (e - em) / (e + em);
  if (adjust) t = fmin(fmax(t, minv), maxv);
  return t;
}

}  // namespace culamp

// culamp_pair_kernel.h -- the templated pair-pass kernel (device code),
// included by the per-precision translation units.
//
// One templated kernel covers every O(n*m) pass of the reference's hot path
// (row/column log-sum-exp, column marginal + cost, row residual + cost,
// barycentric projections, max cost, min-plus).  Layout on the device:
//
//   * each workgroup (256 threads = 4 wave64) owns 256*RPT targets, RPT per
//     thread held in registers;
//   * the sources of one split (grid.y) stream through LDS in tiles of 256
//     points (SoA: x, y, z, exponent term a_s, weight w_s), so every LDS read
//     is a wave-wide broadcast and every HBM read of a source is one
//     coalesced load per tile;
//   * the cost is generated in registers and the max-shifted exp-sum runs on
//     sub-tiles of 8 sources per target;
//   * when a target's sources are split over several workgroups, each writes
//     its partial to HBM and the last workgroup to arrive (device-scope
//     atomic ticket, no spinning) merges the partials in split order, so
//     results are bitwise reproducible run to run;
//   * the merged result goes through a per-target epilogue (log marginal
//     minus LSE, marginal minus sum, fused LAMP dual update, ...).
//
// Mixed precision (PREC == 1): cost and exponent in FP32, in log2 units so
// the exponential is a single v_exp_f32.  The FP64 source term a_s is split
// into an FP32 head (used in the exponent) and its rounding error, which is
// folded into the FP32 source weight w_s = b_s * 2^(a_s - head); per-target
// offsets are subtracted exactly in the FP64 epilogue.  The running max is
// kept integer-valued so rescaling is an exact ldexp, and every sub-tile sum
// is flushed into an FP64 accumulator.
#pragma once
#include "culamp_internal.h"

#include <math.h>
#include <type_traits>

namespace culamp {
namespace {

constexpr double kLog2e = 1.4426950408889634074;
constexpr double kLn2 = 0.69314718055994530942;
constexpr int kSub = 8;  // sources per sub-tile

__device__ __forceinline__ float exp2_fast(float x) {
  return __builtin_amdgcn_exp2f(x);
}

template <int COST>
__device__ __forceinline__ float pair_cost(float ax, float ay, float az,
                                           float bx, float by, float bz,
                                           float p) {
  float dr = ax - bx, dg = ay - by, db = az - bz;
  if constexpr (COST == COST_L1) {
    return fabsf(dr) + fabsf(dg) + fabsf(db);
  } else if constexpr (COST == COST_SQL2) {
    return fmaf(dr, dr, fmaf(dg, dg, db * db));
  } else if constexpr (COST == COST_LINF) {
    return fmaxf(fabsf(dr), fmaxf(fabsf(dg), fabsf(db)));
  } else {
    return powf(fabsf(dr), p) + powf(fabsf(dg), p) + powf(fabsf(db), p);
  }
}

template <int COST>
__device__ __forceinline__ double pair_cost(double ax, double ay, double az,
                                            double bx, double by, double bz,
                                            double p) {
  double dr = ax - bx, dg = ay - by, db = az - bz;
  if constexpr (COST == COST_L1) {
    return fabs(dr) + fabs(dg) + fabs(db);
  } else if constexpr (COST == COST_SQL2) {
    return fma(dr, dr, fma(dg, dg, db * db));
  } else if constexpr (COST == COST_LINF) {
    return fmax(fabs(dr), fmax(fabs(dg), fabs(db)));
  } else {
    return pow(fabs(dr), p) + pow(fabs(dg), p) + pow(fabs(db), p);
  }
}

template <int KIND>
constexpr int kind_fields() {
  return KIND == KIND_BARY ? 5 : (KIND == KIND_MAX || KIND == KIND_MIN) ? 1 : 2;
}

template <int PREC, int KIND, int COST, int RPT>
__global__ __launch_bounds__(kThreads) void pair_pass(const PassDesc d,
                                                      const Work w,
                                                      const int splits,
                                                      const int64_t split_len) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  constexpr int NF = kind_fields<KIND>();
  constexpr bool kExpKind = (KIND == KIND_LSE || KIND == KIND_SUM || KIND == KIND_BARY);
  constexpr bool kMaxShift = (KIND == KIND_LSE || KIND == KIND_BARY);
  // "empty" running max: finite so that differences stay finite
  constexpr double kEmptyD = PREC == 0 ? -1e300 : -1e30;
  const R kEmpty = (R)kEmptyD;

  __shared__ R s_x[kTile], s_y[kTile], s_z[kTile], s_a[kTile], s_w[kTile];
  __shared__ int s_last;

  const int tid = threadIdx.x;
  const int64_t tb = blockIdx.x;
  const int split = blockIdx.y;
  const int64_t s0 = (int64_t)split * split_len;
  const int64_t s1 = min(d.ns, s0 + split_len);
  const int64_t ntp = (int64_t)gridDim.x * kThreads * RPT;
  const R k = PREC == 0 ? (R)d.k : (R)(d.k * kLog2e);
  const R pexp = (R)d.p;

  R tx[RPT], ty[RPT], tz[RPT], tq[RPT];
  double qd[RPT];
  int64_t tidx[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int64_t t = tb * kThreads * RPT + (int64_t)r * kThreads + tid;
    tidx[r] = t;
    const bool ok = t < d.nt;
    tx[r] = ok ? (R)d.tgt[3 * t] : (R)0;
    ty[r] = ok ? (R)d.tgt[3 * t + 1] : (R)0;
    tz[r] = ok ? (R)d.tgt[3 * t + 2] : (R)0;
    double q = 0.0;
    if (KIND == KIND_SUM && d.tgtA != nullptr && ok) q = d.gamma * d.tgtA[t];
    if (PREC == 1 && KIND == KIND_LSE && d.prev_offset && ok) {
      const double prev = d.out0[t];
      const double lse_prev = d.epi == EPI_SK_HALF ? log(d.marg[t]) - prev : prev;
      q = isfinite(lse_prev) ? -lse_prev : 0.0;
    }
    if (PREC == 1) {
      const float q2 = (float)(q * kLog2e);
      tq[r] = (R)q2;
      qd[r] = (double)q2;
    } else {
      tq[r] = (R)q;
      qd[r] = q;
    }
  }

  R m[RPT];
  double S[RPT], Sc[RPT], Sy0[RPT], Sy1[RPT], Sy2[RPT];
  R acc[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    m[r] = kEmpty;
    S[r] = Sc[r] = Sy0[r] = Sy1[r] = Sy2[r] = 0.0;
    acc[r] = KIND == KIND_MIN ? (R)INFINITY : (R)-INFINITY;
  }

  for (int64_t base = s0; base < s1; base += kTile) {
    const int cnt = (int)min((int64_t)kTile, s1 - base);
    const int cntp = (cnt + kSub - 1) / kSub * kSub;
    __syncthreads();
    if (tid < cnt) {
      const int64_t s = base + tid;
      s_x[tid] = (R)d.src[3 * s];
      s_y[tid] = (R)d.src[3 * s + 1];
      s_z[tid] = (R)d.src[3 * s + 2];
      const double av = d.srcA != nullptr ? d.alpha * d.srcA[s] : 0.0;
      const double bw = d.srcB != nullptr ? d.srcB[s] : 1.0;
      if (PREC == 1 && kExpKind) {
        const double a2 = av * kLog2e;
        const float hi = (float)a2;
        const double lo = a2 - (double)hi;
        s_a[tid] = (R)hi;
        s_w[tid] = (R)(isfinite(lo) ? bw * (1.0 + lo * kLn2) : bw);
      } else {
        s_a[tid] = (R)av;
        s_w[tid] = (R)bw;
      }
    } else if (tid < cntp) {
      s_x[tid] = s_y[tid] = s_z[tid] = (R)0;
      s_a[tid] = (R)-INFINITY;
      s_w[tid] = (R)0;
    }
    __syncthreads();

    if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
      for (int j = 0; j < cnt; ++j) {
        const R sx = s_x[j], sy = s_y[j], sz = s_z[j], sa = s_a[j];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const R C = pair_cost<COST>(tx[r], ty[r], tz[r], sx, sy, sz, pexp);
          if constexpr (KIND == KIND_MAX) acc[r] = acc[r] > C ? acc[r] : C;
          else {
            const R v = C + sa;
            acc[r] = acc[r] < v ? acc[r] : v;
          }
        }
      }
    } else {
      for (int j = 0; j < cntp; j += kSub) {
        R v[RPT][kSub];
        R Cv[RPT][kSub];
        R wv[kSub], px[kSub], py[kSub], pz[kSub];
#pragma unroll
        for (int u = 0; u < kSub; ++u) {
          const R sx = s_x[j + u], sy = s_y[j + u], sz = s_z[j + u];
          const R sa = s_a[j + u];
          wv[u] = s_w[j + u];
          px[u] = sx; py[u] = sy; pz[u] = sz;
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            const R C = pair_cost<COST>(tx[r], ty[r], tz[r], sx, sy, sz, pexp);
            Cv[r][u] = C;
            v[r][u] = fma(k, C, sa + tq[r]);
          }
        }
        if constexpr (kMaxShift) {
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            R tm = v[r][0];
#pragma unroll
            for (int u = 1; u < kSub; ++u) tm = tm > v[r][u] ? tm : v[r][u];
            if constexpr (PREC == 1) {
              const float mn = fmaxf(m[r], ceilf(tm));
              const int sh = (int)fmaxf(m[r] - mn, -2000.0f);
              const double f = ldexp(1.0, sh);
              S[r] *= f;
              if (KIND == KIND_BARY) { Sy0[r] *= f; Sy1[r] *= f; Sy2[r] *= f; }
              m[r] = mn;
              float loc = 0.0f;
#pragma unroll
              for (int u = 0; u < kSub; ++u) loc = fmaf(exp2_fast(v[r][u] - mn), wv[u], loc);
              S[r] += (double)loc;
            } else {
              const double mn = fmax((double)m[r], (double)tm);
              if (mn > (double)m[r]) {
                const double f = exp((double)m[r] - mn);
                S[r] *= f;
                if (KIND == KIND_BARY) { Sy0[r] *= f; Sy1[r] *= f; Sy2[r] *= f; }
                m[r] = (R)mn;
              }
              double loc = 0.0, l0 = 0.0, l1 = 0.0, l2 = 0.0;
#pragma unroll
              for (int u = 0; u < kSub; ++u) {
                const double e = exp((double)v[r][u] - mn) * (double)wv[u];
                loc += e;
                if (KIND == KIND_BARY) {
                  l0 = fma(e, (double)px[u], l0);
                  l1 = fma(e, (double)py[u], l1);
                  l2 = fma(e, (double)pz[u], l2);
                }
              }
              S[r] += loc;
              if (KIND == KIND_BARY) { Sy0[r] += l0; Sy1[r] += l1; Sy2[r] += l2; }
            }
          }
        } else {  // KIND_SUM
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            R ls = (R)0, lc = (R)0;
#pragma unroll
            for (int u = 0; u < kSub; ++u) {
              R e;
              if constexpr (PREC == 1) e = exp2_fast(v[r][u]) * wv[u];
              else e = exp(v[r][u]) * wv[u];
              ls += e;
              lc = fma(e, Cv[r][u], lc);
            }
            S[r] += (double)ls;
            Sc[r] += (double)lc;
          }
        }
      }
    }
  }

  // ---- merge split partials (last-arriving workgroup) ----------------------
  if (splits > 1) {
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      double f[5];
      if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) f[0] = (double)acc[r];
      else if constexpr (KIND == KIND_SUM) { f[0] = S[r]; f[1] = Sc[r]; }
      else { f[0] = (double)m[r]; f[1] = S[r]; f[2] = Sy0[r]; f[3] = Sy1[r]; f[4] = Sy2[r]; }
#pragma unroll
      for (int q = 0; q < NF; ++q)
        w.partial[((int64_t)q * splits + split) * ntp + tidx[r]] = f[q];
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) {
      const unsigned prev = atomicAdd(&w.counters[tb], 1u);
      s_last = (prev == (unsigned)(splits - 1));
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if (tid == 0) w.counters[tb] = 0u;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int64_t t = tidx[r];
      if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
        double a = w.partial[t];
        for (int q = 1; q < splits; ++q) {
          const double b = w.partial[(int64_t)q * ntp + t];
          a = KIND == KIND_MAX ? fmax(a, b) : fmin(a, b);
        }
        acc[r] = (R)a;
      } else if constexpr (KIND == KIND_SUM) {
        double a = 0.0, c = 0.0;
        for (int q = 0; q < splits; ++q) {
          a += w.partial[(int64_t)q * ntp + t];
          c += w.partial[((int64_t)splits + q) * ntp + t];
        }
        S[r] = a;
        Sc[r] = c;
      } else {
        double M = -INFINITY;
        for (int q = 0; q < splits; ++q) M = fmax(M, w.partial[(int64_t)q * ntp + t]);
        double a = 0.0, y0 = 0.0, y1 = 0.0, y2 = 0.0;
        for (int q = 0; q < splits; ++q) {
          const double mq = w.partial[(int64_t)q * ntp + t];
          double fct;
          if (PREC == 1) fct = ldexp(1.0, (int)fmax(mq - M, -2000.0));
          else fct = exp(mq - M);
          a += fct * w.partial[((int64_t)splits + q) * ntp + t];
          if (KIND == KIND_BARY) {
            y0 += fct * w.partial[((int64_t)2 * splits + q) * ntp + t];
            y1 += fct * w.partial[((int64_t)3 * splits + q) * ntp + t];
            y2 += fct * w.partial[((int64_t)4 * splits + q) * ntp + t];
          }
        }
        m[r] = (R)M;
        S[r] = a;
        Sy0[r] = y0; Sy1[r] = y1; Sy2[r] = y2;
      }
    }
  }

  // ---- epilogue --------------------------------------------------------------
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int64_t t = tidx[r];
    if (t >= d.nt) continue;
    if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
      d.out0[t] = (double)acc[r];
    } else if constexpr (KIND == KIND_BARY) {
      d.out0[3 * t] = Sy0[r] / S[r];
      d.out0[3 * t + 1] = Sy1[r] / S[r];
      d.out0[3 * t + 2] = Sy2[r] / S[r];
    } else if constexpr (KIND == KIND_LSE) {
      double lse;
      if ((double)m[r] <= (double)kEmpty || S[r] == 0.0) lse = -INFINITY;
      else if (PREC == 1) lse = ((double)m[r] + log2(S[r]) - qd[r]) * kLn2;
      else lse = (double)m[r] + log(S[r]);
      if (d.epi == EPI_SK_HALF) d.out0[t] = log(d.marg[t]) - lse;
      else d.out0[t] = lse;
    } else {  // KIND_SUM
      double a = S[r], c = Sc[r];
      if (PREC == 1 && d.tgtA != nullptr) {
        const double corr = exp2(d.gamma * d.tgtA[t] * kLog2e - qd[r]);
        a *= corr;
        c *= corr;
      }
      if (d.epi == EPI_SK_RESID) {
        d.out0[t] = d.marg[t] - a;
        if (d.out1 != nullptr) d.out1[t] = c;
      } else {
        d.out0[t] = a;
        if (d.out1 != nullptr) d.out1[t] = c;
        if (d.epi == EPI_LAMP_COL && d.fuse_update) {
          const double th0 = d.th0[t];
          const double mixv = d.mix_src[t];
          const double th = theta_update(a, d.cvec[t], d.etamu[t], th0, d.eta_mu,
                                         d.adjust, d.minv, d.maxv, d.upd_winf);
          d.th_out[t] = th;
          d.nu_out[t] = (1.0 - d.mix_t) * d.nu_in[t] + d.mix_t * mixv;
        }
      }
    }
  }
}

template <int PREC, int KIND, int COST>
hipError_t launch_rpt(const PassDesc& d, const LaunchCfg& c, const Work& w, hipStream_t s) {
  dim3 grid((unsigned)c.tblocks, (unsigned)c.splits);
  switch (c.rpt) {
    case 1: hipLaunchKernelGGL((pair_pass<PREC, KIND, COST, 1>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    case 2: hipLaunchKernelGGL((pair_pass<PREC, KIND, COST, 2>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    case 4: hipLaunchKernelGGL((pair_pass<PREC, KIND, COST, 4>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int PREC, int KIND>
hipError_t launch_cost(const PassDesc& d, const LaunchCfg& c, const Work& w, hipStream_t s) {
  switch (cost_of_p(d.p)) {
    case COST_L1: return launch_rpt<PREC, KIND, COST_L1>(d, c, w, s);
    case COST_SQL2: return launch_rpt<PREC, KIND, COST_SQL2>(d, c, w, s);
    case COST_LINF: return launch_rpt<PREC, KIND, COST_LINF>(d, c, w, s);
    default: return launch_rpt<PREC, KIND, COST_GENP>(d, c, w, s);
  }
}

}  // namespace
}  // namespace culamp

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
//   * when a target's sources are split over several workgroups (grid.y),
//     each writes its partial state to HBM and pass_merge, launched next on
//     the same stream, combines the splits in a fixed order, so results are
//     bitwise reproducible run to run;
//   * the merged result goes through a per-target epilogue (log marginal
//     minus LSE, marginal minus sum, fused LAMP dual update, ...);
//   * every kernel body also has a batched entry point (*_batch) whose
//     workgroups each take one (instance, source split, target block) item
//     from a table, so many instances of different sizes share one launch.
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
// fourth word of a raw (stride 0) CDNA buffer resource: 32-bit data format,
// no swizzle; with it the hardware checks each offset against num_records
constexpr int kBufferWord3 = 0x00020000;

// 16-byte copies from global memory straight into LDS (buffer_load ... lds):
// `base` and `nbytes` are wave-uniform (the resource is built in scalar
// registers), `voff` is the lane's byte offset, voff + 16 <= nbytes; lane l
// writes bytes [16 l, 16 l + 16) from `lds`.
__device__ __forceinline__ void lds_copy16(const void* base, const int nbytes, const unsigned voff,
                                           __attribute__((address_space(3))) void* lds) {
  const auto r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, nbytes, kBufferWord3);
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, lds, 16, voff, 0, 0, 0);
}

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

// FP64 fields of a partial state (split partials in HBM, the workgroup-wide
// layout's LDS exchange): MAX / MIN [acc]; SUM [S, Sc, Se, Sce]; LSE
// [m, S, Se]; BARY [m, S, Se, Sy0, Sy1, Sy2].  The compensation terms travel
// with their sums, so a merge loses nothing to a partial that is large
// beside the final sum (cancelling signed terms).
template <int KIND>
constexpr int kind_fields() {
  return KIND == KIND_BARY ? 6 : (KIND == KIND_MAX || KIND == KIND_MIN) ? 1 : KIND == KIND_SUM ? 4 : 3;
}

// Per-target accumulator state of a pass.  R is the pair arithmetic type; in
// mixed precision `m` is an integer-valued FP32 running max in log2 units.
// Compensated FP64 sums: the marginal and cost sums of SUM passes (both
// precisions) and the exp-sums of FP64 passes add with TwoSum, the rounding
// errors going to Se / Sce (scaled with S when the reference moves); in the
// mixed SUM passes the FP32 sub-tile sums of a tile first meet in a plain
// FP64 block sum that is TwoSum-added once per tile.  Merges of partial
// states are TwoSums too, and the epilogue folds the errors in (settle).  A mixed-precision LSE adds its FP32 sub-tile sums to S plainly:
// each such sum carries a relative error near 2^-24, against 2^-53 for the
// FP64 addition, so a compensation could not change its result.
template <typename R>
struct TState {
  R m;
  double S, Sc, Sy0, Sy1, Sy2;
  R acc;
  double qd;  // exact value of the per-target exponent offset used
  double Se = 0.0, Sce = 0.0;
  // Mixed SUM passes with block flushes: the FP32 sub-tile sums of one tile
  // add plainly here (a handful of terms) and flush_blocks moves the block
  // into (S, Se) / (Sc, Sce) with TwoSum once per tile.
  double Sb = 0.0, Scb = 0.0;
};

template <int PREC, int KIND, typename R>
__device__ __forceinline__ void flush_blocks(TState<R>& st) {
  if constexpr (PREC == 1 && KIND == KIND_SUM) {
    two_sum_acc(st.S, st.Se, st.Sb);
    two_sum_acc(st.Sc, st.Sce, st.Scb);
    st.Sb = 0.0;
    st.Scb = 0.0;
  }
}

template <typename R>
__device__ __forceinline__ void settle(TState<R>& st) {
  st.S += st.Se + st.Sb;
  st.Sc += st.Sce + st.Scb;
  st.Se = st.Sb = 0.0;
  st.Sce = st.Scb = 0.0;
}

// A state as its partial fields (see kind_fields) and back.
template <int KIND, typename R>
__device__ __forceinline__ void to_fields(const TState<R>& st, double (&f)[6]) {
  if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
    f[0] = (double)st.acc;
  } else if constexpr (KIND == KIND_SUM) {
    f[0] = st.S; f[1] = st.Sc; f[2] = st.Se; f[3] = st.Sce;
  } else {
    f[0] = (double)st.m; f[1] = st.S; f[2] = st.Se;
    if constexpr (KIND == KIND_BARY) { f[3] = st.Sy0; f[4] = st.Sy1; f[5] = st.Sy2; }
  }
}

template <int KIND, typename R>
__device__ __forceinline__ void from_fields(TState<R>& st, const double (&f)[6]) {
  if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
    st.acc = (R)f[0];
  } else if constexpr (KIND == KIND_SUM) {
    st.S = f[0]; st.Sc = f[1]; st.Se = f[2]; st.Sce = f[3];
  } else {
    st.m = (R)f[0]; st.S = f[1]; st.Se = f[2];
    if constexpr (KIND == KIND_BARY) { st.Sy0 = f[3]; st.Sy1 = f[4]; st.Sy2 = f[5]; }
  }
}

template <int PREC, typename R>
__device__ __forceinline__ R empty_max() {
  return PREC == 0 ? (R)-1e300 : (R)-1e30f;
}

// Combine two partial states of one target (commutative: the same result
// whichever operand comes first).
template <int PREC, int KIND, typename R>
__device__ __forceinline__ void combine_state(TState<R>& a, const TState<R>& b) {
  if constexpr (KIND == KIND_MAX) {
    a.acc = a.acc > b.acc ? a.acc : b.acc;
  } else if constexpr (KIND == KIND_MIN) {
    a.acc = a.acc < b.acc ? a.acc : b.acc;
  } else if constexpr (KIND == KIND_SUM) {
    a.Se += b.Se;
    two_sum_acc(a.S, a.Se, b.S);
    a.Sce += b.Sce;
    two_sum_acc(a.Sc, a.Sce, b.Sc);
  } else {
    const R mn = a.m > b.m ? a.m : b.m;
    double fa, fb;
    if constexpr (PREC == 1) {
      fa = ldexp(1.0, (int)fmaxf(a.m - mn, -2000.0f));
      fb = ldexp(1.0, (int)fmaxf(b.m - mn, -2000.0f));
    } else {
      fa = exp((double)a.m - (double)mn);
      fb = exp((double)b.m - (double)mn);
    }
    {
#pragma clang fp contract(off)
      const double sa = a.S * fa, sb = b.S * fb;
      a.Se = a.Se * fa + b.Se * fb;
      a.S = sa;
      two_sum_acc(a.S, a.Se, sb);
    }
    if (KIND == KIND_BARY) {
      a.Sy0 = a.Sy0 * fa + b.Sy0 * fb;
      a.Sy1 = a.Sy1 * fa + b.Sy1 * fb;
      a.Sy2 = a.Sy2 * fa + b.Sy2 * fb;
    }
    a.m = mn;
  }
}

// Butterfly combine across the G adjacent lanes that share a target.
template <int PREC, int KIND, int G, typename R>
__device__ __forceinline__ void combine_lanes(TState<R>& st) {
  flush_blocks<PREC, KIND, R>(st);
#pragma unroll
  for (int off = 1; off < G; off <<= 1) {
    TState<R> o = st;
    o.m = __shfl_xor(st.m, off);
    o.S = __shfl_xor(st.S, off);
    o.Sc = __shfl_xor(st.Sc, off);
    o.Se = __shfl_xor(st.Se, off);
    if (KIND == KIND_SUM) o.Sce = __shfl_xor(st.Sce, off);
    o.acc = __shfl_xor(st.acc, off);
    if (KIND == KIND_BARY) {
      o.Sy0 = __shfl_xor(st.Sy0, off);
      o.Sy1 = __shfl_xor(st.Sy1, off);
      o.Sy2 = __shfl_xor(st.Sy2, off);
    }
    combine_state<PREC, KIND, R>(st, o);
  }
}

// Epilogue of one target once its state is complete.
// The per-target inputs of the SUM / LSE epilogues, loaded together so their
// round trips overlap (and, in the workgroup-wide layout, ahead of the pair
// loop).
struct EpiLoad {
  double ta;                      // tgtA[t] (mixed SUM exponent correction)
  double mg;                      // marg[t] (EPI_SK_HALF, EPI_SK_RESID)
  double th0, c, emu, nu, mix;    // EPI_LAMP_COL fused update inputs
  double rx, ry, rz, rw;          // record of the next pass: tgt[t], rec_wsrc[t]
};

template <int PREC, int KIND>
__device__ __forceinline__ EpiLoad load_epilogue_inputs(const PassDesc& d, const int64_t t) {
  EpiLoad e;
  e.ta = e.mg = e.th0 = e.c = e.emu = e.nu = e.mix = 0.0;
  e.rx = e.ry = e.rz = 0.0;
  e.rw = 1.0;
  if constexpr (PREC == 1 && (KIND == KIND_SUM || KIND == KIND_LSE)) {
    if (d.rec_out != nullptr) {
      e.rx = d.tgt[3 * t];
      e.ry = d.tgt[3 * t + 1];
      e.rz = d.tgt[3 * t + 2];
      if (d.rec_wsrc != nullptr) e.rw = d.rec_wsrc[t];
    }
  }
  if constexpr (KIND == KIND_SUM) {
    if (PREC == 1 && d.tgtA != nullptr) e.ta = d.tgtA[t];
    if (d.epi == EPI_SK_RESID) e.mg = d.marg[t];
    if (d.epi == EPI_LAMP_COL && d.fuse_update) {
      e.th0 = d.th0[t];
      e.c = d.cvec[t];
      e.emu = d.etamu[t];
      e.nu = d.nu_in[t];
      e.mix = d.mix_src[t];
    }
  } else if constexpr (KIND == KIND_LSE) {
    if (d.epi == EPI_SK_HALF) e.mg = d.marg[t];
  }
  return e;
}

// Mixed precision: the FP32 exponent term and weight of a source whose
// exponent is av (natural units) and weight bw: a = fl32(av log2 e) and
// w = bw (1 + lo ln 2) with lo the FP64 remainder of a, so that
// 2^a w = bw e^av to FP32 accuracy.  (Contraction off: the staging in the
// pair loop and the record written by an epilogue must round alike.)
__device__ __forceinline__ void split_term(const double av, const double bw, float* a, float* w) {
#pragma clang fp contract(off)
  const double a2 = av * kLog2e;
  const float hi = (float)a2;
  const double lo = a2 - (double)hi;
  *a = hi;
  *w = (float)(isfinite(lo) ? bw * (1.0 + lo * kLn2) : bw);
}

// Coordinate scale of a mixed squared-Euclidean LSE pass with cost factor k
// (natural units, k <= 0): sqrt(-k log2 e).
__device__ __forceinline__ double lse_coord_scale(const double k) {
#pragma clang fp contract(off)
  return sqrt(fmax(-(k * kLog2e), 0.0));
}

// Record of target t as a source of the next pass (see PassDesc::rec); v is
// the value this epilogue wrote for t.
__device__ __forceinline__ void emit_record(const PassDesc& d, const int64_t t, const double v,
                                            const EpiLoad& E) {
  const double csc = d.rec_scaled ? lse_coord_scale(d.rec_k) : 1.0;
  float4 r;
  r.x = (float)(d.rec_scaled ? E.rx * csc : E.rx);
  r.y = (float)(d.rec_scaled ? E.ry * csc : E.ry);
  r.z = (float)(d.rec_scaled ? E.rz * csc : E.rz);
  float w;
  split_term(d.rec_alpha * v, E.rw, &r.w, &w);
  d.rec_out[t] = r;
  d.recw_out[t] = w;
}

template <int PREC, int KIND, typename R>
__device__ __forceinline__ void epilogue_from(const PassDesc& d, const int64_t t,
                                              const TState<R>& st_in, const EpiLoad& E) {
  TState<R> st = st_in;
  settle(st);
  if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
    d.out0[t] = (double)st.acc;
  } else if constexpr (KIND == KIND_BARY) {
    if (d.epi == EPI_ADJ) {
      const double p0 = d.tgt[3 * t], p1 = d.tgt[3 * t + 1], p2 = d.tgt[3 * t + 2];
      if (d.adj_norm) {
        const double w = (d.tgtW != nullptr ? d.tgtW[t] : 1.0) * d.adj_c0;
        const double inv = 1.0 / st.S;
        d.out0[3 * t] += w * (p0 - st.Sy0 * inv);
        d.out0[3 * t + 1] += w * (p1 - st.Sy1 * inv);
        d.out0[3 * t + 2] += w * (p2 - st.Sy2 * inv);
      } else {
        const double f = exp((double)st.m + (d.tgtA != nullptr ? d.gamma * d.tgtA[t] : 0.0));
        const double S = st.S * f;
        if (d.out1 != nullptr)
          d.out1[t] = d.adj_beta != 0.0 ? d.adj_beta * d.out1[t] + d.adj_c1 * S : d.adj_c1 * S;
        if (d.out0 != nullptr) {
          d.out0[3 * t] += d.adj_c0 * (p0 * S - st.Sy0 * f);
          d.out0[3 * t + 1] += d.adj_c0 * (p1 * S - st.Sy1 * f);
          d.out0[3 * t + 2] += d.adj_c0 * (p2 * S - st.Sy2 * f);
        }
      }
    } else {
      d.out0[3 * t] = st.Sy0 / st.S;
      d.out0[3 * t + 1] = st.Sy1 / st.S;
      d.out0[3 * t + 2] = st.Sy2 / st.S;
    }
  } else if constexpr (KIND == KIND_LSE) {
    double lse;
    if ((double)st.m <= (double)empty_max<PREC, R>() || st.S == 0.0) lse = -INFINITY;
    else if (PREC == 1) lse = ((double)st.m + log2(st.S) - st.qd) * kLn2;
    else lse = (double)st.m + log(st.S);
    double v = lse;
    if (d.epi == EPI_SK_HALF) {
      asm volatile("" ::: "memory");  // (keeps the logarithm behind the branch)
      v = log(E.mg) - lse;
    }
    d.out0[t] = v;
    if (PREC == 1 && d.rec_out != nullptr) emit_record(d, t, v, E);
  } else {  // KIND_SUM
    double a = st.S, c = st.Sc;
    if (PREC == 1 && d.tgtA != nullptr) {
      const double corr = exp2(d.gamma * E.ta * kLog2e - st.qd);
      a *= corr;
      c *= corr;
    }
    if (d.epi == EPI_SK_RESID) {
      d.out0[t] = E.mg - a;
      if (d.out1 != nullptr) d.out1[t] = c;
    } else {
      d.out0[t] = a;
      if (d.out1 != nullptr) d.out1[t] = c;
      if (d.epi == EPI_LAMP_COL && d.fuse_update) {
        const double th = theta_update(a, E.c, E.emu, E.th0, d.eta_mu, d.adjust, d.minv,
                                       d.maxv, d.upd_winf);
        d.th_out[t] = th;
        const double nu = (1.0 - d.mix_t) * E.nu + d.mix_t * E.mix;
        d.nu_out[t] = nu;
        if (PREC == 1 && d.rec_out != nullptr) emit_record(d, t, nu, E);
      }
    }
  }
}

template <int PREC, int KIND, typename R>
__device__ __forceinline__ void epilogue_one(const PassDesc& d, const int64_t t, const TState<R>& st) {
  epilogue_from<PREC, KIND, R>(d, t, st, load_epilogue_inputs<PREC, KIND>(d, t));
}

// With one split the workgroup applies the epilogue itself; with several it
// writes its partial state ([field][split][target]) and pass_merge, launched
// next on the same stream, combines the splits in split order.
template <int PREC, int KIND, int NT, typename R>
__device__ __forceinline__ void finish_pass(const PassDesc& d, const Work& w,
                                            const int splits, const int split,
                                            const int64_t ntp, const int64_t (&tidx)[NT],
                                            const TState<R> (&st)[NT]) {
  constexpr int NF = kind_fields<KIND>();
  if (splits > 1) {
#pragma unroll
    for (int r = 0; r < NT; ++r) {
      if (tidx[r] >= ntp) continue;
      double f[6];
      to_fields<KIND, R>(st[r], f);
#pragma unroll
      for (int q = 0; q < NF; ++q)
        w.partial[((int64_t)q * splits + split) * ntp + tidx[r]] = f[q];
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < NT; ++r)
    if (tidx[r] < d.nt) epilogue_one<PREC, KIND, R>(d, tidx[r], st[r]);
}

// Per-target exponent offset q_t: the target term of a SUM pass, or (mixed
// LSE passes of the solver loops) minus the target's previous LSE so that
// FP32 exponents stay near zero.  Returns q in natural units.
// The per-target inputs of a point pass, loaded together (t must be a valid
// target index) so that their round trips overlap.
struct TgtLoad {
  double x, y, z;     // coordinates (point passes)
  double ta, prev, mg;  // tgtA[t], out0[t], marg[t] as target_offset needs them
};

template <int PREC, int KIND>
__device__ __forceinline__ TgtLoad load_offset_inputs(const PassDesc& d, const int64_t t) {
  TgtLoad L;
  L.x = L.y = L.z = 0.0;
  L.ta = (KIND == KIND_SUM && d.tgtA != nullptr) ? d.tgtA[t] : 0.0;
  const bool prev = PREC == 1 && KIND == KIND_LSE && d.prev_offset;
  L.prev = prev ? d.out0[t] : 0.0;
  L.mg = (prev && d.epi == EPI_SK_HALF) ? d.marg[t] : 1.0;
  return L;
}

template <int PREC, int KIND>
__device__ __forceinline__ double offset_from(const PassDesc& d, const TgtLoad& L) {
  double q = 0.0;
  if (KIND == KIND_SUM && d.tgtA != nullptr) q = d.gamma * L.ta;
  if (PREC == 1 && KIND == KIND_LSE && d.prev_offset) {
    // (a branch on the uniform epilogue kind, kept a branch by the empty
    // volatile statement, so that the other LSE passes do not evaluate the
    // logarithm)
    double lse_prev = L.prev;
    if (d.epi == EPI_SK_HALF) {
      asm volatile("" ::: "memory");
      lse_prev = log(L.mg) - L.prev;
    }
    q = isfinite(lse_prev) ? -lse_prev : 0.0;
  }
  return q;
}

template <int PREC, int KIND>
__device__ __forceinline__ double target_offset(const PassDesc& d, int64_t t) {
  return offset_from<PREC, KIND>(d, load_offset_inputs<PREC, KIND>(d, t));
}

// Combine the split partials of each target, then apply the epilogue.
// kMergeLanes adjacent lanes share a target; lane `sub` folds splits sub,
// sub + kMergeLanes, ... in order (unrolled, so that their loads go out
// together), then the lanes combine by butterfly: a fixed order, so the
// result is deterministic and equal to the workgroup-wide layout's.

template <int PREC, int KIND>
__device__ __forceinline__ void merge_body(const PassDesc& d, const Work& w, const int64_t block,
                                           const int splits, const int64_t ntp) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  constexpr int NF = kind_fields<KIND>();
  const int sub = threadIdx.x % kMergeLanes;
  const int64_t t = block * (kThreads / kMergeLanes) + threadIdx.x / kMergeLanes;
  const bool ok = t < d.nt;
  // the owner lane's per-target inputs go out with the partial loads
  const bool own = ok && sub == 0;
  TgtLoad L{};
  EpiLoad E{};
  if constexpr (KIND != KIND_BARY) {
    if (own) {
      L = load_offset_inputs<PREC, KIND>(d, t);
      E = load_epilogue_inputs<PREC, KIND>(d, t);
    }
  }
  TState<R> st;
  st.m = empty_max<PREC, R>();
  st.S = st.Sc = st.Sy0 = st.Sy1 = st.Sy2 = 0.0;
  st.acc = KIND == KIND_MIN ? (R)INFINITY : (R)-INFINITY;
  if (ok) {
#pragma unroll 4
    for (int q = sub; q < splits; q += kMergeLanes) {
      TState<R> p = st;
      double f[6];
#pragma unroll
      for (int fi = 0; fi < NF; ++fi) f[fi] = w.partial[((int64_t)fi * splits + q) * ntp + t];
      from_fields<KIND, R>(p, f);
      combine_state<PREC, KIND, R>(st, p);
    }
  }
  combine_lanes<PREC, KIND, kMergeLanes, R>(st);
  if (!own) return;
  if constexpr (KIND == KIND_BARY) {
    const double q = target_offset<PREC, KIND>(d, t);
    st.qd = PREC == 1 ? (double)(float)(q * kLog2e) : q;
    epilogue_one<PREC, KIND, R>(d, t, st);
  } else {
    const double q = offset_from<PREC, KIND>(d, L);
    st.qd = PREC == 1 ? (double)(float)(q * kLog2e) : q;
    epilogue_from<PREC, KIND, R>(d, t, st, E);
  }
}

template <int PREC, int KIND>
__global__ __launch_bounds__(kThreads) void pass_merge(const PassDesc din, const Work w,
                                                       const int splits, const int64_t ntp) {
  const PassDesc d = resolve_desc(din);
  merge_body<PREC, KIND>(d, w, blockIdx.x, splits, ntp);
}

// Batched merge: items[i] = (instance, merge block) for instances with splits.
template <int PREC, int KIND>
__global__ __launch_bounds__(kThreads) void pass_merge_batch(const PassDesc* descs,
                                                             const BatchItem* items,
                                                             const BatchInst* insts,
                                                             const Work w) {
  const BatchItem it = items[blockIdx.x];
  if (it.inst < 0) return;  // padding of an XCD's item sequence
  const BatchInst bi = insts[it.inst];
  const PassDesc d = resolve_desc(descs[it.inst]);
  const Work wi{w.partial + bi.pbase};
  merge_body<PREC, KIND>(d, wi, it.tb, bi.splits, bi.ntp);
}

template <int PREC, int KIND>
hipError_t launch_merge(const PassDesc& d, const LaunchCfg& c, const Work& w, int64_t ntp,
                        hipStream_t s) {
  if (c.splits <= 1) return hipSuccess;
  const int64_t per = kThreads / kMergeLanes;
  const unsigned blocks = (unsigned)((d.nt + per - 1) / per);
  hipLaunchKernelGGL((pass_merge<PREC, KIND>), dim3(blocks), dim3(kThreads), 0, s, d, w,
                     c.splits, ntp);
  return hipGetLastError();
}

// Source exponent term and weight as staged in LDS (see the header comment).
template <int PREC, bool EXPKIND, typename R>
__device__ __forceinline__ void source_terms_v(const PassDesc& d, const double a_raw,
                                               const double b_raw, R* a_out, R* w_out) {
  const double av = d.srcA != nullptr ? d.alpha * a_raw : 0.0;
  const double bw = d.srcB != nullptr ? b_raw : 1.0;
  if (PREC == 1 && EXPKIND) {
    float a, w;
    split_term(av, bw, &a, &w);
    *a_out = (R)a;
    *w_out = (R)w;
  } else {
    *a_out = (R)av;
    *w_out = (R)bw;
  }
}

template <int PREC, bool EXPKIND, typename R>
__device__ __forceinline__ void source_terms(const PassDesc& d, int64_t s, R* a_out, R* w_out) {
  source_terms_v<PREC, EXPKIND, R>(d, d.srcA != nullptr ? d.srcA[s] : 0.0,
                                   d.srcB != nullptr ? d.srcB[s] : 1.0, a_out, w_out);
}

// 2^x for the small arguments of lse_lazy's rescale (a difference of two
// FP32 rounding errors, |x| <= 2^-4 unless an exponent offset exceeds 2^20
// binades): Taylor polynomial of degree 6 in x ln 2, relative error below
// 2^-36 for |x| <= 2^-4; larger arguments take the library exp2.
__device__ __forceinline__ double exp2_small(double x) {
  if (!(fabs(x) <= 0.0625)) return exp2(x);
  const double y = x * kLn2;
  return 1.0 + y * (1.0 + y * (0.5 + y * (1.0 / 6.0 + y * (1.0 / 24.0 + y * (1.0 / 120.0 +
                                                                          y * (1.0 / 720.0))))));
}

// Mixed-precision LSE accumulation with a lazily moved reference.  The
// caller forms each exponent relative to the target's reference as
// v' = k C + (a_s + tqm), tqm = fl(tq - m), with m integer valued; S holds
// sum 2^(v_true + tq - m + dl), dl = tqm - (tq - m) exactly (FP64).  While the
// sub-tile sum stays in [2^-90, 2^100] relative to the reference (the common
// case once the target offset centres the exponents near 0) no per-pair max
// or shift is needed: the FP32 terms of such a sum that could underflow are
// below 2^-30 of it.  Otherwise the reference moves to the ceiling of the
// larger of the sub-tile max and the binary exponent of S (an upper bound of
// log2 S), S is rescaled in FP64 and the sub-tile is re-evaluated.  The
// caller removes dl from S at the end (S *= 2^-dl), which restores the (m, S)
// convention of the other kinds.
template <int NV>
__device__ __forceinline__ void lse_lazy(TState<float>& st, const float (&v)[NV],
                                         const float (&wv)[NV], const float tq, float& tqm,
                                         double& dl) {
  float loc = exp2_fast(v[0]) * wv[0];
#pragma unroll
  for (int u = 1; u < NV; ++u) loc = fmaf(exp2_fast(v[u]), wv[u], loc);
  const double Sn = st.S + (double)loc;
  if (loc < 0x1p100f && Sn > 0x1p-90) {
    st.S = Sn;
    return;
  }
  float tm = v[0];
#pragma unroll
  for (int u = 1; u < NV; ++u) tm = fmaxf(tm, v[u]);
  int es = 0;
  if (st.S > 0.0) (void)frexp(st.S, &es);
  const float sl = st.S > 0.0 ? (float)es : -INFINITY;
  const float top = fmaxf(tm, sl);
  if (!(top > -INFINITY) || !(top < INFINITY)) {
    st.S = Sn;
    return;
  }
  const float c = ceilf(top);
  const float mnew = st.m + c;
  const float tqm_new = tq - mnew;
  const double dnew = (double)tqm_new - ((double)tq - (double)mnew);
  const float dv = tqm_new - tqm;
  st.S = ldexp(st.S, -(int)c) * exp2_small(dnew - dl);
  loc = 0.0f;
#pragma unroll
  for (int u = 0; u < NV; ++u) loc = fmaf(exp2_fast(v[u] + dv), wv[u], loc);
  st.S += (double)loc;
  st.m = mnew;
  tqm = tqm_new;
  dl = dnew;
}

// Fold NV exponents of one target into its running state.
// BLOCK (mixed SUM passes): the sub-tile sums go to the block sums Sb / Scb,
// which the caller flushes into the compensated sums once per tile.
template <int PREC, int KIND, int NV, typename R, bool NEEDC = true, bool BLOCK = false>
__device__ __forceinline__ void accumulate(TState<R>& st, const R (&v)[NV], const R (&wv)[NV],
                                           const R (&Cv)[NV], const R (&px)[NV],
                                           const R (&py)[NV], const R (&pz)[NV]) {
  if constexpr (KIND == KIND_LSE || KIND == KIND_BARY) {
    R tm = v[0];
#pragma unroll
    for (int u = 1; u < NV; ++u) tm = fmax(tm, v[u]);
    if constexpr (PREC == 1) {
      const float mn = fmaxf(st.m, ceilf(tm));
      const int sh = (int)fmaxf(st.m - mn, -2000.0f);
      st.S *= ldexp(1.0, sh);
      st.m = mn;
      float loc = exp2_fast(v[0] - mn) * wv[0];
#pragma unroll
      for (int u = 1; u < NV; ++u) loc = fmaf(exp2_fast(v[u] - mn), wv[u], loc);
      st.S += (double)loc;
    } else {
      const double mn = fmax((double)st.m, (double)tm);
      if (mn > (double)st.m) {
        const double f = exp((double)st.m - mn);
        st.S *= f;
        st.Se *= f;
        if (KIND == KIND_BARY) { st.Sy0 *= f; st.Sy1 *= f; st.Sy2 *= f; }
        st.m = (R)mn;
      }
      // (every term goes into the compensated sum on its own)
      double l0 = 0.0, l1 = 0.0, l2 = 0.0;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const double e = exp((double)v[u] - mn) * (double)wv[u];
        two_sum_acc(st.S, st.Se, e);
        if (KIND == KIND_BARY) {
          l0 = fma(e, (double)px[u], l0);
          l1 = fma(e, (double)py[u], l1);
          l2 = fma(e, (double)pz[u], l2);
        }
      }
      if (KIND == KIND_BARY) { st.Sy0 += l0; st.Sy1 += l1; st.Sy2 += l2; }
    }
  } else if constexpr (KIND == KIND_SUM) {
    if constexpr (PREC == 0) {
      // every term into the compensated sums; the cost term's product is
      // split exactly (e C = p + fma(e, C, -p))
#pragma clang fp contract(off)
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const double e = exp((double)v[u]) * (double)wv[u];
        two_sum_acc(st.S, st.Se, e);
        if constexpr (NEEDC) {
          const double p = e * (double)Cv[u];
          st.Sce += fma(e, (double)Cv[u], -p);
          two_sum_acc(st.Sc, st.Sce, p);
        }
      }
    } else {
      // (the sums start from the first term: no addition of zero)
      float ls = exp2_fast(v[0]) * wv[0];
      float lc = NEEDC ? ls * Cv[0] : 0.0f;
#pragma unroll
      for (int u = 1; u < NV; ++u) {
        const float e = exp2_fast(v[u]) * wv[u];
        ls += e;
        if constexpr (NEEDC) lc = fmaf(e, Cv[u], lc);
      }
      if constexpr (BLOCK) {
        st.Sb += (double)ls;
        if constexpr (NEEDC) st.Scb += (double)lc;
      } else {
        two_sum_acc(st.S, st.Se, (double)ls);
        if constexpr (NEEDC) two_sum_acc(st.Sc, st.Sce, (double)lc);
      }
    }
  }
}

// A source point as staged in LDS: coordinates and exponent term in one
// 16-byte (FP32) or 32-byte (FP64) record, so a lane reads it with one wide
// LDS load; the weight sits in its own array.
template <typename R>
struct alignas(4 * sizeof(R)) SrcPt {
  R x, y, z, a;
};

// Barrier among the NTHR threads that share a staged tile: the workgroup, or
// one wave (a wave's LDS operations complete in issue order, so it needs only
// compiler fences).
template <int NTHR>
__device__ __forceinline__ void group_sync() {
  if constexpr (NTHR == kThreads) {
    __syncthreads();
  } else {
    static_assert(NTHR == 64, "a group is a workgroup of kThreads or one wave");
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// The pair loop of one group of NTHR threads (kLanesPerTarget adjacent lanes
// per target, RPT targets per lane group) over sources [s0, s1), staged in
// tiles of TILE through the group's LDS records s_p / s_w.  Targets are
// t0 + r * (NTHR / kLanesPerTarget) + lane / kLanesPerTarget.  On return the
// lanes have combined their states; the owner lane of each target keeps its
// index in tidx, the others hold `none`.
template <int PREC, int KIND, int COST, int RPT, int NTHR, int TILE, typename R, bool REC = false>
__device__ __forceinline__ void pair_accum(const PassDesc& d, SrcPt<R>* s_p, R* s_w,
                                           const int tid, const int64_t t0, const int64_t s0,
                                           const int64_t s1, const int64_t none,
                                           int64_t (&tidx)[RPT], TState<R> (&st)[RPT]) {
  constexpr bool kExpKind = (KIND == KIND_LSE || KIND == KIND_SUM || KIND == KIND_BARY);
  constexpr int G = kLanesPerTarget, TPB = NTHR / G;
  static_assert(TILE % (kSub * G) == 0, "tile / sub-tile shape");

  const int g = tid % G, tt = tid / G;
  const R k = PREC == 0 ? (R)d.k : (R)(d.k * kLog2e);
  const R pexp = (R)d.p;

  // mixed LSE: lazily moved reference (see lse_lazy)
  constexpr bool kLazy = (PREC == 1 && KIND == KIND_LSE);
  // mixed squared-Euclidean LSE: coordinates are staged pre-multiplied by
  // sqrt(-k) (k <= 0 in log2 units for every LSE pass), so the exponent
  // (a_s + tqm) - |t' - s'|^2 is three fused multiply-adds after the
  // differences, with no separate scaling of the cost
  constexpr bool kScaled = kLazy && COST == COST_SQL2;
  const double csc = kScaled ? lse_coord_scale(d.k) : 1.0;
  // staged records from the previous pass's epilogue (mixed LSE / SUM)
  // (REC: this instantiation reads them; the launcher picks it when d.rec is set)
  constexpr bool use_rec = REC && PREC == 1 && kExpKind;
  // This thread's source of the next tile, held in registers from its load
  // until it is written to LDS: the loads of a tile go out together and one
  // tile ahead of the pair loop (every index is clamped into range, so no
  // load waits on a branch).
  static_assert(TILE == NTHR, "one staged source per thread and tile");
  double px = 0.0, py = 0.0, pz = 0.0, pa = 0.0, pb = 1.0;
  float4 rq = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float rqw = 0.0f;
  // (tile offsets are 32-bit, relative to s0: a split holds fewer than 2^31
  // sources)
  const int span = s1 > s0 ? (int)(s1 - s0) : 0;
  auto fetch = [&](const int off) {
    const int64_t sidx = s0 + min(off + tid, span - 1);
    if constexpr (use_rec) {
      rq = d.rec[sidx];
      rqw = d.recw[sidx];
      return;
    }
    px = d.src[3 * sidx];
    py = d.src[3 * sidx + 1];
    pz = d.src[3 * sidx + 2];
    pa = d.srcA != nullptr ? d.srcA[sidx] : 0.0;
    pb = d.srcB != nullptr ? d.srcB[sidx] : 1.0;
  };
  if (span > 0) fetch(0);

  R tx[RPT], ty[RPT], tz[RPT], tq[RPT], tqm[RPT];
  double dl[RPT];
  TgtLoad tl[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int64_t t = t0 + (int64_t)r * TPB + tt;
    const int64_t tcl = min(t, d.nt - 1);
    tl[r] = load_offset_inputs<PREC, KIND>(d, tcl);
    tl[r].x = d.tgt[3 * tcl];
    tl[r].y = d.tgt[3 * tcl + 1];
    tl[r].z = d.tgt[3 * tcl + 2];
  }
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int64_t t = t0 + (int64_t)r * TPB + tt;
    tidx[r] = t;
    const bool ok = t < d.nt;
    tx[r] = ok ? (R)(kScaled ? tl[r].x * csc : tl[r].x) : (R)0;
    ty[r] = ok ? (R)(kScaled ? tl[r].y * csc : tl[r].y) : (R)0;
    tz[r] = ok ? (R)(kScaled ? tl[r].z * csc : tl[r].z) : (R)0;
    const double q = ok ? offset_from<PREC, KIND>(d, tl[r]) : 0.0;
    if (PREC == 1) {
      const float q2 = (float)(q * kLog2e);
      tq[r] = (R)q2;
      st[r].qd = (double)q2;
    } else {
      tq[r] = (R)q;
      st[r].qd = q;
    }
    st[r].m = kLazy ? (R)0 : empty_max<PREC, R>();
    st[r].S = st[r].Sc = st[r].Sy0 = st[r].Sy1 = st[r].Sy2 = 0.0;
    st[r].acc = KIND == KIND_MIN ? (R)INFINITY : (R)-INFINITY;
    tqm[r] = tq[r];
    dl[r] = 0.0;
  }

  // Planar tiles (squared Euclidean): when every staged source of a tile and
  // every target of the group has z == 0, the z terms of the cost are exact
  // zeros (fma(-0, 0, a) == a, fma(g, g, 0 * 0) == g * g), so the tile runs a
  // loop without them and gives bit-identical results.
  constexpr bool kPlanarOk = COST == COST_SQL2 && kExpKind;
  bool tgt_nz = false;
#pragma unroll
  for (int r = 0; r < RPT; ++r) tgt_nz = tgt_nz || tz[r] != (R)0;
  __shared__ int s_nz[NTHR / 64];

  // One wave's tile (the workgroup-wide layout) knows whether it is planar
  // before it writes LDS: a planar tile then carries each source's weight in
  // the z slot, so the pair loop reads a source with one 16-byte load.
  constexpr bool kPacked = kPlanarOk && NTHR == 64;
  for (int off = 0; off < span; off += TILE) {
    const int cnt = min(TILE, span - off);
    const int cntp = (cnt + kSub * G - 1) / (kSub * G) * (kSub * G);
    group_sync<NTHR>();
    // every thread stages its source (past the split's end, a copy of the
    // last one); in a partial tile the slots from cnt to the sub-tile
    // boundary cntp are then overwritten by padding that adds nothing
    bool planar = false;
    {
      SrcPt<R> sp;
      R wsrc;
      if constexpr (use_rec) {
        sp.x = (R)rq.x;
        sp.y = (R)rq.y;
        sp.z = (R)rq.z;
        sp.a = (R)rq.w;
        wsrc = (R)rqw;
      } else {
        sp.x = (R)(kScaled ? px * csc : px);
        sp.y = (R)(kScaled ? py * csc : py);
        sp.z = (R)(kScaled ? pz * csc : pz);
        source_terms_v<PREC, kExpKind>(d, pa, pb, &sp.a, &wsrc);
      }
      if constexpr (kPlanarOk) {
        // any nonzero z in the wave, then (workgroup groups) across its waves
        // through one LDS word per wave, read after the tile's barrier
        const bool wave_nz = __builtin_amdgcn_ballot_w64(sp.z != (R)0 || tgt_nz) != 0;
        if constexpr (kPacked) {
          planar = !wave_nz;
          if (planar) sp.z = wsrc;
        } else {
          s_nz[tid / 64] = wave_nz ? 1 : 0;
        }
      }
      s_p[tid] = sp;
      if constexpr (!kPacked) s_w[tid] = wsrc;
      else if (!planar) s_w[tid] = wsrc;
    }
    if (tid >= cnt && tid < cntp) {
      // (packed or not, the padding's weight is 0 and its z is 0)
      s_p[tid] = SrcPt<R>{(R)0, (R)0, (R)0, (R)-INFINITY};
      s_w[tid] = (R)0;
    }
    if (off + TILE < span) fetch(off + TILE);
    group_sync<NTHR>();
    if constexpr (kPlanarOk && !kPacked) {
      int any = 0;
#pragma unroll
      for (int q = 0; q < NTHR / 64; ++q) any |= s_nz[q];
      planar = __builtin_amdgcn_readfirstlane(any) == 0;
    }

    if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
      for (int j = g; j < cnt; j += G) {
        const SrcPt<R> sp = s_p[j];
        const R sx = sp.x, sy = sp.y, sz = sp.z, sa = sp.a;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const R C = pair_cost<COST>(tx[r], ty[r], tz[r], sx, sy, sz, pexp);
          if constexpr (KIND == KIND_MAX) st[r].acc = st[r].acc > C ? st[r].acc : C;
          else {
            const R v = C + sa;
            st[r].acc = st[r].acc < v ? st[r].acc : v;
          }
        }
      }
    } else {
      // lane g takes sources j0 + u*G + g: the G lanes of a target read G
      // consecutive LDS words per step
      auto tile_pairs = [&](auto planar_tag, auto cost_tag) {
        constexpr bool kPlanar = decltype(planar_tag)::value;
        constexpr bool kNeedC = decltype(cost_tag)::value;
        for (int j0 = 0; j0 < cntp; j0 += kSub * G) {
          R v[RPT][kSub];
          R Cv[RPT][kSub];
          R wv[kSub], px[kSub], py[kSub], pz[kSub];
#pragma unroll
          for (int u = 0; u < kSub; ++u) {
            const int j = j0 + u * G + g;
            const SrcPt<R> sp = s_p[j];
            const R sx = sp.x, sy = sp.y;
            const R sz = kPacked && kPlanar ? (R)0 : sp.z;
            const R sa = sp.a;
            wv[u] = kPacked && kPlanar ? sp.z : s_w[j];
            px[u] = sx; py[u] = sy; pz[u] = sz;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
              const R dr = tx[r] - sx, dg = ty[r] - sy;
              if constexpr (kScaled) {
                Cv[r][u] = (R)0;
                const R a0 = sa + tqm[r];
                if constexpr (kPlanar) {
                  v[r][u] = fma(-dr, dr, fma(-dg, dg, a0));
                } else {
                  const R db = tz[r] - sz;
                  v[r][u] = fma(-dr, dr, fma(-dg, dg, fma(-db, db, a0)));
                }
              } else {
                const R C = kPlanar ? fma(dr, dr, dg * dg)
                                    : pair_cost<COST>(tx[r], ty[r], tz[r], sx, sy, sz, pexp);
                Cv[r][u] = C;
                v[r][u] = fma(k, C, sa + (kLazy ? tqm[r] : tq[r]));
              }
            }
          }
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            if constexpr (kLazy) lse_lazy<kSub>(st[r], v[r], wv, tq[r], tqm[r], dl[r]);
            else accumulate<PREC, KIND, kSub, R, kNeedC, true>(st[r], v[r], wv, Cv[r], px, py, pz);
          }
        }
      };
      // A marginal pass accumulates the cost sum only when its cost output is
      // asked for (the sum feeds nothing else; the marginals keep their bits).
      auto with_cost = [&](auto planar_tag) {
        if constexpr (KIND == KIND_SUM) {
          if (d.out1 != nullptr) tile_pairs(planar_tag, std::true_type{});
          else tile_pairs(planar_tag, std::false_type{});
        } else {
          tile_pairs(planar_tag, std::true_type{});
        }
      };
      if constexpr (kPlanarOk) {
        if (planar) with_cost(std::true_type{});
        else with_cost(std::false_type{});
      } else {
        with_cost(std::false_type{});
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r) flush_blocks<PREC, KIND, R>(st[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    if constexpr (kLazy) st[r].S *= exp2_small(-dl[r]);
    combine_lanes<PREC, KIND, G, R>(st[r]);
    if (g != 0) tidx[r] = none;  // one owner lane per target
  }
}

template <int PREC, int KIND, int COST, int RPT, bool REC = false>
__device__ __forceinline__ void pair_body(const PassDesc& d, const Work& w, const int64_t tb,
                                          const int split, const int splits,
                                          const int64_t split_len, const int64_t ntp) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  constexpr int TPB = kThreads / kLanesPerTarget;
  __shared__ SrcPt<R> s_p[kTile];
  __shared__ R s_w[kTile];
  const int64_t s0 = (int64_t)split * split_len;
  const int64_t s1 = min(d.ns, s0 + split_len);
  int64_t tidx[RPT];
  TState<R> st[RPT];
  pair_accum<PREC, KIND, COST, RPT, kThreads, kTile, R, REC>(d, s_p, s_w, threadIdx.x,
                                                              tb * TPB * RPT, s0, s1, ntp, tidx, st);
  finish_pass<PREC, KIND, RPT, R>(d, w, splits, split, ntp, tidx, st);
}

// Workgroup-wide splits: a workgroup of kWideWaves waves owns kWideTargets
// targets (one wave's worth at kLanesPerTarget lanes per target); wave q takes
// source split q, so the kWideWaves splits of a target are merged in LDS by
// the same workgroup, with no partials in HBM and no second launch.  Each wave
// stages its own split's sources.  The per-split states and their merge are
// the split path's (pair_body with the same split_len, then merge_body), so
// both give identical results.
template <int PREC, int KIND, int COST, bool REC = false>
__global__ __launch_bounds__(kWideThreads) void pair_pass_wide(const PassDesc din,
                                                               const int64_t split_len) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  constexpr int NF = kind_fields<KIND>();
  constexpr int TILE = 64;  // one staged source per lane and tile
  __shared__ SrcPt<R> s_p[kWideWaves][TILE];
  __shared__ R s_w[kWideWaves][TILE];
  __shared__ double s_part[NF][kWideWaves][kWideTargets];
  const PassDesc d = resolve_desc(din);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t t0 = (int64_t)blockIdx.x * kWideTargets;
  const int64_t s0 = (int64_t)wave * split_len;
  const int64_t s1 = min(d.ns, s0 + split_len);
  const int nsplit = (int)((d.ns + split_len - 1) / split_len);
  const int64_t none = -1;
  // The merge lane of each target (below) loads the target's offset and
  // epilogue inputs now, so that their round trip overlaps the pair loop
  // instead of following the merge.
  const int sub = threadIdx.x % kMergeLanes, tl = threadIdx.x / kMergeLanes;
  const int64_t t = t0 + tl;
  const bool own = threadIdx.x < kWideTargets * kMergeLanes && sub == 0 && t < d.nt;
  TgtLoad L{};
  EpiLoad E{};
  if (own) {
    L = load_offset_inputs<PREC, KIND>(d, t);
    E = load_epilogue_inputs<PREC, KIND>(d, t);
  }
  int64_t tidx[1];
  TState<R> st[1];
  pair_accum<PREC, KIND, COST, 1, 64, TILE, R, REC>(d, s_p[wave], s_w[wave], lane, t0, s0, s1,
                                                     none, tidx, st);
  if (tidx[0] != none) {
    const int tp = (int)(tidx[0] - t0);
    double f[6];
    to_fields<KIND, R>(st[0], f);
#pragma unroll
    for (int q = 0; q < NF; ++q) s_part[q][wave][tp] = f[q];
  }
  __syncthreads();
  // merge_body over the LDS partials, by the first wave: kMergeLanes lanes
  // per target, lane `sub` folding splits sub, sub + kMergeLanes, ... in order
  if (threadIdx.x >= kWideTargets * kMergeLanes) return;
  const bool ok = t < d.nt;
  TState<R> acc;
  acc.m = empty_max<PREC, R>();
  acc.S = acc.Sc = acc.Sy0 = acc.Sy1 = acc.Sy2 = 0.0;
  acc.acc = KIND == KIND_MIN ? (R)INFINITY : (R)-INFINITY;
  if (ok) {
    for (int q = sub; q < nsplit; q += kMergeLanes) {
      TState<R> p = acc;
      double f[6];
#pragma unroll
      for (int fi = 0; fi < NF; ++fi) f[fi] = s_part[fi][q][tl];
      from_fields<KIND, R>(p, f);
      combine_state<PREC, KIND, R>(acc, p);
    }
  }
  combine_lanes<PREC, KIND, kMergeLanes, R>(acc);
  if (!own) return;
  const double q = offset_from<PREC, KIND>(d, L);
  acc.qd = PREC == 1 ? (double)(float)(q * kLog2e) : q;
  epilogue_from<PREC, KIND, R>(d, t, acc, E);
}

template <int PREC, int KIND, int COST, int RPT, bool REC = false>
__global__ __launch_bounds__(kThreads) void pair_pass(const PassDesc din, const Work w,
                                                      const int splits,
                                                      const int64_t split_len) {
  const PassDesc d = resolve_desc(din);
  pair_body<PREC, KIND, COST, RPT, REC>(d, w, blockIdx.x, blockIdx.y, splits, split_len,
                                        (int64_t)gridDim.x * (kThreads / kLanesPerTarget) * RPT);
}

template <int PREC, int KIND, int COST, int RPT>
__global__ __launch_bounds__(kThreads) void pair_pass_batch(const PassDesc* descs,
                                                            const BatchItem* items,
                                                            const BatchInst* insts,
                                                            const Work w) {
  const BatchItem it = items[blockIdx.x];
  if (it.inst < 0) return;  // padding of an XCD's item sequence
  const BatchInst bi = insts[it.inst];
  const PassDesc d = resolve_desc(descs[it.inst]);
  const Work wi{w.partial + bi.pbase};
  pair_body<PREC, KIND, COST, RPT>(d, wi, it.tb, it.split, bi.splits, bi.split_len, bi.ntp);
}

// ---- feature (cell-similarity) costs ------------------------------------------
// raw_row_dot_cuda's `dot += x * y` (src/CellUtilities.jl:202-208): the product
// is rounded before the sum (no fused multiply-add), so that the dot of a row
// with itself equals its squared norm exactly.
__device__ __forceinline__ double dot_step(double acc, double x, double y) {
#pragma clang fp contract(off)
  return acc + x * y;
}

// Cost from the accumulated statistic `acc` (sum |x~ - y~| of the rows divided
// by their sums for l1, the raw dot x . y otherwise) and the two rows'
// constants, with feature_cost_cuda's operations (src/CellUtilities.jl:
// 211-282): dot = acc / (s_i s_j), then the metric.
template <int METRIC, typename R>
__device__ __forceinline__ R feature_cost(R acc, R ni, R mi, R ci, R si, R nj, R mj, R cj, R sj,
                                          R dim, R scale) {
  // each operation rounded on its own, as the reference evaluates it
#pragma clang fp contract(off)
  if constexpr (METRIC == METRIC_L1) {
    return acc / scale;
  } else {
    const R dot = acc / (si * sj);
    if constexpr (METRIC == METRIC_L2) {
      const R v = ni + nj - (R)2 * dot;
      return sqrt(v > (R)0 ? v : (R)0) / scale;
    } else if constexpr (METRIC == METRIC_COSINE) {
      const R den = sqrt(ni * nj);
      if (!(den > (R)0)) return (R)0;
      R sim = dot / den;
      sim = sim < (R)-1 ? (R)-1 : (sim > (R)1 ? (R)1 : sim);
      return ((R)1 - sim) / scale;
    } else {
      const R den = sqrt(ci * cj);
      if (den <= (R)0) {
        const R dm = mi - mj;
        return (dm < (R)0 ? -dm : dm) <= (R)1.4901161193847656e-08 ? (R)0 : (R)1 / scale;
      }
      R cor = (dot - dim * mi * mj) / den;
      cor = cor < (R)-1 ? (R)-1 : (cor > (R)1 ? (R)1 : cor);
      return ((R)1 - cor) / scale;
    }
  }
}

template <int PREC, int KIND, int METRIC>
__device__ __forceinline__ void feature_body(const PassDesc& d, const Work& w, const int64_t tb,
                                             const int split, const int splits,
                                             const int64_t split_len, const int64_t ntp) {
  using R = typename std::conditional<PREC == 0, double, float>::type;
  constexpr int TT = kFeatTargets, ST = 64, KC = 16, RED = 17;
  constexpr bool kExpKind = (KIND == KIND_LSE || KIND == KIND_SUM);
  static_assert(KIND != KIND_BARY, "no barycentric pass for feature costs");
  __shared__ R Xs[KC][TT];
  __shared__ R Ys[KC][ST];
  __shared__ R s_a[ST], s_w[ST], s_n[ST], s_m[ST], s_c[ST], s_s[ST];
  __shared__ int s_valid[ST];
  __shared__ double red0[TT][RED], red1[TT][RED];
  // (compensation terms of the sums in red1 / red0 for the exp kinds)
  __shared__ double red2[kExpKind ? TT : 1][RED], red3[KIND == KIND_SUM ? TT : 1][RED];

  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  const int64_t s0 = (int64_t)split * split_len;
  const int64_t s1 = min(d.ns, s0 + split_len);
  const R k = PREC == 0 ? (R)d.k : (R)(d.k * kLog2e);
  const R dimR = (R)d.dim, scale = (R)d.fscale;
  const R* X = PREC == 0 ? (const R*)d.tf64 : (const R*)d.tf32;
  const R* Y = PREC == 0 ? (const R*)d.sf64 : (const R*)d.sf32;

  R tn[4], tm[4], tcn[4], ts[4], tq[4];
  TState<R> st[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t t = tb * TT + ty * 4 + a;
    const bool ok = t < d.nt;
    tn[a] = ok ? (R)d.trc[t] : (R)0;
    tm[a] = ok ? (R)d.trc[d.nt + t] : (R)0;
    tcn[a] = ok ? (R)d.trc[2 * d.nt + t] : (R)0;
    ts[a] = ok ? (R)d.trc[3 * d.nt + t] : (R)1;
    const double q = ok ? target_offset<PREC, KIND>(d, t) : 0.0;
    if (PREC == 1) {
      const float q2 = (float)(q * kLog2e);
      tq[a] = (R)q2;
      st[a].qd = (double)q2;
    } else {
      tq[a] = (R)q;
      st[a].qd = q;
    }
    st[a].m = empty_max<PREC, R>();
    st[a].S = st[a].Sc = st[a].Sy0 = st[a].Sy1 = st[a].Sy2 = 0.0;
    // max_logsumexp_cellsim! starts its running max at 0 (MirrorProxCellSim.jl:21)
    st[a].acc = KIND == KIND_MIN ? (R)INFINITY : (R)0;
  }

  for (int64_t base = s0; base < s1; base += ST) {
    const int cnt = (int)min((int64_t)ST, s1 - base);
    R acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = (R)0;
    for (int64_t k0 = 0; k0 < d.dim; k0 += KC) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < (KC * TT) / kThreads; ++q) {
        const int e = tid + q * kThreads;
        const int kk = e / TT, tt = e % TT;
        const int64_t kd = k0 + kk;
        const int64_t t = tb * TT + tt;
        const int64_t s = base + tt;
        Xs[kk][tt] = (t < d.nt && kd < d.dim) ? X[t + kd * d.nt] : (R)0;
        Ys[kk][tt] = (tt < cnt && kd < d.dim) ? Y[s + kd * d.ns] : (R)0;
      }
      if (k0 == 0 && tid < ST) {
        const int64_t s = base + tid;
        if (tid < cnt) {
          source_terms<PREC, kExpKind>(d, s, &s_a[tid], &s_w[tid]);
          s_n[tid] = (R)d.src_rc[s];
          s_m[tid] = (R)d.src_rc[d.ns + s];
          s_c[tid] = (R)d.src_rc[2 * d.ns + s];
          s_s[tid] = (R)d.src_rc[3 * d.ns + s];
          s_valid[tid] = 1;
        } else {
          s_a[tid] = (R)-INFINITY;
          s_w[tid] = (R)0;
          s_n[tid] = s_m[tid] = s_c[tid] = (R)0;
          s_s[tid] = (R)1;
          s_valid[tid] = 0;
        }
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < KC; ++kk) {
        R xa[4], yb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) xa[a] = Xs[kk][ty * 4 + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) yb[b] = Ys[kk][tx * 4 + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            if constexpr (METRIC == METRIC_L1) {
              const R df = xa[a] - yb[b];
              acc[a][b] += df < (R)0 ? -df : df;
            } else if constexpr (PREC == 0) {
              acc[a][b] = dot_step(acc[a][b], xa[a], yb[b]);
            } else {
              acc[a][b] = fma(xa[a], yb[b], acc[a][b]);
            }
          }
      }
    }
    // costs and exponents of the 4 x 4 pairs of this thread
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      R v[4], wv[4], Cv[4], dz[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = tx * 4 + b;
        const R C = feature_cost<METRIC, R>(acc[a][b], tn[a], tm[a], tcn[a], ts[a], s_n[j],
                                            s_m[j], s_c[j], s_s[j], dimR, scale);
        const bool valid = s_valid[j] != 0;
        Cv[b] = C;
        wv[b] = s_w[j];
        dz[b] = (R)0;
        if constexpr (KIND == KIND_MAX) {
          if (valid) st[a].acc = st[a].acc > C ? st[a].acc : C;
        } else if constexpr (KIND == KIND_MIN) {
          if (valid) {
            const R vv = C + s_a[j];
            st[a].acc = st[a].acc < vv ? st[a].acc : vv;
          }
        } else {
          v[b] = valid ? fma(k, C, s_a[j] + tq[a]) : (R)-INFINITY;
        }
      }
      if constexpr (kExpKind) accumulate<PREC, KIND, 4, R>(st[a], v, wv, Cv, dz, dz, dz);
    }
  }

  // ---- combine the 16 column threads of each target (fixed order) ----------
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int tt = ty * 4 + a;
    if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
      red0[tt][tx] = (double)st[a].acc;
    } else if constexpr (KIND == KIND_SUM) {
      red0[tt][tx] = st[a].S;
      red1[tt][tx] = st[a].Sc;
      red2[tt][tx] = st[a].Se;
      red3[tt][tx] = st[a].Sce;
    } else {
      red0[tt][tx] = (double)st[a].m;
      red1[tt][tx] = st[a].S;
      red2[tt][tx] = st[a].Se;
    }
  }
  __syncthreads();
  int64_t tidx[1];
  TState<R> fin[1];
  tidx[0] = tid < TT ? tb * TT + tid : ntp;
  fin[0].m = empty_max<PREC, R>();
  fin[0].S = fin[0].Sc = fin[0].Sy0 = fin[0].Sy1 = fin[0].Sy2 = 0.0;
  fin[0].acc = (R)0;
  fin[0].qd = 0.0;
  if (tid < TT) {
    // the target's offset: the same value every column thread used
    const int64_t t = tb * TT + tid;
    const double q = t < d.nt ? target_offset<PREC, KIND>(d, t) : 0.0;
    fin[0].qd = PREC == 1 ? (double)(float)(q * kLog2e) : q;
    if constexpr (KIND == KIND_MAX || KIND == KIND_MIN) {
      double a = red0[tid][0];
      for (int q2 = 1; q2 < 16; ++q2) a = KIND == KIND_MAX ? fmax(a, red0[tid][q2]) : fmin(a, red0[tid][q2]);
      fin[0].acc = (R)a;
    } else if constexpr (KIND == KIND_SUM) {
      double a = 0.0, c = 0.0, ae = 0.0, ce = 0.0;
      for (int q2 = 0; q2 < 16; ++q2) {
        ae += red2[tid][q2];
        two_sum_acc(a, ae, red0[tid][q2]);
        ce += red3[tid][q2];
        two_sum_acc(c, ce, red1[tid][q2]);
      }
      fin[0].S = a;
      fin[0].Sc = c;
      fin[0].Se = ae;
      fin[0].Sce = ce;
    } else {
#pragma clang fp contract(off)
      double M = red0[tid][0];
      for (int q2 = 1; q2 < 16; ++q2) M = fmax(M, red0[tid][q2]);
      double a = 0.0, ae = 0.0;
      for (int q2 = 0; q2 < 16; ++q2) {
        const double f = PREC == 1 ? ldexp(1.0, (int)fmax(red0[tid][q2] - M, -2000.0))
                                   : exp(red0[tid][q2] - M);
        ae += f * red2[tid][q2];
        two_sum_acc(a, ae, f * red1[tid][q2]);
      }
      fin[0].m = (R)M;
      fin[0].S = a;
      fin[0].Se = ae;
    }
  }
  finish_pass<PREC, KIND, 1, R>(d, w, splits, split, ntp, tidx, fin);
}

template <int PREC, int KIND, int METRIC>
__global__ __launch_bounds__(kThreads) void feature_pass(const PassDesc din, const Work w,
                                                         const int splits,
                                                         const int64_t split_len) {
  const PassDesc d = resolve_desc(din);
  feature_body<PREC, KIND, METRIC>(d, w, blockIdx.x, blockIdx.y, splits, split_len,
                                   (int64_t)gridDim.x * kFeatTargets);
}

template <int PREC, int KIND, int METRIC>
__global__ __launch_bounds__(kThreads) void feature_pass_batch(const PassDesc* descs,
                                                               const BatchItem* items,
                                                               const BatchInst* insts,
                                                               const Work w) {
  const BatchItem it = items[blockIdx.x];
  if (it.inst < 0) return;  // padding of an XCD's item sequence
  const BatchInst bi = insts[it.inst];
  const PassDesc d = resolve_desc(descs[it.inst]);
  const Work wi{w.partial + bi.pbase};
  feature_body<PREC, KIND, METRIC>(d, wi, it.tb, it.split, bi.splits, bi.split_len, bi.ntp);
}

template <int PREC, int KIND, int METRIC>
hipError_t launch_feature_t(const PassDesc& d, const LaunchCfg& c, const Work& w, hipStream_t s) {
  dim3 grid((unsigned)c.tblocks, (unsigned)c.splits);
  hipLaunchKernelGGL((feature_pass<PREC, KIND, METRIC>), grid, dim3(kThreads), 0, s, d, w,
                     c.splits, c.split_len);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_merge<PREC, KIND>(d, c, w, c.tblocks * kFeatTargets, s);
}

template <int PREC, int KIND>
hipError_t launch_feature_metric(const PassDesc& d, const LaunchCfg& c, const Work& w, hipStream_t s) {
  switch (d.metric) {
    case METRIC_L1: return launch_feature_t<PREC, KIND, METRIC_L1>(d, c, w, s);
    case METRIC_L2: return launch_feature_t<PREC, KIND, METRIC_L2>(d, c, w, s);
    case METRIC_COSINE: return launch_feature_t<PREC, KIND, METRIC_COSINE>(d, c, w, s);
    default: return launch_feature_t<PREC, KIND, METRIC_CORRELATION>(d, c, w, s);
  }
}


// ---- MFMA feature kernel (cosine / correlation on unit rows) ------------------
// Rows are packed as z_i = sign(s_i) x_i / |x_i| (cosine) or
// sign(s_i) (x_i - mean_i) / |x_i - mean_i| (correlation), so that the
// reference's similarity (CellUtilities.jl:245-282) is the dot z_i . z_j,
// clamped to [-1, 1].  They are stored as bf16 rows (row-major, width
// 16*KS, zero padded), optionally with a bf16 remainder row z - bf16(z) for
// the three-product split (NSPLIT == 3).
//
// One wave owns 32 targets: their rows are the B operand of
// v_mfma_f32_32x32x16_bf16 and stay in registers.  32-source tiles are
// staged in LDS as the A operand and shared by the 4 waves.  The 32x32
// result is laid out so that lane l holds 16 sources of target l % 32, hence
// the log-sum-exp runs per lane; lanes l and l + 32 merge once at the end.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

template <int KIND, int NSPLIT, int KS>
__device__ __forceinline__ void mfma_body(const PassDesc& d, const Work& w, const int64_t tb,
                                          const int split, const int splits,
                                          const int64_t split_len, const int64_t ntp) {
  constexpr int DP = 16 * KS;          // padded row width (bf16)
  constexpr int CPR = 2 * KS;          // 16-byte chunks per row
  constexpr int SUB = 32;              // sources per MFMA tile
  // sources staged per workgroup barrier (two MFMA tiles): 2 x 16 KB of
  // rows per workgroup at d = 128, so that LDS leaves room for three
  // workgroups per CU
  constexpr int TILE = 64;
  // Double-buffered staging: the rows of tile T + 1 go from global memory
  // straight into the other LDS buffer (buffer loads to LDS, no registers) while
  // tile T is computed, and the tile's one barrier waits for them.  Rows are
  // unpadded with their 16-byte chunks rotated by an XOR with the row index
  // (chunk c of row r sits at position c ^ (r mod CPR)), so the MFMA operand
  // reads of 16 consecutive rows hit distinct banks.  A partial tile repeats
  // its last row, whose copies carry the exponent term -inf.
  // (the two buffers are separate arrays, so that the compiler sees that
  // the copies into one never alias the operand reads from the other and
  // lets them stay in flight)
  __shared__ __align__(16) __bf16 Ah0[TILE * DP], Ah1[TILE * DP];
  __shared__ __align__(16) __bf16 Al0[NSPLIT == 3 ? TILE * DP : 8], Al1[NSPLIT == 3 ? TILE * DP : 8];
  __shared__ __align__(16) float s_a[2][TILE];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t s0 = (int64_t)split * split_len;
  const int64_t s1 = min(d.ns, s0 + split_len);
  const __bf16* Th = d.tz_hi;
  const __bf16* Tl = d.tz_lo;
  const __bf16* Sh = d.sz_hi;
  const __bf16* Sl = d.sz_lo;
  // exponent: v = k C + a_s + q_t with C = (1 - sim) / scale
  //           = kc * sim + (a_s + k / scale) + q_t,  kc = -k / scale
  // The rows are scaled by kUnitRowScale, so the products give
  // kUnitRowScale^2 sim: the exponent factor and the cost divide it out.
  const float kc = (float)(-d.k * kLog2e / d.fscale / kUnitRowScale2);
  const float inv_scale = (float)(1.0 / d.fscale);
  const float inv_scale_s = (float)(1.0 / (d.fscale * kUnitRowScale2));
  const float s2 = (float)kUnitRowScale2;

  const int jt = lane & 31;
  const int64_t t = tb * 128 + wave * 32 + jt;
  const bool tok = t < d.nt;
  bf16x8_t bh[KS], bl[NSPLIT == 3 ? KS : 1];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k0 = ks * 16 + 8 * (lane >> 5);
    if (tok) {
      bh[ks] = *(const bf16x8_t*)(Th + t * DP + k0);
      if constexpr (NSPLIT == 3) bl[ks] = *(const bf16x8_t*)(Tl + t * DP + k0);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) bh[ks][e] = (__bf16)0.0f;
      if constexpr (NSPLIT == 3) {
#pragma unroll
        for (int e = 0; e < 8; ++e) bl[ks][e] = (__bf16)0.0f;
      }
    }
  }
  TState<float> st;
  {
    const double q = tok ? target_offset<1, KIND>(d, t) : 0.0;
    const float q2 = (float)(q * kLog2e);
    st.qd = (double)q2;
    st.m = KIND == KIND_LSE ? 0.0f : empty_max<1, float>();
    st.S = st.Sc = st.Sy0 = st.Sy1 = st.Sy2 = 0.0;
    st.acc = 0.0f;
  }
  const float tq = (float)st.qd;
  float tqm = tq;    // LSE: exponent offset relative to the lazily moved reference
  double dl = 0.0;
  // The target's exponent offset rides in the MFMA accumulator: the products
  // start from tvec = tinit = offset / kc (the lane's 16 accumulator entries
  // all belong to its target), so acc = s2 sim + tinit and the exponent is
  // fma(kc, acc, a_s): one addition and no clamp per pair (s2 sim lies in
  // [-s2, s2] by the row scaling; the cost, where a clamp of sim to [-1, 1]
  // is part of the reference's definition, uses med3(acc, tinit - s2,
  // tinit + s2)).  The offset actually applied is kc * tinit, exact in FP64
  // (eoff); the LSE's dl and the sums' final rescale account for its
  // difference from the requested offset.  (kc == 0, an infinite
  // regularisation, keeps the offset out of the accumulator; so does a tiny
  // |kc|, whose tinit could overflow.)
  //
  // tvec is set once per work item: a rescale (lse_lazy_folded) moves tinit
  // and eoff, and the later sub-tiles of that lane add eoff - ebuilt to their
  // exponents (rescales are rare once the offset centres the exponents), so
  // that the accumulator seed is loop-invariant and never copied.
  const bool fold = fabsf(kc) >= 0x1p-40f;
  float tinit = 0.0f, lo = -s2, hi = s2;
  double eoff = 0.0, ebuilt = 0.0;
  auto set_offset = [&](const float off) {
    tinit = off / kc;
    lo = tinit - s2;
    hi = tinit + s2;
    eoff = (double)kc * (double)tinit;
  };
  if (fold) {
    set_offset(tq);
    ebuilt = eoff;
    if constexpr (KIND == KIND_LSE) dl = eoff - (double)tq;  // (reference m = 0)
  }
  f32x16_t tvec;
#pragma unroll
  for (int r = 0; r < 16; ++r) tvec[r] = fold ? tinit : 0.0f;

  // lse_lazy with the offset in the accumulator: the sub-tile's exponents v
  // carry the effective offset eoff (S holds sum 2^(v_true + tq - m + dl),
  // dl = eoff - (tq - m) exactly); when the sum leaves [2^-90, 2^100] the
  // reference m moves as in lse_lazy, tinit follows it and the sub-tile is
  // re-evaluated with the new offset.
  auto lse_lazy_folded = [&](const float (&v)[16]) {
    float loc = exp2_fast(v[0]);
#pragma unroll
    for (int u = 1; u < 16; ++u) loc += exp2_fast(v[u]);
    const double Sn = st.S + (double)loc;
    if (loc < 0x1p100f && Sn > 0x1p-90) {
      st.S = Sn;
      return;
    }
    float tm = v[0];
#pragma unroll
    for (int u = 1; u < 16; ++u) tm = fmaxf(tm, v[u]);
    int es = 0;
    if (st.S > 0.0) (void)frexp(st.S, &es);
    const float sl = st.S > 0.0 ? (float)es : -INFINITY;
    const float top = fmaxf(tm, sl);
    if (!(top > -INFINITY) || !(top < INFINITY)) {
      st.S = Sn;
      return;
    }
    const float c = ceilf(top);
    const float mnew = st.m + c;
    const double e_old = eoff;
    set_offset(tq - mnew);
    const double dnew = eoff - ((double)tq - (double)mnew);
    const float dv = (float)(eoff - e_old);
    st.S = ldexp(st.S, -(int)c) * exp2_small(dnew - dl);
    loc = 0.0f;
#pragma unroll
    for (int u = 0; u < 16; ++u) loc += exp2_fast(v[u] + dv);
    st.S += (double)loc;
    st.m = mnew;
    dl = dnew;
  };

  // ---- staging ----
  // Source term per source: the FP64 exponent term a_s (with the constant
  // k / scale folded in) plus log2 of its weight, rounded once to FP32, so
  // the exp-sum needs no weight: 2^(a'' + ...) = b_s 2^(a_s + ...) within
  // 2^-24 |a''| in the exponent, far below the bf16 products' error.
  double sa_raw = 0.0, sb_raw = 1.0;
  // (global address space: these loads cannot alias the LDS the row copies
  // in flight write, so they need not wait for them)
  typedef const __attribute__((address_space(1))) double* gdp;
  auto load_terms = [&](const int64_t base) {
    const int cnt = (int)min((int64_t)TILE, s1 - base);
    const bool real = tid < cnt;
    sa_raw = (real && d.srcA != nullptr) ? *(gdp)(d.srcA + base + tid) : 0.0;
    sb_raw = (real && d.srcB != nullptr) ? *(gdp)(d.srcB + base + tid) : 1.0;
  };
  auto store_terms = [&](const int64_t base, const int buf) {
    const int cnt = (int)min((int64_t)TILE, s1 - base);
    if (tid < TILE) {
      float a = -INFINITY;
      if (tid < cnt && sb_raw > 0.0) {
        int e2 = 0;
        const double mant = frexp(sb_raw, &e2);  // (no FP32 underflow of tiny weights)
        const double full = (d.srcA != nullptr ? d.alpha * sa_raw : 0.0) + d.k / d.fscale;
        a = (float)(full * kLog2e + (double)e2 + (double)__log2f((float)mant));
      }
      s_a[buf][tid] = a;
    }
  };
  // issues the row loads of the tile at `base` into buffer `buf`: wave w
  // fills the 1 KB blocks w, w + 4, ... (64 lanes x 16 bytes each); lane l
  // of block i brings chunk c = (e mod CPR) ^ (row mod CPR) of row
  // e / CPR, e = 64 i + l, to byte 16 e of the buffer
  const unsigned wv_u = __builtin_amdgcn_readfirstlane(wave);
  constexpr int BLOCKS = TILE * CPR / 64;                      // 1 KB blocks per tile
  constexpr int NB = (BLOCKS + kThreads / 64 - 1) / (kThreads / 64);  // blocks per wave
  unsigned roff[NB];  // byte offset of each block's chunk within a full tile
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const unsigned e = (4u * j + wv_u) * 64u + (unsigned)lane;
    const unsigned row = e / CPR, c = (e % CPR) ^ (row % CPR);
    roff[j] = row * (unsigned)(DP * 2) + c * 16u;
  }
  // The copies are raw buffer loads to LDS: the tile's row base is the
  // resource (scalar), the lane's 32-bit offset the only vector operand.
  constexpr int kRowBytes = DP * 2;
  auto load_rows = [&](const int64_t base, const int buf) {
    const int cnt = (int)min((int64_t)TILE, s1 - base);
    const char* th = (const char*)(Sh + base * DP);
    const char* tl = NSPLIT == 3 ? (const char*)(Sl + base * DP) : nullptr;
    typedef __attribute__((address_space(3))) void* lds_t;
    if (cnt == TILE) {
      // full tile: the precomputed lane offsets as they are (no VALU)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (BLOCKS % (kThreads / 64) != 0 && 4 * j + (int)wv_u >= BLOCKS) break;
        const int blk = (4 * j + (int)wv_u) * 512;  // bf16 elements per 1 KB block
        lds_copy16(th, TILE * kRowBytes, roff[j], (lds_t)(&(buf ? Ah1 : Ah0)[blk]));
        if constexpr (NSPLIT == 3)
          lds_copy16(tl, TILE * kRowBytes, roff[j], (lds_t)(&(buf ? Al1 : Al0)[blk]));
      }
      return;
    }
    // a partial tile repeats its last row (every offset stays below cnt rows)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (BLOCKS % (kThreads / 64) != 0 && 4 * j + (int)wv_u >= BLOCKS) break;
      unsigned off = roff[j];
      const unsigned row = off / (unsigned)kRowBytes;
      if (row >= (unsigned)cnt) off += ((unsigned)cnt - 1u - row) * (unsigned)kRowBytes;
      const int blk = (4 * j + (int)wv_u) * 512;
      lds_copy16(th, cnt * kRowBytes, off, (lds_t)(&(buf ? Ah1 : Ah0)[blk]));
      if constexpr (NSPLIT == 3)
        lds_copy16(tl, cnt * kRowBytes, off, (lds_t)(&(buf ? Al1 : Al0)[blk]));
    }
    // (an empty statement the compiler cannot move: it keeps these copies
    // from being merged with the full tile's into per-lane selects)
    __asm__ volatile("");
  };

  const int ntiles = s1 > s0 ? (int)((s1 - s0 + TILE - 1) / TILE) : 0;
  if (ntiles > 0) {
    load_rows(s0, 0);
    load_terms(s0);
    store_terms(s0, 0);
  }
  __syncthreads();
  // One tile: the buffer index and the sub-tile offsets are compile-time
  // constants (the tile loop below is unrolled by the two buffers), so the
  // LDS operand reads carry them as immediate offsets and the loop spends no
  // VALU on addresses.
  static_assert(SUB % CPR == 0, "row rotation repeats within a sub-tile");
  // lane's element offset of its operand chunk ks in row jt of a tile
  int aoff[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) aoff[ks] = jt * DP + ((2 * ks + (lane >> 5)) ^ (jt & (CPR - 1))) * 8;
  auto tile = [&](const int64_t base, auto bufc, const bool more) {
    constexpr int buf = decltype(bufc)::value;
    if (more) {  // in flight during this tile's work (the terms first, so
                 // that their wait at the tile's end leaves the rows in flight)
      load_terms(base + TILE);
      load_rows(base + TILE, buf ^ 1);
    }
    const int nsub = ((int)min((int64_t)TILE, s1 - base) + SUB - 1) / SUB;
#pragma unroll
    for (int sub = 0; sub < TILE / SUB; ++sub) {
      if (sub >= nsub) break;
      const int r0 = sub * SUB;
      f32x16_t acc = tvec;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int at = r0 * DP + aoff[ks];
        const bf16x8_t ah = *(const bf16x8_t*)(&(buf ? Ah1 : Ah0)[at]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[ks], acc, 0, 0, 0);
        if constexpr (NSPLIT == 3) {
          const bf16x8_t al = *(const bf16x8_t*)(&(buf ? Al1 : Al0)[at]);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[ks], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[ks], acc, 0, 0, 0);
        }
      }
      // the lane's 16 sources are 4 runs of 4 consecutive rows: i = 8 q + 4 (lane / 32) + e;
      // padded sources carry a_s = -inf, w_s = 0
      float sa[16], sw[16];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i0 = r0 + 8 * q + 4 * (lane >> 5);
        const float4 a4 = *(const float4*)(&s_a[buf][i0]);
        sa[4 * q] = a4.x; sa[4 * q + 1] = a4.y; sa[4 * q + 2] = a4.z; sa[4 * q + 3] = a4.w;
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) sw[r] = 1.0f;  // (the weights are in the terms)
      if constexpr (KIND == KIND_LSE) {
        float v[16];
        if (fold) {
#pragma unroll
          for (int r = 0; r < 16; ++r) v[r] = fmaf(kc, acc[r], sa[r]);
          if (eoff != ebuilt) {  // (after a rescale of this lane)
            const float dv = (float)(eoff - ebuilt);
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] += dv;
          }
          lse_lazy_folded(v);
        } else {
#pragma unroll
          for (int r = 0; r < 16; ++r) v[r] = fmaf(kc, acc[r], sa[r] + tqm);
          lse_lazy<16>(st, v, sw, tq, tqm, dl);
        }
      } else if (d.out1 != nullptr) {
        // C = (1 - sim) / scale with sim = (med3(acc, lo, hi) - tinit) / s2,
        // i.e. the reference's clamp of sim to [-1, 1]
        const float c1 = fmaf(tinit, inv_scale_s, inv_scale);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          float v[8], wv[8], Cv[8], dz[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int r = h * 8 + u;
            const float simc = __builtin_amdgcn_fmed3f(acc[r], lo, hi);
            Cv[u] = fmaf(-simc, inv_scale_s, c1);
            wv[u] = sw[r];
            dz[u] = 0.0f;
            v[u] = fold ? fmaf(kc, acc[r], sa[r]) : fmaf(kc, acc[r], sa[r] + tq);
          }
          accumulate<1, KIND, 8, float, true, true>(st, v, wv, Cv, dz, dz, dz);
        }
      } else {  // no cost output: marginal only
        float v[16], dz[16];
        if (fold) {
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            dz[r] = 0.0f;
            v[r] = fmaf(kc, acc[r], sa[r]);
          }
        } else {
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            dz[r] = 0.0f;
            v[r] = fmaf(kc, acc[r], sa[r] + tq);
          }
        }
        accumulate<1, KIND, 16, float, false, true>(st, v, sw, dz, dz, dz, dz);
      }
    }
    flush_blocks<1, KIND, float>(st);
    if (more) store_terms(base + TILE, buf ^ 1);
    __syncthreads();
  };
  for (int T = 0; T < ntiles; T += 2) {
    const int64_t base = s0 + (int64_t)T * TILE;
    tile(base, std::integral_constant<int, 0>{}, T + 1 < ntiles);
    if (T + 1 < ntiles) tile(base + TILE, std::integral_constant<int, 1>{}, T + 2 < ntiles);
  }
  if constexpr (KIND != KIND_LSE) {
    // back to the requested offset tq (as the split merge and epilogue use)
    if (fold) {
      const double f = exp2_small((double)tq - eoff);
      st.S *= f;
      st.Sc *= f;
      st.Se *= f;
      st.Sce *= f;
    }
  }
  if constexpr (KIND == KIND_LSE) st.S *= exp2_small(-dl);
  // merge the two half-waves holding the same target
  {
    flush_blocks<1, KIND, float>(st);
    TState<float> o = st;
    o.m = __shfl_xor(st.m, 32);
    o.S = __shfl_xor(st.S, 32);
    o.Se = __shfl_xor(st.Se, 32);
    if constexpr (KIND != KIND_LSE) {
      o.Sc = __shfl_xor(st.Sc, 32);
      o.Sce = __shfl_xor(st.Sce, 32);
    }
    combine_state<1, KIND, float>(st, o);
  }
  int64_t tidx[1] = {lane < 32 ? t : ntp};
  TState<float> fin[1] = {st};
  finish_pass<1, KIND, 1, float>(d, w, splits, split, ntp, tidx, fin);
}

// (one product: three waves per SIMD, 168 VGPRs each, three 33 KB workgroups
// per CU at d = 128; the three-product split needs more registers and LDS)
template <int KIND, int NSPLIT, int KS>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NSPLIT == 3 ? 2 : 3))) void mfma_feature_pass(const PassDesc din, const Work w,
                                                              const int splits,
                                                              const int64_t split_len) {
  const PassDesc d = resolve_desc(din);
  mfma_body<KIND, NSPLIT, KS>(d, w, blockIdx.x, blockIdx.y, splits, split_len,
                              (int64_t)gridDim.x * 128);
}

template <int KIND, int NSPLIT, int KS>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NSPLIT == 3 ? 2 : 3))) void mfma_feature_pass_batch(const PassDesc* descs,
                                                                    const BatchItem* items,
                                                                    const BatchInst* insts,
                                                                    const Work w) {
  const BatchItem it = items[blockIdx.x];
  if (it.inst < 0) return;  // padding of an XCD's item sequence
  const BatchInst bi = insts[it.inst];
  const PassDesc d = resolve_desc(descs[it.inst]);
  const Work wi{w.partial + bi.pbase};
  mfma_body<KIND, NSPLIT, KS>(d, wi, it.tb, it.split, bi.splits, bi.split_len, bi.ntp);
}

template <int KIND, int NSPLIT>
hipError_t launch_mfma_ks(const PassDesc& d, const LaunchCfg& c, const Work& w, hipStream_t s) {
  dim3 grid((unsigned)c.tblocks, (unsigned)c.splits);
  switch (d.zks) {
    case 1: hipLaunchKernelGGL((mfma_feature_pass<KIND, NSPLIT, 1>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    case 2: hipLaunchKernelGGL((mfma_feature_pass<KIND, NSPLIT, 2>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    case 4: hipLaunchKernelGGL((mfma_feature_pass<KIND, NSPLIT, 4>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    case 8: hipLaunchKernelGGL((mfma_feature_pass<KIND, NSPLIT, 8>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_merge<1, KIND>(d, c, w, c.tblocks * 128, s);
}

template <int PREC, int KIND, int COST, bool REC>
hipError_t launch_rpt_rec(const PassDesc& d, const LaunchCfg& c, const Work& w, hipStream_t s) {
  if constexpr (KIND != KIND_BARY) {  // (five FP64 fields do not fit the wide layout's registers)
    if (c.wide) {
      hipLaunchKernelGGL((pair_pass_wide<PREC, KIND, COST, REC>), dim3((unsigned)c.tblocks),
                         dim3(kWideThreads), 0, s, d, c.split_len);
      return hipGetLastError();
    }
  }
  dim3 grid((unsigned)c.tblocks, (unsigned)c.splits);
  switch (c.rpt) {
    case 1: hipLaunchKernelGGL((pair_pass<PREC, KIND, COST, 1, REC>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    case 2: hipLaunchKernelGGL((pair_pass<PREC, KIND, COST, 2, REC>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    case 4: hipLaunchKernelGGL((pair_pass<PREC, KIND, COST, 4, REC>), grid, dim3(kThreads), 0, s, d, w, c.splits, c.split_len); break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_merge<PREC, KIND>(d, c, w, c.tblocks * (kThreads / kLanesPerTarget) * c.rpt, s);
}

// Passes whose sources come as staged records (mixed LSE / SUM with d.rec set)
// run an instantiation that reads only records; all others convert sources.
template <int PREC, int KIND, int COST>
hipError_t launch_rpt(const PassDesc& d, const LaunchCfg& c, const Work& w, hipStream_t s) {
  if constexpr (PREC == 1 && (KIND == KIND_LSE || KIND == KIND_SUM))
    if (d.rec != nullptr) return launch_rpt_rec<PREC, KIND, COST, true>(d, c, w, s);
  return launch_rpt_rec<PREC, KIND, COST, false>(d, c, w, s);
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

// ---- batched launches: one workgroup per table item --------------------------
template <int PREC, int KIND>
hipError_t launch_merge_batch(const BatchLaunch& b, const Work& w, hipStream_t s) {
  if (b.nmitems <= 0) return hipSuccess;
  hipLaunchKernelGGL((pass_merge_batch<PREC, KIND>), dim3((unsigned)b.nmitems), dim3(kThreads), 0,
                     s, b.descs, b.mitems, b.insts, w);
  return hipGetLastError();
}

template <int PREC, int KIND, int COST>
hipError_t launch_batch_rpt(const BatchLaunch& b, const Work& w, hipStream_t s) {
  const dim3 grid((unsigned)b.nitems);
  switch (b.rpt) {
    case 1: hipLaunchKernelGGL((pair_pass_batch<PREC, KIND, COST, 1>), grid, dim3(kThreads), 0, s, b.descs, b.items, b.insts, w); break;
    case 2: hipLaunchKernelGGL((pair_pass_batch<PREC, KIND, COST, 2>), grid, dim3(kThreads), 0, s, b.descs, b.items, b.insts, w); break;
    case 4: hipLaunchKernelGGL((pair_pass_batch<PREC, KIND, COST, 4>), grid, dim3(kThreads), 0, s, b.descs, b.items, b.insts, w); break;
    default: return hipErrorInvalidValue;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_merge_batch<PREC, KIND>(b, w, s);
}

template <int PREC, int KIND>
hipError_t launch_batch_points(const BatchLaunch& b, const Work& w, hipStream_t s) {
  if (b.nitems <= 0) return hipSuccess;
  switch (b.cost) {
    case COST_L1: return launch_batch_rpt<PREC, KIND, COST_L1>(b, w, s);
    case COST_SQL2: return launch_batch_rpt<PREC, KIND, COST_SQL2>(b, w, s);

    case COST_LINF: return launch_batch_rpt<PREC, KIND, COST_LINF>(b, w, s);
    default: return launch_batch_rpt<PREC, KIND, COST_GENP>(b, w, s);
  }
}

template <int PREC, int KIND, int METRIC>
hipError_t launch_batch_feature_t(const BatchLaunch& b, const Work& w, hipStream_t s) {
  hipLaunchKernelGGL((feature_pass_batch<PREC, KIND, METRIC>), dim3((unsigned)b.nitems),
                     dim3(kThreads), 0, s, b.descs, b.items, b.insts, w);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_merge_batch<PREC, KIND>(b, w, s);
}

template <int PREC, int KIND>
hipError_t launch_batch_features(const BatchLaunch& b, const Work& w, hipStream_t s) {
  if (b.nitems <= 0) return hipSuccess;
  switch (b.metric) {
    case METRIC_L1: return launch_batch_feature_t<PREC, KIND, METRIC_L1>(b, w, s);
    case METRIC_L2: return launch_batch_feature_t<PREC, KIND, METRIC_L2>(b, w, s);
    case METRIC_COSINE: return launch_batch_feature_t<PREC, KIND, METRIC_COSINE>(b, w, s);
    default: return launch_batch_feature_t<PREC, KIND, METRIC_CORRELATION>(b, w, s);
  }
}

template <int KIND, int NSPLIT>
hipError_t launch_batch_mfma_ks(const BatchLaunch& b, const Work& w, hipStream_t s) {
  if (b.nitems <= 0) return hipSuccess;
  const dim3 grid((unsigned)b.nitems);
  switch (b.zks) {
    case 1: hipLaunchKernelGGL((mfma_feature_pass_batch<KIND, NSPLIT, 1>), grid, dim3(kThreads), 0, s, b.descs, b.items, b.insts, w); break;
    case 2: hipLaunchKernelGGL((mfma_feature_pass_batch<KIND, NSPLIT, 2>), grid, dim3(kThreads), 0, s, b.descs, b.items, b.insts, w); break;
    case 4: hipLaunchKernelGGL((mfma_feature_pass_batch<KIND, NSPLIT, 4>), grid, dim3(kThreads), 0, s, b.descs, b.items, b.insts, w); break;
    case 8: hipLaunchKernelGGL((mfma_feature_pass_batch<KIND, NSPLIT, 8>), grid, dim3(kThreads), 0, s, b.descs, b.items, b.insts, w); break;
    default: return hipErrorInvalidValue;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_merge_batch<1, KIND>(b, w, s);
}

}  // namespace
}  // namespace culamp

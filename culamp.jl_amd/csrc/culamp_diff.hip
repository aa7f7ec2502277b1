// culamp_diff.hip -- Sinkhorn divergence and its gradient with respect to the
// point positions and weights (squared Euclidean cost, fixed eps).
//
// Forward: OT(a,b) by Sinkhorn in the reference's log-domain form
// (SinkhornKernel.jl:341-444 half-steps with reg * W_inf = eps); the self
// terms OT(a,a), OT(b,b) by the symmetric update f <- (f + T(f)) / 2 with
// T(f)_i = log a_i - LSE_j(K_ij + f_j), whose fixed point is the symmetric
// potential (the alternating update converges slowly on a self problem).
// Each solve runs until the L1 row infeasibility reaches `tol`, keeping its
// potentials every `checkpoint` iterations in a ring that spans the last
// `unroll` iterations.
//
// Symmetric terms: value eps [2<a,f> - 2<a,log a> - sum P + 1], and per
// iteration f_t = (f_{t-1} + T(f_{t-1})) / 2 the adjoint (R row-normalised
// plan of the LSE, g the cotangent of f_t)
//        g'_j  = g_j / 2 - (1/2) sum_i g_i R_ij
//        g_x_j += 1/eps sum_i g_i R_ij (x_j - x_i)
//        g_x_i += 1/eps g_i (x_i - sum_j R_ij x_j)
//        g_a  += g / (2 a).
//
// Backward (all passes FP64, KIND_BARY state with the EPI_ADJ epilogue):
//   * envelope term: partial derivatives of the dual value at the final
//     potentials, one row and one column pass;
//   * unrolled adjoint: the cotangents of the final potentials are carried
//     back through the last `unroll` iterations.  Each iteration's column
//     half-step psi_t = log b - LSE_i(K + phi_t) and row half-step
//     phi_t = log a - LSE_j(K + psi_{t-1}) (K_ij = -|x_i - y_j|^2 / eps) has
//     the adjoint
//        g_phi_i  -= sum_j g_psi_j Q_ij           Q = col-normalised plan
//        g_x_i    += 2/eps sum_j g_psi_j Q_ij (x_i - y_j)
//        g_y_j    += 2/eps g_psi_j (y_j - sum_i Q_ij x_i)
//        g_psi'_j  = -sum_i g_phi_i R_ij          R = row-normalised plan
//        g_y_j    += 2/eps sum_i g_phi_i R_ij (y_j - x_i)
//        g_x_i    += 2/eps g_phi_i (x_i - sum_j R_ij y_j)
//        g_b += g_psi / b,  g_a += g_phi / a
//     each line one O(n m) pass; segments between checkpoints are recomputed
//     forward (deterministic kernels: bit-identical to the first pass).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "culamp_host.h"

using namespace culamp;
using namespace culamp_host;

namespace {

// g += c * u / w
__global__ void acc_ratio_kernel(double* g, const double* u, const double* w, double c, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] += c * (u[i] / w[i]);
}
// g += eps * (pot - log w - 1)
__global__ void acc_explicit_kernel(double* g, const double* pot, const double* w, double eps,
                                    int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] += eps * (pot[i] - log(w[i]) - 1.0);
}
// out = pot - log w
__global__ void minus_log_kernel(double* out, const double* pot, const double* w, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = pot[i] - log(w[i]);
}

unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

PassDesc adj_desc(const double* tgt, int64_t nt, const double* src, int64_t ns, double eps) {
  PassDesc d = blank();
  d.tgt = tgt; d.nt = nt; d.src = src; d.ns = ns;
  d.p = 2.0;
  d.k = -1.0 / eps;
  d.epi = EPI_ADJ;
  return d;
}

// One OT(a, x; b, y) term: forward solve, value, gradients.
struct Term {
  const double *x, *y, *a, *b;
  int64_t n, m;
};

struct TermState {
  DevBufs B;
  double *phi = nullptr, *psi = nullptr, *res = nullptr;
  double *phiT = nullptr, *psiT = nullptr;
  double *ck_phi = nullptr, *ck_psi = nullptr;  // ring of checkpoints
  int ring = 0, every = 1;
  std::vector<int64_t> ck_iter;
  int64_t T = 0;
  int64_t t_mixed = INT64_MAX;  // iterations after this one ran in FP64 (polish)
};

int store_ck(culamp_ctx* ctx, const Term& tm, TermState& S, int64_t t) {
  const int slot = (int)((t / S.every) % S.ring);
  HIP_TRY(hipMemcpyAsync(S.ck_phi + (int64_t)slot * tm.n, S.phi, sizeof(double) * tm.n,
                         hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(S.ck_psi + (int64_t)slot * tm.m, S.psi, sizeof(double) * tm.m,
                         hipMemcpyDeviceToDevice, ctx->stream));
  S.ck_iter[slot] = t;
  return CULAMP_OK;
}

int find_ck(const TermState& S, int64_t t, int* slot) {
  for (int q = 0; q < S.ring; ++q)
    if (S.ck_iter[q] == t) {
      *slot = q;
      return CULAMP_OK;
    }
  return fail(CULAMP_ERR_INVALID_ARG, "internal: checkpoint %lld not in the ring", (long long)t);
}

struct HalfSteps {
  PassDesc d1, d2;
};

HalfSteps half_steps(const Term& tm, TermState& S, double eps) {
  HalfSteps h;
  h.d1 = sk_half_desc(tm.x, tm.n, tm.y, tm.m, tm.a, S.psi, eps, 1.0, 2.0, S.phi);
  h.d1.prev_offset = 1;
  h.d2 = sk_half_desc(tm.y, tm.m, tm.x, tm.n, tm.b, S.phi, eps, 1.0, 2.0, S.psi);
  h.d2.prev_offset = 1;
  return h;
}

int forward(culamp_ctx* ctx, const Term& tm, double eps, const culamp_divergence_opts& o,
            TermState& S, culamp_result* out) {
  const double t0 = now_s();
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  TRY(S.B.alloc(&S.phi, tm.n)); TRY(S.B.alloc(&S.psi, tm.m)); TRY(S.B.alloc(&S.res, tm.n));
  TRY(S.B.alloc(&S.phiT, tm.n)); TRY(S.B.alloc(&S.psiT, tm.m));
  S.every = o.checkpoint > 0 ? o.checkpoint : 1;
  S.ring = o.unroll > 0 ? (int)((o.unroll + S.every - 1) / S.every) + 2 : 1;
  S.ck_iter.assign(S.ring, -1);
  TRY(S.B.alloc(&S.ck_phi, (int64_t)S.ring * tm.n));
  TRY(S.B.alloc(&S.ck_psi, (int64_t)S.ring * tm.m));
  HIP_TRY(launch_fill(S.phi, 0.0, tm.n, ctx->stream));
  HIP_TRY(launch_fill(S.psi, 0.0, tm.m, ctx->stream));
  const HalfSteps h = half_steps(tm, S, eps);
  const PassDesc d3 = sk_resid_desc(tm.x, tm.n, tm.y, tm.m, tm.a, S.phi, S.psi, eps, 1.0, 2.0,
                                    S.res, nullptr);
  if (o.unroll > 0) TRY(store_ck(ctx, tm, S, 0));
  int64_t t = 0;
  int converged = 0;
  double resid = NAN;
  int prec = o.precision;
  // after a switch to FP64 (at tol, or on a stall) the solve continues to
  // the smaller of tol and the polish tolerance
  const double fp64_target = o.polish_tol > 0.0 ? std::fmin(o.tol, o.polish_tol) : o.tol;
  PrecisionGuard guard(o.precision, ctx->fp64_fallback != 0);
  auto residual = [&]() -> int {
    TRY(run_pass(ctx, d3, prec, KIND_SUM));
    ReduceSpec rs{};
    rs.nops = 1; rs.op[0] = 5; rs.a[0] = S.res; rs.len[0] = tm.n;
    return reduce_to_host(ctx, rs, &resid);
  };
  while (t < o.itermax) {
    TRY(run_pass(ctx, h.d1, prec, KIND_LSE));
    TRY(run_pass(ctx, h.d2, prec, KIND_LSE));
    ++t;
    if (o.unroll > 0 && t % S.every == 0) TRY(store_ck(ctx, tm, S, t));
    if (t % o.frequency == 0) {
      TRY(residual());
      if (resid <= (prec == o.precision ? o.tol : fp64_target)) {
        if (prec != CULAMP_FP64 && o.polish_tol > 0.0) {
          // continue in FP64 from the mixed solution to the polish tolerance
          S.t_mixed = t;
          prec = CULAMP_FP64;
          TRY(residual());
          if (resid > fp64_target) continue;
        }
        converged = 1;
        break;
      }
      if (guard.observe(resid, o.tol, t)) {
        // the mixed residual stalled above tol (its FP32 floor): FP64 from here
        S.t_mixed = t;
        prec = CULAMP_FP64;
      }
      if (now_s() - t0 > o.tmax) break;
    }
  }
  if (!converged) TRY(residual());
  S.T = t;
  HIP_TRY(hipMemcpyAsync(S.phiT, S.phi, sizeof(double) * tm.n, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(S.psiT, S.psi, sizeof(double) * tm.m, hipMemcpyDeviceToDevice, ctx->stream));
  // value: eps [<a,phi> + <b,psi> - <a,log a> - <b,log b> - (sum a - sum res) + 1]
  ReduceSpec rs{};
  rs.nops = 6;
  rs.op[0] = 2; rs.a[0] = tm.a; rs.b[0] = S.phi; rs.len[0] = tm.n;
  rs.op[1] = 2; rs.a[1] = tm.b; rs.b[1] = S.psi; rs.len[1] = tm.m;
  rs.op[2] = 4; rs.a[2] = tm.a; rs.len[2] = tm.n;
  rs.op[3] = 4; rs.a[3] = tm.b; rs.len[3] = tm.m;
  rs.op[4] = 0; rs.a[4] = tm.a; rs.len[4] = tm.n;
  rs.op[5] = 0; rs.a[5] = S.res; rs.len[5] = tm.n;
  double v[6];
  TRY(reduce_to_host(ctx, rs, v));
  const double value = eps * (v[0] + v[1] - v[2] - v[3] - (v[4] - v[5]) + 1.0);
  if (out) {
    std::memset(out, 0, sizeof *out);
    out->iterations = t;
    out->objective = value;
    out->residual = resid;
    out->primal = NAN;
    out->dual = value;
    out->winf = NAN;
    out->eta_t = eps;
    out->converged = converged;
    out->fp64_from = S.t_mixed != INT64_MAX ? S.t_mixed : 0;
    out->nonfinite = !(std::isfinite(value) && std::isfinite(resid));
    out->elapsed_s = now_s() - t0;
    out->pairs = ctx->pairs - pairs0;
    out->passes = ctx->passes - passes0;
  }
  return CULAMP_OK;
}

// Gradient accumulators of one term (device, added to, never cleared here).
struct Grads {
  double *gx, *ga, *gy, *gb;  // gy / gb may alias gx / ga (self terms)
  bool xside, yside;
};

int backward(culamp_ctx* ctx, const Term& tm, double eps, const culamp_divergence_opts& o,
             TermState& S, const Grads& G) {
  const int64_t n = tm.n, m = tm.m;
  DevBufs B;
  double *gphi, *gpsi, *u, *v;
  TRY(B.alloc(&gphi, n)); TRY(B.alloc(&gpsi, m)); TRY(B.alloc(&u, m)); TRY(B.alloc(&v, n));
  const hipStream_t s = ctx->stream;
  // envelope term at the final potentials
  HIP_TRY(launch_axpby(gphi, tm.a, eps, tm.a, 0.0, n, s));
  HIP_TRY(launch_axpby(gpsi, tm.b, eps, tm.b, 0.0, m, s));
  {
    PassDesc d = adj_desc(tm.x, n, tm.y, m, eps);
    d.srcA = S.psiT; d.alpha = 1.0; d.tgtA = S.phiT; d.gamma = 1.0;
    d.out1 = gphi; d.adj_c1 = -eps; d.adj_beta = 1.0;
    d.out0 = G.xside ? G.gx : nullptr; d.adj_c0 = 2.0;
    TRY(run_pass(ctx, d, CULAMP_FP64, KIND_BARY));
    PassDesc e = adj_desc(tm.y, m, tm.x, n, eps);
    e.srcA = S.phiT; e.alpha = 1.0; e.tgtA = S.psiT; e.gamma = 1.0;
    e.out1 = gpsi; e.adj_c1 = -eps; e.adj_beta = 1.0;
    e.out0 = G.yside ? G.gy : nullptr; e.adj_c0 = 2.0;
    TRY(run_pass(ctx, e, CULAMP_FP64, KIND_BARY));
  }
  if (G.xside) {
    hipLaunchKernelGGL(acc_explicit_kernel, dim3(nb(n)), dim3(256), 0, s, G.ga, S.phiT, tm.a, eps, n);
    HIP_TRY(hipGetLastError());
  }
  if (G.yside) {
    hipLaunchKernelGGL(acc_explicit_kernel, dim3(nb(m)), dim3(256), 0, s, G.gb, S.psiT, tm.b, eps, m);
    HIP_TRY(hipGetLastError());
  }
  if (o.unroll <= 0 || S.T == 0) return CULAMP_OK;

  // unrolled window: from the last checkpoint at or before T - unroll
  const int64_t c = S.every;
  const int64_t lo = S.T > o.unroll ? ((S.T - o.unroll) / c) * c : 0;
  double *seg_phi, *seg_psi;
  TRY(B.alloc(&seg_phi, c * n));
  TRY(B.alloc(&seg_psi, c * m));
  const HalfSteps h = half_steps(tm, S, eps);
  const double k2 = 2.0 / eps;
  int64_t end = S.T;
  while (end > lo) {
    const int64_t start = ((end - 1) / c) * c;  // checkpoint at or below end - 1
    int slot = 0;
    TRY(find_ck(S, start, &slot));
    const double* ck_phi = S.ck_phi + (int64_t)slot * n;
    const double* ck_psi = S.ck_psi + (int64_t)slot * m;
    // recompute iterations start+1 .. end from the checkpoint
    HIP_TRY(hipMemcpyAsync(S.phi, ck_phi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.psi, ck_psi, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    for (int64_t t = start + 1; t <= end; ++t) {
      const int pr = t <= S.t_mixed ? o.precision : CULAMP_FP64;
      TRY(run_pass(ctx, h.d1, pr, KIND_LSE));
      TRY(run_pass(ctx, h.d2, pr, KIND_LSE));
      const int64_t q = t - start - 1;
      HIP_TRY(hipMemcpyAsync(seg_phi + q * n, S.phi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      HIP_TRY(hipMemcpyAsync(seg_psi + q * m, S.psi, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    }
    for (int64_t t = end; t > start; --t) {
      const int64_t q = t - start - 1;
      const double* phit = seg_phi + q * n;
      const double* psit = seg_psi + q * m;
      const double* psip = q == 0 ? ck_psi : seg_psi + (q - 1) * m;
      // column half-step psi_t = log b - LSE_i(K + phi_t)
      if (G.yside) {
        hipLaunchKernelGGL(acc_ratio_kernel, dim3(nb(m)), dim3(256), 0, s, G.gb, gpsi, tm.b, 1.0, m);
        HIP_TRY(hipGetLastError());
      }
      hipLaunchKernelGGL(minus_log_kernel, dim3(nb(m)), dim3(256), 0, s, u, psit, tm.b, m);
      HIP_TRY(hipGetLastError());
      {
        PassDesc d = adj_desc(tm.x, n, tm.y, m, eps);
        d.srcA = u; d.alpha = 1.0; d.srcB = gpsi; d.tgtA = phit; d.gamma = 1.0;
        d.out1 = gphi; d.adj_c1 = -1.0; d.adj_beta = 1.0;
        d.out0 = G.xside ? G.gx : nullptr; d.adj_c0 = k2;
        TRY(run_pass(ctx, d, CULAMP_FP64, KIND_BARY));
      }
      if (G.yside) {
        PassDesc d = adj_desc(tm.y, m, tm.x, n, eps);
        d.srcA = phit; d.alpha = 1.0;
        d.adj_norm = 1; d.tgtW = gpsi; d.out0 = G.gy; d.adj_c0 = k2;
        TRY(run_pass(ctx, d, CULAMP_FP64, KIND_BARY));
      }
      // row half-step phi_t = log a - LSE_j(K + psi_{t-1})
      if (G.xside) {
        hipLaunchKernelGGL(acc_ratio_kernel, dim3(nb(n)), dim3(256), 0, s, G.ga, gphi, tm.a, 1.0, n);
        HIP_TRY(hipGetLastError());
      }
      hipLaunchKernelGGL(minus_log_kernel, dim3(nb(n)), dim3(256), 0, s, v, phit, tm.a, n);
      HIP_TRY(hipGetLastError());
      {
        PassDesc d = adj_desc(tm.y, m, tm.x, n, eps);
        d.srcA = v; d.alpha = 1.0; d.srcB = gphi; d.tgtA = psip; d.gamma = 1.0;
        d.out1 = gpsi; d.adj_c1 = -1.0; d.adj_beta = 0.0;
        d.out0 = G.yside ? G.gy : nullptr; d.adj_c0 = k2;
        TRY(run_pass(ctx, d, CULAMP_FP64, KIND_BARY));
      }
      if (G.xside) {
        PassDesc d = adj_desc(tm.x, n, tm.y, m, eps);
        d.srcA = psip; d.alpha = 1.0;
        d.adj_norm = 1; d.tgtW = gphi; d.out0 = G.gx; d.adj_c0 = k2;
        TRY(run_pass(ctx, d, CULAMP_FP64, KIND_BARY));
      }
      HIP_TRY(launch_fill(gphi, 0.0, n, s));
    }
    end = start;
  }
  return CULAMP_OK;
}

// ---- symmetric self terms OT(a, x; a, x) ------------------------------------
struct SymState {
  DevBufs B;
  double *f = nullptr, *Tf = nullptr, *res = nullptr, *fT = nullptr;
  double *ck_f = nullptr, *ck_Tf = nullptr;
  int ring = 0, every = 1;
  std::vector<int64_t> ck_iter;
  int64_t T = 0;
  int64_t t_mixed = INT64_MAX;  // iterations after this one ran in FP64 (polish)
};

int sym_store(culamp_ctx* ctx, int64_t n, SymState& S, int64_t t) {
  const int slot = (int)((t / S.every) % S.ring);
  HIP_TRY(hipMemcpyAsync(S.ck_f + (int64_t)slot * n, S.f, sizeof(double) * n, hipMemcpyDeviceToDevice,
                         ctx->stream));
  HIP_TRY(hipMemcpyAsync(S.ck_Tf + (int64_t)slot * n, S.Tf, sizeof(double) * n,
                         hipMemcpyDeviceToDevice, ctx->stream));
  S.ck_iter[slot] = t;
  return CULAMP_OK;
}

// one symmetric iteration: Tf = T(f) (LSE pass, centred on the previous Tf),
// then f = f / 2 + Tf / 2
int sym_step(culamp_ctx* ctx, const double* x, const double* a, int64_t n, double eps, int prec,
             SymState& S) {
  PassDesc d = sk_half_desc(x, n, x, n, a, S.f, eps, 1.0, 2.0, S.Tf);
  d.prev_offset = 1;
  TRY(run_pass(ctx, d, prec, KIND_LSE));
  HIP_TRY(launch_axpby(S.f, S.f, 0.5, S.Tf, 0.5, n, ctx->stream));
  return CULAMP_OK;
}

int sym_forward(culamp_ctx* ctx, const double* x, const double* a, int64_t n, double eps,
                const culamp_divergence_opts& o, SymState& S, culamp_result* out) {
  const double t0 = now_s();
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  TRY(S.B.alloc(&S.f, n)); TRY(S.B.alloc(&S.Tf, n)); TRY(S.B.alloc(&S.res, n));
  TRY(S.B.alloc(&S.fT, n));
  S.every = o.checkpoint > 0 ? o.checkpoint : 1;
  S.ring = o.unroll > 0 ? (int)((o.unroll + S.every - 1) / S.every) + 2 : 1;
  S.ck_iter.assign(S.ring, -1);
  TRY(S.B.alloc(&S.ck_f, (int64_t)S.ring * n));
  TRY(S.B.alloc(&S.ck_Tf, (int64_t)S.ring * n));
  HIP_TRY(launch_fill(S.f, 0.0, n, ctx->stream));
  HIP_TRY(launch_fill(S.Tf, 0.0, n, ctx->stream));
  const PassDesc d3 = sk_resid_desc(x, n, x, n, a, S.f, S.f, eps, 1.0, 2.0, S.res, nullptr);
  if (o.unroll > 0) TRY(sym_store(ctx, n, S, 0));
  int64_t t = 0;
  int converged = 0;
  double resid = NAN;
  int prec = o.precision;
  // after a switch to FP64 (at tol, or on a stall) the solve continues to
  // the smaller of tol and the polish tolerance
  const double fp64_target = o.polish_tol > 0.0 ? std::fmin(o.tol, o.polish_tol) : o.tol;
  PrecisionGuard guard(o.precision, ctx->fp64_fallback != 0);
  auto residual = [&]() -> int {
    TRY(run_pass(ctx, d3, prec, KIND_SUM));
    ReduceSpec rs{};
    rs.nops = 1; rs.op[0] = 5; rs.a[0] = S.res; rs.len[0] = n;
    return reduce_to_host(ctx, rs, &resid);
  };
  while (t < o.itermax) {
    TRY(sym_step(ctx, x, a, n, eps, prec, S));
    ++t;
    if (o.unroll > 0 && t % S.every == 0) TRY(sym_store(ctx, n, S, t));
    if (t % o.frequency == 0) {
      TRY(residual());
      if (resid <= (prec == o.precision ? o.tol : fp64_target)) {
        if (prec != CULAMP_FP64 && o.polish_tol > 0.0) {
          // continue in FP64 from the mixed solution to the polish tolerance
          S.t_mixed = t;
          prec = CULAMP_FP64;
          TRY(residual());
          if (resid > fp64_target) continue;
        }
        converged = 1;
        break;
      }
      if (guard.observe(resid, o.tol, t)) {
        // the mixed residual stalled above tol (its FP32 floor): FP64 from here
        S.t_mixed = t;
        prec = CULAMP_FP64;
      }
      if (now_s() - t0 > o.tmax) break;
    }
  }
  if (!converged) TRY(residual());
  S.T = t;
  HIP_TRY(hipMemcpyAsync(S.fT, S.f, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  ReduceSpec rs{};
  rs.nops = 4;
  rs.op[0] = 2; rs.a[0] = a; rs.b[0] = S.f; rs.len[0] = n;
  rs.op[1] = 4; rs.a[1] = a; rs.len[1] = n;
  rs.op[2] = 0; rs.a[2] = a; rs.len[2] = n;
  rs.op[3] = 0; rs.a[3] = S.res; rs.len[3] = n;
  double v[4];
  TRY(reduce_to_host(ctx, rs, v));
  const double value = eps * (2.0 * v[0] - 2.0 * v[1] - (v[2] - v[3]) + 1.0);
  if (out) {
    std::memset(out, 0, sizeof *out);
    out->iterations = t;
    out->objective = value;
    out->residual = resid;
    out->primal = NAN;
    out->dual = value;
    out->winf = NAN;
    out->eta_t = eps;
    out->converged = converged;
    out->fp64_from = S.t_mixed != INT64_MAX ? S.t_mixed : 0;
    out->nonfinite = !(std::isfinite(value) && std::isfinite(resid));
    out->elapsed_s = now_s() - t0;
    out->pairs = ctx->pairs - pairs0;
    out->passes = ctx->passes - passes0;
  }
  return CULAMP_OK;
}

// Adds the gradient of one symmetric term to gx (3 x n) and ga (n).
int sym_backward(culamp_ctx* ctx, const double* x, const double* a, int64_t n, double eps,
                 const culamp_divergence_opts& o, SymState& S, double* gx, double* ga) {
  DevBufs B;
  double *g, *gn, *h;
  TRY(B.alloc(&g, n)); TRY(B.alloc(&gn, n)); TRY(B.alloc(&h, n));
  const hipStream_t s = ctx->stream;
  // envelope: dF/dx_i = 4 (x_i S_i - Y_i), dF/df = 2 eps (a - S), dF/da = 2 eps (f - log a - 1)
  HIP_TRY(launch_axpby(g, a, 2.0 * eps, a, 0.0, n, s));
  {
    PassDesc d = adj_desc(x, n, x, n, eps);
    d.srcA = S.fT; d.alpha = 1.0; d.tgtA = S.fT; d.gamma = 1.0;
    d.out1 = g; d.adj_c1 = -2.0 * eps; d.adj_beta = 1.0;
    d.out0 = gx; d.adj_c0 = 4.0;
    TRY(run_pass(ctx, d, CULAMP_FP64, KIND_BARY));
  }
  hipLaunchKernelGGL(acc_explicit_kernel, dim3(nb(n)), dim3(256), 0, s, ga, S.fT, a, 2.0 * eps, n);
  HIP_TRY(hipGetLastError());
  if (o.unroll <= 0 || S.T == 0) return CULAMP_OK;
  const int64_t c = S.every;
  const int64_t lo = S.T > o.unroll ? ((S.T - o.unroll) / c) * c : 0;
  double *seg_f, *seg_Tf;
  TRY(B.alloc(&seg_f, c * n));
  TRY(B.alloc(&seg_Tf, c * n));
  const double k1 = 1.0 / eps;
  int64_t end = S.T;
  while (end > lo) {
    const int64_t start = ((end - 1) / c) * c;
    int slot = -1;
    for (int q = 0; q < S.ring; ++q)
      if (S.ck_iter[q] == start) slot = q;
    if (slot < 0) return fail(CULAMP_ERR_INVALID_ARG, "internal: checkpoint %lld not in the ring", (long long)start);
    const double* ckf = S.ck_f + (int64_t)slot * n;
    HIP_TRY(hipMemcpyAsync(S.f, ckf, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.Tf, S.ck_Tf + (int64_t)slot * n, sizeof(double) * n,
                           hipMemcpyDeviceToDevice, s));
    for (int64_t t = start + 1; t <= end; ++t) {
      TRY(sym_step(ctx, x, a, n, eps, t <= S.t_mixed ? o.precision : CULAMP_FP64, S));
      const int64_t q = t - start - 1;
      HIP_TRY(hipMemcpyAsync(seg_f + q * n, S.f, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      HIP_TRY(hipMemcpyAsync(seg_Tf + q * n, S.Tf, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    for (int64_t t = end; t > start; --t) {
      const int64_t q = t - start - 1;
      const double* fp = q == 0 ? ckf : seg_f + (q - 1) * n;  // f_{t-1}
      const double* Tft = seg_Tf + q * n;                       // T(f_{t-1})
      hipLaunchKernelGGL(acc_ratio_kernel, dim3(nb(n)), dim3(256), 0, s, ga, g, a, 0.5, n);
      HIP_TRY(hipGetLastError());
      // h_i = -LSE_i = T(f)_i - log a_i
      hipLaunchKernelGGL(minus_log_kernel, dim3(nb(n)), dim3(256), 0, s, h, Tft, a, n);
      HIP_TRY(hipGetLastError());
      {
        PassDesc d = adj_desc(x, n, x, n, eps);  // targets j, sources i
        d.srcA = h; d.alpha = 1.0; d.srcB = g; d.tgtA = fp; d.gamma = 1.0;
        d.out1 = gn; d.adj_c1 = -0.5; d.adj_beta = 0.0;
        d.out0 = gx; d.adj_c0 = k1;
        TRY(run_pass(ctx, d, CULAMP_FP64, KIND_BARY));
      }
      {
        PassDesc d = adj_desc(x, n, x, n, eps);  // targets i, sources j
        d.srcA = fp; d.alpha = 1.0;
        d.adj_norm = 1; d.tgtW = g; d.out0 = gx; d.adj_c0 = k1;
        TRY(run_pass(ctx, d, CULAMP_FP64, KIND_BARY));
      }
      HIP_TRY(launch_axpby(g, g, 0.5, gn, 1.0, n, s));
    }
    end = start;
  }
  return CULAMP_OK;
}

int check_opts(const culamp_divergence_opts* o) {
  if (!o) return fail(CULAMP_ERR_INVALID_ARG, "null options");
  if (!(o->eps > 0.0) && !(o->eta_p > 0.0))
    return fail(CULAMP_ERR_INVALID_ARG, "eps or eta_p must be positive");
  if (o->frequency <= 0 || o->itermax <= 0)
    return fail(CULAMP_ERR_INVALID_ARG, "frequency and itermax must be positive");
  if (o->unroll < 0 || o->checkpoint < 0) return fail(CULAMP_ERR_INVALID_ARG, "negative unroll / checkpoint");
  if (o->polish_tol < 0.0) return fail(CULAMP_ERR_INVALID_ARG, "polish_tol must be >= 0");
  if (o->precision != CULAMP_FP64 && o->precision != CULAMP_MIXED)
    return fail(CULAMP_ERR_INVALID_ARG, "precision must be CULAMP_FP64 or CULAMP_MIXED");
  return CULAMP_OK;
}

int divergence(culamp_ctx* ctx, const double* hx, int64_t n, const double* ha, const double* hy,
               int64_t m, const double* hb, const culamp_divergence_opts* o, double scale,
               double* value, double* grad_x, double* grad_a, double* grad_y, double* grad_b,
               culamp_result* terms) {
  TRY(check_ctx(ctx));
  TRY(check_opts(o));
  if (!hx || !ha || !hy || !hb) return fail(CULAMP_ERR_INVALID_ARG, "null input array");
  if (n <= 0 || m <= 0) return fail(CULAMP_ERR_INVALID_ARG, "sizes must be positive");
  const hipStream_t s = ctx->stream;
  DevBufs B;
  double *x, *y, *a, *b, *tmp;
  TRY(B.alloc(&x, 3 * n)); TRY(B.alloc(&y, 3 * m)); TRY(B.alloc(&a, n)); TRY(B.alloc(&b, m));
  TRY(B.alloc(&tmp, n));
  HIP_TRY(hipMemcpyAsync(x, hx, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(y, hy, sizeof(double) * 3 * m, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(a, ha, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b, hb, sizeof(double) * m, hipMemcpyHostToDevice, s));
  double eps = o->eps;
  if (!(eps > 0.0)) {
    PassDesc d = max_desc(x, n, y, m, 2.0, tmp);
    TRY(run_pass(ctx, d, CULAMP_FP64, KIND_MAX));
    ReduceSpec rs{};
    rs.nops = 1; rs.op[0] = 3; rs.a[0] = tmp; rs.len[0] = n;
    double w;
    TRY(reduce_to_host(ctx, rs, &w));
    eps = o->eta_p * w;
  }
  const bool want_x = grad_x || grad_a, want_y = grad_y || grad_b;
  const bool want = want_x || want_y;
  double *gx, *ga, *gy, *gb, *sx, *sa, *sy, *sb;  // OT(a,b) and self-term accumulators
  TRY(B.alloc(&gx, 3 * n)); TRY(B.alloc(&ga, n)); TRY(B.alloc(&gy, 3 * m)); TRY(B.alloc(&gb, m));
  TRY(B.alloc(&sx, 3 * n)); TRY(B.alloc(&sa, n)); TRY(B.alloc(&sy, 3 * m)); TRY(B.alloc(&sb, m));
  for (auto pr : {std::make_pair(gx, 3 * n), std::make_pair(ga, n), std::make_pair(gy, 3 * m),
                  std::make_pair(gb, m), std::make_pair(sx, 3 * n), std::make_pair(sa, n),
                  std::make_pair(sy, 3 * m), std::make_pair(sb, m)})
    HIP_TRY(launch_fill(pr.first, 0.0, pr.second, s));
  culamp_result r[3];
  // OT(a, x; b, y)
  {
    Term tm{x, y, a, b, n, m};
    TermState S;
    TRY(forward(ctx, tm, eps, *o, S, &r[0]));
    if (want) TRY(backward(ctx, tm, eps, *o, S, Grads{gx, ga, gy, gb, want_x, want_y}));
    HIP_TRY(hipStreamSynchronize(s));
  }
  // OT(a, x; a, x) and OT(b, y; b, y): symmetric updates
  {
    SymState S;
    TRY(sym_forward(ctx, x, a, n, eps, *o, S, &r[1]));
    if (want_x) TRY(sym_backward(ctx, x, a, n, eps, *o, S, sx, sa));
    HIP_TRY(hipStreamSynchronize(s));
  }
  {
    SymState S;
    TRY(sym_forward(ctx, y, b, m, eps, *o, S, &r[2]));
    if (want_y) TRY(sym_backward(ctx, y, b, m, eps, *o, S, sy, sb));
    HIP_TRY(hipStreamSynchronize(s));
  }
  // S gradients: OT(a,b) terms minus half the self terms, times `scale`
  if (want_x) {
    HIP_TRY(launch_axpby(gx, gx, scale, sx, -0.5 * scale, 3 * n, s));
    HIP_TRY(launch_axpby(ga, ga, scale, sa, -0.5 * scale, n, s));
  }
  if (want_y) {
    HIP_TRY(launch_axpby(gy, gy, scale, sy, -0.5 * scale, 3 * m, s));
    HIP_TRY(launch_axpby(gb, gb, scale, sb, -0.5 * scale, m, s));
  }
  TRY(copy_out(ctx, grad_x, gx, 3 * n));
  TRY(copy_out(ctx, grad_a, ga, n));
  TRY(copy_out(ctx, grad_y, gy, 3 * m));
  TRY(copy_out(ctx, grad_b, gb, m));
  HIP_TRY(hipStreamSynchronize(s));
  if (value) *value = r[0].objective - 0.5 * r[1].objective - 0.5 * r[2].objective;
  if (terms)
    for (int q = 0; q < 3; ++q) terms[q] = r[q];
  return CULAMP_OK;
}

}  // namespace

extern "C" {

void culamp_divergence_opts_default(culamp_divergence_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->eps = 0.0;
  o->eta_p = 1e-2;
  o->tol = 1e-6;
  o->itermax = 100000;
  o->frequency = 10;
  o->unroll = 64;
  o->checkpoint = 16;
  o->precision = CULAMP_MIXED;
  o->polish_tol = 1e-9;
  o->tmax = INFINITY;
}

int culamp_sinkhorn_divergence(culamp_ctx* ctx, const double* x, int64_t n, const double* a,
                               const double* y, int64_t m, const double* b,
                               const culamp_divergence_opts* opts, double* value, double* grad_x,
                               double* grad_a, double* grad_y, double* grad_b,
                               culamp_result* terms) {
  return divergence(ctx, x, n, a, y, m, b, opts, 1.0, value, grad_x, grad_a, grad_y, grad_b, terms);
}

int culamp_sinkhorn_divergence_vjp(culamp_ctx* ctx, const double* x, int64_t n, const double* a,
                                   const double* y, int64_t m, const double* b,
                                   const culamp_divergence_opts* opts, double dS, double* grad_x,
                                   double* grad_a, double* grad_y, double* grad_b) {
  return divergence(ctx, x, n, a, y, m, b, opts, dS, nullptr, grad_x, grad_a, grad_y, grad_b,
                    nullptr);
}

int culamp_lse_adjoint(culamp_ctx* ctx, const double* tgt, int64_t nt, const double* src,
                       int64_t ns, double eps, const double* srcA, double alpha,
                       const double* srcB, const double* tgtA, double gamma, int normalized,
                       const double* tgtW, double c0, double c1, double beta, double* out0,
                       double* out1) {
  TRY(check_ctx(ctx));
  if (!tgt || !src || nt < 0 || ns <= 0) return fail(CULAMP_ERR_INVALID_ARG, "bad points or sizes");
  if (!(eps > 0.0)) return fail(CULAMP_ERR_INVALID_ARG, "eps must be positive");
  if (normalized && !out0) return fail(CULAMP_ERR_INVALID_ARG, "normalized adjoint needs out0");
  PassDesc d = adj_desc(tgt, nt, src, ns, eps);
  d.srcA = srcA; d.alpha = alpha; d.srcB = srcB; d.tgtA = tgtA; d.gamma = gamma;
  d.adj_norm = normalized ? 1 : 0; d.tgtW = tgtW;
  d.adj_c0 = c0; d.adj_c1 = c1; d.adj_beta = beta;
  d.out0 = out0; d.out1 = out1;
  return run_pass(ctx, d, CULAMP_FP64, KIND_BARY);
}

}  // extern "C"

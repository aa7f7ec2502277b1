// culamp_api.hip -- host side of libculamp: contexts, device-resident
// problems, the kernel-level C ABI and the solver loops (LAMP mirror prox,
// Sinkhorn, annealed Sinkhorn) that drive the pair-pass kernels.
//
// The solver loops restate the reference's Julia host loops
// (src/solvers/MirrorProxKernel.jl:1019-1132, src/solvers/SinkhornKernel.jl:
// 477-559, src/solvers/AnnealedSinkhorn.jl:65-177) with the same stop tests,
// log lines and outputs.  Unlike the reference they do not synchronise the
// device every iteration: launches are queued on the context's stream and the
// host only waits on iterations that log or test for convergence.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "culamp.h"
#include "culamp_host.h"

using namespace culamp;

namespace culamp_host {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Julia's @printf prints non-finite values as Inf/-Inf/NaN.
std::string fmt(const char* f, double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "Inf" : "-Inf";
  char b[64];
  snprintf(b, sizeof b, f, v);
  return b;
}

void emit(culamp_log_fn fn, void* user, const std::string& s) {
  if (fn) fn(user, s.c_str());
}

int ensure_work(culamp_ctx* ctx, size_t partial_doubles) {
  if (partial_doubles > ctx->partial_cap) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return fail(CULAMP_ERR_INVALID_ARG,
                  "call needs a larger work buffer while capturing: issue it once before capturing");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->partial) HIP_TRY(hipFree(ctx->partial));
    ctx->partial = nullptr;
    size_t cap = partial_doubles + partial_doubles / 2;
    HIP_TRY(hipMalloc(&ctx->partial, cap * sizeof(double)));
    ctx->partial_cap = cap;
  }
  return CULAMP_OK;
}

int alias_problem(culamp_ctx* ctx, const culamp_problem* owner, const double* r,
                  const double* c, culamp_problem** out) {
  *out = nullptr;
  culamp_problem* pb = new culamp_problem(*owner);
  pb->geom_owner = owner->geom_owner ? owner->geom_owner : owner;
  pb->r = pb->c = pb->logr = pb->logc = nullptr;
  pb->hr.assign(r, r + pb->n);
  pb->hc.assign(c, c + pb->m);
  hipError_t e = hipSuccess;
  for (double** p : {&pb->r, &pb->logr})
    if (e == hipSuccess) e = hipMalloc(p, sizeof(double) * pb->n);
  for (double** p : {&pb->c, &pb->logc})
    if (e == hipSuccess) e = hipMalloc(p, sizeof(double) * pb->m);
  if (e == hipSuccess) e = hipMemcpyAsync(pb->r, r, sizeof(double) * pb->n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pb->c, c, sizeof(double) * pb->m, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_log(pb->logr, pb->r, pb->n, ctx->stream);
  if (e == hipSuccess) e = launch_log(pb->logc, pb->c, pb->m, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->problems.insert(pb);
  if (e != hipSuccess) {
    culamp_problem_destroy(pb);
    return fail(CULAMP_ERR_HIP, "problem upload failed: %s", hipGetErrorString(e));
  }
  *out = pb;
  return CULAMP_OK;
}

int run_pass(culamp_ctx* ctx, const PassDesc& d, int precision, int kind) {
  if (d.nt <= 0) return CULAMP_OK;
  if (precision < CULAMP_FP64 || precision > CULAMP_BF16X3)
    return fail(CULAMP_ERR_INVALID_ARG, "precision must be CULAMP_FP64, _MIXED, _BF16 or _BF16X3");
  if (d.metric >= 0 && (precision == CULAMP_BF16 || precision == CULAMP_BF16X3) && d.zks > 0 &&
      (kind == KIND_LSE || kind == KIND_SUM)) {
    const int nsplit = precision == CULAMP_BF16X3 ? 3 : 1;
    LaunchCfg cfg = choose_mfma_launch(d.nt, d.ns, ctx->num_cu, ctx->force_splits);
    if (cfg.splits > 1) TRY(ensure_work(ctx, (size_t)pass_fields(kind) * cfg.splits * cfg.tblocks * 128));
    Work w{ctx->partial};
    HIP_TRY(launch_feature_mfma(d, kind, nsplit, cfg, w, ctx->stream));
    ctx->pairs += (double)d.nt * (double)d.ns;
    ctx->passes += 1;
    return CULAMP_OK;
  }
  if (precision > CULAMP_MIXED) precision = CULAMP_MIXED;
  if (d.metric >= 0) {
    if (kind == KIND_BARY) return fail(CULAMP_ERR_INVALID_ARG, "no barycentric pass for feature costs");
    const bool fp32 = precision == CULAMP_MIXED && (kind == KIND_LSE || kind == KIND_SUM);
    LaunchCfg cfg = choose_feature_launch(d.nt, d.ns, ctx->num_cu, ctx->force_splits);
    if (cfg.splits > 1) {
      const size_t ntp = (size_t)cfg.tblocks * kFeatTargets;
      TRY(ensure_work(ctx, (size_t)pass_fields(kind) * cfg.splits * ntp));
    }
    Work w{ctx->partial};
    HIP_TRY(fp32 ? launch_feature_fp32(d, kind, cfg, w, ctx->stream)
                 : launch_feature_fp64(d, kind, cfg, w, ctx->stream));
    ctx->pairs += (double)d.nt * (double)d.ns;
    ctx->passes += 1;
    return CULAMP_OK;
  }
  LaunchCfg cfg = choose_launch(d.nt, d.ns, precision, kind, ctx->num_cu,
                                ctx->force_rpt, ctx->force_splits, ctx->layout);
  if (cfg.splits > 1 && !cfg.wide) {
    const size_t ntp = (size_t)cfg.tblocks * (kThreads / kLanesPerTarget) * cfg.rpt;
    TRY(ensure_work(ctx, (size_t)pass_fields(kind) * cfg.splits * ntp));
  }
  Work w{ctx->partial};
  HIP_TRY(launch_pass(d, precision, kind, cfg, w, ctx->stream));
  ctx->pairs += (double)d.nt * (double)d.ns;
  ctx->passes += 1;
  return CULAMP_OK;
}

PassDesc blank() {
  PassDesc d;
  std::memset(&d, 0, sizeof d);
  d.metric = -1;
  return d;
}

// Point the pass at a feature problem's packed rows: `rows` makes the rows
// (x side) the targets, otherwise the columns (y side).
void feat(PassDesc& d, const culamp_problem* pb, bool rows) {
  if (pb->metric < 0) return;
  d.metric = pb->metric;
  d.dim = pb->d;
  d.fscale = pb->scale;
  d.tgt = d.src = nullptr;
  if (rows) {
    d.nt = pb->n; d.ns = pb->m;
    d.tf64 = pb->xf64; d.sf64 = pb->yf64; d.tf32 = pb->xf32; d.sf32 = pb->yf32;
    d.trc = pb->xrc; d.src_rc = pb->yrc;
    d.tz_hi = (const __bf16*)pb->xz_hi; d.tz_lo = (const __bf16*)pb->xz_lo;
    d.sz_hi = (const __bf16*)pb->yz_hi; d.sz_lo = (const __bf16*)pb->yz_lo;
  } else {
    d.nt = pb->m; d.ns = pb->n;
    d.tf64 = pb->yf64; d.sf64 = pb->xf64; d.tf32 = pb->yf32; d.sf32 = pb->xf32;
    d.trc = pb->yrc; d.src_rc = pb->xrc;
    d.tz_hi = (const __bf16*)pb->yz_hi; d.tz_lo = (const __bf16*)pb->yz_lo;
    d.sz_hi = (const __bf16*)pb->xz_hi; d.sz_lo = (const __bf16*)pb->xz_lo;
  }
  d.zks = pb->zks;
}

int check_ctx(culamp_ctx* ctx) {
  if (!ctx) return fail(CULAMP_ERR_INVALID_ARG, "null context");
  HIP_TRY(hipSetDevice(ctx->device));
  return CULAMP_OK;
}

int check_sizes(int64_t n, int64_t m) {
  if (n < 0 || m < 0) return fail(CULAMP_ERR_INVALID_ARG, "negative size (n=%lld, m=%lld)", (long long)n, (long long)m);
  return CULAMP_OK;
}

// ---- pass builders (kernel-level semantics) ---------------------------------

// MirrorProxKernel.jl:833-938: v = -(C*st/(2W) + theta_j)/reg
PassDesc lse_rows_desc(const double* x, int64_t n, const double* y, int64_t m,
                       const double* theta, double reg, double st, double winf,
                       double p, double* out) {
  PassDesc d = blank();
  const double c1 = st / (2.0 * winf);
  const double invreg = 1.0 / reg;
  d.tgt = x; d.nt = n; d.src = y; d.ns = m; d.p = p;
  d.k = -c1 * invreg;
  d.srcA = theta; d.alpha = -invreg;
  d.epi = EPI_LSE; d.out0 = out;
  return d;
}

// MirrorProxKernel.jl:7-94
PassDesc colmarg_desc(const double* x, int64_t n, const double* y, int64_t m,
                      const double* r, const double* theta, const double* logZ,
                      double reg, double st, double winf, double p, double* res,
                      double* cost) {
  PassDesc d = blank();
  const double c1 = st / (2.0 * winf);
  const double invreg = 1.0 / reg;
  d.tgt = y; d.nt = m; d.src = x; d.ns = n; d.p = p;
  d.k = -c1 * invreg;
  d.srcA = logZ; d.alpha = -1.0;
  d.srcB = r;
  d.tgtA = theta; d.gamma = -invreg;
  d.epi = EPI_SUM; d.out0 = res; d.out1 = cost;
  return d;
}

// SinkhornKernel.jl:341-444: out_i = log(marg_i) - LSE_j(-C/(reg W) + psi_j)
PassDesc sk_half_desc(const double* x, int64_t n, const double* y, int64_t m,
                      const double* marg, const double* psi, double reg,
                      double winf, double p, double* out) {
  PassDesc d = blank();
  d.tgt = x; d.nt = n; d.src = y; d.ns = m; d.p = p;
  d.k = -1.0 / (reg * winf);
  d.srcA = psi; d.alpha = 1.0;
  d.epi = EPI_SK_HALF; d.out0 = out; d.marg = marg;
  return d;
}

// SinkhornKernel.jl:9-81
PassDesc sk_resid_desc(const double* x, int64_t n, const double* y, int64_t m,
                       const double* marg, const double* phi, const double* psi,
                       double reg, double winf, double p, double* res,
                       double* cost) {
  PassDesc d = blank();
  d.tgt = x; d.nt = n; d.src = y; d.ns = m; d.p = p;
  d.k = -1.0 / (reg * winf);
  d.srcA = psi; d.alpha = 1.0;
  d.tgtA = phi; d.gamma = 1.0;
  d.epi = EPI_SK_RESID; d.out0 = res; d.out1 = cost; d.marg = marg;
  return d;
}

PassDesc max_desc(const double* x, int64_t n, const double* y, int64_t m,
                  double p, double* out) {
  PassDesc d = blank();
  d.tgt = x; d.nt = n; d.src = y; d.ns = m; d.p = p;
  d.epi = EPI_MAX; d.out0 = out;
  return d;
}

PassDesc minplus_desc(const double* x, int64_t n, const double* y, int64_t m,
                      const double* theta, double winf, double p, double* out) {
  PassDesc d = blank();
  d.tgt = x; d.nt = n; d.src = y; d.ns = m; d.p = p;
  d.srcA = theta; d.alpha = 2.0 * winf;
  d.epi = EPI_MIN; d.out0 = out;
  return d;
}

// SinkhornKernel.jl:124-162
PassDesc bary_sk_desc(const double* x, int64_t n, const double* y, int64_t m,
                      const double* psi, double reg, double winf, double p,
                      double* out) {
  PassDesc d = blank();
  d.tgt = x; d.nt = n; d.src = y; d.ns = m; d.p = p;
  d.k = -1.0 / (reg * winf);
  d.srcA = psi; d.alpha = 1.0;
  d.epi = EPI_BARY; d.out0 = out;
  return d;
}

// MirrorProxKernel.jl:97-142 (rows): weights exp(-(C/(2W) + nu_j)/reg)
PassDesc bary_lamp_rows_desc(const double* x, int64_t n, const double* y,
                             int64_t m, const double* nu, double reg,
                             double winf, double p, double* out) {
  PassDesc d = blank();
  d.tgt = x; d.nt = n; d.src = y; d.ns = m; d.p = p;
  d.k = -1.0 / (2.0 * winf * reg);
  d.srcA = nu; d.alpha = -1.0 / reg;
  d.epi = EPI_BARY; d.out0 = out;
  return d;
}

// MirrorProxKernel.jl:144-189 (columns): weights r_i exp(-C/(2W reg) - logZ_i)
PassDesc bary_lamp_cols_desc(const double* x, int64_t n, const double* y,
                             int64_t m, const double* r, const double* logZ,
                             double reg, double winf, double p, double* out) {
  PassDesc d = blank();
  d.tgt = y; d.nt = m; d.src = x; d.ns = n; d.p = p;
  d.k = -1.0 / (2.0 * winf * reg);
  d.srcA = logZ; d.alpha = -1.0;
  d.srcB = r;
  d.epi = EPI_BARY; d.out0 = out;
  return d;
}

// W_inf = max_ij C_ij (MirrorProxKernel.jl:1032-1034): max pass, device max,
// one scalar back to the host.
int device_winf(culamp_ctx* ctx, const double* x, int64_t n, const double* y,
                int64_t m, double p, double* tmp_n, double* winf) {
  TRY(run_pass(ctx, max_desc(x, n, y, m, p, tmp_n), CULAMP_FP64, KIND_MAX));
  ReduceSpec rs{};
  rs.nops = 1; rs.op[0] = 3; rs.a[0] = tmp_n; rs.len[0] = n;
  double v[1];
  TRY(reduce_to_host(ctx, rs, v));
  *winf = v[0];
  return CULAMP_OK;
}

// W_inf of a problem: point clouds as max_logsumexp_spp_ct!
// (MirrorProxKernel.jl:280-326); feature problems as max_logsumexp_cellsim!
// (MirrorProxCellSim.jl:1-41), whose running max starts at 0.
int problem_winf(culamp_ctx* ctx, culamp_problem* pb, double p, double* tmp_n, double* winf) {
  if (pb->metric < 0) return device_winf(ctx, pb->x, pb->n, pb->y, pb->m, p, tmp_n, winf);
  PassDesc d = max_desc(nullptr, pb->n, nullptr, pb->m, p, tmp_n);
  feat(d, pb, true);
  TRY(run_pass(ctx, d, CULAMP_FP64, KIND_MAX));
  ReduceSpec rs{};
  rs.nops = 1; rs.op[0] = 3; rs.a[0] = tmp_n; rs.len[0] = pb->n;
  double v[1];
  TRY(reduce_to_host(ctx, rs, v));
  *winf = std::fmax(v[0], 0.0);
  return CULAMP_OK;
}

// The reduction writes its scalars straight into coherent pinned memory:
// reduce_start enqueues it and records ctx->ev_chk, reduce_finish waits for
// that event and reads them.
int reduce_start(culamp_ctx* ctx, const ReduceSpec& rs) {
  HIP_TRY(launch_reduce(rs, ctx->chk_dev, ctx->stream));
  HIP_TRY(hipEventRecord(ctx->ev_chk, ctx->stream));
  return CULAMP_OK;
}

int reduce_finish(culamp_ctx* ctx, int nops, double* out) {
  HIP_TRY(hipEventSynchronize(ctx->ev_chk));
  for (int k = 0; k < nops; ++k) out[k] = ((volatile double*)ctx->chk_host)[k];
  return CULAMP_OK;
}

int reduce_to_host(culamp_ctx* ctx, const ReduceSpec& rs, double* out) {
  TRY(reduce_start(ctx, rs));
  return reduce_finish(ctx, rs.nops, out);
}

int copy_out(culamp_ctx* ctx, double* host, const double* dev, int64_t n) {
  if (!host || n <= 0) return CULAMP_OK;
  HIP_TRY(hipMemcpyAsync(host, dev, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  return CULAMP_OK;
}

// Sinkhorn kernels other than the fused LSE only know p == 1 or squared L2
// (SinkhornKernel.jl:38-42, 259-263, 146-150).
double sk_aux_p(double p) { return p == 1.0 ? 1.0 : 2.0; }

struct SkState {
  double* phi; double* psi; double* res; double* cost;
};

// One sinkhorn_color_transfer solve on device-resident state
// (SinkhornKernel.jl:477-559); phi/psi hold the warm start on entry.
// For feature problems this is sinkhorn_cellsim (SinkhornCellSim.jl:240-310).
int sinkhorn_core(culamp_ctx* ctx, culamp_problem* pb, const culamp_args* a,
                  int64_t frequency, double p, culamp_log_fn log_fn, void* user,
                  SkState& st, culamp_result* out) {
  const int64_t n = pb->n, m = pb->m;
  const bool cellsim = pb->metric >= 0;
  const double t0 = now_s();
  const double eta = a->eta_p;
  double winf;
  TRY(problem_winf(ctx, pb, p, st.res, &winf));
  const double p_lse = (p == INFINITY) ? sk_aux_p(p) : p;
  const double p_aux = sk_aux_p(p);
  const char* tag = cellsim ? ",sinkhorn_cellsim\n" : ",sinkhorn_kernel\n";
  // Staged source records (mixed point problems): each half step's epilogue
  // writes the FP32 source records of the other half step (see PassDesc).
  PrecisionGuard guard(a->precision, ctx->fp64_fallback != 0);
  bool use_rec = ctx->records && a->precision == CULAMP_MIXED && !cellsim;
  DevBufs RB;
  float4 *rec_x = nullptr, *rec_y = nullptr;
  float *recw_x = nullptr, *recw_y = nullptr;
  if (use_rec) {
    double *bx, *by, *wx, *wy;
    TRY(RB.alloc(&bx, 2 * n)); TRY(RB.alloc(&by, 2 * m));
    TRY(RB.alloc(&wx, (n + 1) / 2)); TRY(RB.alloc(&wy, (m + 1) / 2));
    rec_x = reinterpret_cast<float4*>(bx); rec_y = reinterpret_cast<float4*>(by);
    recw_x = reinterpret_cast<float*>(wx); recw_y = reinterpret_cast<float*>(wy);
  }
  bool recy_ok = false;  // rec_y holds the records of psi
  const int rec_scaled = cost_of_p(p_lse) == COST_SQL2;
  auto half_rows = [&]() {
    PassDesc d1 = sk_half_desc(pb->x, n, pb->y, m, pb->r, st.psi, eta, winf, p_lse, st.phi);
    feat(d1, pb, true);
    d1.prev_offset = 1;
    if (use_rec) {
      const PassDesc nx = sk_half_desc(pb->y, m, pb->x, n, pb->c, st.phi, eta, winf, p_lse, st.psi);
      if (recy_ok) { d1.rec = rec_y; d1.recw = recw_y; }
      d1.rec_out = rec_x; d1.recw_out = recw_x;
      d1.rec_alpha = nx.alpha; d1.rec_k = nx.k; d1.rec_scaled = rec_scaled; d1.rec_wsrc = nx.srcB;
    }
    return run_pass(ctx, d1, guard.prec, KIND_LSE);
  };
  auto half_cols = [&]() {
    PassDesc d2 = sk_half_desc(pb->y, m, pb->x, n, pb->c, st.phi, eta, winf, p_lse, st.psi);
    feat(d2, pb, false);
    d2.prev_offset = 1;
    if (use_rec) {
      const PassDesc ny = sk_half_desc(pb->x, n, pb->y, m, pb->r, st.psi, eta, winf, p_lse, st.phi);
      d2.rec = rec_x; d2.recw = recw_x;
      d2.rec_out = rec_y; d2.recw_out = recw_y;
      d2.rec_alpha = ny.alpha; d2.rec_k = ny.k; d2.rec_scaled = rec_scaled; d2.rec_wsrc = ny.srcB;
    }
    const int rc = run_pass(ctx, d2, guard.prec, KIND_LSE);
    recy_ok = use_rec;
    return rc;
  };
  auto resid = [&]() {
    PassDesc d3 = sk_resid_desc(pb->x, n, pb->y, m, pb->r, st.phi, st.psi, eta, winf, p_aux, st.res, st.cost);
    feat(d3, pb, true);
    return run_pass(ctx, d3, guard.prec, KIND_SUM);
  };
  if (a->verbose) emit(log_fn, user, "time(s),iter,infeas,ot_objective,dual\n");
  // A check's sums are read back after the next iteration's half steps are
  // enqueued, so the GPU runs them while the host waits; phi and psi are
  // kept at the check, and a stop (or the time limit right after) restores
  // them, so the returned state is the check's.
  double *phi_keep, *psi_keep;
  TRY(RB.alloc(&phi_keep, n)); TRY(RB.alloc(&psi_keep, m));
  auto restore = [&]() {
    recy_ok = false;  // rec_y describes the undone half step
    return launch_copy3(st.phi, phi_keep, nullptr, nullptr, n, st.psi, psi_keep, m, ctx->stream);
  };
  int64_t num_iter = 0;
  int converged = 0;
  bool ahead = false;  // this iteration's half steps were enqueued at the last check
  for (int64_t it = 1; it <= a->itermax; ++it) {
    const double elapsed = now_s() - t0;
    if (elapsed > a->tmax) {
      if (ahead) HIP_TRY(restore());
      break;
    }
    if (!ahead) {
      TRY(half_rows());
      TRY(half_cols());
    }
    ahead = false;
    if ((it - 1) % frequency == 0) {
      TRY(resid());
      ReduceSpec rs{};
      rs.nops = 4;
      rs.op[0] = 5; rs.a[0] = st.res; rs.len[0] = n;
      rs.op[1] = 0; rs.a[1] = st.cost; rs.len[1] = n;
      rs.op[2] = 2; rs.a[2] = st.psi; rs.b[2] = pb->c; rs.len[2] = m;
      rs.op[3] = 2; rs.a[3] = st.phi; rs.b[3] = pb->r; rs.len[3] = n;
      TRY(reduce_start(ctx, rs));
      const bool spec = it < a->itermax;
      if (spec) {
        HIP_TRY(launch_copy3(phi_keep, st.phi, nullptr, nullptr, n, psi_keep, st.psi, m, ctx->stream));
        TRY(half_rows());
        TRY(half_cols());
      }
      double v[4];
      TRY(reduce_finish(ctx, 4, v));
      const double rr = v[0];
      if (a->verbose) {
        emit(log_fn, user, fmt("%.6e", elapsed) + "," + std::to_string(it) + "," +
                               fmt("%.14e", rr) + "," + fmt("%.14e", v[1]) + "," +
                               fmt("%.14e", v[2] + v[3]) + tag);
      }
      if (rr <= a->epsilon / 6.0) {
        converged = 1;
        if (spec) HIP_TRY(restore());
        break;
      }
      if (guard.observe(rr, a->epsilon / 6.0, it)) {
        // later passes in FP64; the half steps already enqueued stay mixed
        // and still write records, which the next FP64 pass does not read
        use_rec = false;
        recy_ok = false;
      }
      ahead = spec;
    }
    num_iter += 1;
  }
  TRY(resid());
  ReduceSpec rs{};
  rs.nops = 4;
  rs.op[0] = 5; rs.a[0] = st.res; rs.len[0] = n;
  rs.op[1] = 0; rs.a[1] = st.cost; rs.len[1] = n;
  rs.op[2] = 2; rs.a[2] = st.psi; rs.b[2] = pb->c; rs.len[2] = m;
  rs.op[3] = 2; rs.a[3] = st.phi; rs.b[3] = pb->r; rs.len[3] = n;
  double v[4];
  TRY(reduce_to_host(ctx, rs, v));
  if (out) {
    out->iterations = num_iter;
    out->objective = v[1];
    out->residual = v[0];
    out->primal = NAN;
    out->dual = v[2] + v[3];
    out->winf = winf;
    out->eta_t = eta;
    out->converged = converged;
    out->nonfinite = !(std::isfinite(v[0]) && std::isfinite(v[1]));
    out->elapsed_s = now_s() - t0;
    out->fp64_from = guard.fp64_from;
  }
  return CULAMP_OK;
}

int need_features(const culamp_problem* pb) {
  if (!pb) return fail(CULAMP_ERR_INVALID_ARG, "null problem");
  if (pb->metric < 0) return fail(CULAMP_ERR_INVALID_ARG, "not a feature problem (use culamp_feature_problem_create)");
  return CULAMP_OK;
}

}  // namespace culamp_host

using namespace culamp_host;

// ============================================================================
extern "C" {

int culamp_abi_version(void) { return CULAMP_ABI_VERSION; }

const char* culamp_last_error(void) { return g_err.c_str(); }

// EOTArgs defaults (src/Utilities.jl:18-31).
void culamp_args_default(culamp_args* a) {
  if (!a) return;
  a->eta_p = 0.0; a->eta_mu = 0.0; a->tau_p = 1.0; a->tau_mu = 1.0;
  a->alpha = 0.01; a->B = 1.0; a->epsilon = 1e-4; a->anneal_mult = 0.95;
  a->itermax = 10000; a->inner_iter = 10; a->tmax = INFINITY;
  a->verbose = 1; a->precision = CULAMP_FP64;
  a->stop = CULAMP_STOP_REFERENCE; a->reserved = 0;
}

int culamp_device_count(int* count) {
  if (!count) return fail(CULAMP_ERR_INVALID_ARG, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return CULAMP_OK;
}

int culamp_ctx_create(int device, culamp_ctx** out) {
  if (!out) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0)
    return fail(CULAMP_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= cnt)
    return fail(CULAMP_ERR_INVALID_ARG, "device %d out of range [0, %d)", device, cnt);
  culamp_ctx* ctx = new culamp_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
  if (e == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
  if (e == hipSuccess) e = hipMalloc(&ctx->scal, 64 * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&ctx->scal_host, 64 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_chk, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipHostMalloc(&ctx->chk_host, 64 * sizeof(double), hipHostMallocCoherent | hipHostMallocMapped);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&ctx->chk_dev, ctx->chk_host, 0);
  if (e != hipSuccess) {
    culamp_ctx_destroy(ctx);
    return fail(CULAMP_ERR_HIP, "context creation failed: %s", hipGetErrorString(e));
  }
  culamp_noalloc_setup(device);  // (the device is current)
  *out = ctx;
  return CULAMP_OK;
}

int culamp_ctx_destroy(culamp_ctx* ctx) {
  if (!ctx) return CULAMP_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  culamp_destroy_batches(ctx);
  while (!ctx->denses.empty()) culamp_dense_destroy(*ctx->denses.begin());
  while (!ctx->problems.empty()) culamp_problem_destroy(*ctx->problems.begin());
  if (ctx->partial) (void)hipFree(ctx->partial);
  if (ctx->scal) (void)hipFree(ctx->scal);
  if (ctx->scal_host) (void)hipHostFree(ctx->scal_host);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev_chk) (void)hipEventDestroy(ctx->ev_chk);
  if (ctx->chk_host) (void)hipHostFree(ctx->chk_host);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CULAMP_OK;
}

// Launch-shape override for tests and tuning: 0 = automatic.
int culamp_ctx_set_launch(culamp_ctx* ctx, int rows_per_thread, int splits) {
  if (!ctx) return fail(CULAMP_ERR_INVALID_ARG, "null context");
  if (rows_per_thread != 0 && rows_per_thread != 1 && rows_per_thread != 2 && rows_per_thread != 4)
    return fail(CULAMP_ERR_INVALID_ARG, "rows_per_thread must be 0, 1, 2 or 4");
  if (splits < 0) return fail(CULAMP_ERR_INVALID_ARG, "splits must be >= 0");
  ctx->force_rpt = rows_per_thread;
  ctx->force_splits = splits;
  return CULAMP_OK;
}

int culamp_ctx_set_layout(culamp_ctx* ctx, int layout) {
  TRY(check_ctx(ctx));
  if (layout != LAYOUT_AUTO && layout != LAYOUT_SPLIT && layout != LAYOUT_WIDE)
    return fail(CULAMP_ERR_INVALID_ARG, "layout must be 0 (auto), 1 (split) or 2 (wide)");
  ctx->layout = layout;
  return CULAMP_OK;
}

int culamp_ctx_set_source_records(culamp_ctx* ctx, int enable) {
  TRY(check_ctx(ctx));
  if (enable != 0 && enable != 1) return fail(CULAMP_ERR_INVALID_ARG, "enable must be 0 or 1");
  ctx->records = enable;
  return CULAMP_OK;
}

int culamp_ctx_set_fp64_fallback(culamp_ctx* ctx, int enable) {
  if (!ctx) return fail(CULAMP_ERR_INVALID_ARG, "null context");
  ctx->fp64_fallback = enable ? 1 : 0;
  return CULAMP_OK;
}

int culamp_ctx_synchronize(culamp_ctx* ctx) {
  TRY(check_ctx(ctx));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return CULAMP_OK;
}

int culamp_malloc(culamp_ctx* ctx, size_t bytes, void** dptr) {
  TRY(check_ctx(ctx));
  if (!dptr) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  hipError_t e = hipMalloc(dptr, bytes > 0 ? bytes : 1);
  if (e != hipSuccess) return fail(CULAMP_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return CULAMP_OK;
}

int culamp_free(culamp_ctx* ctx, void* dptr) {
  TRY(check_ctx(ctx));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (dptr) HIP_TRY(hipFree(dptr));
  return CULAMP_OK;
}

int culamp_memcpy_htod(culamp_ctx* ctx, void* dst, const void* src, size_t bytes) {
  TRY(check_ctx(ctx));
  if (bytes == 0) return CULAMP_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return CULAMP_OK;
}

int culamp_memcpy_dtoh(culamp_ctx* ctx, void* dst, const void* src, size_t bytes) {
  TRY(check_ctx(ctx));
  if (bytes == 0) return CULAMP_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return CULAMP_OK;
}

int culamp_memcpy_dtod(culamp_ctx* ctx, void* dst, const void* src, size_t bytes) {
  TRY(check_ctx(ctx));
  if (bytes == 0) return CULAMP_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return CULAMP_OK;
}

int culamp_timer_start(culamp_ctx* ctx) {
  TRY(check_ctx(ctx));
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  return CULAMP_OK;
}

int culamp_timer_stop(culamp_ctx* ctx, double* ms) {
  TRY(check_ctx(ctx));
  if (!ms) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(hipEventSynchronize(ctx->ev1));
  float f = 0.f;
  HIP_TRY(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  *ms = (double)f;
  return CULAMP_OK;
}

// ---- stream capture ---------------------------------------------------------
struct culamp_graph {
  hipGraphExec_t exec = nullptr;
};

int culamp_capture_begin(culamp_ctx* ctx) {
  TRY(check_ctx(ctx));
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  return CULAMP_OK;
}

int culamp_capture_end(culamp_ctx* ctx, culamp_graph** out) {
  TRY(check_ctx(ctx));
  if (!out) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamEndCapture(ctx->stream, &g));
  culamp_graph* gr = new culamp_graph();
  const hipError_t e = hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    delete gr;
    return fail(CULAMP_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  *out = gr;
  return CULAMP_OK;
}

int culamp_graph_launch(culamp_ctx* ctx, culamp_graph* g) {
  TRY(check_ctx(ctx));
  if (!g) return fail(CULAMP_ERR_INVALID_ARG, "null graph");
  HIP_TRY(hipGraphLaunch(g->exec, ctx->stream));
  return CULAMP_OK;
}

int culamp_graph_destroy(culamp_graph* g) {
  if (!g) return CULAMP_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  delete g;
  return CULAMP_OK;
}

// ---- kernel level -----------------------------------------------------------

int culamp_max_cost(culamp_ctx* ctx, const double* x, int64_t n, const double* y,
                    int64_t m, double p, double* out) {
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, max_desc(x, n, y, m, p, out), CULAMP_FP64, KIND_MAX);
}

int culamp_lse_rows(culamp_ctx* ctx, const double* x, int64_t n, const double* y,
                    int64_t m, const double* theta, double reg, double st,
                    double winf, double p, int precision, double* out) {
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, lse_rows_desc(x, n, y, m, theta, reg, st, winf, p, out), precision, KIND_LSE);
}

int culamp_col_marginal_cost(culamp_ctx* ctx, const double* x, int64_t n,
                             const double* y, int64_t m, const double* r,
                             const double* theta, const double* logZ, double reg,
                             double st, double winf, double p, int precision,
                             double* res, double* cost) {
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, colmarg_desc(x, n, y, m, r, theta, logZ, reg, st, winf, p, res, cost), precision, KIND_SUM);
}

int culamp_minplus_rows(culamp_ctx* ctx, const double* x, int64_t n,
                        const double* y, int64_t m, const double* theta,
                        double winf, double p, double* out) {
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, minplus_desc(x, n, y, m, theta, winf, p, out), CULAMP_FP64, KIND_MIN);
}

int culamp_lamp_theta_update(culamp_ctx* ctx, double* theta, const double* theta0,
                             const double* res, const double* c,
                             const double* eta_mu_i, int64_t m, double eta_mu,
                             int adjust, double minv, double maxv, double winf) {
  TRY(check_ctx(ctx));
  TRY(check_sizes(m, 0));
  HIP_TRY(launch_theta_update(theta, theta0, res, c, eta_mu_i, m, eta_mu, adjust, minv, maxv, winf, ctx->stream));
  return CULAMP_OK;
}

int culamp_sinkhorn_half_step(culamp_ctx* ctx, const double* x, int64_t n,
                              const double* y, int64_t m, const double* marg,
                              const double* psi, double reg, double winf,
                              double p, int precision, double* out) {
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, sk_half_desc(x, n, y, m, marg, psi, reg, winf, p, out), precision, KIND_LSE);
}

int culamp_row_residual_cost(culamp_ctx* ctx, const double* x, int64_t n,
                             const double* y, int64_t m, const double* marg,
                             const double* phi, const double* psi, double reg,
                             double winf, double p, int precision, double* res,
                             double* cost) {
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, sk_resid_desc(x, n, y, m, marg, phi, psi, reg, winf, p, res, cost), precision, KIND_SUM);
}

int culamp_barycentric_sinkhorn(culamp_ctx* ctx, const double* x, int64_t n,
                                const double* y, int64_t m, const double* phi,
                                const double* psi, double reg, double winf,
                                double p, double* out) {
  (void)phi;  // phi_i scales every weight of row i and cancels in the mean
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, bary_sk_desc(x, n, y, m, psi, reg, winf, p, out), CULAMP_FP64, KIND_BARY);
}

int culamp_barycentric_lamp_rows(culamp_ctx* ctx, const double* x, int64_t n,
                                 const double* y, int64_t m, const double* nu,
                                 const double* r, const double* logZ,
                                 double reg, double winf, double p, double* out) {
  (void)r; (void)logZ;  // per-row factors cancel in the mean
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, bary_lamp_rows_desc(x, n, y, m, nu, reg, winf, p, out), CULAMP_FP64, KIND_BARY);
}

int culamp_barycentric_lamp_cols(culamp_ctx* ctx, const double* x, int64_t n,
                                 const double* y, int64_t m, const double* r,
                                 const double* nu, const double* logZ,
                                 double reg, double winf, double p, double* out) {
  (void)nu;  // nu_j scales every weight of column j and cancels in the mean
  TRY(check_ctx(ctx));
  TRY(check_sizes(n, m));
  return run_pass(ctx, bary_lamp_cols_desc(x, n, y, m, r, logZ, reg, winf, p, out), CULAMP_FP64, KIND_BARY);
}

// ---- problems ---------------------------------------------------------------

int culamp_problem_create(culamp_ctx* ctx, const double* x, int64_t n,
                          const double* y, int64_t m, const double* r,
                          const double* c, culamp_problem** out) {
  TRY(check_ctx(ctx));
  if (!out) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  if (n <= 0 || m <= 0) return fail(CULAMP_ERR_INVALID_ARG, "problem sizes must be positive (n=%lld, m=%lld)", (long long)n, (long long)m);
  if (!x || !y || !r || !c) return fail(CULAMP_ERR_INVALID_ARG, "null input array");
  culamp_problem* pb = new culamp_problem();
  pb->ctx = ctx; pb->n = n; pb->m = m;
  pb->hr.assign(r, r + n);
  pb->hc.assign(c, c + m);
  hipError_t e = hipSuccess;
  auto al = [&](double** p, int64_t len) { if (e == hipSuccess) e = hipMalloc(p, sizeof(double) * (size_t)len); };
  al(&pb->x, 3 * n); al(&pb->y, 3 * m); al(&pb->r, n); al(&pb->c, m); al(&pb->logr, n); al(&pb->logc, m);
  if (e == hipSuccess) e = hipMemcpyAsync(pb->x, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pb->y, y, sizeof(double) * 3 * m, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pb->r, r, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pb->c, c, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_log(pb->logr, pb->r, n, ctx->stream);
  if (e == hipSuccess) e = launch_log(pb->logc, pb->c, m, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    culamp_problem_destroy(pb);
    return fail(CULAMP_ERR_HIP, "problem upload failed: %s", hipGetErrorString(e));
  }
  ctx->problems.insert(pb);
  *out = pb;
  return CULAMP_OK;
}

int culamp_problem_destroy(culamp_problem* pb) {
  if (!pb) return CULAMP_OK;
  pb->ctx->problems.erase(pb);
  (void)hipSetDevice(pb->ctx->device);
  (void)hipStreamSynchronize(pb->ctx->stream);
  if (pb->geom_owner) {
    for (double* p : {pb->r, pb->c, pb->logr, pb->logc})
      if (p) (void)hipFree(p);
    delete pb;
    return CULAMP_OK;
  }
  for (double* p : {pb->x, pb->y, pb->r, pb->c, pb->logr, pb->logc, pb->xf64, pb->xrc})
    if (p) (void)hipFree(p);
  if (pb->xf32) (void)hipFree(pb->xf32);
  for (uint16_t* p : {pb->xz_hi, pb->xz_lo})
    if (p) (void)hipFree(p);
  if (!pb->self) {
    for (double* p : {pb->yf64, pb->yrc})
      if (p) (void)hipFree(p);
    if (pb->yf32) (void)hipFree(pb->yf32);
    for (uint16_t* p : {pb->yz_hi, pb->yz_lo})
      if (p) (void)hipFree(p);
  }
  delete pb;
  return CULAMP_OK;
}

// get_flat_metrics (src/CellUtilities.jl:304-311) on an n x d column-major
// matrix, then the packed forms the feature kernel reads: rows divided by
// their sums, and [squared norm / s^2 | mean / s | centred squared norm / s^2].
static void pack_features(const double* x, int64_t n, int64_t d, bool scaled,
                          std::vector<double>& f64, std::vector<float>& f32,
                          std::vector<double>& rc) {
  f64.resize((size_t)n * d);
  f32.resize((size_t)n * d);
  rc.resize((size_t)4 * n);
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0, q = 0.0;
    for (int64_t k = 0; k < d; ++k) {
      const double v = x[i + k * n];
      s += v;
      q += v * v;
    }
    const double mu = s / (double)d;
    double cs = 0.0;
    for (int64_t k = 0; k < d; ++k) {
      const double v = x[i + k * n] - mu;
      cs += v * v;
    }
    for (int64_t k = 0; k < d; ++k) {
      const double v = scaled ? x[i + k * n] / s : x[i + k * n];
      f64[i + k * n] = v;
      f32[i + k * n] = (float)v;
    }
    rc[i] = q / (s * s);
    rc[n + i] = mu / s;
    rc[2 * n + i] = cs / (s * s);
    rc[3 * n + i] = s;
  }
}

static uint16_t bf16_rne(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float bf16_to_float(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Unit rows for the MFMA path: z_i = sign(s_i) x_i / |x_i| (cosine) or
// sign(s_i) (x_i - mean_i) / |x_i - mean_i| (correlation), whose dot is the
// reference's similarity, each scaled by kUnitRowScale before rounding so
// that the dot of two rounded rows stays inside [-1, 1] with no clamp (the
// kernel divides kUnitRowScale^2 back out).  Returns false when some row has
// no direction (the reference's zero-denominator branches), or d > 128.
static bool pack_unit_rows(const double* x, int64_t n, int64_t d, int metric, int ks,
                           std::vector<uint16_t>& hi, std::vector<uint16_t>& lo) {
  const int64_t dp = 16 * ks;
  hi.assign((size_t)n * dp, 0);
  lo.assign((size_t)n * dp, 0);
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = 0; k < d; ++k) s += x[i + k * n];
    const double mu = metric == CULAMP_METRIC_CORRELATION ? s / (double)d : 0.0;
    double q = 0.0;
    for (int64_t k = 0; k < d; ++k) {
      const double v = x[i + k * n] - mu;
      q += v * v;
    }
    if (!(q > 0.0) || !(std::isfinite(s)) || s == 0.0) return false;
    const double f = (s > 0 ? kUnitRowScale : -kUnitRowScale) / std::sqrt(q);
    for (int64_t k = 0; k < d; ++k) {
      const double z = (x[i + k * n] - mu) * f;
      const uint16_t h = bf16_rne((float)z);
      hi[(size_t)i * dp + k] = h;
      lo[(size_t)i * dp + k] = bf16_rne((float)(z - (double)bf16_to_float(h)));
    }
  }
  return true;
}

int culamp_feature_problem_create(culamp_ctx* ctx, const double* x, int64_t n,
                                  const double* y, int64_t m, int64_t d, int metric,
                                  double scale, const double* r, const double* c,
                                  culamp_problem** out) {
  TRY(check_ctx(ctx));
  if (!out) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  if (!x || !r || !c) return fail(CULAMP_ERR_INVALID_ARG, "null input array");
  const bool self = (y == nullptr || y == x);
  if (self) m = n;
  if (n <= 0 || m <= 0 || d <= 0)
    return fail(CULAMP_ERR_INVALID_ARG, "feature problem sizes must be positive (n=%lld, m=%lld, d=%lld)",
                (long long)n, (long long)m, (long long)d);
  if (metric < CULAMP_METRIC_L1 || metric > CULAMP_METRIC_CORRELATION)
    return fail(CULAMP_ERR_INVALID_ARG, "unknown metric %d", metric);
  if (!(scale > 0.0)) return fail(CULAMP_ERR_INVALID_ARG, "scale must be positive");
  culamp_problem* pb = new culamp_problem();
  pb->ctx = ctx; pb->n = n; pb->m = m; pb->d = d; pb->metric = metric; pb->scale = scale;
  pb->self = self;
  pb->hr.assign(r, r + n);
  pb->hc.assign(c, c + m);
  std::vector<double> xf, xrc, yf, yrc;
  std::vector<float> xf32, yf32;
  // l1 reads x_ik / s_i per element; the dot metrics read raw rows and divide
  // the dot by s_i s_j (CellUtilities.jl:213-220, 233-238)
  const bool scaled = metric == CULAMP_METRIC_L1;
  pack_features(x, n, d, scaled, xf, xf32, xrc);
  if (!self) pack_features(y, m, d, scaled, yf, yf32, yrc);
  hipError_t e = hipSuccess;
  auto al = [&](void** p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(p, bytes); };
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  };
  al((void**)&pb->r, sizeof(double) * n); al((void**)&pb->c, sizeof(double) * m);
  al((void**)&pb->logr, sizeof(double) * n); al((void**)&pb->logc, sizeof(double) * m);
  al((void**)&pb->xf64, sizeof(double) * n * d); al((void**)&pb->xf32, sizeof(float) * n * d);
  al((void**)&pb->xrc, sizeof(double) * 4 * n);
  if (!self) {
    al((void**)&pb->yf64, sizeof(double) * m * d); al((void**)&pb->yf32, sizeof(float) * m * d);
    al((void**)&pb->yrc, sizeof(double) * 4 * m);
  }
  up(pb->r, r, sizeof(double) * n); up(pb->c, c, sizeof(double) * m);
  up(pb->xf64, xf.data(), sizeof(double) * n * d); up(pb->xf32, xf32.data(), sizeof(float) * n * d);
  up(pb->xrc, xrc.data(), sizeof(double) * 4 * n);
  if (!self) {
    up(pb->yf64, yf.data(), sizeof(double) * m * d); up(pb->yf32, yf32.data(), sizeof(float) * m * d);
    up(pb->yrc, yrc.data(), sizeof(double) * 4 * m);
  } else {
    pb->yf64 = pb->xf64; pb->yf32 = pb->xf32; pb->yrc = pb->xrc;
  }
  std::vector<uint16_t> xzh, xzl, yzh, yzl;
  if ((metric == CULAMP_METRIC_COSINE || metric == CULAMP_METRIC_CORRELATION) && d <= 128) {
    int ks = 1;
    while (16 * ks < d) ks *= 2;
    bool ok = pack_unit_rows(x, n, d, metric, ks, xzh, xzl);
    if (ok && !self) ok = pack_unit_rows(y, m, d, metric, ks, yzh, yzl);
    if (ok) {
      const size_t dp = (size_t)16 * ks;
      pb->zks = ks;
      al((void**)&pb->xz_hi, 2 * n * dp); al((void**)&pb->xz_lo, 2 * n * dp);
      up(pb->xz_hi, xzh.data(), 2 * n * dp); up(pb->xz_lo, xzl.data(), 2 * n * dp);
      if (!self) {
        al((void**)&pb->yz_hi, 2 * m * dp); al((void**)&pb->yz_lo, 2 * m * dp);
        up(pb->yz_hi, yzh.data(), 2 * m * dp); up(pb->yz_lo, yzl.data(), 2 * m * dp);
      } else {
        pb->yz_hi = pb->xz_hi; pb->yz_lo = pb->xz_lo;
      }
    }
  }
  if (e == hipSuccess) e = launch_log(pb->logr, pb->r, n, ctx->stream);
  if (e == hipSuccess) e = launch_log(pb->logc, pb->c, m, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    culamp_problem_destroy(pb);
    return fail(CULAMP_ERR_HIP, "feature problem upload failed: %s", hipGetErrorString(e));
  }
  ctx->problems.insert(pb);
  *out = pb;
  return CULAMP_OK;
}


// ---- feature kernel level (device vectors, problem-owned packed rows) -------

// warp_logsumexp_spp_sim_fused! (src/solvers/MirrorProxCellSim.jl:259-342)
int culamp_feature_lse_rows(culamp_ctx* ctx, culamp_problem* pb, const double* theta,
                            double reg, double st, double winf, int precision, double* out) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  PassDesc d = lse_rows_desc(nullptr, pb->n, nullptr, pb->m, theta, reg, st, winf, 2.0, out);
  feat(d, pb, true);
  return run_pass(ctx, d, precision, KIND_LSE);
}

// residual_cellsim_c! (MirrorProxCellSim.jl:190-257)
int culamp_feature_col_marginal_cost(culamp_ctx* ctx, culamp_problem* pb, const double* r,
                                     const double* theta, const double* logZ, double reg,
                                     double st, double winf, int precision, double* res,
                                     double* cost) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  PassDesc d = colmarg_desc(nullptr, pb->n, nullptr, pb->m, r, theta, logZ, reg, st, winf, 2.0, res, cost);
  feat(d, pb, false);
  return run_pass(ctx, d, precision, KIND_SUM);
}

// warp_logsumexp_cellsim_fused! (src/solvers/SinkhornCellSim.jl:153-238):
// rows != 0 updates the row potential from the column potential `pot`,
// rows == 0 the column potential from the row potential.
int culamp_feature_sinkhorn_half_step(culamp_ctx* ctx, culamp_problem* pb, int rows,
                                      const double* marg, const double* pot, double reg,
                                      double winf, int precision, double* out) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  PassDesc d = sk_half_desc(nullptr, rows ? pb->n : pb->m, nullptr, rows ? pb->m : pb->n,
                            marg, pot, reg, winf, 2.0, out);
  feat(d, pb, rows != 0);
  return run_pass(ctx, d, precision, KIND_LSE);
}

// residual_cellsim_opt! (SinkhornCellSim.jl:1-72)
int culamp_feature_row_residual_cost(culamp_ctx* ctx, culamp_problem* pb, const double* marg,
                                     const double* phi, const double* psi, double reg,
                                     double winf, int precision, double* res, double* cost) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  PassDesc d = sk_resid_desc(nullptr, pb->n, nullptr, pb->m, marg, phi, psi, reg, winf, 2.0, res, cost);
  feat(d, pb, true);
  return run_pass(ctx, d, precision, KIND_SUM);
}

// max_logsumexp_cellsim! (MirrorProxCellSim.jl:1-41): out_i = max(0, max_j C_ij)
int culamp_feature_max_cost(culamp_ctx* ctx, culamp_problem* pb, double* out) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  PassDesc d = max_desc(nullptr, pb->n, nullptr, pb->m, 2.0, out);
  feat(d, pb, true);
  return run_pass(ctx, d, CULAMP_FP64, KIND_MAX);
}

// warp_min_reduce_cellsim! (MirrorProxCellSim.jl:43-92)
int culamp_feature_minplus_rows(culamp_ctx* ctx, culamp_problem* pb, const double* theta,
                                double winf, double* out) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  PassDesc d = minplus_desc(nullptr, pb->n, nullptr, pb->m, theta, winf, 2.0, out);
  feat(d, pb, true);
  return run_pass(ctx, d, CULAMP_FP64, KIND_MIN);
}

// ---- solvers ----------------------------------------------------------------

// extragradient_color_transfer (src/solvers/MirrorProxKernel.jl:1019-1132).
int culamp_lamp_solve(culamp_ctx* ctx, culamp_problem* pb, const culamp_args* a,
                      int64_t frequency, double p, culamp_log_fn log_fn,
                      void* user, double* nu_out, double* phi_out,
                      double* psi_out, double* img1, double* img2,
                      culamp_result* out) {
  TRY(check_ctx(ctx));
  if (!pb || !a) return fail(CULAMP_ERR_INVALID_ARG, "null problem or args");
  if (pb->metric >= 0) return fail(CULAMP_ERR_INVALID_ARG, "feature problem: use the cellsim solvers");
  if (frequency <= 0) return fail(CULAMP_ERR_INVALID_ARG, "frequency must be positive");
  const int64_t n = pb->n, m = pb->m;
  PrecisionGuard guard(a->precision, ctx->fp64_fallback != 0);
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  DevBufs B;
  double *theta, *nu, *thetab, *nub, *res, *cost, *etamu, *logZ, *tmp, *nu_keep, *theta_keep;
  TRY(B.alloc(&theta, m)); TRY(B.alloc(&nu, m)); TRY(B.alloc(&thetab, m));
  TRY(B.alloc(&nub, m)); TRY(B.alloc(&res, m)); TRY(B.alloc(&cost, m));
  TRY(B.alloc(&etamu, m)); TRY(B.alloc(&logZ, n)); TRY(B.alloc(&tmp, n > m ? n : m));
  double* logZ_keep;
  TRY(B.alloc(&nu_keep, m)); TRY(B.alloc(&theta_keep, m)); TRY(B.alloc(&logZ_keep, n));
  for (double* v : {theta, nu, thetab, nub, res, cost}) HIP_TRY(launch_fill(v, 0.0, m, ctx->stream));
  HIP_TRY(launch_fill(logZ, 0.0, n, ctx->stream));
  const double t0 = now_s();
  double winf;
  TRY(device_winf(ctx, pb->x, n, pb->y, m, p, tmp, &winf));
  const double eta = a->eta_p / (2.0 * winf);                             // 1035
  HIP_TRY(launch_etamu(etamu, pb->c, a->alpha / (double)n, a->tau_mu, m, ctx->stream));  // 1037
  double negent = 0.0;
  for (double v : pb->hr) if (v > 0.0) negent += v * std::log(v);
  const double hr = eta * negent;                                         // 1066
  emit(log_fn, user, "time(s),iter,infeas,ot_objective,primal,dual,solver\n");
  double st = 1.0, etat = INFINITY;
  const double taup = 1.0;  // hard-coded in the reference (1073)
  const double minv = std::tanh(-a->B / 2.0), maxv = std::tanh(a->B / 2.0);
  double primal = NAN, dual = NAN, resid = NAN, objective = NAN;
  int converged = 0;
  int64_t steps = 0;

  // Staged source records (mixed precision): each iteration pass's epilogue
  // writes the FP32 source records of the pass after it (the row LSE for the
  // column pass over x, the column pass for the next row LSE over y), so
  // those passes skip converting their sources.  recy_ok: rec_y holds the
  // records of lse(nu or nub, etat, st) for the current etat and st.
  bool use_rec = ctx->records && a->precision == CULAMP_MIXED;
  float4 *rec_x = nullptr, *rec_y = nullptr;
  float *recw_x = nullptr, *recw_y = nullptr;
  if (use_rec) {
    double *bx, *by, *wx, *wy;
    TRY(B.alloc(&bx, 2 * n)); TRY(B.alloc(&by, 2 * m));
    TRY(B.alloc(&wx, (n + 1) / 2)); TRY(B.alloc(&wy, (m + 1) / 2));
    rec_x = reinterpret_cast<float4*>(bx); rec_y = reinterpret_cast<float4*>(by);
    recw_x = reinterpret_cast<float*>(wx); recw_y = reinterpret_cast<float*>(wy);
  }
  bool recy_ok = false;
  // this row LSE's epilogue: records of the column pass over x that follows
  auto emit_x = [&](PassDesc& d) {
    if (!use_rec) return;
    const PassDesc nx = colmarg_desc(pb->x, n, pb->y, m, pb->r, nu, logZ, 1.0, 1.0, winf, p,
                                     res, cost);
    d.rec_out = rec_x; d.recw_out = recw_x;
    d.rec_alpha = nx.alpha; d.rec_k = nx.k; d.rec_scaled = 0; d.rec_wsrc = nx.srcB;
  };
  // this column pass's epilogue: records of lse(., reg, s) over y
  auto emit_y = [&](PassDesc& d, double reg, double s) {
    if (!use_rec) return;
    const PassDesc ny = lse_rows_desc(pb->x, n, pb->y, m, nu, reg, s, winf, p, logZ);
    d.rec_out = rec_y; d.recw_out = recw_y;
    d.rec_alpha = ny.alpha; d.rec_k = ny.k; d.rec_scaled = cost_of_p(p) == COST_SQL2;
    d.rec_wsrc = ny.srcB;
  };
  auto lse = [&](const double* pot, double reg, double s, double* dst, bool offset) {
    PassDesc d = lse_rows_desc(pb->x, n, pb->y, m, pot, reg, s, winf, p, dst);
    d.prev_offset = offset ? 1 : 0;
    return run_pass(ctx, d, guard.prec, KIND_LSE);
  };
  // the row LSE of the loop: lse(pot, etat, st) -> logZ, reading rec_y when
  // it is current and writing rec_x
  auto lse_step = [&](const double* pot, double reg, double s) {
    PassDesc d = lse_rows_desc(pb->x, n, pb->y, m, pot, reg, s, winf, p, logZ);
    d.prev_offset = 1;
    if (use_rec && recy_ok) { d.rec = rec_y; d.recw = recw_y; }
    emit_x(d);
    return run_pass(ctx, d, guard.prec, KIND_LSE);
  };
  auto use_x = [&](PassDesc& d) {
    if (use_rec) { d.rec = rec_x; d.recw = recw_x; }
  };
  // The column pass of the first half step, fused with the theta update and
  // the nu mixing (1103-1109); `with_cost` also writes the column costs
  // (a check needs them).
  // The dual (and its pass) is needed for the log and the reference's gap
  // stop; stopping on infeasibility without logging needs neither.
  const bool need_dual = a->verbose || a->stop != CULAMP_STOP_INFEAS;
  auto col_step1 = [&](double etat_new, double st_new, bool with_cost) {
    PassDesc d = colmarg_desc(pb->x, n, pb->y, m, pb->r, nu, logZ, etat, st, winf, p, res, cost);
    d.epi = EPI_LAMP_COL; d.fuse_update = 1; d.adjust = 0;
    d.th0 = theta; d.cvec = pb->c; d.etamu = etamu; d.eta_mu = a->eta_mu;
    d.minv = minv; d.maxv = maxv; d.upd_winf = 1.0; d.th_out = thetab;
    d.nu_in = nu; d.mix_src = theta; d.nu_out = nub; d.mix_t = etat_new;
    if (!with_cost) d.out1 = nullptr;  // the column costs are only read at checks
    use_x(d);
    emit_y(d, etat_new, st_new);  // nub, for the next row LSE
    return run_pass(ctx, d, guard.prec, KIND_SUM);
  };
  for (int64_t it = 1; it <= a->itermax; ++it) {
    const double elapsed = now_s() - t0;
    if (elapsed > a->tmax) break;
    const bool check_it = a->verbose || a->stop == CULAMP_STOP_INFEAS;
    const double etat_new = 1.0 / (taup + (1.0 / etat) * (1.0 - eta));
    const double st_new = (1.0 - etat_new) * st + etat_new;
    bool half_done = false, pending = false;
    const double etat_prev = etat, st_prev = st;
    if (check_it && (it - 1) % frequency == 0) {  // 1081-1101
      // The check's row LSE and column pass are those of the first half
      // step (same nu, eta_t and s_t): the column pass also applies the
      // step's fused update (to thetab / nub, so theta and nu are unchanged
      // if the loop stops here) and the step skips both.
      TRY(lse_step(nu, etat, st));
      TRY(col_step1(etat_new, st_new, true));
      half_done = true;
      if (need_dual) {
        if (eta > 0) {
          TRY(lse(theta, eta, 1.0, tmp, false));
        } else {
          TRY(run_pass(ctx, minplus_desc(pb->x, n, pb->y, m, theta, winf, p, tmp), CULAMP_FP64, KIND_MIN));
        }
      }
      ReduceSpec rs{};
      rs.nops = need_dual ? 6 : 4;
      rs.op[0] = 1; rs.a[0] = res; rs.b[0] = pb->c; rs.len[0] = m;
      rs.op[1] = 0; rs.a[1] = cost; rs.len[1] = m;
      rs.op[2] = 2; rs.a[2] = res; rs.b[2] = nu; rs.len[2] = m;
      rs.op[3] = 2; rs.a[3] = pb->r; rs.b[3] = logZ; rs.len[3] = n;
      rs.op[4] = 2; rs.a[4] = pb->r; rs.b[4] = tmp; rs.len[4] = n;
      rs.op[5] = 2; rs.a[5] = pb->c; rs.b[5] = theta; rs.len[5] = m;
      TRY(reduce_start(ctx, rs));
      // The stop test is read back after the second half step is enqueued,
      // so the GPU runs that half step while the host waits; nu, theta and
      // logZ (the next row LSE's centring offsets) are kept here so that a
      // stop can restore them.
      HIP_TRY(launch_copy3(nu_keep, nu, theta_keep, theta, m, logZ_keep, logZ, n, ctx->stream));
      pending = true;
    }
    // extragradient step (1103-1115); the theta update and the nu mixing are
    // fused into the column-marginal pass epilogue.
    if (!half_done) {
      TRY(lse_step(nu, etat, st));
      TRY(col_step1(etat_new, st_new, false));
    }
    etat = etat_new;
    st = st_new;
    recy_ok = use_rec;
    TRY(lse_step(nub, etat, st));
    {
      PassDesc d = colmarg_desc(pb->x, n, pb->y, m, pb->r, nub, logZ, etat, st, winf, p, res, cost);
      d.epi = EPI_LAMP_COL; d.fuse_update = 1; d.adjust = 1;
      d.th0 = theta; d.cvec = pb->c; d.etamu = etamu; d.eta_mu = a->eta_mu;
      d.minv = minv; d.maxv = maxv; d.upd_winf = 1.0; d.th_out = theta;
      d.nu_in = nu; d.mix_src = thetab; d.nu_out = nu; d.mix_t = etat;
      d.out1 = nullptr;
      use_x(d);
      emit_y(d, etat, st);  // nu, for the next iteration's row LSE
      TRY(run_pass(ctx, d, guard.prec, KIND_SUM));
    }
    if (pending) {
      double v[6];
      TRY(reduce_finish(ctx, need_dual ? 6 : 4, v));
      resid = v[0];
      objective = v[1];
      const double q = eta / etat_prev;
      primal = objective * (1.0 - q) - q * v[2] + q * v[3] + 2.0 * winf * resid;
      if (!need_dual) dual = NAN;
      else if (eta > 0) dual = -2.0 * eta * v[4] - 2.0 * winf * v[5] + 2.0 * hr;
      else dual = v[4] - 2.0 * winf * v[5];
      if (a->verbose)
        emit(log_fn, user, fmt("%.6e", elapsed) + "," + std::to_string(it) + "," +
                               fmt("%.14e", resid) + "," + fmt("%.14e", objective) + "," +
                               fmt("%.14e", primal) + "," + fmt("%.14e", dual) + ",lamp_kernel\n");
      // CULAMP_STOP_INFEAS never stops at the pre-step check (eta_t = inf:
      // the plan is r (x) 1/m whatever the costs, feasible when c is uniform)
      const bool met = a->stop == CULAMP_STOP_INFEAS
                           ? (std::isfinite(etat_prev) && resid < a->epsilon / 6.0)
                           : (primal - dual < a->epsilon / 6.0 && resid < a->epsilon / 6.0);
      if (met) {
        // undo the half step run past the stop: the state is the check's
        converged = 1;
        HIP_TRY(launch_copy3(nu, nu_keep, theta, theta_keep, m, logZ, logZ_keep, n, ctx->stream));
        etat = etat_prev;
        st = st_prev;
        recy_ok = false;  // rec_y describes the undone state
        break;
      }
      if (std::isfinite(etat_prev) && guard.observe(resid, a->epsilon / 6.0, it)) {
        use_rec = false;  // later passes in FP64, which read no records
        recy_ok = false;
      }
    }
    steps += 1;
  }
  // epilogue (1124-1130)
  {
    PassDesc d = lse_rows_desc(pb->x, n, pb->y, m, nu, etat, st, winf, p, logZ);
    d.prev_offset = 1;
    if (use_rec && recy_ok) { d.rec = rec_y; d.recw = recw_y; }
    TRY(run_pass(ctx, d, guard.prec, KIND_LSE));
  }
  double* img_dev = nullptr;
  if (img1 || img2) TRY(B.alloc(&img_dev, 3 * (n > m ? n : m)));
  if (img1) {
    TRY(run_pass(ctx, bary_lamp_rows_desc(pb->x, n, pb->y, m, nu, etat, winf, p, img_dev), CULAMP_FP64, KIND_BARY));
    TRY(copy_out(ctx, img1, img_dev, 3 * n));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (img2) {
    TRY(run_pass(ctx, bary_lamp_cols_desc(pb->x, n, pb->y, m, pb->r, logZ, etat, winf, p, img_dev), CULAMP_FP64, KIND_BARY));
    TRY(copy_out(ctx, img2, img_dev, 3 * m));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (psi_out) {
    HIP_TRY(launch_axpby(tmp, nu, -2.0 * winf / a->eta_p, nu, 0.0, m, ctx->stream));
    TRY(copy_out(ctx, psi_out, tmp, m));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (phi_out) {
    HIP_TRY(launch_log_minus(tmp, pb->r, logZ, n, ctx->stream));
    TRY(copy_out(ctx, phi_out, tmp, n));
  }
  TRY(copy_out(ctx, nu_out, nu, m));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (out) {
    out->iterations = steps;
    out->objective = objective;
    out->residual = resid;
    out->primal = primal;
    out->dual = dual;
    out->winf = winf;
    out->eta_t = etat;
    out->elapsed_s = now_s() - t0;
    out->converged = converged;
    out->pairs = ctx->pairs - pairs0;
    out->passes = ctx->passes - passes0;
    out->nonfinite = (a->verbose || a->stop == CULAMP_STOP_INFEAS) &&
                     !(std::isfinite(resid) && std::isfinite(objective));
    out->fp64_from = guard.fp64_from;
  }
  return CULAMP_OK;
}

int culamp_sinkhorn_solve(culamp_ctx* ctx, culamp_problem* pb, const culamp_args* a,
                          int64_t frequency, double p, const double* phi0,
                          const double* psi0, culamp_log_fn log_fn, void* user,
                          double* phi_out, double* psi_out, double* img1,
                          double* img2, culamp_result* out) {
  TRY(check_ctx(ctx));
  if (!pb || !a) return fail(CULAMP_ERR_INVALID_ARG, "null problem or args");
  if (pb->metric >= 0) return fail(CULAMP_ERR_INVALID_ARG, "feature problem: use the cellsim solvers");
  if (frequency <= 0) return fail(CULAMP_ERR_INVALID_ARG, "frequency must be positive");
  const int64_t n = pb->n, m = pb->m;
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  DevBufs B;
  SkState st;
  TRY(B.alloc(&st.phi, n)); TRY(B.alloc(&st.psi, m));
  TRY(B.alloc(&st.res, n > m ? n : m)); TRY(B.alloc(&st.cost, n));
  if (phi0) HIP_TRY(hipMemcpyAsync(st.phi, phi0, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  else HIP_TRY(launch_fill(st.phi, 0.0, n, ctx->stream));
  if (psi0) HIP_TRY(hipMemcpyAsync(st.psi, psi0, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
  else HIP_TRY(launch_fill(st.psi, 0.0, m, ctx->stream));
  culamp_result r{};
  TRY(sinkhorn_core(ctx, pb, a, frequency, p, log_fn, user, st, &r));
  const double p_aux = sk_aux_p(p);
  double* img_dev = nullptr;
  if (img1 || img2) TRY(B.alloc(&img_dev, 3 * (n > m ? n : m)));
  if (img1) {
    TRY(run_pass(ctx, bary_sk_desc(pb->x, n, pb->y, m, st.psi, a->eta_p, r.winf, p_aux, img_dev), CULAMP_FP64, KIND_BARY));
    TRY(copy_out(ctx, img1, img_dev, 3 * n));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (img2) {
    TRY(run_pass(ctx, bary_sk_desc(pb->y, m, pb->x, n, st.phi, a->eta_p, r.winf, p_aux, img_dev), CULAMP_FP64, KIND_BARY));
    TRY(copy_out(ctx, img2, img_dev, 3 * m));
  }
  TRY(copy_out(ctx, phi_out, st.phi, n));
  TRY(copy_out(ctx, psi_out, st.psi, m));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  r.pairs = ctx->pairs - pairs0;
  r.passes = ctx->passes - passes0;
  if (out) *out = r;
  return CULAMP_OK;
}

// annealed_sinkhorn_euclidean (src/solvers/AnnealedSinkhorn.jl:65-177).
int culamp_annealed_sinkhorn_solve(culamp_ctx* ctx, culamp_problem* pb,
                                   const culamp_args* a_in, int64_t frequency,
                                   double p, culamp_log_fn log_fn, void* user,
                                   double* phi_out, double* psi_out,
                                   double* img1, double* img2,
                                   culamp_result* out) {
  TRY(check_ctx(ctx));
  if (!pb || !a_in) return fail(CULAMP_ERR_INVALID_ARG, "null problem or args");
  if (pb->metric >= 0) return fail(CULAMP_ERR_INVALID_ARG, "feature problem: use the cellsim solvers");
  if (frequency <= 0) return fail(CULAMP_ERR_INVALID_ARG, "frequency must be positive");
  const int64_t n = pb->n, m = pb->m;
  culamp_args a = *a_in;
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  const double eta_final = a.epsilon;
  double eta = std::fmax(a.eta_p, eta_final);
  const double t0 = now_s();
  DevBufs B;
  SkState st;
  double *phiws, *psiws, *phiprev, *psiprev;
  TRY(B.alloc(&st.phi, n)); TRY(B.alloc(&st.psi, m));
  TRY(B.alloc(&st.res, n > m ? n : m)); TRY(B.alloc(&st.cost, n));
  TRY(B.alloc(&phiws, n)); TRY(B.alloc(&psiws, m));
  TRY(B.alloc(&phiprev, n)); TRY(B.alloc(&psiprev, m));
  HIP_TRY(hipMemcpyAsync(phiws, pb->logr, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(psiws, pb->logc, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(phiprev, pb->logr, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(psiprev, pb->logc, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
  double objective = NAN, residual = NAN;
  const double q = a.anneal_mult;
  double preveta = 1.0;
  const int verbose = a.verbose;
  a.verbose = 0;
  int64_t total = 0;
  int64_t fp64_from = 0;  // first stage's FP64 takeover, in overall iterations
  const double tmax = a.tmax;
  culamp_result r{};
  if (verbose) emit(log_fn, user, "time(s),iter,infeas,ot_objective,eta,solver\n");
  for (;;) {
    a.eta_p = eta;
    HIP_TRY(hipMemcpyAsync(st.phi, phiws, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(st.psi, psiws, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
    TRY(sinkhorn_core(ctx, pb, &a, frequency, p, nullptr, nullptr, st, &r));
    if (fp64_from == 0 && r.fp64_from > 0) fp64_from = total + r.fp64_from;
    total += r.iterations;
    objective = r.objective;
    residual = r.residual;
    const double elapsed = now_s() - t0;
    a.tmax = tmax - elapsed;
    if (verbose)
      emit(log_fn, user, fmt("%.6e", elapsed) + "," + std::to_string(total) + "," +
                             fmt("%.14e", residual) + "," + fmt("%.14e", objective) + "," +
                             fmt("%.14e", eta) + ",annealed_sinkhorn_kernel\n");
    if (eta == eta_final || elapsed >= tmax) break;
    const double f = eta / preveta;
    // phi_ws = phi + f (phi - phi_prev); phi_prev = phi   (123-126)
    HIP_TRY(launch_axpby(phiws, st.phi, 1.0 + f, phiprev, -f, n, ctx->stream));
    HIP_TRY(launch_axpby(psiws, st.psi, 1.0 + f, psiprev, -f, m, ctx->stream));
    HIP_TRY(hipMemcpyAsync(phiprev, st.phi, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(psiprev, st.psi, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
    preveta = eta;
    eta = std::fmax(q * eta, eta_final);
  }
  double winf = r.winf;
  const double p_aux = sk_aux_p(p);
  double* img_dev = nullptr;
  if (img1 || img2) TRY(B.alloc(&img_dev, 3 * (n > m ? n : m)));
  if (img1) {
    TRY(run_pass(ctx, bary_sk_desc(pb->x, n, pb->y, m, st.psi, eta_final, winf, p_aux, img_dev), CULAMP_FP64, KIND_BARY));
    TRY(copy_out(ctx, img1, img_dev, 3 * n));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (img2) {
    TRY(run_pass(ctx, bary_sk_desc(pb->y, m, pb->x, n, st.phi, eta_final, winf, p_aux, img_dev), CULAMP_FP64, KIND_BARY));
    TRY(copy_out(ctx, img2, img_dev, 3 * m));
  }
  TRY(copy_out(ctx, phi_out, st.phi, n));
  TRY(copy_out(ctx, psi_out, st.psi, m));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (out) {
    out->iterations = total;
    out->objective = objective;
    out->residual = residual;
    out->primal = NAN;
    out->dual = NAN;
    out->winf = winf;
    out->eta_t = eta;
    out->elapsed_s = now_s() - t0;
    out->converged = residual <= eta_final / 6.0;
    out->fp64_from = fp64_from;
    out->pairs = ctx->pairs - pairs0;
    out->passes = ctx->passes - passes0;
    out->nonfinite = !(std::isfinite(residual) && std::isfinite(objective));
  }
  return CULAMP_OK;
}

// ---- cell-similarity solvers ----------------------------------------------

// extragradient_cellsim (src/solvers/MirrorProxCellSim.jl:94-188).  Differs
// from the point-cloud LAMP loop: eta = eta_p (118), its own primal/dual
// formulas (150-159), the averaged gamma for the eta = 0 dual (140, 179), and
// phi = log r - (last row values computed), objective = sum of the last
// column costs (184-187).
int culamp_lamp_cellsim_solve(culamp_ctx* ctx, culamp_problem* pb, const culamp_args* a,
                              int64_t frequency, culamp_log_fn log_fn, void* user,
                              double* nu_out, double* phi_out, double* psi_out,
                              culamp_result* out) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  if (!a) return fail(CULAMP_ERR_INVALID_ARG, "null args");
  if (frequency <= 0) return fail(CULAMP_ERR_INVALID_ARG, "frequency must be positive");
  const int64_t n = pb->n, m = pb->m;
  PrecisionGuard guard(a->precision, ctx->fp64_fallback != 0);
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  DevBufs B;
  double *theta, *nu, *thetab, *nub, *gam, *res, *cost, *etamu, *sumvals;
  TRY(B.alloc(&theta, m)); TRY(B.alloc(&nu, m)); TRY(B.alloc(&thetab, m));
  TRY(B.alloc(&nub, m)); TRY(B.alloc(&gam, m)); TRY(B.alloc(&res, m));
  TRY(B.alloc(&cost, m)); TRY(B.alloc(&etamu, m)); TRY(B.alloc(&sumvals, n > m ? n : m));
  for (double* v : {theta, nu, thetab, nub, gam, res, cost}) HIP_TRY(launch_fill(v, 0.0, m, ctx->stream));
  HIP_TRY(launch_fill(sumvals, 0.0, n, ctx->stream));
  const double t0 = now_s();
  double winf;
  TRY(problem_winf(ctx, pb, 2.0, sumvals, &winf));
  HIP_TRY(launch_fill(sumvals, 0.0, n, ctx->stream));
  const double eta = a->eta_p;
  HIP_TRY(launch_etamu(etamu, pb->c, a->alpha / (double)n, a->tau_mu, m, ctx->stream));
  double negent = 0.0;
  for (double v : pb->hr) if (v > 0.0) negent += v * std::log(v);
  const double hr = eta * negent;
  emit(log_fn, user, "time(s),iter,infeas,ot_objective,primal,dual,solver\n");
  double st = 1.0, etat = INFINITY;
  const double minv = std::tanh(-a->B / 2.0), maxv = std::tanh(a->B / 2.0);
  double primal = NAN, dual = NAN, resid = NAN, objective = NAN;
  int converged = 0;
  int64_t steps = 0;
  auto lse = [&](const double* pot, double reg, double s) {
    PassDesc d = lse_rows_desc(nullptr, n, nullptr, m, pot, reg, s, winf, 2.0, sumvals);
    feat(d, pb, true);
    d.prev_offset = 1;
    return run_pass(ctx, d, guard.prec, KIND_LSE);
  };
  auto colmarg = [&](const double* pot, double reg, double s, PassDesc* proto) {
    PassDesc d = colmarg_desc(nullptr, n, nullptr, m, pb->r, pot, sumvals, reg, s, winf, 2.0, res, cost);
    if (proto) {
      d.epi = EPI_LAMP_COL; d.fuse_update = 1; d.adjust = proto->adjust;
      d.th0 = proto->th0; d.cvec = pb->c; d.etamu = etamu; d.eta_mu = a->eta_mu;
      d.minv = minv; d.maxv = maxv; d.upd_winf = 1.0; d.th_out = proto->th_out;
      d.nu_in = proto->nu_in; d.mix_src = proto->mix_src; d.nu_out = proto->nu_out;
      d.mix_t = proto->mix_t;
    }
    feat(d, pb, false);
    return run_pass(ctx, d, guard.prec, KIND_SUM);
  };
  for (int64_t it = 1; it <= a->itermax; ++it) {
    const double elapsed = now_s() - t0;
    if (elapsed > a->tmax) break;
    const bool check_it = a->verbose || a->stop == CULAMP_STOP_INFEAS;
    if (check_it && (it - 1) % frequency == 0) {  // 144-166
      TRY(lse(nu, etat, st));
      TRY(colmarg(nu, etat, st, nullptr));
      ReduceSpec rs{};
      rs.nops = 4;
      rs.op[0] = 1; rs.a[0] = res; rs.b[0] = pb->c; rs.len[0] = m;
      rs.op[1] = 0; rs.a[1] = cost; rs.len[1] = m;
      rs.op[2] = 2; rs.a[2] = pb->r; rs.b[2] = sumvals; rs.len[2] = n;
      rs.op[3] = 2; rs.a[3] = pb->c; rs.b[3] = nu; rs.len[3] = m;
      double v[4];
      TRY(reduce_to_host(ctx, rs, v));
      resid = v[0];
      objective = v[1];
      ReduceSpec rd{};
      rd.nops = 2;
      if (eta > 0) {
        primal = objective * (1.0 - eta / etat) - eta * v[2] - 2.0 * winf * eta / etat * v[3] +
                 hr + 2.0 * winf * resid;
        TRY(lse(theta, eta, 1.0));
        rd.op[1] = 2; rd.a[1] = pb->c; rd.b[1] = theta; rd.len[1] = m;
      } else {
        primal = objective + 2.0 * winf * resid;
        PassDesc dm = minplus_desc(nullptr, n, nullptr, m, gam, winf, 2.0, sumvals);
        feat(dm, pb, true);
        TRY(run_pass(ctx, dm, CULAMP_FP64, KIND_MIN));
        rd.op[1] = 2; rd.a[1] = pb->c; rd.b[1] = gam; rd.len[1] = m;
      }
      rd.op[0] = 2; rd.a[0] = pb->r; rd.b[0] = sumvals; rd.len[0] = n;
      double w[2];
      TRY(reduce_to_host(ctx, rd, w));
      if (eta > 0) dual = -eta * w[0] - 2.0 * winf * w[1] + hr;
      else dual = w[0] - 2.0 * winf * w[1];
      if (a->verbose)
        emit(log_fn, user, fmt("%.6e", elapsed) + "," + std::to_string(it) + "," +
                               fmt("%.14e", resid) + "," + fmt("%.14e", objective) + "," +
                               fmt("%.14e", primal) + "," + fmt("%.14e", dual) + ",lamp_kernel\n");
      const bool met = a->stop == CULAMP_STOP_INFEAS
                           ? (std::isfinite(etat) && resid < a->epsilon / 6.0)
                           : (primal - dual < a->epsilon / 6.0 && resid < a->epsilon / 6.0);
      if (met) {
        converged = 1;
        break;
      }
      if (std::isfinite(etat)) (void)guard.observe(resid, a->epsilon / 6.0, it);
    }
    // 168-180
    const double etat_new = 1.0 / (1.0 + (1.0 / etat) * (1.0 - eta));
    TRY(lse(nu, etat, st));
    {
      PassDesc pr = blank();
      pr.adjust = 0; pr.th0 = theta; pr.th_out = thetab;
      pr.nu_in = nu; pr.mix_src = theta; pr.nu_out = nub; pr.mix_t = etat_new;
      TRY(colmarg(nu, etat, st, &pr));
    }
    etat = etat_new;
    st = (1.0 - etat) * st + etat;
    TRY(lse(nub, etat, st));
    {
      PassDesc pr = blank();
      pr.adjust = 1; pr.th0 = theta; pr.th_out = theta;
      pr.nu_in = nu; pr.mix_src = thetab; pr.nu_out = nu; pr.mix_t = etat;
      TRY(colmarg(nub, etat, st, &pr));
    }
    HIP_TRY(launch_axpby(gam, gam, 1.0 - etat, theta, etat, m, ctx->stream));
    steps += 1;
  }
  ReduceSpec ro{};
  ro.nops = 1; ro.op[0] = 0; ro.a[0] = cost; ro.len[0] = m;
  double objf;
  TRY(reduce_to_host(ctx, ro, &objf));
  if (psi_out) {
    HIP_TRY(launch_axpby(res, nu, -2.0 * winf / a->eta_p, nu, 0.0, m, ctx->stream));
    TRY(copy_out(ctx, psi_out, res, m));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (phi_out) {
    double* phi_dev;
    TRY(B.alloc(&phi_dev, n));
    HIP_TRY(launch_log_minus(phi_dev, pb->r, sumvals, n, ctx->stream));
    TRY(copy_out(ctx, phi_out, phi_dev, n));
  }
  TRY(copy_out(ctx, nu_out, nu, m));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (out) {
    out->iterations = steps;
    out->objective = objf;
    out->residual = resid;
    out->primal = primal;
    out->dual = dual;
    out->winf = winf;
    out->eta_t = etat;
    out->elapsed_s = now_s() - t0;
    out->converged = converged;
    out->pairs = ctx->pairs - pairs0;
    out->passes = ctx->passes - passes0;
    out->nonfinite = !std::isfinite(objf);
    out->fp64_from = guard.fp64_from;
  }
  return CULAMP_OK;
}

// sinkhorn_cellsim (src/solvers/SinkhornCellSim.jl:240-310); the cost of a
// pair is between row i of x and row j of y (the reference evaluates it
// within one data matrix and is only called with data1 == data2).
int culamp_sinkhorn_cellsim_solve(culamp_ctx* ctx, culamp_problem* pb, const culamp_args* a,
                                  int64_t frequency, const double* phi0, const double* psi0,
                                  culamp_log_fn log_fn, void* user, double* phi_out,
                                  double* psi_out, culamp_result* out) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  if (!a) return fail(CULAMP_ERR_INVALID_ARG, "null args");
  if (frequency <= 0) return fail(CULAMP_ERR_INVALID_ARG, "frequency must be positive");
  const int64_t n = pb->n, m = pb->m;
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  DevBufs B;
  SkState st;
  TRY(B.alloc(&st.phi, n)); TRY(B.alloc(&st.psi, m));
  TRY(B.alloc(&st.res, n > m ? n : m)); TRY(B.alloc(&st.cost, n));
  if (phi0) HIP_TRY(hipMemcpyAsync(st.phi, phi0, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  else HIP_TRY(launch_fill(st.phi, 0.0, n, ctx->stream));
  if (psi0) HIP_TRY(hipMemcpyAsync(st.psi, psi0, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
  else HIP_TRY(launch_fill(st.psi, 0.0, m, ctx->stream));
  culamp_result r{};
  TRY(sinkhorn_core(ctx, pb, a, frequency, 2.0, log_fn, user, st, &r));
  TRY(copy_out(ctx, phi_out, st.phi, n));
  TRY(copy_out(ctx, psi_out, st.psi, m));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  r.pairs = ctx->pairs - pairs0;
  r.passes = ctx->passes - passes0;
  if (out) *out = r;
  return CULAMP_OK;
}

// annealed_sinkhorn_cellsim (src/solvers/AnnealedSinkhornCellSim.jl:1-63).
// No tmax test between stages.  The stage solve updates the warm-start arrays
// in place (SinkhornCellSim.jl:257-259), so after the extrapolation the
// "previous" potentials are the extrapolated warm start (49-52).
int culamp_annealed_sinkhorn_cellsim_solve(culamp_ctx* ctx, culamp_problem* pb,
                                           const culamp_args* a_in, int64_t frequency,
                                           culamp_log_fn log_fn, void* user, double* phi_out,
                                           double* psi_out, culamp_result* out) {
  TRY(check_ctx(ctx));
  TRY(need_features(pb));
  if (!a_in) return fail(CULAMP_ERR_INVALID_ARG, "null args");
  if (frequency <= 0) return fail(CULAMP_ERR_INVALID_ARG, "frequency must be positive");
  const int64_t n = pb->n, m = pb->m;
  const double pairs0 = ctx->pairs;
  const int64_t passes0 = ctx->passes;
  culamp_args a = *a_in;
  const double eta_final = a.epsilon;
  double eta = std::fmax(a.eta_p, eta_final);
  const double t0 = now_s();
  DevBufs B;
  SkState st;
  double *phiprev, *psiprev;
  TRY(B.alloc(&st.phi, n)); TRY(B.alloc(&st.psi, m));
  TRY(B.alloc(&st.res, n > m ? n : m)); TRY(B.alloc(&st.cost, n));
  TRY(B.alloc(&phiprev, n)); TRY(B.alloc(&psiprev, m));
  HIP_TRY(hipMemcpyAsync(st.phi, pb->logr, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(st.psi, pb->logc, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(phiprev, pb->logr, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(psiprev, pb->logc, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
  const double q = a.anneal_mult;
  double preveta = 1.0;
  const int verbose = a.verbose;
  a.verbose = 0;
  int64_t total = 0;
  int64_t fp64_from = 0;  // first stage's FP64 takeover, in overall iterations
  culamp_result r{};
  if (verbose) emit(log_fn, user, "time(s),iter,infeas,ot_objective,eta,solver\n");
  for (;;) {
    a.eta_p = eta;
    TRY(sinkhorn_core(ctx, pb, &a, frequency, 2.0, nullptr, nullptr, st, &r));
    if (fp64_from == 0 && r.fp64_from > 0) fp64_from = total + r.fp64_from;
    total += r.iterations;
    if (verbose)
      emit(log_fn, user, fmt("%.6e", now_s() - t0) + "," + std::to_string(total) + "," +
                             fmt("%.14e", r.residual) + "," + fmt("%.14e", r.objective) + "," +
                             fmt("%.14e", eta) + ",annealed_sinkhorn_cellsim\n");
    if (eta == eta_final) break;
    const double f = eta / preveta;
    HIP_TRY(launch_axpby(st.phi, st.phi, 1.0 + f, phiprev, -f, n, ctx->stream));
    HIP_TRY(launch_axpby(st.psi, st.psi, 1.0 + f, psiprev, -f, m, ctx->stream));
    HIP_TRY(hipMemcpyAsync(phiprev, st.phi, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(psiprev, st.psi, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
    preveta = eta;
    eta = std::fmax(q * eta, eta_final);
  }
  TRY(copy_out(ctx, phi_out, st.phi, n));
  TRY(copy_out(ctx, psi_out, st.psi, m));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (out) {
    *out = r;
    out->iterations = total;
    out->fp64_from = fp64_from;
    out->eta_t = eta;
    out->elapsed_s = now_s() - t0;
    out->pairs = ctx->pairs - pairs0;
    out->passes = ctx->passes - passes0;
  }
  return CULAMP_OK;
}

}  // extern "C"

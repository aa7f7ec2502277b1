// culamp_batch.hip -- batched solves: many independent OT instances of
// different sizes advanced together, one launch per pass for all of them.
//
// Per instance the loops are those of the single-instance solvers
// (Sinkhorn: src/solvers/SinkhornKernel.jl:477-559 and
// src/solvers/SinkhornCellSim.jl:240-310; LAMP: src/solvers/
// MirrorProxKernel.jl:1019-1132 and src/solvers/MirrorProxCellSim.jl:94-188):
// same passes, same scalar schedules, same checks every `frequency`
// iterations, same stop tests and outputs.  What changes is how the work is
// issued:
//
//   * a pass of every active instance is one kernel launch over an item
//     table (instance, source split, target block), instances ordered by the
//     work of their items, largest first, so the hardware dispatcher drains
//     the table as a longest-first work queue;
//   * LAMP's iteration scalars (eta_t, s_t and the mixing weights) live on the
//     device, one slot per pass and instance, advanced by a small kernel, so
//     the passes of an iteration take no per-iteration host arguments;
//   * the iterations between two checks are therefore identical launches and
//     are captured once per active set in a HIP graph and replayed;
//   * at a check, one segmented reduction and one device-to-host copy give
//     every instance's infeasibility, cost and dual terms; converged instances
//     leave the tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <map>
#include <tuple>
#include <vector>

#include "culamp_host.h"

using namespace culamp;
using namespace culamp_host;

struct culamp_batch {
  culamp_ctx* ctx = nullptr;
  std::vector<culamp_problem*> pbs;
  int32_t count = 0;
  bool feature = false;
  int metric = -1;
  int64_t d = 0;
  double scale = 1.0;
  int zks = 0;          // MFMA row width / 16 when every instance is eligible, else 0
  std::vector<int64_t> n, m, off_n, off_m;
  int64_t sum_n = 0, sum_m = 0;
  double *r = nullptr, *c = nullptr, *logr = nullptr;  // pooled marginals
  culamp_batch_tuning tune{};
  // W_inf of every instance per cost exponent p (feature batches: one entry),
  // computed by the first solve that needs it: it depends on the data only
  std::vector<std::pair<double, std::vector<double>>> winf_cache;
};

namespace {

constexpr int kDyn = 20;  // doubles per instance: slots S, A, B, C, D of 4

// Segment of one instance in a pooled vector, plus one per-instance scalar.
struct Seg {
  int64_t off, len;
  double s0;
};

__global__ __launch_bounds__(kReduceThreads) void batch_reduce_kernel(const ReduceSpec* specs,
                                                            const int32_t* act, double* out) {
  // workgroup (i, q): reduction q of active instance i
  const int32_t k = act[blockIdx.x];
  if ((int)blockIdx.y >= specs[k].nops) return;
  reduce_op(specs[k], blockIdx.y, out + 8 * (int64_t)k);
}

// LAMP scalar schedule of one iteration (MirrorProxKernel.jl:1073, 1108-1115;
// MirrorProxCellSim.jl:168-180): slot S holds the current (eta_t, s_t); the
// four passes of the iteration read slots A..D.
__global__ void advance_kernel(double* dyn, const double* eta, const int32_t* act, int nact) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nact) return;
  const int32_t k = act[i];
  double* D = dyn + (int64_t)kDyn * k;
  const double et = D[0], st = D[1], e = eta[k];
  const double et_new = 1.0 / (1.0 + (1.0 / et) * (1.0 - e));
  const double st_new = (1.0 - et_new) * st + et_new;
  D[4] = et; D[5] = st; D[6] = 0.0;                // A: LSE of nu
  D[8] = et; D[9] = st; D[10] = et_new;            // B: column pass, nu-bar mixing
  D[12] = et_new; D[13] = st_new; D[14] = 0.0;     // C: LSE of nu-bar
  D[16] = et_new; D[17] = st_new; D[18] = et_new;  // D: column pass, nu mixing
  D[0] = et_new; D[1] = st_new;
}

// gamma = (1 - eta_t) gamma + eta_t theta (MirrorProxCellSim.jl:179) with the
// instance's new eta_t.
__global__ void seg_gamma_kernel(double* gam, const double* theta, const Seg* seg,
                                 const int32_t* act, const double* dyn) {
  const int32_t k = act[blockIdx.y];
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= seg[k].len) return;
  const int64_t i = seg[k].off + j;
  const double et = dyn[(int64_t)kDyn * k + 12];
  gam[i] = axpby_elem(1.0 - et, gam[i], et, theta[i]);
}

// eta_mu_i = (c + alpha / n) / tau_mu (MirrorProxKernel.jl:1037)
__global__ void seg_etamu_kernel(double* out, const double* c, const Seg* seg,
                                 const int32_t* act, double div) {
  const int32_t k = act[blockIdx.y];
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= seg[k].len) return;
  const int64_t i = seg[k].off + j;
  out[i] = (c[i] + seg[k].s0) / div;
}

__global__ void pooled_fill_kernel(double* out, double v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}
__global__ void pooled_log_minus_kernel(double* out, const double* r, const double* a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = log(r[i]) - a[i];
}

unsigned blocks_of(int64_t n) { return (unsigned)((n + 255) / 256); }

// Device buffer owned by an engine.
template <typename T>
struct DevArr {
  T* p = nullptr;
  size_t cap = 0;
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
  int reserve(size_t n) {
    if (n <= cap && p) return CULAMP_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = n > 0 ? n : 1;
    hipError_t e = hipMalloc(&p, sizeof(T) * want);
    if (e != hipSuccess) return fail(CULAMP_ERR_ALLOC, "hipMalloc(%zu bytes): %s", sizeof(T) * want, hipGetErrorString(e));
    cap = want;
    return CULAMP_OK;
  }
  int upload(const std::vector<T>& v, hipStream_t s) {
    TRY(reserve(v.size()));
    if (!v.empty()) HIP_TRY(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return CULAMP_OK;
  }
};

struct GraphExec {
  hipGraphExec_t exec = nullptr;
  ~GraphExec() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
};

// Item tables of one kernel family, direction and targets-per-thread.
struct Plan {
  int family = FAMILY_POINTS;
  bool rows = true;
  int rpt = 1;
  int64_t split = 0;  // sources per item chosen for the current active set
  std::vector<BatchInst> insts;
  std::vector<BatchItem> items, mitems;
  DevArr<BatchInst> d_insts;
  DevArr<BatchItem> d_items, d_mitems;
  size_t partial = 0;  // doubles of split partials
};

// XCDs of the device: gfx950 has 32 CUs per XCD (256 CUs in 8 XCDs when
// the whole device is one partition).
int xcd_count(int num_cu) { return num_cu >= 64 ? num_cu / 32 : 1; }

int targets_per_block(int family, int rpt) {
  return family == FAMILY_POINTS ? (kThreads / kLanesPerTarget) * rpt
                                 : family == FAMILY_FEATURES ? kFeatTargets : 128;
}
int64_t split_granule(int family) { return family == FAMILY_POINTS ? 8 : family == FAMILY_MFMA ? 32 : 1; }
int64_t min_split(int family) { return family == FAMILY_POINTS ? kTile : 64; }

struct Engine {
  culamp_ctx* ctx;
  culamp_batch* b;
  std::vector<int32_t> active;
  DevArr<int32_t> d_active;
  DevArr<Seg> d_seg_n, d_seg_m;
  std::vector<std::unique_ptr<Plan>> plans;
  int64_t want_items = 0;

  Engine(culamp_ctx* c, culamp_batch* bb) : ctx(c), b(bb) {
    want_items = b->tune.min_items > 0 ? b->tune.min_items : (int64_t)4 * ctx->num_cu;
  }

  int64_t nt_of(int32_t k, bool rows) const { return rows ? b->n[k] : b->m[k]; }
  int64_t ns_of(int32_t k, bool rows) const { return rows ? b->m[k] : b->n[k]; }

  int init_segments(double etamu_add_num = 0.0) {
    std::vector<Seg> sn(b->count), sm(b->count);
    for (int32_t k = 0; k < b->count; ++k) {
      sn[k] = Seg{b->off_n[k], b->n[k], 0.0};
      sm[k] = Seg{b->off_m[k], b->m[k], etamu_add_num / (double)b->n[k]};
    }
    TRY(d_seg_n.upload(sn, ctx->stream));
    TRY(d_seg_m.upload(sm, ctx->stream));
    return CULAMP_OK;
  }

  int64_t max_m_active() const {
    int64_t v = 0;
    for (int32_t k : active) v = std::max(v, b->m[k]);
    return v;
  }

  // Launch shape of a plan for the current active set.
  void shape(Plan& P) {
    const int tpb = targets_per_block(P.family, P.rpt);
    const int64_t g = split_granule(P.family);
    // Split length: the tuned one (or the longest source range), halved while
    // the active instances give fewer items than want_items, so that a
    // shrinking active set (instances converge one by one) still fills the
    // device.
    int64_t L = 0;
    for (int32_t k : active) L = std::max(L, ns_of(k, P.rows));
    if (b->tune.max_split > 0) L = std::min(L, b->tune.max_split);
    auto count = [&](int64_t len) {
      int64_t tot = 0;
      for (int32_t k : active) {
        const int64_t nt = nt_of(k, P.rows), ns = ns_of(k, P.rows);
        tot += ((nt + tpb - 1) / tpb) * ((ns + len - 1) / len);
      }
      return tot;
    };
    while (count(L) < want_items && L / 2 >= min_split(P.family)) L /= 2;
    if (L < g) L = g;
    P.split = L;
    P.insts.assign(b->count, BatchInst{0, 0, 0, 1, 0});
    std::vector<int32_t> order = active;
    std::vector<int64_t> item_work(b->count, 0);
    int64_t pbase = 0;
    for (int32_t k : active) {
      const int64_t nt = nt_of(k, P.rows), ns = ns_of(k, P.rows);
      const int64_t tblocks = std::max<int64_t>(1, (nt + tpb - 1) / tpb);
      int64_t splits = std::max<int64_t>(1, (ns + L - 1) / L);
      int64_t len = (ns + splits - 1) / splits;
      len = std::max<int64_t>(g, (len + g - 1) / g * g);
      splits = std::max<int64_t>(1, (ns + len - 1) / len);
      BatchInst bi{};
      bi.ntp = tblocks * tpb;
      bi.split_len = len;
      bi.splits = (int32_t)splits;
      bi.pbase = pbase;
      if (splits > 1) pbase += (int64_t)pass_fields(KIND_SUM) * splits * bi.ntp;  // (the most fields of LSE / SUM)
      P.insts[k] = bi;
      item_work[k] = len;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t c) { return item_work[a] > item_work[c]; });
    // XCD-aware table: workgroup i of a launch runs on XCD i mod X and each
    // XCD has its own L2, so all items of an instance go to one XCD, where
    // they run side by side and read the instance's sources from that L2
    // (spread over every XCD, each instance's sources were fetched from
    // memory once per XCD, and the many instances in flight per XCD evicted
    // each other).  Instances are assigned longest-processing-time first by
    // total work; each XCD's sequence keeps the largest-item-first order and
    // shorter sequences are padded with items that do nothing.
    const int X = xcd_count(ctx->num_cu);
    std::vector<int32_t> by_work = order;
    auto work_of = [&](int32_t k) {
      const BatchInst& bi = P.insts[k];
      return (double)(bi.ntp / tpb) * (double)bi.splits * (double)bi.split_len;
    };
    std::stable_sort(by_work.begin(), by_work.end(),
                     [&](int32_t a, int32_t c) { return work_of(a) > work_of(c); });
    std::vector<int> home(b->count, 0);
    std::vector<double> load(X, 0.0);
    for (int32_t k : by_work) {
      int x = 0;
      for (int q = 1; q < X; ++q)
        if (load[q] < load[x]) x = q;
      home[k] = x;
      load[x] += work_of(k);
    }
    std::vector<std::vector<BatchItem>> seq(X);
    P.items.clear();
    P.mitems.clear();
    for (int32_t k : order) {
      const BatchInst& bi = P.insts[k];
      const int64_t tblocks = bi.ntp / tpb;
      for (int32_t sp = 0; sp < bi.splits; ++sp)
        for (int64_t tb = 0; tb < tblocks; ++tb) seq[home[k]].push_back(BatchItem{k, sp, tb});
      if (bi.splits > 1) {
        const int64_t per = kThreads / kMergeLanes;
        const int64_t mb = (nt_of(k, P.rows) + per - 1) / per;
        for (int64_t t = 0; t < mb; ++t) P.mitems.push_back(BatchItem{k, 0, t});
      }
    }
    size_t longest = 0;
    for (const auto& q : seq) longest = std::max(longest, q.size());
    P.items.assign(longest * (size_t)X, BatchItem{-1, 0, 0});
    for (int x = 0; x < X; ++x)
      for (size_t j = 0; j < seq[x].size(); ++j) P.items[j * (size_t)X + (size_t)x] = seq[x][j];
    P.partial = (size_t)pbase;
  }

  int upload_plan(Plan& P) {
    TRY(P.d_insts.upload(P.insts, ctx->stream));
    TRY(P.d_items.upload(P.items, ctx->stream));
    TRY(P.d_mitems.upload(P.mitems, ctx->stream));
    if (P.partial > 0) TRY(ensure_work(ctx, P.partial));
    return CULAMP_OK;
  }

  // Sets the active instances (stream idle) and rebuilds every plan.
  int set_active(const std::vector<int32_t>& act) {
    active = act;
    TRY(d_active.upload(active, ctx->stream));
    for (auto& P : plans) {
      shape(*P);
      TRY(upload_plan(*P));
    }
    return CULAMP_OK;
  }

  int need(int family, bool rows, int rpt) {
    for (auto& P : plans)
      if (P->family == family && P->rows == rows && P->rpt == rpt) return CULAMP_OK;
    plans.emplace_back(new Plan());
    Plan& P = *plans.back();
    P.family = family;
    P.rows = rows;
    P.rpt = rpt;
    shape(P);
    return upload_plan(P);
  }

  Plan* find(int family, bool rows, int rpt) {
    for (auto& P : plans)
      if (P->family == family && P->rows == rows && P->rpt == rpt) return P.get();
    return nullptr;
  }

  // Kernel family and precision of a pass, as run_pass picks them for one
  // instance.
  struct Route {
    int family, rpt, fp32, mfma_split;
  };
  Route route(int precision, int kind) const {
    Route r{};
    const bool expkind = kind == KIND_LSE || kind == KIND_SUM;
    if (b->feature && (precision == CULAMP_BF16 || precision == CULAMP_BF16X3) && b->zks > 0 && expkind) {
      r.family = FAMILY_MFMA;
      r.rpt = 1;
      r.mfma_split = precision == CULAMP_BF16X3 ? 3 : 1;
      return r;
    }
    const int prec = precision > CULAMP_MIXED ? CULAMP_MIXED : precision;
    r.fp32 = prec == CULAMP_MIXED && expkind;
    if (b->feature) {
      r.family = FAMILY_FEATURES;
      r.rpt = 1;
      return r;
    }
    r.family = FAMILY_POINTS;
    if (!expkind) {
      r.rpt = 1;
    } else {
      const int max_rpt = r.fp32 ? 4 : 2;
      int want = b->tune.rows_per_thread;
      if (want == 4 && max_rpt < 4) want = 2;
      if (want == 0) {
        // the largest rpt whose unsplit pass still has enough workgroups
        want = 1;
        for (int cand = max_rpt; cand >= 1; cand /= 2) {
          int64_t tot = 0;
          for (int32_t k = 0; k < b->count; ++k)
            tot += (b->n[k] + 64 * cand - 1) / (64 * cand);
          if (tot >= want_items) {
            want = cand;
            break;
          }
        }
      }
      r.rpt = want;
    }
    return r;
  }

  int prepare(int precision, int kind, bool rows) {
    const Route r = route(precision, kind);
    return need(r.family, rows, r.rpt);
  }

  // One batched pass over the active instances.
  int launch(const PassDesc* d_descs, int precision, int kind, bool rows, double p) {
    if (active.empty()) return CULAMP_OK;
    const Route r = route(precision, kind);
    Plan* P = find(r.family, rows, r.rpt);
    if (!P) return fail(CULAMP_ERR_INVALID_ARG, "internal: batch pass not prepared");
    BatchLaunch bl{};
    bl.descs = d_descs;
    bl.insts = P->d_insts.p;
    bl.items = P->d_items.p;
    bl.mitems = P->d_mitems.p;
    bl.nitems = (int64_t)P->items.size();
    bl.nmitems = (int64_t)P->mitems.size();
    bl.family = r.family;
    bl.rpt = r.rpt;
    bl.cost = cost_of_p(p);
    bl.metric = b->metric;
    bl.zks = b->zks;
    bl.nsplit = r.mfma_split;
    Work w{ctx->partial};
    if (r.family == FAMILY_MFMA) HIP_TRY(launch_batch_mfma(bl, kind, w, ctx->stream));
    else if (r.fp32) HIP_TRY(launch_batch_fp32(bl, kind, w, ctx->stream));
    else HIP_TRY(launch_batch_fp64(bl, kind, w, ctx->stream));
    return CULAMP_OK;
  }

  // nops: the largest reduction count of the specs (one workgroup each)
  int reduce(const ReduceSpec* d_specs, double* d_out, int nops) {
    if (active.empty() || nops <= 0) return CULAMP_OK;
    hipLaunchKernelGGL(batch_reduce_kernel, dim3((unsigned)active.size(), (unsigned)nops),
                       dim3(kReduceThreads), 0, ctx->stream, d_specs, d_active.p, d_out);
    HIP_TRY(hipGetLastError());
    return CULAMP_OK;
  }

  // Runs `body` captured into `g` (built on first use), or eagerly.
  template <typename F>
  int replay(GraphExec& g, bool use_graph, F&& body) {
    if (!use_graph) return body();
    if (!g.exec) {
      HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      const int rc = body();
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
      if (rc != CULAMP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      HIP_TRY(e);
      const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP_TRY(ei);
    }
    HIP_TRY(hipGraphLaunch(g.exec, ctx->stream));
    return CULAMP_OK;
  }
};

// Pooled device vectors of a solve.
struct Pool {
  DevBufs bufs;
  int vec(double** out, int64_t n) { return bufs.alloc(out, n); }
};

int upload_descs(DevArr<PassDesc>& dst, const std::vector<PassDesc>& v, hipStream_t s) {
  return dst.upload(v, s);
}

// W_inf of every instance: max pass (FP64), segmented max, one copy back.
// Points as max_logsumexp_spp_ct! (MirrorProxKernel.jl:280-326); features as
// max_logsumexp_cellsim! (MirrorProxCellSim.jl:1-41), whose running max
// starts at 0.
int batch_winf(Engine& E, double p, double* tmp_n, std::vector<double>& winf) {
  culamp_batch* b = E.b;
  culamp_ctx* ctx = E.ctx;
  for (const auto& e : b->winf_cache)
    if (e.first == p || (std::isnan(e.first) && std::isnan(p))) {
      winf = e.second;
      return CULAMP_OK;
    }
  std::vector<PassDesc> dm(b->count);
  std::vector<ReduceSpec> rs(b->count);
  for (int32_t k = 0; k < b->count; ++k) {
    culamp_problem* pb = b->pbs[k];
    PassDesc d = max_desc(pb->x, pb->n, pb->y, pb->m, p, tmp_n + b->off_n[k]);
    feat(d, pb, true);
    dm[k] = d;
    ReduceSpec r{};
    r.nops = 1; r.op[0] = 3; r.a[0] = tmp_n + b->off_n[k]; r.len[0] = pb->n;
    rs[k] = r;
  }
  DevArr<PassDesc> d_dm;
  DevArr<ReduceSpec> d_rs;
  DevArr<double> d_out;
  TRY(upload_descs(d_dm, dm, ctx->stream));
  TRY(d_rs.upload(rs, ctx->stream));
  TRY(d_out.reserve(8 * (size_t)b->count));
  TRY(E.prepare(CULAMP_FP64, KIND_MAX, true));
  TRY(E.launch(d_dm.p, CULAMP_FP64, KIND_MAX, true, p));
  TRY(E.reduce(d_rs.p, d_out.p, 1));
  std::vector<double> h(8 * (size_t)b->count);
  HIP_TRY(hipMemcpyAsync(h.data(), d_out.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  winf.resize(b->count);
  for (int32_t k = 0; k < b->count; ++k) winf[k] = b->feature ? std::fmax(h[8 * k], 0.0) : h[8 * k];
  b->winf_cache.emplace_back(p, winf);
  return CULAMP_OK;
}

std::vector<int32_t> all_instances(const culamp_batch* b) {
  std::vector<int32_t> v(b->count);
  for (int32_t k = 0; k < b->count; ++k) v[k] = k;
  return v;
}

int check_batch(culamp_ctx* ctx, culamp_batch* b, const culamp_args* a, int64_t frequency) {
  TRY(check_ctx(ctx));
  if (!b || b->ctx != ctx) return fail(CULAMP_ERR_INVALID_ARG, "null batch or batch of another context");
  if (!a) return fail(CULAMP_ERR_INVALID_ARG, "null args");
  if (frequency <= 0) return fail(CULAMP_ERR_INVALID_ARG, "frequency must be positive");
  if (a->precision < CULAMP_FP64 || a->precision > CULAMP_BF16X3)
    return fail(CULAMP_ERR_INVALID_ARG, "precision must be CULAMP_FP64, _MIXED, _BF16 or _BF16X3");
  return CULAMP_OK;
}

// ---- Sinkhorn ----------------------------------------------------------------
int batch_sinkhorn(culamp_ctx* ctx, culamp_batch* b, const culamp_args* a, int64_t frequency,
                   double p, const double* phi0, const double* psi0, double* phi_out,
                   double* psi_out, culamp_result* results) {
  const double t0 = now_s();
  const int32_t K = b->count;
  const int prec = a->precision;
  const double eta = a->eta_p;
  Engine E(ctx, b);
  TRY(E.set_active(all_instances(b)));
  Pool pool;
  double *phi, *psi, *res, *cost, *tmp, *d_out;
  TRY(pool.vec(&phi, b->sum_n)); TRY(pool.vec(&psi, b->sum_m));
  TRY(pool.vec(&res, b->sum_n)); TRY(pool.vec(&cost, b->sum_n));
  TRY(pool.vec(&tmp, b->sum_n)); TRY(pool.vec(&d_out, 8 * (int64_t)K));
  if (phi0) HIP_TRY(hipMemcpyAsync(phi, phi0, sizeof(double) * b->sum_n, hipMemcpyHostToDevice, ctx->stream));
  else hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(b->sum_n)), dim3(256), 0, ctx->stream, phi, 0.0, b->sum_n);
  if (psi0) HIP_TRY(hipMemcpyAsync(psi, psi0, sizeof(double) * b->sum_m, hipMemcpyHostToDevice, ctx->stream));
  else hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(b->sum_m)), dim3(256), 0, ctx->stream, psi, 0.0, b->sum_m);
  HIP_TRY(hipGetLastError());
  std::vector<double> winf;
  const double pw = b->feature ? 2.0 : p;
  TRY(batch_winf(E, pw, tmp, winf));
  const double p_lse = b->feature ? 2.0 : (p == INFINITY ? sk_aux_p(p) : p);
  const double p_aux = b->feature ? 2.0 : sk_aux_p(p);
  std::vector<PassDesc> h1(K), h2(K), h3(K);
  std::vector<ReduceSpec> hr(K);
  for (int32_t k = 0; k < K; ++k) {
    culamp_problem* pb = b->pbs[k];
    const int64_t n = pb->n, m = pb->m, on = b->off_n[k], om = b->off_m[k];
    double* rk = b->r + on;
    double* ck = b->c + om;
    PassDesc d1 = sk_half_desc(pb->x, n, pb->y, m, rk, psi + om, eta, winf[k], p_lse, phi + on);
    feat(d1, pb, true);
    d1.prev_offset = 1;
    PassDesc d2 = sk_half_desc(pb->y, m, pb->x, n, ck, phi + on, eta, winf[k], p_lse, psi + om);
    feat(d2, pb, false);
    d2.prev_offset = 1;
    PassDesc d3 = sk_resid_desc(pb->x, n, pb->y, m, rk, phi + on, psi + om, eta, winf[k], p_aux,
                                res + on, cost + on);
    feat(d3, pb, true);
    h1[k] = d1; h2[k] = d2; h3[k] = d3;
    ReduceSpec rs{};
    rs.nops = 4;
    rs.op[0] = 5; rs.a[0] = res + on; rs.len[0] = n;
    rs.op[1] = 0; rs.a[1] = cost + on; rs.len[1] = n;
    rs.op[2] = 2; rs.a[2] = psi + om; rs.b[2] = ck; rs.len[2] = m;
    rs.op[3] = 2; rs.a[3] = phi + on; rs.b[3] = rk; rs.len[3] = n;
    hr[k] = rs;
  }
  DevArr<PassDesc> D1, D2, D3;
  DevArr<ReduceSpec> DR;
  TRY(upload_descs(D1, h1, ctx->stream)); TRY(upload_descs(D2, h2, ctx->stream));
  TRY(upload_descs(D3, h3, ctx->stream)); TRY(DR.upload(hr, ctx->stream));
  TRY(E.prepare(prec, KIND_LSE, true));
  TRY(E.prepare(prec, KIND_LSE, false));
  TRY(E.prepare(prec, KIND_SUM, true));

  std::vector<int64_t> iters(K, 0), npass(K, 0);
  std::vector<int32_t> conv(K, 0);
  std::vector<double> fin(8 * (size_t)K, NAN), h(8 * (size_t)K);
  const bool use_graph = b->tune.graph != 0;
  GraphExec g;
  auto step = [&]() -> int {
    TRY(E.launch(D1.p, prec, KIND_LSE, true, p_lse));
    return E.launch(D2.p, prec, KIND_LSE, false, p_lse);
  };
  auto resid = [&]() -> int {
    TRY(E.launch(D3.p, prec, KIND_SUM, true, p_aux));
    TRY(E.reduce(DR.p, d_out, 4));
    HIP_TRY(hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int32_t k : E.active) npass[k] += 1;
    return CULAMP_OK;
  };
  int64_t it = 0;  // completed iterations of the active instances
  while (it < a->itermax && !E.active.empty()) {
    if (now_s() - t0 > a->tmax) break;
    // iterations up to and including the next one that checks ((it-1) % f == 0)
    const int64_t nsteps = std::min<int64_t>(it == 0 ? 1 : frequency, a->itermax - it);
    if (nsteps == frequency) {
      TRY(E.replay(g, use_graph, [&]() -> int {
        for (int64_t s = 0; s < nsteps; ++s) TRY(step());
        return CULAMP_OK;
      }));
    } else {
      for (int64_t s = 0; s < nsteps; ++s) TRY(step());
    }
    for (int32_t k : E.active) npass[k] += 2 * nsteps;
    it += nsteps;
    if ((it - 1) % frequency != 0) continue;
    TRY(resid());
    std::vector<int32_t> still;
    for (int32_t k : E.active) {
      if (h[8 * k] <= a->epsilon / 6.0) {  // SinkhornKernel.jl:535
        conv[k] = 1;
        iters[k] = it - 1;
        for (int q = 0; q < 4; ++q) fin[8 * k + q] = h[8 * k + q];
      } else {
        still.push_back(k);
      }
    }
    if (still.size() != E.active.size()) {
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      TRY(E.set_active(still));
      g.reset();
    }
  }
  // instances that did not converge: iterations run, then the final residual
  for (int32_t k : E.active) iters[k] = it;
  if (!E.active.empty()) {
    TRY(resid());
    for (int32_t k : E.active)
      for (int q = 0; q < 4; ++q) fin[8 * k + q] = h[8 * k + q];
  }
  if (phi_out)
    HIP_TRY(hipMemcpyAsync(phi_out, phi, sizeof(double) * b->sum_n, hipMemcpyDeviceToHost, ctx->stream));
  if (psi_out)
    HIP_TRY(hipMemcpyAsync(psi_out, psi, sizeof(double) * b->sum_m, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double elapsed = now_s() - t0;
  for (int32_t k = 0; k < K; ++k) {
    if (!results) break;
    culamp_result& o = results[k];
    std::memset(&o, 0, sizeof o);
    o.iterations = iters[k];
    o.residual = fin[8 * k];
    o.objective = fin[8 * k + 1];
    o.primal = NAN;
    o.dual = fin[8 * k + 2] + fin[8 * k + 3];
    o.winf = winf[k];
    o.eta_t = eta;
    o.elapsed_s = elapsed;
    o.converged = conv[k];
    o.nonfinite = !(std::isfinite(o.residual) && std::isfinite(o.objective));
    o.passes = npass[k];
    o.pairs = (double)npass[k] * (double)b->n[k] * (double)b->m[k];
  }
  return CULAMP_OK;
}

// ---- LAMP ----------------------------------------------------------------------
int batch_lamp(culamp_ctx* ctx, culamp_batch* b, const culamp_args* a, int64_t frequency,
               double p, double* nu_out, double* phi_out, double* psi_out,
               culamp_result* results) {
  const double t0 = now_s();
  const int32_t K = b->count;
  const int prec = a->precision;
  const bool cellsim = b->feature;
  const double pp = cellsim ? 2.0 : p;
  Engine E(ctx, b);
  TRY(E.set_active(all_instances(b)));
  TRY(E.init_segments(a->alpha));
  Pool pool;
  const int64_t M = b->sum_m, N = b->sum_n;
  double *theta, *nu, *thetab, *nub, *gam, *res, *cost, *etamu, *logZ, *tmp, *d_out, *dyn, *d_eta;
  TRY(pool.vec(&theta, M)); TRY(pool.vec(&nu, M)); TRY(pool.vec(&thetab, M));
  TRY(pool.vec(&nub, M)); TRY(pool.vec(&gam, M)); TRY(pool.vec(&res, M));
  TRY(pool.vec(&cost, M)); TRY(pool.vec(&etamu, M));
  TRY(pool.vec(&logZ, N)); TRY(pool.vec(&tmp, N));
  TRY(pool.vec(&d_out, 16 * (int64_t)K)); TRY(pool.vec(&dyn, kDyn * (int64_t)K));
  TRY(pool.vec(&d_eta, K));
  for (double* v : {theta, nu, thetab, nub, gam, res, cost})
    hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(M)), dim3(256), 0, ctx->stream, v, 0.0, M);
  hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(N)), dim3(256), 0, ctx->stream, logZ, 0.0, N);
  HIP_TRY(hipGetLastError());
  std::vector<double> winf;
  TRY(batch_winf(E, pp, tmp, winf));
  if (cellsim)
    hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(N)), dim3(256), 0, ctx->stream, logZ, 0.0, N);
  // per-instance scalars (MirrorProxKernel.jl:1035, 1066; MirrorProxCellSim.jl:118)
  std::vector<double> eta(K), hr(K), etat(K, INFINITY), st(K, 1.0);
  for (int32_t k = 0; k < K; ++k) {
    eta[k] = cellsim ? a->eta_p : a->eta_p / (2.0 * winf[k]);
    double negent = 0.0;
    for (double v : b->pbs[k]->hr) if (v > 0.0) negent += v * std::log(v);
    hr[k] = eta[k] * negent;
  }
  {
    std::vector<double> hd((size_t)kDyn * K, 0.0);
    for (int32_t k = 0; k < K; ++k) { hd[(size_t)kDyn * k] = INFINITY; hd[(size_t)kDyn * k + 1] = 1.0; }
    HIP_TRY(hipMemcpyAsync(dyn, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_eta, eta.data(), sizeof(double) * K, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  const int64_t maxm = E.max_m_active();
  hipLaunchKernelGGL(seg_etamu_kernel, dim3(blocks_of(maxm), (unsigned)E.active.size()), dim3(256), 0,
                     ctx->stream, etamu, b->c, E.d_seg_m.p, E.d_active.p, a->tau_mu);
  HIP_TRY(hipGetLastError());
  const double minv = std::tanh(-a->B / 2.0), maxv = std::tanh(a->B / 2.0);

  // descriptor tables: check passes E1 (LSE of nu), E2 (column pass), E3
  // (dual); iteration passes A..D
  std::vector<PassDesc> hE1(K), hE2(K), hE3(K), hA(K), hB(K), hC(K), hD(K), hF(K);
  std::vector<ReduceSpec> hR1(K), hR2(K);
  for (int32_t k = 0; k < K; ++k) {
    culamp_problem* pb = b->pbs[k];
    const int64_t n = pb->n, m = pb->m, on = b->off_n[k], om = b->off_m[k];
    const double* rk = b->r + on;
    const double* ck = b->c + om;
    double* dk = dyn + (int64_t)kDyn * k;
    auto lse = [&](const double* pot, double* dst, const double* slot) {
      PassDesc d = lse_rows_desc(pb->x, n, pb->y, m, pot, 1.0, 1.0, winf[k], pp, dst);
      feat(d, pb, true);
      d.prev_offset = 1;
      d.dyn = slot; d.dyn_kind = DYN_LSE; d.dyn_winf = winf[k];
      return d;
    };
    auto col = [&](const double* pot, const double* slot) {
      PassDesc d = colmarg_desc(pb->x, n, pb->y, m, rk, pot, logZ + on, 1.0, 1.0, winf[k], pp,
                                res + om, cost + om);
      feat(d, pb, false);
      d.dyn = slot; d.dyn_kind = DYN_COL; d.dyn_winf = winf[k];
      return d;
    };
    auto fuse = [&](PassDesc& d, int adjust, double* th_out, const double* mix_src, double* nu_o) {
      d.epi = EPI_LAMP_COL; d.fuse_update = 1; d.adjust = adjust;
      d.th0 = theta + om; d.cvec = ck; d.etamu = etamu + om; d.eta_mu = a->eta_mu;
      d.minv = minv; d.maxv = maxv; d.upd_winf = 1.0; d.th_out = th_out;
      d.nu_in = nu + om; d.mix_src = mix_src; d.nu_out = nu_o;
    };
    hE1[k] = lse(nu + om, logZ + on, dk);
    hE2[k] = col(nu + om, dk);
    if (eta[k] > 0) {
      // points: into tmp without offset; cellsim: into sumvals (logZ) with
      // offset, as the single-instance loops do
      PassDesc d = lse_rows_desc(pb->x, n, pb->y, m, theta + om, eta[k], 1.0, winf[k], pp,
                                 cellsim ? logZ + on : tmp + on);
      feat(d, pb, true);
      d.prev_offset = cellsim ? 1 : 0;
      hE3[k] = d;
    } else {
      PassDesc d = minplus_desc(pb->x, n, pb->y, m, cellsim ? gam + om : theta + om, winf[k], pp,
                                cellsim ? logZ + on : tmp + on);
      feat(d, pb, true);
      hE3[k] = d;
    }
    hA[k] = lse(nu + om, logZ + on, dk + 4);
    hB[k] = col(nu + om, dk + 8);
    fuse(hB[k], 0, thetab + om, theta + om, nub + om);
    hB[k].out1 = nullptr;  // the iteration passes need no column costs
    hC[k] = lse(nub + om, logZ + on, dk + 12);
    hD[k] = col(nub + om, dk + 16);
    hF[k] = hD[k];  // cellsim: the last iteration's column costs, recomputed at the end
    fuse(hD[k], 1, theta + om, thetab + om, nu + om);
    hD[k].out1 = nullptr;
    ReduceSpec r1{};
    r1.nops = 4;
    r1.op[0] = 1; r1.a[0] = res + om; r1.b[0] = ck; r1.len[0] = m;
    r1.op[1] = 0; r1.a[1] = cost + om; r1.len[1] = m;
    if (cellsim) {
      r1.op[2] = 2; r1.a[2] = rk; r1.b[2] = logZ + on; r1.len[2] = n;
      r1.op[3] = 2; r1.a[3] = ck; r1.b[3] = nu + om; r1.len[3] = m;
    } else {
      r1.op[2] = 2; r1.a[2] = res + om; r1.b[2] = nu + om; r1.len[2] = m;
      r1.op[3] = 2; r1.a[3] = rk; r1.b[3] = logZ + on; r1.len[3] = n;
    }
    hR1[k] = r1;
    ReduceSpec r2{};
    r2.nops = 2;
    r2.op[0] = 2; r2.a[0] = rk; r2.b[0] = cellsim ? logZ + on : tmp + on; r2.len[0] = n;
    r2.op[1] = 2; r2.a[1] = ck; r2.b[1] = (cellsim && !(eta[k] > 0)) ? gam + om : theta + om; r2.len[1] = m;
    hR2[k] = r2;
  }
  DevArr<PassDesc> DE1, DE2, DE3, DA, DB, DC, DD, DF;
  DevArr<ReduceSpec> DR1, DR2;
  TRY(upload_descs(DE1, hE1, ctx->stream)); TRY(upload_descs(DE2, hE2, ctx->stream));
  TRY(upload_descs(DE3, hE3, ctx->stream)); TRY(upload_descs(DA, hA, ctx->stream));
  TRY(upload_descs(DB, hB, ctx->stream)); TRY(upload_descs(DC, hC, ctx->stream));
  TRY(upload_descs(DD, hD, ctx->stream));
  TRY(upload_descs(DF, hF, ctx->stream));
  TRY(DR1.upload(hR1, ctx->stream)); TRY(DR2.upload(hR2, ctx->stream));
  // the dual pass: LSE when eta > 0 (every instance of a batch shares the
  // sign of eta_p), min-plus (FP64) otherwise
  const bool dual_lse = a->eta_p > 0;
  TRY(E.prepare(prec, KIND_LSE, true));
  TRY(E.prepare(prec, KIND_SUM, false));
  if (!dual_lse) TRY(E.prepare(CULAMP_FP64, KIND_MIN, true));

  std::vector<int64_t> steps(K, 0), npass(K, 0);
  std::vector<int32_t> conv(K, 0);
  std::vector<double> primal(K, NAN), dual(K, NAN), resid(K, NAN), objective(K, NAN);
  std::vector<double> h(16 * (size_t)K);
  const bool use_graph = b->tune.graph != 0;
  GraphExec g;
  auto step = [&]() -> int {
    hipLaunchKernelGGL(advance_kernel, dim3(blocks_of((int64_t)E.active.size())), dim3(256), 0,
                       ctx->stream, dyn, d_eta, E.d_active.p, (int)E.active.size());
    HIP_TRY(hipGetLastError());
    TRY(E.launch(DA.p, prec, KIND_LSE, true, pp));
    TRY(E.launch(DB.p, prec, KIND_SUM, false, pp));
    TRY(E.launch(DC.p, prec, KIND_LSE, true, pp));
    TRY(E.launch(DD.p, prec, KIND_SUM, false, pp));
    if (cellsim) {
      hipLaunchKernelGGL(seg_gamma_kernel, dim3(blocks_of(maxm), (unsigned)E.active.size()), dim3(256),
                         0, ctx->stream, gam, theta, E.d_seg_m.p, E.d_active.p, dyn);
      HIP_TRY(hipGetLastError());
    }
    return CULAMP_OK;
  };
  // check (MirrorProxKernel.jl:1081-1101, MirrorProxCellSim.jl:144-166)
  auto check = [&](int64_t it, std::vector<int32_t>& still) -> int {
    TRY(E.launch(DE1.p, prec, KIND_LSE, true, pp));
    TRY(E.launch(DE2.p, prec, KIND_SUM, false, pp));
    TRY(E.reduce(DR1.p, d_out, 4));
    if (dual_lse) TRY(E.launch(DE3.p, prec, KIND_LSE, true, pp));
    else TRY(E.launch(DE3.p, CULAMP_FP64, KIND_MIN, true, pp));
    TRY(E.reduce(DR2.p, d_out + 8 * (int64_t)K, 2));
    HIP_TRY(hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    still.clear();
    for (int32_t k : E.active) {
      npass[k] += 3;
      const double* v = &h[8 * (size_t)k];
      const double* w = &h[8 * (size_t)K + 8 * (size_t)k];
      resid[k] = v[0];
      objective[k] = v[1];
      const double W = winf[k], et = etat[k], e = eta[k];
      if (cellsim) {
        if (e > 0) {
          primal[k] = objective[k] * (1.0 - e / et) - e * v[2] - 2.0 * W * e / et * v[3] + hr[k] +
                      2.0 * W * resid[k];
          dual[k] = -e * w[0] - 2.0 * W * w[1] + hr[k];
        } else {
          primal[k] = objective[k] + 2.0 * W * resid[k];
          dual[k] = w[0] - 2.0 * W * w[1];
        }
      } else {
        const double q = e / et;
        primal[k] = objective[k] * (1.0 - q) - q * v[2] + q * v[3] + 2.0 * W * resid[k];
        if (e > 0) dual[k] = -2.0 * e * w[0] - 2.0 * W * w[1] + 2.0 * hr[k];
        else dual[k] = w[0] - 2.0 * W * w[1];
      }
      // CULAMP_STOP_INFEAS skips the pre-step check (eta_t = inf)
      const bool met = a->stop == CULAMP_STOP_INFEAS
                           ? (std::isfinite(et) && resid[k] < a->epsilon / 6.0)
                           : (primal[k] - dual[k] < a->epsilon / 6.0 && resid[k] < a->epsilon / 6.0);
      if (met) {
        conv[k] = 1;
        steps[k] = it - 1;
      } else {
        still.push_back(k);
      }
    }
    return CULAMP_OK;
  };
  const bool check_it = a->verbose || a->stop == CULAMP_STOP_INFEAS;
  int64_t it = 1;  // the iteration about to run
  std::vector<int32_t> still;
  while (it <= a->itermax && !E.active.empty()) {
    if (now_s() - t0 > a->tmax) break;
    if (check_it && (it - 1) % frequency == 0) {
      TRY(check(it, still));
      if (still.size() != E.active.size()) {
        TRY(E.set_active(still));
        g.reset();
        if (E.active.empty()) break;
      }
    }
    const int64_t nsteps = std::min<int64_t>(frequency, a->itermax - it + 1);
    if (nsteps == frequency) {
      TRY(E.replay(g, use_graph, [&]() -> int {
        for (int64_t s = 0; s < nsteps; ++s) TRY(step());
        return CULAMP_OK;
      }));
    } else {
      for (int64_t s = 0; s < nsteps; ++s) TRY(step());
    }
    // host mirror of the device schedule (same operations, same order)
    for (int32_t k : E.active) {
      for (int64_t s = 0; s < nsteps; ++s) {
        const double etn = 1.0 / (1.0 + (1.0 / etat[k]) * (1.0 - eta[k]));
        etat[k] = etn;
        st[k] = (1.0 - etat[k]) * st[k] + etat[k];
      }
      npass[k] += 4 * nsteps;
    }
    it += nsteps;
  }
  for (int32_t k : E.active) steps[k] = it - 1;
  // epilogue over every instance
  TRY(E.set_active(all_instances(b)));
  std::vector<double> objf(K, NAN);
  if (!cellsim) {
    // final LSE of nu with each instance's own (eta_t, s_t) (MirrorProxKernel.jl:1124)
    TRY(E.launch(DE1.p, prec, KIND_LSE, true, pp));
    for (int32_t k = 0; k < K; ++k) npass[k] += 1;
  } else {
    // objective = sum of the last column-pass costs (MirrorProxCellSim.jl:184):
    // instances stopped by a check have them from the check's column pass,
    // the others recompute the last iteration's column pass with costs
    std::vector<int32_t> unconv;
    for (int32_t k = 0; k < K; ++k)
      if (!conv[k]) unconv.push_back(k);
    if (!unconv.empty()) {
      TRY(E.set_active(unconv));
      TRY(E.launch(DF.p, prec, KIND_SUM, false, pp));
      for (int32_t k : unconv) npass[k] += 1;
      TRY(E.set_active(all_instances(b)));
    }
    std::vector<ReduceSpec> ro(K);
    for (int32_t k = 0; k < K; ++k) {
      ReduceSpec r{};
      r.nops = 1; r.op[0] = 0; r.a[0] = cost + b->off_m[k]; r.len[0] = b->m[k];
      ro[k] = r;
    }
    DevArr<ReduceSpec> DRO;
    TRY(DRO.upload(ro, ctx->stream));
    TRY(E.reduce(DRO.p, d_out, 1));
    HIP_TRY(hipMemcpyAsync(h.data(), d_out, sizeof(double) * 8 * K, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int32_t k = 0; k < K; ++k) objf[k] = h[8 * (size_t)k];
  }
  if (phi_out) {
    hipLaunchKernelGGL(pooled_log_minus_kernel, dim3(blocks_of(N)), dim3(256), 0, ctx->stream, tmp, b->r,
                       logZ, N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(phi_out, tmp, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
  }
  std::vector<double> hnu(M);
  HIP_TRY(hipMemcpyAsync(hnu.data(), nu, sizeof(double) * M, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (nu_out) std::memcpy(nu_out, hnu.data(), sizeof(double) * M);
  if (psi_out) {
    // psi = -2 winf nu / eta_p (MirrorProxKernel.jl:1127)
    for (int32_t k = 0; k < K; ++k) {
      const double sc = -2.0 * winf[k] / a->eta_p;
      for (int64_t j = 0; j < b->m[k]; ++j) psi_out[b->off_m[k] + j] = sc * hnu[b->off_m[k] + j];
    }
  }
  const double elapsed = now_s() - t0;
  for (int32_t k = 0; k < K; ++k) {
    if (!results) break;
    culamp_result& o = results[k];
    std::memset(&o, 0, sizeof o);
    o.iterations = steps[k];
    o.objective = cellsim ? objf[k] : objective[k];
    o.residual = resid[k];
    o.primal = primal[k];
    o.dual = dual[k];
    o.winf = winf[k];
    o.eta_t = etat[k];
    o.elapsed_s = elapsed;
    o.converged = conv[k];
    o.passes = npass[k];
    o.pairs = (double)npass[k] * (double)b->n[k] * (double)b->m[k];
    o.nonfinite = cellsim ? !std::isfinite(objf[k])
                          : (check_it && !(std::isfinite(resid[k]) && std::isfinite(objective[k])));
  }
  return CULAMP_OK;
}

}  // namespace

void culamp_destroy_batches(culamp_ctx* ctx) {
  while (!ctx->batches.empty()) culamp_batch_destroy(*ctx->batches.begin());
}

extern "C" {

int culamp_batch_create(culamp_ctx* ctx, const culamp_instance* inst, int32_t count, int64_t d,
                        int metric, double scale, culamp_batch** out) {
  TRY(check_ctx(ctx));
  if (!out) return fail(CULAMP_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  if (!inst || count <= 0 || count > 65535)
    return fail(CULAMP_ERR_INVALID_ARG, "batch needs 1..65535 instances (got %d)", (int)count);
  if (d < 0) return fail(CULAMP_ERR_INVALID_ARG, "d must be >= 0");
  culamp_batch* b = new culamp_batch();
  b->ctx = ctx;
  b->count = count;
  b->feature = d > 0;
  b->metric = d > 0 ? metric : -1;
  b->d = d;
  b->scale = scale;
  b->tune.graph = 1;
  ctx->batches.insert(b);
  int rc = CULAMP_OK;
  // instances given the same x / y arrays (same host pointers and sizes)
  // share one device copy of their points or feature rows
  std::map<std::tuple<const double*, int64_t, const double*, int64_t>, culamp_problem*> geom;
  for (int32_t k = 0; k < count && rc == CULAMP_OK; ++k) {
    const culamp_instance& I = inst[k];
    culamp_problem* pb = nullptr;
    const double* ykey = (d > 0 && I.y == nullptr) ? I.x : I.y;
    const auto key = std::make_tuple(I.x, I.n, ykey, (d > 0 && I.y == nullptr) ? I.n : I.m);
    const auto it = geom.find(key);
    if (it != geom.end() && I.r != nullptr && I.c != nullptr) {
      rc = alias_problem(ctx, it->second, I.r, I.c, &pb);
    } else {
      if (d > 0) rc = culamp_feature_problem_create(ctx, I.x, I.n, I.y, I.m, d, metric, scale, I.r, I.c, &pb);
      else rc = culamp_problem_create(ctx, I.x, I.n, I.y, I.m, I.r, I.c, &pb);
      if (rc == CULAMP_OK) geom.emplace(key, pb);
    }
    if (rc != CULAMP_OK) {
      const std::string msg = g_err;
      culamp_batch_destroy(b);
      return fail(rc, "instance %d: %s", (int)k, msg.c_str());
    }
    b->pbs.push_back(pb);
    b->off_n.push_back(b->sum_n);
    b->off_m.push_back(b->sum_m);
    b->n.push_back(pb->n);
    b->m.push_back(pb->m);
    b->sum_n += pb->n;
    b->sum_m += pb->m;
  }
  if (b->feature) {
    b->zks = b->pbs[0]->zks;
    for (culamp_problem* pb : b->pbs)
      if (pb->zks != b->zks) b->zks = 0;
  }
  hipError_t e = hipMalloc(&b->r, sizeof(double) * b->sum_n);
  if (e == hipSuccess) e = hipMalloc(&b->c, sizeof(double) * b->sum_m);
  if (e == hipSuccess) e = hipMalloc(&b->logr, sizeof(double) * b->sum_n);
  for (int32_t k = 0; k < count && e == hipSuccess; ++k) {
    e = hipMemcpyAsync(b->r + b->off_n[k], b->pbs[k]->r, sizeof(double) * b->n[k], hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->c + b->off_m[k], b->pbs[k]->c, sizeof(double) * b->m[k], hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->logr + b->off_n[k], b->pbs[k]->logr, sizeof(double) * b->n[k], hipMemcpyDeviceToDevice, ctx->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    culamp_batch_destroy(b);
    return fail(CULAMP_ERR_HIP, "batch upload failed: %s", hipGetErrorString(e));
  }
  *out = b;
  return CULAMP_OK;
}

int culamp_batch_destroy(culamp_batch* b) {
  if (!b) return CULAMP_OK;
  culamp_ctx* ctx = b->ctx;
  ctx->batches.erase(b);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (culamp_problem* pb : b->pbs) culamp_problem_destroy(pb);
  for (double* p : {b->r, b->c, b->logr})
    if (p) (void)hipFree(p);
  delete b;
  return CULAMP_OK;
}

int culamp_batch_set_tuning(culamp_batch* b, const culamp_batch_tuning* t) {
  if (!b || !t) return fail(CULAMP_ERR_INVALID_ARG, "null batch or tuning");
  if (t->rows_per_thread != 0 && t->rows_per_thread != 1 && t->rows_per_thread != 2 && t->rows_per_thread != 4)
    return fail(CULAMP_ERR_INVALID_ARG, "rows_per_thread must be 0, 1, 2 or 4");
  if (t->max_split < 0 || t->min_items < 0) return fail(CULAMP_ERR_INVALID_ARG, "negative tuning value");
  b->tune = *t;
  return CULAMP_OK;
}

int culamp_batch_get_tuning(const culamp_batch* b, culamp_batch_tuning* t) {
  if (!b || !t) return fail(CULAMP_ERR_INVALID_ARG, "null batch or tuning");
  *t = b->tune;
  return CULAMP_OK;
}

int culamp_batch_autotune(culamp_ctx* ctx, culamp_batch* b, int precision, double p, int reps,
                          culamp_batch_tuning* chosen, double* ms_per_pass) {
  TRY(check_ctx(ctx));
  if (!b || b->ctx != ctx) return fail(CULAMP_ERR_INVALID_ARG, "null batch or batch of another context");
  if (precision < CULAMP_FP64 || precision > CULAMP_BF16X3)
    return fail(CULAMP_ERR_INVALID_ARG, "precision must be CULAMP_FP64, _MIXED, _BF16 or _BF16X3");
  if (reps <= 0) reps = 5;
  const int32_t K = b->count;
  DevBufs bufs;
  double *theta, *out;
  TRY(bufs.alloc(&theta, b->sum_m));
  TRY(bufs.alloc(&out, b->sum_n));
  hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(b->sum_m)), dim3(256), 0, ctx->stream, theta, 0.0, b->sum_m);
  hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(b->sum_n)), dim3(256), 0, ctx->stream, out, 0.0, b->sum_n);
  HIP_TRY(hipGetLastError());
  const double pp = b->feature ? 2.0 : p;
  std::vector<PassDesc> hd(K);
  for (int32_t k = 0; k < K; ++k) {
    culamp_problem* pb = b->pbs[k];
    PassDesc d = lse_rows_desc(pb->x, pb->n, pb->y, pb->m, theta + b->off_m[k], 0.05, 1.0, 1.0, pp,
                               out + b->off_n[k]);
    feat(d, pb, true);
    hd[k] = d;
  }
  DevArr<PassDesc> D;
  TRY(upload_descs(D, hd, ctx->stream));
  const culamp_batch_tuning saved = b->tune;
  std::vector<culamp_batch_tuning> cands;
  const bool points = !b->feature;
  const bool bf = b->feature && (precision == CULAMP_BF16 || precision == CULAMP_BF16X3) && b->zks > 0;
  const int max_rpt = points ? (precision == CULAMP_FP64 ? 2 : 4) : 1;
  int64_t maxs = 0;
  for (int32_t k = 0; k < K; ++k) maxs = std::max(maxs, b->m[k]);
  const int64_t floor_split = points ? kTile : (bf ? 64 : 64);
  for (int rpt = max_rpt; rpt >= 1; rpt /= 2) {
    for (int64_t L = maxs; ; L /= 2) {
      culamp_batch_tuning t = saved;
      t.rows_per_thread = points ? rpt : 0;
      t.max_split = L;
      cands.push_back(t);
      if (L / 2 < floor_split || cands.size() > 24) break;
    }
  }
  double best = INFINITY;
  culamp_batch_tuning best_t = saved;
  for (const culamp_batch_tuning& t : cands) {
    b->tune = t;
    Engine E(ctx, b);
    int rc = E.set_active(all_instances(b));
    if (rc == CULAMP_OK) rc = E.prepare(precision, KIND_LSE, true);
    if (rc == CULAMP_OK) rc = E.launch(D.p, precision, KIND_LSE, true, pp);  // warm-up
    if (rc != CULAMP_OK) {
      b->tune = saved;
      return rc;
    }
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    for (int q = 0; q < reps && rc == CULAMP_OK; ++q) rc = E.launch(D.p, precision, KIND_LSE, true, pp);
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    if (rc != CULAMP_OK) {
      b->tune = saved;
      return rc;
    }
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    const double per = (double)ms / reps;
    if (per < best) {
      best = per;
      best_t = t;
    }
  }
  b->tune = best_t;
  if (chosen) *chosen = best_t;
  if (ms_per_pass) *ms_per_pass = best;
  return CULAMP_OK;
}

int culamp_batch_time_pass(culamp_ctx* ctx, culamp_batch* b, int precision, double p, int reps,
                           double* ms_per_pass, int32_t* launches_per_pass) {
  TRY(check_ctx(ctx));
  if (!b || b->ctx != ctx) return fail(CULAMP_ERR_INVALID_ARG, "null batch or batch of another context");
  if (precision < CULAMP_FP64 || precision > CULAMP_BF16X3)
    return fail(CULAMP_ERR_INVALID_ARG, "precision must be CULAMP_FP64, _MIXED, _BF16 or _BF16X3");
  if (reps <= 0) return fail(CULAMP_ERR_INVALID_ARG, "reps must be positive");
  const int32_t K = b->count;
  DevBufs bufs;
  double *theta, *out;
  TRY(bufs.alloc(&theta, b->sum_m));
  TRY(bufs.alloc(&out, b->sum_n));
  hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(b->sum_m)), dim3(256), 0, ctx->stream, theta, 0.0, b->sum_m);
  hipLaunchKernelGGL(pooled_fill_kernel, dim3(blocks_of(b->sum_n)), dim3(256), 0, ctx->stream, out, 0.0, b->sum_n);
  HIP_TRY(hipGetLastError());
  const double pp = b->feature ? 2.0 : p;
  std::vector<PassDesc> hd(K);
  for (int32_t k = 0; k < K; ++k) {
    culamp_problem* pb = b->pbs[k];
    PassDesc d = lse_rows_desc(pb->x, pb->n, pb->y, pb->m, theta + b->off_m[k], 0.05, 1.0, 1.0, pp,
                               out + b->off_n[k]);
    feat(d, pb, true);
    hd[k] = d;
  }
  DevArr<PassDesc> D;
  TRY(upload_descs(D, hd, ctx->stream));
  Engine E(ctx, b);
  TRY(E.set_active(all_instances(b)));
  TRY(E.prepare(precision, KIND_LSE, true));
  // warm-up: as many launches as are timed (the device's clocks settle over
  // the first tens of milliseconds of a sustained load)
  for (int q = 0; q < std::max(reps, 1); ++q) TRY(E.launch(D.p, precision, KIND_LSE, true, pp));
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  for (int q = 0; q < reps; ++q) TRY(E.launch(D.p, precision, KIND_LSE, true, pp));
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(hipEventSynchronize(ctx->ev1));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (ms_per_pass) *ms_per_pass = (double)ms / reps;
  if (launches_per_pass) {
    const Engine::Route r = E.route(precision, KIND_LSE);
    const Plan* P = E.find(r.family, true, r.rpt);
    *launches_per_pass = (P && !P->mitems.empty()) ? 2 : 1;
  }
  return CULAMP_OK;
}

int culamp_batch_sinkhorn_solve(culamp_ctx* ctx, culamp_batch* b, const culamp_args* a,
                                int64_t frequency, double p, const double* phi0, const double* psi0,
                                double* phi, double* psi, culamp_result* results) {
  TRY(check_batch(ctx, b, a, frequency));
  return batch_sinkhorn(ctx, b, a, frequency, p, phi0, psi0, phi, psi, results);
}

int culamp_batch_lamp_solve(culamp_ctx* ctx, culamp_batch* b, const culamp_args* a,
                            int64_t frequency, double p, double* nu, double* phi, double* psi,
                            culamp_result* results) {
  TRY(check_batch(ctx, b, a, frequency));
  return batch_lamp(ctx, b, a, frequency, p, nu, phi, psi, results);
}

}  // extern "C"

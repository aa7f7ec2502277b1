// culamp_host.h -- host-side state shared by the translation units of
// libculamp (contexts, problems, batches) and the helpers the solver loops
// use to build and issue passes.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include "culamp.h"
#include "culamp_internal.h"

struct culamp_ctx {
  int device = 0;
  int num_cu = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_chk = nullptr;   // end of a solver check's reduction
  double* chk_host = nullptr;    // coherent pinned scalars the reductions write directly
  double* chk_dev = nullptr;     // device address of chk_host
  double* partial = nullptr;
  size_t partial_cap = 0;  // doubles
  double* scal = nullptr;      // device scalars for reductions
  double* scal_host = nullptr; // pinned mirror
  int force_rpt = 0, force_splits = 0;
  int layout = 0;              // culamp::Layout of point passes
  int records = 1;             // solver loops pass staged source records between passes
  int fp64_fallback = 1;       // reduced-precision solves hand over to FP64 when they stall
  double pairs = 0.0;          // pair evaluations issued (throughput accounting)
  int64_t passes = 0;
  std::set<culamp_problem*> problems;  // destroyed with the context
  std::set<culamp_batch*> batches;     // destroyed with the context (first)
  std::set<culamp_dense*> denses;      // destroyed with the context
};

struct culamp_problem {
  culamp_ctx* ctx = nullptr;
  int64_t n = 0, m = 0;
  double *x = nullptr, *y = nullptr, *r = nullptr, *c = nullptr;
  double *logr = nullptr, *logc = nullptr;
  std::vector<double> hr, hc;
  // feature problems (cell-similarity costs); metric < 0 for point clouds
  int metric = -1;
  int64_t d = 0;
  double scale = 1.0;
  double *xf64 = nullptr, *yf64 = nullptr, *xrc = nullptr, *yrc = nullptr;
  float *xf32 = nullptr, *yf32 = nullptr;
  bool self = false;  // y is x
  // shares the point / feature buffers of another problem (same inputs, own
  // marginals); only r, c, logr, logc belong to this problem
  const culamp_problem* geom_owner = nullptr;
  // MFMA path (cosine / correlation): unit rows in bf16, width 16 * zks
  uint16_t *xz_hi = nullptr, *xz_lo = nullptr, *yz_hi = nullptr, *yz_lo = nullptr;
  int zks = 0;  // 0: not eligible
};

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return culamp_host::fail(CULAMP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define TRY(expr)                      \
  do {                                 \
    int rc_ = (expr);                  \
    if (rc_ != CULAMP_OK) return rc_;  \
  } while (0)

namespace culamp_host {

using culamp::PassDesc;

extern thread_local std::string g_err;
int fail(int code, const char* fmt, ...);
double now_s();
std::string fmt(const char* f, double v);
void emit(culamp_log_fn fn, void* user, const std::string& s);
int ensure_work(culamp_ctx* ctx, size_t partial_doubles);
// A problem on `owner`'s geometry with its own marginals (n, m as owner's).
int alias_problem(culamp_ctx* ctx, const culamp_problem* owner, const double* r,
                  const double* c, culamp_problem** out);
int run_pass(culamp_ctx* ctx, const PassDesc& d, int precision, int kind);
PassDesc blank();
void feat(PassDesc& d, const culamp_problem* pb, bool rows);
int check_ctx(culamp_ctx* ctx);
int need_features(const culamp_problem* pb);
int reduce_to_host(culamp_ctx* ctx, const culamp::ReduceSpec& rs, double* out);
int reduce_start(culamp_ctx* ctx, const culamp::ReduceSpec& rs);
int reduce_finish(culamp_ctx* ctx, int nops, double* out);
int copy_out(culamp_ctx* ctx, double* host, const double* dev, int64_t n);
double sk_aux_p(double p);
// Looks up the device address of the no-allocation entry point's scratch
// array on `device` (current), once per device.
void culamp_noalloc_setup(int device);

PassDesc lse_rows_desc(const double* x, int64_t n, const double* y, int64_t m,
                       const double* theta, double reg, double st, double winf,
                       double p, double* out);
PassDesc colmarg_desc(const double* x, int64_t n, const double* y, int64_t m,
                      const double* r, const double* theta, const double* logZ,
                      double reg, double st, double winf, double p, double* res,
                      double* cost);
PassDesc sk_half_desc(const double* x, int64_t n, const double* y, int64_t m,
                      const double* marg, const double* psi, double reg,
                      double winf, double p, double* out);
PassDesc sk_resid_desc(const double* x, int64_t n, const double* y, int64_t m,
                       const double* marg, const double* phi, const double* psi,
                       double reg, double winf, double p, double* res,
                       double* cost);
PassDesc max_desc(const double* x, int64_t n, const double* y, int64_t m,
                  double p, double* out);
PassDesc minplus_desc(const double* x, int64_t n, const double* y, int64_t m,
                      const double* theta, double winf, double p, double* out);

// The FP64 takeover of a reduced-precision solve (culamp_ctx_set_fp64_fallback):
// fed the infeasibility of every check with a finite eta_t, it switches `prec`
// to FP64 once the infeasibility has stayed at or above the stop tolerance for
// kStaleChecks consecutive checks without reaching kGain times its best value.
struct PrecisionGuard {
  static constexpr int kStaleChecks = 10;
  static constexpr double kGain = 0.97;
  int prec;
  bool enabled;
  double best = INFINITY;
  int stale = 0;
  int64_t fp64_from = 0;  // iteration of the switch (0: none)
  PrecisionGuard(int precision, bool on) : prec(precision), enabled(on && precision != CULAMP_FP64) {}
  // Returns true when the caller's later passes must run in FP64.
  bool observe(double resid, double tol, int64_t it) {
    if (!enabled || prec == CULAMP_FP64 || !(resid >= tol)) return false;
    if (resid < kGain * best) {
      best = resid;
      stale = 0;
      return false;
    }
    if (++stale < kStaleChecks) return false;
    prec = CULAMP_FP64;
    fp64_from = it;
    return true;
  }
};

// RAII bundle of device buffers for one solve.
struct DevBufs {
  std::vector<void*> ptrs;
  ~DevBufs() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  int alloc(double** out, int64_t n) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (e != hipSuccess) return fail(CULAMP_ERR_ALLOC, "hipMalloc(%lld doubles): %s", (long long)n, hipGetErrorString(e));
    ptrs.push_back(p);
    *out = (double*)p;
    return CULAMP_OK;
  }
  int alloc_bytes(void** out, size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes > 0 ? bytes : 1);
    if (e != hipSuccess) return fail(CULAMP_ERR_ALLOC, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    ptrs.push_back(p);
    *out = p;
    return CULAMP_OK;
  }
};

}  // namespace culamp_host

// Destroys the batches of a context (defined with the batch engine).
void culamp_destroy_batches(culamp_ctx* ctx);

// culamp_noalloc.hip -- the row log-sum-exp entry point that allocates
// nothing (no reference counterpart), for callers that cannot call hipMalloc:
// code under HIP graph capture, tight serving loops, harnesses that forbid
// allocation.
//
// culamp_lse_rows keeps the split partials of its passes in a per-context
// buffer that grows with hipMalloc on demand.  culamp_lse_rows_noalloc runs
// the same kernels with the same launch shape, so its results are the same
// bits, but its partials live in `scratch`, a file-scope device array of
// kScratchDoubles doubles that every GPU the library runs on holds until the
// process ends.  Its device address is looked up once per device when a
// context is created on it (culamp_noalloc_setup, called from
// culamp_ctx_create with that device current) and kept per device; host code
// passes that address, never the array's name, to the kernels.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two contexts' streams, or two host threads) race.  Each pass writes
// every partial slot that its merge reads before the merge runs, so nothing
// needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "culamp.h"
#include "culamp_host.h"

using namespace culamp;
using namespace culamp_host;

namespace {

// 16 MiB: an automatic launch keeps splits * target blocks below twice its
// target of 8 workgroups per CU, so its partials (3 fields -- m, S and the
// compensation of S -- x splits x 128 padded targets per block) stay below
// 3 x 4096 x 128 < 2^21 doubles on a 256-CU device.
constexpr size_t kScratchDoubles = size_t(1) << 21;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

}  // namespace

namespace culamp_host {

// Looks up this device's address of `scratch` (once per device; the device
// must be current).  A failure is kept and reported by the entry point, so a
// context still works for every other call.
void culamp_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_done[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_status[device] = e;
  g_ptr[device] = e == hipSuccess ? static_cast<double*>(p) : nullptr;
  g_done[device] = true;
}

}  // namespace culamp_host

extern "C" {

int culamp_lse_rows_noalloc(culamp_ctx* ctx, const double* x, int64_t n, const double* y,
                            int64_t m, const double* theta, double reg, double st,
                            double winf, double p, int precision, double* out) {
  TRY(check_ctx(ctx));
  // (the same argument rules as culamp_lse_rows: no targets is a no-op)
  if (n < 0 || m < 0)
    return fail(CULAMP_ERR_INVALID_ARG, "negative size (n=%lld, m=%lld)", (long long)n, (long long)m);
  if (n == 0) return CULAMP_OK;
  if (precision < CULAMP_FP64 || precision > CULAMP_BF16X3)
    return fail(CULAMP_ERR_INVALID_ARG, "precision must be CULAMP_FP64, _MIXED, _BF16 or _BF16X3");
  double* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_done[ctx->device])
      return fail(CULAMP_ERR_HIP, "culamp_lse_rows_noalloc: no scratch address for device %d",
                  ctx->device);
    if (g_ptr[ctx->device] == nullptr)
      return fail(CULAMP_ERR_HIP, "culamp_lse_rows_noalloc: hipGetSymbolAddress failed on device %d: %s",
                  ctx->device, hipGetErrorString(g_status[ctx->device]));
    ptr = g_ptr[ctx->device];
  }
  // (point passes treat the bf16 settings as mixed, as culamp_lse_rows does)
  const int prec = precision > CULAMP_MIXED ? CULAMP_MIXED : precision;
  const PassDesc d = lse_rows_desc(x, n, y, m, theta, reg, st, winf, p, out);
  const LaunchCfg cfg = choose_launch(d.nt, d.ns, prec, KIND_LSE, ctx->num_cu, ctx->force_rpt,
                                      ctx->force_splits, ctx->layout);
  if (cfg.splits > 1 && !cfg.wide) {
    const size_t ntp = (size_t)cfg.tblocks * (kThreads / kLanesPerTarget) * cfg.rpt;
    const size_t need = (size_t)pass_fields(KIND_LSE) * (size_t)cfg.splits * ntp;
    if (need > kScratchDoubles)
      return fail(CULAMP_ERR_INVALID_ARG,
                  "culamp_lse_rows_noalloc: this launch (%d splits x %zu targets) needs %zu doubles "
                  "of split partials, above the %zu-double scratch array",
                  cfg.splits, ntp, need, kScratchDoubles);
  }
  const Work w{ptr};
  HIP_TRY(launch_pass(d, prec, KIND_LSE, cfg, w, ctx->stream));
  ctx->pairs += (double)d.nt * (double)d.ns;
  ctx->passes += 1;
  return CULAMP_OK;
}

}  // extern "C"

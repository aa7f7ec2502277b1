/*
 * culamp.h -- C ABI of libculamp, the MI355X (gfx950) implementation of
 * CuLAMP.jl's fused on-the-fly cost + row/column log-sum-exp hot path.
 *
 * Every entry point returns an int status (CULAMP_OK == 0).  On failure
 * culamp_last_error() returns a thread-local message.  No entry point retains a
 * caller pointer after it returns.
 *
 * Two levels mirror the reference:
 *   - kernel level: one symbol per device kernel of the reference, same math
 *     and argument meaning, operating on DEVICE pointers (allocated with
 *     culamp_malloc or any HIP allocation on the context's device) on the
 *     context's stream.  Calls are stream-ordered; culamp_ctx_synchronize
 *     waits for them.
 *   - solver level: one symbol per reference solver, taking HOST arrays (or a
 *     device-resident culamp_problem) and writing HOST outputs.  Solver calls
 *     return when their results are in the caller's buffers.
 *
 * Point clouds are 3 x n column-major FP64 (x[3*i + k]) as the reference's
 * `img1::CuDeviceMatrix{Float64}` / `locations` arrays.  The cost is the
 * reference kernels' `l2dist`: p == 1 sum|d|, p == 2 SQUARED Euclidean,
 * p == +Inf max|d|, otherwise sum|d|^p.
 *
 * precision: CULAMP_FP64 runs every pair in IEEE double.  CULAMP_MIXED runs
 * the cost and the exp-sum per pair in FP32 while potentials, partial-sum
 * merges, marginals and infeasibility stay FP64.
 *
 * One context owns one HIP stream; calls on one context are not reentrant.
 * Different contexts may be used from different threads.
 */
#ifndef CULAMP_H
#define CULAMP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CULAMP_ABI_VERSION 1

enum {
  CULAMP_OK = 0,
  CULAMP_ERR_INVALID_ARG = 1,
  CULAMP_ERR_HIP = 2,
  CULAMP_ERR_NONFINITE = 3,
  CULAMP_ERR_NOT_CONVERGED = 4,
  CULAMP_ERR_NO_DEVICE = 5,
  CULAMP_ERR_ALLOC = 6
};

/* precision: CULAMP_FP64 and CULAMP_MIXED apply to every pass.  For the
 * cosine and correlation feature costs, CULAMP_BF16 evaluates the row dots
 * as one bf16 MFMA product with FP32 accumulation and CULAMP_BF16X3 as the
 * three-product split (bf16 head and remainder); other passes treat both as
 * CULAMP_MIXED. */
enum { CULAMP_FP64 = 0, CULAMP_MIXED = 1, CULAMP_BF16 = 2, CULAMP_BF16X3 = 3 };
enum { CULAMP_STOP_REFERENCE = 0, CULAMP_STOP_INFEAS = 1 };
/* Feature metrics of src/CellUtilities.jl:5 ("pearson" = correlation). */
enum {
  CULAMP_METRIC_L1 = 0,
  CULAMP_METRIC_L2 = 1,
  CULAMP_METRIC_COSINE = 2,
  CULAMP_METRIC_CORRELATION = 3
};

typedef struct culamp_ctx culamp_ctx;
typedef struct culamp_problem culamp_problem;
typedef struct culamp_batch culamp_batch;
typedef struct culamp_graph culamp_graph;
typedef struct culamp_dense culamp_dense;

/* Mirrors EOTArgs (src/Utilities.jl:18-31) field for field, same defaults via
 * culamp_args_default.  Two fields are new and default to the reference's
 * behaviour: `precision` (CULAMP_FP64) and `stop` (CULAMP_STOP_REFERENCE).
 * CULAMP_STOP_INFEAS stops LAMP when the L1 infeasibility is below
 * epsilon/6, tested every `frequency` iterations whether or not `verbose` is
 * set, except at the first check (eta_t = inf there, so the plan is the
 * product r x 1/m whatever the costs, and it is feasible whenever c is
 * uniform).  The reference tests primal - dual and infeasibility, and only
 * when verbose (MirrorProxKernel.jl:1081,1098); its gap term is what keeps it
 * from stopping at that first check. */
typedef struct {
  double eta_p, eta_mu, tau_p, tau_mu, alpha, B, epsilon, anneal_mult;
  int64_t itermax, inner_iter;
  double tmax;
  int32_t verbose;
  int32_t precision;
  int32_t stop;
  int32_t reserved;
} culamp_args;

/* Solver summary (fields the reference returns or logs). */
typedef struct {
  int64_t iterations;  /* completed iterations (Sinkhorn: num_iter) */
  double objective;    /* transport cost <P, C> at exit */
  double residual;     /* L1 marginal infeasibility at exit */
  double primal, dual; /* primal / dual at the last check (NaN if none; the
                        * single-instance LAMP solver skips the dual pass,
                        * leaving dual NaN, when it stops on infeasibility
                        * without logging) */
  double winf;         /* W_inf = max_ij C_ij */
  double eta_t;        /* LAMP: final eta_t; Sinkhorn: eta */
  double elapsed_s;    /* wall time of the solve */
  double pairs;        /* cost + exp evaluations issued (sum of n*m over
                          O(n*m) passes), for throughput accounting */
  int64_t passes;      /* O(n*m) passes issued */
  int32_t converged;   /* stop test met */
  int32_t nonfinite;   /* a NaN/Inf was seen in the residual or objective */
  int64_t fp64_from;   /* reduced-precision solves: the iteration from which
                        * the passes ran in FP64 because the infeasibility
                        * stopped improving above the stop tolerance (see
                        * culamp_ctx_set_fp64_fallback) or, for a Sinkhorn
                        * divergence term, the iteration after which the
                        * forward ran in FP64 (polish or stall); 0 if they
                        * never did */
} culamp_result;

/* Log sink: called synchronously on the calling thread with one CSV line in
 * the reference's exact format (header and rows, trailing newline). */
typedef void (*culamp_log_fn)(void *user, const char *line);

/* ---- library / context -------------------------------------------------- */
int culamp_abi_version(void);
const char *culamp_last_error(void);
void culamp_args_default(culamp_args *args);
int culamp_device_count(int *count);
int culamp_ctx_create(int device, culamp_ctx **out);
int culamp_ctx_destroy(culamp_ctx *ctx);
int culamp_ctx_synchronize(culamp_ctx *ctx);
/* Launch-shape override (no reference counterpart; used by tests and the
 * tuner): targets per thread in {1, 2, 4} and number of source splits; 0
 * selects automatically. */
int culamp_ctx_set_launch(culamp_ctx *ctx, int rows_per_thread, int splits);
/* Layout of the source splits of point-cloud passes (no reference
 * counterpart; tests and tuning): 0 automatic, 1 splits over workgroups
 * merged by a second launch, 2 up to 16 splits inside one workgroup of 16
 * waves, merged in LDS.  With the same split length both give identical
 * results.  With layout 2 a nonzero `splits` of culamp_ctx_set_launch caps
 * the in-workgroup split count. */
int culamp_ctx_set_layout(culamp_ctx *ctx, int layout);
/* Staged source records (no reference counterpart; tests and tuning): with
 * 1 (the default) the mixed-precision LAMP loop has each pass's epilogue
 * write the FP32 source records (scaled coordinates, exponent head and
 * weight) of the pass after it, which then skips converting its sources; 0
 * makes every pass convert its own.  Both give identical results. */
int culamp_ctx_set_source_records(culamp_ctx *ctx, int enable);
/* FP64 takeover of reduced-precision solves (no reference counterpart): with
 * 1 (the default) the single-instance LAMP, cell-similarity LAMP and Sinkhorn
 * loops (and the annealed stages) in CULAMP_MIXED, _BF16 or _BF16X3 switch
 * every later pass to FP64 once the L1 infeasibility has stayed above the
 * stop tolerance epsilon/6 for 10 consecutive checks without reaching 0.97
 * times its best value so far.  FP32 exponents carry an absolute error of
 * about 2^-24 times their magnitude, which at small eta_p (C / eta_t of the
 * transported pairs in the thousands) puts a floor above 1e-6 under the
 * infeasibility.  culamp_result.fp64_from reports the switch.  0 keeps the
 * requested precision throughout (to measure that floor). */
int culamp_ctx_set_fp64_fallback(culamp_ctx *ctx, int enable);
/* Compute-rate probe (no reference counterpart; measurement): the device's
 * throughput of the FP32 exponential (v_exp_f32, one per pair in every pass)
 * and of the FP32 fused multiply-add, in lane operations per second, from
 * independent-chain micro-kernels timed on the context's stream (best of 3). */
int culamp_probe_rates(culamp_ctx *ctx, double *exp_per_s, double *fma_per_s);
/* MFMA-rate probe (no reference counterpart; measurement): sustained dense
 * bf16 matrix-core throughput in flop/s, from independent chains of
 * v_mfma_f32_32x32x16_bf16 at three waves per SIMD (the occupancy of the
 * MFMA pair pass) on every CU, timed on the context's stream (best of 3). */
int culamp_probe_mfma(culamp_ctx *ctx, double *bf16_flops_per_s);
/* Device memory helpers (stream-ordered on the context's stream). */
int culamp_malloc(culamp_ctx *ctx, size_t bytes, void **dptr);
int culamp_free(culamp_ctx *ctx, void *dptr);
int culamp_memcpy_htod(culamp_ctx *ctx, void *dst, const void *src, size_t bytes);
int culamp_memcpy_dtoh(culamp_ctx *ctx, void *dst, const void *src, size_t bytes);
int culamp_memcpy_dtod(culamp_ctx *ctx, void *dst, const void *src, size_t bytes);
/* HIP-event timer on the context's stream: start, then stop returns the
 * elapsed milliseconds between the two events (waits for the stop event). */
int culamp_timer_start(culamp_ctx *ctx);
int culamp_timer_stop(culamp_ctx *ctx, double *ms);

/* Stream capture (no reference counterpart): kernel-level calls issued on the
 * context between begin and end are recorded into a HIP graph instead of
 * running, and run each time the graph is launched (stream-ordered).  Issue
 * each call once before capturing it, so that its work buffers exist (nothing
 * is allocated while capturing); solver-level calls cannot be captured. */
int culamp_capture_begin(culamp_ctx *ctx);
int culamp_capture_end(culamp_ctx *ctx, culamp_graph **graph);
int culamp_graph_launch(culamp_ctx *ctx, culamp_graph *graph);
int culamp_graph_destroy(culamp_graph *graph);

/* ---- kernel level (device pointers) -------------------------------------- */
/* max_logsumexp_spp_ct! (src/solvers/MirrorProxKernel.jl:280-326):
 * out[i] = max_j C(x_i, y_j). */
int culamp_max_cost(culamp_ctx *ctx, const double *x, int64_t n,
                    const double *y, int64_t m, double p, double *out);
/* warp_logsumexp_spp_ct_fused! (MirrorProxKernel.jl:833-938), same math as
 * warp_logsumexp_spp_ct_opt! (702-829):
 * out[i] = log sum_j exp(-(C(x_i,y_j) * st/(2 winf) + theta[j]) / reg). */
int culamp_lse_rows(culamp_ctx *ctx, const double *x, int64_t n,
                    const double *y, int64_t m, const double *theta,
                    double reg, double st, double winf, double p,
                    int precision, double *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): culamp_lse_rows with the same arguments, kernels, launch
 * shape and result bits, except that it calls no hipMalloc / hipFree and no
 * allocating routine.  Its split partials live in a static 16 MiB device array
 * (2^21 doubles) that each GPU the library runs on holds until the process
 * ends; a launch that would need more (only a forced shape can, see
 * culamp_ctx_set_launch) is rejected with CULAMP_ERR_INVALID_ARG.  The array
 * is shared by every call on one GPU and keeps the last call's contents:
 * calls that overlap on one GPU (several contexts or threads) race. */
int culamp_lse_rows_noalloc(culamp_ctx *ctx, const double *x, int64_t n,
                            const double *y, int64_t m, const double *theta,
                            double reg, double st, double winf, double p,
                            int precision, double *out);
/* residual_spp_c! (MirrorProxKernel.jl:7-94): for each column j
 *   w_ij = r[i] exp(-(C_ij st/(2 winf) + theta[j]) / reg - logZ[i])
 *   res[j] = sum_i w_ij,  cost[j] = sum_i w_ij C_ij   (cost may be NULL;
 *   the pass then skips the cost sum and res keeps its bits). */
int culamp_col_marginal_cost(culamp_ctx *ctx, const double *x, int64_t n,
                             const double *y, int64_t m, const double *r,
                             const double *theta, const double *logZ,
                             double reg, double st, double winf, double p,
                             int precision, double *res, double *cost);
/* warp_min_reduce! (MirrorProxKernel.jl:941-1016):
 * out[i] = min_j C(x_i,y_j) + 2 winf theta[j]. */
int culamp_minplus_rows(culamp_ctx *ctx, const double *x, int64_t n,
                        const double *y, int64_t m, const double *theta,
                        double winf, double p, double *out);
/* update_theta_residual (src/solvers/MirrorProx.jl:159-180), length m:
 * theta = tanh((res-c) winf/eta_mu_i... ) in the reference's overflow-safe
 * form; clamped to [minv, maxv] when adjust != 0.  theta may alias theta0. */
int culamp_lamp_theta_update(culamp_ctx *ctx, double *theta,
                             const double *theta0, const double *res,
                             const double *c, const double *eta_mu_i,
                             int64_t m, double eta_mu, int adjust, double minv,
                             double maxv, double winf);
/* warp_logsumexp_ct_fused! (src/solvers/SinkhornKernel.jl:341-444):
 * out[i] = log(marg[i]) - log sum_j exp(-C(x_i,y_j)/(reg winf) + psi[j]).
 * The reference's p == Inf Sinkhorn uses warp_logsumexp_ct_opt! (230-338),
 * which evaluates squared L2: pass p = 2 for that behaviour. */
int culamp_sinkhorn_half_step(culamp_ctx *ctx, const double *x, int64_t n,
                              const double *y, int64_t m, const double *marg,
                              const double *psi, double reg, double winf,
                              double p, int precision, double *out);
/* residual_opt! (SinkhornKernel.jl:9-81):
 *   w_ij = exp(-C_ij/(reg winf) + phi[i] + psi[j])
 *   res[i] = marg[i] - sum_j w_ij,  cost[i] = sum_j w_ij C_ij (may be NULL;
 *   the pass then skips the cost sum and res keeps its bits). */
int culamp_row_residual_cost(culamp_ctx *ctx, const double *x, int64_t n,
                             const double *y, int64_t m, const double *marg,
                             const double *phi, const double *psi, double reg,
                             double winf, double p, int precision, double *res,
                             double *cost);
/* Barycentric projections, out is 3 x n:
 * naive_findmaxindex_ct! (SinkhornKernel.jl:124-162):
 *   out_i = sum_j y_j w_ij / sum_j w_ij, w_ij = exp(-C/(reg winf) + phi_i + psi_j) */
int culamp_barycentric_sinkhorn(culamp_ctx *ctx, const double *x, int64_t n,
                                const double *y, int64_t m, const double *phi,
                                const double *psi, double reg, double winf,
                                double p, double *out);
/* naive_findmaxindex_spp_ct! (MirrorProxKernel.jl:97-142), rows (out: 3 x n):
 *   w_ij = r_i exp(-(C/(2 winf) + nu_j)/reg - logZ_i) */
int culamp_barycentric_lamp_rows(culamp_ctx *ctx, const double *x, int64_t n,
                                 const double *y, int64_t m, const double *nu,
                                 const double *r, const double *logZ,
                                 double reg, double winf, double p,
                                 double *out);
/* naive_findmaxindex_spp_ct_t! (MirrorProxKernel.jl:144-189), columns
 * (out: 3 x m): mean of x_i under w_ij for fixed j. */
int culamp_barycentric_lamp_cols(culamp_ctx *ctx, const double *x, int64_t n,
                                 const double *y, int64_t m, const double *r,
                                 const double *nu, const double *logZ,
                                 double reg, double winf, double p,
                                 double *out);

/* ---- problems (device-resident inputs) ----------------------------------- */
/* Copies x (3 x n), y (3 x m), r (n), c (m) from host memory to the device. */
int culamp_problem_create(culamp_ctx *ctx, const double *x, int64_t n,
                          const double *y, int64_t m, const double *r,
                          const double *c, culamp_problem **out);
int culamp_problem_destroy(culamp_problem *prob);
/* Cell-similarity (feature-cost) problem: rows of x (n x d) and y (m x d),
 * both COLUMN-major like the reference's `data::Matrix` (x[i + k*n]); y may
 * be NULL or equal to x for the reference's self-cost (data1 == data2).  The
 * cost is feature_cost_cuda (src/CellUtilities.jl:187-301) divided by
 * `scale` (_cost_scale, 78-84: l1 2, l2 sqrt 2, cosine 2, correlation 2). */
int culamp_feature_problem_create(culamp_ctx *ctx, const double *x, int64_t n,
                                  const double *y, int64_t m, int64_t d,
                                  int metric, double scale, const double *r,
                                  const double *c, culamp_problem **out);

/* ---- feature kernel level (device vectors on a feature problem) ---------- */
/* warp_logsumexp_spp_sim_fused! (src/solvers/MirrorProxCellSim.jl:259-342) */
int culamp_feature_lse_rows(culamp_ctx *ctx, culamp_problem *prob,
                            const double *theta, double reg, double st,
                            double winf, int precision, double *out);
/* residual_cellsim_c! (MirrorProxCellSim.jl:190-257) */
int culamp_feature_col_marginal_cost(culamp_ctx *ctx, culamp_problem *prob,
                                     const double *r, const double *theta,
                                     const double *logZ, double reg, double st,
                                     double winf, int precision, double *res,
                                     double *cost);
/* warp_logsumexp_cellsim_fused! (src/solvers/SinkhornCellSim.jl:153-238);
 * rows != 0: out (n) from the column potential pot (m), else the reverse. */
int culamp_feature_sinkhorn_half_step(culamp_ctx *ctx, culamp_problem *prob,
                                      int rows, const double *marg,
                                      const double *pot, double reg,
                                      double winf, int precision, double *out);
/* residual_cellsim_opt! (SinkhornCellSim.jl:1-72) */
int culamp_feature_row_residual_cost(culamp_ctx *ctx, culamp_problem *prob,
                                     const double *marg, const double *phi,
                                     const double *psi, double reg, double winf,
                                     int precision, double *res, double *cost);
/* max_logsumexp_cellsim! (MirrorProxCellSim.jl:1-41): max(0, max_j C_ij) */
int culamp_feature_max_cost(culamp_ctx *ctx, culamp_problem *prob, double *out);
/* warp_min_reduce_cellsim! (MirrorProxCellSim.jl:43-92) */
int culamp_feature_minplus_rows(culamp_ctx *ctx, culamp_problem *prob,
                                const double *theta, double winf, double *out);

/* ---- solver level --------------------------------------------------------- */
/* extragradient_color_transfer (src/solvers/MirrorProxKernel.jl:1019-1132):
 * LAMP mirror prox with on-the-fly cost.  Outputs (any may be NULL):
 *   nu (m), phi (n) = log r - logZ, psi (m) = -2 winf nu / eta_p,
 *   img1 (3 x n) and img2 (3 x m) barycentric projections. */
int culamp_lamp_solve(culamp_ctx *ctx, culamp_problem *prob,
                      const culamp_args *args, int64_t frequency, double p,
                      culamp_log_fn log_fn, void *user, double *nu, double *phi,
                      double *psi, double *img1, double *img2,
                      culamp_result *result);
/* sinkhorn_color_transfer (src/solvers/SinkhornKernel.jl:477-559).  phi0 (n)
 * and psi0 (m) are optional warm starts (NULL = zeros).  img1/img2 may be
 * NULL (return_assignments = false). */
int culamp_sinkhorn_solve(culamp_ctx *ctx, culamp_problem *prob,
                          const culamp_args *args, int64_t frequency, double p,
                          const double *phi0, const double *psi0,
                          culamp_log_fn log_fn, void *user, double *phi,
                          double *psi, double *img1, double *img2,
                          culamp_result *result);
/* annealed_sinkhorn_euclidean (src/solvers/AnnealedSinkhorn.jl:65-177). */
int culamp_annealed_sinkhorn_solve(culamp_ctx *ctx, culamp_problem *prob,
                                   const culamp_args *args, int64_t frequency,
                                   double p, culamp_log_fn log_fn, void *user,
                                   double *phi, double *psi, double *img1,
                                   double *img2, culamp_result *result);

/* extragradient_cellsim (src/solvers/MirrorProxCellSim.jl:94-188) on a
 * feature problem; nu, psi: m; phi: n (any may be NULL). */
int culamp_lamp_cellsim_solve(culamp_ctx *ctx, culamp_problem *prob,
                              const culamp_args *args, int64_t frequency,
                              culamp_log_fn log_fn, void *user, double *nu,
                              double *phi, double *psi, culamp_result *result);
/* sinkhorn_cellsim (src/solvers/SinkhornCellSim.jl:240-310) */
int culamp_sinkhorn_cellsim_solve(culamp_ctx *ctx, culamp_problem *prob,
                                  const culamp_args *args, int64_t frequency,
                                  const double *phi0, const double *psi0,
                                  culamp_log_fn log_fn, void *user, double *phi,
                                  double *psi, culamp_result *result);
/* annealed_sinkhorn_cellsim (src/solvers/AnnealedSinkhornCellSim.jl:1-63) */
int culamp_annealed_sinkhorn_cellsim_solve(culamp_ctx *ctx,
                                           culamp_problem *prob,
                                           const culamp_args *args,
                                           int64_t frequency,
                                           culamp_log_fn log_fn, void *user,
                                           double *phi, double *psi,
                                           culamp_result *result);

/* ---- batches: many independent instances per launch ----------------------
 * No single reference symbol: the reference solves one instance per call and
 * its only multi-instance workload loops over pairs on the host
 * (src/CellSimilarity.jl:164-255).  A batch runs every instance's passes in
 * one launch per pass (one workgroup per (instance, source split, target
 * block) item, largest items first), checks all instances every `frequency`
 * iterations and drops converged ones; the iterations between checks are
 * captured once per active set in a HIP graph and replayed.  Per instance the
 * arithmetic, stop test and outputs are those of the single-instance solver
 * named below (FP64: equal to it up to the summation order of split passes). */

/* One instance (host arrays, copied at culamp_batch_create). */
typedef struct {
  int64_t n, m;
  const double *x; /* point clouds: 3 x n; features: n x d column-major */
  const double *y; /* 3 x m, or m x d; features may pass NULL for y = x */
  const double *r; /* n */
  const double *c; /* m */
} culamp_instance;

/* Launch shape of a batch.  0 selects automatically. */
typedef struct {
  int32_t rows_per_thread; /* point kernel targets per thread: 1, 2 or 4 */
  int32_t graph;           /* 1: replay captured iterations (default), 0: eager */
  int64_t max_split;       /* most sources one workgroup reads per pass */
  int64_t min_items;       /* fewest workgroups a pass should have */
} culamp_batch_tuning;

/* d == 0: point clouds (3-D, cost chosen by p at solve time).  d > 0:
 * feature instances with `metric` and `scale` as culamp_feature_problem_create.
 * count <= 65535. */
int culamp_batch_create(culamp_ctx *ctx, const culamp_instance *inst,
                        int32_t count, int64_t d, int metric, double scale,
                        culamp_batch **out);
int culamp_batch_destroy(culamp_batch *batch);
int culamp_batch_set_tuning(culamp_batch *batch, const culamp_batch_tuning *t);
int culamp_batch_get_tuning(const culamp_batch *batch, culamp_batch_tuning *t);
/* Times candidate shapes of the batch's row log-sum-exp pass at `precision`
 * (HIP events, `reps` launches each) and keeps the fastest; *chosen (may be
 * NULL) receives it and *ms_per_pass (may be NULL) its time. */
int culamp_batch_autotune(culamp_ctx *ctx, culamp_batch *batch, int precision,
                          double p, int reps, culamp_batch_tuning *chosen,
                          double *ms_per_pass);
/* Measurement (no reference counterpart): times the batch's row
 * log-sum-exp pass over all instances at its current tuning, `reps` launches
 * between two HIP events on the context's stream after `reps` untimed
 * warm-up launches;
 * *ms_per_pass gets the mean and *launches_per_pass (may be NULL) the kernel
 * launches per pass (2 when some instance's source splits need the merge
 * kernel, else 1). */
int culamp_batch_time_pass(culamp_ctx *ctx, culamp_batch *batch, int precision,
                           double p, int reps, double *ms_per_pass,
                           int32_t *launches_per_pass);
/* Per instance: sinkhorn_color_transfer (points, SinkhornKernel.jl:477-559)
 * or sinkhorn_cellsim (features, SinkhornCellSim.jl:240-310).  phi0/psi0
 * (optional) and phi/psi are the instances' vectors concatenated in order
 * (sum n and sum m doubles); results has `count` entries.  `verbose` keeps
 * its meaning for the stop test; no log lines are written.  tmax is tested
 * between checks. */
int culamp_batch_sinkhorn_solve(culamp_ctx *ctx, culamp_batch *batch,
                                const culamp_args *args, int64_t frequency,
                                double p, const double *phi0,
                                const double *psi0, double *phi, double *psi,
                                culamp_result *results);
/* Per instance: extragradient_color_transfer without images (points,
 * MirrorProxKernel.jl:1019-1132) or extragradient_cellsim (features,
 * MirrorProxCellSim.jl:94-188).  nu, psi: sum m; phi: sum n (any may be
 * NULL). */
int culamp_batch_lamp_solve(culamp_ctx *ctx, culamp_batch *batch,
                            const culamp_args *args, int64_t frequency,
                            double p, double *nu, double *phi, double *psi,
                            culamp_result *results);

/* ---- differentiation: Sinkhorn divergence and its gradient ---------------
 * No reference counterpart (the reference has no gradients).  For squared
 * Euclidean costs C_ij = |x_i - y_j|^2 and a fixed regularisation eps,
 *   OT(a,x; b,y) = eps [<a,phi> + <b,psi> - <a,log a> - <b,log b>
 *                       - sum_ij P_ij + 1],   P_ij = e^{-C_ij/eps + phi_i + psi_j}
 * at the Sinkhorn potentials (the reference's log-domain convention,
 * SinkhornKernel.jl:341-444, with reg * W_inf = eps), and
 *   S = OT(a,x; b,y) - OT(a,x; a,x)/2 - OT(b,y; b,y)/2.
 * Gradients: the envelope term (partials at fixed potentials) plus the
 * adjoint of the last `unroll` Sinkhorn iterations of each solve
 * (truncated back-propagation through the log-sum-exp passes), recomputing
 * the forward from checkpoints stored every `checkpoint` iterations.  The
 * gradient in a is the unprojected partial derivative; only its projection
 * on {sum a = 0} is meaningful for normalised weights. */
typedef struct {
  double eps;         /* > 0: regularisation; <= 0: eta_p * max_ij C(x_i, y_j) */
  double eta_p;       /* used when eps <= 0 */
  double tol;         /* stop: L1 row infeasibility <= tol */
  double polish_tol;  /* mixed forward: once converged, continue in FP64 until the FP64
                         infeasibility is <= polish_tol (0: no FP64 stage); the FP32
                         rounding of the potentials otherwise limits the weight gradient */
  int64_t itermax;
  int64_t frequency;  /* infeasibility tested every `frequency` iterations */
  int32_t unroll;     /* iterations back-propagated at the end of each solve (0: envelope only) */
  int32_t checkpoint; /* checkpoint interval inside the unrolled window (>= 1) */
  int32_t precision;  /* forward passes: CULAMP_FP64 or CULAMP_MIXED; adjoint passes are FP64
                         (a mixed forward whose infeasibility stalls above tol continues in
                         FP64, see culamp_ctx_set_fp64_fallback) */
  int32_t reserved;
  double tmax;        /* each forward solve stops (not converged) after tmax seconds,
                         tested at every infeasibility check (default: no limit) */
} culamp_divergence_opts;

void culamp_divergence_opts_default(culamp_divergence_opts *opts);
/* value (may be NULL); gradients (any may be NULL): grad_x 3 x n, grad_a n,
 * grad_y 3 x m, grad_b m; terms (3 entries or NULL): the OT(a,b), OT(a,a)
 * and OT(b,b) solves (objective = the term's value). */
int culamp_sinkhorn_divergence(culamp_ctx *ctx, const double *x, int64_t n,
                               const double *a, const double *y, int64_t m,
                               const double *b,
                               const culamp_divergence_opts *opts,
                               double *value, double *grad_x, double *grad_a,
                               double *grad_y, double *grad_b,
                               culamp_result *terms);
/* Vector-Jacobian product: the gradients above times the cotangent dS. */
int culamp_sinkhorn_divergence_vjp(culamp_ctx *ctx, const double *x, int64_t n,
                                   const double *a, const double *y, int64_t m,
                                   const double *b,
                                   const culamp_divergence_opts *opts, double dS,
                                   double *grad_x, double *grad_a,
                                   double *grad_y, double *grad_b);
/* Kernel level (device pointers): one adjoint pass over targets t (3 x nt)
 * and sources s (3 x ns) with v_ts = -|t - s|^2 / eps + alpha srcA[s]
 * + gamma tgtA[t] and weights srcB[s] (NULL: 1):
 *   S_t = sum_s srcB[s] e^{v_ts},  Y_t = sum_s srcB[s] e^{v_ts} s
 *   normalized == 0: out1[t] = beta out1[t] + c1 S_t (beta 0: overwrite),
 *                    out0[3t+k] += c0 (t_k S_t - Y_tk)
 *   normalized != 0: out0[3t+k] += c0 tgtW[t] (t_k - Y_tk / S_t)
 * srcA, tgtA, tgtW, out0, out1 may be NULL. */
int culamp_lse_adjoint(culamp_ctx *ctx, const double *tgt, int64_t nt,
                       const double *src, int64_t ns, double eps,
                       const double *srcA, double alpha, const double *srcB,
                       const double *tgtA, double gamma, int normalized,
                       const double *tgtW, double c0, double c1, double beta,
                       double *out0, double *out1);

/* ---- dense cost matrix (the plain dense GPU path) -----------------------
 * The reference's dense-W kernels, kept as the HBM-streaming comparator of
 * the on-the-fly passes: the same row LSE and column marginal computed from
 * a stored n x m cost matrix.  culamp_dense_create uploads W (host, n x m
 * row-major FP64, i.e. the bytes of the reference's column-major Wt) and its
 * transpose, as FP64 (CULAMP_FP64) or FP32 (CULAMP_MIXED: FP32 costs and
 * exponents, FP64 sums); the passes take device vectors. */
int culamp_dense_create(culamp_ctx *ctx, const double *w, int64_t n, int64_t m,
                        int precision, culamp_dense **out);
int culamp_dense_destroy(culamp_dense *dm);
/* warp_logsumexp_fused! (src/solvers/MirrorProx.jl:236-296):
 * out_i = LSE_j(-(st W_ij / (2 W_inf) + theta_j) / reg), i < n. */
int culamp_dense_lse_rows(culamp_ctx *ctx, const culamp_dense *dm, const double *theta,
                          double reg, double st, double winf, double *out);
/* residual_c! (src/solvers/MirrorProx.jl:299-333):
 * out_j = sum_i r_i exp(-(st W_ij / (2 W_inf) + theta_j) / reg - logZ_i), j < m. */
int culamp_dense_col_marginal(culamp_ctx *ctx, const culamp_dense *dm, const double *r,
                              const double *theta, const double *logZ, double reg,
                              double st, double winf, double *out);

/* LAMP over the stored matrix, LAMP(r, c, W::CuArray, args, frequency;
 * _θ) (src/solvers/MirrorProx.jl:397-503): r (n) and c (m) host vectors,
 * theta0 (m, host) or NULL for zeros; returns theta (m) and, when plan is
 * not NULL, the final plan r .* softmax(...) (n x m row-major).  Logs, when
 * args->verbose, the reference's "time(s),iter,infeas,ot_objective,primal,
 * dual,pdgap,solver" lines; stops when infeas and primal - dual are below
 * epsilon / 2. */
int culamp_dense_lamp_solve(culamp_ctx *ctx, const culamp_dense *dm, const double *r,
                            const double *c, const culamp_args *args, int64_t frequency,
                            const double *theta0, culamp_log_fn log_fn, void *user,
                            double *theta, double *plan, culamp_result *out);

#ifdef __cplusplus
}
#endif
#endif

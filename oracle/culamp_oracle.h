/*
 * culamp_oracle.h -- FP64 CPU restatement of CuLAMP.jl's on-the-fly cost +
 * log-sum-exp hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Tests, the smoke check and the CPU-baseline leg
 * of the benchmark load this library to compare the HIP library against it.  It is never linked into, called from, or used as a
 * fallback by the product library (libculamp).
 *
 * Every function states the reference file:line it restates (paths relative to
 * the CuLAMP.jl repository root).  Conventions:
 *   - point clouds are 3 x n column-major (x[3*i + k]), as the reference's
 *     `img1::CuDeviceMatrix{T}` / `locations` arrays;
 *   - p == 1: sum |d|, p == 2: SQUARED Euclidean, p == +Inf: max |d|,
 *     otherwise sum |d|^p (no root), exactly the kernels' `l2dist`
 *     (src/solvers/MirrorProxKernel.jl:869-878);
 *   - all arithmetic is IEEE double.  Loops over targets are OpenMP-parallel
 *     (or_set_threads), loops over sources are sequential, so results do not
 *     depend on the thread count.
 */
#ifndef CULAMP_ORACLE_H
#define CULAMP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors EOTArgs (src/Utilities.jl:18-31) field for field. */
typedef struct {
  double eta_p, eta_mu, tau_p, tau_mu, alpha, B, epsilon, anneal_mult;
  int64_t itermax, inner_iter;
  double tmax;
  int32_t verbose;
  int32_t stop;  /* 0: the reference's stop test; 1: infeasibility < eps/6
                    only, tested every `frequency` iterations (LAMP) */
} or_args;

typedef void (*or_log_fn)(void *user, const char *line);

typedef struct {
  int64_t iterations;   /* loop index at exit (LAMP) or num_iter (Sinkhorn) */
  double objective;     /* transport cost at exit */
  double residual;      /* L1 infeasibility at exit */
  double primal, dual;  /* last logged primal/dual (LAMP), NaN if none */
  double winf;          /* W_inf used */
  double eta_t;         /* LAMP: final eta_t */
  int32_t converged;    /* stop test met */
  int32_t _pad;
} or_result;

void or_set_threads(int n);
int or_get_threads(void);

double or_cost3(const double *a, const double *b, double p);

/* max_logsumexp_spp_ct! (src/solvers/MirrorProxKernel.jl:280-326): out_i = max_j C_ij */
void or_max_cost_rows(const double *x, int64_t n, const double *y, int64_t m,
                      double p, double *out);
/* warp_logsumexp_spp_ct_fused! (MirrorProxKernel.jl:833-938) and the two-pass
 * warp_logsumexp_spp_ct_opt! (702-829): out_i = LSE_j(-(C_ij*st/(2 W) + th_j)/reg) */
void or_lse_rows(const double *x, int64_t n, const double *y, int64_t m,
                 const double *theta, double reg, double st, double winf,
                 double p, double *out);
/* residual_spp_c! (MirrorProxKernel.jl:7-94): column marginal and cost */
void or_col_marginal_cost(const double *x, int64_t n, const double *y, int64_t m,
                          const double *r, const double *theta,
                          const double *logZ, double reg, double st,
                          double winf, double p, double *res, double *cost);
/* warp_min_reduce! (MirrorProxKernel.jl:941-1016): out_i = min_j C_ij + 2 W th_j */
void or_minplus_rows(const double *x, int64_t n, const double *y, int64_t m,
                     const double *theta, double winf, double p, double *out);
/* update_theta_residual (src/solvers/MirrorProx.jl:159-180) */
void or_theta_update(double *theta, const double *theta0, const double *res,
                     const double *c, const double *eta_mu_i, int64_t m,
                     double eta_mu, int adjust, double minv, double maxv,
                     double winf);
/* warp_logsumexp_ct_fused! / _ct_opt! (src/solvers/SinkhornKernel.jl:341-444,
 * 230-338): out_i = log(marg_i) - LSE_j(-C_ij/(reg W) + psi_j) */
void or_sinkhorn_half_step(const double *x, int64_t n, const double *y,
                           int64_t m, const double *marg, const double *psi,
                           double reg, double winf, double p, double *out);
/* residual_opt! (SinkhornKernel.jl:9-81): row residual and row cost */
void or_row_residual_cost(const double *x, int64_t n, const double *y,
                          int64_t m, const double *marg, const double *phi,
                          const double *psi, double reg, double winf, double p,
                          double *res, double *cost);
/* naive_findmaxindex_ct! (SinkhornKernel.jl:124-162) */
void or_barycentric_sinkhorn(const double *x, int64_t n, const double *y,
                             int64_t m, const double *phi, const double *psi,
                             double reg, double winf, double p, double *out);
/* naive_findmaxindex_spp_ct! (MirrorProxKernel.jl:97-142) */
void or_barycentric_lamp_rows(const double *x, int64_t n, const double *y,
                              int64_t m, const double *nu, const double *r,
                              const double *logZ, double reg, double st,
                              double winf, double p, double *out);
/* naive_findmaxindex_spp_ct_t! (MirrorProxKernel.jl:144-189) */
void or_barycentric_lamp_cols(const double *x, int64_t n, const double *y,
                              int64_t m, const double *r, const double *nu,
                              const double *logZ, double reg, double st,
                              double winf, double p, double *out);

/* extragradient_color_transfer (MirrorProxKernel.jl:1019-1132).  Outputs may be
 * NULL.  nu, psi: m; phi: n; img1: 3n; img2: 3m. */
int or_lamp_solve(const double *x, int64_t n, const double *y, int64_t m,
                  const double *r, const double *c, const or_args *args,
                  int64_t frequency, double p, or_log_fn logfn, void *user,
                  double *nu, double *phi, double *psi, double *img1,
                  double *img2, or_result *res);
/* sinkhorn_color_transfer (SinkhornKernel.jl:477-559).  phi0/psi0 may be NULL
 * (zeros).  phi: n, psi: m, img1: 3n, img2: 3m (may be NULL). */
int or_sinkhorn_solve(const double *x, int64_t n, const double *y, int64_t m,
                      const double *r, const double *c, const or_args *args,
                      int64_t frequency, double p, const double *phi0,
                      const double *psi0, or_log_fn logfn, void *user,
                      double *phi, double *psi, double *img1, double *img2,
                      or_result *res);
/* annealed_sinkhorn_euclidean (src/solvers/AnnealedSinkhorn.jl:65-177) */
int or_annealed_sinkhorn_solve(const double *x, int64_t n, const double *y,
                               int64_t m, const double *r, const double *c,
                               const or_args *args, int64_t frequency, double p,
                               or_log_fn logfn, void *user, double *phi,
                               double *psi, double *img1, double *img2,
                               or_result *res);
/* sinkhorn_log on a dense W (src/solvers/Sinkhorn.jl:8-84).  W is n x m
 * row-major here (W[i*m + j]); P (n*m) may be NULL. */
int or_sinkhorn_log_dense(const double *r, int64_t n, const double *c,
                          int64_t m, const double *W, const or_args *args,
                          int64_t frequency, or_log_fn logfn, void *user,
                          double *phi, double *psi, double *P, or_result *res);
/* get_euclidean_distance (src/Utilities.jl:145-157) on an h x w grid,
 * N = h*w, W row-major N x N (symmetric). */
void or_grid_cost_matrix(int64_t height, int64_t width, double p, double *W);


/* ---- feature (cell-similarity) costs -------------------------------------
 * Points are rows of an n x d matrix stored COLUMN-major as the reference's
 * `data::Matrix` (x[i + k*n]); metric: 0 l1, 1 l2, 2 cosine, 3 correlation
 * (src/CellUtilities.jl:5, 78-84, 187-301). */
#ifndef OR_FEATURE_METRICS
#define OR_FEATURE_METRICS
enum { OR_L1 = 0, OR_L2 = 1, OR_COSINE = 2, OR_CORRELATION = 3 };
#endif
/* get_flat_metrics (src/CellUtilities.jl:304-311) */
void or_feature_stats(const double *x, int64_t n, int64_t d, double *sums,
                      double *sqnorms, double *means, double *csqnorms);
/* feature_cost_cuda (src/CellUtilities.jl:187-301) between row i of x and row
 * j of y (each with its own stats) -> C, divided by `scale`. */
double or_feature_cost(int metric, const double *x, int64_t n, int64_t i,
                       const double *xs, const double *y, int64_t m, int64_t j,
                       const double *ys, int64_t d, double scale);
/* C (n x m row-major) for all pairs. */
void or_feature_cost_matrix(int metric, const double *x, int64_t n,
                            const double *y, int64_t m, int64_t d, double scale,
                            double *C);
/* extragradient_cellsim (src/solvers/MirrorProxCellSim.jl:94-188) on one
 * data matrix (N = M = n).  nu, phi, psi: n.  res->objective is the
 * reference's returned objective (sum of the last column costs). */
int or_lamp_cellsim_solve(const double *data, int64_t n, int64_t d,
                          double scale, int metric, const double *r,
                          const double *c, const or_args *args,
                          int64_t frequency, or_log_fn logfn, void *user,
                          double *nu, double *phi, double *psi, or_result *res);
int or_lamp_cellsim2_solve(const double *data1, int64_t n, const double *data2,
                           int64_t m, int64_t d, double scale, int metric,
                           const double *r, const double *c, const or_args *args,
                           int64_t frequency, or_log_fn logfn, void *user,
                           double *nu_out, double *phi_out, double *psi_out,
                           or_result *out);

/* LAMP(r, c, W::AbstractMatrix, args, frequency) (src/solvers/MirrorProx.jl:
 * 506-575), the reference's CPU path: dense W (n x n row-major), FP64.
 * plan (n x n, may be NULL) = r .* p of the final state; theta (n). */
int or_cpu_lamp_solve(const double *r, const double *c, int64_t n, const double *W,
                      const or_args *args, int64_t frequency, or_log_fn logfn, void *user,
                      double *plan, double *theta, or_result *res);

/* sinkhorn_cellsim (src/solvers/SinkhornCellSim.jl:240-310): within each
 * half-step the cost is evaluated among the rows of that call's first data
 * argument (SinkhornCellSim.jl:43, 193); the reference calls it with
 * data1 == data2. */
int or_sinkhorn_cellsim_solve(const double *data1, int64_t n,
                              const double *data2, int64_t m, int64_t d,
                              double scale, int metric, const double *r,
                              const double *c, const or_args *args,
                              int64_t frequency, const double *phi0,
                              const double *psi0, or_log_fn logfn, void *user,
                              double *phi, double *psi, or_result *res);
/* annealed_sinkhorn_cellsim (src/solvers/AnnealedSinkhornCellSim.jl:1-63) */
int or_annealed_sinkhorn_cellsim_solve(const double *data1, int64_t n,
                                       const double *data2, int64_t m,
                                       int64_t d, double scale, int metric,
                                       const double *r, const double *c,
                                       const or_args *args, int64_t frequency,
                                       or_log_fn logfn, void *user, double *phi,
                                       double *psi, or_result *res);

#ifdef __cplusplus
}
#endif
#endif

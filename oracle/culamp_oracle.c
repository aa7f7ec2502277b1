/*
 * culamp_oracle.c -- FP64 CPU restatement of CuLAMP.jl's hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the results are checked against it; never part
 * of the product path).
 * See culamp_oracle.h for conventions.  Each function cites the reference
 * file:line it restates.  Reductions over sources are sequential in index
 * order; log-sum-exps are two-pass (max, then sum of exp(v - max)), which is
 * the arithmetic of warp_logsumexp_spp_ct_opt! (MirrorProxKernel.jl:702-829)
 * and equal in exact arithmetic to the online form of the fused kernels.
 */
#include "culamp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void or_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int or_get_threads(void) { return g_threads; }

#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(g_threads)")
#else
#define OMP_FOR
#endif

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Julia's @printf prints non-finite values as Inf/-Inf/NaN. */
static void fmt_e(char *buf, size_t n, const char *fmt, double v) {
  if (isnan(v)) snprintf(buf, n, "NaN");
  else if (isinf(v)) snprintf(buf, n, v > 0 ? "Inf" : "-Inf");
  else snprintf(buf, n, fmt, v);
}

/* Cost of MirrorProxKernel.jl:869-878 (and every other kernel): p == 2 is the
 * squared Euclidean distance evaluated as muladd(dr,dr,muladd(dg,dg,db*db)). */
double or_cost3(const double *a, const double *b, double p) {
  double dr = a[0] - b[0], dg = a[1] - b[1], db = a[2] - b[2];
  if (p == 1.0) return fabs(dr) + fabs(dg) + fabs(db);
  if (p == 2.0) return dr * dr + (dg * dg + db * db);
  if (p == INFINITY) return fmax(fabs(dr), fmax(fabs(dg), fabs(db)));
  return pow(fabs(dr), p) + pow(fabs(dg), p) + pow(fabs(db), p);
}

/* Sinkhorn kernels other than the fused LSE only know p == 1 or squared L2
 * (SinkhornKernel.jl:38-42, 259-263, 146-150). */
static double sk_aux_p(double p) { return p == 1.0 ? 1.0 : 2.0; }

/* MirrorProxKernel.jl:280-326 */
void or_max_cost_rows(const double *x, int64_t n, const double *y, int64_t m,
                      double p, double *out) {
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double mx = -INFINITY;
    for (int64_t j = 0; j < m; ++j) {
      double c = or_cost3(x + 3 * i, y + 3 * j, p);
      if (c > mx) mx = c;
    }
    out[i] = mx;
  }
}

static double lse_finish(double mx, double s) {
  if (mx == -INFINITY) return -INFINITY;
  if (mx == INFINITY) return INFINITY;
  return log(s) + mx;
}

/* MirrorProxKernel.jl:833-938 / 702-829: v = -(C*c1 + th_j) * invreg with
 * c1 = st/(2W), invreg = 1/reg.  reg = Inf gives invreg = 0 (first LAMP
 * iteration, where eta_t = Inf). */
void or_lse_rows(const double *x, int64_t n, const double *y, int64_t m,
                 const double *theta, double reg, double st, double winf,
                 double p, double *out) {
  const double c1 = st / (2.0 * winf);
  const double invreg = 1.0 / reg;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double mx = -INFINITY;
    for (int64_t j = 0; j < m; ++j) {
      double c = or_cost3(x + 3 * i, y + 3 * j, p);
      double v = -(c * c1 + theta[j]) * invreg;
      if (v > mx) mx = v;
    }
    double s = 0.0;
    if (isfinite(mx))
      for (int64_t j = 0; j < m; ++j) {
        double c = or_cost3(x + 3 * i, y + 3 * j, p);
        double v = -(c * c1 + theta[j]) * invreg;
        s += exp(v - mx);
      }
    out[i] = lse_finish(mx, s);
  }
}

/* MirrorProxKernel.jl:7-94: for column j,
 *   value = (C_ij*c1 + th_j)*(-invreg) - logZ_i
 *   res_j = sum_i exp(value) r_i,  cost_j = sum_i exp(value) C_ij r_i */
void or_col_marginal_cost(const double *x, int64_t n, const double *y, int64_t m,
                          const double *r, const double *theta,
                          const double *logZ, double reg, double st,
                          double winf, double p, double *res, double *cost) {
  const double c1 = st / (2.0 * winf);
  const double invreg = 1.0 / reg;
  OMP_FOR
  for (int64_t j = 0; j < m; ++j) {
    double acc = 0.0, cacc = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double c = or_cost3(y + 3 * j, x + 3 * i, p);
      double value = (c * c1 + theta[j]) * (-invreg) - logZ[i];
      double e = exp(value);
      acc += e * r[i];
      cacc += e * c * r[i];
    }
    res[j] = acc;
    if (cost) cost[j] = cacc;
  }
}

/* MirrorProxKernel.jl:941-1016: v = muladd(th_j, 2W, C_ij), out_i = min_j v */
void or_minplus_rows(const double *x, int64_t n, const double *y, int64_t m,
                     const double *theta, double winf, double p, double *out) {
  const double c1 = 2.0 * winf;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double mn = INFINITY;
    for (int64_t j = 0; j < m; ++j) {
      double v = theta[j] * c1 + or_cost3(x + 3 * i, y + 3 * j, p);
      if (v < mn) mn = v;
    }
    out[i] = mn;
  }
}

/* MirrorProx.jl:159-180, written in the same overflow-safe form:
 *   d = 2W(res - c)/eta_mu_i;  e = exp(d - |d|) * ((th0+1)/(1-th0))^(1-eta_mu)
 *   th = (e - exp(-|d|)) / (e + exp(-|d|)),  optionally clamped. */
void or_theta_update(double *theta, const double *theta0, const double *res,
                     const double *c, const double *eta_mu_i, int64_t m,
                     double eta_mu, int adjust, double minv, double maxv,
                     double winf) {
  for (int64_t j = 0; j < m; ++j) {
    double d = 2.0 * winf * (res[j] - c[j]) / eta_mu_i[j];
    double th0 = theta0[j];
    double mx = fmax(d, -d);
    double e = exp(d - mx) * pow((th0 + 1.0) / (1.0 - th0), 1.0 - eta_mu);
    double em = exp(-mx);
    double t = (e - em) / (e + em);
    if (adjust) t = t < minv ? minv : (t > maxv ? maxv : t);
    theta[j] = t;
  }
}

/* SinkhornKernel.jl:341-444 (and the two-pass 230-338):
 *   v = muladd(C_ij, -1/(reg W), psi_j);  out_i = log(marg_i) - LSE_j v */
void or_sinkhorn_half_step(const double *x, int64_t n, const double *y,
                           int64_t m, const double *marg, const double *psi,
                           double reg, double winf, double p, double *out) {
  const double invreg = -1.0 / (reg * winf);
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double mx = -INFINITY;
    for (int64_t j = 0; j < m; ++j) {
      double v = or_cost3(x + 3 * i, y + 3 * j, p) * invreg + psi[j];
      if (v > mx) mx = v;
    }
    double s = 0.0;
    if (isfinite(mx))
      for (int64_t j = 0; j < m; ++j) {
        double v = or_cost3(x + 3 * i, y + 3 * j, p) * invreg + psi[j];
        s += exp(v - mx);
      }
    out[i] = log(marg[i]) - lse_finish(mx, s);
  }
}

/* SinkhornKernel.jl:9-81: value = muladd(C, invreg, phi_i) + psi_j,
 *   res_i = marg_i - sum_j exp(value),  cost_i = sum_j exp(value) C */
void or_row_residual_cost(const double *x, int64_t n, const double *y,
                          int64_t m, const double *marg, const double *phi,
                          const double *psi, double reg, double winf, double p,
                          double *res, double *cost) {
  const double invreg = -1.0 / (reg * winf);
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double acc = 0.0, cacc = 0.0;
    for (int64_t j = 0; j < m; ++j) {
      double c = or_cost3(x + 3 * i, y + 3 * j, p);
      double e = exp((c * invreg + phi[i]) + psi[j]);
      acc += e;
      cacc += e * c;
    }
    res[i] = marg[i] - acc;
    if (cost) cost[i] = cacc;
  }
}

/* SinkhornKernel.jl:124-162: prob = exp(-C/(reg W) + phi_i + psi_j),
 * out_i = sum_j y_j prob / sum_j prob */
void or_barycentric_sinkhorn(const double *x, int64_t n, const double *y,
                             int64_t m, const double *phi, const double *psi,
                             double reg, double winf, double p, double *out) {
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double a0 = 0, a1 = 0, a2 = 0, ps = 0;
    for (int64_t j = 0; j < m; ++j) {
      double c = or_cost3(x + 3 * i, y + 3 * j, p);
      double prob = exp(-c / (reg * winf) + phi[i] + psi[j]);
      ps += prob;
      a0 += y[3 * j] * prob;
      a1 += y[3 * j + 1] * prob;
      a2 += y[3 * j + 2] * prob;
    }
    out[3 * i] = a0 / ps;
    out[3 * i + 1] = a1 / ps;
    out[3 * i + 2] = a2 / ps;
  }
}

/* MirrorProxKernel.jl:97-142: prob = exp(-(C/(2W) + nu_j)/reg - logZ_i) r_i */
void or_barycentric_lamp_rows(const double *x, int64_t n, const double *y,
                              int64_t m, const double *nu, const double *r,
                              const double *logZ, double reg, double st,
                              double winf, double p, double *out) {
  (void)st;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double a0 = 0, a1 = 0, a2 = 0, ps = 0;
    for (int64_t j = 0; j < m; ++j) {
      double c = or_cost3(x + 3 * i, y + 3 * j, p);
      double prob = exp(-(c / (2.0 * winf) + nu[j]) / reg - logZ[i]) * r[i];
      a0 += y[3 * j] * prob;
      a1 += y[3 * j + 1] * prob;
      a2 += y[3 * j + 2] * prob;
      ps += prob;
    }
    out[3 * i] = a0 / ps;
    out[3 * i + 1] = a1 / ps;
    out[3 * i + 2] = a2 / ps;
  }
}

/* MirrorProxKernel.jl:144-189: for column j, prob_i as above, mean of x. */
void or_barycentric_lamp_cols(const double *x, int64_t n, const double *y,
                              int64_t m, const double *r, const double *nu,
                              const double *logZ, double reg, double st,
                              double winf, double p, double *out) {
  (void)st;
  OMP_FOR
  for (int64_t j = 0; j < m; ++j) {
    double a0 = 0, a1 = 0, a2 = 0, ps = 0;
    for (int64_t i = 0; i < n; ++i) {
      double c = or_cost3(y + 3 * j, x + 3 * i, p);
      double prob = exp(-(c / (2.0 * winf) + nu[j]) / reg - logZ[i]) * r[i];
      a0 += x[3 * i] * prob;
      a1 += x[3 * i + 1] * prob;
      a2 += x[3 * i + 2] * prob;
      ps += prob;
    }
    out[3 * j] = a0 / ps;
    out[3 * j + 1] = a1 / ps;
    out[3 * j + 2] = a2 / ps;
  }
}

static double dotv(const double *a, const double *b, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* neg_entropy (src/Utilities.jl:137-143): sum y log y over y > 0 */
static double neg_entropy_sum(const double *a, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i)
    if (a[i] > 0.0) s += a[i] * log(a[i]);
  return s;
}

static double max_cost(const double *x, int64_t n, const double *y, int64_t m,
                       double p) {
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  or_max_cost_rows(x, n, y, m, p, tmp);
  double w = -INFINITY;
  for (int64_t i = 0; i < n; ++i)
    if (tmp[i] > w) w = tmp[i];
  free(tmp);
  return w;
}

static void emit(or_log_fn logfn, void *user, const char *line) {
  if (logfn) logfn(user, line);
}

/* extragradient_color_transfer (MirrorProxKernel.jl:1019-1132). */
int or_lamp_solve(const double *x, int64_t n, const double *y, int64_t m,
                  const double *r, const double *c, const or_args *args,
                  int64_t frequency, double p, or_log_fn logfn, void *user,
                  double *nu_out, double *phi_out, double *psi_out,
                  double *img1, double *img2, or_result *out) {
  if (n <= 0 || m <= 0 || frequency <= 0) return 1;
  double *theta = calloc((size_t)m, sizeof(double));
  double *nu = calloc((size_t)m, sizeof(double));
  double *thetab = calloc((size_t)m, sizeof(double));
  double *nub = calloc((size_t)m, sizeof(double));
  double *res = calloc((size_t)m, sizeof(double));
  double *cost = calloc((size_t)m, sizeof(double));
  double *etamu = calloc((size_t)m, sizeof(double));
  double *logZ = calloc((size_t)n, sizeof(double));
  /* 1032-1035 */
  const double winf = max_cost(x, n, y, m, p);
  const double eta = args->eta_p / (2.0 * winf);
  /* 1037: note alpha / N with N = #rows */
  for (int64_t j = 0; j < m; ++j)
    etamu[j] = (c[j] + args->alpha / (double)n) / args->tau_mu;
  const double t0 = now_s();
  const double hr = eta * neg_entropy_sum(r, n); /* 1066 */
  emit(logfn, user, "time(s),iter,infeas,ot_objective,primal,dual,solver\n");
  double st = 1.0, etat = INFINITY;
  const double taup = 1.0; /* hard-coded, 1073 */
  const double minv = tanh(-args->B / 2.0), maxv = tanh(args->B / 2.0);
  double primal = NAN, dual = NAN, resid = NAN, objective = NAN;
  int converged = 0;
  int64_t steps = 0;
  for (int64_t it = 1; it <= args->itermax; ++it) {
    double elapsed = now_s() - t0;
    if (elapsed > args->tmax) break;
    if ((args->verbose || args->stop == 1) && (it - 1) % frequency == 0) { /* 1081-1101 */
      or_lse_rows(x, n, y, m, nu, etat, st, winf, p, logZ);
      or_col_marginal_cost(x, n, y, m, r, nu, logZ, etat, st, winf, p, res, cost);
      resid = 0.0;
      for (int64_t j = 0; j < m; ++j) resid += fabs(res[j] - c[j]);
      objective = 0.0;
      for (int64_t j = 0; j < m; ++j) objective += cost[j];
      double q = eta / etat;
      primal = objective * (1.0 - q) - q * dotv(res, nu, m) +
               q * dotv(r, logZ, n) + 2.0 * winf * resid;
      if (eta > 0) {
        or_lse_rows(x, n, y, m, theta, eta, 1.0, winf, p, logZ);
        dual = -2.0 * eta * dotv(r, logZ, n) - 2.0 * winf * dotv(c, theta, m) +
               2.0 * hr;
      } else {
        or_minplus_rows(x, n, y, m, theta, winf, p, logZ);
        dual = dotv(r, logZ, n) - 2.0 * winf * dotv(c, theta, m);
      }
      char b0[64], b1[64], b2[64], b3[64], b4[64], line[512];
      fmt_e(b0, 64, "%.6e", elapsed);
      fmt_e(b1, 64, "%.14e", resid);
      fmt_e(b2, 64, "%.14e", objective);
      fmt_e(b3, 64, "%.14e", primal);
      fmt_e(b4, 64, "%.14e", dual);
      snprintf(line, sizeof line, "%s,%lld,%s,%s,%s,%s,lamp_kernel\n", b0,
               (long long)it, b1, b2, b3, b4);
      if (args->verbose) emit(logfn, user, line);
      /* stop == 1: the infeasibility alone, never at the pre-step check
       * (eta_t = inf: the plan is r (x) 1/m whatever the costs, exactly
       * feasible when c is uniform) */
      const int met = args->stop == 1 ? (isfinite(etat) && resid < args->epsilon / 6.0)
                                       : (primal - dual < args->epsilon / 6.0 && resid < args->epsilon / 6.0);
      if (met) {
        converged = 1;
        break;
      }
    }
    /* extragradient step, 1103-1115 */
    or_lse_rows(x, n, y, m, nu, etat, st, winf, p, logZ);
    or_col_marginal_cost(x, n, y, m, r, nu, logZ, etat, st, winf, p, res, cost);
    or_theta_update(thetab, theta, res, c, etamu, m, args->eta_mu, 0, minv, maxv, 1.0);
    etat = 1.0 / (taup + (1.0 / etat) * (1.0 - eta));
    for (int64_t j = 0; j < m; ++j) nub[j] = (1.0 - etat) * nu[j] + etat * theta[j];
    st = (1.0 - etat) * st + etat;
    or_lse_rows(x, n, y, m, nub, etat, st, winf, p, logZ);
    or_col_marginal_cost(x, n, y, m, r, nub, logZ, etat, st, winf, p, res, cost);
    or_theta_update(theta, theta, res, c, etamu, m, args->eta_mu, 1, minv, maxv, 1.0);
    for (int64_t j = 0; j < m; ++j) nu[j] = (1.0 - etat) * nu[j] + etat * thetab[j];
    steps += 1;
  }
  /* epilogue, 1124-1130 */
  or_lse_rows(x, n, y, m, nu, etat, st, winf, p, logZ);
  if (img1) or_barycentric_lamp_rows(x, n, y, m, nu, r, logZ, etat, st, winf, p, img1);
  if (img2) or_barycentric_lamp_cols(x, n, y, m, r, nu, logZ, etat, st, winf, p, img2);
  if (psi_out)
    for (int64_t j = 0; j < m; ++j) psi_out[j] = -2.0 * winf * nu[j] / args->eta_p;
  if (phi_out)
    for (int64_t i = 0; i < n; ++i) phi_out[i] = log(r[i]) - logZ[i];
  if (nu_out) memcpy(nu_out, nu, sizeof(double) * (size_t)m);
  if (out) {
    out->iterations = steps;
    out->objective = objective;
    out->residual = resid;
    out->primal = primal;
    out->dual = dual;
    out->winf = winf;
    out->eta_t = etat;
    out->converged = converged;
  }
  free(theta); free(nu); free(thetab); free(nub); free(res); free(cost);
  free(etamu); free(logZ);
  return 0;
}

/* sinkhorn_color_transfer (SinkhornKernel.jl:477-559).  p == Inf runs the
 * two-pass kernel, which evaluates the squared-L2 cost (SinkhornKernel.jl:
 * 259-263) against the Chebyshev W_inf; the residual and barycentric kernels
 * only know p == 1 or squared L2. */
int or_sinkhorn_solve(const double *x, int64_t n, const double *y, int64_t m,
                      const double *r, const double *c, const or_args *args,
                      int64_t frequency, double p, const double *phi0,
                      const double *psi0, or_log_fn logfn, void *user,
                      double *phi_out, double *psi_out, double *img1,
                      double *img2, or_result *out) {
  if (n <= 0 || m <= 0 || frequency <= 0) return 1;
  double *phi = calloc((size_t)n, sizeof(double));
  double *psi = calloc((size_t)m, sizeof(double));
  double *res = calloc((size_t)n, sizeof(double));
  double *cost = calloc((size_t)n, sizeof(double));
  if (phi0) memcpy(phi, phi0, sizeof(double) * (size_t)n);
  if (psi0) memcpy(psi, psi0, sizeof(double) * (size_t)m);
  const double t0 = now_s();
  const double eta = args->eta_p;
  const double winf = max_cost(x, n, y, m, p); /* 500-502 */
  const double p_lse = (p == INFINITY) ? sk_aux_p(p) : p;
  const double p_aux = sk_aux_p(p);
  if (args->verbose) emit(logfn, user, "time(s),iter,infeas,ot_objective,dual\n");
  int64_t num_iter = 0;
  int converged = 0;
  for (int64_t it = 1; it <= args->itermax; ++it) {
    double elapsed = now_s() - t0;
    if (elapsed > args->tmax) break;
    or_sinkhorn_half_step(x, n, y, m, r, psi, eta, winf, p_lse, phi);
    or_sinkhorn_half_step(y, m, x, n, c, phi, eta, winf, p_lse, psi);
    if ((it - 1) % frequency == 0) {
      or_row_residual_cost(x, n, y, m, r, phi, psi, eta, winf, p_aux, res, cost);
      double rr = 0.0;
      for (int64_t i = 0; i < n; ++i) rr += fabs(res[i]);
      if (args->verbose) {
        double ot = 0.0;
        for (int64_t i = 0; i < n; ++i) ot += cost[i];
        double obj = dotv(psi, c, m) + dotv(phi, r, n);
        char b0[64], b1[64], b2[64], b3[64], line[512];
        fmt_e(b0, 64, "%.6e", elapsed);
        fmt_e(b1, 64, "%.14e", rr);
        fmt_e(b2, 64, "%.14e", ot);
        fmt_e(b3, 64, "%.14e", obj);
        snprintf(line, sizeof line, "%s,%lld,%s,%s,%s,sinkhorn_kernel\n", b0,
                 (long long)it, b1, b2, b3);
        emit(logfn, user, line);
      }
      if (rr <= args->epsilon / 6.0) {
        converged = 1;
        break;
      }
    }
    num_iter += 1;
  }
  or_row_residual_cost(x, n, y, m, r, phi, psi, eta, winf, p_aux, res, cost);
  double rv = 0.0, ob = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    rv += fabs(res[i]);
    ob += cost[i];
  }
  if (img1) or_barycentric_sinkhorn(x, n, y, m, phi, psi, eta, winf, p_aux, img1);
  if (img2) or_barycentric_sinkhorn(y, m, x, n, psi, phi, eta, winf, p_aux, img2);
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)n);
  if (psi_out) memcpy(psi_out, psi, sizeof(double) * (size_t)m);
  if (out) {
    out->iterations = num_iter;
    out->objective = ob;
    out->residual = rv;
    out->primal = NAN;
    out->dual = dotv(psi, c, m) + dotv(phi, r, n);
    out->winf = winf;
    out->eta_t = eta;
    out->converged = converged;
  }
  free(phi); free(psi); free(res); free(cost);
  return 0;
}

/* annealed_sinkhorn_euclidean (AnnealedSinkhorn.jl:65-177) */
int or_annealed_sinkhorn_solve(const double *x, int64_t n, const double *y,
                               int64_t m, const double *r, const double *c,
                               const or_args *args_in, int64_t frequency,
                               double p, or_log_fn logfn, void *user,
                               double *phi_out, double *psi_out, double *img1,
                               double *img2, or_result *out) {
  if (n <= 0 || m <= 0 || frequency <= 0) return 1;
  or_args args = *args_in;
  const double eta_final = args.epsilon;
  double eta = fmax(args.eta_p, eta_final);
  const double t0 = now_s();
  double *phi = malloc(sizeof(double) * (size_t)n);
  double *psi = malloc(sizeof(double) * (size_t)m);
  double *phiws = malloc(sizeof(double) * (size_t)n);
  double *psiws = malloc(sizeof(double) * (size_t)m);
  double *phiprev = malloc(sizeof(double) * (size_t)n);
  double *psiprev = malloc(sizeof(double) * (size_t)m);
  for (int64_t i = 0; i < n; ++i) phi[i] = phiws[i] = phiprev[i] = log(r[i]);
  for (int64_t j = 0; j < m; ++j) psi[j] = psiws[j] = psiprev[j] = log(c[j]);
  double objective = NAN, residual = NAN;
  const double q = args.anneal_mult;
  double preveta = 1.0;
  const int verbose = args.verbose;
  args.verbose = 0;
  int64_t total = 0;
  const double tmax = args.tmax;
  if (verbose) emit(logfn, user, "time(s),iter,infeas,ot_objective,eta,solver\n");
  for (;;) {
    args.eta_p = eta;
    or_result rs;
    or_sinkhorn_solve(x, n, y, m, r, c, &args, frequency, p, phiws, psiws, NULL,
                      NULL, phi, psi, NULL, NULL, &rs);
    total += rs.iterations;
    objective = rs.objective;
    residual = rs.residual;
    double elapsed = now_s() - t0;
    args.tmax = tmax - elapsed;
    if (verbose) {
      char b0[64], b1[64], b2[64], b3[64], line[512];
      fmt_e(b0, 64, "%.6e", elapsed);
      fmt_e(b1, 64, "%.14e", residual);
      fmt_e(b2, 64, "%.14e", objective);
      fmt_e(b3, 64, "%.14e", eta);
      snprintf(line, sizeof line, "%s,%lld,%s,%s,%s,annealed_sinkhorn_kernel\n",
               b0, (long long)total, b1, b2, b3);
      emit(logfn, user, line);
    }
    if (eta == eta_final || elapsed >= tmax) break;
    for (int64_t i = 0; i < n; ++i) {
      phiws[i] = phi[i] + (eta / preveta) * (phi[i] - phiprev[i]);
      phiprev[i] = phi[i];
    }
    for (int64_t j = 0; j < m; ++j) {
      psiws[j] = psi[j] + (eta / preveta) * (psi[j] - psiprev[j]);
      psiprev[j] = psi[j];
    }
    preveta = eta;
    eta = fmax(q * eta, eta_final);
  }
  const double winf = max_cost(x, n, y, m, p);
  const double p_aux = sk_aux_p(p);
  if (img1) or_barycentric_sinkhorn(x, n, y, m, phi, psi, eta_final, winf, p_aux, img1);
  if (img2) or_barycentric_sinkhorn(y, m, x, n, psi, phi, eta_final, winf, p_aux, img2);
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)n);
  if (psi_out) memcpy(psi_out, psi, sizeof(double) * (size_t)m);
  if (out) {
    out->iterations = total;
    out->objective = objective;
    out->residual = residual;
    out->primal = NAN;
    out->dual = NAN;
    out->winf = winf;
    out->eta_t = eta;
    out->converged = residual <= eta_final / 6.0;
  }
  free(phi); free(psi); free(phiws); free(psiws); free(phiprev); free(psiprev);
  return 0;
}

/* logsumexp(x, dims) of src/Utilities.jl:34-41 on rows / columns of a dense
 * n x m row-major matrix K with additive row/column terms. */
static void lse_cols_dense(const double *K, int64_t n, int64_t m,
                           const double *rowterm, double *out) {
  /* out_j = LSE_i (K_ij + rowterm_i)  (dims = 1); columns in parallel, each
   * summed over i in order */
  OMP_FOR
  for (int64_t j = 0; j < m; ++j) {
    double mx = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
      double v = K[i * m + j] + rowterm[i];
      if (v > mx) mx = v;
    }
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += exp(K[i * m + j] + rowterm[i] - mx);
    out[j] = lse_finish(mx, s);
  }
}

static void lse_rows_dense(const double *K, int64_t n, int64_t m,
                           const double *colterm, double *out) {
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double mx = -INFINITY;
    for (int64_t j = 0; j < m; ++j) {
      double v = K[i * m + j] + colterm[j];
      if (v > mx) mx = v;
    }
    double s = 0.0;
    for (int64_t j = 0; j < m; ++j) s += exp(K[i * m + j] + colterm[j] - mx);
    out[i] = lse_finish(mx, s);
  }
}

/* sinkhorn_log (src/solvers/Sinkhorn.jl:8-84) on a dense cost. */
int or_sinkhorn_log_dense(const double *r, int64_t n, const double *c,
                          int64_t m, const double *W, const or_args *args,
                          int64_t frequency, or_log_fn logfn, void *user,
                          double *phi_out, double *psi_out, double *P_out,
                          or_result *out) {
  if (n <= 0 || m <= 0 || frequency <= 0) return 1;
  const size_t nm = (size_t)n * (size_t)m;
  double *K = malloc(sizeof(double) * nm);
  for (size_t k = 0; k < nm; ++k) K[k] = -W[k] / args->eta_p; /* 19 */
  double *phi = calloc((size_t)n, sizeof(double));
  double *psi = calloc((size_t)m, sizeof(double));
  double *tmp_m = malloc(sizeof(double) * (size_t)m);
  double *tmp_n = malloc(sizeof(double) * (size_t)n);
  double *colsum = malloc(sizeof(double) * (size_t)m);
  if (args->verbose)
    emit(logfn, user, "time(s),iter,infeas,ot_objective,primal,dual,solver\n");
  const double t0 = now_s();
  int64_t num_iter = 0;
  int converged = 0;
  double feas = NAN, obj = NAN, pobj = NAN, dobj = NAN;
  for (int64_t it = 1; it <= args->itermax; ++it) {
    lse_cols_dense(K, n, m, phi, tmp_m); /* 44 */
    for (int64_t j = 0; j < m; ++j) psi[j] = log(c[j]) - tmp_m[j];
    lse_rows_dense(K, n, m, psi, tmp_n); /* 45 */
    for (int64_t i = 0; i < n; ++i) phi[i] = log(r[i]) - tmp_n[i];
    double elapsed = now_s() - t0;
    if (elapsed > args->tmax) break;
    if ((it - 1) % frequency == 0) { /* 55-71 */
      double rowl1 = 0.0, ent = 0.0, mx = -INFINITY;
      obj = 0.0;
      for (int64_t j = 0; j < m; ++j) colsum[j] = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        double rs = 0.0;
        for (int64_t j = 0; j < m; ++j) {
          double v = K[i * m + j] + phi[i] + psi[j];
          double pij = exp(v);
          rs += pij;
          colsum[j] += pij;
          obj += pij * W[i * m + j];
          if (pij > 0.0) ent += pij * log(pij);
          if (v > mx) mx = v;
        }
        rowl1 += fabs(rs - r[i]);
      }
      double coll1 = 0.0;
      for (int64_t j = 0; j < m; ++j) coll1 += fabs(colsum[j] - c[j]);
      feas = coll1 + rowl1;
      pobj = obj + args->eta_p * ent;
      double s = 0.0;
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < m; ++j) s += exp(K[i * m + j] + phi[i] + psi[j] - mx);
      double lse_all = lse_finish(mx, s);
      dobj = -args->eta_p * (-lse_all - dotv(c, psi, m) - dotv(r, phi, n));
      if (args->verbose) {
        char b1[64], b2[64], b3[64], b4[64], line[512];
        fmt_e(b1, 64, "%.14e", feas);
        fmt_e(b2, 64, "%.14e", obj);
        fmt_e(b3, 64, "%.14e", pobj);
        fmt_e(b4, 64, "%.14e", dobj);
        snprintf(line, sizeof line, "%.6g,%lld,%s,%s,%s,%s,sinkhorn\n", elapsed,
                 (long long)it, b1, b2, b3, b4);
        emit(logfn, user, line);
      }
      if (pobj - dobj < args->epsilon / 6.0 && feas < args->epsilon / 6.0) {
        converged = 1;
        break;
      }
    }
    num_iter += 1;
  }
  /* 74-76 */
  lse_cols_dense(K, n, m, phi, tmp_m);
  for (int64_t j = 0; j < m; ++j) psi[j] = log(c[j]) - tmp_m[j];
  lse_rows_dense(K, n, m, psi, tmp_n);
  for (int64_t i = 0; i < n; ++i) phi[i] = log(r[i]) - tmp_n[i];
  double objf = 0.0;
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < m; ++j) {
      double pij = exp(K[i * m + j] + phi[i] + psi[j]);
      objf += pij * W[i * m + j];
      if (P_out) P_out[i * m + j] = pij;
    }
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)n);
  if (psi_out) memcpy(psi_out, psi, sizeof(double) * (size_t)m);
  if (out) {
    out->iterations = num_iter;
    out->objective = objf;
    out->residual = feas;
    out->primal = pobj;
    out->dual = dobj;
    out->winf = NAN;
    out->eta_t = args->eta_p;
    out->converged = converged;
  }
  free(K); free(phi); free(psi); free(tmp_m); free(tmp_n); free(colsum);
  return 0;
}

/* get_euclidean_distance (src/Utilities.jl:145-157): grid index i -> (i / h,
 * i % h); p >= 10 is treated as the Chebyshev distance. */
void or_grid_cost_matrix(int64_t height, int64_t width, double p, double *W) {
  const int64_t N = height * width;
  for (int64_t i = 0; i < N; ++i)
    for (int64_t j = 0; j < N; ++j) {
      double a = fabs((double)(i / height - j / height));
      double b = fabs((double)(i % height - j % height));
      W[i * N + j] = p < 10 ? pow(a, p) + pow(b, p) : fmax(a, b);
    }
}

/* ===================== feature (cell-similarity) costs ===================== */

/* get_flat_metrics (src/CellUtilities.jl:304-311).  x is n x d column-major. */
void or_feature_stats(const double *x, int64_t n, int64_t d, double *sums,
                      double *sqnorms, double *means, double *csqnorms) {
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0, q = 0.0;
    for (int64_t k = 0; k < d; ++k) {
      double v = x[i + k * n];
      s += v;
      q += v * v;
    }
    double mu = s / (double)d, cs = 0.0;
    for (int64_t k = 0; k < d; ++k) {
      double v = x[i + k * n] - mu;
      cs += v * v;
    }
    sums[i] = s;
    sqnorms[i] = q;
    means[i] = mu;
    csqnorms[i] = cs;
  }
}

static double clamp1(double v) { return v < -1.0 ? -1.0 : (v > 1.0 ? 1.0 : v); }

/* feature_cost_cuda and the four metric functions (src/CellUtilities.jl:
 * 187-301).  xs / ys hold [sums | sqnorms | means | csqnorms] (4 x n). */
double or_feature_cost(int metric, const double *x, int64_t n, int64_t i,
                       const double *xs, const double *y, int64_t m, int64_t j,
                       const double *ys, int64_t d, double scale) {
  const double si = xs[i], sj = ys[j];
  if (metric == OR_L1) { /* 211-226 */
    double dist = 0.0;
    for (int64_t k = 0; k < d; ++k) dist += fabs(x[i + k * n] / si - y[j + k * m] / sj);
    return dist / scale;
  }
  double raw = 0.0;
  for (int64_t k = 0; k < d; ++k) raw += x[i + k * n] * y[j + k * m];
  const double dot = raw / (si * sj);
  if (metric == OR_L2) { /* 230-243 */
    double ni = xs[n + i] / (si * si), nj = ys[m + j] / (sj * sj);
    return sqrt(fmax(ni + nj - 2.0 * dot, 0.0)) / scale;
  }
  if (metric == OR_COSINE) { /* 245-259 */
    double denom = sqrt((xs[n + i] / (si * si)) * (ys[m + j] / (sj * sj)));
    if (!(denom > 0.0)) return 0.0;
    return (1.0 - clamp1(dot / denom)) / scale;
  }
  if (metric == OR_CORRELATION) { /* 261-282 */
    double mi = xs[2 * n + i] / si, mj = ys[2 * m + j] / sj;
    double ci = xs[3 * n + i] / (si * si), cj = ys[3 * m + j] / (sj * sj);
    double denom = sqrt(ci * cj);
    if (denom <= 0.0) return fabs(mi - mj) <= sqrt(2.220446049250313e-16) ? 0.0 : 1.0 / scale;
    return (1.0 - clamp1((dot - (double)d * mi * mj) / denom)) / scale;
  }
  return INFINITY; /* 300 */
}

typedef struct {
  const double *x, *y, *xs, *ys;
  int64_t n, m, d;
  int metric;
  double scale;
} fctx;

static double fcost(const fctx *f, int64_t i, int64_t j) {
  return or_feature_cost(f->metric, f->x, f->n, i, f->xs, f->y, f->m, j, f->ys, f->d, f->scale);
}

static double *stats_alloc(const double *x, int64_t n, int64_t d) {
  double *s = malloc(sizeof(double) * 4 * (size_t)n);
  or_feature_stats(x, n, d, s, s + n, s + 2 * n, s + 3 * n);
  return s;
}

void or_feature_cost_matrix(int metric, const double *x, int64_t n,
                            const double *y, int64_t m, int64_t d, double scale,
                            double *C) {
  double *xs = stats_alloc(x, n, d), *ys = stats_alloc(y, m, d);
  fctx f = {x, y, xs, ys, n, m, d, metric, scale};
  OMP_FOR
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < m; ++j) C[i * m + j] = fcost(&f, i, j);
  free(xs);
  free(ys);
}

/* Dense-cost helpers over a precomputed C (n x m row-major); the feature
 * kernels evaluate the same C_ij per pair. */
static void d_lse_rows(const double *C, int64_t n, int64_t m, const double *theta,
                       double reg, double st, double winf, double *out) {
  const double c1 = st / (2.0 * winf), invreg = 1.0 / reg;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double mx = -INFINITY;
    for (int64_t j = 0; j < m; ++j) {
      double v = -(C[i * m + j] * c1 + theta[j]) * invreg;
      if (v > mx) mx = v;
    }
    double s = 0.0;
    if (isfinite(mx))
      for (int64_t j = 0; j < m; ++j) s += exp(-(C[i * m + j] * c1 + theta[j]) * invreg - mx);
    out[i] = lse_finish(mx, s);
  }
}

static void d_colmarg(const double *C, int64_t n, int64_t m, const double *r,
                      const double *theta, const double *logZ, double reg,
                      double st, double winf, double *res, double *cost) {
  const double c1 = st / (2.0 * winf), invreg = 1.0 / reg;
  OMP_FOR
  for (int64_t j = 0; j < m; ++j) {
    double a = 0.0, ca = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double c = C[i * m + j];
      double w = exp((c * c1 + theta[j]) * (-invreg) - logZ[i]) * r[i];
      a += w;
      ca += w * c;
    }
    res[j] = a;
    cost[j] = ca;
  }
}

static void d_minplus(const double *C, int64_t n, int64_t m, const double *theta,
                      double winf, double *out) {
  for (int64_t i = 0; i < n; ++i) {
    double mn = INFINITY;
    for (int64_t j = 0; j < m; ++j) {
      double v = theta[j] * (2.0 * winf) + C[i * m + j];
      if (v < mn) mn = v;
    }
    out[i] = mn;
  }
}

/* Sinkhorn half-steps on dense C: rows (phi) and columns (psi). */
static void d_sk_rows(const double *C, int64_t n, int64_t m, const double *marg,
                      const double *psi, double reg, double winf, double *out) {
  const double k = -1.0 / (reg * winf);
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double mx = -INFINITY;
    for (int64_t j = 0; j < m; ++j) {
      double v = C[i * m + j] * k + psi[j];
      if (v > mx) mx = v;
    }
    double s = 0.0;
    if (isfinite(mx))
      for (int64_t j = 0; j < m; ++j) s += exp(C[i * m + j] * k + psi[j] - mx);
    out[i] = log(marg[i]) - lse_finish(mx, s);
  }
}

static void d_sk_cols(const double *C, int64_t n, int64_t m, const double *marg,
                      const double *phi, double reg, double winf, double *out) {
  const double k = -1.0 / (reg * winf);
  OMP_FOR
  for (int64_t j = 0; j < m; ++j) {
    double mx = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
      double v = C[i * m + j] * k + phi[i];
      if (v > mx) mx = v;
    }
    double s = 0.0;
    if (isfinite(mx))
      for (int64_t i = 0; i < n; ++i) s += exp(C[i * m + j] * k + phi[i] - mx);
    out[j] = log(marg[j]) - lse_finish(mx, s);
  }
}

static void d_sk_resid(const double *C, int64_t n, int64_t m, const double *marg,
                       const double *phi, const double *psi, double reg,
                       double winf, double *res, double *cost) {
  const double k = -1.0 / (reg * winf);
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double a = 0.0, ca = 0.0;
    for (int64_t j = 0; j < m; ++j) {
      double c = C[i * m + j];
      double w = exp((c * k + phi[i]) + psi[j]);
      a += w;
      ca += w * c;
    }
    res[i] = marg[i] - a;
    cost[i] = ca;
  }
}

/* max_logsumexp_cellsim! (MirrorProxCellSim.jl:1-41): running max starts at 0 */
static double d_winf(const double *C, int64_t n, int64_t m) {
  double w = -INFINITY;
  for (int64_t i = 0; i < n; ++i) {
    double mx = 0.0;
    for (int64_t j = 0; j < m; ++j) mx = fmax(mx, C[i * m + j]);
    w = fmax(w, mx);
  }
  return w;
}

/* extragradient_cellsim (src/solvers/MirrorProxCellSim.jl:94-188) */
int or_lamp_cellsim_solve(const double *data, int64_t n, int64_t d,
                          double scale, int metric, const double *r,
                          const double *c, const or_args *args,
                          int64_t frequency, or_log_fn logfn, void *user,
                          double *nu_out, double *phi_out, double *psi_out,
                          or_result *out) {
  return or_lamp_cellsim2_solve(data, n, data, n, d, scale, metric, r, c, args, frequency,
                                logfn, user, nu_out, phi_out, psi_out, out);
}

/* The same loop with the cost between row i of data1 (n x d) and row j of
 * data2 (m x d); the reference calls it with one matrix (data1 == data2). */
int or_lamp_cellsim2_solve(const double *data1, int64_t n, const double *data2,
                           int64_t m, int64_t d, double scale, int metric,
                           const double *r, const double *c, const or_args *args,
                           int64_t frequency, or_log_fn logfn, void *user,
                           double *nu_out, double *phi_out, double *psi_out,
                           or_result *out) {
  if (n <= 0 || m <= 0 || d <= 0 || frequency <= 0) return 1;
  double *C = malloc(sizeof(double) * (size_t)n * (size_t)m);
  or_feature_cost_matrix(metric, data1, n, data2, m, d, scale, C);
  double *theta = calloc((size_t)m, sizeof(double)), *nu = calloc((size_t)m, sizeof(double));
  double *thetab = calloc((size_t)m, sizeof(double)), *nub = calloc((size_t)m, sizeof(double));
  double *gam = calloc((size_t)m, sizeof(double)), *res = calloc((size_t)m, sizeof(double));
  double *cost = calloc((size_t)m, sizeof(double)), *etamu = calloc((size_t)m, sizeof(double));
  double *sumvals = calloc((size_t)n, sizeof(double));
  const double winf = d_winf(C, n, m); /* 115-117 */
  const double eta = args->eta_p;       /* 118: not divided by 2 W */
  for (int64_t j = 0; j < m; ++j) etamu[j] = (c[j] + args->alpha / (double)n) / args->tau_mu;
  const double t0 = now_s();
  const double hr = eta * neg_entropy_sum(r, n);
  emit(logfn, user, "time(s),iter,infeas,ot_objective,primal,dual,solver\n");
  double st = 1.0, etat = INFINITY;
  const double minv = tanh(-args->B / 2.0), maxv = tanh(args->B / 2.0);
  double primal = NAN, dual = NAN, resid = NAN, objective = NAN;
  int converged = 0;
  int64_t steps = 0;
  for (int64_t it = 1; it <= args->itermax; ++it) {
    double elapsed = now_s() - t0;
    if (elapsed > args->tmax) break;
    if ((args->verbose || args->stop == 1) && (it - 1) % frequency == 0) { /* 144-166 */
      d_lse_rows(C, n, m, nu, etat, st, winf, sumvals);
      d_colmarg(C, n, m, r, nu, sumvals, etat, st, winf, res, cost);
      resid = 0.0;
      objective = 0.0;
      for (int64_t j = 0; j < m; ++j) {
        resid += fabs(res[j] - c[j]);
        objective += cost[j];
      }
      if (eta > 0) {
        primal = objective * (1.0 - eta / etat) - eta * dotv(r, sumvals, n) -
                 2.0 * winf * eta / etat * dotv(c, nu, m) + hr + 2.0 * winf * resid;
        d_lse_rows(C, n, m, theta, eta, 1.0, winf, sumvals);
        dual = -eta * dotv(r, sumvals, n) - 2.0 * winf * dotv(c, theta, m) + hr;
      } else {
        primal = objective + 2.0 * winf * resid;
        d_minplus(C, n, m, gam, winf, sumvals);
        dual = dotv(r, sumvals, n) - 2.0 * winf * dotv(c, gam, m);
      }
      if (args->verbose) {
        char b0[64], b1[64], b2[64], b3[64], b4[64], line[512];
        fmt_e(b0, 64, "%.6e", elapsed);
        fmt_e(b1, 64, "%.14e", resid);
        fmt_e(b2, 64, "%.14e", objective);
        fmt_e(b3, 64, "%.14e", primal);
        fmt_e(b4, 64, "%.14e", dual);
        snprintf(line, sizeof line, "%s,%lld,%s,%s,%s,%s,lamp_kernel\n", b0, (long long)it,
                 b1, b2, b3, b4);
        emit(logfn, user, line);
      }
      /* stop == 1: the infeasibility alone, never at the pre-step check
       * (eta_t = inf: the plan is r (x) 1/m whatever the costs, exactly
       * feasible when c is uniform) */
      const int met = args->stop == 1 ? (isfinite(etat) && resid < args->epsilon / 6.0)
                                       : (primal - dual < args->epsilon / 6.0 && resid < args->epsilon / 6.0);
      if (met) {
        converged = 1;
        break;
      }
    }
    /* 168-180 */
    d_lse_rows(C, n, m, nu, etat, st, winf, sumvals);
    d_colmarg(C, n, m, r, nu, sumvals, etat, st, winf, res, cost);
    or_theta_update(thetab, theta, res, c, etamu, m, args->eta_mu, 0, minv, maxv, 1.0);
    etat = 1.0 / (1.0 + (1.0 / etat) * (1.0 - eta));
    for (int64_t j = 0; j < m; ++j) nub[j] = (1.0 - etat) * nu[j] + etat * theta[j];
    st = (1.0 - etat) * st + etat;
    d_lse_rows(C, n, m, nub, etat, st, winf, sumvals);
    d_colmarg(C, n, m, r, nub, sumvals, etat, st, winf, res, cost);
    or_theta_update(theta, theta, res, c, etamu, m, args->eta_mu, 1, minv, maxv, 1.0);
    for (int64_t j = 0; j < m; ++j) {
      nu[j] = (1.0 - etat) * nu[j] + etat * thetab[j];
      gam[j] = (1.0 - etat) * gam[j] + etat * theta[j];
    }
    steps += 1;
  }
  /* 184-187: phi uses whatever sumvals holds last */
  double objf = 0.0;
  for (int64_t j = 0; j < m; ++j) objf += cost[j];
  if (psi_out)
    for (int64_t j = 0; j < m; ++j) psi_out[j] = -2.0 * winf * nu[j] / args->eta_p;
  if (phi_out)
    for (int64_t i = 0; i < n; ++i) phi_out[i] = log(r[i]) - sumvals[i];
  if (nu_out) memcpy(nu_out, nu, sizeof(double) * (size_t)m);
  if (out) {
    out->iterations = steps;
    out->objective = objf;
    out->residual = resid;
    out->primal = primal;
    out->dual = dual;
    out->winf = winf;
    out->eta_t = etat;
    out->converged = converged;
  }
  free(C); free(theta); free(nu); free(thetab); free(nub); free(gam); free(res);
  free(cost); free(etamu); free(sumvals);
  return 0;
}

/* sinkhorn_cellsim (src/solvers/SinkhornCellSim.jl:240-310) with the cost
 * C_ij between row i of data1 and row j of data2 (equal to the reference's
 * per-call cost when data1 == data2, its only call pattern). */
static int sk_cellsim_core(const double *C, int64_t n, int64_t m, const double *r,
                           const double *c, const or_args *args, int64_t frequency,
                           double *phi, double *psi, double *res, double *cost,
                           or_log_fn logfn, void *user, or_result *out) {
  const double t0 = now_s();
  const double winf = d_winf(C, n, m); /* 269-270 */
  const double eta = args->eta_p;
  int64_t num_iter = 0;
  int converged = 0;
  if (args->verbose) emit(logfn, user, "time(s),iter,infeas,ot_objective,dual\n");
  for (int64_t it = 1; it <= args->itermax; ++it) {
    double elapsed = now_s() - t0;
    if (elapsed > args->tmax) break;
    d_sk_rows(C, n, m, r, psi, eta, winf, phi);
    d_sk_cols(C, n, m, c, phi, eta, winf, psi);
    if ((it - 1) % frequency == 0) {
      d_sk_resid(C, n, m, r, phi, psi, eta, winf, res, cost);
      double rr = 0.0;
      for (int64_t i = 0; i < n; ++i) rr += fabs(res[i]);
      if (args->verbose) {
        double ot = 0.0;
        for (int64_t i = 0; i < n; ++i) ot += cost[i];
        char b0[64], b1[64], b2[64], b3[64], line[512];
        fmt_e(b0, 64, "%.6e", elapsed);
        fmt_e(b1, 64, "%.14e", rr);
        fmt_e(b2, 64, "%.14e", ot);
        fmt_e(b3, 64, "%.14e", dotv(psi, c, m) + dotv(phi, r, n));
        snprintf(line, sizeof line, "%s,%lld,%s,%s,%s,sinkhorn_cellsim\n", b0, (long long)it,
                 b1, b2, b3);
        emit(logfn, user, line);
      }
      if (rr <= args->epsilon / 6.0) {
        converged = 1;
        break;
      }
    }
    num_iter += 1;
  }
  d_sk_resid(C, n, m, r, phi, psi, eta, winf, res, cost);
  double rv = 0.0, ob = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    rv += fabs(res[i]);
    ob += cost[i];
  }
  if (out) {
    out->iterations = num_iter;
    out->objective = ob;
    out->residual = rv;
    out->primal = NAN;
    out->dual = dotv(psi, c, m) + dotv(phi, r, n);
    out->winf = winf;
    out->eta_t = eta;
    out->converged = converged;
  }
  return 0;
}

int or_sinkhorn_cellsim_solve(const double *data1, int64_t n,
                              const double *data2, int64_t m, int64_t d,
                              double scale, int metric, const double *r,
                              const double *c, const or_args *args,
                              int64_t frequency, const double *phi0,
                              const double *psi0, or_log_fn logfn, void *user,
                              double *phi_out, double *psi_out, or_result *out) {
  if (n <= 0 || m <= 0 || d <= 0 || frequency <= 0) return 1;
  double *C = malloc(sizeof(double) * (size_t)n * (size_t)m);
  or_feature_cost_matrix(metric, data1, n, data2, m, d, scale, C);
  double *phi = calloc((size_t)n, sizeof(double)), *psi = calloc((size_t)m, sizeof(double));
  double *res = calloc((size_t)n, sizeof(double)), *cost = calloc((size_t)n, sizeof(double));
  if (phi0) memcpy(phi, phi0, sizeof(double) * (size_t)n);
  if (psi0) memcpy(psi, psi0, sizeof(double) * (size_t)m);
  sk_cellsim_core(C, n, m, r, c, args, frequency, phi, psi, res, cost, logfn, user, out);
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)n);
  if (psi_out) memcpy(psi_out, psi, sizeof(double) * (size_t)m);
  free(C); free(phi); free(psi); free(res); free(cost);
  return 0;
}

/* annealed_sinkhorn_cellsim (src/solvers/AnnealedSinkhornCellSim.jl:1-63):
 * no tmax test between stages, unlike the point-cloud version. */
int or_annealed_sinkhorn_cellsim_solve(const double *data1, int64_t n,
                                       const double *data2, int64_t m,
                                       int64_t d, double scale, int metric,
                                       const double *r, const double *c,
                                       const or_args *args_in, int64_t frequency,
                                       or_log_fn logfn, void *user, double *phi_out,
                                       double *psi_out, or_result *out) {
  if (n <= 0 || m <= 0 || d <= 0 || frequency <= 0) return 1;
  double *C = malloc(sizeof(double) * (size_t)n * (size_t)m);
  or_feature_cost_matrix(metric, data1, n, data2, m, d, scale, C);
  or_args args = *args_in;
  const double eta_final = args.epsilon;
  double eta = fmax(args.eta_p, eta_final);
  const double t0 = now_s();
  double *phi = malloc(sizeof(double) * (size_t)n), *psi = malloc(sizeof(double) * (size_t)m);
  double *phiprev = malloc(sizeof(double) * (size_t)n), *psiprev = malloc(sizeof(double) * (size_t)m);
  double *res = calloc((size_t)n, sizeof(double)), *cost = calloc((size_t)n, sizeof(double));
  for (int64_t i = 0; i < n; ++i) phi[i] = phiprev[i] = log(r[i]);
  for (int64_t j = 0; j < m; ++j) psi[j] = psiprev[j] = log(c[j]);
  const double q = args.anneal_mult;
  double preveta = 1.0, objective = NAN, residual = NAN;
  const int verbose = args.verbose;
  args.verbose = 0;
  int64_t total = 0;
  if (verbose) emit(logfn, user, "time(s),iter,infeas,ot_objective,eta,solver\n");
  or_result rs;
  for (;;) {
    args.eta_p = eta;
    /* phi, psi hold the warm start phi_ws, psi_ws and are updated in place */
    sk_cellsim_core(C, n, m, r, c, &args, frequency, phi, psi, res, cost, NULL, NULL, &rs);
    total += rs.iterations;
    objective = rs.objective;
    residual = rs.residual;
    if (verbose) {
      char b0[64], b1[64], b2[64], b3[64], line[512];
      fmt_e(b0, 64, "%.6e", now_s() - t0);
      fmt_e(b1, 64, "%.14e", residual);
      fmt_e(b2, 64, "%.14e", objective);
      fmt_e(b3, 64, "%.14e", eta);
      snprintf(line, sizeof line, "%s,%lld,%s,%s,%s,annealed_sinkhorn_cellsim\n", b0,
               (long long)total, b1, b2, b3);
      emit(logfn, user, line);
    }
    if (eta == eta_final) break;
    /* 49-52: sinkhorn_cellsim updates the warm-start arrays in place
     * (SinkhornCellSim.jl:257-259 binds phi = phi0), so after
     * `phi_ws .= phi + f (phi - phi_prev)` the assignment `phi_prev .= phi`
     * copies the extrapolated warm start, not the stage solution. */
    for (int64_t i = 0; i < n; ++i) {
      phi[i] = phi[i] + (eta / preveta) * (phi[i] - phiprev[i]);
      phiprev[i] = phi[i];
    }
    for (int64_t j = 0; j < m; ++j) {
      psi[j] = psi[j] + (eta / preveta) * (psi[j] - psiprev[j]);
      psiprev[j] = psi[j];
    }
    preveta = eta;
    eta = fmax(q * eta, eta_final);
  }
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)n);
  if (psi_out) memcpy(psi_out, psi, sizeof(double) * (size_t)m);
  if (out) {
    *out = rs;
    out->iterations = total;
    out->eta_t = eta;
  }
  free(C); free(phi); free(psi); free(phiprev); free(psiprev); free(res); free(cost);
  return 0;
}

/* ---- the reference's CPU path: LAMP on a dense cost matrix ----------------
 * LAMP(r, c, W::AbstractMatrix, args, frequency) (src/solvers/MirrorProx.jl:
 * 506-575), FP64, W n x n row-major (W[i n + j], the reference's W[i, j]).
 * Row reductions run in parallel over rows and column sums over blocks of
 * columns, each summing over the rows in order, so the results do not depend
 * on the thread count. */

/* rows of v_ij = -(W_ij / 2 W_inf + th_j) / etat: max and log sum exp(v - max) */
static void cl_rows(const double *W, int64_t n, const double *th, double winf, double etat,
                    double *mx, double *ls) {
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    const double *w = W + i * n;
    double m = -INFINITY;
    for (int64_t j = 0; j < n; ++j) {
      const double v = -(w[j] / (2.0 * winf) + th[j]) / etat;
      if (v > m) m = v;
    }
    double s = 0.0;
    for (int64_t j = 0; j < n; ++j) s += exp(-(w[j] / (2.0 * winf) + th[j]) / etat - m);
    mx[i] = m;
    ls[i] = log(s);
  }
}

/* infeas(theta, eta_t) (521-526): column sums of r_i softmax_j(v_ij) */
static void cl_infeas(const double *W, int64_t n, const double *r, const double *th,
                      double winf, double etat, double *mx, double *ls, double *res) {
  cl_rows(W, n, th, winf, etat, mx, ls);
  const int64_t BS = 64;
  const int64_t nb = (n + BS - 1) / BS;
  OMP_FOR
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t j0 = b * BS, j1 = j0 + BS < n ? j0 + BS : n;
    for (int64_t j = j0; j < j1; ++j) res[j] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double *w = W + i * n;
      const double off = -mx[i] - ls[i] + log(r[i]);
      for (int64_t j = j0; j < j1; ++j) res[j] += exp(-(w[j] / (2.0 * winf) + th[j]) / etat + off);
    }
  }
}

/* theta <- tanh((res - c) ./ eta_mu + (1 - tau_mu eta_mu) / 2 log((th0 + 1) / (1 - th0))) (541, 545) */
static void cl_theta(double *out, const double *th0, const double *res, const double *c,
                     const double *etamu, int64_t n, double tau_mu, double eta_mu) {
  for (int64_t j = 0; j < n; ++j)
    out[j] = tanh((res[j] - c[j]) / etamu[j] +
                  0.5 * (1.0 - tau_mu * eta_mu) * log((th0[j] + 1.0) / (1.0 - th0[j])));
}

int or_cpu_lamp_solve(const double *r, const double *c, int64_t n, const double *W,
                      const or_args *args, int64_t frequency, or_log_fn logfn, void *user,
                      double *plan_out, double *theta_out, or_result *out) {
  if (n <= 0 || frequency <= 0) return 1;
  double winf = 0.0; /* norm(W, Inf): the largest |W_ij| (508) */
  for (int64_t k = 0; k < n * n; ++k) winf = fmax(winf, fabs(W[k]));
  const double etap = args->eta_p / (2.0 * winf); /* 509 */
  double *theta = calloc((size_t)n, sizeof(double)), *nu = calloc((size_t)n, sizeof(double));
  double *thetab = calloc((size_t)n, sizeof(double)), *nub = calloc((size_t)n, sizeof(double));
  double *gam = calloc((size_t)n, sizeof(double)), *res = calloc((size_t)n, sizeof(double));
  double *mx = calloc((size_t)n, sizeof(double)), *ls = calloc((size_t)n, sizeof(double));
  double *etamu = calloc((size_t)n, sizeof(double)), *lse = calloc((size_t)n, sizeof(double));
  double *P = malloc(sizeof(double) * (size_t)n * (size_t)n);
  double *rs = calloc((size_t)n, sizeof(double)), *cs = calloc((size_t)n, sizeof(double));
  for (int64_t j = 0; j < n; ++j) etamu[j] = (c[j] + args->alpha / (double)n) / args->tau_mu; /* 523 */
  double etat = INFINITY;
  const double taup = args->tau_p;
  const double minv = tanh(-args->B / 2.0), maxv = tanh(args->B / 2.0);
  double rlogr = 0.0;
  for (int64_t i = 0; i < n; ++i) rlogr += r[i] * log(r[i]);
  emit(logfn, user, "time(s),iter,infeas,ot_objective,primal,dual,solver\n"); /* 527 */
  const double t0 = now_s();
  double feas = NAN, obj = NAN, pobj = NAN, dobj = NAN;
  int converged = 0;
  int64_t it_exit = 0;
  for (int64_t it = 1; it <= args->itermax; ++it) {
    it_exit = it;
    const double elapsed = now_s() - t0;
    if (elapsed > args->tmax) break;
    if ((it - 1) % frequency == 0) { /* 543-557 */
      /* p = softmax(-(W / 2 W_inf .+ nu') ./ etat, norm_dims = 2) */
      cl_rows(W, n, nu, winf, etat, mx, ls);
      OMP_FOR
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j)
          P[i * n + j] = exp(-(W[i * n + j] / (2.0 * winf) + nu[j]) / etat - mx[i] - ls[i]);
      /* feas = norm(sum(r .* p, dims=1)' - c, 1) */
      feas = 0.0;
      for (int64_t j = 0; j < n; ++j) cs[j] = 0.0;
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j) cs[j] += r[i] * P[i * n + j];
      for (int64_t j = 0; j < n; ++j) feas += fabs(cs[j] - c[j]);
      /* primalv(p, W, W_inf, etap, r, c) (MirrorProx.jl:28-32) */
      double wp = 0.0, ent = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        double e = 0.0;
        for (int64_t j = 0; j < n; ++j) {
          const double p = P[i * n + j];
          wp += W[i * n + j] * r[i] * p;
          if (p > 0.0) e += p * log(p);
        }
        ent += r[i] * e;
      }
      pobj = 2.0 * (0.5 * wp + winf * feas + etap * ent + etap * rlogr);
      /* dualv(gamma, W, W_inf, etap, r, c) (MirrorProx.jl:7-13) */
      if (etap == 0.0) {
        double a = 0.0;
        for (int64_t i = 0; i < n; ++i) {
          double m = INFINITY;
          for (int64_t j = 0; j < n; ++j) m = fmin(m, W[i * n + j] + 2.0 * winf * gam[j]);
          a += r[i] * m;
        }
        dobj = -dotv(c, gam, n) * 2.0 * winf + a;
      } else {
        OMP_FOR
        for (int64_t i = 0; i < n; ++i) {
          double m = -INFINITY;
          for (int64_t j = 0; j < n; ++j) m = fmax(m, -(0.5 * W[i * n + j] + winf * gam[j]) / etap);
          double s = 0.0;
          for (int64_t j = 0; j < n; ++j) s += exp(-(0.5 * W[i * n + j] + winf * gam[j]) / etap - m);
          lse[i] = log(s) + m;
        }
        dobj = 2.0 * ((-winf * dotv(c, gam, n) - etap * dotv(r, lse, n)) + etap * rlogr);
      }
      /* obj = dot(round(r .* p, r, c), W) (Utilities.jl:158-165) */
      for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) s += r[i] * P[i * n + j];
        rs[i] = fmin(r[i] / s, 1.0);  /* row scaling of gamma+ */
      }
      for (int64_t j = 0; j < n; ++j) cs[j] = 0.0;
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j) cs[j] += r[i] * P[i * n + j] * rs[i];
      for (int64_t j = 0; j < n; ++j) cs[j] = fmin(c[j] / cs[j], 1.0);  /* column scaling */
      double l1 = 0.0;
      double *rmu = lse, *rnu = mx;  /* (reused as scratch) */
      for (int64_t j = 0; j < n; ++j) rnu[j] = c[j];
      for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) {
          const double g = r[i] * P[i * n + j] * rs[i] * cs[j];
          s += g;
          rnu[j] -= g;
        }
        rmu[i] = r[i] - s;
        l1 += fabs(rmu[i]);
      }
      obj = 0.0;
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j)
          obj += W[i * n + j] * (r[i] * P[i * n + j] * rs[i] * cs[j] + rmu[i] * rnu[j] / l1);
      if (args->verbose) {
        char b0[64], b1[64], b2[64], b3[64], b4[64], line[512];
        fmt_e(b0, 64, "%.6e", elapsed);
        fmt_e(b1, 64, "%.14e", feas);
        fmt_e(b2, 64, "%.14e", obj);
        fmt_e(b3, 64, "%.14e", pobj);
        fmt_e(b4, 64, "%.14e", dobj);
        snprintf(line, sizeof line, "%s,%lld,%s,%s,%s,%s,lamp\n", b0, (long long)it, b1, b2, b3, b4);
        emit(logfn, user, line);
      }
      if (pobj - dobj < args->epsilon / 6.0 && feas < args->epsilon / 6.0) { /* 554 */
        converged = 1;
        break;
      }
    }
    cl_infeas(W, n, r, nu, winf, etat, mx, ls, res);                         /* 560 */
    cl_theta(thetab, theta, res, c, etamu, n, args->tau_mu, args->eta_mu);  /* 561 */
    etat = 1.0 / (taup + (1.0 / etat) * (1.0 - etap));                      /* 562 */
    for (int64_t j = 0; j < n; ++j) nub[j] = nu[j] + taup * etat * (theta[j] - nu[j]); /* 563 */
    cl_infeas(W, n, r, nub, winf, etat, mx, ls, res);                       /* 564 */
    cl_theta(theta, theta, res, c, etamu, n, args->tau_mu, args->eta_mu);   /* 565 */
    for (int64_t j = 0; j < n; ++j) gam[j] = theta[j] + taup * etat * (theta[j] - gam[j]); /* 566 */
    for (int64_t j = 0; j < n; ++j) theta[j] = fmin(fmax(theta[j], minv), maxv);            /* 567 */
    for (int64_t j = 0; j < n; ++j) nu[j] = nu[j] + taup * etat * (thetab[j] - nu[j]);       /* 568 */
  }
  /* 570-571: the plan r .* p of the final nu, eta_t */
  if (plan_out) {
    cl_rows(W, n, nu, winf, etat, mx, ls);
    OMP_FOR
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j)
        plan_out[i * n + j] = r[i] * exp(-(W[i * n + j] / (2.0 * winf) + nu[j]) / etat - mx[i] - ls[i]);
  }
  if (theta_out) memcpy(theta_out, theta, sizeof(double) * (size_t)n);
  if (out) {
    out->iterations = it_exit;
    out->objective = obj;
    out->residual = feas;
    out->primal = pobj;
    out->dual = dobj;
    out->winf = winf;
    out->eta_t = etat;
    out->converged = converged;
  }
  free(theta); free(nu); free(thetab); free(nub); free(gam); free(res); free(mx); free(ls);
  free(etamu); free(lse); free(P); free(rs); free(cs);
  return 0;
}

"""Parity at the benchmark configurations' full sizes, where the oracle cannot
run whole solves in seconds: the oracle on a random subset of rows / columns
against the GPU pass over the full problem, size-independent identities, and
mixed precision against the (oracle-pinned) FP64 path on the same instance."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(ctx, a):
    import culamp_amd
    return culamp_amd.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


@pytest.mark.parametrize("p", [1.0, 2.0, math.inf])
def test_color_transfer_clouds_rows_vs_oracle(gpu_ctx, oracle, p):
    """16384 x 16384 Gaussian-mixture RGB clouds: the full-size row LSE and
    column marginal against the oracle on 64 random rows / columns."""
    import culamp_amd
    from culamp_amd import FP64, MIXED, workloads
    from culamp_amd import kernels as K
    x, y, r, c = workloads.gmm_clouds()
    n, m = len(x), len(y)
    rng = np.random.default_rng(31)
    th = rng.uniform(-0.5, 0.5, m)
    winf = float(oracle.max_cost_rows(x[:256], y, p).max())  # any positive scale
    reg, st = 0.02, 1.0
    rows = rng.choice(n, 64, replace=False)
    ref = oracle.lse_rows(x[rows], y, th, reg, st, winf, p)
    dx, dy, dth = dev(gpu_ctx, x), dev(gpu_ctx, y), dev(gpu_ctx, th)
    out = culamp_amd.DeviceArray(gpu_ctx, n)
    K.warp_logsumexp_spp_ct_fused(gpu_ctx, out, dx, dy, dth, reg, st, winf, p, FP64)
    full = out.get()
    assert np.max(np.abs(full[rows] - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-12
    K.warp_logsumexp_spp_ct_fused(gpu_ctx, out, dx, dy, dth, reg, st, winf, p, MIXED)
    vmax = (st * oracle.max_cost_rows(x[rows], y, p).max() / (2 * winf) + 0.5) / reg
    err = np.max(np.abs(out.get()[rows] - ref))
    say("p", p, "mixed lse err", err, "bound", 4 * 2.0 ** -24 * vmax + 1e-9)
    assert err <= 4 * 2.0 ** -24 * vmax + 1e-9
    # column marginals with the FP64 logZ, 64 random columns
    cols = rng.choice(m, 64, replace=False)
    dr = dev(gpu_ctx, r)
    dlz = dev(gpu_ctx, full)
    res, cost = culamp_amd.DeviceArray(gpu_ctx, m), culamp_amd.DeviceArray(gpu_ctx, m)
    K.residual_spp_c(gpu_ctx, res, cost, dx, dy, dr, dth, dlz, reg, st, winf, p, FP64)
    rres, rcost = oracle.col_marginal_cost(x, y[cols], r, th[cols], full, reg, st, winf, p)
    assert np.allclose(res.get()[cols], rres, rtol=1e-11, atol=1e-300)
    assert np.allclose(cost.get()[cols], rcost, rtol=1e-11, atol=1e-300)
    # mass invariant over all columns: sum_j res_j = sum_i r_i
    assert abs(res.get().sum() - r.sum()) < 1e-12


def test_color_transfer_solves_mixed_vs_fp64(gpu_ctx):
    """16384-point clouds with uniform weights, eta_p = 1e-2.  Sinkhorn: mixed
    precision reaches infeasibility 1e-6 and its transport cost stays within
    1e-5 relative of the FP64 path's.  LAMP: the infeasibility stop runs real
    iterations (uniform marginals make the eta_t = inf product plan feasible,
    so a stop at that first check would return it), stops where the
    reference's own gap-and-infeasibility stop does, and after the same
    number of iterations the mixed cost is within 1e-5 of the FP64 one."""
    import culamp_amd
    from culamp_amd import workloads
    x, y, r, c = workloads.gmm_clouds()
    pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
    base = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False,
                              stop="infeas")
    info = {}
    for prec in ("fp64", "mixed"):
        info[prec] = pb.sinkhorn(base.copy(precision=prec), 10, 2.0, log=None, assignments=False)[-1]
    f, mx = info["fp64"], info["mixed"]
    say("sinkhorn fp64", f["iterations"], f["residual"], f["objective"], "mixed", mx["iterations"],
        mx["residual"], mx["objective"])
    assert f["converged"] == 1 and mx["converged"] == 1 and mx["residual"] <= 1e-6
    assert abs(mx["objective"] - f["objective"]) / f["objective"] < 1e-5

    f = pb.lamp(base.copy(precision="fp64"), 10, 2.0, log=None, assignments=False)[-1]
    mx = pb.lamp(base.copy(precision="mixed"), 10, 2.0, log=None, assignments=False)[-1]
    ref = pb.lamp(base.copy(precision="fp64", stop="reference", verbose=True), 10, 2.0,
                  log=None, assignments=False)[-1]
    say("lamp fp64", f["iterations"], f["residual"], f["objective"], "mixed", mx["iterations"],
        mx["residual"], mx["objective"], "reference stop", ref["iterations"], ref["objective"])
    assert f["iterations"] > 0 and mx["iterations"] > 0
    assert f["converged"] == 1 and mx["converged"] == 1
    assert f["residual"] < 1e-6 and mx["residual"] < 1e-6
    # far from the product plan's cost (the eta_t = inf check)
    prod = float(r @ (((x[:, None, :] - y[None, ::64, :]) ** 2).sum(-1)).mean(1))
    assert f["objective"] < 0.6 * prod
    assert ref["converged"] == 1 and ref["iterations"] == f["iterations"]
    assert abs(ref["objective"] - f["objective"]) <= 1e-12 * f["objective"]
    # the same iteration count in both precisions: the check at it = K + 1
    # reports the state after K iterations (epsilon 0 never stops)
    k = f["iterations"]
    fixed = base.copy(stop="reference", verbose=True, epsilon=0.0, itermax=k + 1)
    f2 = pb.lamp(fixed.copy(precision="fp64"), 10, 2.0, log=None, assignments=False)[-1]
    g2 = pb.lamp(fixed.copy(precision="mixed"), 10, 2.0, log=None, assignments=False)[-1]
    say("lamp after", k, "iterations: fp64", f2["objective"], f2["residual"], "mixed",
        g2["objective"], g2["residual"])
    assert abs(f2["objective"] - f["objective"]) <= 1e-12 * f["objective"]
    assert abs(g2["objective"] - f2["objective"]) / f2["objective"] < 1e-5
    # FP32 noise in the mixed marginals: a few 1e-8 at this scale
    assert abs(g2["residual"] - f2["residual"]) < 2e-7
    pb.close()


def test_grid_histograms_lamp_pnorms(gpu_ctx):
    """64 x 64 grid histograms (n = m = 4096), p in {1, 2, inf}, eta_p = 1e-2:
    mixed LAMP reaches infeasibility 1e-6, and after the same 600 iterations
    its cost and infeasibility agree with the FP64 path's (stop iterations may
    differ by a check, and LAMP's cost still moves at that point)."""
    import culamp_amd
    from culamp_amd import workloads
    x, y, r, c = workloads.grid_histograms(64)
    pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
    for p in (1.0, 2.0, math.inf):
        base = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False,
                                  stop="infeas")
        mx = pb.lamp(base.copy(precision="mixed"), 10, p, log=None, assignments=False)[-1]
        assert mx["converged"] == 1 and mx["residual"] < 1e-6
        fixed = base.copy(stop="reference", verbose=True, itermax=600)
        f = pb.lamp(fixed.copy(precision="fp64"), 10, p, log=None, assignments=False)[-1]
        g = pb.lamp(fixed.copy(precision="mixed"), 10, p, log=None, assignments=False)[-1]
        say("p", p, "mixed stop", mx["iterations"], "600 it: fp64", f["objective"], f["residual"],
            "mixed", g["objective"], g["residual"])
        assert f["iterations"] == g["iterations"] == 600
        assert abs(g["objective"] - f["objective"]) / f["objective"] < 1e-5
        # the mixed infeasibility carries FP32 noise of a few 1e-7 at this scale (p = 1:
        # C / (reg W_inf) up to ~100 at eta_t -> eta_p / (2 W_inf))
        assert abs(g["residual"] - f["residual"]) < 5e-7
    pb.close()


@pytest.mark.parametrize("p", [1.0, 2.0, math.inf])
def test_grid_histograms_lamp_eta_1e4(gpu_ctx, oracle, p):
    """64 x 64 grid histograms, eta_p = 1e-4 (the small-regularisation half of
    the 4096-point grid configuration).  FP64 LAMP stops at infeasibility 1e-6
    and its returned state is checked with the oracle at full size: the row
    LSE on 64 random rows, and the column marginal over all 4096 columns,
    whose L1 infeasibility and cost must be the ones the solver reported.
    Mixed precision: FP32 exponents reach C / eta_t ~ 1e3 here, so the mixed
    infeasibility has a floor of 1e-6 to 4e-6; the solve with its FP64
    takeover reaches 1e-6, and after the FP64 solve's iteration count the
    pure-mixed cost is within 1e-5 of the FP64 cost."""
    import culamp_amd
    from culamp_amd import workloads
    x, y, r, c = workloads.grid_histograms(64)
    pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
    base = culamp_amd.EOTArgs(eta_p=1e-4, epsilon=6e-6, B=1.1, itermax=20000, verbose=False,
                              stop="infeas")
    nu, phi, psi, _, _, f = pb.lamp(base.copy(precision="fp64"), 10, p, log=None,
                                    assignments=False)
    say("p", p, "fp64", f["iterations"], f["residual"], f["objective"], f["elapsed_s"])
    assert f["converged"] == 1 and 0 < f["iterations"] and f["residual"] < 1e-6
    assert f["fp64_from"] == 0
    # the returned state, from the oracle
    winf, reg = f["winf"], f["eta_t"]
    # the schedule 1 / eta_t = (1 - (1 - eta)^K) / eta, eta = eta_p / (2 W_inf)
    # (MirrorProxKernel.jl:1035, 1108): far above eta after K ~ 1e3 steps
    eta = 1e-4 / (2 * winf)
    assert abs(reg * (1 - (1 - eta) ** f["iterations"]) / eta - 1) < 1e-9
    logz = np.log(r) - phi
    rows = np.random.default_rng(41).choice(len(x), 64, replace=False)
    ref = oracle.lse_rows(x[rows], y, nu, reg, 1.0, winf, p)
    assert np.max(np.abs(logz[rows] - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-11
    res, cost = oracle.col_marginal_cost(x, y, r, nu, logz, reg, 1.0, winf, p)
    infeas = float(np.abs(res - c).sum())
    say("   oracle infeasibility", infeas, "cost", cost.sum())
    # (|res_j - c_j| cancels: 4096 terms of ~1e-19 rounding each)
    assert abs(infeas - f["residual"]) < 1e-14
    assert abs(cost.sum() - f["objective"]) < 1e-11 * f["objective"]
    # mixed with the FP64 takeover
    g = pb.lamp(base.copy(precision="mixed"), 10, p, log=None, assignments=False)[-1]
    say("   mixed+takeover", g["iterations"], g["residual"], g["objective"], "fp64 from",
        g["fp64_from"], g["elapsed_s"])
    assert g["converged"] == 1 and g["residual"] < 1e-6
    if g["fp64_from"]:
        assert 0 < g["fp64_from"] <= g["iterations"] + 1
    # pure mixed after the FP64 solve's iterations (the check at it = K + 1)
    k = f["iterations"]
    fixed = base.copy(stop="reference", verbose=True, epsilon=0.0, itermax=k + 1)
    gpu_ctx.set_fp64_fallback(False)
    try:
        h = pb.lamp(fixed.copy(precision="mixed"), 10, p, log=None, assignments=False)[-1]
    finally:
        gpu_ctx.set_fp64_fallback(True)
    say("   mixed after", k, "iterations:", h["objective"], h["residual"], "cost rel err",
        abs(h["objective"] - f["objective"]) / f["objective"])
    assert h["fp64_from"] == 0
    assert abs(h["objective"] - f["objective"]) / f["objective"] < 1e-5
    assert h["residual"] < 1e-5
    pb.close()


def test_uniform_grid_sinkhorn_vs_dense_reference(gpu_ctx, oracle):
    """32 x 32 grid (n = m = 1024), the reference's CPU path: dense log-domain
    Sinkhorn on W / max W (src/solvers/Sinkhorn.jl:8-84) and the on-the-fly
    GPU kernel solve reach the same fixed point (same potentials up to the
    constant shift, same plan and cost)."""
    import culamp_amd
    from culamp_amd import workloads
    x, y, r, c = workloads.uniform_grid_pair(32)
    W = oracle.grid_cost_matrix(32, 32, 2.0)
    W = W / W.max()
    eta = 1e-2
    d = oracle.sinkhorn_log_dense(r, c, W, frequency=10, eta_p=eta, epsilon=6e-12, itermax=200000,
                                  verbose=False)
    pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
    a = culamp_amd.EOTArgs(eta_p=eta, epsilon=6e-12, itermax=200000, verbose=False)
    phi, psi, _, _, info = pb.sinkhorn(a, 10, 2.0, log=None, assignments=False)
    K = -W / eta
    P_gpu = np.exp(K + phi[:, None] + psi[None, :])
    say("dense iters", d["iterations"], "gpu iters", info["iterations"])
    assert np.max(np.abs(P_gpu - d["P"])) < 1e-10 * d["P"].max() + 1e-14
    # the kernel path reports <P, C> in the cost's own units: C = W * W_inf
    assert abs(info["objective"] - (P_gpu * W).sum() * info["winf"]) < 1e-10 * info["objective"]
    shift = np.mean(phi - d["phi"])
    assert np.max(np.abs(phi - d["phi"] - shift)) < 1e-7
    pb.close()

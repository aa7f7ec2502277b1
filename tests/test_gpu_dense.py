"""Dense cost-matrix passes (culamp_dense_*, the reference's plain dense GPU
path): against the oracle's on-the-fly restatement of the same functions on
W_ij = C(x_i, y_j), against the on-the-fly GPU kernels, FP64 and mixed, over
ragged shapes."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_REL = 1e-12
MIXED_LSE_ULPS = 4 * 2.0 ** -24


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(ctx, a):
    import culamp_amd
    return culamp_amd.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


def cost_matrix(x, y, p):
    d = np.abs(x[:, None, :] - y[None, :, :])
    if p == 1.0:
        return d.sum(-1)
    if p == 2.0:
        return (d ** 2).sum(-1)
    if math.isinf(p):
        return d.max(-1)
    return (d ** p).sum(-1)


@pytest.mark.parametrize("n,m", [(1, 1), (5, 3), (257, 129), (1000, 777), (700, 2100)])
@pytest.mark.parametrize("p", [1.0, 2.0, math.inf])
def test_dense_passes_vs_oracle(gpu_ctx, oracle, n, m, p):
    import culamp_amd
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    say("dense n", n, "m", m, "p", p)
    rng = np.random.default_rng(n + 13 * m)
    x, y = rng.uniform(size=(n, 3)), rng.uniform(size=(m, 3))
    th = rng.uniform(-1, 1, m)
    r = rng.uniform(0.1, 1, n)
    r /= r.sum()
    W = cost_matrix(x, y, p)
    winf = float(np.abs(W).max())
    reg, st = 0.02, 1.0
    ref = oracle.lse_rows(x, y, th, reg, st, winf, p)
    res_ref, _ = oracle.col_marginal_cost(x, y, r, th, ref, reg, st, winf, p)
    dth, dr, dlz = dev(gpu_ctx, th), dev(gpu_ctx, r), dev(gpu_ctx, ref)
    o, res = culamp_amd.DeviceArray(gpu_ctx, n), culamp_amd.DeviceArray(gpu_ctx, m)
    d64 = culamp_amd.DenseProblem(W, "fp64", gpu_ctx)
    d64.lse_rows(o, dth, reg, st, winf)
    dense_lse = o.get()
    assert np.max(np.abs(dense_lse - ref) / np.maximum(np.abs(ref), 1.0)) < FP64_REL
    d64.col_marginal(res, dr, dth, dlz, reg, st, winf)
    assert np.max(np.abs(res.get() - res_ref) / np.maximum(np.abs(res_ref), 1e-300)) < 1e-11
    # the on-the-fly kernel computes the same function of the same costs
    dx, dy = dev(gpu_ctx, x), dev(gpu_ctx, y)
    K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, reg, st, winf, p, FP64)
    assert np.max(np.abs(dense_lse - o.get()) / np.maximum(np.abs(o.get()), 1.0)) < FP64_REL
    d64.close()
    # mixed: FP32 matrix and exponents, FP64 sums
    d32 = culamp_amd.DenseProblem(W, "mixed", gpu_ctx)
    d32.lse_rows(o, dth, reg, st, winf)
    vmax = (W.max() / (2 * winf) + 1.0) / reg
    err = np.max(np.abs(o.get() - ref))
    say("mixed dense lse err", err)
    assert err < MIXED_LSE_ULPS * vmax + 1e-9
    d32.col_marginal(res, dr, dth, dlz, reg, st, winf)
    tol = 2e-5 * max(1.0, vmax * 2.0 ** -24 * 1e4)
    assert np.max(np.abs(res.get() - res_ref)) <= tol * max(res_ref.max(), 1e-300)
    # mass invariant over all columns: sum_j res_j = sum_i r_i
    d64 = culamp_amd.DenseProblem(W, "fp64", gpu_ctx)
    d64.col_marginal(res, dr, dth, dlz, reg, st, winf)
    assert abs(res.get().sum() - r.sum()) < 1e-12
    d32.close()
    d64.close()


def test_dense_grid_4096_matches_on_the_fly(gpu_ctx, oracle):
    """64 x 64 grid (n = m = 4096), the dense matrix of the reference's
    `run` path: rows against the oracle on 48 sampled rows, and the dense
    FP64 pass against the on-the-fly FP64 pass everywhere."""
    import culamp_amd
    from culamp_amd import FP64, workloads
    from culamp_amd import kernels as K
    x, y, r, c = workloads.grid_histograms(64)
    W = cost_matrix(x, y, 2.0)
    winf = float(W.max())
    rng = np.random.default_rng(3)
    th = rng.uniform(-0.01, 0.01, len(y))
    dth = dev(gpu_ctx, th)
    o = culamp_amd.DeviceArray(gpu_ctx, len(x))
    o2 = culamp_amd.DeviceArray(gpu_ctx, len(x))
    dm = culamp_amd.DenseProblem(W, "fp64", gpu_ctx)
    dm.lse_rows(o, dth, 0.0025, 1.0, winf)
    K.warp_logsumexp_spp_ct_fused(gpu_ctx, o2, dev(gpu_ctx, x), dev(gpu_ctx, y), dth, 0.0025,
                                  1.0, winf, 2.0, FP64)
    a, b = o.get(), o2.get()
    assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)) < FP64_REL
    rows = rng.choice(len(x), 48, replace=False)
    ref = oracle.lse_rows(x[rows], y, th, 0.0025, 1.0, winf, 2.0)
    assert np.max(np.abs(a[rows] - ref) / np.maximum(np.abs(ref), 1.0)) < FP64_REL
    dm.close()


def _parse(lines):
    out = []
    for line in lines:
        f = line.strip().split(",")
        if f[0].startswith("time"):
            continue
        out.append((int(f[1]), float(f[2]), float(f[3]), float(f[4]), float(f[5]), float(f[6])))
    return out


@pytest.mark.parametrize("eta", [1e-2, 0.0])
def test_dense_lamp_fp64_vs_oracle(gpu_ctx, oracle, eta):
    """The dense-matrix LAMP loop (MirrorProx.jl:397-503) on a 10 x 10 grid:
    every logged check (infeasibility, rounded-plan cost, primal, dual, gap)
    and the final theta and plan against the oracle's restatement."""
    import culamp_amd
    from culamp_amd import workloads
    x, y, r, c = workloads.uniform_grid_pair(10)
    W = oracle.grid_cost_matrix(10, 10, 2.0)
    kw = dict(eta_p=eta, epsilon=1e-9, B=1.1, itermax=400, verbose=True)
    o = oracle.dense_lamp_solve(r, c, W, frequency=20, **kw)
    lines = []
    dm = culamp_amd.DenseProblem(W, "fp64", gpu_ctx)
    P, theta, info = dm.lamp(r, c, culamp_amd.EOTArgs(**kw), 20, log=lines.append)
    dm.close()
    got = _parse(lines)
    assert len(got) == len(o["log"])
    for g, (it, feas, obj, pv, dv) in zip(got, o["log"]):
        assert g[0] == it
        assert abs(g[1] - feas) <= 1e-9 * max(feas, 1e-6), (it, g[1], feas)
        assert abs(g[2] - obj) <= 1e-9 * abs(obj), (it, g[2], obj)
        assert abs(g[3] - pv) <= 1e-9 * abs(pv) and abs(g[4] - dv) <= 1e-9 * abs(dv), it
    assert info["iterations"] == o["iterations"]
    assert np.max(np.abs(theta - o["theta"])) < 1e-9
    assert np.max(np.abs(P - o["plan"])) < 1e-12


def test_dense_lamp_ragged_and_mixed(gpu_ctx, oracle):
    """A non-square random cost (n = 70, m = 45): FP64 against the oracle;
    the FP32 matrix (mixed) reaches the same plan cost within 1e-5."""
    import culamp_amd
    rng = np.random.default_rng(9)
    n, m = 70, 45
    W = rng.uniform(0, 1, (n, m)) ** 2
    r, c = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(m))
    kw = dict(eta_p=5e-2, epsilon=1e-9, B=1.1, itermax=600, verbose=False)
    o = oracle.dense_lamp_solve(r, c, W, frequency=25, **kw)
    dm = culamp_amd.DenseProblem(W, "fp64", gpu_ctx)
    P, theta, info = dm.lamp(r, c, culamp_amd.EOTArgs(**kw), 25, log=None)
    dm.close()
    assert info["iterations"] == o["iterations"]
    assert np.max(np.abs(P - o["plan"])) < 1e-12
    assert np.max(np.abs(theta - o["theta"])) < 1e-9
    d32 = culamp_amd.DenseProblem(W, "mixed", gpu_ctx)
    P32, _, _ = d32.lamp(r, c, culamp_amd.EOTArgs(**kw), 25, log=None)
    d32.close()
    cost, cost32 = float((W * o["plan"]).sum()), float((W * P32).sum())
    say("dense lamp mixed cost rel err", abs(cost32 - cost) / cost)
    assert abs(cost32 - cost) < 1e-5 * cost

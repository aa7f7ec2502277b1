"""Feature-cost (cell-similarity) parity: the feature kernels against the
oracle's cost matrix, and the cellsim LAMP / Sinkhorn / annealed Sinkhorn
loops against the oracle's restatement of the reference loops."""
import sys

import numpy as np
import pytest
from scipy.special import logsumexp

pytestmark = pytest.mark.gpu

METRICS = ["l1", "l2", "cosine", "correlation"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cells(seed, n, d, density=0.5):
    rng = np.random.default_rng(seed)
    x = np.abs(rng.normal(size=(n, d))) * (rng.uniform(size=(n, d)) < density) + 1e-3
    return x


def dev(ctx, a):
    import culamp_amd
    return culamp_amd.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("n,m,d", [(1, 1, 1), (70, 45, 13), (300, 257, 128)])
def test_feature_kernels_vs_oracle(gpu_ctx, oracle, metric, n, m, d):
    import culamp_amd
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    say("feature kernels", metric, n, m, d)
    x, y = cells(n + d, n, d), cells(m + 7 * d, m, d)
    rng = np.random.default_rng(n * m)
    r, c = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(m))
    scale = culamp_amd.cost_scale(metric)
    C = oracle.feature_cost_matrix(x, y, metric, scale)
    pb = culamp_amd.FeatureProblem(x, y, scale, metric, r, c, gpu_ctx)
    th = rng.uniform(-1, 1, m)
    winf, reg, st = (C.max() if C.max() > 0 else 1.0), 0.03, 1.0
    V = -(C * st / (2 * winf) + th[None, :]) / reg
    ref = logsumexp(V, axis=1)
    out = culamp_amd.DeviceArray(gpu_ctx, n)
    dth = dev(gpu_ctx, th)
    for rpt_splits in ((0, 0), (0, 1), (0, 3)):
        gpu_ctx.set_launch(*rpt_splits)
        try:
            K.warp_logsumexp_spp_sim_fused(gpu_ctx, out, pb, dth, reg, st, winf, FP64)
            assert np.max(np.abs(out.get() - ref) / np.maximum(np.abs(ref), 1)) < 1e-12
            K.warp_logsumexp_spp_sim_fused(gpu_ctx, out, pb, dth, reg, st, winf, MIXED)
            vmax = (0.5 + 1.0) / reg
            assert np.max(np.abs(out.get() - ref)) < 1e-5 * vmax + 1e-7
        finally:
            gpu_ctx.set_launch(0, 0)
    # column marginal + cost (residual_cellsim_c!)
    W = np.exp(V - ref[:, None]) * r[:, None]
    res_ref, cost_ref = W.sum(0), (W * C).sum(0)
    res, cost = culamp_amd.DeviceArray(gpu_ctx, m), culamp_amd.DeviceArray(gpu_ctx, m)
    K.residual_cellsim_c(gpu_ctx, res, cost, pb, dev(gpu_ctx, r), dth, dev(gpu_ctx, ref), reg, st,
                         winf, FP64)
    assert np.allclose(res.get(), res_ref, rtol=1e-11, atol=1e-300)
    assert np.allclose(cost.get(), cost_ref, rtol=1e-11, atol=1e-300)
    # Sinkhorn half-steps both ways and the row residual
    psi, phi = rng.normal(0, 2, m), rng.normal(0, 2, n)
    k = -1.0 / (reg * winf)
    half_r = np.log(r) - logsumexp(C * k + psi[None, :], axis=1)
    half_c = np.log(c) - logsumexp(C.T * k + phi[None, :], axis=1)
    K.warp_logsumexp_cellsim_fused(gpu_ctx, out, pb, True, dev(gpu_ctx, r), dev(gpu_ctx, psi), reg, winf, FP64)
    assert np.allclose(out.get(), half_r, rtol=1e-12, atol=1e-12)
    oc = culamp_amd.DeviceArray(gpu_ctx, m)
    K.warp_logsumexp_cellsim_fused(gpu_ctx, oc, pb, False, dev(gpu_ctx, c), dev(gpu_ctx, phi), reg, winf, FP64)
    assert np.allclose(oc.get(), half_c, rtol=1e-12, atol=1e-12)
    E = np.exp(C * k + phi[:, None] + psi[None, :])
    rr, rc = culamp_amd.DeviceArray(gpu_ctx, n), culamp_amd.DeviceArray(gpu_ctx, n)
    K.residual_cellsim_opt(gpu_ctx, rr, rc, pb, dev(gpu_ctx, r), dev(gpu_ctx, phi), dev(gpu_ctx, psi), reg, winf, FP64)
    assert np.allclose(rr.get(), r - E.sum(1), rtol=1e-10, atol=1e-12 * E.sum(1).max())
    assert np.allclose(rc.get(), (E * C).sum(1), rtol=1e-11, atol=1e-300)
    # max (running max starts at 0) and min-plus
    K.max_logsumexp_cellsim(gpu_ctx, out, pb)
    assert np.array_equal(out.get(), np.maximum(C.max(1), 0.0)) or np.allclose(out.get(), np.maximum(C.max(1), 0), rtol=1e-14)
    K.warp_min_reduce_cellsim(gpu_ctx, out, pb, dth, winf)
    assert np.allclose(out.get(), (C + 2 * winf * th[None, :]).min(1), rtol=1e-13, atol=1e-14)
    pb.close()


def compare_logs(gpu_lines, ref_lines, rtol):
    assert len(gpu_lines) == len(ref_lines), (len(gpu_lines), len(ref_lines))
    assert gpu_lines[0] == ref_lines[0]
    for a, b in zip(gpu_lines[1:], ref_lines[1:]):
        fa, fb = a.rstrip("\n").split(","), b.rstrip("\n").split(",")
        assert fa[1] == fb[1] and fa[-1] == fb[-1], (a, b)
        for u, v in zip(fa[2:-1], fb[2:-1]):
            assert abs(float(u) - float(v)) <= rtol * max(abs(float(v)), 1e-12) + 1e-13, (a, b)


@pytest.mark.parametrize("metric", METRICS)
def test_lamp_cellsim_matches_oracle(gpu_ctx, oracle, metric):
    import culamp_amd
    say("lamp cellsim", metric)
    X = cells(3, 60, 16)
    rng = np.random.default_rng(4)
    r, c = rng.dirichlet(np.ones(60)), rng.dirichlet(np.ones(60))
    scale = culamp_amd.cost_scale(metric)
    for eta, stop, itermax in ((1e-2, "infeas", 20000), (0.0, "reference", 600)):
        kw = dict(eta_p=eta, epsilon=6e-6, B=1.1, itermax=itermax, verbose=True, stop=stop)
        lines = []
        nu, phi, psi, obj, info = culamp_amd.extragradient_cellsim(
            X, scale, metric, r, c, culamp_amd.EOTArgs(**kw), 10, log=lines.append, ctx=gpu_ctx,
            return_info=True)
        o = oracle.lamp_cellsim_solve(X, r, c, metric, scale, frequency=10, **kw)
        assert info["iterations"] == o["iterations"], (eta, info["iterations"], o["iterations"])
        compare_logs(lines, o["log"], 1e-7)
        assert np.allclose(nu, o["nu"], atol=1e-8)
        assert np.allclose(phi, o["phi"], rtol=1e-8, atol=1e-8)
        assert abs(obj - o["objective"]) <= 1e-9 * abs(o["objective"]) + 1e-14


@pytest.mark.parametrize("metric", METRICS)
def test_sinkhorn_cellsim_matches_oracle(gpu_ctx, oracle, metric):
    import culamp_amd
    say("sinkhorn cellsim", metric)
    X = cells(5, 80, 20)
    rng = np.random.default_rng(6)
    r, c = rng.dirichlet(np.ones(80)), rng.dirichlet(np.ones(80))
    scale = culamp_amd.cost_scale(metric)
    args = culamp_amd.EOTArgs(eta_p=0.05, epsilon=6e-8, itermax=50000, verbose=True)
    lines = []
    phi, psi, obj, res, nit, info = culamp_amd.sinkhorn_cellsim(
        X, X, scale, metric, r, c, args, 10, return_cuda=True, log=lines.append, ctx=gpu_ctx,
        return_info=True)
    o = oracle.sinkhorn_cellsim_solve(X, X, r, c, metric, scale, frequency=10, eta_p=0.05,
                                      epsilon=6e-8, itermax=50000, verbose=True)
    assert nit == o["iterations"]
    compare_logs(lines, o["log"], 1e-7)
    assert np.allclose(phi, o["phi"], atol=1e-9) and np.allclose(psi, o["psi"], atol=1e-9)
    # cross problem (x != y), ragged
    Y = cells(7, 55, 20)
    c2 = rng.dirichlet(np.ones(55))
    out = culamp_amd.sinkhorn_cellsim(X, Y, scale, metric, r, c2,
                                      culamp_amd.EOTArgs(eta_p=0.05, epsilon=6e-8, verbose=False),
                                      5, return_cuda=True, log=None, ctx=gpu_ctx)
    o2 = oracle.sinkhorn_cellsim_solve(X, Y, r, c2, metric, scale, frequency=5, eta_p=0.05,
                                       epsilon=6e-8, verbose=False)
    assert out[4] == o2["iterations"]
    assert np.allclose(out[0], o2["phi"], atol=1e-9)


def test_annealed_sinkhorn_cellsim_matches_oracle(gpu_ctx, oracle):
    import culamp_amd
    X = cells(8, 50, 12)
    rng = np.random.default_rng(9)
    r, c = rng.dirichlet(np.ones(50)), rng.dirichlet(np.ones(50))
    args = culamp_amd.EOTArgs(eta_p=0.2, epsilon=1e-2, anneal_mult=0.5, verbose=True)
    lines = []
    phi, psi, obj, info = culamp_amd.annealed_sinkhorn_cellsim(X, X, 2.0, "cosine", r, c, args, 5,
                                                               log=lines.append, ctx=gpu_ctx,
                                                               return_info=True)
    o = oracle.annealed_sinkhorn_cellsim_solve(X, X, r, c, "cosine", 2.0, frequency=5, eta_p=0.2,
                                               epsilon=1e-2, anneal_mult=0.5, verbose=True)
    assert info["iterations"] == o["iterations"]
    compare_logs(lines, o["log"], 1e-7)
    assert np.allclose(phi, o["phi"], atol=1e-8) and abs(obj - o["objective"]) < 1e-10


def test_mixed_cellsim_sinkhorn(gpu_ctx, oracle):
    """Mixed precision (FP32 feature dots) converges to infeasibility 1e-6 and
    its cost stays within 1e-5 relative of the FP64 oracle."""
    import culamp_amd
    X = cells(10, 500, 128, 0.3)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    Y = cells(11, 400, 128, 0.3)
    Y /= np.linalg.norm(Y, axis=1, keepdims=True)
    rng = np.random.default_rng(12)
    r, c = rng.dirichlet(np.ones(500)), rng.dirichlet(np.ones(400))
    args = culamp_amd.EOTArgs(eta_p=5e-2, epsilon=6e-6, verbose=False, precision="mixed")
    phi, psi, obj, res, nit = culamp_amd.sinkhorn_cellsim(X, Y, 2.0, "cosine", r, c, args, 10,
                                                          return_cuda=True, log=None, ctx=gpu_ctx)
    o = oracle.sinkhorn_cellsim_solve(X, Y, r, c, "cosine", 2.0, frequency=10, eta_p=5e-2,
                                      epsilon=6e-6, verbose=False)
    say("mixed cellsim", nit, res, obj, "fp64", o["iterations"], o["residual"], o["objective"])
    assert res <= 1e-6
    assert abs(obj - o["objective"]) / o["objective"] < 1e-5


@pytest.mark.parametrize("metric", ["cosine", "correlation"])
@pytest.mark.parametrize("n,m,d", [(37, 300, 7), (1000, 700, 128)])
def test_mfma_feature_lse_and_marginal(gpu_ctx, metric, n, m, d):
    """bf16 MFMA dots (single product, and the 3-product split) against the
    FP64 kernel on the same problem.  Tolerances: the similarity error of a
    bf16 dot of unit rows is <= 2^-7 (single) / 2^-15 (split); the exponent
    error is that times |k| / scale."""
    import culamp_amd
    from culamp_amd import BF16, BF16X3, FP64
    from culamp_amd import kernels as K
    say("mfma", metric, n, m, d)
    rng = np.random.default_rng(n + m + d)
    x = np.abs(rng.normal(size=(n, d))) * (rng.uniform(size=(n, d)) < 0.3) + 0.01
    y = np.abs(rng.normal(size=(m, d))) * (rng.uniform(size=(m, d)) < 0.3) + 0.01
    r, c = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(m))
    pb = culamp_amd.FeatureProblem(x, y, 2.0, metric, r, c, gpu_ctx)
    th = dev(gpu_ctx, rng.uniform(-1, 1, m))
    reg, st, winf = 0.05, 1.0, 0.7
    ref, out = culamp_amd.DeviceArray(gpu_ctx, n), culamp_amd.DeviceArray(gpu_ctx, n)
    K.warp_logsumexp_spp_sim_fused(gpu_ctx, ref, pb, th, reg, st, winf, FP64)
    kscale = 1.0 / (2 * winf * reg * 2.0)
    for prec, bound in ((BF16, 2.0 ** -7), (BF16X3, 2.0 ** -15)):
        for splits in (0, 1, 5):
            gpu_ctx.set_launch(0, splits)
            try:
                K.warp_logsumexp_spp_sim_fused(gpu_ctx, out, pb, th, reg, st, winf, prec)
            finally:
                gpu_ctx.set_launch(0, 0)
            err = np.max(np.abs(out.get() - ref.get()))
            say("  prec", prec, "splits", splits, "lse err", err)
            assert err <= bound * kscale + 1e-6
    # column marginal with the FP64 logZ
    res0, res1 = culamp_amd.DeviceArray(gpu_ctx, m), culamp_amd.DeviceArray(gpu_ctx, m)
    c0, c1 = culamp_amd.DeviceArray(gpu_ctx, m), culamp_amd.DeviceArray(gpu_ctx, m)
    rd = dev(gpu_ctx, r)
    K.residual_cellsim_c(gpu_ctx, res0, c0, pb, rd, th, ref, reg, st, winf, FP64)
    K.residual_cellsim_c(gpu_ctx, res1, c1, pb, rd, th, ref, reg, st, winf, BF16X3)
    rel = np.max(np.abs(res1.get() - res0.get()) / res0.get())
    say("  marginal rel err (bf16x3)", rel)
    assert rel < 2.0 ** -15 * kscale * 4 + 1e-6
    pb.close()


def test_bf16_sinkhorn_cellsim_converges(gpu_ctx, oracle):
    """Config-4-shaped instance: d = 128 sparse unit rows, cosine, eta_p = 5e-2.
    bf16x3 and bf16 reach infeasibility 1e-6 under their own arithmetic; the
    transport cost error against the FP64 oracle is stated by the bounds."""
    import culamp_amd
    rng = np.random.default_rng(40)
    X = np.abs(rng.normal(size=(700, 128))) * (rng.uniform(size=(700, 128)) < 0.3)
    Y = np.abs(rng.normal(size=(500, 128))) * (rng.uniform(size=(500, 128)) < 0.3)
    X[X.sum(1) == 0, 0] = 1.0
    Y[Y.sum(1) == 0, 0] = 1.0
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    Y /= np.linalg.norm(Y, axis=1, keepdims=True)
    r, c = rng.dirichlet(np.ones(700)), rng.dirichlet(np.ones(500))
    o = oracle.sinkhorn_cellsim_solve(X, Y, r, c, "cosine", 2.0, frequency=10, eta_p=5e-2,
                                      epsilon=6e-6, itermax=20000, verbose=False)
    # (observed: bf16 3.4e-6 on this instance; bf16x3 below 1e-7 on the batch instances)
    for prec, tol in (("bf16x3", 1e-6), ("bf16", 2e-4)):
        args = culamp_amd.EOTArgs(eta_p=5e-2, epsilon=6e-6, itermax=20000, verbose=False, precision=prec)
        phi, psi, obj, res, nit = culamp_amd.sinkhorn_cellsim(X, Y, 2.0, "cosine", r, c, args, 10,
                                                              return_cuda=True, log=None, ctx=gpu_ctx)
        say(prec, nit, res, obj, "fp64", o["iterations"], o["objective"])
        assert res <= 1e-6
        assert abs(obj - o["objective"]) / o["objective"] < tol

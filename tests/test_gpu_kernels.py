"""Kernel-level parity: every libculamp kernel entry point against the golden
vectors and the FP64 oracle, FP64 and mixed precision, ragged sizes, forced
launch shapes (targets per thread x source splits) and edge cases."""
import math
import sys

import numpy as np
import pytest

from conftest import pval, unhex

pytestmark = pytest.mark.gpu

# Mixed precision tolerance on an LSE output: FP32 cost/exponent rounding is
# relative 2^-24 per term; the bound is in units of the largest |exponent|
# of the row (4 ulps) plus an FP64-merge floor.
MIXED_LSE_ULPS = 4 * 2.0 ** -24
# Mixed precision tolerance on sums (marginals, costs): relative to the sum.
MIXED_SUM_REL = 2e-5
FP64_REL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(ctx, a):
    import culamp_amd
    return culamp_amd.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


def out(ctx, n):
    import culamp_amd
    return culamp_amd.DeviceArray(ctx, n)


def rel(a, b, floor=1e-300):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def test_golden_lamp(gpu_ctx, golden):
    from culamp_amd import kernels as K
    for c in golden:
        if c["kind"] != "lamp":
            continue
        say("golden lamp p", c["p"], "reg", c["reg"])
        x, y = unhex(c["x"], (-1, 3)), unhex(c["y"], (-1, 3))
        n, m = len(x), len(y)
        p, reg = pval(c["p"]), pval(c["reg"])
        dx, dy, th, r = dev(gpu_ctx, x), dev(gpu_ctx, y), dev(gpu_ctx, unhex(c["theta"])), dev(gpu_ctx, unhex(c["r"]))
        o = out(gpu_ctx, n)
        K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, th, reg, c["st"], c["winf"], p)
        assert rel(o.get(), c["logZ"]) < FP64_REL
        lz = dev(gpu_ctx, unhex(c["logZ_in"]))
        res, cost = out(gpu_ctx, m), out(gpu_ctx, m)
        K.residual_spp_c(gpu_ctx, res, cost, dx, dy, r, th, lz, reg, c["st"], c["winf"], p)
        assert rel(res.get(), c["res"]) < FP64_REL
        assert rel(cost.get(), c["cost"]) < FP64_REL
        K.warp_min_reduce(gpu_ctx, o, dx, dy, th, c["winf"], p)
        assert rel(o.get(), c["minplus"]) < FP64_REL
        K.max_logsumexp_spp_ct(gpu_ctx, o, dx, dy, p)
        assert rel(o.get(), c["maxcost"]) < 1e-14
        if "bary_rows" in c:
            b1, b2 = out(gpu_ctx, 3 * n), out(gpu_ctx, 3 * m)
            K.naive_findmaxindex_spp_ct(gpu_ctx, b1, dx, dy, th, r, lz, reg, 1.0, c["winf"], p)
            K.naive_findmaxindex_spp_ct_t(gpu_ctx, b2, dx, dy, r, th, lz, reg, 1.0, c["winf"], p)
            assert rel(b1.get().reshape(-1, 3), c["bary_rows"]) < 1e-11
            assert rel(b2.get().reshape(-1, 3), c["bary_cols"]) < 1e-11


def test_golden_sinkhorn(gpu_ctx, golden):
    from culamp_amd import kernels as K
    for c in golden:
        if c["kind"] != "sinkhorn":
            continue
        say("golden sinkhorn p", c["p"], "reg", c["reg"])
        x, y = unhex(c["x"], (-1, 3)), unhex(c["y"], (-1, 3))
        n = len(x)
        p = pval(c["p"])
        dx, dy = dev(gpu_ctx, x), dev(gpu_ctx, y)
        mg, phi, psi = dev(gpu_ctx, unhex(c["marg"])), dev(gpu_ctx, unhex(c["phi"])), dev(gpu_ctx, unhex(c["psi"]))
        o = out(gpu_ctx, n)
        K.warp_logsumexp_ct_fused(gpu_ctx, o, dx, dy, mg, psi, c["reg"], c["winf"], p)
        assert rel(o.get(), c["half"]) < FP64_REL
        if p in (1.0, 2.0):  # residual_opt / naive_findmaxindex_ct know p = 1 or squared L2
            res, cost = out(gpu_ctx, n), out(gpu_ctx, n)
            K.residual_opt(gpu_ctx, res, cost, dx, dy, mg, phi, psi, c["reg"], c["winf"], p)
            scale = max(1.0, float(np.max(np.abs(c["resid"]))))
            assert np.max(np.abs(res.get() - np.array(c["resid"]))) < 1e-12 * scale
            assert rel(cost.get(), c["rcost"]) < FP64_REL
            b = out(gpu_ctx, 3 * n)
            K.naive_findmaxindex_ct(gpu_ctx, b, dx, dy, phi, psi, c["reg"], c["winf"], p)
            assert rel(b.get().reshape(-1, 3), c["bary"]) < 1e-11


def test_golden_theta(gpu_ctx, golden):
    from culamp_amd import kernels as K
    for c in golden:
        if c["kind"] != "theta":
            continue
        m = len(c["out"])
        th = out(gpu_ctx, m)
        K.update_theta_residual(gpu_ctx, th, dev(gpu_ctx, unhex(c["theta0"])), dev(gpu_ctx, unhex(c["res"])),
                                dev(gpu_ctx, unhex(c["c"])), dev(gpu_ctx, unhex(c["etamu"])), c["eta_mu"],
                                bool(c["adjust"]), c["minv"], c["maxv"], 1.0)
        assert rel(th.get(), c["out"]) < 1e-13


def _inputs(seed, n, m):
    rng = np.random.default_rng(seed)
    x = rng.uniform(size=(n, 3))
    y = rng.uniform(size=(m, 3))
    th = rng.uniform(-1, 1, m)
    r = rng.uniform(0.1, 1, n)
    r /= r.sum()
    return x, y, th, r


SHAPES = [(1, 1), (3, 5), (7, 300), (257, 129), (1000, 777), (2048, 4096)]
LAUNCH = [(0, 0), (1, 1), (1, 3), (2, 17), (4, 5)]


@pytest.mark.parametrize("n,m", SHAPES)
@pytest.mark.parametrize("p", [1.0, 2.0, math.inf, 3.0])
def test_lse_and_colmarg_vs_oracle(gpu_ctx, oracle, n, m, p):
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    say("lse/colmarg n", n, "m", m, "p", p)
    x, y, th, r = _inputs(n * 7 + m, n, m)
    winf = float(oracle.max_cost_rows(x, y, p).max())
    reg, st = 0.02, 1.0
    ref = oracle.lse_rows(x, y, th, reg, st, winf, p)
    res_ref, cost_ref = oracle.col_marginal_cost(x, y, r, th, ref, reg, st, winf, p)
    dx, dy, dth, dr = dev(gpu_ctx, x), dev(gpu_ctx, y), dev(gpu_ctx, th), dev(gpu_ctx, r)
    dlz = dev(gpu_ctx, ref)
    o, res, cost = out(gpu_ctx, n), out(gpu_ctx, m), out(gpu_ctx, m)
    # exponent magnitude bound per row for the mixed tolerance
    vmax = (np.max(oracle.max_cost_rows(x, y, p)) / (2 * winf) + 1.0) / reg
    for rpt, splits in LAUNCH:
        gpu_ctx.set_launch(rpt, splits)
        try:
            K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, reg, st, winf, p, FP64)
            assert rel(o.get(), ref) < FP64_REL, (rpt, splits)
            K.residual_spp_c(gpu_ctx, res, cost, dx, dy, dr, dth, dlz, reg, st, winf, p, FP64)
            assert rel(res.get(), res_ref, 1e-300) < 1e-11, (rpt, splits)
            assert rel(cost.get(), cost_ref, 1e-300) < 1e-11, (rpt, splits)
            K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, reg, st, winf, p, MIXED)
            err = np.max(np.abs(o.get() - ref))
            assert err < MIXED_LSE_ULPS * vmax + 1e-9, (rpt, splits, err)
            K.residual_spp_c(gpu_ctx, res, cost, dx, dy, dr, dth, dlz, reg, st, winf, p, MIXED)
            # marginal sums: relative to the total mass, scaled by the exponent size
            tol = MIXED_SUM_REL * max(1.0, vmax * 2.0 ** -24 * 1e4)
            assert np.max(np.abs(res.get() - res_ref)) <= tol * max(res_ref.max(), 1e-300), (rpt, splits)
        finally:
            gpu_ctx.set_launch(0, 0)


def test_deterministic_repeat(gpu_ctx, oracle):
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    x, y, th, r = _inputs(11, 3000, 5000)
    dx, dy, dth = dev(gpu_ctx, x), dev(gpu_ctx, y), dev(gpu_ctx, th)
    o = out(gpu_ctx, 3000)
    for prec in (FP64, MIXED):
        K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, 0.01, 1.0, 3.0, 2.0, prec)
        a = o.get()
        for _ in range(3):
            K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, 0.01, 1.0, 3.0, 2.0, prec)
            assert np.array_equal(a, o.get())


@pytest.mark.parametrize("planar", [False, True])
@pytest.mark.parametrize("n,m", [(4096, 4096), (4100, 333), (700, 5000), (16, 9), (1, 7)])
def test_wide_layout_matches_split_layout(gpu_ctx, oracle, n, m, planar):
    """Source splits merged inside one workgroup (wide layout) and over
    workgroups with a merge launch (split layout) give identical bits for the
    same split length, for every kind the wide layout runs, both precisions
    and every cost, for 3-D points and for points with z = 0 (whose wide
    tiles carry the weight in the z slot); the wide FP64 row LSE also matches
    the oracle."""
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    say("wide vs split n", n, "m", m, "planar", planar)
    x, y, th, r = _inputs(3 * n + m, n, m)
    if planar:
        x[:, 2] = 0.0
        y[:, 2] = 0.0
    dx, dy, dth, dr = dev(gpu_ctx, x), dev(gpu_ctx, y), dev(gpu_ctx, th), dev(gpu_ctx, r)
    lz = dev(gpu_ctx, np.linspace(-1.0, 2.0, n))
    o, res, cost = out(gpu_ctx, n), out(gpu_ctx, m), out(gpu_ctx, m)

    def run_all(p):
        got = []
        for prec in (FP64, MIXED):
            K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, 0.02, 1.0, 3.0, p, prec)
            got.append(o.get())
            K.residual_spp_c(gpu_ctx, res, cost, dx, dy, dr, dth, lz, 0.02, 1.0, 3.0, p, prec)
            got += [res.get(), cost.get()]
        K.max_logsumexp_spp_ct(gpu_ctx, o, dx, dy, p)
        got.append(o.get())
        K.warp_min_reduce(gpu_ctx, o, dx, dy, dth, 3.0, p)
        got.append(o.get())
        return got

    try:
        for p in (1.0, 2.0, math.inf, 3.0):
            for splits in (16, 5):
                gpu_ctx.set_layout("split")
                gpu_ctx.set_launch(1, splits)
                want = run_all(p)
                gpu_ctx.set_layout("wide")
                gpu_ctx.set_launch(0, splits)
                for rep in range(2):
                    got = run_all(p)
                    for k, (w, g) in enumerate(zip(want, got)):
                        assert np.array_equal(w, g), (p, splits, rep, k)
            rows = np.arange(min(n, 24))
            ref = oracle.lse_rows(x[rows], y, th, 0.02, 1.0, 3.0, p)
            K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, 0.02, 1.0, 3.0, p, FP64)
            assert rel(o.get()[rows], ref) < FP64_REL, p
    finally:
        gpu_ctx.set_launch(0, 0)
        gpu_ctx.set_layout("auto")


@pytest.mark.parametrize("n,m", [(4096, 4096), (3000, 5000), (700, 333)])
def test_planar_tiles_match_general_loop(gpu_ctx, n, m):
    """Squared-Euclidean passes over points with z == 0 run a loop without the
    z terms.  Giving some sources (or one target) a z so small that its square
    vanishes in the pass's arithmetic (FP64: 1e-170; FP32: 1e-30) sends those
    tiles through the general loop instead; every output must keep its bits
    (wide and split layouts, FP64 and mixed, row LSE, column marginal and
    cost, Sinkhorn half step and residual, barycentric x / y)."""
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    say("planar tiles n", n, "m", m)
    rng = np.random.default_rng(n + m)
    x, y = rng.uniform(size=(n, 3)), rng.uniform(size=(m, 3))
    x[:, 2] = 0.0
    y[:, 2] = 0.0
    th = rng.uniform(-1, 1, m)
    r = rng.uniform(0.1, 1, n)
    r /= r.sum()
    lz = rng.uniform(-1, 2, n)
    phi = rng.normal(0, 1, n)
    dth, dr, dlz, dphi = dev(gpu_ctx, th), dev(gpu_ctx, r), dev(gpu_ctx, lz), dev(gpu_ctx, phi)
    o, res, cost = out(gpu_ctx, n), out(gpu_ctx, m), out(gpu_ctx, m)
    rres, rcost, b = out(gpu_ctx, n), out(gpu_ctx, n), out(gpu_ctx, 3 * n)

    def run(xa, ya, prec):
        dx, dy = dev(gpu_ctx, xa), dev(gpu_ctx, ya)
        got = []
        K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, 0.02, 1.0, 3.0, 2.0, prec)
        got.append(o.get())
        K.residual_spp_c(gpu_ctx, res, cost, dx, dy, dr, dth, dlz, 0.02, 1.0, 3.0, 2.0, prec)
        got += [res.get(), cost.get()]
        K.warp_logsumexp_ct_fused(gpu_ctx, o, dx, dy, dr, dth, 0.02, 3.0, 2.0, prec)
        got.append(o.get())
        K.residual_opt(gpu_ctx, rres, rcost, dx, dy, dr, dphi, dth, 0.02, 3.0, 2.0, prec)
        got += [rres.get(), rcost.get()]
        if prec == FP64:
            K.naive_findmaxindex_ct(gpu_ctx, b, dx, dy, dphi, dth, 0.02, 3.0, 2.0)
            got.append(b.get().reshape(-1, 3)[:, :2].copy())
        return got

    for prec, tiny in ((FP64, 1e-170), (MIXED, 1e-30)):
        want = run(x, y, prec)
        ys = y.copy()
        ys[::97, 2] = tiny
        xs = x.copy()
        xs[min(5, n - 1), 2] = tiny
        for label, (xa, ya) in (("sources", (x, ys)), ("target", (xs, y))):
            got = run(xa, ya, prec)
            for k, (w, g) in enumerate(zip(want, got)):
                assert np.array_equal(w, g), (prec, label, k)


@pytest.mark.parametrize("n,m", [(4096, 4096), (3000, 5000), (70, 33)])
def test_marginals_same_without_cost(gpu_ctx, n, m):
    """A marginal pass asked for no cost output skips the cost sum; its
    marginals (LAMP column pass, Sinkhorn row residual) keep their bits, for
    planar and 3-D points, FP64 and mixed, every cost."""
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    say("marginals without cost n", n, "m", m)
    rng = np.random.default_rng(7 * n + m)
    x, y = rng.uniform(size=(n, 3)), rng.uniform(size=(m, 3))
    th = rng.uniform(-1, 1, m)
    r = rng.uniform(0.1, 1, n)
    r /= r.sum()
    dth, dr = dev(gpu_ctx, th), dev(gpu_ctx, r)
    dlz, dphi = dev(gpu_ctx, rng.uniform(-1, 2, n)), dev(gpu_ctx, rng.normal(0, 1, n))
    res, cost, rres, rcost = out(gpu_ctx, m), out(gpu_ctx, m), out(gpu_ctx, n), out(gpu_ctx, n)
    for planar in (True, False):
        xa, ya = x.copy(), y.copy()
        if planar:
            xa[:, 2] = 0.0
            ya[:, 2] = 0.0
        dx, dy = dev(gpu_ctx, xa), dev(gpu_ctx, ya)
        for p in (2.0, 1.0, math.inf):
            for prec in (FP64, MIXED):
                K.residual_spp_c(gpu_ctx, res, cost, dx, dy, dr, dth, dlz, 0.02, 1.0, 3.0, p, prec)
                want = res.get()
                K.residual_spp_c(gpu_ctx, res, None, dx, dy, dr, dth, dlz, 0.02, 1.0, 3.0, p, prec)
                assert np.array_equal(res.get(), want), (planar, p, prec)
                K.residual_opt(gpu_ctx, rres, rcost, dx, dy, dr, dphi, dth, 0.02, 3.0, p, prec)
                want = rres.get()
                K.residual_opt(gpu_ctx, rres, None, dx, dy, dr, dphi, dth, 0.02, 3.0, p, prec)
                assert np.array_equal(rres.get(), want), (planar, p, prec)


def test_captured_graph_replays_match_eager(gpu_ctx, oracle):
    """Kernel-level calls captured into a graph: every replay equals the eager
    result, with new inputs uploaded and eager work run between the replays."""
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    x, y, th, r = _inputs(13, 2000, 3000)
    dx, dy, dth, dr = dev(gpu_ctx, x), dev(gpu_ctx, y), dev(gpu_ctx, th), dev(gpu_ctx, r)
    lz, res, cost = out(gpu_ctx, 2000), out(gpu_ctx, 3000), out(gpu_ctx, 3000)

    def step(prec):
        K.warp_logsumexp_spp_ct_fused(gpu_ctx, lz, dx, dy, dth, 0.02, 1.0, 3.0, 2.0, prec)
        K.residual_spp_c(gpu_ctx, res, cost, dx, dy, dr, dth, lz, 0.02, 1.0, 3.0, 2.0, prec)

    rng = np.random.default_rng(4)
    for prec in (FP64, MIXED):
        step(prec)  # work buffers exist before capture
        graph = gpu_ctx.capture(lambda: step(prec))
        for rep in range(4):
            th_r = th + rng.normal(0, 0.3, th.size)
            dth.upload(th_r)
            step(prec)
            want = (lz.get(), res.get(), cost.get())
            lz.upload(np.zeros(2000))
            res.upload(np.zeros(3000))
            graph.launch()
            got = (lz.get(), res.get(), cost.get())
            for w, g in zip(want, got):
                assert np.array_equal(w, g), (prec, rep)
            if prec == FP64:
                ref = oracle.lse_rows(x[:32], y, th_r, 0.02, 1.0, 3.0, 2.0)
                assert rel(got[0][:32], ref) < FP64_REL
        graph.close()


@pytest.mark.parametrize("n,m", [(1, 1), (5, 3), (700, 1100), (4096, 4096)])
@pytest.mark.parametrize("p", [1.0, 2.0, math.inf, 3.0])
def test_sinkhorn_kernels_vs_oracle(gpu_ctx, oracle, n, m, p):
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    say("sinkhorn kernels n", n, "m", m, "p", p)
    rng = np.random.default_rng(n + 3 * m)
    x, y = rng.uniform(size=(n, 3)), rng.uniform(size=(m, 3))
    marg = rng.uniform(0.1, 1, n)
    marg /= marg.sum()
    phi, psi = rng.normal(0, 3, n), rng.normal(0, 3, m)
    winf = float(oracle.max_cost_rows(x, y, p).max())
    reg = 0.03
    ref = oracle.sinkhorn_half_step(x, y, marg, psi, reg, winf, p)
    dx, dy, dm, dphi, dpsi = (dev(gpu_ctx, a) for a in (x, y, marg, phi, psi))
    o = out(gpu_ctx, n)
    K.warp_logsumexp_ct_fused(gpu_ctx, o, dx, dy, dm, dpsi, reg, winf, p, FP64)
    assert rel(o.get(), ref) < FP64_REL
    K.warp_logsumexp_ct_fused(gpu_ctx, o, dx, dy, dm, dpsi, reg, winf, p, MIXED)
    vmax = 1.0 / reg + np.max(np.abs(psi))
    assert np.max(np.abs(o.get() - ref)) < MIXED_LSE_ULPS * vmax + 1e-9
    pa = 1.0 if p == 1 else 2.0
    res_ref, cost_ref = oracle.row_residual_cost(x, y, marg, phi, psi, reg, winf, pa)
    res, cost = out(gpu_ctx, n), out(gpu_ctx, n)
    K.residual_opt(gpu_ctx, res, cost, dx, dy, dm, dphi, dpsi, reg, winf, p, FP64)
    scale = np.maximum(np.abs(res_ref), marg)
    assert np.max(np.abs(res.get() - res_ref) / scale) < 1e-10
    assert rel(cost.get(), cost_ref, 1e-300) < 1e-10
    b_ref = oracle.barycentric_sinkhorn(x, y, phi, psi, reg, winf, pa)
    b = out(gpu_ctx, 3 * n)
    K.naive_findmaxindex_ct(gpu_ctx, b, dx, dy, dphi, dpsi, reg, winf, p)
    assert np.max(np.abs(b.get().reshape(-1, 3) - b_ref)) < 1e-10


def test_edge_cases(gpu_ctx, oracle):
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    x, y, th, r = _inputs(5, 9, 6)
    dx, dy, dth = dev(gpu_ctx, x), dev(gpu_ctx, y), dev(gpu_ctx, th)
    o = out(gpu_ctx, 9)
    # reg = Inf (first LAMP iteration): log m exactly
    for prec in (FP64, MIXED):
        K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, math.inf, 1.0, 2.0, 2.0, prec)
        assert np.allclose(o.get(), np.log(6), rtol=1e-12, atol=0)
    # m = 0: empty log-sum-exp is -Inf
    import culamp_amd
    empty = culamp_amd.DeviceArray(gpu_ctx, 0)
    K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, empty, empty, 0.1, 1.0, 2.0, 2.0, FP64)
    assert np.all(np.isneginf(o.get()))
    # n = 0 is a no-op
    K.warp_logsumexp_spp_ct_fused(gpu_ctx, empty, empty, dy, dth, 0.1, 1.0, 2.0, 2.0, FP64)
    # invalid precision is reported, not run
    with pytest.raises(culamp_amd.CulampError):
        K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, dth, 0.1, 1.0, 2.0, 2.0, 7)
    # huge potentials: no overflow in either precision
    big = dev(gpu_ctx, np.array([-1e3, 5.0, 1e3, 0.0, 2.0, -7.0]))
    for prec in (FP64, MIXED):
        K.warp_logsumexp_spp_ct_fused(gpu_ctx, o, dx, dy, big, 0.01, 1.0, 2.0, 2.0, prec)
        ref = oracle.lse_rows(x, y, big.get(), 0.01, 1.0, 2.0, 2.0)
        assert np.all(np.isfinite(o.get()))
        assert np.max(np.abs(o.get() - ref) / np.abs(ref)) < (1e-12 if prec == FP64 else 1e-6)


def test_probe_rates(gpu_ctx):
    """The compute-rate probe returns finite, plausible device rates (lane
    operations per second): exponentials at most as fast as FMAs."""
    r = gpu_ctx.probe_rates()
    say("probe", r)
    assert 1e11 < r["exp_per_s"] < 1e15 and 1e11 < r["fma_per_s"] < 1e15
    assert r["exp_per_s"] <= 1.05 * r["fma_per_s"]
    # bf16 MFMA: between a tenth of and the published dense peak (2.5 PFLOP/s)
    m = gpu_ctx.probe_mfma()
    say("probe mfma", m / 1e12, "TFLOP/s")
    assert 2.5e14 < m <= 2.6e15

"""Sinkhorn divergence and its gradients.  The reference has no gradients, so
parity is pinned by mathematics: an independent dense numpy Sinkhorn for the
value (dual and primal forms), brute force for the adjoint kernel, central
finite differences for the gradients, the one-point closed form and the
zero of S(a, x; a, x)."""
import sys

import numpy as np
import pytest
from scipy.special import logsumexp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sqdist(x, y):
    return ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)


def dense_ot(x, a, y, b, eps, iters=20000, sym=False):
    """Log-domain Sinkhorn on the dense cost (symmetric averaged updates for a
    self problem); returns (dual value, primal value)."""
    K = -sqdist(x, y) / eps
    phi, psi = np.zeros(len(a)), np.zeros(len(b))
    for _ in range(iters):
        if sym:
            phi_new = 0.5 * (phi + np.log(a) - logsumexp(K + phi[None, :], axis=1))
            done = np.max(np.abs(phi_new - phi)) < 1e-15
            phi = psi = phi_new
        else:
            phi = np.log(a) - logsumexp(K + psi[None, :], axis=1)
            psi_new = np.log(b) - logsumexp(K + phi[:, None], axis=0)
            done = np.max(np.abs(psi_new - psi)) < 1e-15
            psi = psi_new
        if done:
            break
    P = np.exp(K + phi[:, None] + psi[None, :])
    dual = eps * (a @ phi + b @ psi - a @ np.log(a) - b @ np.log(b) - P.sum() + 1.0)
    primal = (P * sqdist(x, y)).sum() + eps * (P * (np.log(P) - np.log(np.outer(a, b)))).sum()
    return dual, primal


def clouds(seed, n, m):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3)) * 0.5
    y = rng.normal(size=(m, 3)) * 0.4 + 0.3
    a = rng.uniform(0.5, 1.5, n)
    b = rng.uniform(0.5, 1.5, m)
    return x, a / a.sum(), y, b / b.sum()


def test_lse_adjoint_kernel(gpu_ctx):
    import culamp_amd
    from culamp_amd import DeviceArray
    rng = np.random.default_rng(1)
    nt, ns, eps = 37, 53, 0.07
    t, s = rng.normal(size=(nt, 3)), rng.normal(size=(ns, 3))
    srcA, srcB = rng.normal(size=ns), rng.normal(size=ns)
    tgtA, tgtW = rng.normal(size=nt), rng.normal(size=nt)
    alpha, gamma, c0, c1, beta = 0.8, -0.6, 1.7, -0.7, 0.4
    V = -sqdist(t, s) / eps + alpha * srcA[None, :] + gamma * tgtA[:, None]
    E = np.exp(V)
    S = (E * srcB[None, :]).sum(1)
    Y = (E * srcB[None, :]) @ s
    o1_0, o0_0 = rng.normal(size=nt), rng.normal(size=(nt, 3))
    lib = culamp_amd.load()
    d = {k: DeviceArray.from_host(gpu_ctx, np.ascontiguousarray(v)) for k, v in
         dict(t=t, s=s, srcA=srcA, srcB=srcB, tgtA=tgtA, tgtW=tgtW, o1=o1_0, o0=o0_0).items()}
    for splits in (0, 1, 4):
        gpu_ctx.set_launch(0, splits)
        try:
            d["o1"].upload(o1_0); d["o0"].upload(o0_0)
            culamp_amd._lib.check(lib.culamp_lse_adjoint(
                gpu_ctx.handle, d["t"].ptr, nt, d["s"].ptr, ns, eps, d["srcA"].ptr, alpha,
                d["srcB"].ptr, d["tgtA"].ptr, gamma, 0, None, c0, c1, beta, d["o0"].ptr, d["o1"].ptr))
            want1 = beta * o1_0 + c1 * S
            want0 = o0_0 + c0 * (t * S[:, None] - Y)
            assert np.allclose(d["o1"].get(), want1, rtol=1e-12, atol=1e-12 * np.abs(want1).max())
            assert np.allclose(d["o0"].get().reshape(nt, 3), want0, rtol=1e-12,
                               atol=1e-12 * np.abs(want0).max())
            # normalized: out0 += c0 tgtW (t - Y/S) with unit source weights
            d["o0"].upload(o0_0)
            culamp_amd._lib.check(lib.culamp_lse_adjoint(
                gpu_ctx.handle, d["t"].ptr, nt, d["s"].ptr, ns, eps, d["srcA"].ptr, alpha, None,
                None, 0.0, 1, d["tgtW"].ptr, c0, 0.0, 0.0, d["o0"].ptr, None))
            E1 = np.exp(-sqdist(t, s) / eps + alpha * srcA[None, :])
            bary = (E1 @ s) / E1.sum(1)[:, None]
            want0 = o0_0 + c0 * tgtW[:, None] * (t - bary)
            assert np.allclose(d["o0"].get().reshape(nt, 3), want0, rtol=1e-11, atol=1e-12)
        finally:
            gpu_ctx.set_launch(0, 0)


def test_divergence_value_vs_dense(gpu_ctx):
    import culamp_amd
    x, a, y, b = clouds(2, 60, 45)
    eps = 0.1
    for prec, tol in (("fp64", 1e-9), ("mixed", 1e-5)):
        val, terms = culamp_amd.sinkhorn_divergence(x, a, y, b, grad=False, ctx=gpu_ctx, eps=eps,
                                                     tol=1e-12, precision=prec, return_terms=True)
        ab = dense_ot(x, a, y, b, eps)
        aa = dense_ot(x, a, x, a, eps, sym=True)
        bb = dense_ot(y, b, y, b, eps, sym=True)
        for (dual, primal), t in zip((ab, aa, bb), terms):
            # (tol 1e-12 is below the mixed forward's FP32 floor: its stall
            # guard hands the solve to FP64, which reaches it)
            assert t["converged"] == 1, (prec, t)
            if prec == "mixed":
                assert t["fp64_from"] > 0
            assert abs(dual - primal) < 1e-10 * abs(primal) + 1e-13  # dual value == primal value
            assert abs(t["objective"] - dual) < tol * abs(dual) + 1e-12, (prec, t["objective"], dual)
        want = ab[0] - 0.5 * aa[0] - 0.5 * bb[0]
        say(prec, "S", val, want)
        assert abs(val - want) < tol * max(abs(ab[0]), 1e-12)


def fd_check(gpu_ctx, x, a, y, b, eps, g_x, g_a, coords, pairs, h=1e-3, h_a=1e-6):
    import culamp_amd
    kw = dict(grad=False, ctx=gpu_ctx, eps=eps, tol=1e-13, precision="fp64", itermax=20000)
    fd_x = []
    for q, (i, k) in enumerate(coords):
        if q % 8 == 0:
            say("  fd x", q)
        xp, xm = x.copy(), x.copy()
        xp[i, k] += h
        xm[i, k] -= h
        fd_x.append((culamp_amd.sinkhorn_divergence(xp, a, y, b, **kw) -
                     culamp_amd.sinkhorn_divergence(xm, a, y, b, **kw)) / (2 * h))
    fd_x = np.array(fd_x)
    gx = np.array([g_x[i, k] for i, k in coords])
    # mass-preserving perturbations of a: a_i + h, a_j - h -> g_a[i] - g_a[j]; the
    # weights are ~1/n, so the step is 1e-6 (at 1e-3 the central difference's own
    # truncation error, from the a log a terms, is ~1e-4 relative)
    fd_a = []
    for i, j in pairs:
        ap, am = a.copy(), a.copy()
        ap[i] += h_a; ap[j] -= h_a
        am[i] -= h_a; am[j] += h_a
        fd_a.append((culamp_amd.sinkhorn_divergence(x, ap, y, b, **kw) -
                     culamp_amd.sinkhorn_divergence(x, am, y, b, **kw)) / (2 * h_a))
    fd_a = np.array(fd_a)
    ga = np.array([g_a[i] - g_a[j] for i, j in pairs])
    ex = np.max(np.abs(gx - fd_x)) / np.max(np.abs(fd_x))
    ea = np.max(np.abs(ga - fd_a)) / np.max(np.abs(fd_a))
    return ex, ea


def test_divergence_gradient_vs_finite_differences(gpu_ctx):
    """Central differences (FP64 forward to infeasibility 1e-13; h = 1e-3 on 32
    random coordinates of x, h = 1e-6 on 16 mass-preserving pairs of a); the
    gradient (mixed forward to 1e-6 then FP64 to 1e-9, unrolled adjoint over
    the last 64 iterations, checkpoints every 16; and the FP64 forward) within
    1e-4 relative, also without the FP64 stage."""
    import culamp_amd
    x, a, y, b = clouds(3, 300, 260)
    eps = 0.05
    rng = np.random.default_rng(4)
    coords = [(int(rng.integers(300)), int(rng.integers(3))) for _ in range(32)]
    pairs = [tuple(int(v) for v in rng.choice(300, 2, replace=False)) for _ in range(16)]
    for prec, polish, tol_x, tol_a in (("mixed", 1e-9, 1e-4, 1e-4), ("fp64", 0.0, 1e-4, 1e-4),
                                       ("mixed", 0.0, 1e-4, 1e-4)):
        (val, gx, ga, gy, gb), terms = culamp_amd.sinkhorn_divergence(
            x, a, y, b, ctx=gpu_ctx, eps=eps, tol=1e-6, itermax=5000, precision=prec, unroll=64,
            checkpoint=16, polish_tol=polish, return_terms=True)
        say(prec, polish, "terms", [(t["iterations"], t["residual"]) for t in terms])
        assert all(t["converged"] for t in terms)
        ex, ea = fd_check(gpu_ctx, x, a, y, b, eps, gx, ga, coords, pairs)
        say(prec, polish, "grad rel err x", ex, "a", ea)
        assert ex < tol_x and ea < tol_a
    # y side by symmetry: S(a,x;b,y) = S(b,y;a,x)
    _, gx2, ga2, gy2, gb2 = culamp_amd.sinkhorn_divergence(y, b, x, a, ctx=gpu_ctx, eps=eps,
                                                           tol=1e-10, precision="fp64", itermax=20000)
    _, gx1, ga1, gy1, gb1 = culamp_amd.sinkhorn_divergence(x, a, y, b, ctx=gpu_ctx, eps=eps,
                                                           tol=1e-10, precision="fp64", itermax=20000)
    assert np.allclose(gy1, gx2, atol=1e-8 * np.abs(gx2).max())
    assert np.allclose(gb1 - gb1.mean(), ga2 - ga2.mean(), atol=1e-8 * np.abs(ga2).max())


def test_unrolled_adjoint_corrects_the_envelope(gpu_ctx):
    """Stopped early (infeasibility 1e-4), the envelope gradient misses the
    potentials' sensitivity; back-propagating through the iterations brings
    the gradient closer to the converged one, and checkpointing (recomputing
    segments forward) changes nothing."""
    import culamp_amd
    x, a, y, b = clouds(5, 200, 180)
    eps = 0.05
    ref = culamp_amd.sinkhorn_divergence(x, a, y, b, ctx=gpu_ctx, eps=eps, tol=1e-13,
                                         precision="fp64", itermax=20000)[1]
    errs = {}
    for unroll, ck in ((0, 16), (64, 16), (100000, 16), (100000, 1)):
        g = culamp_amd.sinkhorn_divergence(x, a, y, b, ctx=gpu_ctx, eps=eps, tol=1e-4, frequency=1,
                                           precision="fp64", unroll=unroll, checkpoint=ck,
                                           itermax=20000)[1]
        errs[(unroll, ck)] = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
    say("errors", errs)
    assert errs[(64, 16)] < 0.5 * errs[(0, 16)]
    # full back-propagation from the start: every checkpoint interval gives the
    # same gradient (segments are recomputed bit for bit)
    assert errs[(100000, 16)] == errs[(100000, 1)]


def test_single_point_closed_form_and_zero(gpu_ctx):
    import culamp_amd
    rng = np.random.default_rng(6)
    x = rng.normal(size=(1, 3))
    y = rng.normal(size=(40, 3))
    b = rng.dirichlet(np.ones(40))
    val, gx, ga, gy, gb = culamp_amd.sinkhorn_divergence(x, np.ones(1), y, b, ctx=gpu_ctx,
                                                         eps=0.3, tol=1e-13, precision="fp64",
                                                         itermax=20000)
    want = 2 * (b[:, None] * (x - y)).sum(0)
    assert np.allclose(gx[0], want, rtol=1e-10, atol=1e-12)
    # S(a, x; a, x) = 0 and its gradient vanishes
    x, a, _, _ = clouds(7, 80, 10)
    val, gx, ga, gy, gb = culamp_amd.sinkhorn_divergence(x, a, x, a, ctx=gpu_ctx, eps=0.5,
                                                         tol=1e-13, precision="fp64",
                                                         itermax=20000)
    assert abs(val) < 1e-12
    assert np.max(np.abs(gx)) < 1e-9 and np.max(np.abs(ga - ga.mean())) < 1e-9


def test_autograd_function(gpu_ctx):
    import torch
    import culamp_amd
    x, a, y, b = clouds(8, 50, 40)
    X = torch.tensor(x, requires_grad=True)
    A = torch.tensor(a, requires_grad=True)
    S = culamp_amd.diff.SinkhornDivergence.apply(X, A, torch.tensor(y), torch.tensor(b),
                                                 dict(ctx=gpu_ctx, eps=0.1, tol=1e-12,
                                                      precision="fp64", itermax=20000))
    (3.0 * S).backward()
    val, gx, ga, _, _ = culamp_amd.sinkhorn_divergence(x, a, y, b, ctx=gpu_ctx, eps=0.1, tol=1e-12,
                                                       precision="fp64", itermax=20000)
    assert abs(S.item() - val) < 1e-14
    assert np.allclose(X.grad.numpy(), 3.0 * gx) and np.allclose(A.grad.numpy(), 3.0 * ga)

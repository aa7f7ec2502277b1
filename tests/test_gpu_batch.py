"""Batched solves against the oracle's single-instance restatement of the
reference loops: every instance of a batch must stop at the oracle's
iteration with the oracle's outputs (FP64), whatever the other instances do;
mixed and bf16 batches must reach the stop test and state their error."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hist(rng, n):
    a = rng.uniform(size=n) + 1e-3
    return a / a.sum()


def point_instances(seed, sizes):
    rng = np.random.default_rng(seed)
    out = []
    for n, m in sizes:
        x, y = rng.uniform(size=(n, 3)), rng.uniform(size=(m, 3))
        out.append((x, y, hist(rng, n), hist(rng, m)))
    return out


SIZES = [(50, 37), (90, 61), (64, 64), (120, 30), (17, 200), (1, 5), (300, 260)]


@pytest.mark.parametrize("p", [1.0, 2.0, math.inf])
@pytest.mark.parametrize("graph", [1, 0])
def test_batch_sinkhorn_points_fp64(gpu_ctx, oracle, p, graph):
    import culamp_amd
    say("batch sinkhorn points p", p, "graph", graph)
    inst = point_instances(1, SIZES)
    b = culamp_amd.Batch(inst, gpu_ctx)
    b.set_tuning(graph=graph)
    args = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-8, itermax=20000, verbose=False)
    out = b.sinkhorn(args, 10, p)
    for k, ((x, y, r, c), (phi, psi, info)) in enumerate(zip(inst, out)):
        o = oracle.sinkhorn_solve(x, y, r, c, frequency=10, p=p, eta_p=1e-2, epsilon=6e-8,
                                  itermax=20000, verbose=False)
        assert info["iterations"] == o["iterations"], (k, info["iterations"], o["iterations"])
        assert info["converged"] == o["converged"]
        assert np.allclose(phi, o["phi"], rtol=1e-9, atol=1e-9), k
        assert np.allclose(psi, o["psi"], rtol=1e-9, atol=1e-9), k
        assert abs(info["objective"] - o["objective"]) <= 1e-9 * abs(o["objective"]) + 1e-14
        assert abs(info["residual"] - o["residual"]) <= 1e-9 * o["residual"] + 1e-14
    b.close()


def test_batch_sinkhorn_forced_splits_and_itermax(gpu_ctx, oracle):
    """Source splits (merge kernel on the batch tables), warm starts, and an
    itermax that stops some instances before they converge."""
    import culamp_amd
    inst = point_instances(2, [(700, 900), (300, 1100), (40, 33)])
    b = culamp_amd.Batch(inst, gpu_ctx)
    b.set_tuning(max_split=256)
    rng = np.random.default_rng(3)
    phi0 = [rng.normal(0, 0.1, x.shape[0]) for x, _, _, _ in inst]
    psi0 = [rng.normal(0, 0.1, y.shape[0]) for _, y, _, _ in inst]
    args = culamp_amd.EOTArgs(eta_p=5e-3, epsilon=6e-9, itermax=57, verbose=False)
    out = b.sinkhorn(args, 5, 2.0, phi0=phi0, psi0=psi0)
    for k, ((x, y, r, c), (phi, psi, info)) in enumerate(zip(inst, out)):
        o = oracle.sinkhorn_solve(x, y, r, c, frequency=5, p=2.0, phi0=phi0[k], psi0=psi0[k],
                                  eta_p=5e-3, epsilon=6e-9, itermax=57, verbose=False)
        say("  inst", k, info["iterations"], o["iterations"], info["converged"])
        assert info["iterations"] == o["iterations"] and info["converged"] == o["converged"]
        assert np.allclose(phi, o["phi"], atol=1e-9) and np.allclose(psi, o["psi"], atol=1e-9)
    b.close()


@pytest.mark.parametrize("eta,stop,itermax,verbose", [(1e-2, "infeas", 20000, False),
                                                      (0.0, "reference", 800, True),
                                                      (1e-2, "reference", 150, False)])
def test_batch_lamp_points_fp64(gpu_ctx, oracle, eta, stop, itermax, verbose):
    import culamp_amd
    say("batch lamp points", eta, stop)
    inst = point_instances(4, SIZES[:5])
    b = culamp_amd.Batch(inst, gpu_ctx)
    kw = dict(eta_p=eta, epsilon=6e-6, B=1.1, itermax=itermax, verbose=verbose, stop=stop)
    out = b.lamp(culamp_amd.EOTArgs(**kw), 10, 2.0)
    for k, ((x, y, r, c), (nu, phi, psi, info)) in enumerate(zip(inst, out)):
        o = oracle.lamp_solve(x, y, r, c, frequency=10, p=2.0, **kw)
        assert info["iterations"] == o["iterations"], (k, info["iterations"], o["iterations"])
        assert info["converged"] == o["converged"]
        assert np.allclose(nu, o["nu"], rtol=0, atol=1e-8), k
        assert np.allclose(phi, o["phi"], rtol=1e-8, atol=1e-8), k
        if eta > 0:
            assert np.allclose(psi, o["psi"], rtol=1e-7, atol=1e-7), k
        if stop == "infeas" or verbose:
            for f in ("residual", "objective", "primal", "dual"):
                assert abs(info[f] - o[f]) <= 1e-7 * abs(o[f]) + 1e-12, (k, f, info[f], o[f])
    b.close()


def cells(rng, n, d, density=0.5):
    return np.abs(rng.normal(size=(n, d))) * (rng.uniform(size=(n, d)) < density) + 1e-3


@pytest.mark.parametrize("metric", ["l1", "l2", "cosine", "correlation"])
def test_batch_cellsim_fp64(gpu_ctx, oracle, metric):
    import culamp_amd
    say("batch cellsim", metric)
    rng = np.random.default_rng(20)
    scale = culamp_amd.cost_scale(metric)
    sizes = [(60, 45), (33, 80), (70, 70)]
    inst = []
    for n, m in sizes:
        inst.append((cells(rng, n, 16), cells(rng, m, 16), rng.dirichlet(np.ones(n)),
                     rng.dirichlet(np.ones(m))))
    b = culamp_amd.Batch(inst, gpu_ctx, metric=metric, scale=scale)
    out = b.sinkhorn(culamp_amd.EOTArgs(eta_p=0.05, epsilon=6e-8, verbose=False), 5)
    for k, ((x, y, r, c), (phi, psi, info)) in enumerate(zip(inst, out)):
        o = oracle.sinkhorn_cellsim_solve(x, y, r, c, metric, scale, frequency=5, eta_p=0.05,
                                          epsilon=6e-8, verbose=False)
        assert info["iterations"] == o["iterations"], (k, info["iterations"], o["iterations"])
        assert np.allclose(phi, o["phi"], atol=1e-9) and np.allclose(psi, o["psi"], atol=1e-9)
    b.close()
    # LAMP cellsim: the reference's extragradient_cellsim takes one data matrix
    inst = []
    for n in (60, 41, 75):
        inst.append((cells(rng, n, 16), None, rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))))
    b = culamp_amd.Batch(inst, gpu_ctx, metric=metric, scale=scale)
    for eta, stop, itermax in ((1e-2, "infeas", 20000), (0.0, "reference", 300)):
        kw = dict(eta_p=eta, epsilon=6e-6, B=1.1, itermax=itermax, verbose=True, stop=stop)
        out = b.lamp(culamp_amd.EOTArgs(**kw), 10)
        for k, ((x, _, r, c), (nu, phi, psi, info)) in enumerate(zip(inst, out)):
            o = oracle.lamp_cellsim_solve(x, r, c, metric, scale, frequency=10, **kw)
            assert info["iterations"] == o["iterations"], (eta, k, info["iterations"], o["iterations"])
            assert np.allclose(nu, o["nu"], atol=1e-8), (eta, k)
            assert np.allclose(phi, o["phi"], rtol=1e-8, atol=1e-8), (eta, k)
            assert abs(info["objective"] - o["objective"]) <= 1e-9 * abs(o["objective"]) + 1e-14
    b.close()
    # LAMP cellsim between two point sets (the batched cell-similarity
    # configuration's instances), against the oracle's two-matrix loop
    inst = []
    for n, m in ((60, 45), (33, 70)):
        inst.append((cells(rng, n, 16), cells(rng, m, 16), rng.dirichlet(np.ones(n)),
                     rng.dirichlet(np.ones(m))))
    b = culamp_amd.Batch(inst, gpu_ctx, metric=metric, scale=scale)
    kw = dict(eta_p=5e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False, stop="infeas")
    out = b.lamp(culamp_amd.EOTArgs(**kw), 10)
    for k, ((x, y, r, c), (nu, phi, psi, info)) in enumerate(zip(inst, out)):
        o = oracle.lamp_cellsim2_solve(x, y, r, c, metric, scale, frequency=10, **kw)
        assert info["iterations"] == o["iterations"] > 0, (k, info["iterations"], o["iterations"])
        assert np.allclose(nu, o["nu"], atol=1e-8), k
        assert abs(info["objective"] - o["objective"]) <= 1e-9 * abs(o["objective"]) + 1e-14
    b.close()


def unit_cells(rng, n, d=128, density=0.3):
    x = np.abs(rng.normal(size=(n, d))) * (rng.uniform(size=(n, d)) < density)
    x[x.sum(1) == 0, 0] = 1.0
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def test_batch_bf16_cellsim_and_autotune(gpu_ctx, oracle):
    """Cosine on d = 128 unit rows through the batched MFMA kernel: every
    instance reaches infeasibility 1e-6 under its own arithmetic and its cost
    is within the bf16x3 / bf16 bounds of the FP64 oracle; the autotuned
    shape gives the same iterations."""
    import culamp_amd
    rng = np.random.default_rng(41)
    sizes = [(700, 500), (300, 900), (256, 256), (1200, 333)]
    inst = [(unit_cells(rng, n), unit_cells(rng, m), rng.dirichlet(np.ones(n)),
             rng.dirichlet(np.ones(m))) for n, m in sizes]
    b = culamp_amd.Batch(inst, gpu_ctx, metric="cosine", scale=2.0)
    refs = [oracle.sinkhorn_cellsim_solve(x, y, r, c, "cosine", 2.0, frequency=10, eta_p=5e-2,
                                          epsilon=6e-6, itermax=20000, verbose=False)
            for x, y, r, c in inst]
    # bf16: the cost moves ~1.4e-5 on these instances and up to 8e-5 over the
    # 256 cell-similarity instances of the bench; bf16x3 ~4e-8
    for prec, tol in (("bf16x3", 1e-6), ("bf16", 2e-4)):
        args = culamp_amd.EOTArgs(eta_p=5e-2, epsilon=6e-6, itermax=20000, verbose=False,
                                  precision=prec)
        out = b.sinkhorn(args, 10)
        for k, (phi, psi, info) in enumerate(out):
            o = refs[k]
            say(prec, k, info["iterations"], info["residual"], info["objective"], "fp64",
                o["iterations"], o["objective"])
            assert info["converged"] == 1 and info["residual"] <= 1e-6
            assert abs(info["objective"] - o["objective"]) / o["objective"] < tol
        if prec == "bf16x3":
            first = [i["iterations"] for _, _, i in out]
    tune, ms = b.autotune("bf16x3")
    say("autotune", tune, ms)
    assert ms > 0 and tune["max_split"] > 0
    args = culamp_amd.EOTArgs(eta_p=5e-2, epsilon=6e-6, itermax=20000, verbose=False,
                              precision="bf16x3")
    again = b.sinkhorn(args, 10)
    assert [i["iterations"] for _, _, i in again] == first
    lout = b.lamp(args.copy(stop="infeas", B=1.1), 10)
    for k, (nu, phi, psi, info) in enumerate(lout):
        assert info["converged"] == 1 and info["residual"] < 1e-6, (k, info)
    b.close()


def test_batch_mixed_points(gpu_ctx, oracle):
    import culamp_amd
    inst = point_instances(9, [(600, 500), (257, 999), (1024, 100)])
    b = culamp_amd.Batch(inst, gpu_ctx)
    args = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False,
                              precision="mixed", stop="infeas")
    for solver in ("sinkhorn", "lamp"):
        out = b.sinkhorn(args, 10, 2.0) if solver == "sinkhorn" else b.lamp(args, 10, 2.0)
        for k, (x, y, r, c) in enumerate(inst):
            info = out[k][-1]
            fn = oracle.sinkhorn_solve if solver == "sinkhorn" else oracle.lamp_solve
            kw = dict(frequency=10, p=2.0, eta_p=1e-2, epsilon=6e-6, itermax=20000, verbose=False)
            if solver == "lamp":
                kw.update(B=1.1, stop="infeas")
            o = fn(x, y, r, c, **kw)
            say(solver, k, info["iterations"], info["residual"], "fp64", o["iterations"])
            assert info["converged"] == 1 and info["residual"] <= 1e-6
            assert abs(info["objective"] - o["objective"]) / o["objective"] < 1e-5
    b.close()


def test_cellsim_lamp_ragged_single_vs_batch(gpu_ctx):
    """extragradient_cellsim generalised to x != y (n != m): the single-instance
    solver and the batch agree (FP64), including phi (length n)."""
    import culamp_amd
    rng = np.random.default_rng(30)
    x, y = cells(rng, 70, 12), cells(rng, 45, 12)
    r, c = rng.dirichlet(np.ones(70)), rng.dirichlet(np.ones(45))
    kw = dict(eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False, stop="infeas")
    pb = culamp_amd.FeatureProblem(x, y, 2.0, "cosine", r, c, gpu_ctx)
    nu, phi, psi, info = pb.lamp(culamp_amd.EOTArgs(**kw), 10, log=None)
    b = culamp_amd.Batch([(x, y, r, c)], gpu_ctx, metric="cosine", scale=2.0)
    (bnu, bphi, bpsi, binfo), = b.lamp(culamp_amd.EOTArgs(**kw), 10)
    assert info["converged"] == 1 and info["iterations"] == binfo["iterations"]
    assert phi.shape == (70,) and np.all(np.isfinite(phi))
    assert np.allclose(nu, bnu, atol=1e-12) and np.allclose(phi, bphi, atol=1e-10)
    assert np.allclose(psi, bpsi, atol=1e-10)
    b.close()
    pb.close()

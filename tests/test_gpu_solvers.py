"""Solver-level parity: libculamp's LAMP, Sinkhorn and annealed Sinkhorn loops
against the oracle's restatement of the reference loops on the same seeded
inputs: same stop iteration, same log lines (time column excepted), same
outputs.  Mixed precision is checked against the FP64 oracle with a stated
tolerance."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def grid(g):
    k = np.arange(g * g)
    return np.stack([(k // g) / (g - 1), (k % g) / (g - 1), np.zeros(g * g)], axis=1)


def hist(seed, n):
    rng = np.random.default_rng(seed)
    a = rng.uniform(size=n) + 1e-3
    return a / a.sum()


def compare_logs(gpu_lines, ref_lines, rtol):
    assert len(gpu_lines) == len(ref_lines), (len(gpu_lines), len(ref_lines))
    assert gpu_lines[0] == ref_lines[0]
    for a, b in zip(gpu_lines[1:], ref_lines[1:]):
        fa, fb = a.rstrip("\n").split(","), b.rstrip("\n").split(",")
        assert fa[1] == fb[1] and fa[-1] == fb[-1], (a, b)
        for u, v in zip(fa[2:-1], fb[2:-1]):
            u, v = float(u), float(v)
            assert abs(u - v) <= rtol * max(abs(v), 1e-12) + 1e-13, (a, b)


def run_lamp(pb, **kw):
    import culamp_amd
    lines = []
    args = culamp_amd.EOTArgs(**{k: v for k, v in kw.items() if k not in ("frequency", "p")})
    nu, phi, psi, a1, a2, info = pb.lamp(args, kw.get("frequency", 10), kw.get("p", 2.0), lines.append)
    return dict(nu=nu, phi=phi, psi=psi, img1=a1, img2=a2, log=lines, **info)


@pytest.mark.parametrize("p", [1.0, 2.0, math.inf])
def test_lamp_fp64_matches_oracle(gpu_ctx, oracle, p):
    import culamp_amd
    say("lamp fp64 p", p)
    x = grid(8)
    r, c = hist(1, 64), hist(2, 64)
    pb = culamp_amd.Problem(x, x, r, c, gpu_ctx)
    kw = dict(eta_p=0.0, epsilon=6e-6, B=1.1, itermax=20000, verbose=True, frequency=10, p=p)
    g = run_lamp(pb, **kw)
    o = oracle.lamp_solve(x, x, r, c, **kw)
    assert g["iterations"] == o["iterations"] and g["converged"] == o["converged"] == 1
    compare_logs(g["log"], o["log"], 1e-7)
    assert np.allclose(g["nu"], o["nu"], rtol=0, atol=1e-8)
    assert np.allclose(g["phi"], o["phi"], rtol=1e-8, atol=1e-8)
    assert np.allclose(g["img1"], o["img1"], atol=1e-8)
    assert np.allclose(g["img2"], o["img2"], atol=1e-8)


def test_lamp_fp64_eta_positive_ragged(gpu_ctx, oracle):
    """N != M, eta_p > 0, the infeasibility stop, point clouds."""
    import culamp_amd
    rng = np.random.default_rng(7)
    x, y = rng.uniform(size=(50, 3)), rng.uniform(size=(37, 3))
    r, c = hist(3, 50), hist(4, 37)
    pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
    kw = dict(eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=True, frequency=25, p=2.0,
              stop="infeas")
    g = run_lamp(pb, **kw)
    o = oracle.lamp_solve(x, y, r, c, **kw)
    assert g["iterations"] == o["iterations"] and g["converged"] == 1
    compare_logs(g["log"], o["log"], 1e-7)
    assert np.allclose(g["psi"], o["psi"], rtol=1e-7, atol=1e-7)
    # reference-mode run for a fixed number of iterations (its stop never fires for eta_p > 0)
    kw2 = dict(kw, stop="reference", itermax=300)
    g2, o2 = run_lamp(pb, **kw2), oracle.lamp_solve(x, y, r, c, **kw2)
    assert g2["iterations"] == o2["iterations"] == 300
    compare_logs(g2["log"], o2["log"], 1e-7)


@pytest.mark.parametrize("p", [1.0, 2.0, math.inf, 3.0])
def test_sinkhorn_fp64_matches_oracle(gpu_ctx, oracle, p):
    import culamp_amd
    say("sinkhorn fp64 p", p)
    x = grid(8)
    r, c = hist(5, 64), hist(6, 64)
    pb = culamp_amd.Problem(x, x, r, c, gpu_ctx)
    args = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, itermax=20000, verbose=True)
    lines = []
    phi, psi, a1, a2, info = pb.sinkhorn(args, 10, p, log=lines.append)
    o = oracle.sinkhorn_solve(x, x, r, c, frequency=10, p=p, eta_p=1e-2, epsilon=6e-6, itermax=20000,
                              verbose=True)
    assert info["iterations"] == o["iterations"] and info["converged"] == o["converged"]
    compare_logs(lines, o["log"], 1e-7)
    assert np.allclose(phi, o["phi"], rtol=1e-8, atol=1e-8)
    assert np.allclose(psi, o["psi"], rtol=1e-8, atol=1e-8)
    assert np.allclose(a1, o["img1"], atol=1e-8) and np.allclose(a2, o["img2"], atol=1e-8)


def test_sinkhorn_warm_start_and_ragged(gpu_ctx, oracle):
    import culamp_amd
    rng = np.random.default_rng(8)
    x, y = rng.uniform(size=(90, 3)), rng.uniform(size=(61, 3))
    r, c = hist(9, 90), hist(10, 61)
    phi0, psi0 = rng.normal(0, 0.1, 90), rng.normal(0, 0.1, 61)
    out = culamp_amd.sinkhorn_color_transfer(x, y, r, c, culamp_amd.EOTArgs(eta_p=0.05, epsilon=6e-8, verbose=False),
                                             5, 2.0, return_assignments=False, return_cuda=True,
                                             φ0=phi0, ψ0=psi0, ctx=gpu_ctx, log=None)
    o = oracle.sinkhorn_solve(x, y, r, c, frequency=5, p=2.0, phi0=phi0, psi0=psi0, eta_p=0.05,
                              epsilon=6e-8, verbose=False)
    assert out[4] == o["iterations"]
    assert np.allclose(out[0], o["phi"], atol=1e-9) and np.allclose(out[1], o["psi"], atol=1e-9)
    assert abs(out[2] - o["objective"]) < 1e-10 and abs(out[3] - o["residual"]) < 1e-12


def test_annealed_sinkhorn_matches_oracle(gpu_ctx, oracle):
    import culamp_amd
    x = grid(6)
    r, c = hist(11, 36), hist(12, 36)
    args = culamp_amd.EOTArgs(eta_p=0.1, epsilon=5e-3, anneal_mult=0.5, itermax=20000, verbose=True)
    lines = []
    phi, psi, obj, info = culamp_amd.annealed_sinkhorn_euclidean(r, c, x, x, "", "", "", args, 10, 2.0,
                                                                  log=lines.append, ctx=gpu_ctx,
                                                                  return_info=True)
    o = oracle.annealed_sinkhorn_solve(x, x, r, c, frequency=10, p=2.0, eta_p=0.1, epsilon=5e-3,
                                       anneal_mult=0.5, itermax=20000, verbose=True)
    assert info["iterations"] == o["iterations"]
    compare_logs(lines, o["log"], 1e-7)
    assert np.allclose(phi, o["phi"], atol=1e-8) and abs(obj - o["objective"]) < 1e-10


def test_mixed_precision_solvers(gpu_ctx, oracle):
    """Mixed precision reaches infeasibility 1e-6 by its own (FP64) check, and
    its transport cost stays within 1e-5 relative of the FP64 oracle's."""
    import culamp_amd
    rng = np.random.default_rng(13)
    x, y = rng.uniform(size=(600, 3)), rng.uniform(size=(500, 3))
    r, c = hist(14, 600), hist(15, 500)
    pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
    args = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, itermax=20000, verbose=False, precision="mixed")
    phi, psi, _, _, info = pb.sinkhorn(args, 10, 2.0, log=None, assignments=False)
    o = oracle.sinkhorn_solve(x, y, r, c, frequency=10, p=2.0, eta_p=1e-2, epsilon=6e-6,
                              itermax=20000, verbose=False)
    say("sinkhorn mixed", info["iterations"], info["residual"], "fp64", o["iterations"], o["residual"])
    assert info["converged"] == 1 and info["residual"] <= 1e-6
    assert abs(info["objective"] - o["objective"]) / o["objective"] < 1e-5
    largs = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False,
                               precision="mixed", stop="infeas")
    nu, lphi, lpsi, _, _, linfo = pb.lamp(largs, 10, 2.0, log=None, assignments=False)
    lo = oracle.lamp_solve(x, y, r, c, frequency=10, p=2.0, eta_p=1e-2, epsilon=6e-6, B=1.1,
                           itermax=20000, verbose=False, stop="infeas")
    say("lamp mixed", linfo["iterations"], linfo["residual"], "fp64", lo["iterations"], lo["residual"])
    assert linfo["converged"] == 1 and linfo["residual"] <= 1e-6
    assert abs(linfo["objective"] - lo["objective"]) / lo["objective"] < 1e-5


@pytest.mark.parametrize("layout,p,eta,stop", [("auto", 2.0, 1e-2, "reference"),
                                               ("split", 1.0, 0.0, "reference"),
                                               ("wide", math.inf, 5e-2, "reference"),
                                               ("split", 2.0, 0.0, "reference"),
                                               ("auto", 2.0, 1e-2, "infeas"),
                                               ("split", 1.0, 5e-2, "infeas")])
def test_lamp_source_records_identical(gpu_ctx, layout, p, eta, stop):
    """Mixed-precision LAMP with the staged source records passed between
    passes gives bit-identical logs and outputs to every pass converting its
    own sources (64 x 64 grid pair and a ragged cloud pair), also when the
    infeasibility stop ends the loop at a check."""
    import culamp_amd
    from culamp_amd import workloads
    say("lamp source records", layout, p, eta, stop)
    rng = np.random.default_rng(21)
    cases = [workloads.grid_histograms(64),
             (rng.uniform(size=(700, 3)), rng.uniform(size=(450, 3)), hist(22, 700), hist(23, 450))]
    args = culamp_amd.EOTArgs(eta_p=eta, epsilon=1e-6 if stop == "reference" else 6e-6, B=1.1,
                              itermax=300 if stop == "reference" else 20000, verbose=True,
                              precision="mixed", stop=stop)
    gpu_ctx.set_layout(layout)
    try:
        for x, y, r, c in cases:
            pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
            out = {}
            for rec in (True, False):
                gpu_ctx.set_source_records(rec)
                lines = []
                res = pb.lamp(args, 25, p, lines.append)
                out[rec] = ([l.split(",", 1)[1] for l in lines[1:]], res)
            pb.close()
            assert out[True][0] == out[False][0]
            for a, b in zip(out[True][1][:5], out[False][1][:5]):
                assert np.array_equal(a, b)
            assert out[True][1][5]["iterations"] == out[False][1][5]["iterations"]
            if stop == "infeas":
                assert out[True][1][5]["converged"] == 1
    finally:
        gpu_ctx.set_source_records(True)
        gpu_ctx.set_layout("auto")


@pytest.mark.parametrize("n,m", [(4096, 4096), (700, 450)])
def test_planar_solves_match_general_loop(gpu_ctx, n, m):
    """Mixed LAMP and Sinkhorn on planar clouds (z == 0, whose squared-
    Euclidean passes run the loop without z terms) against the same clouds
    with z = 1e-30 on every 61st target point, a z whose square vanishes in
    FP32 (and is absorbed by every FP64 cost of these distinct points) but
    which sends the tiles holding it through the general loop: identical
    logs and outputs."""
    import culamp_amd
    say("planar solves", n, m)
    rng = np.random.default_rng(41)
    x, y = rng.uniform(size=(n, 3)), rng.uniform(size=(m, 3))
    x[:, 2] = 0.0
    y[:, 2] = 0.0
    yz = y.copy()
    yz[::61, 2] = 1e-30
    r, c = hist(42, n), hist(43, m)
    lamp = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=True,
                              precision="mixed", stop="infeas")
    sk = culamp_amd.EOTArgs(eta_p=2e-2, epsilon=1e-6, itermax=2000, verbose=True,
                            precision="mixed", stop="infeas")
    out = []
    for ya in (y, yz):
        pb = culamp_amd.Problem(x, ya, r, c, gpu_ctx)
        try:
            l1, l2 = [], []
            a = pb.lamp(lamp, 10, 2.0, l1.append)
            b = pb.sinkhorn(sk, 10, 2.0, log=l2.append)
        finally:
            pb.close()
        out.append(([l.split(",", 1)[1] for l in l1[1:]], [l.split(",", 1)[1] for l in l2[1:]],
                    a, b))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    for k, nk in ((2, 5), (3, 4)):
        for a, b in zip(out[0][k][:nk], out[1][k][:nk]):
            if a is not None and a.ndim == 2 and a.shape[1] == 3:
                # barycentric projections: z averages the points' own z
                a, b = a[:, :2], b[:, :2]
            assert np.array_equal(a, b)
    assert out[0][2][5]["iterations"] == out[1][2][5]["iterations"]


@pytest.mark.parametrize("p", [2.0, 1.0, math.inf])
def test_sinkhorn_source_records_identical(gpu_ctx, p):
    """Mixed-precision Sinkhorn and annealed Sinkhorn with staged source
    records passed between the half steps: bit-identical to every pass
    converting its own sources (ragged point clouds)."""
    import culamp_amd
    say("sinkhorn source records", p)
    rng = np.random.default_rng(31)
    x, y = rng.uniform(size=(900, 3)), rng.uniform(size=(610, 3))
    r, c = hist(32, 900), hist(33, 610)
    pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
    args = culamp_amd.EOTArgs(eta_p=2e-2, epsilon=1e-7, itermax=400, verbose=True,
                              precision="mixed")
    out = {}
    try:
        for rec in (True, False):
            gpu_ctx.set_source_records(rec)
            l1, l2 = [], []
            sk = pb.sinkhorn(args, 10, p, log=l1.append)
            an = pb.annealed_sinkhorn(args.copy(eta_p=0.5, anneal_mult=0.5, epsilon=2e-2), 10, p,
                                      log=l2.append)
            out[rec] = ([l.split(",", 1)[1] for l in l1[1:]], [l.split(",", 1)[1] for l in l2[1:]],
                        sk, an)
    finally:
        gpu_ctx.set_source_records(True)
        pb.close()
    assert out[True][0] == out[False][0] and out[True][1] == out[False][1]
    for k in (2, 3):
        for a, b in zip(out[True][k][:4], out[False][k][:4]):
            assert np.array_equal(a, b)
